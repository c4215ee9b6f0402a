"""Serving loop: batched grouped launches of DGEMM / SGEMM / DGEMV / left
DTRSM calls from many threads, against the CPU oracle.

Bars:
  * every served result meets the per-call bars: GEMM the netlib DMMCH gauge
    ratio < 16 (tests/netlib/dblat3.f:2538-2546); GEMV the same gauge,
    |y - yref| <= 16 eps (|alpha| |op(A)| |x| + |beta| |y0|) elementwise;
    TRSM the normwise residual ||op(A) X - alpha B|| / (||A|| ||X|| m eps) <= 4
    and agreement with the oracle's solve;
  * a call's result is bit-identical alone and in a mix of concurrent calls
    from 16 threads (tolerance 0: the batch kernels give every problem the
    same work decomposition whatever else is in the batch);
  * graph replay and eager launches give identical bits; every GEMM tile
    shape gives identical bits (the k order does not depend on the tile);
  * a small HBM budget forces evictions and re-uploads without changing
    results; the statistics show the batching (served, batches, max_batch).
"""
import os
import tempfile
import threading

import numpy as np
import pytest

from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


@pytest.fixture()
def serve(gpu):
    assert gpu.set_options("serve=64G") == 0
    yield gpu
    gpu.set_options("serve=0")


def _gemm_case(rng, seed, single=False, maxdim=300):
    m, n, k = (int(v) for v in rng.integers(1, maxdim, size=3))
    ta, tb = rng.choice(["N", "T"]), rng.choice(["N", "T"])
    dt = np.float32 if single else np.float64
    A = uniform(seed, (m, k) if ta == "N" else (k, m)).astype(dt)
    B = uniform(seed + 1, (k, n) if tb == "N" else (n, k)).astype(dt)
    C0 = uniform(seed + 2, (m, n)).astype(dt)
    alpha, beta = float(rng.choice([1.0, -0.5, 2.0])), float(rng.choice([0.0, 0.5, -1.0]))
    return dict(ta=str(ta), tb=str(tb), m=m, n=n, k=k, A=np.asfortranarray(A),
                B=np.asfortranarray(B), C0=np.asfortranarray(C0), alpha=alpha, beta=beta,
                single=single)


def _gemm_ratio(ora, cs, Cgot):
    A, B, C0 = cs["A"], cs["B"], cs["C0"]
    if cs["single"]:
        A, B, C0, Cgot = (x.astype(np.float64, order="F") for x in (A, B, C0, Cgot))
    eps = EPS32 if cs["single"] else EPS64
    Cgot = np.asfortranarray(Cgot)
    return ora.ora_gemm_gauge_ratio(c(cs["ta"]), c(cs["tb"]), cs["m"], cs["n"], cs["k"],
                                    cs["alpha"], p(A), A.shape[0], p(B), B.shape[0], cs["beta"],
                                    p(C0), p(Cgot), cs["m"], eps)


def _run_gemm(gb, cs):
    A, B = gb.tracked_copy(cs["A"]), gb.tracked_copy(cs["B"])
    C = gb.tracked_copy(cs["C0"])
    f = gb.sgemm if cs["single"] else gb.dgemm
    f(cs["ta"], cs["tb"], cs["m"], cs["n"], cs["k"], cs["alpha"], A, A.shape[0], B, B.shape[0],
      cs["beta"], C, cs["m"])
    return np.array(C)


def _gemv_case(rng, seed):
    m, n = (int(v) for v in rng.integers(1, 700, size=2))
    tr = str(rng.choice(["N", "T"]))
    incx, incy = int(rng.choice([1, -1, 2, -3])), int(rng.choice([1, -1, 2]))
    lenx, leny = (n, m) if tr == "N" else (m, n)
    A = uniform(seed, (m, n))
    x = uniform(seed + 1, (1 + (lenx - 1) * abs(incx),))
    y0 = uniform(seed + 2, (1 + (leny - 1) * abs(incy),))
    alpha, beta = float(rng.choice([1.0, -0.7])), float(rng.choice([0.0, 0.5]))
    return dict(tr=tr, m=m, n=n, A=A, x=x, y0=y0, incx=incx, incy=incy, alpha=alpha, beta=beta)


def _gemv_check(ora, cs, ygot):
    yref = cs["y0"].copy()
    ora.ora_dgemv(c(cs["tr"]), cs["m"], cs["n"], cs["alpha"], p(cs["A"]), cs["m"], p(cs["x"]),
                  cs["incx"], cs["beta"], p(yref), cs["incy"])
    # gauge: the same product on absolute values (arrays kept alive across the call)
    g = np.abs(cs["y0"]).copy()
    absA, absx = np.asfortranarray(np.abs(cs["A"])), np.abs(cs["x"])
    ora.ora_dgemv(c(cs["tr"]), cs["m"], cs["n"], abs(cs["alpha"]), p(absA), cs["m"], p(absx),
                  cs["incx"], abs(cs["beta"]), p(g), cs["incy"])
    assert np.all(np.abs(ygot - yref) <= 16 * EPS64 * g + 1e-300)


def _run_gemv(gb, cs):
    A, x, y = gb.tracked_copy(cs["A"]), gb.tracked_copy(cs["x"]), gb.tracked_copy(cs["y0"])
    gb.dgemv(cs["tr"], cs["m"], cs["n"], cs["alpha"], A, cs["m"], x, cs["incx"], cs["beta"], y,
             cs["incy"])
    return np.array(y)


def _trsm_case(rng, seed, uplo=None, ta=None, diag=None):
    m, n = int(rng.integers(1, 400)), int(rng.integers(1, 90))
    uplo = uplo or str(rng.choice(["L", "U"]))
    ta = ta or str(rng.choice(["N", "T"]))
    diag = diag or str(rng.choice(["N", "U"]))
    A = uniform(seed, (m, m))
    A[np.diag_indices(m)] = m + 0.5 * (uniform(seed + 1, (m,)) + 1.0)
    B0 = uniform(seed + 2, (m, n))
    return dict(uplo=uplo, ta=ta, diag=diag, m=m, n=n, A=np.asfortranarray(A), B0=B0,
                alpha=float(rng.choice([1.0, -2.0])))


def _trsm_check(ora, cs, X):
    m = cs["m"]
    ref = cs["B0"].copy(order="F")
    ora.ora_dtrsm(c("L"), c(cs["uplo"]), c(cs["ta"]), c(cs["diag"]), m, cs["n"], cs["alpha"],
                  p(cs["A"]), m, p(ref), m)
    T = np.tril(cs["A"]) if cs["uplo"] == "L" else np.triu(cs["A"])
    if cs["diag"] == "U":
        T[np.diag_indices(m)] = 1.0
    opA = T if cs["ta"] == "N" else T.T
    res = np.linalg.norm(opA @ X - cs["alpha"] * cs["B0"])
    assert res / (np.linalg.norm(T) * np.linalg.norm(X) * m * EPS64) <= 4
    assert np.linalg.norm(X - ref) <= 64 * m * EPS64 * np.linalg.norm(ref)


def _run_trsm(gb, cs):
    A, B = gb.tracked_copy(cs["A"]), gb.tracked_copy(cs["B0"])
    gb.dtrsm("L", cs["uplo"], cs["ta"], cs["diag"], cs["m"], cs["n"], cs["alpha"], A, cs["m"], B,
             cs["m"])
    return np.array(B)


def test_served_gemm_matches_oracle(serve, ora):
    rng = np.random.default_rng(1)
    serve.reset_stats()
    for i in range(16):
        cs = _gemm_case(rng, 100 + 10 * i, single=bool(i % 2))
        got = _run_gemm(serve, cs)
        assert _gemm_ratio(ora, cs, got) < 16, cs
    assert serve.stats()["served"] == 16


def test_served_gemv_matches_oracle(serve, ora):
    rng = np.random.default_rng(2)
    serve.reset_stats()
    for i in range(16):
        cs = _gemv_case(rng, 300 + 10 * i)
        _gemv_check(ora, cs, _run_gemv(serve, cs))
    assert serve.stats()["served"] == 16


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_served_trsm_matches_oracle(serve, ora, uplo, ta, diag):
    rng = np.random.default_rng(3)
    serve.reset_stats()
    for i in range(3):
        cs = _trsm_case(rng, 500 + 10 * i, uplo, ta, diag)
        _trsm_check(ora, cs, _run_trsm(serve, cs))
    assert serve.stats()["served"] == 3


def test_right_side_trsm_takes_the_per_call_path(serve, ora):
    m, n = 70, 50
    A = uniform(41, (n, n))
    A[np.diag_indices(n)] = n + 1.0
    B0 = uniform(42, (m, n))
    At, B = serve.tracked_copy(A), serve.tracked_copy(B0)
    serve.reset_stats()
    serve.dtrsm("R", "L", "N", "N", m, n, 1.0, At, n, B, m)
    assert serve.stats()["served"] == 0 and serve.stats()["device_calls"] == 1
    ref = B0.copy(order="F")
    ora.ora_dtrsm(c("R"), c("L"), c("N"), c("N"), m, n, 1.0, p(A), n, p(ref), m)
    assert np.linalg.norm(np.array(B) - ref) <= 1e-13 * np.linalg.norm(ref)


def _mixed_cases(seed, count, maxdim=300):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        r = rng.random()
        s = 1000 * seed + 10 * i
        if r < 0.6:
            out.append(("gemm", _gemm_case(rng, s, single=bool(rng.random() < 0.5), maxdim=maxdim)))
        elif r < 0.9:
            out.append(("gemv", _gemv_case(rng, s)))
        else:
            out.append(("trsm", _trsm_case(rng, s)))
    return out


def _run(gb, kind, cs):
    return {"gemm": _run_gemm, "gemv": _run_gemv, "trsm": _run_trsm}[kind](gb, cs)


def test_same_result_alone_and_in_the_mix(serve, ora):
    """Each probe call alone (one thread, batch of one) and again while 16
    threads keep the loop busy with other calls: bit-identical."""
    probes = _mixed_cases(7, 24)
    alone = [_run(serve, k, cs) for k, cs in probes]
    for (k, cs), got in zip(probes, alone):
        if k == "gemm":
            assert _gemm_ratio(ora, cs, got) < 16
        elif k == "gemv":
            _gemv_check(ora, cs, got)
        else:
            _trsm_check(ora, cs, got)

    serve.reset_stats()
    mixed = [None] * len(probes)
    errors = []

    def noise(tid):
        try:
            for k, cs in _mixed_cases(100 + tid, 12):
                _run(serve, k, cs)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    def prober(idx):
        try:
            for i in range(idx, len(probes), 4):
                mixed[i] = _run(serve, *probes[i])
        except Exception as e:  # pragma: no cover
            errors.append(e)

    ths = [threading.Thread(target=noise, args=(t,)) for t in range(12)]
    ths += [threading.Thread(target=prober, args=(t,)) for t in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b, equal_nan=True)
    st = serve.stats()
    assert st["served"] == len(probes) + 12 * 12
    assert st["max_batch"] >= 2  # calls really shared launches
    assert st["batches"] < st["served"]


def test_graph_and_eager_launches_agree(gpu, ora):
    cases = _mixed_cases(11, 12)
    outs = []
    for graph in (1, 0):
        assert gpu.set_options(f"serve=64G;serve_graph={graph}") == 0
        outs.append([_run(gpu, k, cs) for k, cs in cases])
    gpu.set_options("serve=0;serve_graph=0")
    for a, b in zip(*outs):
        assert np.array_equal(a, b, equal_nan=True)


def test_tile_shapes_give_identical_bits(gpu):
    """Device-level batch API: one GEMM computed with each tile shape."""
    gb = gpu
    m, n, k = 200, 170, 333
    A0, B0 = uniform(61, (m, k)), uniform(62, (n, k))
    res = []
    for cfg in (0, 1, 2):
        A, B = gb.tracked_copy(A0), gb.tracked_copy(B0)
        C = gb.tracked_copy(uniform(63, (m, n)))
        b = gb.Batch(use_graph=False)
        b.force_cfg(cfg)
        assert b.dgemm("N", "T", m, n, k, 1.5, gb.device_ptr(A), m, gb.device_ptr(B), n, 0.25,
                       gb.device_ptr(C), m) == 0
        b.run()
        b.close()
        res.append(np.array(C))
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[1], res[2])
    ref = A0 @ B0.T * 1.5 + 0.25 * uniform(63, (m, n))
    assert np.max(np.abs(res[0] - ref)) < 1e-12


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
@pytest.mark.parametrize("single", [False, True])
def test_split_k_products(gpu, ora, ta, tb, single):
    """Few tiles, long k: the batch kernels split k into chunks and fold them
    in order.  Same bits for every tile shape and alone or batched with
    other products; DMMCH gauge against the oracle."""
    gb = gpu
    rng = np.random.default_rng(17)
    m, n, k = 48, 40, 3000
    dt = np.float32 if single else np.float64
    A0 = uniform(91, (m, k) if ta == "N" else (k, m)).astype(dt)
    B0 = uniform(92, (k, n) if tb == "N" else (n, k)).astype(dt)
    C0 = uniform(93, (m, n)).astype(dt)
    cs = dict(ta=ta, tb=tb, m=m, n=n, k=k, A=np.asfortranarray(A0), B=np.asfortranarray(B0),
              C0=np.asfortranarray(C0), alpha=1.5, beta=-0.5, single=single)
    outs = []
    for cfg, others in ((0, 0), (1, 0), (2, 0), (1, 5)):
        if single and cfg == 0:
            continue
        A, B, C = gb.tracked_copy(cs["A"]), gb.tracked_copy(cs["B"]), gb.tracked_copy(cs["C0"])
        b = gb.Batch(use_graph=False)
        b.force_cfg(cfg)
        add = b.sgemm if single else b.dgemm
        keep = []
        for i in range(others):  # unrelated products in the same launch
            oc = _gemm_case(rng, 900 + 10 * i, single=single, maxdim=200)
            oA, oB, oC = (gb.tracked_copy(oc[x]) for x in ("A", "B", "C0"))
            keep += [oA, oB, oC]
            assert add(oc["ta"], oc["tb"], oc["m"], oc["n"], oc["k"], oc["alpha"],
                       gb.device_ptr(oA), oA.shape[0], gb.device_ptr(oB), oB.shape[0],
                       oc["beta"], gb.device_ptr(oC), oc["m"]) == 0
        assert add(ta, tb, m, n, k, 1.5, gb.device_ptr(A), A.shape[0], gb.device_ptr(B),
                   B.shape[0], -0.5, gb.device_ptr(C), m) == 0
        b.run()
        b.close()
        outs.append(np.array(C))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    assert _gemm_ratio(ora, cs, outs[0]) < 16


def test_batch_rejects_what_it_cannot_run(gpu):
    gb = gpu
    b = gb.Batch(use_graph=False)
    A = gb.tracked_copy(uniform(71, (8, 8)))
    d = gb.device_ptr(A)
    assert b.dgemm("N", "N", 8, 8, 8, 0.0, d, 8, d, 8, 1.0, d, 8) == -1  # alpha == 0
    assert b.dgemm("N", "N", 8, 8, 0, 1.0, d, 8, d, 8, 1.0, d, 8) == -1  # k == 0
    assert b.dtrsm("R", "L", "N", "N", 8, 8, 1.0, d, 8, d, 8) == -1      # right side
    assert b.dtrsm("L", "L", "N", "N", 1025, 8, 1.0, d, 1025, d, 1025) == -1  # m > 1024
    for _ in range(32):
        assert b.dgemv("N", 8, 8, 1.0, d, 8, d, 1, 0.0, d, 1) == 0
    assert b.dgemv("N", 8, 8, 1.0, d, 8, d, 1, 0.0, d, 1) == 1          # slot full
    assert b.count() == 32
    b.close()


def test_small_budget_evicts_and_stays_correct(gpu, ora):
    """Budget of 3 MiB with eight 1 MiB operands in rotation: mirrors are
    evicted (host copies are current) and re-uploaded; results unchanged."""
    gb = gpu
    n = 362  # 362^2 * 8 B ~ 1 MiB
    mats = [gb.tracked_copy(uniform(80 + i, (n, n))) for i in range(8)]
    host = [np.array(a) for a in mats]
    assert gb.set_options("serve=3M") == 0
    try:
        gb.reset_stats()
        for i in range(8):
            a, b = mats[i], mats[(i + 3) % 8]
            C = gb.tracked_empty((n, n))
            gb.dgemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, C, n)
            np.testing.assert_allclose(np.array(C), host[i] @ host[(i + 3) % 8], rtol=0,
                                       atol=1e-11)
            del C
        st = gb.stats()
        assert st["served"] == 8 and st["evictions"] > 0
    finally:
        gb.set_options("serve=0")


def test_tuner_writes_and_reloads_the_table(gpu):
    gb = gpu
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.txt")
        assert gb.tune_gemm(path, force=True) == 0
        lines = open(path).read().split("\n")
        rows = [ln.split() for ln in lines if ln.strip()]
        assert len(rows) == 2 * 7 * 7
        assert all(r[0] in "ds" and 0 <= int(r[3]) <= 2 for r in rows)
        assert all(not (r[0] == "s" and int(r[3]) == 0) for r in rows)  # no 32x32 FP32 tile
        choice = gb.tune_choice(False, 2048, 2048, 2048)
        assert gb.tune_gemm(path, force=False) == 0
        assert gb.tune_choice(False, 2048, 2048, 2048) == choice


SCIPY_OPENBLAS = "/usr/local/lib/python3.10/dist-packages/scipy.libs/libscipy_openblas-68440149.so"


@pytest.mark.parametrize("mode", ["serve", "plain"])
def test_serving_trace_verified(gb, mode):
    """The serving trace itself (serve_bench: 16 threads, the 60/30/10
    GEMM / DGEMV / DTRSM mix with M, N, K log-uniform in [32, 2048], DGEMV
    m, n in [256, 16384], DTRSM orders in [64, 1024]), 1,600 calls with every
    call checked against the host BLAS: GEMM and GEMV normwise error
    ||C - Cref|| / (||A|| ||B|| k eps) <= 4, TRSM residual
    ||A X - alpha B0|| / (||A|| ||X|| m eps) <= 4."""
    import json
    import subprocess
    if not os.path.exists(SCIPY_OPENBLAS):
        pytest.skip("the host BLAS (scipy's OpenBLAS) is not installed here")
    cmd = [gb.SERVE_BENCH, "--lib", gb.LIB_PATH, "--threads", "16", "--calls", "1600",
           "--warmup", "2", "--verify", "1", "--verify-blas", SCIPY_OPENBLAS,
           "--verify-prefix", "scipy_"]
    if mode == "plain":
        cmd.append("--plain")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode in (0, 4), out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    total = 0
    for op in ("dgemm", "sgemm", "dgemv", "dtrsm"):
        v = res[f"verify_{op}"]
        assert v["failed"] == 0 and v["max_err"] <= 4.0, (op, v)
        assert v["checked"] > 0, op
        total += v["checked"]
    assert total == 1600
    assert out.returncode == 0
    if mode == "serve":
        assert res["served"] >= 1600
