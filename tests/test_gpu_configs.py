"""GPU parity at the benchmark workloads' own sizes.

Every call goes through the Fortran-ABI entry points with the host
fall-through off (cutoff 0), so the gfx950 kernels compute every result; the
large shapes select the 128 x 128 GEMM tile kernels (FP64 and FP32, all four
trans patterns), the triangular-grid SYRK and the multi-level TRSM recursion.

Bars (stated per test):
  * integer-valued operands whose every partial sum stays below 2^24 (FP32)
    or 2^53 (FP64): bit-exact against float64 NumPy products, whatever the
    summation order;
  * uniform[-1,1] operands: the netlib DMMCH gauge ratio
    max|C - Cref| / (eps * (|alpha| |op(A)| |op(B)| + |beta| |C0|)) < 16
    (tests/netlib/dblat3.f:2538-2546) on sampled columns, Cref from the oracle;
  * DTRSM: the normwise residual ||A X - alpha B|| / (||A|| ||X|| m eps) <= 4;
  * the Level-1/2 chain at n = 8192 under LD_PRELOAD: idamax bit-exact with the
    oracle on the same SplitMix64 inputs, ddot / dnrm2 within the rounding
    bound of their sums.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import libgpublas_amd as gbmod
from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def _ints(rng, shape, lo=-4, hi=4, dtype=np.float64):
    return np.asfortranarray(rng.integers(lo, hi + 1, size=shape).astype(dtype))


def _op(a, t):
    return a if t == "N" else a.T


# ---------------------------------------------------------------------------
# Level-3 sweep: GEMM at M, N, K in {4096, 8192}, all four trans, FP64 + FP32
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(4096, 4096, 4096), (8192, 8192, 4096)])
def test_gemm_sweep_integer_exact(gpu, prec, ta, tb, m, n, k):
    """|entries| <= 4: every partial sum is an integer below 16 k + 4 |beta| < 2^24,
    exact in FP32 and FP64 in any order, so the GPU result must equal the
    float64 product bit for bit."""
    dt = np.float64 if prec == "d" else np.float32
    rng = np.random.default_rng(1000 + m // 1024 + 7 * (ta == "T") + 13 * (tb == "T") + 29 * (prec == "s"))
    A = _ints(rng, (m, k) if ta == "N" else (k, m), dtype=dt)
    B = _ints(rng, (k, n) if tb == "N" else (n, k), dtype=dt)
    C0 = _ints(rng, (m, n), dtype=dt)
    C = C0.copy(order="F")
    before = gpu.stats()["device_calls"]
    fn = gpu.dgemm if prec == "d" else gpu.sgemm
    fn(ta, tb, m, n, k, 1.0, A, A.shape[0], B, B.shape[0], -2.0, C, m)
    assert gpu.stats()["device_calls"] == before + 1
    ref = _op(A, ta).astype(np.float64) @ _op(B, tb).astype(np.float64) - 2.0 * C0.astype(np.float64)
    assert np.array_equal(C.astype(np.float64), ref)


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_gemm_sweep_uniform_gauge(gpu, ora, prec, ta, tb):
    """4096^3, U[-1,1], alpha 1, beta 0.5: the long-k steady state of the
    128 x 128 kernels; DMMCH gauge ratio < 16 on 48 sampled columns of C
    against the oracle (dblat3.f:2457-2578)."""
    n = 4096
    dt = np.float64 if prec == "d" else np.float32
    A = uniform(301 + (prec == "s"), (n, n), dt)
    B = uniform(303 + (prec == "s"), (n, n), dt)
    C0 = uniform(305 + (prec == "s"), (n, n), dt)
    C = C0.copy(order="F")
    (gpu.dgemm if prec == "d" else gpu.sgemm)(ta, tb, n, n, n, 1.0, A, n, B, n, 0.5, C, n)
    cols = np.random.default_rng(7).choice(n, 48, replace=False)
    # columns j of C need columns j of op(B): rows j of B when tb = 'T'
    Bs = np.asfortranarray((B[:, cols] if tb == "N" else B[cols, :]).astype(np.float64))
    ldb = Bs.shape[0]
    A64 = np.asfortranarray(A.astype(np.float64))
    C0s = np.asfortranarray(C0[:, cols].astype(np.float64))
    Cs = np.asfortranarray(C[:, cols].astype(np.float64))
    eps = EPS64 if prec == "d" else EPS32
    r = ora.ora_gemm_gauge_ratio(c(ta), c(tb), n, len(cols), n, 1.0, p(A64), n, p(Bs), ldb, 0.5,
                                 p(C0s), p(Cs), n, eps)
    assert r < 16.0, r


# ---------------------------------------------------------------------------
# DSYRK lower, n = 8192, k = 2048
# ---------------------------------------------------------------------------
def test_dsyrk_sweep_integer_exact(gpu):
    n, k = 8192, 2048
    rng = np.random.default_rng(77)
    A = _ints(rng, (n, k))
    C0 = _ints(rng, (n, n))
    C = C0.copy(order="F")
    before = gpu.stats()["device_calls"]
    gpu.dsyrk("L", "N", n, k, 1.0, A, n, 3.0, C, n)
    assert gpu.stats()["device_calls"] == before + 1
    ref = A @ A.T + 3.0 * C0
    low = np.tril(np.ones((n, n), bool))
    assert np.array_equal(C[low], ref[low])
    assert np.array_equal(C[~low], C0[~low])  # the upper triangle is not touched


# ---------------------------------------------------------------------------
# DTRSM left / lower / no-trans / non-unit, m = 8192, nrhs = 4096
# ---------------------------------------------------------------------------
def test_dtrsm_sweep_residual(gpu):
    """A diagonally dominant (diag = m + U[0,1], off-diagonal U[-1,1]), the
    strict upper triangle NaN (never read); normwise residual <= 4."""
    m, nrhs = 8192, 4096
    A = uniform(401, (m, m))
    A[np.diag_indices(m)] = m + 0.5 * (uniform(402, (m,)) + 1.0)
    A[np.triu_indices(m, 1)] = np.nan
    A = np.asfortranarray(A)
    B0 = uniform(403, (m, nrhs))
    B = B0.copy(order="F")
    alpha = 0.75
    before = gpu.stats()["device_calls"]
    gpu.dtrsm("L", "L", "N", "N", m, nrhs, alpha, A, m, B, m)
    assert gpu.stats()["device_calls"] == before + 1
    L = np.tril(np.nan_to_num(A, nan=0.0))
    res = np.linalg.norm(L @ B - alpha * B0) / (np.linalg.norm(L) * np.linalg.norm(B) * m * EPS64)
    assert np.all(np.isfinite(B))
    assert res <= 4.0, res


@pytest.mark.parametrize("side,uplo,ta", [("L", "L", "N"), ("R", "U", "T")])
def test_dtrsm_integer_exact_multilevel(gpu, side, uplo, ta):
    """Unit triangle with entries in {-1, 0, 1} and an integer solution X:
    B = op(A) X (or X op(A)) is exact, and so is every step of the blocked
    solve (2048 rows: three halvings above the 256-row base solves)."""
    rng = np.random.default_rng(88)
    na, nrhs = 2048, 384
    T = rng.integers(-1, 2, size=(na, na)).astype(float)
    T = np.tril(T, -1) if uplo == "L" else np.triu(T, 1)
    T += np.eye(na)
    A = T.copy()
    A[np.diag_indices(na)] = np.nan  # unit diagonal: never read
    A = np.asfortranarray(A)
    opT = T if ta == "N" else T.T
    if side == "L":
        X = rng.integers(-3, 4, size=(na, nrhs)).astype(float)
        B = np.asfortranarray(opT @ X)
        m, n = na, nrhs
    else:
        X = rng.integers(-3, 4, size=(nrhs, na)).astype(float)
        B = np.asfortranarray(X @ opT)
        m, n = nrhs, na
    gpu.dtrsm(side, uplo, ta, "U", m, n, 1.0, A, na, B, m)
    assert np.array_equal(B, X)


# ---------------------------------------------------------------------------
# Level-1/2 chain, n = 8192
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgemv_8192_gauge(gpu, ora, trans):
    n = 8192
    A = uniform(501, (n, n))
    x = uniform(502, (n,))
    y0 = uniform(503, (n,))
    y = y0.copy()
    R = y0.copy()
    gpu.dgemv(trans, n, n, 1.3, A, n, x, 1, 0.5, y, 1)
    ora.ora_dgemv(c(trans), n, n, 1.3, p(A), n, p(x), 1, 0.5, p(R), 1)
    G = 1.3 * (np.abs(_op(A, trans)) @ np.abs(x)) + 0.5 * np.abs(y0)
    assert np.max(np.abs(y - R) / (EPS64 * G)) < 16.0
    err = np.linalg.norm(y - R) / (np.linalg.norm(A) * np.linalg.norm(x) * n * EPS64)
    assert err <= 4.0, err


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgemv_8192_integer_exact(gpu, trans):
    n = 8192
    rng = np.random.default_rng(504)
    A = _ints(rng, (n, n))
    x = rng.integers(-4, 5, size=n).astype(float)
    y = rng.integers(-4, 5, size=n).astype(float)
    ref = 2.0 * (_op(A, trans) @ x) - y
    gpu.dgemv(trans, n, n, 2.0, A, n, x, 1, -1.0, y, 1)
    assert np.array_equal(y, ref)


def _oracle_chain(ora, n, iters, seed):
    A = uniform(seed, (n, n))
    x = uniform(seed + 1, (n,))
    y = np.zeros(n)
    z = uniform(seed + 3, (n,))
    d = r = im = None
    for _ in range(iters):
        ora.ora_dgemv(b"N", n, n, 1.0, p(A), n, p(x), 1, 0.0, p(y), 1)
        ora.ora_daxpy(n, 0.3, p(y), 1, p(z), 1)
        d = ora.ora_ddot(n, p(y), 1, p(z), 1)
        r = ora.ora_dnrm2(n, p(z), 1)
        im = ora.ora_idamax(n, p(z), 1)
    return A, x, y, z, d, r, im


def test_chain_8192_preload_vs_oracle(ora):
    """The chain driver (malloc + Fortran BLAS calls, nothing GPU-aware) under
    LD_PRELOAD with the object-tracking file naming its four operands, 3
    iterations at n = 8192: idamax equals the oracle's, ddot and dnrm2 within
    their rounding bounds, and the operands stay tracked."""
    n, iters, seed = 8192, 3, 20240601
    with tempfile.TemporaryDirectory() as td:
        ot = os.path.join(td, "chain.objtrack")
        with open(ot, "w") as f:
            f.write("D #0 A\nD #1 x\nD #2 y\nD #3 z\n")
        env = dict(os.environ)
        env["LD_PRELOAD"] = gbmod.PRELOAD_PATH
        env["GPUBLAS_OPTIONS"] = f"heuristic=objtrack:{ot};cutoff=0"
        out = subprocess.run([gbmod.CHAIN_BENCH, "--n", str(n), "--iters", str(iters),
                              "--warmup", "0", "--seed", str(seed)],
                             env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["tracked"] == 4
    A, x, y, z, d, r, im = _oracle_chain(ora, n, iters, seed)
    az = np.abs(z)
    top = np.sort(az)[-2:]
    assert top[1] - top[0] > 1e-9 * top[1]  # the maximum is not a near-tie
    assert res["idamax"] == im
    assert abs(res["ddot"] - d) <= 1e-11 * float(np.abs(y) @ az)
    assert abs(res["dnrm2"] - r) <= 1e-12 * r
    # the driver's checksum sum(y + z) read back through the host view
    cs = [ln for ln in out.stderr.splitlines() if "checksum" in ln]
    got = float(cs[-1].split()[-1])
    assert abs(got - float(np.sum(y + z))) <= 1e-10 * float(np.sum(np.abs(y) + az))


def test_level1_8192_vs_oracle(gpu, ora):
    n = 8192
    y = uniform(601, (n,)) * 50.0
    z = uniform(602, (n,))
    z[1234] = 3.5
    z[4321] = -3.5  # equal magnitudes: the first index wins
    yt, zt = gpu.tracked_copy(y), gpu.tracked_copy(z)
    z0 = z.copy()
    gpu.daxpy(n, 0.3, yt, 1, zt, 1)
    ora.ora_daxpy(n, 0.3, p(y), 1, p(z), 1)
    # one fused multiply-add against a multiply and an add: at most one rounding apart
    assert np.all(np.abs(np.array(zt) - z) <= EPS64 * (np.abs(0.3 * y) + np.abs(z0)))
    assert gpu.idamax(n, zt, 1) == ora.ora_idamax(n, p(z), 1)
    d = gpu.ddot(n, yt, 1, zt, 1)
    assert abs(d - ora.ora_ddot(n, p(y), 1, p(z), 1)) <= n * EPS64 * float(np.abs(y) @ np.abs(z))
    r = gpu.dnrm2(n, zt, 1)
    assert abs(r - ora.ora_dnrm2(n, p(z), 1)) <= n * EPS64 * r
