"""cutoff=auto on the GPU (route.c).

A call whose flops and flops-per-byte sit exactly at its routine's crossover
row runs on the device; with min_flops raised by one part in a million it runs
on the host BLAS.  Both results are checked: the host one bit-identical to
scipy's OpenBLAS (the same routine), the device one within the kernels' usual
tolerance of it.  The measurement (gpublas_tune_crossover) must write a table
that loads back unchanged through xover=<file>."""
import math

import numpy as np
import pytest
import scipy.linalg.blas as sb

import libgpublas_amd as gb
from _oracle import uniform

pytestmark = pytest.mark.gpu


@pytest.fixture
def auto():
    if not gb.device_ready():
        pytest.fail("no gfx950 device visible to libgpublas")
    saved = {r: gb.get_crossover(r) for r in gb.CROSSOVER_ROUTINES}
    old = gb.get_cutoff()
    gb.set_cutoff(gb.CUTOFF_AUTO)
    for r in gb.CROSSOVER_ROUTINES:
        gb.set_crossover(r, math.inf, math.inf)
    yield
    gb.set_cutoff(old)
    for r, (f, i) in saved.items():
        gb.set_crossover(r, f, i)


M, N, K = 70, 50, 40


def _case(name):
    """(call(), reference, flops, operand bytes, tolerance) of one routine."""
    A, B, C = uniform(1, (M, K)), uniform(2, (K, N)), uniform(3, (M, N))
    S = uniform(4, (M, M))
    T = np.asfortranarray(uniform(5, (M, M)) + np.diag(np.full(M, 4.0)))
    Bm = uniform(6, (M, N))
    if name == "dgemm":
        out = C.copy(order="F")
        return (lambda: gb.dgemm("N", "N", M, N, K, 0.5, A, M, B, K, 1.5, out, M), out,
                sb.dgemm(0.5, A, B, 1.5, C.copy(order="F")), 2.0 * M * N * K,
                8.0 * (M * K + K * N + 2 * M * N), 1e-13)
    if name == "sgemm":
        As, Bs, Cs = (x.astype(np.float32, order="F") for x in (A, B, C))
        out = Cs.copy(order="F")
        return (lambda: gb.sgemm("N", "N", M, N, K, 0.5, As, M, Bs, K, 1.5, out, M), out,
                sb.sgemm(0.5, As, Bs, 1.5, Cs.copy(order="F")), 2.0 * M * N * K,
                4.0 * (M * K + K * N + 2 * M * N), 2e-5)
    if name == "dsyrk":
        out = S.copy(order="F")
        return (lambda: gb.dsyrk("L", "N", M, K, 0.5, A, M, 1.5, out, M), out,
                sb.dsyrk(0.5, A, 1.5, S.copy(order="F"), lower=1), 1.0 * M * M * K,
                8.0 * (M * K + 2 * M * M), 1e-13)
    if name == "dsyr2k":
        A2 = uniform(7, (M, K))
        out = S.copy(order="F")
        return (lambda: gb.dsyr2k("U", "N", M, K, 0.5, A, M, A2, M, 1.5, out, M), out,
                sb.dsyr2k(0.5, A, A2, 1.5, S.copy(order="F")), 2.0 * M * M * K,
                8.0 * (2 * M * K + 2 * M * M), 1e-13)
    if name == "dsymm":
        out = Bm.copy(order="F")
        return (lambda: gb.dsymm("L", "U", M, N, 0.5, S, M, C, M, 1.5, out, M), out,
                sb.dsymm(0.5, S, C, 1.5, Bm.copy(order="F")), 2.0 * M * N * M,
                8.0 * (M * M + 3 * M * N), 1e-13)
    if name == "dtrsm":
        out = Bm.copy(order="F")
        return (lambda: gb.dtrsm("L", "L", "N", "N", M, N, 0.5, T, M, out, M), out,
                sb.dtrsm(0.5, T, Bm.copy(order="F"), lower=1), 1.0 * M * N * M,
                8.0 * (M * M + 2 * M * N), 1e-13)
    if name == "dtrmm":
        out = Bm.copy(order="F")
        return (lambda: gb.dtrmm("R", "U", "N", "N", M, N, 0.5, uniform(8, (N, N)), N, out, M),
                out, sb.dtrmm(0.5, uniform(8, (N, N)), Bm.copy(order="F"), side=1), 1.0 * M * N * N,
                8.0 * (N * N + 2 * M * N), 1e-13)
    if name == "dgemv":
        x, y = uniform(9, (N,)), uniform(10, (M,))
        out = y.copy()
        return (lambda: gb.dgemv("N", M, N, 1.0, C, M, x, 1, 0.5, out, 1), out,
                sb.dgemv(1.0, C, x, 0.5, y.copy()), 2.0 * M * N, 8.0 * (M * N + M + N), 1e-13)
    if name == "daxpy":
        x, y = uniform(11, (5000,)), uniform(12, (5000,))
        out = y.copy()
        return (lambda: gb.daxpy(5000, 0.25, x, 1, out, 1), out,
                sb.daxpy(x, y.copy(), a=0.25), 2.0 * 5000, 24.0 * 5000, 1e-15)
    raise KeyError(name)


ROUTED = ["dgemm", "sgemm", "dsyrk", "dsyr2k", "dsymm", "dtrsm", "dtrmm", "dgemv", "daxpy"]


@pytest.mark.parametrize("name", ROUTED)
@pytest.mark.parametrize("side", ["gpu", "host"])
def test_auto_routes_at_the_threshold(auto, name, side):
    call, out, ref, flops, nbytes, tol = _case(name)
    gb.set_crossover(name, flops * (1.0 if side == "gpu" else 1.000001), flops / nbytes)
    gb.reset_stats()
    call()
    st = gb.stats()
    if side == "gpu":
        assert st["device_calls"] == 1 and st["fallthrough"] == 0, st
        np.testing.assert_allclose(out, ref, rtol=tol, atol=tol)
    else:
        assert st["fallthrough"] == 1 and st["device_calls"] == 0, st
        assert np.array_equal(out, ref)


def test_tune_crossover_table_loads_back(auto, tmp_path):
    path = tmp_path / "xover.txt"
    gb.tune_crossover(path, 128)
    text = path.read_text()
    rows = {}
    timed = set()
    for line in text.splitlines():
        if line.startswith("# ") and len(line.split()) == 7 and line.split()[2].isdigit():
            _, r, n, fl, by, th, tg = line.split()
            timed.add(r)
            assert float(th) > 0 and float(tg) > 0
        elif line and not line.startswith("#"):
            r, f, i = line.split()
            rows[r] = (float(f), float(i))
    assert set(rows) == set(gb.CROSSOVER_ROUTINES) == timed
    for r, (f, i) in rows.items():
        gf, gi = gb.get_crossover(r)  # the measurement installs what it wrote
        assert math.isclose(f, gf, rel_tol=1e-8) and math.isclose(i, gi, rel_tol=1e-8)
        gb.set_crossover(r, 0.0, 0.0)
    assert gb.set_options(f"xover={path}") == 0
    for r, (f, i) in rows.items():
        gf, gi = gb.get_crossover(r)
        assert math.isclose(f, gf, rel_tol=1e-8) and math.isclose(i, gi, rel_tol=1e-8)
    # an 8^3 GEMM through the host link never beats the host BLAS
    assert not gb.crossover_gpu("dgemm", 2.0 * 8 ** 3, 32.0 * 8 ** 2)
