"""Host-or-GPU routing of calls with no resident operand (route.c), on the CPU.

cutoff=<n> keeps the reference's fixed rule (blas_level3/gemm.cc:129-141);
cutoff=auto decides by a crossover table: a call goes to the GPU only when
its flops reach the routine's min_flops and its flops per operand byte reach
min_intensity.  Here every call the table sends to the host runs on the host
BLAS (scipy's OpenBLAS) and must be bit-identical to calling it directly; the
one call the table sends to the GPU is made in a child process, where the
missing GPU turns it into the library's fatal "no usable gfx950 GPU" abort --
proof that it was routed there."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.linalg.blas as sb

import libgpublas_amd as gb
from _oracle import uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def auto():
    saved = {r: gb.get_crossover(r) for r in gb.CROSSOVER_ROUTINES}
    old = gb.get_cutoff()
    assert gb.set_options("cutoff=auto") == 0
    assert gb.get_cutoff() == gb.CUTOFF_AUTO
    yield
    gb.set_cutoff(old)
    for r, (f, i) in saved.items():
        gb.set_crossover(r, f, i)


def test_options_and_api():
    old = gb.get_cutoff()
    assert gb.set_options("cutoff=auto") == 0 and gb.get_cutoff() == -1
    assert gb.set_options("cutoff=64") == 0 and gb.get_cutoff() == 64
    assert gb.set_options("cutoff=12x") != 0 and gb.get_cutoff() == 64
    assert gb.set_options("cutoff=-3") != 0
    gb.set_cutoff(-5)
    assert gb.get_cutoff() == gb.CUTOFF_AUTO
    gb.set_cutoff(old)
    f0 = gb.get_crossover("dgemm")
    gb.set_crossover("dgemm", 123.0, 4.5)
    assert gb.get_crossover("dgemm") == (123.0, 4.5)
    gb.set_crossover("dgemm", *f0)
    for bad in [("zgemm", 1.0, 1.0), ("dgemm", -1.0, 1.0), ("dgemm", 1.0, float("nan"))]:
        with pytest.raises(ValueError):
            gb.set_crossover(*bad)
    with pytest.raises(ValueError):
        gb.get_crossover("dgemmx")


def test_table_rule_needs_both_work_and_intensity():
    f0 = gb.get_crossover("dgemm")
    try:
        # crossover at the square n = 192: 2n^3 flops over 32 n^2 bytes
        n0 = 192
        gb.set_crossover("dgemm", 2.0 * n0 ** 3, 2.0 * n0 / 32)

        def gpu(m, n, k):
            return gb.crossover_gpu("dgemm", 2.0 * m * n * k, 8.0 * (m * k + k * n + 2 * m * n))

        assert gpu(192, 192, 192) and gpu(4096, 4096, 4096)
        assert not gpu(191, 191, 191)
        assert not gpu(64, 64, 100000)   # enough flops, but 8 flops/byte: copy-bound
        assert not gpu(4096, 4096, 8)    # C dominates the bytes: copy-bound
        assert gpu(1024, 1024, 256)      # 64 flops/byte
        gb.set_crossover("dgemm", math.inf, math.inf)
        assert not gpu(1 << 14, 1 << 14, 1 << 14)
        gb.set_crossover("dgemm", 0.0, 0.0)
        assert gpu(1, 1, 1)
    finally:
        gb.set_crossover("dgemm", *f0)


def test_builtin_table_is_sane():
    for r in gb.CROSSOVER_ROUTINES:
        f, i = gb.get_crossover(r)
        assert f >= 0 and i >= 0
    # a 16^3 GEMM is launch-bound, a 4096^3 one is not
    assert not gb.crossover_gpu("dgemm", 2.0 * 16 ** 3, 32.0 * 16 ** 2)
    assert gb.crossover_gpu("dgemm", 2.0 * 4096 ** 3, 32.0 * 4096 ** 2)


def test_xover_file_round_trip(tmp_path, auto):
    good = tmp_path / "x.txt"
    good.write_text("# comment line\n"
                    "dgemm 1e6 3.5   # trailing comment\n"
                    "\n"
                    "dtrsm inf inf\n")
    assert gb.set_options(f"xover={good}") == 0
    assert gb.get_crossover("dgemm") == (1e6, 3.5)
    assert gb.get_crossover("dtrsm") == (math.inf, math.inf)
    before = {r: gb.get_crossover(r) for r in gb.CROSSOVER_ROUTINES}
    for text in ["dgemm 1e6\n", "zgemm 1 1\n", "dgemm -1 2\n", "dgemm 1 2x\n", "# empty\n"]:
        bad = tmp_path / "bad.txt"
        bad.write_text(text)
        assert gb.set_options(f"xover={bad}") != 0, text
        assert {r: gb.get_crossover(r) for r in gb.CROSSOVER_ROUTINES} == before
    assert gb.set_options(f"xover={tmp_path / 'missing.txt'}") != 0


def _all_host():
    for r in gb.CROSSOVER_ROUTINES:
        gb.set_crossover(r, math.inf, math.inf)


def test_auto_level3_on_host_bit_identical(auto):
    _all_host()
    m, n, k = 37, 29, 23
    A, B, C = uniform(1, (m, k)), uniform(2, (k, n)), uniform(3, (m, n))
    gb.reset_stats()
    calls = 0

    C1 = C.copy(order="F")
    gb.dgemm("N", "N", m, n, k, 0.5, A, m, B, k, 1.5, C1, m)
    assert np.array_equal(C1, sb.dgemm(0.5, A, B, 1.5, C.copy(order="F")))
    calls += 1

    As, Bs, Cs = (x.astype(np.float32, order="F") for x in (A, B, C))
    C2 = Cs.copy(order="F")
    gb.sgemm("N", "N", m, n, k, 0.5, As, m, Bs, k, 1.5, C2, m)
    assert np.array_equal(C2, sb.sgemm(0.5, As, Bs, 1.5, Cs.copy(order="F")))
    calls += 1

    S = uniform(4, (m, m))
    S1 = S.copy(order="F")
    gb.dsyrk("L", "N", m, k, 0.5, A, m, 1.5, S1, m)
    assert np.array_equal(S1, sb.dsyrk(0.5, A, 1.5, S.copy(order="F"), lower=1))
    calls += 1

    A2 = uniform(5, (m, k))
    S2 = S.copy(order="F")
    gb.dsyr2k("U", "N", m, k, 0.5, A, m, A2, m, 1.5, S2, m)
    assert np.array_equal(S2, sb.dsyr2k(0.5, A, A2, 1.5, S.copy(order="F")))
    calls += 1

    Sym = uniform(6, (m, m))
    Bm = uniform(7, (m, n))
    C3 = C.copy(order="F")
    gb.dsymm("L", "U", m, n, 0.5, Sym, m, Bm, m, 1.5, C3, m)
    assert np.array_equal(C3, sb.dsymm(0.5, Sym, Bm, 1.5, C.copy(order="F")))
    calls += 1

    T = np.asfortranarray(uniform(8, (m, m)) + np.diag(np.full(m, 4.0)))
    X = Bm.copy(order="F")
    gb.dtrsm("L", "L", "N", "N", m, n, 0.5, T, m, X, m)
    assert np.array_equal(X, sb.dtrsm(0.5, T, Bm.copy(order="F"), lower=1))
    calls += 1

    Y = Bm.copy(order="F")
    gb.dtrmm("L", "U", "T", "N", m, n, 0.5, T, m, Y, m)
    assert np.array_equal(Y, sb.dtrmm(0.5, T, Bm.copy(order="F"), trans_a=1))
    calls += 1

    st = gb.stats()
    assert st["fallthrough"] == calls and st["device_calls"] == 0


def test_auto_level12_on_host(auto):
    _all_host()
    n = 4000  # above the fixed cutoff: under the fixed rule these would go to the GPU
    x, y = uniform(9, (n,)), uniform(10, (n,))
    A = uniform(11, (300, 200))
    gb.reset_stats()
    assert gb.ddot(n, x, 1, y, 1) == sb.ddot(x, y)
    assert gb.dasum(n, x, 1) == sb.dasum(x)
    y1 = y.copy()
    gb.daxpy(n, 0.25, x, 1, y1, 1)
    assert np.array_equal(y1, sb.daxpy(x, y.copy(), a=0.25))
    xs = uniform(12, (200,))
    ys = uniform(13, (300,))
    y2 = ys.copy()
    gb.dgemv("N", 300, 200, 1.0, A, 300, xs, 1, 0.5, y2, 1)
    assert np.array_equal(y2, sb.dgemv(1.0, A, xs, 0.5, ys.copy()))
    assert gb.stats()["fallthrough"] == 4 and gb.stats()["device_calls"] == 0


_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, {root!r})
import libgpublas_amd as gb
for r in gb.CROSSOVER_ROUTINES:
    gb.set_crossover(r, float("inf"), float("inf"))
gb.set_cutoff(gb.CUTOFF_AUTO)
m, n, k = 40, 30, 20
A = np.ones((m, k), order="F"); B = np.ones((k, n), order="F"); C = np.zeros((m, n), order="F")
flops = 2.0 * m * n * k
nbytes = 8.0 * (m * k + k * n + 2 * m * n)
gb.set_crossover("dgemm", flops * {fscale}, flops / nbytes * {iscale})
gb.reset_stats()
gb.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
print("HOST", gb.stats()["fallthrough"], float(C[0, 0]), flush=True)
"""


@pytest.mark.parametrize("fscale,iscale,want", [(1.0, 1.0, "gpu"), (1.000001, 1.0, "host"),
                                                (1.0, 1.000001, "host")])
def test_auto_gemm_threshold_both_sides(fscale, iscale, want):
    """A DGEMM exactly at the table's thresholds goes to the GPU, just below
    either one to the host: the entry point's work model is the table's."""
    env = dict(os.environ)
    env["GPUBLAS_OPTIONS"] = ""
    env["HIP_VISIBLE_DEVICES"] = ""  # no GPU even where one exists
    code = _CHILD.format(root=ROOT, fscale=fscale, iscale=iscale)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=120)
    if want == "host":
        assert r.returncode == 0, r.stderr
        assert r.stdout.split() == ["HOST", "1", "20.0"]
    else:
        assert r.returncode != 0 and "no usable gfx950 GPU" in r.stderr, (r.returncode, r.stderr)
