"""The drop-in boundary on the CPU (no GPU needed): exported symbols, netlib
argument checking and xerbla, quick returns, the host fall-through cutoff,
CBLAS layout handling, options, tracked objects and the malloc interposer.

Nothing here reaches a kernel: every call either fails its argument check,
returns early, or is below the fall-through cutoff and runs on the host BLAS
(scipy's OpenBLAS in this image), exactly as the reference's gemm_perf_check
(blas_level3/gemm.cc:129-141) routes small problems."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.linalg.blas as sb

import libgpublas_amd as gb
from _oracle import p, uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"if", "for", "while", "return", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_stats_struct_mirror_matches_header():
    """The Python mirror of gpublas_stats_t has the header's fields in order
    (gpublas_get_stats fills the whole C struct: a shorter mirror is a heap
    overrun in the caller)."""
    src = open(os.path.join(ROOT, "include", "gpublas.h")).read()
    body = re.search(r"typedef struct \{(.*?)\} gpublas_stats_t;", src, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in re.findall(r"uint64_t\s+([^;]+);", body):
        fields += [f.strip() for f in decl.split(",")]
    assert [n for n, _ in gb.Stats._fields_] == fields


@pytest.mark.parametrize("header", ["gpublas.h", "gpublas_dev.h"])
def test_library_exports_every_declared_symbol(header):
    gb.library()
    syms = _exports(gb.LIB_PATH)
    missing = [n for n in _declared(header) if n not in syms]
    assert not missing, missing


def test_preload_library_adds_the_malloc_family():
    syms = _exports(gb.PRELOAD_PATH)
    for s in ("malloc", "calloc", "realloc", "free", "malloc_usable_size", "dgemm_", "cblas_dgemm"):
        assert s in syms
    assert "malloc" not in _exports(gb.LIB_PATH)


def _ba(*vals):
    return [ctypes.byref(v) for v in vals]


GEMM_CASES = [  # (transa, transb, m, n, k, lda, ldb, ldc) -> INFO
    (("X", "N", 2, 2, 2, 2, 2, 2), 1), (("N", "Y", 2, 2, 2, 2, 2, 2), 2),
    (("N", "N", -1, 2, 2, 2, 2, 2), 3), (("N", "N", 2, -1, 2, 2, 2, 2), 4),
    (("N", "N", 2, 2, -1, 2, 2, 2), 5), (("N", "N", 3, 2, 2, 2, 2, 3), 8),
    (("N", "N", 2, 2, 3, 2, 2, 2), 10), (("N", "N", 3, 2, 2, 3, 2, 2), 13)]


@pytest.mark.parametrize("args,info", GEMM_CASES)
def test_dgemm_xerbla(args, info):
    ta, tb, m, n, k, lda, ldb, ldc = args
    dummy = np.zeros(64)
    gb.dgemm(ta, tb, m, n, k, 1.0, dummy, lda, dummy, ldb, 0.0, dummy, ldc)
    assert gb.last_xerbla() == ("DGEMM ", info)


def test_level3_xerbla_positions(ora):
    d = np.zeros(64)
    gb.dsyrk("L", "N", 3, 2, 1.0, d, 2, 0.0, d, 3)
    assert gb.last_xerbla() == ("DSYRK ", ora.ora_syrk_info(b"L", b"N", 3, 2, 2, 3))
    gb.dtrsm("L", "L", "N", "X", 2, 2, 1.0, d, 2, d, 2)
    assert gb.last_xerbla() == ("DTRSM ", 4)
    gb.dtrsm("R", "L", "N", "N", 2, 3, 1.0, d, 2, d, 2)
    assert gb.last_xerbla() == ("DTRSM ", 9)
    gb.dsymm("L", "L", 3, 2, 1.0, d, 3, d, 3, 0.0, d, 2)
    assert gb.last_xerbla() == ("DSYMM ", 12)
    gb.ssyrk("Q", "N", 3, 2, 1.0, d, 3, 0.0, d, 3)
    assert gb.last_xerbla() == ("SSYRK ", 1)
    gb.dsyr2k("L", "T", 3, 4, 1.0, d, 4, d, 3, 0.0, d, 3)
    assert gb.last_xerbla() == ("DSYR2K", ora.ora_syr2k_info(b"L", b"T", 3, 4, 4, 3, 3))
    gb.dsyr2k("L", "N", 3, 2, 1.0, d, 3, d, 3, 0.0, d, 2)
    assert gb.last_xerbla() == ("DSYR2K", 12)
    gb.ssyr2k("L", "Q", 3, 2, 1.0, d, 3, d, 3, 0.0, d, 3)
    assert gb.last_xerbla() == ("SSYR2K", 2)
    gb.dtrmm("L", "L", "N", "X", 2, 2, 1.0, d, 2, d, 2)
    assert gb.last_xerbla() == ("DTRMM ", 4)
    gb.dtrmm("R", "L", "N", "N", 2, 3, 1.0, d, 2, d, 2)
    assert gb.last_xerbla() == ("DTRMM ", ora.ora_trmm_info(b"R", b"L", b"N", b"N", 2, 3, 2, 2))
    gb.strmm("L", "U", "T", "U", 3, 2, 1.0, d, 3, d, 2)
    assert gb.last_xerbla() == ("STRMM ", 11)


def test_level12_xerbla_positions():
    d = np.zeros(64)
    gb.dgemv("N", 3, 2, 1.0, d, 2, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DGEMV ", 6)
    gb.dgemv("N", 2, 2, 1.0, d, 2, d, 0, 0.0, d, 1)
    assert gb.last_xerbla() == ("DGEMV ", 8)
    gb.dger(2, 2, 1.0, d, 1, d, 0, d, 2)
    assert gb.last_xerbla() == ("DGER  ", 7)
    gb.dtrsv("L", "N", "N", 3, d, 2, d, 1)
    assert gb.last_xerbla() == ("DTRSV ", 6)


def test_cblas_error_positions():
    lib = gb.library()
    d = np.zeros(64)
    lib.cblas_dgemm(999, 111, 111, 2, 2, 2, 1.0, p(d), 2, p(d), 2, 0.0, p(d), 2)
    assert gb.last_xerbla() == ("DGEMM ", 1)
    lib.cblas_dgemm(102, 111, 7, 2, 2, 2, 1.0, p(d), 2, p(d), 2, 0.0, p(d), 2)
    assert gb.last_xerbla() == ("DGEMM ", 3)
    # column-major m < 0 is CBLAS parameter 4; row-major it is the Fortran n -> CBLAS 4 too
    lib.cblas_dgemm(102, 111, 111, -1, 2, 2, 1.0, p(d), 2, p(d), 2, 0.0, p(d), 2)
    assert gb.last_xerbla() == ("DGEMM ", 4)
    lib.cblas_dgemm(101, 111, 111, -1, 2, 2, 1.0, p(d), 2, p(d), 2, 0.0, p(d), 2)
    assert gb.last_xerbla() == ("DGEMM ", 4)
    lib.cblas_dgemm(101, 111, 111, 2, 3, 2, 1.0, p(d), 2, p(d), 3, 0.0, p(d), 2)
    assert gb.last_xerbla() == ("DGEMM ", 14)  # row-major: ldc < n


def test_quick_returns_touch_nothing():
    gb.reset_stats()
    # null operands: quick returns must not dereference them
    gb.dgemm("N", "N", 0, 5, 5, 1.0, None, 1, None, 5, 0.0, None, 1)
    gb.dgemm("N", "N", 5, 5, 0, 1.0, None, 5, None, 1, 1.0, None, 5)
    gb.dsyrk("L", "N", 0, 5, 1.0, None, 1, 0.0, None, 1)
    gb.dtrsm("L", "L", "N", "N", 0, 5, 1.0, None, 1, None, 1)
    gb.dsymm("L", "L", 4, 0, 1.0, None, 4, None, 4, 0.0, None, 4)
    gb.dgemv("N", 0, 4, 1.0, None, 1, None, 1, 0.0, None, 1)
    gb.daxpy(0, 1.0, None, 1, None, 1)
    assert gb.ddot(0, None, 1, None, 1) == 0.0
    assert gb.dnrm2(-1, None, 1) == 0.0
    assert gb.idamax(0, None, 1) == 0
    s = gb.stats()
    assert s["device_calls"] == 0 and s["fallthrough"] == 0


def test_alpha_zero_host_path_is_netlib():
    C = np.full((5, 4), np.nan, order="F")
    A, B = uniform(1, (5, 3)), uniform(2, (3, 4))
    gb.dgemm("N", "N", 5, 4, 3, 0.0, A, 5, B, 3, 0.0, C, 5)
    assert np.all(C == 0.0)  # the reference's C *= beta would keep NaN (gemm.cc:121-126)
    C = uniform(3, (5, 4))
    want = 2.0 * C
    gb.dgemm("N", "N", 5, 4, 0, 1.0, A, 5, B, 1, 2.0, C, 5)
    assert np.array_equal(C, want)
    S0 = uniform(4, (4, 4))
    S = S0.copy(order="F")
    gb.dsyrk("U", "N", 4, 3, 0.0, A, 4, 3.0, S, 4)
    tri = np.triu(np.ones((4, 4), bool))
    assert np.array_equal(S[tri], 3.0 * S0[tri]) and np.array_equal(S[~tri], S0[~tri])


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_small_gemm_falls_through_to_host_blas(ta, tb):
    gb.set_cutoff(512)
    m, n, k = 9, 7, 5
    A = uniform(5, (m, k) if ta == "N" else (k, m))
    B = uniform(6, (k, n) if tb == "N" else (n, k))
    C = uniform(7, (m, n))
    H = sb.dgemm(0.5, A, B, 1.5, C.copy(order="F"), trans_a=ta == "T", trans_b=tb == "T")
    gb.reset_stats()
    gb.dgemm(ta, tb, m, n, k, 0.5, A, A.shape[0], B, B.shape[0], 1.5, C, m)
    assert gb.stats()["fallthrough"] == 1
    assert np.array_equal(C, H)  # the same host routine: bit-identical


def test_small_level12_fall_through():
    gb.set_cutoff(512)
    x, y = uniform(8, (100,)), uniform(9, (100,))
    gb.reset_stats()
    assert gb.ddot(100, x, 1, y, 1) == sb.ddot(x, y)
    assert gb.dnrm2(100, x, 1) == sb.dnrm2(x)
    assert gb.idamax(100, x, 1) == sb.idamax(x) + 1
    lib = gb.library()
    assert lib.cblas_idamax(100, p(x), 1) == sb.idamax(x)  # CBLAS is 0-based
    assert gb.stats()["fallthrough"] == 4


def test_cblas_row_major_gemm_via_fall_through():
    gb.set_cutoff(512)
    lib = gb.library()
    m, n, k = 6, 5, 4
    A = np.ascontiguousarray(uniform(10, (m, k)))  # row-major
    B = np.ascontiguousarray(uniform(11, (k, n)))
    C = np.zeros((m, n))
    lib.cblas_dgemm(101, 111, 111, m, n, k, 1.0, p(A), k, p(B), n, 0.0, p(C), n)
    np.testing.assert_allclose(C, A @ B, rtol=1e-14, atol=1e-15)
    x = uniform(12, (k,))
    y = np.zeros(m)
    lib.cblas_dgemv(101, 111, m, k, 1.0, p(A), k, p(x), 1, 0.0, p(y), 1)
    np.testing.assert_allclose(y, A @ x, rtol=1e-14, atol=1e-15)


def test_options_parsing():
    assert gb.set_options("cutoff=128;min_track=64k") == 0
    assert gb.get_cutoff() == 128
    assert gb.set_options("bogus_option") != 0
    assert gb.set_options("heuristic=objtrack:/nonexistent/file") != 0
    gb.set_options("cutoff=512;heuristic=false")


def test_tracked_object_host_side():
    a = gb.tracked_empty((1000,))
    a[:] = np.arange(1000.0)
    assert gb.is_tracked(a) and gb.is_tracked(a[500:])
    assert float(a.sum()) == 499500.0
    assert not gb.is_tracked(np.zeros(10))
    s = gb.stats()
    assert s["tracked_objects"] >= 1


def _run_preloaded(objtrack_text, extra=""):
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "ot.txt")
        open(f, "w").write(objtrack_text)
        env = dict(os.environ)
        env["LD_PRELOAD"] = gb.PRELOAD_PATH
        env["GPUBLAS_OPTIONS"] = f"heuristic=objtrack:{f};min_track=0{extra}"
        out = subprocess.run([gb.CHAIN_BENCH, "--n", "64", "--iters", "0", "--warmup", "0"],
                             env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out


def test_objtrack_file_selects_allocations_by_ordinal():
    """The reference's "%1[HD] #nth" format (lib/oracle.c:41): the driver's
    first four mallocs are its operands A, x, y, z."""
    r = _run_preloaded("D #0 A\nD #1 x\nD #2 y\nD #3 z\n")
    assert '"tracked": 4' in r.stdout
    r = _run_preloaded("D #0\nH #1\nD #2\n# comment line\n")
    assert '"tracked": 2' in r.stdout
    r = _run_preloaded("")
    assert '"tracked": 0' in r.stdout


def test_objtrack_trace_names_call_sites():
    """The trace option logs every tracked allocation with its call site
    (module+offset), the key a call-site objtrack line uses."""
    with tempfile.TemporaryDirectory() as td:
        tr = os.path.join(td, "trace.txt")
        _run_preloaded("D #0\nD #1\n", extra=f";trace={tr}")
        lines = open(tr).read().splitlines()
    t = [ln for ln in lines if ln.startswith("T ")]
    assert len(t) == 2 and all("site=[chain_bench+0x" in ln for ln in t)
    site = re.search(r"site=\[(chain_bench\+0x[0-9a-f]+)\]", t[0]).group(1)
    r = _run_preloaded(f"D @{site}\n")
    assert '"tracked": 1' in r.stdout


def test_device_required_message_without_gpu():
    if gb.device_ready():
        pytest.skip("a GPU is present")
    code = ("import numpy as np, libgpublas_amd as gb; gb.set_cutoff(0); "
            "a=np.ones((4,4),order='F'); gb.dgemm('N','N',4,4,4,1.0,a,4,a,4,0.0,a,4)")
    out = subprocess.run(["python3", "-c", code], cwd=ROOT, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode != 0
    assert "no usable gfx950 GPU" in out.stderr


def test_netlib_dblat3_error_exits_and_fallthrough_gemm():
    """The reference's own Level-3 tester (tests/netlib/dblat3.f, built from
    the reference tree into oracle/_ref) run against libgpublas: error exits
    of DGEMM (DCHKE) and DGEMM results through the host fall-through."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dblat3")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/dblat3 not built")
    inp = open(os.path.join(ROOT, "tests", "golden", "input.dblat3.gemm")).read()
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ)
        env["GPUBLAS_OPTIONS"] = "cutoff=512"
        env["GPUBLAS_HOST_BLAS"] = gb.__dict__.get("SCIPY_OPENBLAS", "")
        out = subprocess.run([exe], input=inp, cwd=td, env=env, capture_output=True, text=True,
                             timeout=300)
        summ = open(os.path.join(td, "DBLAT3.SUMM")).read()
    assert out.returncode == 0, out.stderr[-2000:]
    assert "DGEMM  PASSED THE TESTS OF ERROR-EXITS" in summ
    assert "DGEMM  PASSED THE COMPUTATIONAL TESTS" in summ
    assert "FAIL" not in summ and "**" not in summ


def test_symv_syr_xerbla_and_fall_through(ora):
    """DSYMV / DSYR: netlib INFO codes, the CBLAS positions they map to, and
    the small-problem fall-through (n < cutoff) bit-identical to the host
    BLAS -- row-major CBLAS calls take the other triangle."""
    lib = gb.library()
    d = np.zeros(64)
    gb.dsymv("X", 2, 1.0, d, 2, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DSYMV ", ora.ora_symv_info(b"X", 2, 2, 1, 1))
    gb.dsymv("L", 3, 1.0, d, 2, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DSYMV ", 5)
    gb.dsymv("U", 2, 1.0, d, 2, d, 1, 0.0, d, 0)
    assert gb.last_xerbla() == ("DSYMV ", 10)
    gb.dsyr("L", 2, 1.0, d, 0, d, 2)
    assert gb.last_xerbla() == ("DSYR  ", ora.ora_syr_info(b"L", 2, 0, 2))
    gb.dsyr("U", 3, 1.0, d, 1, d, 2)
    assert gb.last_xerbla() == ("DSYR  ", 7)
    lib.cblas_dsymv(102, 121, 3, 1.0, p(d), 2, p(d), 1, 0.0, p(d), 1)
    assert gb.last_xerbla() == ("DSYMV ", 6)  # lda is CBLAS parameter 6
    lib.cblas_dsyr(102, 999, 2, 1.0, p(d), 1, p(d), 2)
    assert gb.last_xerbla() == ("DSYR  ", 2)

    gb.set_cutoff(512)
    n = 50
    A = uniform(61, (n, n))
    x, y0 = uniform(62, (n,)), uniform(63, (n,))
    gb.reset_stats()
    y = y0.copy()
    gb.dsymv("U", n, 0.5, A, n, x, 1, 2.0, y, 1)
    assert np.array_equal(y, sb.dsymv(0.5, A, x, beta=2.0, y=y0.copy(), lower=0))
    R = A.copy(order="F")
    gb.dsyr("L", n, -0.75, x, 1, R, n)
    assert np.array_equal(R, sb.dsyr(-0.75, x, lower=1, a=A.copy(order="F")))
    assert gb.stats()["fallthrough"] == 2
    # row-major: the memory of A read as row-major is A.T column-major; upper
    # row-major = lower column-major of the same bytes
    y = np.zeros(n)
    lib.cblas_dsymv(101, 121, n, 1.0, p(A), n, p(x), 1, 0.0, p(y), 1)  # 121 = CblasUpper
    S = np.triu(A.T) + np.triu(A.T, 1).T  # the row-major upper triangle, symmetrised
    np.testing.assert_allclose(y, S @ x, rtol=1e-13, atol=1e-14)


def test_objtrack_capture_without_uses_writes_an_empty_file():
    """objtrack_out writes the object-tracking file at exit; a run that makes
    no BLAS call leaves it with its header and no rule."""
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "cap.txt")
        env = dict(os.environ)
        env["LD_PRELOAD"] = gb.PRELOAD_PATH
        env["GPUBLAS_OPTIONS"] = f"heuristic=true;min_track=0;objtrack_out={out}"
        r = subprocess.run([gb.CHAIN_BENCH, "--n", "64", "--iters", "0", "--warmup", "0"],
                           env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        text = open(out).read()
    assert text.startswith("# object-tracking file written by libgpublas")
    assert "chain_bench" in text.splitlines()[0]
    assert [ln for ln in text.splitlines() if ln and not ln.startswith("#")] == []


def test_rank2_and_packed_updates_xerbla_and_fall_through(ora):
    """DSYR2 / DSPR / DSPR2: netlib INFO codes and the host fall-through for
    small problems, bit-identical to the host BLAS (scipy's OpenBLAS)."""
    d = np.zeros(64)
    gb.dsyr2("L", 3, 1.0, d, 1, d, 1, d, 2)
    assert gb.last_xerbla() == ("DSYR2 ", ora.ora_syr2_info(b"L", 3, 1, 1, 2))
    gb.dsyr2("U", 2, 1.0, d, 1, d, 0, d, 2)
    assert gb.last_xerbla() == ("DSYR2 ", 7)
    gb.dspr("X", 2, 1.0, d, 1, d)
    assert gb.last_xerbla() == ("DSPR  ", 1)
    gb.dspr2("L", 2, 1.0, d, 0, d, 1, d)
    assert gb.last_xerbla() == ("DSPR2 ", 5)
    gb.set_cutoff(512)
    n = 40
    A = uniform(71, (n, n))
    x, y = uniform(72, (n,)), uniform(73, (n,))
    AP = np.concatenate([A[j:, j] for j in range(n)])
    gb.reset_stats()
    R = A.copy(order="F")
    gb.dsyr2("U", n, 0.5, x, 1, y, 1, R, n)
    assert np.array_equal(R, sb.dsyr2(0.5, x, y, lower=0, a=A.copy(order="F")))
    P = AP.copy()
    gb.dspr("L", n, -1.5, x, 1, P)
    assert np.array_equal(P, sb.dspr(n, -1.5, x, AP.copy(), lower=1))
    P = AP.copy()
    gb.dspr2("L", n, 0.25, x, 1, y, 1, P)
    assert np.array_equal(P, sb.dspr2(n, 0.25, x, y, AP.copy(), lower=1))
    assert gb.stats()["fallthrough"] == 3


def test_triangular_and_packed_mv_xerbla_and_fall_through(ora):
    """DTRMV / DTPMV / DTPSV / DSPMV: netlib INFO codes (and their CBLAS
    positions), and the small-problem fall-through bit-identical to the host
    BLAS (scipy's OpenBLAS)."""
    lib = gb.library()
    d = np.zeros(64)
    gb.dtrmv("L", "X", "N", 2, d, 2, d, 1)
    assert gb.last_xerbla() == ("DTRMV ", ora.ora_trmv_info(b"L", b"X", b"N", 2, 2, 1))
    gb.dtrmv("L", "N", "N", 3, d, 2, d, 1)
    assert gb.last_xerbla() == ("DTRMV ", 6)
    gb.dtrmv("U", "T", "U", 2, d, 2, d, 0)
    assert gb.last_xerbla() == ("DTRMV ", 8)
    gb.dtpmv("U", "N", "Q", 2, d, d, 1)
    assert gb.last_xerbla() == ("DTPMV ", ora.ora_tpmv_info(b"U", b"N", b"Q", 2, 1))
    gb.dtpsv("L", "N", "N", 2, d, d, 0)
    assert gb.last_xerbla() == ("DTPSV ", 7)
    gb.dspmv("L", -1, 1.0, d, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DSPMV ", ora.ora_spmv_info(b"L", -1, 1, 1))
    gb.dspmv("U", 2, 1.0, d, d, 1, 0.0, d, 0)
    assert gb.last_xerbla() == ("DSPMV ", 9)
    lib.cblas_dtrmv(102, 122, 111, 131, 3, p(d), 2, p(d), 1)
    assert gb.last_xerbla() == ("DTRMV ", 7)  # lda is CBLAS parameter 7
    lib.cblas_dtpsv(102, 122, 111, 131, 2, p(d), p(d), 0)
    assert gb.last_xerbla() == ("DTPSV ", 8)
    lib.cblas_dspmv(101, 122, 2, ctypes.c_double(1.0), p(d), p(d), 0, ctypes.c_double(0.0), p(d), 1)
    assert gb.last_xerbla() == ("DSPMV ", 7)

    gb.set_cutoff(512)
    n = 40
    A = uniform(81, (n, n)) + np.diag(np.full(n, 4.0))
    A = np.asfortranarray(A)
    x0, y0 = uniform(82, (n,)), uniform(83, (n,))
    APl = np.concatenate([A[j:, j] for j in range(n)])
    APu = np.concatenate([A[: j + 1, j] for j in range(n)])
    gb.reset_stats()
    x = x0.copy()
    gb.dtrmv("U", "T", "N", n, A, n, x, 1)
    assert np.array_equal(x, sb.dtrmv(A, x0.copy(), lower=0, trans=1))
    x = x0.copy()
    gb.dtpmv("L", "N", "U", n, APl, x, 1)
    assert np.array_equal(x, sb.dtpmv(n, APl, x0.copy(), lower=1, diag=1))
    x = x0.copy()
    gb.dtpsv("U", "N", "N", n, APu, x, 1)
    assert np.array_equal(x, sb.dtpsv(n, APu, x0.copy(), lower=0))
    y = y0.copy()
    gb.dspmv("L", n, 0.5, APl, x0, 1, 2.0, y, 1)
    assert np.array_equal(y, sb.dspmv(n, 0.5, APl, x0, beta=2.0, y=y0.copy(), lower=1))
    assert gb.stats()["fallthrough"] == 4


def _band_of(D, kl, ku, lda):
    """netlib column-major band storage of D (entries outside the band: 0)."""
    m, n = D.shape
    AB = np.zeros((lda, n), order="F")
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            AB[ku + i - j, j] = D[i, j]
    return AB


def test_band_routines_xerbla_and_fall_through(ora):
    """DGBMV / DSBMV / DTBMV / DTBSV: netlib INFO codes and their CBLAS
    positions (row-major swaps m/n and kl/ku), the host fall-through
    bit-identical to scipy's OpenBLAS, and the row-major CBLAS mappings
    checked against dense closed forms."""
    lib = gb.library()
    d = np.zeros(64)
    gb.dgbmv("N", 2, 2, 1, 1, 1.0, d, 2, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DGBMV ", ora.ora_gbmv_info(b"N", 2, 2, 1, 1, 2, 1, 1))
    gb.dgbmv("N", 2, 2, 0, -1, 1.0, d, 2, d, 1, 0.0, d, 1)
    assert gb.last_xerbla() == ("DGBMV ", 5)
    gb.dsbmv("U", 2, 1, 1.0, d, 2, d, 1, 0.0, d, 0)
    assert gb.last_xerbla() == ("DSBMV ", ora.ora_sbmv_info(b"U", 2, 1, 2, 1, 0))
    gb.dtbmv("L", "N", "N", 2, 2, d, 2, d, 1)
    assert gb.last_xerbla() == ("DTBMV ", ora.ora_tbmv_info(b"L", b"N", b"N", 2, 2, 2, 1))
    gb.dtbsv("L", "N", "N", 2, 1, d, 2, d, 0)
    assert gb.last_xerbla() == ("DTBSV ", 9)
    dd = ctypes.c_double
    lib.cblas_dgbmv(101, 111, 2, 3, 1, 0, dd(1.0), p(d), 1, p(d), 1, dd(0.0), p(d), 1)
    assert gb.last_xerbla() == ("DGBMV ", 9)  # lda < kl + ku + 1: CBLAS parameter 9
    lib.cblas_dgbmv(101, 111, 2, 3, -1, 0, dd(1.0), p(d), 2, p(d), 1, dd(0.0), p(d), 1)
    assert gb.last_xerbla() == ("DGBMV ", 5)  # row-major: kl stays CBLAS parameter 5
    lib.cblas_dgbmv(101, 111, -1, 3, 0, 0, dd(1.0), p(d), 2, p(d), 1, dd(0.0), p(d), 1)
    assert gb.last_xerbla() == ("DGBMV ", 3)
    lib.cblas_dsbmv(102, 121, 2, 3, dd(1.0), p(d), 3, p(d), 1, dd(0.0), p(d), 1)
    assert gb.last_xerbla() == ("DSBMV ", 7)
    lib.cblas_dtbsv(102, 121, 111, 131, 2, 0, p(d), 1, p(d), 0)
    assert gb.last_xerbla() == ("DTBSV ", 10)

    gb.set_cutoff(512)
    rng = np.random.default_rng(85)
    m, n, kl, ku = 30, 25, 3, 2
    D = rng.standard_normal((m, n))
    D[np.subtract.outer(np.arange(m), np.arange(n)) > kl] = 0
    D[np.subtract.outer(np.arange(n), np.arange(m)).T > ku] = 0
    AB = _band_of(D, kl, ku, kl + ku + 1)
    x, y0 = rng.standard_normal(n), rng.standard_normal(m)
    gb.reset_stats()
    y = y0.copy()
    gb.dgbmv("N", m, n, kl, ku, 0.5, AB, kl + ku + 1, x, 1, 2.0, y, 1)
    assert np.array_equal(y, sb.dgbmv(m, n, kl, ku, 0.5, AB, x, beta=2.0, y=y0.copy()))
    # row-major: the band array read row by row holds D' band-stored column-major,
    # i.e. the row-major band storage of D' (n x m, kl' = ku, ku' = kl)
    y = np.zeros(n)
    lib.cblas_dgbmv(101, 111, n, m, ku, kl, dd(1.0), p(AB), kl + ku + 1, p(y0), 1, dd(0.0),
                    p(y), 1)
    np.testing.assert_allclose(y, D.T @ y0, rtol=1e-13, atol=1e-13)
    k = 3
    S = rng.standard_normal((n, n))
    S = np.triu(S) - np.triu(S, k + 1)
    AS = _band_of(S, 0, k, k + 1)  # upper band
    y = y0[:n].copy()
    gb.dsbmv("U", n, k, 0.75, AS, k + 1, x, 1, -1.0, y, 1)
    assert np.array_equal(y, sb.dsbmv(k, 0.75, AS, x, beta=-1.0, y=y0[:n].copy()))
    y = np.zeros(n)
    lib.cblas_dsbmv(101, 122, n, k, dd(1.0), p(AS), k + 1, p(x), 1, dd(0.0), p(y), 1)
    Sym = S + S.T - np.diag(np.diag(S))
    np.testing.assert_allclose(y, Sym @ x, rtol=1e-13, atol=1e-13)  # row-major lower = same bytes
    T = S + np.diag(np.full(n, 3.0))
    AT = _band_of(T, 0, k, k + 1)
    xt = x.copy()
    gb.dtbmv("U", "T", "N", n, k, AT, k + 1, xt, 1)
    assert np.array_equal(xt, sb.dtbmv(k, AT, x.copy(), trans=1))
    xt = x.copy()
    gb.dtbsv("U", "N", "N", n, k, AT, k + 1, xt, 1)
    assert np.array_equal(xt, sb.dtbsv(k, AT, x.copy()))
    assert gb.stats()["fallthrough"] == 6  # the two row-major CBLAS calls included
    xt = x.copy()  # row-major lower band triangle = T' column-major upper: solve T' z = x
    lib.cblas_dtbsv(101, 122, 111, 131, n, k, p(AT), k + 1, p(xt), 1)
    np.testing.assert_allclose(T.T @ xt, x, rtol=1e-12, atol=1e-12)


def test_idamin_drotm_fall_through(ora):
    """IDAMIN / DROTM below the cutoff run on the host BLAS: bit-identical to
    scipy's OpenBLAS (drotm) and to the oracle on finite data (idamin);
    cblas_idamin is 0-based."""
    lib = gb.library()
    gb.set_cutoff(512)
    rng = np.random.default_rng(86)
    x = rng.standard_normal(200)
    gb.reset_stats()
    assert gb.idamin(200, x, 1) == ora.ora_idamin(200, p(x), 1) == np.argmin(np.abs(x)) + 1
    assert lib.cblas_idamin(200, p(x), 1) == np.argmin(np.abs(x))
    assert gb.idamin(0, x, 1) == 0 and gb.idamin(5, x, -1) == 0
    xs, ys = rng.standard_normal(100), rng.standard_normal(100)
    for flag in (-1.0, 0.0, 1.0):
        param = np.array([flag, 0.5, -0.25, 1.5, 0.75])
        X, Y = xs.copy(), ys.copy()
        gb.drotm(100, X, 1, Y, 1, param)
        hx, hy = sb.drotm(xs.copy(), ys.copy(), param)
        assert np.array_equal(X, hx) and np.array_equal(Y, hy)
    assert gb.stats()["fallthrough"] == 5  # two idamin calls + three drotm calls
    X = xs.copy()
    gb.drotm(100, X, 1, Y, 1, np.array([-2.0, 9, 9, 9, 9]))  # identity flag: nothing happens
    assert np.array_equal(X, xs)


# ----------------------------------------------------------------------------- errors
def test_cblas_error_positions_fp32_level3():
    """The FP32 Level-3 twins and a row-major DGEMV report CBLAS positions
    (checks run before any device work, so no GPU is needed)."""
    from test_gpu_cblas import (COL, LEFT, NONUNIT, NOTRANS, ROW, UPPER, _fn, _ptr)
    lib = gb.library()
    d = np.zeros(16, np.float32)
    _fn(gb, "syrk", "s")(999, UPPER, NOTRANS, 2, 2, 1.0, _ptr(d), 2, 0.0, _ptr(d), 2)
    assert gb.last_xerbla() == ("SSYRK ", 1)
    _fn(gb, "syrk", "s")(COL, 7, NOTRANS, 2, 2, 1.0, _ptr(d), 2, 0.0, _ptr(d), 2)
    assert gb.last_xerbla() == ("SSYRK ", 2)
    _fn(gb, "syrk", "s")(ROW, UPPER, NOTRANS, 2, -1, 1.0, _ptr(d), 2, 0.0, _ptr(d), 2)
    assert gb.last_xerbla() == ("SSYRK ", 5)
    _fn(gb, "trsm", "s")(ROW, LEFT, UPPER, NOTRANS, 5, 2, 2, 1.0, _ptr(d), 2, _ptr(d), 2)
    assert gb.last_xerbla() == ("STRSM ", 5)
    _fn(gb, "trsm", "s")(ROW, LEFT, UPPER, NOTRANS, NONUNIT, 3, 2, 1.0, _ptr(d), 2, _ptr(d), 2)
    assert gb.last_xerbla() == ("STRSM ", 10)  # row-major: lda < m for the left side
    _fn(gb, "trmm", "s")(COL, 9, UPPER, NOTRANS, NONUNIT, 2, 2, 1.0, _ptr(d), 2, _ptr(d), 2)
    assert gb.last_xerbla() == ("STRMM ", 2)
    _fn(gb, "symm", "s")(ROW, LEFT, UPPER, 2, 3, 1.0, _ptr(d), 2, _ptr(d), 3, 0.0, _ptr(d), 2)
    assert gb.last_xerbla() == ("SSYMM ", 13)  # row-major: ldc < n
    _fn(gb, "syr2k", "s")(COL, UPPER, NOTRANS, 3, 2, 1.0, _ptr(d), 3, _ptr(d), 2, 0.0, _ptr(d), 3)
    assert gb.last_xerbla() == ("SSYR2K", 10)
    dd = np.zeros(16)
    lib.cblas_dgemv(ROW, NOTRANS, 3, 4, 1.0, _ptr(dd), 3, _ptr(dd), 1, 0.0, _ptr(dd), 1)
    assert gb.last_xerbla() == ("DGEMV ", 7)  # row-major: lda < n


def test_packed_triangle_length_beyond_int_range(tmp_path):
    """Packed routines acquire n(n+1)/2 elements; from n = 65536 that count
    passes INT_MAX, so the host layer computes it in size_t (a 32-bit cast
    would hand the device a negative length and a missing operand)."""
    src = tmp_path / "packed_len.c"
    src.write_text('#include "internal.h"\n#include <stdio.h>\n'
                   'int main(void) { printf("%zu %zu %zu %zu %zu\\n", gb_packed_len(-1), '
                   'gb_packed_len(1), gb_packed_len(65535), gb_packed_len(65536), '
                   'gb_packed_len(2000000000)); return 0; }\n')
    exe = tmp_path / "packed_len"
    inc = [os.path.join(ROOT, "include"), os.path.join(ROOT, "libgpublas_amd", "csrc", "host")]
    subprocess.run(["gcc", "-std=gnu11", "-I", inc[0], "-I", inc[1], "-o", str(exe), str(src)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    n = 2000000000
    assert got == [0, 1, 65535 * 65536 // 2, 65536 * 65537 // 2, n * (n + 1) // 2]
    assert got[3] > 2**31 - 1
