"""GPU parity of the Level-3 routines against the CPU oracle.

Every call goes through the Fortran-ABI entry points of libgpublas.so with the
host fall-through disabled (cutoff 0), so the gfx950 kernels compute every
result; `device_calls` is checked to prove it.

Bars (stated per test):
  * netlib DMMCH gauge ratio  max|C-Cref|/(eps*G) < 16   (tests/netlib/dblat3.f:2538-2546)
    with G = |alpha| sum|op(A)||op(B)| + |beta||C0|, the reference's own threshold;
  * bit-exact results on integer-valued data whose sums stay below 2^24 (FP32)
    or 2^53 (FP64), where every summation order gives the same answer;
  * untouched regions (rows beyond m, the unreferenced triangle) bit-identical
    to their input (dblat3.f:2611-2672 LDERES).
"""
import numpy as np
import pytest

import _oracle
from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps
ROGUE = -1e10


def _padded(a, ld, fill=ROGUE):
    """Embed a (r x c) in an ld x c Fortran array whose extra rows hold `fill`."""
    r, cols = a.shape
    big = np.full((ld, cols), fill, dtype=a.dtype, order="F")
    big[:r, :] = a
    return big


def gauge_ratio(ora, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0, Cgot, ldc, eps):
    return ora.ora_gemm_gauge_ratio(c(ta), c(tb), m, n, k, alpha, p(A), lda, p(B), ldb, beta,
                                    p(C0), p(Cgot), ldc, eps)


def _device_calls(gb):
    return gb.stats()["device_calls"]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (64, 64, 16), (65, 63, 17),
                                   (129, 200, 100), (300, 257, 513)])
def test_dgemm_shapes(gpu, ora, ta, tb, m, n, k):
    gb = gpu
    rows_a, cols_a = (m, k) if ta == "N" else (k, m)
    rows_b, cols_b = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = rows_a + 3, rows_b + 1, m + 2
    A = _padded(uniform(11, (rows_a, cols_a)), lda)
    B = _padded(uniform(12, (rows_b, cols_b)), ldb)
    C0 = _padded(uniform(13, (m, n)), ldc)
    alpha, beta = 0.7, -1.3
    C = C0.copy(order="F")
    before = _device_calls(gb)
    gb.dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    assert _device_calls(gb) == before + 1
    r = gauge_ratio(ora, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0, C, ldc, EPS64)
    assert r < 16.0, r
    assert np.array_equal(C[m:, :], C0[m:, :])  # padding rows untouched


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_dgemm_integer_exact(gpu, ora, ta, tb):
    rng = np.random.default_rng(5)
    m, n, k = 150, 130, 170
    A = np.asfortranarray(rng.integers(-8, 9, size=(m, k) if ta == "N" else (k, m)).astype(float))
    B = np.asfortranarray(rng.integers(-8, 9, size=(k, n) if tb == "N" else (n, k)).astype(float))
    C = np.asfortranarray(rng.integers(-8, 9, size=(m, n)).astype(float))
    R = C.copy(order="F")
    gpu.dgemm(ta, tb, m, n, k, 2.0, A, A.shape[0], B, B.shape[0], 3.0, C, m)
    ora.ora_dgemm(c(ta), c(tb), m, n, k, 2.0, p(A), A.shape[0], p(B), B.shape[0], 3.0, p(R), m)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(512, 512, 2048), (1024, 1024, 1024), (576, 448, 1024)])
def test_dgemm_splitk_exact(gpu, ora, ta, tb, m, n, k):
    """Few 64 x 64 tiles and a long k take the split-K path (8, 2 and 4 k
    chunks here; partial sums folded in chunk order); ldc padded, beta != 0.
    Integer data: bit-exact, rows beyond m untouched."""
    rng = np.random.default_rng(41)
    A = np.asfortranarray(rng.integers(-4, 5, size=(m, k) if ta == "N" else (k, m)).astype(float))
    B = np.asfortranarray(rng.integers(-4, 5, size=(k, n) if tb == "N" else (n, k)).astype(float))
    ldc = m + 3
    C0 = _padded(rng.integers(-4, 5, size=(m, n)).astype(float), ldc)
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dgemm(ta, tb, m, n, k, 2.0, A, A.shape[0], B, B.shape[0], -3.0, C, ldc)
    ora.ora_dgemm(c(ta), c(tb), m, n, k, 2.0, p(A), A.shape[0], p(B), B.shape[0], -3.0, p(R), ldc)
    assert np.array_equal(C, R)
    assert np.array_equal(C[m:, :], C0[m:, :])


def test_dgemm_beta_zero_ignores_nan(gpu):
    m = n = k = 70
    A = uniform(1, (m, k))
    B = uniform(2, (k, n))
    C = np.full((m, n), np.nan, order="F")
    gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.all(np.isfinite(C))
    np.testing.assert_allclose(C, A @ B, rtol=0, atol=1e-12)


def test_dgemm_alpha_zero_scales(gpu):
    """alpha = 0 on a device-resident C: C := beta*C on the GPU (netlib)."""
    gb = gpu
    C = gb.tracked_copy(uniform(3, (40, 30)))
    ref = 2.5 * np.array(C)
    A = uniform(4, (40, 10))
    B = uniform(5, (10, 30))
    gb.dgemm("N", "N", 40, 30, 10, 0.0, A, 40, B, 10, 2.5, C, 40)
    assert np.array_equal(np.array(C), ref)


def test_dgemm_config1_normwise(gpu, ora):
    """M=N=K=1024 NN, alpha=1, beta=0.5, U[-1,1]: the normwise bound of the 1024^3 DGEMM workload
    ||C - Cref||_F / (||A||_F ||B||_F k eps) <= 4 against the oracle."""
    n = 1024
    A, B, C0 = uniform(101, (n, n)), uniform(102, (n, n)), uniform(103, (n, n))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dgemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.5, C, n)
    ora.ora_dgemm(b"N", b"N", n, n, n, 1.0, p(A), n, p(B), n, 0.5, p(R), n)
    err = np.linalg.norm(C - R) / (np.linalg.norm(A) * np.linalg.norm(B) * n * EPS64)
    assert err <= 4.0, err


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T"), ("N", "T")])
@pytest.mark.parametrize("m,n,k", [(33, 65, 17), (256, 256, 300)])
def test_sgemm(gpu, ora, ta, tb, m, n, k):
    rows_a, cols_a = (m, k) if ta == "N" else (k, m)
    rows_b, cols_b = (k, n) if tb == "N" else (n, k)
    A = uniform(21, (rows_a, cols_a), np.float32)
    B = uniform(22, (rows_b, cols_b), np.float32)
    C0 = uniform(23, (m, n), np.float32)
    C = C0.copy(order="F")
    gpu.sgemm(ta, tb, m, n, k, 0.5, A, rows_a, B, rows_b, 2.0, C, m)
    # gauge in float64 arithmetic on the float32 data, eps of float32
    r = gauge_ratio(ora, ta, tb, m, n, k, 0.5, A.astype(np.float64, order="F"), rows_a,
                    B.astype(np.float64, order="F"), rows_b, 2.0,
                    C0.astype(np.float64, order="F"), C.astype(np.float64, order="F"), m, EPS32)
    assert r < 16.0, r


def test_sgemm_integer_exact(gpu, ora):
    rng = np.random.default_rng(9)
    m, n, k = 96, 80, 64
    A = np.asfortranarray(rng.integers(-4, 5, size=(m, k)).astype(np.float32))
    B = np.asfortranarray(rng.integers(-4, 5, size=(k, n)).astype(np.float32))
    C = np.zeros((m, n), np.float32, order="F")
    R = C.copy(order="F")
    gpu.sgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    ora.ora_sgemm(b"N", b"N", m, n, k, 1.0, p(A), m, p(B), k, 0.0, p(R), m)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(1, 1), (63, 20), (130, 257)])
def test_dsyrk(gpu, ora, uplo, trans, n, k):
    rows, cols = (n, k) if trans == "N" else (k, n)
    A = uniform(31, (rows, cols))
    C0 = uniform(32, (n, n))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dsyrk(uplo, trans, n, k, 0.9, A, rows, 0.4, C, n)
    ora.ora_dsyrk(c(uplo), c(trans), n, k, 0.9, p(A), rows, 0.4, p(R), n)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(C[~tri], C0[~tri])  # other triangle untouched
    G = 0.9 * (np.abs(A) @ np.abs(A).T if trans == "N" else np.abs(A).T @ np.abs(A)) + 0.4 * np.abs(C0)
    ratio = np.max(np.abs(C - R)[tri] / (EPS64 * G[tri]))
    assert ratio < 16.0, ratio


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
def test_dgemm_short_last_round_exact(gpu, ora, ta, tb):
    """m = n = 2944: 23 x 23 = 529 full 128 x 128 tiles, one more than a round
    of 512 workgroups, so the last 17 tiles run as 64 x 64 quarters.  Integer
    data: bit-exact against the oracle, quarters and big tiles alike."""
    rng = np.random.default_rng(21)
    m = n = 2944
    k = 48
    A = np.asfortranarray(rng.integers(-4, 5, size=(m, k) if ta == "N" else (k, m)).astype(float))
    B = np.asfortranarray(rng.integers(-4, 5, size=(k, n) if tb == "N" else (n, k)).astype(float))
    C = np.asfortranarray(rng.integers(-4, 5, size=(m, n)).astype(float))
    R = C.copy(order="F")
    gpu.dgemm(ta, tb, m, n, k, 1.0, A, A.shape[0], B, B.shape[0], 2.0, C, m)
    ora.ora_dgemm(c(ta), c(tb), m, n, k, 1.0, p(A), A.shape[0], p(B), B.shape[0], 2.0, p(R), m)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dsyrk_short_last_round_exact(gpu, ora, uplo):
    """n = 4096: 32 * 33 / 2 = 528 triangle tiles, so the last 16 (the tail of
    the last tile row, diagonal tile included) run as 64 x 64 quarters, the
    quarter above (below) the diagonal skipped.  Integer data: bit-exact, and
    the other triangle untouched."""
    rng = np.random.default_rng(22)
    n, k = 4096, 40
    A = np.asfortranarray(rng.integers(-4, 5, size=(n, k)).astype(float))
    C0 = np.asfortranarray(rng.integers(-4, 5, size=(n, n)).astype(float))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dsyrk(uplo, "N", n, k, 1.0, A, n, 3.0, C, n)
    ora.ora_dsyrk(c(uplo), c("N"), n, k, 1.0, p(A), n, 3.0, p(R), n)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("m,n", [(5, 9), (70, 130), (200, 64)])
def test_dsymm(gpu, ora, side, uplo, m, n):
    na = m if side == "L" else n
    S = uniform(41, (na, na))
    S = (S + S.T) / 2
    A = np.asfortranarray(S.copy())
    junk = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    A[junk] = ROGUE  # the unreferenced triangle must not be read
    B = uniform(42, (m, n))
    C0 = uniform(43, (m, n))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dsymm(side, uplo, m, n, 1.1, A, na, B, m, 0.3, C, m)
    ora.ora_dsymm(c(side), c(uplo), m, n, 1.1, p(A), na, p(B), m, 0.3, p(R), m)
    full = S @ B if side == "L" else B @ S
    G = 1.1 * (np.abs(S) @ np.abs(B) if side == "L" else np.abs(B) @ np.abs(S)) + 0.3 * np.abs(C0)
    assert np.max(np.abs(C - R) / (EPS64 * G)) < 16.0
    np.testing.assert_allclose(C, 1.1 * full + 0.3 * C0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("m,n", [(300, 257), (257, 320)])
def test_dsymm_expanded_exact(gpu, ora, side, uplo, m, n):
    """From order 256 the stored triangle is expanded into a full symmetric
    matrix before the product (ragged 64-tiles on the edge, padded lda, the
    other triangle ROGUE so any read of it shows).  Integer data: bit-exact."""
    rng = np.random.default_rng(31)
    na = m if side == "L" else n
    lda = na + 5
    S = rng.integers(-4, 5, size=(na, na)).astype(float)
    S = np.tril(S) + np.tril(S, -1).T
    A = np.full((lda, na), ROGUE, order="F")
    A[:na, :] = S
    junk = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    A[:na, :][junk] = ROGUE
    B = np.asfortranarray(rng.integers(-4, 5, size=(m, n)).astype(float))
    C0 = np.asfortranarray(rng.integers(-4, 5, size=(m, n)).astype(float))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    gpu.dsymm(side, uplo, m, n, 2.0, A, lda, B, m, -1.0, C, m)
    ora.ora_dsymm(c(side), c(uplo), m, n, 2.0, p(A), lda, p(B), m, -1.0, p(R), m)
    assert np.array_equal(C, R)
    full = S @ B if side == "L" else B @ S
    assert np.array_equal(C, 2.0 * full - C0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(7, 5), (130, 70), (300, 257)])
def test_dsyr2k(gpu, ora, uplo, trans, n, k):
    rows, cols = (n, k) if trans == "N" else (k, n)
    A = uniform(35, (rows, cols))
    B = uniform(36, (rows, cols))
    C0 = uniform(37, (n, n))
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    before = _device_calls(gpu)
    gpu.dsyr2k(uplo, trans, n, k, 0.9, A, rows, B, rows, 0.4, C, n)
    assert _device_calls(gpu) == before + 1
    ora.ora_dsyr2k(c(uplo), c(trans), n, k, 0.9, p(A), rows, p(B), rows, 0.4, p(R), n)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(C[~tri], C0[~tri])  # other triangle untouched
    aA, aB = (np.abs(A), np.abs(B)) if trans == "N" else (np.abs(A).T, np.abs(B).T)
    G = 0.9 * (aA @ aB.T + aB @ aA.T) + 0.4 * np.abs(C0)
    ratio = np.max(np.abs(C - R)[tri] / (EPS64 * G[tri]))
    assert ratio < 16.0, ratio


def test_dsyr2k_integer_exact(gpu, ora):
    rng = np.random.default_rng(51)
    n, k = 290, 150
    A = np.asfortranarray(rng.integers(-5, 6, size=(n, k)).astype(float))
    B = np.asfortranarray(rng.integers(-5, 6, size=(n, k)).astype(float))
    C = np.asfortranarray(rng.integers(-5, 6, size=(n, n)).astype(float))
    R = C.copy(order="F")
    gpu.dsyr2k("U", "N", n, k, 2.0, A, n, B, n, -1.0, C, n)
    ora.ora_dsyr2k(c("U"), c("N"), n, k, 2.0, p(A), n, p(B), n, -1.0, p(R), n)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("m,n", [(7, 5), (100, 70), (300, 257)])
def test_dtrmm(gpu, ora, side, uplo, ta, diag, m, n):
    na = m if side == "L" else n
    A = uniform(55, (na, na))
    junk = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    A[junk] = ROGUE  # the other triangle must not be read
    if diag == "U":
        A[np.diag_indices(na)] = ROGUE  # nor the diagonal of a unit triangle
    A = np.asfortranarray(A)
    ldb = m + 2
    B0 = _padded(uniform(56, (m, n)), ldb)
    B = B0.copy(order="F")
    R = B0.copy(order="F")
    before = _device_calls(gpu)
    gpu.dtrmm(side, uplo, ta, diag, m, n, 0.8, A, na, B, ldb)
    assert _device_calls(gpu) == before + 1
    ora.ora_dtrmm(c(side), c(uplo), c(ta), c(diag), m, n, 0.8, p(A), na, p(R), ldb)
    assert np.array_equal(B[m:, :], B0[m:, :])  # padding rows untouched
    T = np.where(junk, 0.0, A)
    if diag == "U":
        T[np.diag_indices(na)] = 1.0
    opT = np.abs(T if ta == "N" else T.T)
    G = 0.8 * (opT @ np.abs(B0[:m]) if side == "L" else np.abs(B0[:m]) @ opT)
    ratio = np.max(np.abs(B[:m] - R[:m]) / (EPS64 * np.maximum(G, np.finfo(float).tiny)))
    assert ratio < 16.0, ratio


def test_dtrmm_integer_exact(gpu, ora):
    rng = np.random.default_rng(57)
    m, n = 260, 130
    A = np.asfortranarray(rng.integers(-5, 6, size=(m, m)).astype(float))
    B = np.asfortranarray(rng.integers(-5, 6, size=(m, n)).astype(float))
    R = B.copy(order="F")
    gpu.dtrmm("L", "L", "T", "N", m, n, 3.0, A, m, B, m)
    ora.ora_dtrmm(c("L"), c("L"), c("T"), c("N"), m, n, 3.0, p(A), m, p(R), m)
    assert np.array_equal(B, R)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrmm_triangular_k_exact(gpu, ora, side, uplo, ta, diag):
    """Full 64 x 64 tiles: every tile runs only over the k range where op(A)
    is nonzero (all four range rules across side x op(A) lower/upper).
    Integer data, the other triangle ROGUE: bit-exact."""
    rng = np.random.default_rng(61)
    m, n = 256, 192
    na = m if side == "L" else n
    A = rng.integers(-5, 6, size=(na, na)).astype(float)
    junk = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    A[junk] = ROGUE
    A = np.asfortranarray(A)
    B = np.asfortranarray(rng.integers(-5, 6, size=(m, n)).astype(float))
    R = B.copy(order="F")
    gpu.dtrmm(side, uplo, ta, diag, m, n, 2.0, A, na, B, m)
    ora.ora_dtrmm(c(side), c(uplo), c(ta), c(diag), m, n, 2.0, p(A), na, p(R), m)
    assert np.array_equal(B, R)


def test_dtrmm_triangular_k_big_tiles_exact(gpu, ora):
    """2048 x 2048: 256 tiles of 128 x 128 with the triangular k range."""
    rng = np.random.default_rng(62)
    m = n = 2048
    A = np.asfortranarray(np.tril(rng.integers(-3, 4, size=(m, m)).astype(float)))
    B = np.asfortranarray(rng.integers(-3, 4, size=(m, n)).astype(float))
    R = B.copy(order="F")
    gpu.dtrmm("L", "L", "N", "N", m, n, 1.0, A, m, B, m)
    ora.ora_dtrmm(c("L"), c("L"), c("N"), c("N"), m, n, 1.0, p(A), m, p(R), m)
    assert np.array_equal(B, R)


def _tri_matrix(seed, n, uplo, diag_boost=True):
    A = uniform(seed, (n, n))
    if diag_boost:
        A[np.diag_indices(n)] = n + np.abs(uniform(seed + 1, (n,)))
    junk = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
    A[junk] = ROGUE
    return np.asfortranarray(A)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("m,n", [(7, 5), (100, 70), (200, 150), (300, 260)])
def test_dtrsm(gpu, ora, side, uplo, ta, diag, m, n):
    """Orders up to 64 go to the 64 x 64 leaf, up to 256 to the one-launch
    base solve, larger ones recurse into those with GEMM updates."""
    na = m if side == "L" else n
    A = _tri_matrix(51, na, uplo, diag_boost=(diag == "N"))
    if diag == "U":
        # unit diagonal: keep the strict triangle small so the solve stays well conditioned
        tri = np.tril(np.ones((na, na), bool), -1) if uplo == "L" else np.triu(np.ones((na, na), bool), 1)
        A[tri] *= 1.0 / na
    B0 = uniform(53, (m, n))
    B = B0.copy(order="F")
    R = B0.copy(order="F")
    gpu.dtrsm(side, uplo, ta, diag, m, n, 0.75, A, na, B, m)
    ora.ora_dtrsm(c(side), c(uplo), c(ta), c(diag), m, n, 0.75, p(A), na, p(R), m)
    np.testing.assert_allclose(B, R, rtol=1e-11, atol=1e-13)


def test_dtrsm_integer_exact(gpu, ora):
    """Unit lower A with small integers and B = A @ X_int: X is recovered
    exactly (every intermediate is an integer below 2^53)."""
    rng = np.random.default_rng(3)
    n, nrhs = 130, 40
    L = np.tril(rng.integers(-2, 3, size=(n, n)).astype(float), -1) + np.eye(n)
    X = rng.integers(-3, 4, size=(n, nrhs)).astype(float)
    B = np.asfortranarray(L @ X)
    A = np.asfortranarray(L)
    gpu.dtrsm("L", "L", "N", "U", n, nrhs, 1.0, A, n, B, n)
    assert np.array_equal(B, X)


def test_strsm(gpu, ora):
    m, n = 90, 33
    A = _tri_matrix(61, m, "U").astype(np.float32, order="F")
    B0 = uniform(63, (m, n), np.float32)
    B = B0.copy(order="F")
    R = B0.copy(order="F")
    gpu.strsm("L", "U", "N", "N", m, n, 1.0, A, m, B, m)
    ora.ora_strsm(b"L", b"U", b"N", b"N", m, n, 1.0, p(A), m, p(R), m)
    np.testing.assert_allclose(B, R, rtol=2e-5, atol=1e-6)
