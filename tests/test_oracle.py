"""Pins of the CPU oracle (oracle/refblas.c) -- CPU only.

The oracle restates the netlib semantics the reference's host fall-through
computes.  It is checked here against
  * the reference's own known-answer data: the DMMCH exact-data self-test of
    tests/netlib/dblat3.f:195-256 (all four trans forms of GEMM, bit-exact);
  * the host BLAS the reference falls through to on this machine (scipy's
    OpenBLAS 0.3.29, through scipy.linalg.blas), within the DMMCH gauge;
  * closed forms (nrm2 of constant vectors, idamax tie/NaN rules, identity
    and integer-exact products, triangular solves with integer solutions);
  * the published SplitMix64 test vector for the input generator.
Each check is chosen so that a dropped term, a swapped index or a wrong
transpose in the oracle would fail it.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.blas as sb

import _oracle
from _oracle import c, p, uniform

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def L():
    return _oracle.load()


def test_splitmix64_published_vector(L):
    # Vigna's splitmix64.c, seed 1234567: the first five outputs
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [L.ora_splitmix64_at(1234567, i) for i in range(5)] == want


def test_uniform_generator_exact_grid(L):
    u = uniform(77, (100000,))
    assert u.min() >= -1.0 and u.max() < 1.0
    assert np.all(np.floor(u * 2.0 ** 52) == u * 2.0 ** 52)  # on the 2^-52 grid
    assert abs(u.mean()) < 0.01


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_dmmch_exact_data_selftest(L, ta, tb):
    """dblat3.f:195-256: AB(i,j) = max(i-j+1, 0), b_j = j (NN/NT) or N-j+1
    (TN/TT); the exact product is CC(j) = j((j+1)j)/2 - ((j+1)j(j-1))/3,
    reversed for the transposed-A forms."""
    n = 32
    i = np.arange(1, n + 1)[:, None]
    j = np.arange(1, n + 1)[None, :]
    AB = np.asfortranarray(np.maximum(i - j + 1, 0).astype(float))
    jj = np.arange(1, n + 1)
    cc = (jj * ((jj + 1) * jj) // 2 - ((jj + 1) * jj * (jj - 1)) // 3).astype(float)
    if ta == "N":
        bvec = jj.astype(float)
        want = cc
    else:
        bvec = (n - jj + 1).astype(float)
        want = cc[::-1].copy()
    B = np.asfortranarray(bvec.reshape(n, 1) if tb == "N" else bvec.reshape(1, n))
    C = np.zeros((n, 1), order="F")
    L.ora_dgemm(c(ta), c(tb), n, 1, n, 1.0, p(AB), n, p(B), B.shape[0], 0.0, p(C), n)
    assert np.array_equal(C[:, 0], want)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (3, 5, 2), (17, 9, 33)])
def test_dgemm_vs_host_blas(L, ta, tb, m, n, k):
    A = uniform(1, (m, k) if ta == "N" else (k, m))
    B = uniform(2, (k, n) if tb == "N" else (n, k))
    C0 = uniform(3, (m, n))
    R = C0.copy(order="F")
    L.ora_dgemm(c(ta), c(tb), m, n, k, 0.7, p(A), A.shape[0], p(B), B.shape[0], 1.3, p(R), m)
    H = np.asfortranarray(sb.dgemm(0.7, A, B, 1.3, C0.copy(order="F"), trans_a=ta == "T",
                                   trans_b=tb == "T"))
    ratio = L.ora_gemm_gauge_ratio(c(ta), c(tb), m, n, k, 0.7, p(A), A.shape[0], p(B),
                                   B.shape[0], 1.3, p(C0), p(H), m, EPS)
    assert ratio < 16
    ratio = L.ora_gemm_gauge_ratio(c(ta), c(tb), m, n, k, 0.7, p(A), A.shape[0], p(B),
                                   B.shape[0], 1.3, p(C0), p(R), m, EPS)
    assert ratio < 16


def test_gauge_checker_catches_planted_error(L):
    m = n = k = 20
    A, B, C0 = uniform(4, (m, k)), uniform(5, (k, n)), uniform(6, (m, n))
    R = C0.copy(order="F")
    L.ora_dgemm(b"N", b"N", m, n, k, 1.0, p(A), m, p(B), k, 1.0, p(R), m)
    G = np.abs(A) @ np.abs(B) + np.abs(C0)
    bad = R.copy(order="F")
    bad[3, 7] += 100 * EPS * G[3, 7]
    assert L.ora_gemm_gauge_ratio(b"N", b"N", m, n, k, 1.0, p(A), m, p(B), k, 1.0, p(C0),
                                  p(bad), m, EPS) > 16
    # a transposed operand is a gross error
    wrong = np.asfortranarray(A.T @ B + C0)
    assert L.ora_gemm_gauge_ratio(b"N", b"N", m, n, k, 1.0, p(A), m, p(B), k, 1.0, p(C0),
                                  p(wrong), m, EPS) > 1e6


def test_dgemm_beta_zero_and_alpha_zero_netlib(L):
    A, B = uniform(7, (4, 3)), uniform(8, (3, 5))
    C = np.full((4, 5), np.nan, order="F")
    L.ora_dgemm(b"N", b"N", 4, 5, 3, 1.0, p(A), 4, p(B), 3, 0.0, p(C), 4)
    assert np.all(np.isfinite(C))  # beta == 0: C is not read
    C = np.full((4, 5), np.nan, order="F")
    L.ora_dgemm(b"N", b"N", 4, 5, 3, 0.0, p(A), 4, p(B), 3, 0.0, p(C), 4)
    assert np.all(C == 0.0)  # alpha == beta == 0 gives exact zeros


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dsyrk_vs_host_blas(L, uplo, trans):
    n, k = 13, 7
    A = uniform(9, (n, k) if trans == "N" else (k, n))
    C0 = uniform(10, (n, n))
    R = C0.copy(order="F")
    L.ora_dsyrk(c(uplo), c(trans), n, k, 0.5, p(A), A.shape[0], 2.0, p(R), n)
    H = sb.dsyrk(0.5, A, 2.0, C0.copy(order="F"), trans=int(trans == "T"), lower=int(uplo == "L"))
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    np.testing.assert_allclose(R[tri], H[tri], rtol=1e-14, atol=1e-14)
    assert np.array_equal(R[~tri], C0[~tri])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 2.0])
def test_dsyr2k_vs_host_blas_and_closed_form(L, uplo, trans, beta):
    n, k = 13, 7
    A = uniform(19, (n, k) if trans == "N" else (k, n))
    B = uniform(20, (n, k) if trans == "N" else (k, n))
    C0 = uniform(21, (n, n))
    R = C0.copy(order="F")
    L.ora_dsyr2k(c(uplo), c(trans), n, k, 0.5, p(A), A.shape[0], p(B), B.shape[0], beta, p(R), n)
    H = sb.dsyr2k(0.5, A, B, beta, C0.copy(order="F"), trans=int(trans == "T"),
                  lower=int(uplo == "L"))
    opA, opB = (A, B) if trans == "N" else (A.T, B.T)
    exact = 0.5 * (opA @ opB.T + opB @ opA.T) + beta * C0
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    np.testing.assert_allclose(R[tri], H[tri], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(R[tri], exact[tri], rtol=1e-13, atol=1e-13)
    assert np.array_equal(R[~tri], C0[~tri])


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrmm_vs_host_blas_and_closed_form(L, side, uplo, ta, diag):
    m, n = 11, 9
    na = m if side == "L" else n
    A = uniform(23, (na, na))
    B0 = uniform(24, (m, n))
    R = B0.copy(order="F")
    L.ora_dtrmm(c(side), c(uplo), c(ta), c(diag), m, n, 0.7, p(A), na, p(R), m)
    H = sb.dtrmm(0.7, A, B0.copy(order="F"), side=int(side == "R"), lower=int(uplo == "L"),
                 trans_a=int(ta == "T"), diag=int(diag == "U"))
    T = np.tril(A) if uplo == "L" else np.triu(A)
    if diag == "U":
        T[np.diag_indices(na)] = 1.0
    opT = T if ta == "N" else T.T
    exact = 0.7 * (opT @ B0 if side == "L" else B0 @ opT)
    np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(R, exact, rtol=1e-13, atol=1e-13)


def test_syr2k_trmm_info_codes(L):
    s = L.ora_syr2k_info
    assert s(b"X", b"N", 2, 2, 2, 2, 2) == 1
    assert s(b"L", b"X", 2, 2, 2, 2, 2) == 2
    assert s(b"L", b"N", -1, 2, 2, 2, 2) == 3
    assert s(b"L", b"N", 2, -1, 2, 2, 2) == 4
    assert s(b"L", b"N", 3, 2, 2, 3, 3) == 7
    assert s(b"L", b"T", 3, 4, 4, 3, 3) == 9
    assert s(b"L", b"N", 3, 2, 3, 3, 2) == 12
    assert s(b"U", b"T", 3, 4, 4, 4, 3) == 0
    t = L.ora_trmm_info
    assert t(b"X", b"L", b"N", b"N", 2, 2, 2, 2) == 1
    assert t(b"L", b"X", b"N", b"N", 2, 2, 2, 2) == 2
    assert t(b"L", b"L", b"X", b"N", 2, 2, 2, 2) == 3
    assert t(b"L", b"L", b"N", b"X", 2, 2, 2, 2) == 4
    assert t(b"L", b"L", b"N", b"N", -1, 2, 2, 2) == 5
    assert t(b"L", b"L", b"N", b"N", 2, -1, 2, 2) == 6
    assert t(b"R", b"L", b"N", b"N", 2, 3, 2, 2) == 9
    assert t(b"L", b"L", b"N", b"N", 3, 2, 3, 2) == 11


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrsm_vs_host_blas(L, side, uplo, ta, diag):
    m, n = 11, 6
    na = m if side == "L" else n
    A = uniform(11, (na, na))
    A[np.diag_indices(na)] += 4.0
    B0 = uniform(12, (m, n))
    R = B0.copy(order="F")
    L.ora_dtrsm(c(side), c(uplo), c(ta), c(diag), m, n, 0.5, p(A), na, p(R), m)
    H = sb.dtrsm(0.5, A, B0.copy(order="F"), side=int(side == "R"), lower=int(uplo == "L"),
                 trans_a=int(ta == "T"), diag=int(diag == "U"))
    np.testing.assert_allclose(R, H, rtol=1e-13, atol=1e-14)


def test_dtrsm_integer_solution(L):
    rng = np.random.default_rng(0)
    n, r = 12, 4
    U = np.triu(rng.integers(-3, 4, (n, n)).astype(float), 1) + np.eye(n)
    X = rng.integers(-5, 6, (n, r)).astype(float)
    B = np.asfortranarray(U @ X)
    Uf = np.asfortranarray(U)
    L.ora_dtrsm(b"L", b"U", b"N", b"U", n, r, 1.0, p(Uf), n, p(B), n)
    assert np.array_equal(B, X)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dsymm_vs_host_blas(L, side, uplo):
    m, n = 9, 7
    na = m if side == "L" else n
    A = uniform(13, (na, na))
    B, C0 = uniform(14, (m, n)), uniform(15, (m, n))
    R = C0.copy(order="F")
    L.ora_dsymm(c(side), c(uplo), m, n, 1.5, p(A), na, p(B), m, 0.25, p(R), m)
    H = sb.dsymm(1.5, A, B, 0.25, C0.copy(order="F"), side=int(side == "R"), lower=int(uplo == "L"))
    np.testing.assert_allclose(R, H, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_dgemv_vs_host_blas(L, trans, incx, incy):
    m, n = 15, 11
    A = uniform(16, (m, n))
    lx, ly = (n, m) if trans == "N" else (m, n)
    x = uniform(17, (1 + (lx - 1) * abs(incx),))
    y0 = uniform(18, (1 + (ly - 1) * abs(incy),))
    R = y0.copy()
    L.ora_dgemv(c(trans), m, n, 0.9, p(A), m, p(x), incx, 0.4, p(R), incy)
    H = sb.dgemv(0.9, A, x, beta=0.4, y=y0.copy(), incx=abs(incx), incy=abs(incy),
                 offx=0, offy=0, trans=int(trans == "T"))
    if incx > 0 and incy > 0:
        np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-15)
    # independent closed form with BLAS stride conventions
    xs = x[:: abs(incx)][:lx]
    xs = xs if incx > 0 else xs[::-1]
    ys = y0[:: abs(incy)][:ly]
    ys = ys if incy > 0 else ys[::-1]
    want = 0.9 * ((A if trans == "N" else A.T) @ xs) + 0.4 * ys
    got = R[:: abs(incy)][:ly]
    got = got if incy > 0 else got[::-1]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


def test_dger_and_axpy_closed_forms(L):
    m, n = 8, 6
    A0 = uniform(19, (m, n))
    x, y = uniform(20, (m,)), uniform(21, (n,))
    R = A0.copy(order="F")
    L.ora_dger(m, n, 2.0, p(x), 1, p(y), 1, p(R), m)
    np.testing.assert_allclose(R, A0 + 2.0 * np.outer(x, y), rtol=1e-15, atol=1e-15)
    yy = uniform(22, (m,))
    r = yy.copy()
    L.ora_daxpy(m, -0.5, p(x), 1, p(r), 1)
    assert np.array_equal(r, yy + (-0.5) * x)
    r = yy.copy()
    L.ora_daxpy(4, 2.0, p(x), -2, p(r), 1)  # x walked from its far end
    assert np.array_equal(r[:4], yy[:4] + 2.0 * x[6::-2])  # x(7), x(5), x(3), x(1)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtrsv_vs_host_blas(L, uplo, trans):
    n = 10
    A = uniform(23, (n, n))
    A[np.diag_indices(n)] += 5.0
    x0 = uniform(24, (n,))
    R = x0.copy()
    L.ora_dtrsv(c(uplo), c(trans), b"N", n, p(A), n, p(R), 1)
    H = sb.dtrsv(A, x0.copy(), lower=int(uplo == "L"), trans=int(trans == "T"), diag=0)
    np.testing.assert_allclose(R, H, rtol=1e-13, atol=1e-14)


def test_ddot_nrm2_vs_host_blas(L):
    x, y = uniform(25, (1000,)), uniform(26, (1000,))
    assert abs(L.ora_ddot(1000, p(x), 1, p(y), 1) - sb.ddot(x, y)) <= 1e3 * EPS * np.abs(x * y).sum()
    assert L.ora_dnrm2(1000, p(x), 1) == pytest.approx(sb.dnrm2(x), rel=1e-14)
    assert L.ora_dnrm2(0, p(x), 1) == 0.0
    assert L.ora_dnrm2(1, p(x), 1) == abs(x[0])


@pytest.mark.parametrize("scale", [1e300, 1e-300, 1.0])
def test_dnrm2_closed_form(L, scale):
    x = np.full(400, scale)
    assert L.ora_dnrm2(400, p(x), 1) == pytest.approx(scale * 20.0, rel=1e-15)


def test_idamax_rules(L):
    x = np.array([1.0, -3.0, 3.0, 2.0])
    assert L.ora_idamax(4, p(x), 1) == 2  # first of the tie, 1-based
    assert sb.idamax(x) + 1 == 2  # scipy returns 0-based
    x = np.array([np.nan, 5.0, 1.0])
    assert L.ora_idamax(3, p(x), 1) == 1
    x = np.array([1.0, np.nan, 5.0])
    assert L.ora_idamax(3, p(x), 1) == 3
    assert L.ora_idamax(0, p(x), 1) == 0
    assert L.ora_idamax(3, p(x), 0) == 0
    y = uniform(27, (5000,))
    assert L.ora_idamax(5000, p(y), 1) == int(np.argmax(np.abs(y))) + 1
    assert L.ora_idamax(5000, p(y), 1) == sb.idamax(y) + 1


def test_info_codes(L):
    # gemm.cc:87-116: 1 transa, 2 transb, 3 m, 4 n, 5 k, 8 lda, 10 ldb, 13 ldc
    g = L.ora_gemm_info
    assert g(b"X", b"N", 2, 2, 2, 2, 2, 2) == 1
    assert g(b"N", b"Q", 2, 2, 2, 2, 2, 2) == 2
    assert g(b"N", b"N", -1, 2, 2, 2, 2, 2) == 3
    assert g(b"N", b"N", 2, -1, 2, 2, 2, 2) == 4
    assert g(b"N", b"N", 2, 2, -1, 2, 2, 2) == 5
    assert g(b"N", b"N", 3, 2, 2, 2, 2, 3) == 8
    assert g(b"T", b"N", 3, 2, 4, 3, 3, 3) == 8  # nrowa = k for 'T'
    assert g(b"N", b"N", 2, 2, 3, 2, 2, 2) == 10
    assert g(b"N", b"T", 2, 3, 2, 2, 2, 2) == 10  # nrowb = n for 'T'
    assert g(b"N", b"N", 3, 2, 2, 3, 2, 2) == 13
    assert g(b"c", b"t", 2, 2, 2, 2, 2, 2) == 0
    s = L.ora_syrk_info  # syrk.cc:87-115
    assert [s(b"X", b"N", 2, 2, 2, 2), s(b"L", b"X", 2, 2, 2, 2), s(b"L", b"N", -1, 2, 2, 2),
            s(b"L", b"N", 2, -1, 2, 2), s(b"L", b"N", 3, 2, 2, 3), s(b"L", b"T", 2, 3, 2, 2),
            s(b"U", b"N", 3, 2, 3, 2)] == [1, 2, 3, 4, 7, 7, 10]
    t = L.ora_trsm_info  # trsm.cc:84-116
    assert [t(b"X", b"L", b"N", b"N", 2, 2, 2, 2), t(b"L", b"X", b"N", b"N", 2, 2, 2, 2),
            t(b"L", b"L", b"X", b"N", 2, 2, 2, 2), t(b"L", b"L", b"N", b"X", 2, 2, 2, 2),
            t(b"L", b"L", b"N", b"N", -1, 2, 2, 2), t(b"L", b"L", b"N", b"N", 2, -1, 2, 2),
            t(b"L", b"L", b"N", b"N", 3, 2, 2, 3), t(b"R", b"L", b"N", b"N", 2, 3, 2, 2),
            t(b"L", b"L", b"N", b"N", 3, 2, 3, 2)] == [1, 2, 3, 4, 5, 6, 9, 9, 11]
    y = L.ora_symm_info  # symm.cc:81-104
    assert [y(b"X", b"L", 2, 2, 2, 2, 2), y(b"L", b"X", 2, 2, 2, 2, 2), y(b"L", b"L", -1, 2, 2, 2, 2),
            y(b"L", b"L", 2, -1, 2, 2, 2), y(b"L", b"L", 3, 2, 2, 3, 3), y(b"R", b"L", 2, 3, 2, 2, 2),
            y(b"L", b"L", 3, 2, 3, 2, 3), y(b"L", b"L", 3, 2, 3, 3, 2)] == [1, 2, 3, 4, 7, 7, 9, 12]
    v = L.ora_gemv_info
    assert [v(b"X", 2, 2, 2, 1, 1), v(b"N", -1, 2, 2, 1, 1), v(b"N", 2, -1, 2, 1, 1),
            v(b"N", 3, 2, 2, 1, 1), v(b"N", 2, 2, 2, 0, 1), v(b"N", 2, 2, 2, 1, 0)] == [1, 2, 3, 6, 8, 11]
    r = L.ora_ger_info
    assert [r(-1, 2, 1, 1, 2), r(2, -1, 1, 1, 2), r(2, 2, 0, 1, 2), r(2, 2, 1, 0, 2),
            r(3, 2, 1, 1, 2)] == [1, 2, 5, 7, 9]
    tv = L.ora_trsv_info
    assert [tv(b"X", b"N", b"N", 2, 2, 1), tv(b"L", b"X", b"N", 2, 2, 1), tv(b"L", b"N", b"X", 2, 2, 1),
            tv(b"L", b"N", b"N", -1, 2, 1), tv(b"L", b"N", b"N", 3, 2, 1),
            tv(b"L", b"N", b"N", 2, 2, 0)] == [1, 2, 3, 4, 6, 8]


@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -3), (-1, 2), (-2, -1)])
def test_level1_more_vs_host_blas(L, incx, incy):
    """DSCAL/DCOPY/DSWAP/DROT/DASUM restatements against scipy's OpenBLAS,
    negative increments included (netlib walks those vectors backwards)."""
    n = 37
    x0 = uniform(81, (n * abs(incx),))
    y0 = uniform(82, (n * abs(incy),))
    # dcopy / dswap / drot: the host routine gets the same strided views
    X, Y = x0.copy(), y0.copy()
    L.ora_dcopy(n, p(X), incx, p(Y), incy)
    Hy = sb.dcopy(x0, y0.copy(), n=n, incx=incx, incy=incy)
    assert np.array_equal(Y, Hy)
    X, Y = x0.copy(), y0.copy()
    L.ora_dswap(n, p(X), incx, p(Y), incy)
    Hx, Hy = sb.dswap(x0.copy(), y0.copy(), n=n, incx=incx, incy=incy)
    assert np.array_equal(X, Hx) and np.array_equal(Y, Hy)
    X, Y = x0.copy(), y0.copy()
    L.ora_drot(n, p(X), incx, p(Y), incy, 0.6, 0.8)
    Hx, Hy = sb.drot(x0.copy(), y0.copy(), 0.6, 0.8, n=n, incx=incx, incy=incy)
    np.testing.assert_allclose(X, Hx, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(Y, Hy, rtol=1e-15, atol=1e-15)
    if incx > 0:
        X = x0.copy()
        L.ora_dscal(n, -1.5, p(X), incx)
        assert np.array_equal(X, sb.dscal(-1.5, x0.copy(), n=n, incx=incx))
        a = L.ora_dasum(n, p(x0), incx)
        np.testing.assert_allclose(a, sb.dasum(x0, n=n, incx=incx), rtol=1e-14)
        assert a == pytest.approx(np.sum(np.abs(x0[::incx][:n])), rel=1e-14)
    else:
        X = x0.copy()
        L.ora_dscal(n, -1.5, p(X), incx)  # netlib: no-op for incx <= 0
        assert np.array_equal(X, x0)
        assert L.ora_dasum(n, p(x0), incx) == 0.0


def _sym_from(A, uplo):
    """Full symmetric matrix from the uplo triangle of A (the other ignored)."""
    T = np.tril(A) if uplo == "L" else np.triu(A)
    return T + T.T - np.diag(np.diag(A))


def _strided_get(buf, n, inc):
    v = buf[:: abs(inc)][:n]
    return v if inc > 0 else v[::-1]


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_dsymv_vs_host_blas_and_closed_form(L, uplo, incx, incy):
    """netlib DSYMV restatement: scipy's OpenBLAS on unit strides, and the
    closed form alpha*S@x + beta*y with S built from the uplo triangle while
    the other triangle holds NaN (it must never be read)."""
    n = 13
    A = uniform(91, (n, n))
    other = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    An = A.copy(order="F")
    An[other] = np.nan
    x = uniform(92, (1 + (n - 1) * abs(incx),))
    y0 = uniform(93, (1 + (n - 1) * abs(incy),))
    R = y0.copy()
    L.ora_dsymv(c(uplo), n, 0.7, p(An), n, p(x), incx, -0.3, p(R), incy)
    want = 0.7 * (_sym_from(A, uplo) @ _strided_get(x, n, incx)) - 0.3 * _strided_get(y0, n, incy)
    np.testing.assert_allclose(_strided_get(R, n, incy), want, rtol=1e-13, atol=1e-14)
    if incx > 0 and incy > 0:
        H = sb.dsymv(0.7, A, x, beta=-0.3, y=y0.copy(), incx=incx, incy=incy,
                     lower=int(uplo == "L"))
        np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-15)
    # beta = 0 overwrites y (NaN in y must not survive); alpha = 0 only scales
    Y = np.full_like(y0, np.nan)
    L.ora_dsymv(c(uplo), n, 1.0, p(An), n, p(x), incx, 0.0, p(Y), incy)
    assert np.all(np.isfinite(_strided_get(Y, n, incy)))
    Y = y0.copy()
    L.ora_dsymv(c(uplo), n, 0.0, p(An), n, p(x), incx, 2.0, p(Y), incy)
    assert np.array_equal(_strided_get(Y, n, incy), 2.0 * _strided_get(y0, n, incy))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("incx", [1, 2, -2])
def test_dsyr_vs_host_blas_and_closed_form(L, uplo, incx):
    n = 11
    A0 = uniform(94, (n, n))
    x = uniform(95, (1 + (n - 1) * abs(incx),))
    x[0] = 0.0  # a zero entry: netlib skips its column
    R = A0.copy(order="F")
    L.ora_dsyr(c(uplo), n, -1.25, p(x), incx, p(R), n)
    xs = _strided_get(x, n, incx)
    full = A0 - 1.25 * np.outer(xs, xs)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    np.testing.assert_allclose(R[tri], full[tri], rtol=1e-14, atol=1e-15)
    assert np.array_equal(R[~tri], A0[~tri])  # the other triangle is untouched
    if incx == 1:
        H = sb.dsyr(-1.25, x, lower=int(uplo == "L"), a=A0.copy(order="F"))
        # OpenBLAS fuses the update (one rounding), netlib rounds twice
        np.testing.assert_allclose(R[tri], H[tri], rtol=1e-15, atol=1e-16)


def test_symv_syr_info_codes(L):
    assert [L.ora_symv_info(b"X", 2, 2, 1, 1), L.ora_symv_info(b"L", -1, 2, 1, 1),
            L.ora_symv_info(b"U", 3, 2, 1, 1), L.ora_symv_info(b"L", 2, 2, 0, 1),
            L.ora_symv_info(b"L", 2, 2, 1, 0), L.ora_symv_info(b"l", 2, 2, -1, -1)] == [1, 2, 5, 7, 10, 0]
    assert [L.ora_syr_info(b"X", 2, 1, 2), L.ora_syr_info(b"U", -1, 1, 2),
            L.ora_syr_info(b"U", 2, 0, 2), L.ora_syr_info(b"L", 3, 1, 2),
            L.ora_syr_info(b"u", 3, -2, 3)] == [1, 2, 5, 7, 0]


def _pack(A, uplo):
    """netlib packed storage of the uplo triangle of A (column by column)."""
    n = A.shape[0]
    if uplo == "U":
        return np.concatenate([A[: j + 1, j] for j in range(n)])
    return np.concatenate([A[j:, j] for j in range(n)])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1)])
def test_dsyr2_dspr_dspr2_vs_host_blas_and_closed_form(L, uplo, incx, incy):
    """netlib DSYR2 / DSPR / DSPR2: closed forms A + a(xy' + yx') and A + a xx'
    on the stored triangle (full or packed), the other triangle untouched, and
    scipy's OpenBLAS on unit strides; a zero column of x (and y) is skipped."""
    n = 12
    A0 = uniform(101, (n, n))
    x = uniform(102, (1 + (n - 1) * abs(incx),))
    y = uniform(103, (1 + (n - 1) * abs(incy),))
    x[0] = 0.0
    y[0] = 0.0
    xs, ys = _strided_get(x, n, incx), _strided_get(y, n, incy)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    R = A0.copy(order="F")
    L.ora_dsyr2(c(uplo), n, 0.5, p(x), incx, p(y), incy, p(R), n)
    want = A0 + 0.5 * (np.outer(xs, ys) + np.outer(ys, xs))
    np.testing.assert_allclose(R[tri], want[tri], rtol=1e-14, atol=1e-15)
    assert np.array_equal(R[~tri], A0[~tri])
    if incx == 1 and incy == 1:
        H = sb.dsyr2(0.5, x, y, lower=int(uplo == "L"), a=A0.copy(order="F"))
        np.testing.assert_allclose(R[tri], H[tri], rtol=1e-15, atol=1e-16)
    AP0 = _pack(A0, uplo)
    P = AP0.copy()
    L.ora_dspr(c(uplo), n, -0.75, p(x), incx, p(P))
    np.testing.assert_allclose(P, _pack(A0 - 0.75 * np.outer(xs, xs), uplo), rtol=1e-14, atol=1e-15)
    P2 = AP0.copy()
    L.ora_dspr2(c(uplo), n, 0.5, p(x), incx, p(y), incy, p(P2))
    np.testing.assert_allclose(P2, _pack(want, uplo), rtol=1e-14, atol=1e-15)
    if incx == 1 and incy == 1:
        np.testing.assert_allclose(P, sb.dspr(n, -0.75, x, AP0.copy(), lower=int(uplo == "L")),
                                   rtol=1e-15, atol=1e-16)
        np.testing.assert_allclose(P2, sb.dspr2(n, 0.5, x, y, AP0.copy(), lower=int(uplo == "L")),
                                   rtol=1e-15, atol=1e-16)


def test_rank_update_info_codes(L):
    assert [L.ora_syr2_info(b"X", 2, 1, 1, 2), L.ora_syr2_info(b"L", -1, 1, 1, 2),
            L.ora_syr2_info(b"L", 2, 0, 1, 2), L.ora_syr2_info(b"L", 2, 1, 0, 2),
            L.ora_syr2_info(b"U", 3, 1, 1, 2)] == [1, 2, 5, 7, 9]
    assert [L.ora_spr_info(b"X", 2, 1), L.ora_spr_info(b"U", -2, 1), L.ora_spr_info(b"U", 2, 0),
            L.ora_spr_info(b"l", 2, -1)] == [1, 2, 5, 0]
    assert [L.ora_spr2_info(b"U", 2, 1, 0), L.ora_spr2_info(b"U", 2, 0, 1)] == [7, 5]


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("incx", [1, -2])
def test_trmv_tpmv_tpsv_vs_host_blas_and_closed_form(L, uplo, trans, diag, incx):
    """netlib DTRMV / DTPMV / DTPSV: closed forms op(T) x and op(T)^-1 x with
    T the uplo triangle (unit diagonal for diag U; the other triangle NaN in
    the full-storage call, never read), packed storage giving the same
    result, and scipy's OpenBLAS on unit strides."""
    n = 11
    A = uniform(111, (n, n)) + np.diag(np.full(n, 3.0))
    T = np.tril(A) if uplo == "L" else np.triu(A)
    if diag == "U":
        T = T - np.diag(np.diag(T)) + np.eye(n)
    opT = T if trans == "N" else T.T
    An = A.copy(order="F")
    An[np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)] = np.nan
    x0 = uniform(112, (1 + (n - 1) * abs(incx),))
    xs = _strided_get(x0, n, incx)
    X = x0.copy()
    L.ora_dtrmv(c(uplo), c(trans), c(diag), n, p(An), n, p(X), incx)
    np.testing.assert_allclose(_strided_get(X, n, incx), opT @ xs, rtol=1e-13, atol=1e-14)
    AP = _pack(A, uplo)
    Xp = x0.copy()
    L.ora_dtpmv(c(uplo), c(trans), c(diag), n, p(AP), p(Xp), incx)
    np.testing.assert_allclose(Xp, X, rtol=1e-15, atol=1e-15)
    Xs = x0.copy()
    L.ora_dtpsv(c(uplo), c(trans), c(diag), n, p(AP), p(Xs), incx)
    np.testing.assert_allclose(_strided_get(Xs, n, incx), np.linalg.solve(opT, xs), rtol=1e-12,
                               atol=1e-13)
    if incx == 1:
        kw = dict(lower=int(uplo == "L"), trans=int(trans == "T"), diag=int(diag == "U"))
        np.testing.assert_allclose(X, sb.dtrmv(A.copy(order="F"), x0.copy(), **kw), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(Xp, sb.dtpmv(n, AP, x0.copy(), **kw), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(Xs, sb.dtpsv(n, AP, x0.copy(), **kw), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1)])
def test_dspmv_vs_dsymv_and_host_blas(L, uplo, incx, incy):
    """netlib DSPMV equals DSYMV on the same triangle (packed vs full), and
    scipy's OpenBLAS on unit strides."""
    n = 13
    A = uniform(113, (n, n))
    x = uniform(114, (1 + (n - 1) * abs(incx),))
    y0 = uniform(115, (1 + (n - 1) * abs(incy),))
    R, S = y0.copy(), y0.copy()
    L.ora_dspmv(c(uplo), n, 0.7, p(_pack(A, uplo)), p(x), incx, -0.3, p(R), incy)
    L.ora_dsymv(c(uplo), n, 0.7, p(A), n, p(x), incx, -0.3, p(S), incy)
    np.testing.assert_allclose(R, S, rtol=1e-15, atol=1e-15)
    if incx == 1 and incy == 1:
        H = sb.dspmv(n, 0.7, _pack(A, uplo), x, beta=-0.3, y=y0.copy(), lower=int(uplo == "L"))
        np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-15)


def test_trmv_tpmv_spmv_info_codes(L):
    assert [L.ora_trmv_info(b"X", b"N", b"N", 2, 2, 1), L.ora_trmv_info(b"L", b"X", b"N", 2, 2, 1),
            L.ora_trmv_info(b"L", b"N", b"X", 2, 2, 1), L.ora_trmv_info(b"L", b"N", b"N", -1, 2, 1),
            L.ora_trmv_info(b"L", b"N", b"N", 3, 2, 1), L.ora_trmv_info(b"L", b"N", b"N", 2, 2, 0)] \
        == [1, 2, 3, 4, 6, 8]
    assert [L.ora_tpmv_info(b"U", b"T", b"U", -1, 1), L.ora_tpmv_info(b"U", b"T", b"U", 2, 0)] == [4, 7]
    assert [L.ora_spmv_info(b"X", 2, 1, 1), L.ora_spmv_info(b"U", -1, 1, 1),
            L.ora_spmv_info(b"U", 2, 0, 1), L.ora_spmv_info(b"U", 2, 1, 0)] == [1, 2, 6, 9]


def _gb_dense(AB, m, n, kl, ku):
    """Dense m x n matrix of netlib general band storage A(ku+i-j, j)."""
    D = np.zeros((m, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            D[i, j] = AB[ku + i - j, j]
    return D


def _band_storage(D, kl, ku, lda, fill=np.nan):
    """netlib band storage of the band of D; entries outside the matrix (the
    corners of the band array) and rows beyond kl + ku hold `fill`."""
    m, n = D.shape
    AB = np.full((lda, n), fill, order="F")
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            AB[ku + i - j, j] = D[i, j]
    return AB


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,n,kl,ku", [(9, 9, 2, 1), (12, 7, 0, 3), (6, 11, 4, 0), (5, 5, 7, 6),
                                       (1, 4, 1, 2)])
@pytest.mark.parametrize("incx,incy", [(1, 1), (-2, 3)])
def test_dgbmv_vs_dense_and_host_blas(L, trans, m, n, kl, ku, incx, incy):
    """netlib DGBMV equals alpha*op(D)*x + beta*y for the dense D of the band
    (the band array's unreferenced corners hold NaN and are never read), and
    scipy's OpenBLAS on the same band with finite corners."""
    D = np.zeros((m, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            D[i, j] = uniform(120 + i, (n,))[j]
    lda = kl + ku + 2
    AB = _band_storage(D, kl, ku, lda)
    assert np.array_equal(_gb_dense(AB, m, n, kl, ku), D)
    lx, ly = (n, m) if trans == "N" else (m, n)
    xs, ys = uniform(121, (lx,)), uniform(122, (ly,))
    x = np.full(1 + (lx - 1) * abs(incx), 9.0)
    x[:: abs(incx)] = xs if incx > 0 else xs[::-1]
    y0 = np.full(1 + (ly - 1) * abs(incy), 9.0)
    y0[:: abs(incy)] = ys if incy > 0 else ys[::-1]
    R = y0.copy()
    L.ora_dgbmv(c(trans), m, n, kl, ku, 0.6, p(AB), lda, p(x), incx, -1.25, p(R), incy)
    Op = D if trans == "N" else D.T
    want = 0.6 * Op @ xs - 1.25 * ys
    got = R[:: abs(incy)][:ly]
    got = got if incy > 0 else got[::-1]
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14)
    gaps = np.ones(len(R), bool)
    gaps[:: abs(incy)][:ly] = False
    assert np.all(R[gaps] == 9.0)
    if incx == 1 and incy == 1 and m >= kl + ku + 1:  # scipy's wrapper requires the latter
        ABf = np.where(np.isnan(AB), 0.0, AB)
        H = sb.dgbmv(m, n, kl, ku, 0.6, ABf, x, beta=-1.25, y=y0.copy(), trans=int(trans == "T"))
        np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,k", [(10, 3), (6, 0), (5, 8), (1, 2)])
def test_dsbmv_vs_dense_and_host_blas(L, uplo, n, k):
    """netlib DSBMV: the symmetric matrix of the stored band triangle times x;
    scipy's OpenBLAS agrees."""
    A = uniform(123, (n, n))
    S = np.triu(A) if uplo == "U" else np.tril(A)
    S = S + S.T - np.diag(np.diag(A))
    S[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > k] = 0.0
    T = np.triu(S) if uplo == "U" else np.tril(S)
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = _band_storage(T, kl, ku, k + 1)
    x, y0 = uniform(124, (n,)), uniform(125, (n,))
    R = y0.copy()
    L.ora_dsbmv(c(uplo), n, k, 0.8, p(AB), k + 1, p(x), 1, 0.5, p(R), 1)
    np.testing.assert_allclose(R, 0.8 * S @ x + 0.5 * y0, rtol=1e-14, atol=1e-14)
    H = sb.dsbmv(k, 0.8, np.where(np.isnan(AB), 0.0, AB), x, beta=0.5, y=y0.copy(),
                 lower=int(uplo == "L"))
    np.testing.assert_allclose(R, H, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,k,incx", [(11, 3, 1), (7, 0, 1), (6, 9, -2), (1, 1, 1)])
def test_dtbmv_dtbsv_vs_dense_and_host_blas(L, uplo, trans, diag, n, k, incx):
    """netlib DTBMV / DTBSV: op(T) x and the solve of op(T) x = b with the
    dense band triangle T (a unit diagonal's band slots hold NaN and are not
    read); the solve is checked on integer data whose solution is exact, and
    both against scipy's OpenBLAS."""
    rng = np.random.default_rng(126)
    T = rng.integers(-3, 4, (n, n)).astype(float)
    T = np.triu(T) if uplo == "U" else np.tril(T)
    T[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > k] = 0.0
    T[np.diag_indices(n)] = 1.0 if diag == "U" else 2.0
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = _band_storage(T, kl, ku, k + 1)
    if diag == "U":
        AB[ku, :] = np.nan
    Op = T if trans == "N" else T.T
    xs = rng.integers(-5, 6, n).astype(float)
    x = np.full(1 + (n - 1) * abs(incx), 9.0)
    x[:: abs(incx)] = xs if incx > 0 else xs[::-1]

    def get(v):
        g = v[:: abs(incx)][:n]
        return g if incx > 0 else g[::-1]

    R = x.copy()
    L.ora_dtbmv(c(uplo), c(trans), c(diag), n, k, p(AB), k + 1, p(R), incx)
    assert np.array_equal(get(R), Op @ xs)  # small integers: exact
    b = x.copy()
    b[:: abs(incx)] = (Op @ xs) if incx > 0 else (Op @ xs)[::-1]
    L.ora_dtbsv(c(uplo), c(trans), c(diag), n, k, p(AB), k + 1, p(b), incx)
    np.testing.assert_allclose(get(b), xs, rtol=1e-13, atol=1e-13)
    if incx == 1:
        ABf = np.where(np.isnan(AB), 1.0, AB)
        kw = dict(lower=int(uplo == "L"), trans=int(trans == "T"), diag=int(diag == "U"))
        assert np.array_equal(sb.dtbmv(k, ABf, xs.copy(), **kw), get(R))
        np.testing.assert_allclose(sb.dtbsv(k, ABf, Op @ xs, **kw), get(b), rtol=1e-14, atol=1e-14)


def test_band_info_codes(L):
    assert [L.ora_gbmv_info(b"X", 1, 1, 0, 0, 1, 1, 1), L.ora_gbmv_info(b"N", -1, 1, 0, 0, 1, 1, 1),
            L.ora_gbmv_info(b"N", 1, -1, 0, 0, 1, 1, 1), L.ora_gbmv_info(b"N", 1, 1, -1, 0, 1, 1, 1),
            L.ora_gbmv_info(b"N", 1, 1, 0, -1, 1, 1, 1), L.ora_gbmv_info(b"N", 1, 1, 1, 1, 2, 1, 1),
            L.ora_gbmv_info(b"N", 1, 1, 0, 0, 1, 0, 1), L.ora_gbmv_info(b"N", 1, 1, 0, 0, 1, 1, 0)] \
        == [1, 2, 3, 4, 5, 8, 10, 13]
    assert [L.ora_sbmv_info(b"X", 1, 0, 1, 1, 1), L.ora_sbmv_info(b"U", -1, 0, 1, 1, 1),
            L.ora_sbmv_info(b"U", 1, -1, 1, 1, 1), L.ora_sbmv_info(b"U", 1, 1, 1, 1, 1),
            L.ora_sbmv_info(b"U", 1, 0, 1, 0, 1), L.ora_sbmv_info(b"U", 1, 0, 1, 1, 0)] \
        == [1, 2, 3, 6, 8, 11]
    assert [L.ora_tbmv_info(b"U", b"N", b"N", -1, 0, 1, 1), L.ora_tbmv_info(b"U", b"N", b"N", 1, -1, 1, 1),
            L.ora_tbmv_info(b"U", b"N", b"N", 1, 2, 2, 1), L.ora_tbmv_info(b"U", b"N", b"N", 1, 0, 1, 0)] \
        == [4, 5, 7, 9]


def _host_openblas():
    """scipy's bundled OpenBLAS (the host BLAS of this image), for routines
    scipy.linalg.blas does not wrap."""
    import glob
    import os
    import scipy
    d = os.path.dirname(scipy.__file__)
    libs = glob.glob(d + ".libs/libscipy_openblas*.so") + glob.glob(d + ".libs/libopenblas*.so")
    if not libs:
        pytest.skip("no bundled OpenBLAS next to scipy")
    return ctypes.CDLL(libs[0])


def test_idamin_closed_form_and_host_blas(L):
    """IDAMIN: first index of the minimal |x_i| (numpy argmin of |x| on finite
    data); ties keep the smaller index; NaN is never selected after position
    1 and a NaN at position 1 wins (the netlib IDAMAX loop with '<').  The
    host OpenBLAS idamin_ agrees on finite data; on NaN it follows no single
    rule (0.3.29 returns 1 for [2, NaN, 1, NaN] and 9, the NaN, for
    [5, 4, 3, 2, 1, .5, .25, .1, NaN, .01]), so NaN cases pin the loop only."""
    ob = _host_openblas()
    f = getattr(ob, "scipy_idamin_", None) or getattr(ob, "idamin_")
    f.restype = ctypes.c_int
    I = ctypes.c_int

    def host(x, inc=1):
        return f(ctypes.byref(I(len(x[::inc]))), p(x), ctypes.byref(I(inc)))

    rng = np.random.default_rng(130)
    for n in (1, 2, 7, 1000):
        x = rng.standard_normal(n)
        assert L.ora_idamin(n, p(x), 1) == np.argmin(np.abs(x)) + 1 == host(x)
    x = np.array([3.0, -1.0, 2.0, 1.0, -1.0])  # tie: first of the two -1 / 1 minima
    assert L.ora_idamin(5, p(x), 1) == 2 == host(x)
    x = np.array([2.0, np.nan, 1.0, np.nan])
    assert L.ora_idamin(4, p(x), 1) == 3
    x = np.array([np.nan, 0.5, 0.25])
    assert L.ora_idamin(3, p(x), 1) == 1
    x = np.array([5.0, 9.0, 0.0, 7.0, -0.0, 1.0])
    assert L.ora_idamin(3, p(x), 2) == 2  # elements 5, 0, -0: first zero, the 2nd element
    assert L.ora_idamin(0, p(x), 1) == 0 and L.ora_idamin(3, p(x), 0) == 0


@pytest.mark.parametrize("flag", [-2.0, -1.0, 0.0, 1.0])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_drotm_vs_host_blas_and_closed_form(L, flag, incx, incy):
    """netlib DROTM for every flag: the 2x2 H of the flag applied to each
    (x_i, y_i) pair (closed form), and scipy's OpenBLAS drotm."""
    n = 11
    rng = np.random.default_rng(131)
    xs, ys = rng.standard_normal(n), rng.standard_normal(n)
    param = np.array([flag, 0.7, -0.4, 1.3, 0.2])
    H = {-2.0: [[1, 0], [0, 1]], -1.0: [[0.7, 1.3], [-0.4, 0.2]], 0.0: [[1, 1.3], [-0.4, 1]],
         1.0: [[0.7, 1], [-1, 0.2]]}[flag]
    x = np.full(1 + (n - 1) * abs(incx), 9.0)
    x[:: abs(incx)] = xs if incx > 0 else xs[::-1]
    y = np.full(1 + (n - 1) * abs(incy), 9.0)
    y[:: abs(incy)] = ys if incy > 0 else ys[::-1]
    X, Y = x.copy(), y.copy()
    L.ora_drotm(n, p(X), incx, p(Y), incy, p(param))

    def get(v, inc):
        g = v[:: abs(inc)][:n]
        return g if inc > 0 else g[::-1]

    np.testing.assert_allclose(get(X, incx), H[0][0] * xs + H[0][1] * ys, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(get(Y, incy), H[1][0] * xs + H[1][1] * ys, rtol=1e-15, atol=1e-15)
    assert np.all(np.delete(X, np.s_[:: abs(incx)]) == 9.0)
    hx, hy = sb.drotm(xs.copy(), ys.copy(), param)
    np.testing.assert_allclose(get(X, incx), hx, rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(get(Y, incy), hy, rtol=1e-15, atol=1e-16)


# ---- complex GEMM (netlib ZGEMM / CGEMM) ----
def _cplx(seed, shape, dtype=np.complex128):
    """Deterministic complex input: real and imaginary parts from two streams."""
    return np.asfortranarray((uniform(seed, shape) + 1j * uniform(seed + 1000, shape)).astype(dtype))


def _op(X, t):
    return X if t == "N" else (X.T if t == "T" else X.conj().T)


@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
@pytest.mark.parametrize("prec", ["z", "c"])
def test_complex_gemm_vs_closed_form_and_host_blas(L, ta, tb, prec):
    """op(A) op(B) against numpy's product of the same conjugated /
    transposed arrays (the definition), and against the host BLAS; a form
    whose conjugation is wrong misses by the size of the imaginary parts."""
    dt = np.complex128 if prec == "z" else np.complex64
    eps = np.finfo(np.float64 if prec == "z" else np.float32).eps
    m, n, k = 7, 5, 9
    A = _cplx(40, (m, k) if ta == "N" else (k, m), dt)
    B = _cplx(42, (k, n) if tb == "N" else (n, k), dt)
    C0 = _cplx(44, (m, n), dt)
    alpha, beta = np.array([0.7 - 0.2j], dt), np.array([1.3 + 0.4j], dt)
    R = C0.copy(order="F")
    fn = L.ora_zgemm if prec == "z" else L.ora_cgemm
    fn(c(ta), c(tb), m, n, k, p(alpha), p(A), A.shape[0], p(B), B.shape[0], p(beta), p(R), m)
    oa, ob = _op(A.astype(np.complex128), ta), _op(B.astype(np.complex128), tb)
    want = alpha[0] * (oa @ ob) + beta[0] * C0
    gauge = abs(alpha[0]) * (np.abs(oa) @ np.abs(ob)) + abs(beta[0]) * np.abs(C0)
    assert np.all(np.abs(R - want) <= 8 * (k + 2) * eps * gauge)
    # the conjugation is not a no-op on these data
    other = {"C": "T", "T": "C", "N": "N"}
    if ta != "N" or tb != "N":
        wrong = alpha[0] * (_op(A, other[ta]) @ _op(B, other[tb])) + beta[0] * C0
        assert np.max(np.abs(R - wrong)) > 1e3 * 8 * (k + 2) * eps * np.max(gauge)
    hb = (sb.zgemm if prec == "z" else sb.cgemm)(alpha[0], A, B, beta[0], C0.copy(order="F"),
                                                  trans_a="NTC".index(ta), trans_b="NTC".index(tb))
    assert np.all(np.abs(R - hb) <= 16 * (k + 2) * eps * gauge)


@pytest.mark.parametrize("prec", ["z", "c"])
def test_complex_gemm_integer_exact_and_netlib_quick_paths(L, prec):
    dt = np.complex128 if prec == "z" else np.complex64
    fn = L.ora_zgemm if prec == "z" else L.ora_cgemm
    rng = np.random.default_rng(45)
    A = np.asfortranarray((rng.integers(-4, 5, (6, 8)) + 1j * rng.integers(-4, 5, (6, 8))).astype(dt))
    B = np.asfortranarray((rng.integers(-4, 5, (6, 5)) + 1j * rng.integers(-4, 5, (6, 5))).astype(dt))
    C = np.zeros((8, 5), dt, order="F")
    one, zero = np.array([1], dt), np.array([0], dt)
    fn(b"C", b"N", 8, 5, 6, p(one), p(A), 6, p(B), 6, p(zero), p(C), 8)
    assert np.array_equal(C, A.conj().T @ B)  # every partial sum an exact integer
    C = np.full((8, 5), np.nan, dt, order="F")
    fn(b"C", b"N", 8, 5, 6, p(one), p(A), 6, p(B), 6, p(zero), p(C), 8)
    assert np.all(np.isfinite(C))  # beta == 0: C is not read
    C = np.full((8, 5), np.nan, dt, order="F")
    fn(b"N", b"N", 8, 5, 6, p(zero), p(A), 8, p(B), 6, p(zero), p(C), 8)
    assert np.all(C == 0)  # alpha == beta == 0: exact zeros


# ---- the rest of the complex Level-3 in the oracle ----
def _tr(X, uplo):
    return np.triu(X) if uplo == "U" else np.tril(X)


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("kind", ["syrk", "herk", "syr2k", "her2k"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("nt", [True, False])
def test_complex_rank_k_vs_closed_form_and_host_blas(L, prec, kind, uplo, nt):
    dt, rdt = (np.complex128, np.float64) if prec == "z" else (np.complex64, np.float32)
    eps = np.finfo(rdt).eps
    herm = kind.startswith("he")
    tr = "N" if nt else ("C" if herm else "T")
    n, k = 7, 5
    shape = (n, k) if nt else (k, n)
    A, B, C0 = _cplx(46, shape, dt), _cplx(47, shape, dt), _cplx(48, (n, n), dt)
    al = np.array([0.8], rdt) if kind == "herk" else np.array([0.8 - 0.4j], dt)
    be = np.array([0.6], rdt) if herm else np.array([0.6 + 0.2j], dt)
    R = C0.copy(order="F")
    fn = getattr(L, f"ora_{prec}{kind}")
    if kind.endswith("2k"):
        fn(c(uplo), c(tr), n, k, p(al), p(A), shape[0], p(B), shape[0], p(be), p(R), n)
    else:
        fn(c(uplo), c(tr), n, k, p(al), p(A), shape[0], p(be), p(R), n)
    X = _op(A.astype(np.complex128), tr) if not nt else A.astype(np.complex128)
    Y = _op(B.astype(np.complex128), tr) if not nt else B.astype(np.complex128)
    a = complex(al[0])
    P = {"syrk": a * X @ X.T, "herk": a * X @ X.conj().T, "syr2k": a * (X @ Y.T + Y @ X.T),
         "her2k": a * X @ Y.conj().T + np.conj(a) * Y @ X.conj().T}[kind]
    Cin = C0.astype(np.complex128)
    if herm:
        Cin = Cin - 1j * np.diag(np.diag(Cin).imag)
    want = P + complex(be[0]) * Cin
    G = abs(a) * (np.abs(X) @ np.abs(Y if kind.endswith("2k") else X).T) * (2 if kind.endswith("2k") else 1) + abs(be[0]) * np.abs(C0)
    msk = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
    assert np.all(np.abs(R - want)[msk] <= 8 * (k + 2) * eps * G[msk])
    assert np.array_equal(R[~msk], C0[~msk])
    if herm:
        assert np.all(np.diag(R).imag == 0)
    hb = getattr(sb, prec + kind)
    lower = int(uplo == "L")
    tflag = 0 if nt else (2 if herm else 1)
    if kind.endswith("2k"):
        H = hb(complex(al[0]), A, B, beta=complex(be[0]) if not herm else float(be[0]), c=C0.copy(order="F"), trans=tflag, lower=lower)
    else:
        H = hb(al[0] if herm else complex(al[0]), A, beta=be[0] if herm else complex(be[0]), c=C0.copy(order="F"), trans=tflag, lower=lower)
    assert np.all(np.abs(R - H)[msk] <= 16 * (k + 2) * eps * G[msk])


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("herm", [0, 1])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_complex_symm_hemm_vs_closed_form(L, prec, herm, side, uplo):
    dt, rdt = (np.complex128, np.float64) if prec == "z" else (np.complex64, np.float32)
    eps = np.finfo(rdt).eps
    m, n = 6, 5
    na = m if side == "L" else n
    A, B, C0 = _cplx(49, (na, na), dt), _cplx(50, (m, n), dt), _cplx(51, (m, n), dt)
    al, be = np.array([0.7 + 0.1j], dt), np.array([-0.3 + 0.9j], dt)
    R = C0.copy(order="F")
    getattr(L, f"ora_{prec}{'hemm' if herm else 'symm'}")(c(side), c(uplo), m, n, p(al), p(A), na,
                                                           p(B), m, p(be), p(R), m)
    T = _tr(A.astype(np.complex128), uplo)
    off = T - np.diag(np.diag(T))
    F = T + (off.conj().T if herm else off.T)
    if herm:
        F[np.diag_indices(na)] = F.diagonal().real
    Bd = B.astype(np.complex128)
    want = complex(al[0]) * (F @ Bd if side == "L" else Bd @ F) + complex(be[0]) * C0
    G = abs(al[0]) * (np.abs(F) @ np.abs(Bd) if side == "L" else np.abs(Bd) @ np.abs(F)) + abs(be[0]) * np.abs(C0)
    assert np.all(np.abs(R - want) <= 8 * (na + 2) * eps * G)


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_complex_trmm_trsm_vs_closed_form(L, prec, side, uplo, ta, diag):
    """TRMM against the product of the expanded triangle; TRSM solves the
    system TRMM then undoes (a round trip to the original B)."""
    dt, rdt = (np.complex128, np.float64) if prec == "z" else (np.complex64, np.float32)
    eps = np.finfo(rdt).eps
    m, n = 6, 5
    na = m if side == "L" else n
    A = np.asfortranarray(_cplx(52, (na, na), dt) / na)
    A[np.diag_indices(na)] = 1.0 + 0.5j
    B0 = _cplx(53, (m, n), dt)
    al = np.array([0.6 - 0.8j], dt)
    T = _tr(A.astype(np.complex128), uplo)
    if diag == "U":
        T = T - np.diag(np.diag(T)) + np.eye(na)
    opT = _op(T, ta)
    Y = B0.copy(order="F")
    getattr(L, f"ora_{prec}trmm")(c(side), c(uplo), c(ta), c(diag), m, n, p(al), p(A), na, p(Y), m)
    Bd = B0.astype(np.complex128)
    want = complex(al[0]) * (opT @ Bd if side == "L" else Bd @ opT)
    G = abs(al[0]) * (np.abs(opT) @ np.abs(Bd) if side == "L" else np.abs(Bd) @ np.abs(opT))
    assert np.all(np.abs(Y - want) <= 8 * (na + 2) * eps * G)
    X = Y.copy(order="F")
    inv = np.array([1 / complex(al[0])], dt)
    getattr(L, f"ora_{prec}trsm")(c(side), c(uplo), c(ta), c(diag), m, n, p(inv), p(A), na, p(X), m)
    assert np.allclose(X, B0, rtol=0, atol=64 * na * eps)
    # and against the host BLAS
    H = getattr(sb, prec + "trsm")(complex(al[0]), A, B0.copy(order="F"), side=int(side == "R"),
                                     lower=int(uplo == "L"), trans_a="NTC".index(ta),
                                     diag=int(diag == "U"))
    Z = B0.copy(order="F")
    getattr(L, f"ora_{prec}trsm")(c(side), c(uplo), c(ta), c(diag), m, n, p(al), p(A), na, p(Z), m)
    assert np.allclose(Z, H, rtol=64 * na * eps, atol=64 * na * eps)
