"""GPU parity of the Level-1/2 routines against the CPU oracle.

Bars:
  * idamax: bit-exact (1-based index, first-index tie rule, netlib NaN rule);
  * integer-valued data: bit-exact for daxpy / ddot / dgemv / dger;
  * random data: |got - ref| <= 2 * n * eps * sum|terms| (the standard
    backward-error bound of a length-n inner product, with margin for the two
    different summation orders) for ddot / dgemv / dtrsv, one rounding for daxpy
    (the GPU fuses a*x+y), relative 4 * eps * sqrt(n) for dnrm2.
"""
import numpy as np
import pytest

import _oracle
from _oracle import c, p, uniform
from libgpublas_amd import library as gb_library

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _strided(vec, inc, fill=-7.0):
    """Lay `vec` out with stride |inc| (BLAS order: inc < 0 starts at the end)."""
    n = len(vec)
    a = abs(inc)
    buf = np.full(1 + (n - 1) * a, fill)
    if inc > 0:
        buf[::a] = vec
    else:
        buf[::a] = vec[::-1]
    return buf


@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (1000, 1, 1), (1001, 1, 1), (777, 2, 3),
                                         (500, -1, 2), (300, -3, -2), (8192, 1, 1)])
def test_daxpy(gpu, ora, n, incx, incy):
    x = _strided(uniform(1, (n,)), incx)
    y0 = _strided(uniform(2, (n,)), incy)
    y = y0.copy()
    r = y0.copy()
    gpu.daxpy(n, 0.3, x, incx, y, incy)
    ora.ora_daxpy(n, 0.3, p(x), incx, p(r), incy)
    # one rounding (fma) vs two (mul, add): within a couple of ulps of the terms
    assert np.all(np.abs(y - r) <= 2 * EPS * (np.abs(y0) + 0.3 * np.max(np.abs(x)) + np.abs(r)))


def test_daxpy_integer_exact(gpu, ora):
    rng = np.random.default_rng(1)
    x = rng.integers(-100, 100, 5000).astype(float)
    y = rng.integers(-100, 100, 5000).astype(float)
    r = y.copy()
    gpu.daxpy(5000, 3.0, x, 1, y, 1)
    ora.ora_daxpy(5000, 3.0, p(x), 1, p(r), 1)
    assert np.array_equal(y, r)


@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (5, 1, 1), (2048, 1, 1), (8192, 1, 1),
                                         (100000, 1, 1), (999, 3, -2), (4097, -1, -1)])
def test_ddot(gpu, ora, n, incx, incy):
    xv, yv = uniform(3, (n,)), uniform(4, (n,))
    x, y = _strided(xv, incx), _strided(yv, incy)
    got = gpu.ddot(n, x, incx, y, incy)
    ref = ora.ora_ddot(n, p(x), incx, p(y), incy)
    assert abs(got - ref) <= 2 * n * EPS * np.sum(np.abs(xv * yv))


@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (1000, 1, 1), (70001, 1, 1), (777, 2, -3),
                                         (500, -1, 2), (300, -3, -2)])
def test_dcopy_dswap_drot(gpu, ora, n, incx, incy):
    """Exact element moves (copy, swap) and one rotation per pair, against the
    netlib restatements; the fill between strided elements stays put."""
    x0 = _strided(uniform(5, (n,)), incx)
    y0 = _strided(uniform(6, (n,)), incy)
    x, y, rx, ry = x0.copy(), y0.copy(), x0.copy(), y0.copy()
    gpu.dcopy(n, x, incx, y, incy)
    ora.ora_dcopy(n, p(rx), incx, p(ry), incy)
    assert np.array_equal(y, ry) and np.array_equal(x, x0)
    x, y, rx, ry = x0.copy(), y0.copy(), x0.copy(), y0.copy()
    gpu.dswap(n, x, incx, y, incy)
    ora.ora_dswap(n, p(rx), incx, p(ry), incy)
    assert np.array_equal(x, rx) and np.array_equal(y, ry)
    x, y, rx, ry = x0.copy(), y0.copy(), x0.copy(), y0.copy()
    gpu.drot(n, x, incx, y, incy, 0.6, 0.8)
    ora.ora_drot(n, p(rx), incx, p(ry), incy, 0.6, 0.8)
    bound = 2 * EPS * (np.abs(x0).max() + np.abs(y0).max())
    assert np.max(np.abs(x - rx)) <= bound and np.max(np.abs(y - ry)) <= bound


@pytest.mark.parametrize("n,incx", [(1, 1), (1000, 1), (8192, 1), (70001, 1), (777, 3), (50, -1)])
def test_dscal_dasum(gpu, ora, n, incx):
    xv = uniform(7, (n,))
    x0 = _strided(xv, incx)
    x, r = x0.copy(), x0.copy()
    gpu.dscal(n, -2.5, x, incx)
    ora.ora_dscal(n, -2.5, p(r), incx)
    assert np.array_equal(x, r)  # one multiplication per element; incx < 0: no-op
    got = gpu.dasum(n, x0, incx)
    ref = ora.ora_dasum(n, p(x0), incx)
    if incx <= 0:
        assert got == 0.0 and ref == 0.0
    else:
        assert abs(got - ref) <= 2 * n * EPS * np.sum(np.abs(xv))


def test_dasum_integer_exact_and_repeatable(gpu):
    rng = np.random.default_rng(9)
    x = rng.integers(-1000, 1000, 200001).astype(float)
    a = gpu.dasum(len(x), x, 1)
    assert a == float(np.sum(np.abs(x)))
    assert all(gpu.dasum(len(x), x, 1) == a for _ in range(3))


def test_ddot_integer_exact_and_repeatable(gpu):
    rng = np.random.default_rng(2)
    x = rng.integers(-1000, 1000, 300000).astype(float)
    y = rng.integers(-1000, 1000, 300000).astype(float)
    got = gpu.ddot(300000, x, 1, y, 1)
    assert got == float(np.dot(x.astype(np.int64), y.astype(np.int64)))
    z, w = uniform(5, (123457,)), uniform(6, (123457,))
    vals = {gpu.ddot(123457, z, 1, w, 1) for _ in range(5)}
    assert len(vals) == 1  # fixed reduction order: bitwise repeatable


@pytest.mark.parametrize("n,incx", [(2, 1), (1000, 1), (8192, 1), (70001, 1), (500, 2), (400, -3)])
def test_dnrm2(gpu, ora, n, incx):
    xv = uniform(7, (n,))
    x = _strided(xv, incx)
    got = gpu.dnrm2(n, x, incx)
    ref = ora.ora_dnrm2(n, p(x), incx)
    assert abs(got - ref) <= 4 * EPS * np.sqrt(n) * ref


@pytest.mark.parametrize("scale", [1e300, 1e-300, 1e-160, 1e160, 1.0])
def test_dnrm2_no_overflow_underflow(gpu, scale):
    """Closed form: ||c * 1_n|| = |c| sqrt(n), at magnitudes where a plain
    sum of squares would overflow or flush to zero."""
    n = 10000
    x = np.full(n, scale)
    got = gpu.dnrm2(n, x, 1)
    assert got == pytest.approx(scale * np.sqrt(n), rel=1e-14)


def test_dnrm2_special(gpu):
    x = np.array([3.0, -4.0] + [0.0] * 1000)
    assert gpu.dnrm2(1002, x, 1) == 5.0
    x = uniform(8, (3000,))
    x[1234] = np.inf
    assert gpu.dnrm2(3000, x, 1) == np.inf
    x[1234] = np.nan
    assert np.isnan(gpu.dnrm2(3000, x, 1))
    assert gpu.dnrm2(0, x, 1) == 0.0
    # incx == 0 repeats x(1): |x1| sqrt(n)
    y = np.array([2.0])
    assert gpu.dnrm2(1000, y, 0) == pytest.approx(2.0 * np.sqrt(1000), rel=1e-15)


@pytest.mark.parametrize("n,incx", [(1, 1), (2, 1), (1000, 1), (8192, 1), (100003, 1), (777, 3)])
def test_idamax_random(gpu, ora, n, incx):
    x = _strided(uniform(9, (n,)), incx)
    assert gpu.idamax(n, x, incx) == ora.ora_idamax(n, p(x), incx)


def test_idamax_ties_first_index(gpu, ora):
    n = 50000
    x = uniform(10, (n,)) * 0.5
    for pos in (40000, 7, 12345, 49999):
        x[pos] = 0.9 if pos % 2 else -0.9  # equal |x|, both signs
    want = 8  # 1-based index of position 7
    assert ora.ora_idamax(n, p(x), 1) == want
    assert gpu.idamax(n, x, 1) == want
    y = np.full(9000, 2.5)
    y[::2] *= -1
    assert gpu.idamax(9000, y, 1) == 1


def test_idamax_nan_rule(gpu, ora):
    """netlib: strict '>' never moves to a NaN; a NaN in x(1) gives 1."""
    n = 5000
    x = uniform(11, (n,))
    x[100] = np.nan
    x[4000] = 3.0
    assert gpu.idamax(n, x, 1) == ora.ora_idamax(n, p(x), 1) == 4001
    x[0] = np.nan
    assert gpu.idamax(n, x, 1) == ora.ora_idamax(n, p(x), 1) == 1
    assert gpu.idamax(0, x, 1) == 0
    assert gpu.idamax(10, x, 0) == 0
    assert gpu.idamax(10, x, -1) == 0


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,n,incx,incy", [(1, 1, 1, 1), (100, 70, 1, 1), (513, 1025, 1, 1),
                                           (300, 200, 2, -1), (129, 4000, -2, 3), (5000, 37, 1, 1)])
def test_dgemv(gpu, ora, trans, m, n, incx, incy):
    lda = m + 1
    A = np.full((lda, n), -9.0, order="F")
    A[:m, :] = uniform(12, (m, n))
    lenx, leny = (n, m) if trans == "N" else (m, n)
    xv, yv = uniform(13, (lenx,)), uniform(14, (leny,))
    x, y0 = _strided(xv, incx), _strided(yv, incy)
    y, r = y0.copy(), y0.copy()
    gpu.dgemv(trans, m, n, 0.8, A, lda, x, incx, 0.6, y, incy)
    ora.ora_dgemv(c(trans), m, n, 0.8, p(A), lda, p(x), incx, 0.6, p(r), incy)
    absA = np.abs(A[:m, :]) if trans == "N" else np.abs(A[:m, :]).T
    G = 0.8 * absA @ np.abs(xv) + 0.6 * np.abs(yv)
    k = n if trans == "N" else m
    got = y[:: abs(incy)] if incy > 0 else y[:: abs(incy)][::-1]
    ref = r[:: abs(incy)] if incy > 0 else r[:: abs(incy)][::-1]
    assert np.all(np.abs(got - ref) <= 2 * (k + 2) * EPS * G)
    assert np.array_equal(y[np.arange(len(y)) % abs(incy) != 0], y0[np.arange(len(y)) % abs(incy) != 0])


def test_dgemv_integer_exact(gpu, ora):
    rng = np.random.default_rng(4)
    m, n = 700, 900
    A = np.asfortranarray(rng.integers(-5, 6, size=(m, n)).astype(float))
    x = rng.integers(-5, 6, n).astype(float)
    y = np.zeros(m)
    r = np.zeros(m)
    gpu.dgemv("N", m, n, 1.0, A, m, x, 1, 0.0, y, 1)
    ora.ora_dgemv(b"N", m, n, 1.0, p(A), m, p(x), 1, 0.0, p(r), 1)
    assert np.array_equal(y, r)


@pytest.mark.parametrize("m,n,incx,incy", [(1, 1, 1, 1), (300, 200, 1, 1), (257, 33, -2, 3)])
def test_dger(gpu, ora, m, n, incx, incy):
    lda = m + 2
    A0 = np.full((lda, n), 5.0, order="F")
    A0[:m, :] = uniform(15, (m, n))
    x, y = _strided(uniform(16, (m,)), incx), _strided(uniform(17, (n,)), incy)
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dger(m, n, 1.7, x, incx, y, incy, A, lda)
    ora.ora_dger(m, n, 1.7, p(x), incx, p(y), incy, p(R), lda)
    np.testing.assert_allclose(A, R, rtol=4 * EPS, atol=4 * EPS)
    assert np.array_equal(A[m:, :], A0[m:, :])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,incx", [(1, 1), (50, 1), (300, 1), (200, -2)])
def test_dtrsv(gpu, ora, uplo, trans, diag, n, incx):
    A = uniform(18, (n, n))
    if diag == "N":
        A[np.diag_indices(n)] = n + 1.0
    else:
        A *= 1.0 / n
        A[np.diag_indices(n)] = np.nan  # a unit diagonal is not read
    A[_other_triangle(n, uplo)] = np.nan  # nor is the other triangle
    A = np.asfortranarray(A)
    x0 = _strided(uniform(19, (n,)), incx)
    x, r = x0.copy(), x0.copy()
    gpu.dtrsv(uplo, trans, diag, n, A, n, x, incx)
    ora.ora_dtrsv(c(uplo), c(trans), c(diag), n, p(A), n, p(r), incx)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14)


def _other_triangle(n, uplo):
    return np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (2, 1, 1), (127, 1, 1), (128, 1, 1),
                                         (129, 2, -3), (1000, -1, 2), (4096, 1, 1), (4099, 1, 1)])
def test_dsymv(gpu, ora, uplo, n, incx, incy):
    """Only the uplo triangle is read (the other holds NaN); bound: 2n eps
    times the sum of |terms| per element.  Even leading dimensions take the
    16-B load path on interior tiles, odd ones (n = 127, 129, 4099) the
    per-element path everywhere."""
    A = uniform(31, (n, n))
    A[_other_triangle(n, uplo)] = np.nan
    x = _strided(uniform(32, (n,)), incx)
    y0 = _strided(uniform(33, (n,)), incy)
    y, r = y0.copy(), y0.copy()
    gpu.dsymv(uplo, n, 0.75, A, n, x, incx, -0.5, y, incy)
    ora.ora_dsymv(c(uplo), n, 0.75, p(A), n, p(x), incx, -0.5, p(r), incy)
    S = np.where(np.isnan(A), 0.0, A)
    S = S + S.T - np.diag(np.diag(S))
    xs = x[:: abs(incx)][:n] if incx > 0 else x[:: abs(incx)][:n][::-1]
    ys = y0[:: abs(incy)][:n] if incy > 0 else y0[:: abs(incy)][:n][::-1]
    mag = 0.75 * (np.abs(S) @ np.abs(xs)) + 0.5 * np.abs(ys)
    got = y[:: abs(incy)][:n] if incy > 0 else y[:: abs(incy)][:n][::-1]
    ref = r[:: abs(incy)][:n] if incy > 0 else r[:: abs(incy)][:n][::-1]
    assert np.all(np.abs(got - ref) <= 2 * n * EPS * mag + 1e-300)
    gaps = np.ones(len(y), bool)
    gaps[:: abs(incy)][:n] = False
    assert np.array_equal(y[gaps], y0[gaps])  # stride gaps untouched


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dsymv_integer_exact(gpu, ora, uplo):
    """Integer data with exactly representable sums: bit-exact, any order."""
    n = 1537
    rng = np.random.default_rng(34)
    A = np.asfortranarray(rng.integers(-8, 9, (n, n)).astype(float))
    x = rng.integers(-8, 9, n).astype(float)
    y0 = rng.integers(-8, 9, n).astype(float)
    y, r = y0.copy(), y0.copy()
    gpu.dsymv(uplo, n, 2.0, A, n, x, 1, 3.0, y, 1)
    ora.ora_dsymv(c(uplo), n, 2.0, p(A), n, p(x), 1, 3.0, p(r), 1)
    assert np.array_equal(y, r)


def test_dsymv_alpha_beta_special(gpu, ora):
    n = 300
    A = uniform(35, (n, n))
    x = uniform(36, (n,))
    y = np.full(n, np.nan)
    gpu.dsymv("L", n, 1.0, A, n, x, 1, 0.0, y, 1)  # beta = 0: y overwritten
    r = np.full(n, np.nan)
    ora.ora_dsymv(b"L", n, 1.0, p(A), n, p(x), 1, 0.0, p(r), 1)
    assert np.all(np.isfinite(y))
    np.testing.assert_allclose(y, r, rtol=1e-12, atol=1e-12)
    y0 = uniform(37, (n,))
    y = y0.copy()
    gpu.dsymv("U", n, 0.0, A, n, x, 1, -2.0, y, 1)  # alpha = 0: scale only
    assert np.array_equal(y, -2.0 * y0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,incx", [(1, 1), (255, 1), (256, 2), (1000, -3), (3001, 1)])
def test_dsyr(gpu, ora, uplo, n, incx):
    A0 = uniform(38, (n, n))
    x = _strided(uniform(39, (n,)), incx)
    x[0] = 0.0  # one zero entry: its column is left as it is
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dsyr(uplo, n, -0.625, x, incx, A, n)
    ora.ora_dsyr(c(uplo), n, -0.625, p(x), incx, p(R), n)
    other = _other_triangle(n, uplo)
    assert np.array_equal(A[other], A0[other])  # never touched
    # one fused rounding on the GPU against netlib's two
    np.testing.assert_allclose(A, R, rtol=2 * EPS, atol=2 * EPS)


def test_dsyr_integer_exact(gpu, ora):
    n = 777
    rng = np.random.default_rng(40)
    A0 = np.asfortranarray(rng.integers(-50, 51, (n, n)).astype(float))
    x = rng.integers(-9, 10, n).astype(float)
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dsyr("U", n, 3.0, x, 1, A, n)
    ora.ora_dsyr(b"U", n, 3.0, p(x), 1, p(R), n)
    assert np.array_equal(A, R)


def _pack(A, uplo):
    n = A.shape[0]
    if uplo == "U":
        return np.concatenate([A[: j + 1, j] for j in range(n)])
    return np.concatenate([A[j:, j] for j in range(n)])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (255, 1, 1), (513, 2, -3), (2048, 1, 1),
                                         (3001, -1, 2)])
def test_dsyr2(gpu, ora, uplo, n, incx, incy):
    A0 = uniform(41, (n, n))
    x = _strided(uniform(42, (n,)), incx)
    y = _strided(uniform(43, (n,)), incy)
    x[0] = y[0] = 0.0  # a column with x(j) = y(j) = 0 is left as it is
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dsyr2(uplo, n, 0.375, x, incx, y, incy, A, n)
    ora.ora_dsyr2(c(uplo), n, 0.375, p(x), incx, p(y), incy, p(R), n)
    other = _other_triangle(n, uplo)
    assert np.array_equal(A[other], A0[other])
    np.testing.assert_allclose(A, R, rtol=4 * EPS, atol=4 * EPS)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (300, 1, 1), (1000, -2, 3)])
def test_dspr_dspr2(gpu, ora, uplo, n, incx, incy):
    A0 = uniform(44, (n, n))
    AP0 = _pack(A0, uplo)
    x = _strided(uniform(45, (n,)), incx)
    y = _strided(uniform(46, (n,)), incy)
    P, R = AP0.copy(), AP0.copy()
    gpu.dspr(uplo, n, -0.5, x, incx, P)
    ora.ora_dspr(c(uplo), n, -0.5, p(x), incx, p(R))
    np.testing.assert_allclose(P, R, rtol=2 * EPS, atol=2 * EPS)
    P, R = AP0.copy(), AP0.copy()
    gpu.dspr2(uplo, n, 0.75, x, incx, y, incy, P)
    ora.ora_dspr2(c(uplo), n, 0.75, p(x), incx, p(y), incy, p(R))
    np.testing.assert_allclose(P, R, rtol=4 * EPS, atol=4 * EPS)


def test_rank2_updates_integer_exact(gpu, ora):
    n = 700
    rng = np.random.default_rng(47)
    A0 = np.asfortranarray(rng.integers(-50, 51, (n, n)).astype(float))
    x = rng.integers(-9, 10, n).astype(float)
    y = rng.integers(-9, 10, n).astype(float)
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dsyr2("L", n, 2.0, x, 1, y, 1, A, n)
    ora.ora_dsyr2(b"L", n, 2.0, p(x), 1, p(y), 1, p(R), n)
    assert np.array_equal(A, R)
    AP0 = _pack(A0, "U")
    P, Q = AP0.copy(), AP0.copy()
    gpu.dspr2("U", n, 3.0, x, 1, y, 1, P)
    ora.ora_dspr2(b"U", n, 3.0, p(x), 1, p(y), 1, p(Q))
    assert np.array_equal(P, Q)


def _tri_mag(A, uplo, trans, diag, xs):
    """sum |terms| per element of op(T) x (T the uplo triangle, unit diagonal
    if diag == 'U'): the scale of the inner-product error bound."""
    T = np.tril(A) if uplo == "L" else np.triu(A)
    T = np.where(np.isnan(T), 0.0, T)
    if diag == "U":
        T[np.diag_indices(len(xs))] = 1.0
    if trans != "N":
        T = T.T
    return np.abs(T) @ np.abs(xs)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,incx,pad", [(1, 1, 0), (50, 1, 3), (257, -2, 0), (1000, 1, 1),
                                        (3000, 3, 0)])
def test_dtrmv(gpu, ora, uplo, trans, diag, n, incx, pad):
    """Only the uplo triangle is read (the other holds NaN, and so does a unit
    diagonal); bound 2 n eps sum|terms|; stride gaps and padding rows kept."""
    lda = n + pad
    A = np.full((lda, n), -5.0, order="F")
    A[:n] = uniform(91, (n, n))
    A[:n][_other_triangle(n, uplo)] = np.nan
    if diag == "U":
        A[np.arange(n), np.arange(n)] = np.nan
    x0 = _strided(uniform(92, (n,)), incx)
    x, r = x0.copy(), x0.copy()
    gpu.dtrmv(uplo, trans, diag, n, A, lda, x, incx)
    ora.ora_dtrmv(c(uplo), c(trans), c(diag), n, p(A), lda, p(r), incx)
    mag = _tri_mag(A[:n], uplo, trans, diag, _strided_get(x0, n, incx))
    got, ref = _strided_get(x, n, incx), _strided_get(r, n, incx)
    assert np.all(np.abs(got - ref) <= 2 * n * EPS * mag + 1e-300)
    gaps = np.ones(len(x), bool)
    gaps[:: abs(incx)][:n] = False
    assert np.array_equal(x[gaps], x0[gaps])


def _strided_get(buf, n, inc):
    v = buf[:: abs(inc)][:n]
    return v if inc > 0 else v[::-1]


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtrmv_integer_exact(gpu, ora, uplo, trans):
    n = 1300
    rng = np.random.default_rng(93)
    A = np.asfortranarray(rng.integers(-9, 10, (n, n)).astype(float))
    x0 = rng.integers(-9, 10, n).astype(float)
    for diag in "NU":
        x, r = x0.copy(), x0.copy()
        gpu.dtrmv(uplo, trans, diag, n, A, n, x, 1)
        ora.ora_dtrmv(c(uplo), c(trans), c(diag), n, p(A), n, p(r), 1)
        assert np.array_equal(x, r)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,incx", [(1, 1), (64, 1), (300, -2), (2049, 1)])
def test_dtpmv_dtpsv(gpu, ora, uplo, trans, diag, n, incx):
    """Packed triangular product (bound 2 n eps sum|terms|) and solve
    (diagonally dominant, relative 1e-12); a unit diagonal's packed slots hold
    NaN and must not be read."""
    A = uniform(94, (n, n)) * (1.0 / n)
    A[np.diag_indices(n)] = np.nan if diag == "U" else 2.0
    AP = _pack(A, uplo)
    x0 = _strided(uniform(95, (n,)), incx)
    x, r = x0.copy(), x0.copy()
    gpu.dtpmv(uplo, trans, diag, n, AP, x, incx)
    ora.ora_dtpmv(c(uplo), c(trans), c(diag), n, p(AP), p(r), incx)
    mag = _tri_mag(A, uplo, trans, diag, _strided_get(x0, n, incx))
    assert np.all(np.abs(_strided_get(x, n, incx) - _strided_get(r, n, incx))
                  <= 2 * n * EPS * mag + 1e-300)
    x, r = x0.copy(), x0.copy()
    gpu.dtpsv(uplo, trans, diag, n, AP, x, incx)
    ora.ora_dtpsv(c(uplo), c(trans), c(diag), n, p(AP), p(r), incx)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14)
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (129, 2, -3), (1000, -1, 2), (4099, 1, 1),
                                         (1024, 1, 1), (8192, 1, -1)])
def test_dspmv(gpu, ora, uplo, n, incx, incy):
    A = uniform(96, (n, n))
    AP = _pack(A, uplo)
    x = _strided(uniform(97, (n,)), incx)
    y0 = _strided(uniform(98, (n,)), incy)
    y, r = y0.copy(), y0.copy()
    gpu.dspmv(uplo, n, 0.75, AP, x, incx, -0.5, y, incy)
    ora.ora_dspmv(c(uplo), n, 0.75, p(AP), p(x), incx, -0.5, p(r), incy)
    T = np.tril(A) if uplo == "L" else np.triu(A)
    S = T + T.T - np.diag(np.diag(A))
    mag = 0.75 * (np.abs(S) @ np.abs(_strided_get(x, n, incx))) + 0.5 * np.abs(_strided_get(y0, n, incy))
    assert np.all(np.abs(_strided_get(y, n, incy) - _strided_get(r, n, incy)) <= 2 * n * EPS * mag + 1e-300)
    gaps = np.ones(len(y), bool)
    gaps[:: abs(incy)][:n] = False
    assert np.array_equal(y[gaps], y0[gaps])


def test_dspmv_alpha_zero_and_integer_exact(gpu, ora):
    """Integer entries: every partial sum is exact, so the tile kernel's
    summation order cannot show; n = 900 has ragged edge tiles, n = 1280 only
    whole 128 x 128 tiles (the diagonal tiles load the packed neighbours of
    the triangle and select them away)."""
    rng = np.random.default_rng(99)
    for n, uplo in ((900, "L"), (1280, "L"), (1280, "U")):
        A = rng.integers(-9, 10, (n, n)).astype(float)
        AP = _pack(A, uplo)
        x = rng.integers(-9, 10, n).astype(float)
        y0 = rng.integers(-9, 10, n).astype(float)
        y, r = y0.copy(), y0.copy()
        gpu.dspmv(uplo, n, 2.0, AP, x, 1, 3.0, y, 1)
        ora.ora_dspmv(c(uplo), n, 2.0, p(AP), p(x), 1, 3.0, p(r), 1)
        assert np.array_equal(y, r), (n, uplo)
    n = 900
    AP = _pack(rng.integers(-9, 10, (n, n)).astype(float), "U")
    x = rng.integers(-9, 10, n).astype(float)
    y0 = rng.integers(-9, 10, n).astype(float)
    y = y0.copy()
    gpu.dspmv("U", n, 0.0, AP, x, 1, 0.5, y, 1)  # alpha = 0: y := beta y, AP unread
    assert np.array_equal(y, 0.5 * y0)


def _band_array(m, n, kl, ku, lda, seed, fill=np.nan):
    """Band storage (lda x n) with random band entries; the corners of the
    band array outside the matrix and the rows beyond kl + ku hold `fill`."""
    AB = np.full((lda, n), fill, order="F")
    vals = uniform(seed, (kl + ku + 1, n))
    for j in range(n):
        i0, i1 = max(0, j - ku), min(m, j + kl + 1)
        if i1 > i0:
            AB[ku + i0 - j: ku + i1 - j, j] = vals[ku + i0 - j: ku + i1 - j, j]
    return AB


def _gb_abs_product(AB, m, n, kl, ku, v, trans):
    """sum |A(i,j)| |v| per output element (the error-bound scale)."""
    A = np.where(np.isnan(AB), 0.0, np.abs(AB))
    out = np.zeros(n if trans else m)
    for r in range(kl + ku + 1):
        j = np.arange(n)
        i = j - ku + r
        ok = (i >= 0) & (i < m)
        if trans:
            np.add.at(out, j[ok], A[r, j[ok]] * np.abs(v[i[ok]]))
        else:
            np.add.at(out, i[ok], A[r, j[ok]] * np.abs(v[j[ok]]))
    return out


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,n,kl,ku,incx,incy", [
    (1, 1, 0, 0, 1, 1), (2000, 2000, 5, 5, 1, 1),       # one chunk: direct write
    (5000, 3000, 0, 40, 2, -1), (3000, 5000, 70, 0, -3, 2),
    (257, 300, 300, 2, 1, 1),                            # kl >= m
    (300000, 300000, 8, 8, 1, 1),                        # two chunks per workgroup
    (100000, 90000, 16, 16, 1, 1)])                      # chunks split, partial fold
def test_dgbmv(gpu, ora, trans, m, n, kl, ku, incx, incy):
    """Band corners and padding rows hold NaN (never read); bound
    2 w eps sum|terms|, w = kl + ku + 1; stride gaps of y untouched."""
    lda = kl + ku + 2
    AB = _band_array(m, n, kl, ku, lda, 101)
    lx, ly = (m, n) if trans == "T" else (n, m)
    x = _strided(uniform(102, (lx,)), incx)
    y0 = _strided(uniform(103, (ly,)), incy)
    y, r = y0.copy(), y0.copy()
    gpu.dgbmv(trans, m, n, kl, ku, 0.5, AB, lda, x, incx, -1.5, y, incy)
    ora.ora_dgbmv(c(trans), m, n, kl, ku, 0.5, p(AB), lda, p(x), incx, -1.5, p(r), incy)
    mag = 0.5 * _gb_abs_product(AB, m, n, kl, ku, _strided_get(x, lx, incx), trans == "T") \
        + 1.5 * np.abs(_strided_get(y0, ly, incy))
    w = kl + ku + 1
    err = np.abs(_strided_get(y, ly, incy) - _strided_get(r, ly, incy))
    assert np.all(err <= 2 * w * EPS * mag + 1e-300)
    gaps = np.ones(len(y), bool)
    gaps[:: abs(incy)][:ly] = False
    assert np.array_equal(y[gaps], y0[gaps])


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgbmv_integer_exact_and_alpha_zero(gpu, ora, trans):
    m, n, kl, ku = 70000, 65000, 20, 11
    rng = np.random.default_rng(104)
    AB = np.asfortranarray(rng.integers(-9, 10, (kl + ku + 1, n)).astype(float))
    lx, ly = (m, n) if trans == "T" else (n, m)
    x = rng.integers(-9, 10, lx).astype(float)
    y0 = rng.integers(-9, 10, ly).astype(float)
    y, r = y0.copy(), y0.copy()
    gpu.dgbmv(trans, m, n, kl, ku, 2.0, AB, kl + ku + 1, x, 1, 3.0, y, 1)
    ora.ora_dgbmv(c(trans), m, n, kl, ku, 2.0, p(AB), kl + ku + 1, p(x), 1, 3.0, p(r), 1)
    assert np.array_equal(y, r)
    y = y0.copy()
    gpu.dgbmv(trans, m, n, kl, ku, 0.0, AB, kl + ku + 1, x, 1, 0.5, y, 1)
    assert np.array_equal(y, 0.5 * y0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,k,incx,incy", [(1, 0, 1, 1), (1000, 7, 1, 1), (5000, 100, -2, 3),
                                           (70000, 20, 1, 1), (300, 400, 1, -1)])
def test_dsbmv(gpu, ora, uplo, n, k, incx, incy):
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = _band_array(n, n, kl, ku, k + 1, 105)
    x = _strided(uniform(106, (n,)), incx)
    y0 = _strided(uniform(107, (n,)), incy)
    y, r = y0.copy(), y0.copy()
    gpu.dsbmv(uplo, n, k, 0.75, AB, k + 1, x, incx, 0.25, y, incy)
    ora.ora_dsbmv(c(uplo), n, k, 0.75, p(AB), k + 1, p(x), incx, 0.25, p(r), incy)
    xs = _strided_get(x, n, incx)
    mag = 0.75 * (_gb_abs_product(AB, n, n, kl, ku, xs, False)
                  + _gb_abs_product(AB, n, n, kl, ku, xs, True)) \
        + 0.25 * np.abs(_strided_get(y0, n, incy))
    err = np.abs(_strided_get(y, n, incy) - _strided_get(r, n, incy))
    assert np.all(err <= 4 * (k + 1) * EPS * mag + 1e-300)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,k,incx", [(1, 0, 1), (1000, 5, 1), (3000, 64, -2), (20000, 200, 1)])
def test_dtbmv(gpu, ora, uplo, trans, diag, n, k, incx):
    """A unit diagonal's band row holds NaN (not read); bound 2 (k+1) eps
    sum|terms|."""
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = _band_array(n, n, kl, ku, k + 1, 108)
    if diag == "U":
        AB[ku, :] = np.nan
    x0 = _strided(uniform(109, (n,)), incx)
    x, r = x0.copy(), x0.copy()
    gpu.dtbmv(uplo, trans, diag, n, k, AB, k + 1, x, incx)
    ora.ora_dtbmv(c(uplo), c(trans), c(diag), n, k, p(AB), k + 1, p(r), incx)
    A1 = AB.copy(order="F")
    if diag == "U":
        A1[ku, :] = 1.0
    mag = _gb_abs_product(A1, n, n, kl, ku, _strided_get(x0, n, incx), trans == "T")
    err = np.abs(_strided_get(x, n, incx) - _strided_get(r, n, incx))
    assert np.all(err <= 2 * (k + 1) * EPS * mag + 1e-300)
    gaps = np.ones(len(x), bool)
    gaps[:: abs(incx)][:n] = False
    assert np.array_equal(x[gaps], x0[gaps])


TBSV_CASES = [(1, 0, 1), (100, 1, 1), (5000, 5, 1), (3001, 63, -2),    # register window
              (3000, 64, 1), (4000, 100, 3), (1200, 200, 1),           # LDS window, prefetched
              (1100, 300, 2), (2500, 700, 1),
              (3000, 1500, 1)]                                          # LDS window, k >= 1024


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k,incx", TBSV_CASES)
def test_dtbsv_integer_exact(gpu, ora, uplo, trans, n, k, incx):
    """b = T x with small integers and a diagonal of ones (unit or stored):
    every partial sum of the solve is an integer, so the solution must come
    back exactly; the oracle's solve too.  Band corners hold NaN."""
    rng = np.random.default_rng(110 + k)
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = np.full((k + 1, n), np.nan, order="F")
    for j in range(n):
        i0, i1 = max(0, j - ku), min(n, j + kl + 1)
        AB[ku + i0 - j: ku + i1 - j, j] = rng.integers(-1, 2, i1 - i0)
    for diag in "NU":
        A = AB.copy(order="F")
        A[ku, :] = 1.0 if diag == "N" else np.nan
        xt = rng.integers(-3, 4, n).astype(float)
        # b = op(T) xt through the oracle's DTBMV (exact on these integers)
        b0 = _strided(xt, incx)
        ora.ora_dtbmv(c(uplo), c(trans), c(diag), n, k, p(A), k + 1, p(b0), incx)
        assert np.abs(b0).max() < 2 ** 40
        x, r = b0.copy(), b0.copy()
        gpu.dtbsv(uplo, trans, diag, n, k, A, k + 1, x, incx)
        ora.ora_dtbsv(c(uplo), c(trans), c(diag), n, k, p(A), k + 1, p(r), incx)
        assert np.array_equal(_strided_get(r, n, incx), xt)
        assert np.array_equal(_strided_get(x, n, incx), xt), (diag, np.abs(_strided_get(x, n, incx) - xt).max())


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(4000, 10), (3000, 90), (2000, 1100)])
def test_dtbsv_random(gpu, ora, uplo, trans, n, k):
    """Diagonally dominant random band (diagonal in [2, 3], off-diagonal
    scaled by 1 / (k + 1)): relative agreement with the oracle to 1e-12."""
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = _band_array(n, n, kl, ku, k + 1, 111) * (1.0 / (k + 1))
    AB[ku, :] = 2.5 + 0.5 * uniform(112, (n,))
    x0 = uniform(113, (n,))
    x, r = x0.copy(), x0.copy()
    gpu.dtbsv(uplo, trans, "N", n, k, AB, k + 1, x, 1)
    ora.ora_dtbsv(c(uplo), c(trans), b"N", n, k, p(AB), k + 1, p(r), 1)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14)


def _tb_dominant(n, k, uplo, lda, seed, unit):
    """Triangular band (k super- or sub-diagonals) in band storage with
    off-diagonal entries in [-1, 1] / (2 (k + 1)) and a diagonal in [1.5, 2.5]
    (or a unit diagonal, whose stored row holds NaN): every unit response of a
    chunk stays far inside [-1, 1].  Band-array corners outside the matrix
    hold NaN."""
    ku = k if uplo == "U" else 0
    AB = np.full((lda, n), np.nan, order="F")
    AB[: k + 1, :] = uniform(seed, (k + 1, n)) * (0.5 / (k + 1))
    AB[ku, :] = np.nan if unit else 2.0 + 0.5 * uniform(seed + 1, (n,))
    for j in range(min(k, n)):
        if uplo == "U":
            AB[: k - j, j] = np.nan             # rows above the matrix
        else:
            AB[j + 1: k + 1, n - 1 - j] = np.nan  # rows below the matrix
    return AB


TB_CHUNK_CASES = [(2048, 0, 1), (2048, 1, 1), (4099, 7, -2), (100_003, 15, 1), (100_003, 16, 3),
                  (1 << 20, 32, 1), (70_001, 32, -1), (65_537, 33, 1), (50_000, 63, 2)]


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k,incx", TB_CHUNK_CASES)
def test_dtbsv_chunked(gpu, ora, uplo, trans, n, k, incx):
    """k <= 63, n >= 2048 runs chunked: per-chunk responses, the incoming
    unknowns propagated chunk to chunk, per-chunk re-solves.  Against the
    oracle's sequential solve to 1e-12 relative (diagonally dominant bands,
    stored and unit diagonal); the first 64 unknowns in solve order (chunk 0 is the
    whole-length recurrence's own first steps) bit-identical to the one-wave
    recurrence.  (With responses this small the propagated unknowns often
    come out bit-identical to the one-wave ones.)"""
    lib = gb_library()
    assert lib.gpublas_dev_tbsv_chunked(-1) == 1
    for diag in "NU":
        AB = _tb_dominant(n, k, uplo, k + 2, 120 + k, diag == "U")
        b = _strided(uniform(121 + k, (n,)), incx)
        x, r, q = b.copy(), b.copy(), b.copy()
        gpu.dtbsv(uplo, trans, diag, n, k, AB, k + 2, x, incx)
        ora.ora_dtbsv(c(uplo), c(trans), c(diag), n, k, p(AB), k + 2, p(r), incx)
        xs, rs = _strided_get(x, n, incx), _strided_get(r, n, incx)
        np.testing.assert_allclose(xs, rs, rtol=1e-12, atol=1e-14 * np.abs(rs).max(),
                                   err_msg=str((uplo, trans, diag, n, k, incx)))
        gaps = np.ones(len(b), bool)
        gaps[:: abs(incx)][:n] = False
        assert np.array_equal(x[gaps], b[gaps])
        lib.gpublas_dev_tbsv_chunked(0)
        try:
            gpu.dtbsv(uplo, trans, diag, n, k, AB, k + 2, q, incx)
        finally:
            lib.gpublas_dev_tbsv_chunked(1)
        qs = _strided_get(q, n, incx)
        first = slice(0, 64) if (uplo == "L") == (trans == "N") else slice(n - 64, n)  # solve order
        assert np.array_equal(xs[first], qs[first])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtbsv_chunked_slow_decay(gpu, ora, uplo, trans):
    """x_j = b_j + 0.999 x_(j-1) (unit diagonal, one off-diagonal of -0.999):
    a unit incoming unknown still weighs 0.999^L ~ 0.4 at a chunk's end, so
    the propagated unknowns carry their own roundings -- the result differs
    from the one-wave recurrence in the last bits (the chunked passes ran and
    were not declined) and agrees with the oracle to 1e-12 relative."""
    lib = gb_library()
    n, k = 1 << 20, 1
    ku = k if uplo == "U" else 0
    AB = np.full((2, n), -0.999, order="F")
    AB[ku, :] = np.nan
    if uplo == "U":
        AB[0, 0] = np.nan
    else:
        AB[1, n - 1] = np.nan
    b = uniform(140, (n,))
    x, r, q = b.copy(), b.copy(), b.copy()
    gpu.dtbsv(uplo, trans, "U", n, k, AB, 2, x, 1)
    ora.ora_dtbsv(c(uplo), c(trans), b"U", n, k, p(AB), 2, p(r), 1)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())
    lib.gpublas_dev_tbsv_chunked(0)
    try:
        gpu.dtbsv(uplo, trans, "U", n, k, AB, 2, q, 1)
    finally:
        lib.gpublas_dev_tbsv_chunked(1)
    assert not np.array_equal(x, q)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtbsv_chunked_declined(gpu, ora, uplo, trans):
    """Bands the chunked passes must not take: (a) integer entries in
    {-1, 0, 1} under a unit diagonal (responses grow past 1; the integer
    solution must come back exactly); (b) a zero on the diagonal (infinities
    and NaN from there on).  Both give the one-wave recurrence's bits."""
    lib = gb_library()
    n, k = 9000, 24
    rng = np.random.default_rng(130)
    kl, ku = (0, k) if uplo == "U" else (k, 0)
    AB = np.full((k + 1, n), np.nan, order="F")
    for j in range(n):
        i0, i1 = max(0, j - ku), min(n, j + kl + 1)
        AB[ku + i0 - j: ku + i1 - j, j] = rng.integers(-1, 2, i1 - i0)
    xt = rng.integers(-3, 4, n).astype(float)
    b = xt.copy()
    ora.ora_dtbmv(c(uplo), c(trans), b"U", n, k, p(AB), k + 1, p(b), 1)
    assert np.abs(b).max() < 2 ** 40
    x = b.copy()
    gpu.dtbsv(uplo, trans, "U", n, k, AB, k + 1, x, 1)
    assert np.array_equal(x, xt)
    A2 = _tb_dominant(n, k, uplo, k + 1, 131, False)
    A2[ku, n // 2] = 0.0
    b2 = uniform(132, (n,))
    x2, q2 = b2.copy(), b2.copy()
    gpu.dtbsv(uplo, trans, "N", n, k, A2, k + 1, x2, 1)
    lib.gpublas_dev_tbsv_chunked(0)
    try:
        gpu.dtbsv(uplo, trans, "N", n, k, A2, k + 1, q2, 1)
    finally:
        lib.gpublas_dev_tbsv_chunked(1)
    assert not np.isfinite(x2).all()
    assert np.array_equal(x2, q2, equal_nan=True)


@pytest.mark.parametrize("n,incx", [(2, 1), (1000, 1), (1001, 3), (65536, 1), (65537, 1),
                                    (3_000_000, 1), (200_000, 2)])
def test_idamin(gpu, ora, n, incx):
    """Bit-exact index: first minimal |x_i|, with a planted tie (the earlier
    of two equal minima wins) on both reduction paths (one workgroup up to
    2^16 elements, multi-block beyond)."""
    x = _strided(uniform(140, (n,)) + 2.0, incx)  # |x| in [1, 3]
    a, b = n // 3, (2 * n) // 3
    if n > 2:
        x[a * incx] = -0.5
        x[b * incx] = 0.5  # tie with element a: a must win
    assert gpu.idamin(n, x, incx) == ora.ora_idamin(n, p(x), incx)
    if n > 2:
        assert gpu.idamin(n, x, incx) == a + 1


@pytest.mark.parametrize("n", [1000, 200_000])
def test_idamin_nan_rules(gpu, ora, n):
    x = uniform(141, (n,)) + 2.0
    x[n // 2] = 0.25
    x[n // 4] = np.nan  # never selected after position 1
    assert gpu.idamin(n, x, 1) == ora.ora_idamin(n, p(x), 1) == n // 2 + 1
    x[0] = np.nan  # a NaN at position 1 wins
    assert gpu.idamin(n, x, 1) == ora.ora_idamin(n, p(x), 1) == 1
    x = np.full(n, np.nan)
    assert gpu.idamin(n, x, 1) == 1


@pytest.mark.parametrize("flag", [-2.0, -1.0, 0.0, 1.0])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (1000, 1, 1), (777, 2, -3), (1_000_001, 1, 1)])
def test_drotm(gpu, ora, flag, n, incx, incy):
    """Every flag of the modified-Givens H; the device may fuse a product
    into the sum: bound 2 eps (|h_a w| + |h_b z|) per element."""
    param = np.array([flag, 0.7, -0.4, 1.3, 0.2])
    x0 = _strided(uniform(142, (n,)), incx)
    y0 = _strided(uniform(143, (n,)), incy)
    X, Y, RX, RY = x0.copy(), y0.copy(), x0.copy(), y0.copy()
    gpu.drotm(n, X, incx, Y, incy, param)
    ora.ora_drotm(n, p(RX), incx, p(RY), incy, p(param))
    w, z = np.abs(_strided_get(x0, n, incx)), np.abs(_strided_get(y0, n, incy))
    H = {-2.0: [[1, 0], [0, 1]], -1.0: [[0.7, 1.3], [0.4, 0.2]], 0.0: [[1, 1.3], [0.4, 1]],
         1.0: [[0.7, 1], [1, 0.2]]}[flag]
    bx, by = 2 * EPS * (H[0][0] * w + H[0][1] * z), 2 * EPS * (H[1][0] * w + H[1][1] * z)
    assert np.all(np.abs(_strided_get(X, n, incx) - _strided_get(RX, n, incx)) <= bx)
    assert np.all(np.abs(_strided_get(Y, n, incy) - _strided_get(RY, n, incy)) <= by)
    for got, orig, inc in ((X, x0, incx), (Y, y0, incy)):
        gaps = np.ones(len(got), bool)
        gaps[:: abs(inc)][:n] = False
        assert np.array_equal(got[gaps], orig[gaps])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrsv_blocked_sizes(gpu, ora, uplo, trans, diag):
    """n = 1000: seven full 128-blocks and a partial one, solved in either
    direction by one launch per block; the other triangle (and a unit
    diagonal) NaN, so any read of them shows."""
    n = 1000
    A = uniform(26, (n, n))
    if diag == "N":
        A[np.diag_indices(n)] = n + np.abs(uniform(27, (n,)))
    else:
        A *= 1.0 / n
        A[np.diag_indices(n)] = np.nan
    A[_other_triangle(n, uplo)] = np.nan
    A = np.asfortranarray(A)
    x = uniform(28, (n,))
    r = x.copy()
    gpu.dtrsv(uplo, trans, diag, n, A, n, x, 1)
    ora.ora_dtrsv(c(uplo), c(trans), c(diag), n, p(A), n, p(r), 1)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,lda", [(1537, 1537), (2176, 2180), (640, 640)])
def test_dtrsv_lookback_shapes(gpu, ora, uplo, trans, n, lda):
    """Single-launch solve with helper workgroups sharing each block's earlier
    tiles: 13 blocks with a 1-wide last block and an odd leading dimension
    (element loads), 17 full blocks with padded columns, 5 blocks (one helper
    per block).  Padding rows and the other triangle are NaN, so any read of
    them shows; a second call gives the same bits."""
    A = np.full((lda, n), np.nan)
    T = uniform(35, (n, n))
    T[np.diag_indices(n)] = n + np.abs(uniform(36, (n,)))
    T[_other_triangle(n, uplo)] = np.nan
    A[:n, :] = T
    A = np.asfortranarray(A)
    b = uniform(37, (n,))
    x, x2, r = b.copy(), b.copy(), b.copy()
    gpu.dtrsv(uplo, trans, "N", n, A, lda, x, 1)
    ora.ora_dtrsv(c(uplo), c(trans), c("N"), n, p(A), lda, p(r), 1)
    np.testing.assert_allclose(x, r, rtol=1e-12, atol=1e-14)
    gpu.dtrsv(uplo, trans, "N", n, A, lda, x2, 1)
    assert np.array_equal(x, x2)


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T"), ("U", "N"), ("L", "T")])
def test_dtrsv_nonfinite_reaches_only_later_unknowns(gpu, ora, uplo, trans):
    """n = 1000 with one Inf in the triangle, coupling unknown 300 into 600
    (forward solves) or 600 into 300 (backward): the unknowns solved before
    the Inf appears, including those of the same 128-block, stay finite and
    match the oracle; the rest are non-finite exactly where the oracle's
    are (0 * Inf must not reach a row that was already solved)."""
    n = 1000
    A = uniform(38, (n, n))
    A[np.diag_indices(n)] = n + 1.0
    A[_other_triangle(n, uplo)] = 0.0
    fwd = (uplo == "L") == (trans == "N")
    early, late = (300, 600) if fwd else (600, 300)
    # op(A)[late, early] is the coefficient of the earlier unknown in the later row
    A[(late, early) if trans == "N" else (early, late)] = np.inf
    A = np.asfortranarray(A)
    b = uniform(39, (n,))
    x, r = b.copy(), b.copy()
    gpu.dtrsv(uplo, trans, "N", n, A, n, x, 1)
    ora.ora_dtrsv(c(uplo), c(trans), c("N"), n, p(A), n, p(r), 1)
    solved_before = np.arange(n) < late if fwd else np.arange(n) > late
    assert np.all(np.isfinite(r[solved_before])) and not np.any(np.isfinite(r[~solved_before]))
    assert np.array_equal(np.isfinite(x), np.isfinite(r))
    np.testing.assert_allclose(x[solved_before], r[solved_before], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T"), ("U", "N"), ("L", "T")])
def test_dtrsv_integer_exact_8192(gpu, uplo, trans):
    """n = 8192, unit triangle with entries in {-1, 0, 1} and an integer
    solution: every partial sum of the blocked substitution is an integer
    below 2^53, so x is recovered exactly (64 launches, each block's panel
    applied to every open unknown)."""
    n = 8192
    rng = np.random.default_rng(29)
    T = rng.integers(-1, 2, size=(n, n)).astype(float)
    T = np.tril(T, -1) if uplo == "L" else np.triu(T, 1)
    T[np.diag_indices(n)] = 1.0
    xs = rng.integers(-3, 4, size=n).astype(float)
    b = (T if trans == "N" else T.T) @ xs
    A = np.asfortranarray(T)
    A[np.diag_indices(n)] = np.nan  # unit diagonal: not read
    gpu.dtrsv(uplo, trans, "U", n, A, n, b, 1)
    assert np.array_equal(b, xs)


def test_dger_8192(gpu, ora):
    """DGER at 8192 x 8192 (16-B loads and stores along columns): against the
    oracle within one rounding per element, and bit-exact on integer data;
    a column with y(j) = 0 is left untouched (netlib)."""
    m = n = 8192
    A0 = uniform(31, (m, n))
    x, y = uniform(32, (m,)), uniform(33, (n,))
    y[77] = 0.0
    x[5] = 0.5
    A0[5, 77] = -0.0  # -0.0 + 0.5 * 0.0 would be +0.0 if the zero column were updated
    A, R = A0.copy(order="F"), A0.copy(order="F")
    gpu.dger(m, n, 0.9, x, 1, y, 1, A, m)
    ora.ora_dger(m, n, 0.9, p(x), 1, p(y), 1, p(R), m)
    scale = np.maximum(np.abs(R) + np.abs(np.outer(x, 0.9 * y)), np.finfo(float).tiny)
    assert np.max(np.abs(A - R) / scale) <= 2 * EPS
    assert np.array_equal(A[:, 77], A0[:, 77]) and np.signbit(A[5, 77])
    rng = np.random.default_rng(34)
    Ai = np.asfortranarray(rng.integers(-8, 9, size=(m, n)).astype(float))
    xi, yi = rng.integers(-8, 9, size=m).astype(float), rng.integers(-8, 9, size=n).astype(float)
    want = Ai + 2.0 * np.outer(xi, yi)
    gpu.dger(m, n, 2.0, xi, 1, yi, 1, Ai, m)
    assert np.array_equal(Ai, want)
