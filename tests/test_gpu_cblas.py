"""Every CBLAS twin on the device path, in both layouts (cblas.h:57-822 of the
reference; enum values cblas.h:21-25).

Row-major calls are answered by the column-major routine on the transposed
problem (operand and flag swaps, no data movement).  Each test builds the
logical matrices, stores them in the layout under test with a padded leading
dimension (padding filled with a value no result may contain), calls the twin
with the host fall-through off, and checks the logical result.

Bars: all data are small integers, so every result is an integer far below
2^24 and is exact in any summation order -- the check is bit-exact against
float64 NumPy on the logical problem (dnrm2: the rounding bound of one square
root).  Triangular solves use diagonals of +-1 and integer solutions, so they
are exact too.  Every call must run on the device (`device_calls` + 1) and
leave the padding and the unreferenced triangle untouched.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102
NOTRANS, TRANS = 111, 112
UPPER, LOWER = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
PAD = -7777.0

_SIGS = {
    "gemm": "iiiiiiSpipiSpi", "syrk": "iiiiiSpiSpi", "syr2k": "iiiiiSpipiSpi",
    "trsm": "iiiiiiiSpipi", "trmm": "iiiiiiiSpipi", "symm": "iiiiiSpipiSpi",
    "dgemv": "iiiidpipidpi", "dger": "iiidpipipi", "dtrsv": "iiiiipipi", "dtrmv": "iiiiipipi",
    "dsymv": "iiidpipidpi", "dsyr": "iiidpipi", "dsyr2": "iiidpipipi", "dspr": "iiidpip",
    "dspr2": "iiidpipip", "dtpmv": "iiiiippi", "dtpsv": "iiiiippi", "dspmv": "iiidppidpi",
    "dgbmv": "iiiiiidpipidpi", "dsbmv": "iiiidpipidpi", "dtbmv": "iiiiiipipi",
    "dtbsv": "iiiiiipipi",
}
_CT = {"i": ctypes.c_int, "d": ctypes.c_double, "f": ctypes.c_float, "p": ctypes.c_void_p}


def _fn(gb, name, prec="d"):
    """cblas_<name> for the Level-2 names ("dgemv", ...), cblas_<prec><name>
    for the Level-3 ones ("gemm", ...); S in a signature is the precision."""
    sig = _SIGS[name].replace("S", "d" if prec == "d" else "f")
    full = "cblas_" + (name if name.startswith("d") else prec + name)
    f = getattr(gb.library(), full)
    f.argtypes = [_CT[ch] for ch in sig]
    f.restype = None
    return f


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _rng(seed):
    return np.random.default_rng(seed)


def _ints(rng, shape, lo=-3, hi=3):
    return rng.integers(lo, hi + 1, size=shape).astype(np.float64)


class Stored:
    """A logical r x c matrix stored in `layout` with leading dimension ld."""

    def __init__(self, M, layout, ld_extra=2, dtype=np.float64):
        r, c = M.shape
        self.layout, self.shape = layout, (r, c)
        if layout == COL:
            self.ld = max(1, r + ld_extra)
            self.buf = np.full((self.ld, c), PAD, dtype=dtype, order="F")
            self.buf[:r, :] = M
        else:
            self.ld = max(1, c + ld_extra)
            self.buf = np.full((r, self.ld), PAD, dtype=dtype, order="C")
            self.buf[:, :c] = M
        self.orig = self.buf.copy(order="A")

    def logical(self):
        r, c = self.shape
        return (self.buf[:r, :] if self.layout == COL else self.buf[:, :c]).astype(np.float64)

    def padding_intact(self):
        r, c = self.shape
        if self.layout == COL:
            return np.array_equal(self.buf[r:, :], self.orig[r:, :])
        return np.array_equal(self.buf[:, c:], self.orig[:, c:])


def _tri_mask(n, uplo):
    return np.tril(np.ones((n, n), bool)) if uplo == LOWER else np.triu(np.ones((n, n), bool))


def _tri(rng, n, uplo, diag):
    """Triangle with entries in {-1,0,1} and a +-1 (or unit) diagonal; the other
    triangle (and a unit diagonal) filled with PAD in the stored matrix."""
    T = _ints(rng, (n, n), -1, 1)
    T = np.tril(T, -1) if uplo == LOWER else np.triu(T, 1)
    d = rng.choice([-1.0, 1.0], size=n) if diag == NONUNIT else np.ones(n)
    T += np.diag(d)
    S = np.where(_tri_mask(n, uplo), T, PAD)
    if diag == UNIT:
        S[np.diag_indices(n)] = PAD
    return T, S


def _op(M, t):
    return M if t == NOTRANS else M.T


def _vec(v, inc):
    """Strided storage of v with increment inc (negative: stored backwards)."""
    n = len(v)
    buf = np.full(1 + (n - 1) * abs(inc) if n else 1, PAD)
    idx = np.arange(n) * abs(inc)
    buf[idx if inc > 0 else idx[::-1]] = v
    return buf


def _unvec(buf, n, inc):
    idx = np.arange(n) * abs(inc)
    return buf[idx if inc > 0 else idx[::-1]]


def _pack(M, uplo, layout):
    n = M.shape[0]
    out = []
    # row-major upper = column-major lower of the transpose (and vice versa)
    colwise_lower = (uplo == LOWER) == (layout == COL)
    A = M if layout == COL else M.T
    for j in range(n):
        out.extend(A[j:, j] if colwise_lower else A[:j + 1, j])
    return np.array(out, dtype=np.float64)


def _unpack(ap, n, uplo, layout):
    A = np.zeros((n, n))
    colwise_lower = (uplo == LOWER) == (layout == COL)
    q = 0
    for j in range(n):
        if colwise_lower:
            A[j:, j] = ap[q:q + n - j]
            q += n - j
        else:
            A[:j + 1, j] = ap[q:q + j + 1]
            q += j + 1
    return A if layout == COL else A.T


def _band(M, kl, ku, layout, ld_extra=1):
    """General band storage (symmetric / triangular bands: kl or ku = 0)."""
    m, n = M.shape
    lda = kl + ku + 1 + ld_extra
    if layout == COL:
        S = np.full((lda, n), PAD, order="F")
        for j in range(n):
            for i in range(max(0, j - ku), min(m, j + kl + 1)):
                S[ku + i - j, j] = M[i, j]
    else:
        S = np.full((m, lda), PAD, order="C")
        for i in range(m):
            for j in range(max(0, i - kl), min(n, i + ku + 1)):
                S[i, kl + j - i] = M[i, j]
    return S, lda


def _banded(rng, m, n, kl, ku, lo=-3, hi=3):
    M = _ints(rng, (m, n), lo, hi)
    i, j = np.indices((m, n))
    return np.where((j - i <= ku) & (i - j <= kl), M, 0.0)


def _calls(gb):
    return gb.stats()["device_calls"]


# ----------------------------------------------------------------------------- Level 3
@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("ta,tb", [(NOTRANS, NOTRANS), (NOTRANS, TRANS), (TRANS, NOTRANS), (TRANS, TRANS)])
def test_cblas_gemm(gpu, prec, layout, ta, tb):
    rng = _rng(1)
    m, n, k = 37, 29, 23
    dt = np.float64 if prec == "d" else np.float32
    A = Stored(_ints(rng, (m, k) if ta == NOTRANS else (k, m)), layout, dtype=dt)
    B = Stored(_ints(rng, (k, n) if tb == NOTRANS else (n, k)), layout, dtype=dt)
    C0 = _ints(rng, (m, n))
    C = Stored(C0, layout, dtype=dt)
    before = _calls(gpu)
    _fn(gpu, "gemm", prec)(layout, ta, tb, m, n, k, 2.0, _ptr(A.buf), A.ld, _ptr(B.buf), B.ld,
                           -1.0, _ptr(C.buf), C.ld)
    assert _calls(gpu) == before + 1
    ref = 2.0 * _op(A.logical(), ta) @ _op(B.logical(), tb) - C0
    assert np.array_equal(C.logical(), ref)
    assert C.padding_intact()


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("uplo,trans", [(UPPER, NOTRANS), (LOWER, TRANS), (LOWER, NOTRANS)])
def test_cblas_syrk_syr2k(gpu, prec, layout, uplo, trans):
    rng = _rng(2)
    n, k = 33, 21
    dt = np.float64 if prec == "d" else np.float32
    A = Stored(_ints(rng, (n, k) if trans == NOTRANS else (k, n)), layout, dtype=dt)
    B = Stored(_ints(rng, (n, k) if trans == NOTRANS else (k, n)), layout, dtype=dt)
    C0 = _ints(rng, (n, n))
    tri = _tri_mask(n, uplo)
    for which in ("syrk", "syr2k"):
        C = Stored(C0, layout, dtype=dt)
        before = _calls(gpu)
        if which == "syrk":
            _fn(gpu, "syrk", prec)(layout, uplo, trans, n, k, 3.0, _ptr(A.buf), A.ld, 2.0,
                                   _ptr(C.buf), C.ld)
            a = _op(A.logical(), trans)
            full = 3.0 * a @ a.T + 2.0 * C0
        else:
            _fn(gpu, "syr2k", prec)(layout, uplo, trans, n, k, 3.0, _ptr(A.buf), A.ld,
                                    _ptr(B.buf), B.ld, 2.0, _ptr(C.buf), C.ld)
            a, b = _op(A.logical(), trans), _op(B.logical(), trans)
            full = 3.0 * (a @ b.T + b @ a.T) + 2.0 * C0
        assert _calls(gpu) == before + 1, which
        got = C.logical()
        assert np.array_equal(got[tri], full[tri]), which
        assert np.array_equal(got[~tri], C0[~tri]), which  # other triangle untouched
        assert C.padding_intact()


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("side,uplo", [(LEFT, UPPER), (LEFT, LOWER), (RIGHT, UPPER), (RIGHT, LOWER)])
def test_cblas_symm(gpu, prec, layout, side, uplo):
    rng = _rng(3)
    m, n = 31, 26
    dt = np.float64 if prec == "d" else np.float32
    na = m if side == LEFT else n
    S = _ints(rng, (na, na))
    S = np.tril(S) + np.tril(S, -1).T
    A = Stored(np.where(_tri_mask(na, uplo), S, PAD), layout, dtype=dt)
    B = Stored(_ints(rng, (m, n)), layout, dtype=dt)
    C0 = _ints(rng, (m, n))
    C = Stored(C0, layout, dtype=dt)
    before = _calls(gpu)
    _fn(gpu, "symm", prec)(layout, side, uplo, m, n, 2.0, _ptr(A.buf), A.ld, _ptr(B.buf), B.ld,
                           1.0, _ptr(C.buf), C.ld)
    assert _calls(gpu) == before + 1
    ref = 2.0 * (S @ B.logical() if side == LEFT else B.logical() @ S) + C0
    assert np.array_equal(C.logical(), ref)
    assert C.padding_intact()


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("side,uplo,ta,diag", [
    (LEFT, LOWER, NOTRANS, NONUNIT), (LEFT, UPPER, TRANS, UNIT),
    (RIGHT, UPPER, NOTRANS, NONUNIT), (RIGHT, LOWER, TRANS, UNIT),
    (LEFT, UPPER, NOTRANS, UNIT), (RIGHT, LOWER, NOTRANS, NONUNIT)])
def test_cblas_trsm_trmm(gpu, prec, layout, side, uplo, ta, diag):
    rng = _rng(4)
    m, n = 35, 27
    dt = np.float64 if prec == "d" else np.float32
    na = m if side == LEFT else n
    T, S = _tri(rng, na, uplo, diag)
    A = Stored(S, layout, dtype=dt)
    opT = _op(T, ta)
    X = _ints(rng, (m, n))
    # TRSM: B = op(A) X (left) or X op(A) (right), solved back to X (alpha 2: B/2 stored)
    Bv = 2.0 * (opT @ X if side == LEFT else X @ opT)
    B = Stored(Bv, layout, dtype=dt)
    before = _calls(gpu)
    _fn(gpu, "trsm", prec)(layout, side, uplo, ta, diag, m, n, 0.5, _ptr(A.buf), A.ld,
                           _ptr(B.buf), B.ld)
    assert _calls(gpu) == before + 1
    assert np.array_equal(B.logical(), X)
    assert B.padding_intact()
    # TRMM: X := 3 op(A) X (left) or 3 X op(A) (right)
    B = Stored(X, layout, dtype=dt)
    _fn(gpu, "trmm", prec)(layout, side, uplo, ta, diag, m, n, 3.0, _ptr(A.buf), A.ld,
                           _ptr(B.buf), B.ld)
    assert _calls(gpu) == before + 2
    assert np.array_equal(B.logical(), 3.0 * (opT @ X if side == LEFT else X @ opT))
    assert B.padding_intact()


# ----------------------------------------------------------------------------- Level 2
@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("trans,incx,incy", [(NOTRANS, 1, 1), (TRANS, 2, -1), (NOTRANS, -3, 2)])
def test_cblas_dgemv_dger(gpu, layout, trans, incx, incy):
    rng = _rng(5)
    m, n = 41, 30
    A = Stored(_ints(rng, (m, n)), layout)
    lx, ly = (n, m) if trans == NOTRANS else (m, n)
    x, y0 = _ints(rng, lx), _ints(rng, ly)
    xb, yb = _vec(x, incx), _vec(y0, incy)
    before = _calls(gpu)
    _fn(gpu, "dgemv")(layout, trans, m, n, 2.0, _ptr(A.buf), A.ld, _ptr(xb), incx, -1.0,
                      _ptr(yb), incy)
    assert _calls(gpu) == before + 1
    assert np.array_equal(_unvec(yb, ly, incy), 2.0 * _op(A.logical(), trans) @ x - y0)
    # DGER: A := 2 u v^T + A
    u, v = _ints(rng, m), _ints(rng, n)
    ub, vb = _vec(u, incx), _vec(v, incy)
    A0 = A.logical()
    _fn(gpu, "dger")(layout, m, n, 2.0, _ptr(ub), incx, _ptr(vb), incy, _ptr(A.buf), A.ld)
    assert _calls(gpu) == before + 2
    assert np.array_equal(A.logical(), A0 + 2.0 * np.outer(u, v))
    assert A.padding_intact()


@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("uplo,trans,diag,inc", [
    (LOWER, NOTRANS, NONUNIT, 1), (UPPER, TRANS, UNIT, -2), (UPPER, NOTRANS, NONUNIT, 3),
    (LOWER, TRANS, UNIT, 1)])
def test_cblas_triangular_vectors(gpu, layout, uplo, trans, diag, inc):
    """DTRMV / DTRSV, DTPMV / DTPSV (packed) and DTBMV / DTBSV (band)."""
    rng = _rng(6)
    n, k = 40, 3
    T, S = _tri(rng, n, uplo, diag)
    A = Stored(S, layout)
    x = _ints(rng, n)
    opT = _op(T, trans)
    before = _calls(gpu)
    xb = _vec(x, inc)
    _fn(gpu, "dtrmv")(layout, uplo, trans, diag, n, _ptr(A.buf), A.ld, _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), opT @ x)
    xb = _vec(opT @ x, inc)
    _fn(gpu, "dtrsv")(layout, uplo, trans, diag, n, _ptr(A.buf), A.ld, _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), x)
    assert A.padding_intact()
    ap = _pack(np.where(_tri_mask(n, uplo), T, 0.0), uplo, layout)
    if diag == UNIT:  # the unit diagonal's slots are not read
        dpos = _pack(np.eye(n) * 1.0, uplo, layout) != 0
        ap[dpos] = PAD
    xb = _vec(x, inc)
    _fn(gpu, "dtpmv")(layout, uplo, trans, diag, n, _ptr(ap), _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), opT @ x)
    xb = _vec(opT @ x, inc)
    _fn(gpu, "dtpsv")(layout, uplo, trans, diag, n, _ptr(ap), _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), x)
    # band triangle of half-width k
    Tb = _banded(rng, n, n, k if uplo == LOWER else 0, 0 if uplo == LOWER else k, -1, 1)
    Tb[np.diag_indices(n)] = rng.choice([-1.0, 1.0], size=n) if diag == NONUNIT else 1.0
    Sb, ldb = _band(Tb, k if uplo == LOWER else 0, 0 if uplo == LOWER else k, layout)
    if diag == UNIT:
        kl = k if uplo == LOWER else 0
        if layout == COL:
            Sb[0 if uplo == LOWER else k, :] = PAD
        else:
            Sb[:, kl] = PAD
    opb = _op(Tb, trans)
    xb = _vec(x, inc)
    _fn(gpu, "dtbmv")(layout, uplo, trans, diag, n, k, _ptr(Sb), ldb, _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), opb @ x)
    xb = _vec(opb @ x, inc)
    _fn(gpu, "dtbsv")(layout, uplo, trans, diag, n, k, _ptr(Sb), ldb, _ptr(xb), inc)
    assert np.array_equal(_unvec(xb, n, inc), x)
    assert _calls(gpu) == before + 6


@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("uplo,incx,incy", [(UPPER, 1, 1), (LOWER, -2, 3)])
def test_cblas_symmetric_vectors(gpu, layout, uplo, incx, incy):
    """DSYMV / DSYR / DSYR2, their packed forms DSPMV / DSPR / DSPR2, DSBMV."""
    rng = _rng(7)
    n, k = 38, 4
    S = _ints(rng, (n, n))
    S = np.tril(S) + np.tril(S, -1).T
    tri = _tri_mask(n, uplo)
    x, y0 = _ints(rng, n), _ints(rng, n)
    before = _calls(gpu)
    A = Stored(np.where(tri, S, PAD), layout)
    xb, yb = _vec(x, incx), _vec(y0, incy)
    _fn(gpu, "dsymv")(layout, uplo, n, 2.0, _ptr(A.buf), A.ld, _ptr(xb), incx, 3.0, _ptr(yb), incy)
    assert np.array_equal(_unvec(yb, n, incy), 2.0 * S @ x + 3.0 * y0)
    _fn(gpu, "dsyr")(layout, uplo, n, 2.0, _ptr(xb), incx, _ptr(A.buf), A.ld)
    S1 = S + 2.0 * np.outer(x, x)
    assert np.array_equal(A.logical()[tri], S1[tri])
    yb = _vec(y0, incy)
    _fn(gpu, "dsyr2")(layout, uplo, n, -1.0, _ptr(xb), incx, _ptr(yb), incy, _ptr(A.buf), A.ld)
    S2 = S1 - (np.outer(x, y0) + np.outer(y0, x))
    got = A.logical()
    assert np.array_equal(got[tri], S2[tri])
    assert np.all(got[~tri] == PAD) and A.padding_intact()
    ap = _pack(S, uplo, layout)
    yb = _vec(y0, incy)
    _fn(gpu, "dspmv")(layout, uplo, n, 2.0, _ptr(ap), _ptr(xb), incx, 3.0, _ptr(yb), incy)
    assert np.array_equal(_unvec(yb, n, incy), 2.0 * S @ x + 3.0 * y0)
    _fn(gpu, "dspr")(layout, uplo, n, 2.0, _ptr(xb), incx, _ptr(ap))
    assert np.array_equal(_unpack(ap, n, uplo, layout)[tri], S1[tri])
    yb = _vec(y0, incy)
    _fn(gpu, "dspr2")(layout, uplo, n, -1.0, _ptr(xb), incx, _ptr(yb), incy, _ptr(ap))
    assert np.array_equal(_unpack(ap, n, uplo, layout)[tri], S2[tri])
    # symmetric band of half-width k, the `uplo` half stored
    Sb = _banded(rng, n, n, k, k)
    Sb = np.tril(Sb) + np.tril(Sb, -1).T
    kl, ku = (k, 0) if uplo == LOWER else (0, k)
    St, ldb = _band(np.where(tri, Sb, 0.0), kl, ku, layout)
    yb = _vec(y0, incy)
    _fn(gpu, "dsbmv")(layout, uplo, n, k, 2.0, _ptr(St), ldb, _ptr(xb), incx, -1.0, _ptr(yb), incy)
    assert np.array_equal(_unvec(yb, n, incy), 2.0 * Sb @ x - y0)
    assert _calls(gpu) == before + 7


@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("trans,kl,ku", [(NOTRANS, 2, 5), (TRANS, 4, 1), (NOTRANS, 0, 0)])
def test_cblas_dgbmv(gpu, layout, trans, kl, ku):
    rng = _rng(8)
    m, n = 45, 37
    M = _banded(rng, m, n, kl, ku)
    S, lda = _band(M, kl, ku, layout)
    lx, ly = (n, m) if trans == NOTRANS else (m, n)
    x, y0 = _ints(rng, lx), _ints(rng, ly)
    yb = y0.copy()
    before = _calls(gpu)
    _fn(gpu, "dgbmv")(layout, trans, m, n, kl, ku, 2.0, _ptr(S), lda, _ptr(x), 1, -2.0, _ptr(yb), 1)
    assert _calls(gpu) == before + 1
    assert np.array_equal(yb, 2.0 * _op(M, trans) @ x - 2.0 * y0)


# ----------------------------------------------------------------------------- Level 1
def test_cblas_level1(gpu):
    lib = gpu.library()
    rng = _rng(9)
    n = 301
    x, y = _ints(rng, n), _ints(rng, n)
    x[17], x[230] = 9.0, -9.0  # equal magnitudes: the first index wins
    before = _calls(gpu)
    assert lib.cblas_idamax(n, _ptr(x), 1) == 17  # CBLAS: 0-based
    x2 = x.copy()
    x2[5] = 0.0
    assert lib.cblas_idamin(n, _ptr(x2), 1) == int(np.argmin(np.abs(x2)))
    assert lib.cblas_ddot(n, _ptr(x), 1, _ptr(y), 1) == float(x @ y)
    assert lib.cblas_dasum(n, _ptr(x), 1) == float(np.sum(np.abs(x)))
    r = lib.cblas_dnrm2(n, _ptr(x), 1)
    exact = float(np.sqrt(x @ x))
    assert abs(r - exact) <= 4 * np.finfo(float).eps * exact
    yy = y.copy()
    lib.cblas_daxpy(n, ctypes.c_double(2.0), _ptr(x), 1, _ptr(yy), 1)
    assert np.array_equal(yy, 2.0 * x + y)
    xs = x.copy()
    lib.cblas_dscal(n, ctypes.c_double(-3.0), _ptr(xs), 1)
    assert np.array_equal(xs, -3.0 * x)
    yc = np.zeros(n)
    lib.cblas_dcopy(n, _ptr(x), 1, _ptr(yc), 1)
    assert np.array_equal(yc, x)
    a, b = x.copy(), y.copy()
    lib.cblas_dswap(n, _ptr(a), 1, _ptr(b), 1)
    assert np.array_equal(a, y) and np.array_equal(b, x)
    a, b = x.copy(), y.copy()
    lib.cblas_drot(n, _ptr(a), 1, _ptr(b), 1, ctypes.c_double(2.0), ctypes.c_double(-1.0))
    assert np.array_equal(a, 2.0 * x - y) and np.array_equal(b, 2.0 * y + x)
    a, b = x.copy(), y.copy()
    param = np.array([-1.0, 2.0, -1.0, 3.0, 1.0])  # full H = [[2, 3], [-1, 1]]
    lib.cblas_drotm(n, _ptr(a), 1, _ptr(b), 1, _ptr(param))
    assert np.array_equal(a, 2.0 * x + 3.0 * y) and np.array_equal(b, -x + y)
    assert _calls(gpu) == before + 11
