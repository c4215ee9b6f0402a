"""GPU parity of ZGEMM / CGEMM (complex GEMM on the real MFMA kernels) against
the CPU oracle (oracle/refblas.c ora_zgemm / ora_cgemm, netlib ZGEMM).

Every call goes through the Fortran-ABI entry points with the host
fall-through disabled (cutoff 0); `device_calls` proves the GPU ran it.

Bars:
  * elementwise gauge, the complex form of netlib's DMMCH test
    (tests/netlib/zblat3.f ZMMCH): |C - Cref| <= 16 (k + 2) eps G with
    G = |alpha| |op(A)| |op(B)| + |beta| |C0|;
  * the north star's normwise bound ||C - Cref|| / (||A|| ||B|| k eps) <= 4
    at the larger shape;
  * bit-exact results on small-integer data, where every summation order
    gives the same answer;
  * rows beyond m (padding) untouched; C not read when beta = 0.
"""
import numpy as np
import pytest

from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu

ROGUE = -1e10


def _cplx(seed, shape, dtype):
    return np.asfortranarray((uniform(seed, shape) + 1j * uniform(seed + 1000, shape)).astype(dtype))


def _padded(a, ld):
    r, cols = a.shape
    big = np.full((ld, cols), ROGUE + 1j * ROGUE, dtype=a.dtype, order="F")
    big[:r, :] = a
    return big


def _op(X, t):
    return X if t == "N" else (X.T if t == "T" else X.conj().T)


def _run(gb, ora, prec, ta, tb, m, n, k, alpha, beta, A, lda, B, ldb, C0, ldc, device=True):
    """The library's result and the oracle's on copies of C0.  device: the
    call must have run on the GPU (a scale-only call on an untracked C, alpha
    = 0 or k = 0, is done on the host, as for DGEMM)."""
    dt = np.complex128 if prec == "z" else np.complex64
    C = C0.copy(order="F")
    R = C0.copy(order="F")
    before = gb.stats()["device_calls"]
    (gb.zgemm if prec == "z" else gb.cgemm)(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    if device:
        assert gb.stats()["device_calls"] == before + 1
    fn = ora.ora_zgemm if prec == "z" else ora.ora_cgemm
    al, be = np.array([alpha], dt), np.array([beta], dt)  # alive across the call
    fn(c(ta), c(tb), m, n, k, p(al), p(A), lda, p(B), ldb, p(be), p(R), ldc)
    return C, R


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (33, 47, 29), (300, 257, 129)])
def test_complex_gemm_forms(gpu, ora, prec, ta, tb, m, n, k):
    dt = np.complex128 if prec == "z" else np.complex64
    eps = np.finfo(np.float64 if prec == "z" else np.float32).eps
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ra + 3, rb + 1, m + 2
    A = _padded(_cplx(50, (ra, ca), dt), lda)
    B = _padded(_cplx(52, (rb, cb), dt), ldb)
    C0 = _padded(_cplx(54, (m, n), dt), ldc)
    alpha, beta = 0.7 - 0.3j, -1.1 + 0.5j
    C, R = _run(gpu, ora, prec, ta, tb, m, n, k, alpha, beta, A, lda, B, ldb, C0, ldc)
    oa = _op(A[:ra, :].astype(np.complex128), ta)
    ob = _op(B[:rb, :].astype(np.complex128), tb)
    G = abs(alpha) * (np.abs(oa) @ np.abs(ob)) + abs(beta) * np.abs(C0[:m, :])
    assert np.all(np.abs(C[:m, :] - R[:m, :]) <= 16 * (k + 2) * eps * G)
    assert np.array_equal(C[m:, :], C0[m:, :])  # padding rows untouched


@pytest.mark.parametrize("prec", ["z", "c"])
def test_complex_gemm_integer_exact(gpu, ora, prec):
    """Small-integer complex data (partial sums below 2^24): bit-exact in
    every form, which pins the pack layout of each trans / conj case."""
    dt = np.complex128 if prec == "z" else np.complex64
    rng = np.random.default_rng(56)
    m, n, k = 130, 70, 90
    for ta in "NTC":
        for tb in "NTC":
            ra, ca = (m, k) if ta == "N" else (k, m)
            rb, cb = (k, n) if tb == "N" else (n, k)
            A = np.asfortranarray((rng.integers(-3, 4, (ra, ca)) + 1j * rng.integers(-3, 4, (ra, ca))).astype(dt))
            B = np.asfortranarray((rng.integers(-3, 4, (rb, cb)) + 1j * rng.integers(-3, 4, (rb, cb))).astype(dt))
            C0 = np.asfortranarray((rng.integers(-3, 4, (m, n)) + 1j * rng.integers(-3, 4, (m, n))).astype(dt))
            C, R = _run(gpu, ora, prec, ta, tb, m, n, k, 2 - 1j, 1 + 1j, A, ra, B, rb, C0, m)
            assert np.array_equal(C, R), (ta, tb)
            assert np.array_equal(R, (2 - 1j) * (_op(A, ta) @ _op(B, tb)) + (1 + 1j) * C0)


@pytest.mark.parametrize("prec", ["z", "c"])
def test_complex_gemm_netlib_scalars(gpu, ora, prec):
    """beta = 0: C is not read (NaN there does not survive); alpha = 0:
    C := beta C; alpha = 0, beta = 0: exact zeros; k = 0: C := beta C."""
    dt = np.complex128 if prec == "z" else np.complex64
    m, n, k = 40, 30, 20
    A, B = _cplx(57, (m, k), dt), _cplx(58, (k, n), dt)
    nanC = np.full((m, n), np.nan + 1j * np.nan, dt, order="F")
    C, R = _run(gpu, ora, prec, "N", "N", m, n, k, 1.5 + 0.5j, 0, A, m, B, k, nanC, m)
    assert np.all(np.isfinite(C))
    C0 = _cplx(59, (m, n), dt)
    C, R = _run(gpu, ora, prec, "N", "N", m, n, k, 0, 0.5 - 2j, A, m, B, k, C0, m, device=False)
    np.testing.assert_allclose(C, R, rtol=4 * np.finfo(dt).eps, atol=0)
    C, R = _run(gpu, ora, prec, "N", "N", m, n, k, 0, 0, A, m, B, k, nanC, m, device=False)
    assert np.all(C == 0)
    C, R = _run(gpu, ora, prec, "N", "N", m, n, 0, 1.0, 0.5 - 2j, A, m, B, 1, C0, m, device=False)
    np.testing.assert_allclose(C, R, rtol=4 * np.finfo(dt).eps, atol=0)


def test_zgemm_normwise_2048(gpu):
    """ZGEMM 2048^3 in the conjugate-transpose form: the north star's
    normwise bound against numpy's product of the same data."""
    n = 2048
    A, B, C0 = (_cplx(s, (n, n), np.complex128) for s in (60, 61, 62))
    alpha, beta = 0.9 + 0.1j, 0.5 - 0.5j
    C = C0.copy(order="F")
    gpu.zgemm("C", "N", n, n, n, alpha, A, n, B, n, beta, C, n)
    want = alpha * (A.conj().T @ B) + beta * C0
    err = np.linalg.norm(C - want) / (np.linalg.norm(A) * np.linalg.norm(B) * n * np.finfo(float).eps)
    assert err <= 4, err


def test_cblas_zgemm_row_major(gpu, ora):
    """cblas_zgemm in row-major order is the column-major product of the
    transposed operands: against the oracle on the same bytes."""
    m, n, k = 37, 23, 41
    dt = np.complex128
    A = np.ascontiguousarray(_cplx(63, (m, k), dt))  # row-major m x k
    B = np.ascontiguousarray(_cplx(64, (n, k), dt))  # row-major n x k, used transposed
    C0 = np.ascontiguousarray(_cplx(65, (m, n), dt))
    C, R = C0.copy(), C0.copy()
    alpha, beta = np.array([0.3 + 0.8j]), np.array([-0.4 + 0.1j])
    lib = gpu.library()
    RowMajor, NoTrans, ConjTrans = 101, 111, 113
    lib.cblas_zgemm(RowMajor, NoTrans, ConjTrans, m, n, k, alpha.ctypes.data, A.ctypes.data, k,
                    B.ctypes.data, k, beta.ctypes.data, C.ctypes.data, n)
    # column-major view: C^T (n x m) = conj(B) (n x k, stored k x n) ... as ZGEMM('C', 'N')
    ora.ora_zgemm(c("C"), c("N"), n, m, k, p(alpha), p(B), k, p(A), k, p(beta), p(R), n)
    eps = np.finfo(float).eps
    G = np.abs(alpha[0]) * (np.abs(A) @ np.abs(B).T) + np.abs(beta[0]) * np.abs(C0)
    assert np.all(np.abs(C - R) <= 16 * (k + 2) * eps * G)
    want = alpha[0] * (A @ B.conj().T) + beta[0] * C0
    assert np.all(np.abs(C - want) <= 16 * (k + 2) * eps * G)
