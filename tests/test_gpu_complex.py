"""GPU parity of ZGEMM / CGEMM (complex GEMM on the real MFMA kernels) against
the CPU oracle (oracle/refblas.c ora_zgemm / ora_cgemm, netlib ZGEMM).

Every call goes through the Fortran-ABI entry points with the host
fall-through disabled (cutoff 0); `device_calls` proves the GPU ran it.

Bars:
  * elementwise gauge, the complex form of netlib's DMMCH test
    (tests/netlib/zblat3.f ZMMCH): |C - Cref| <= 16 (k + 2) eps G with
    G = |alpha| |op(A)| |op(B)| + |beta| |C0|;
  * the normwise bound ||C - Cref|| / (||A|| ||B|| k eps) <= 4
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
    """ZGEMM 2048^3 in the conjugate-transpose form: the normwise bound
    ||C - Cref|| / (||A|| ||B|| k eps) <= 4 against numpy's product of the same data."""
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


# ---- the rest of the complex Level-3, at shapes past the netlib testers' N <= 9 ----
def _tri(X, uplo):
    return np.triu(X) if uplo == "U" else np.tril(X)


def _herm_full(A, uplo, herm):
    """The full matrix a SYMM / HEMM reads from the stored triangle."""
    T = _tri(A, uplo)
    off = T - np.diag(np.diag(T))
    F = T + (off.conj().T if herm else off.T)
    if herm:
        F[np.diag_indices(len(F))] = F.diagonal().real  # a Hermitian diagonal is real
    return F


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("kind", ["syrk", "herk", "syr2k", "her2k"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "X"])  # X: T for syrk / syr2k, C for herk / her2k
def test_complex_rank_k_updates(gpu, prec, kind, uplo, trans):
    dt = np.complex128 if prec == "z" else np.complex64
    eps = np.finfo(np.float64 if prec == "z" else np.float32).eps
    herm = kind.startswith("he")
    tr = "N" if trans == "N" else ("C" if herm else "T")
    n, k = 200, 150
    shape = (n, k) if tr == "N" else (k, n)
    A, B = _cplx(70, shape, dt), _cplx(72, shape, dt)
    C0 = _cplx(74, (n, n), dt)
    alpha = 0.6 if kind == "herk" else 0.6 - 0.7j
    beta = 0.8 if herm else 0.8 + 0.3j
    C = C0.copy(order="F")
    fn = getattr(gpu.library(), prec + kind + "_")
    rdt = np.float64 if prec == "z" else np.float32
    al = np.array([alpha], rdt) if kind == "herk" else np.array([alpha], dt)
    be = np.array([beta], rdt) if herm else np.array([beta], dt)
    before = gpu.stats()["device_calls"]
    args = [gpu._c(uplo), gpu._c(tr), gpu._i(n), gpu._i(k), p(al), p(A), gpu._i(shape[0])]
    if kind.endswith("2k"):
        args += [p(B), gpu._i(shape[0])]
    fn(*args, p(be), p(C), gpu._i(n))
    assert gpu.stats()["device_calls"] == before + 1
    X = A.astype(np.complex128) if tr == "N" else (A.T if tr == "T" else A.conj().T).astype(np.complex128)
    Y = B.astype(np.complex128) if tr == "N" else (B.T if tr == "T" else B.conj().T).astype(np.complex128)
    if kind == "syrk":
        P, G = alpha * X @ X.T, abs(alpha) * np.abs(X) @ np.abs(X).T
    elif kind == "herk":
        P, G = alpha * X @ X.conj().T, abs(alpha) * np.abs(X) @ np.abs(X).T
    elif kind == "syr2k":
        P = alpha * X @ Y.T + alpha * Y @ X.T
        G = abs(alpha) * (np.abs(X) @ np.abs(Y).T + np.abs(Y) @ np.abs(X).T)
    else:
        P = alpha * X @ Y.conj().T + np.conj(alpha) * Y @ X.conj().T
        G = abs(alpha) * (np.abs(X) @ np.abs(Y).T + np.abs(Y) @ np.abs(X).T)
    C0d = C0.astype(np.complex128)
    if herm:
        C0d = C0d - 1j * np.diag(np.diag(C0d).imag)  # the diagonal's imaginary part is not read
    want = P + beta * C0d
    G = G + abs(beta) * np.abs(C0)
    m = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.all(np.abs(C - want)[m] <= 16 * (k + 2) * eps * G[m])
    assert np.array_equal(C[~m], C0[~m])  # the other triangle untouched
    if herm:
        assert np.all(np.diag(C).imag == 0)


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("herm", [0, 1])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_complex_symm_hemm(gpu, prec, herm, side, uplo):
    dt = np.complex128 if prec == "z" else np.complex64
    eps = np.finfo(np.float64 if prec == "z" else np.float32).eps
    m, n = 150, 170
    na = m if side == "L" else n
    A = _cplx(76, (na, na), dt)
    B, C0 = _cplx(77, (m, n), dt), _cplx(78, (m, n), dt)
    alpha, beta = 0.9 + 0.2j, -0.4 + 0.6j
    C = C0.copy(order="F")
    fn = getattr(gpu.library(), prec + ("hemm_" if herm else "symm_"))
    al, be = np.array([alpha], dt), np.array([beta], dt)
    fn(gpu._c(side), gpu._c(uplo), gpu._i(m), gpu._i(n), p(al), p(A), gpu._i(na), p(B),
       gpu._i(m), p(be), p(C), gpu._i(m))
    F = _herm_full(A.astype(np.complex128), uplo, herm)
    P = F @ B if side == "L" else B @ F
    G = abs(alpha) * (np.abs(F) @ np.abs(B) if side == "L" else np.abs(B) @ np.abs(F)) + abs(beta) * np.abs(C0)
    assert np.all(np.abs(C - (alpha * P + beta * C0)) <= 16 * (na + 2) * eps * G)


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_complex_trsm_trmm(gpu, prec, side, uplo, ta, diag):
    """TRSM at orders past the substitution base (recursive halving with
    complex GEMM updates): the residual op(A) X - alpha B against the
    triangular-solve error bound; TRMM against the product of the expanded
    triangle."""
    dt = np.complex128 if prec == "z" else np.complex64
    eps = np.finfo(np.float64 if prec == "z" else np.float32).eps
    m, n = 230, 170
    na = m if side == "L" else n
    A = np.asfortranarray(_cplx(80, (na, na), dt) / na)  # small off-diagonal: well conditioned
    A[np.diag_indices(na)] = 1.0 + 0.5j  # (a unit diagonal is not read)
    B0 = _cplx(81, (m, n), dt)
    alpha = 0.8 - 0.5j
    al = np.array([alpha], dt)
    T = _tri(A.astype(np.complex128), uplo)
    if diag == "U":
        T = T - np.diag(np.diag(T)) + np.eye(na)
    opT = T if ta == "N" else (T.T if ta == "T" else T.conj().T)
    lib = gpu.library()
    args = lambda Bx: (gpu._c(side), gpu._c(uplo), gpu._c(ta), gpu._c(diag), gpu._i(m), gpu._i(n),
                       p(al), p(A), gpu._i(na), p(Bx), gpu._i(m))
    X = B0.copy(order="F")
    getattr(lib, prec + "trsm_")(*args(X))
    R = opT @ X - alpha * B0 if side == "L" else X @ opT - alpha * B0
    scale = np.linalg.norm(opT) * np.linalg.norm(X) + abs(alpha) * np.linalg.norm(B0)
    assert np.linalg.norm(R) <= 8 * na * eps * scale
    Y = B0.copy(order="F")
    getattr(lib, prec + "trmm_")(*args(Y))
    want = alpha * (opT @ B0 if side == "L" else B0 @ opT)
    G = abs(alpha) * (np.abs(opT) @ np.abs(B0) if side == "L" else np.abs(B0) @ np.abs(opT))
    assert np.all(np.abs(Y - want) <= 16 * (na + 2) * eps * G)


def test_cblas_complex_row_major_twins(gpu):
    """Row-major cblas_zherk / cblas_zher2k (alpha conjugated on the way to
    the column-major call) / cblas_zhemm / cblas_ztrsm against numpy's
    closed forms on row-major arrays."""
    import ctypes
    lib = gpu.library()
    RowMajor, NoTrans, ConjTrans = 101, 111, 113
    Upper, Lower, Left, Right, NonUnit = 121, 122, 141, 142, 131
    eps = np.finfo(float).eps
    n, k = 90, 70
    rm = lambda s, shape: np.ascontiguousarray(_cplx(s, shape, np.complex128))
    A, B, C0 = rm(90, (n, k)), rm(91, (n, k)), rm(92, (n, n))
    up = np.triu(np.ones((n, n), bool))
    # ZHERK: C = alpha A A^H + beta C, upper triangle of the row-major C
    C = C0.copy()
    lib.cblas_zherk(RowMajor, Upper, NoTrans, n, k, ctypes.c_double(0.7), A.ctypes.data, k,
                    ctypes.c_double(0.3), C.ctypes.data, n)
    want = 0.7 * A @ A.conj().T + 0.3 * (C0 - 1j * np.diag(np.diag(C0).imag))
    G = 0.7 * np.abs(A) @ np.abs(A).T + 0.3 * np.abs(C0)
    assert np.all(np.abs(C - want)[up] <= 16 * (k + 2) * eps * G[up])
    assert np.array_equal(C[~up], C0[~up])
    # ZHER2K: C = alpha A B^H + conj(alpha) B A^H + beta C
    al = np.array([0.4 - 0.9j])
    C = C0.copy()
    lib.cblas_zher2k(RowMajor, Upper, NoTrans, n, k, al.ctypes.data, A.ctypes.data, k,
                     B.ctypes.data, k, ctypes.c_double(0.5), C.ctypes.data, n)
    want = al[0] * A @ B.conj().T + np.conj(al[0]) * B @ A.conj().T + 0.5 * (C0 - 1j * np.diag(np.diag(C0).imag))
    G = abs(al[0]) * (np.abs(A) @ np.abs(B).T + np.abs(B) @ np.abs(A).T) + 0.5 * np.abs(C0)
    assert np.all(np.abs(C - want)[up] <= 16 * (k + 2) * eps * G[up])
    # ZHEMM, left: C = alpha H B2 + beta C with H Hermitian from the lower triangle
    m2, n2 = 60, 50
    H = rm(93, (m2, m2))
    B2, C2 = rm(94, (m2, n2)), rm(95, (m2, n2))
    Lt = np.tril(H)
    Hf = Lt + np.tril(H, -1).conj().T
    Hf[np.diag_indices(m2)] = Hf.diagonal().real
    be = np.array([0.2 + 0.1j])
    C = C2.copy()
    lib.cblas_zhemm(RowMajor, Left, Lower, m2, n2, al.ctypes.data, H.ctypes.data, m2,
                    B2.ctypes.data, n2, be.ctypes.data, C.ctypes.data, n2)
    want = al[0] * Hf @ B2 + be[0] * C2
    G = abs(al[0]) * np.abs(Hf) @ np.abs(B2) + abs(be[0]) * np.abs(C2)
    assert np.all(np.abs(C - want) <= 16 * (m2 + 2) * eps * G)
    # ZTRSM, right, conjugate transpose: X U^H = alpha B2
    U = rm(96, (n2, n2)) / n2
    U[np.diag_indices(n2)] = 1.0 + 0.25j
    X = B2.copy()
    lib.cblas_ztrsm(RowMajor, Right, Upper, ConjTrans, NonUnit, m2, n2, al.ctypes.data,
                    U.ctypes.data, n2, X.ctypes.data, n2)
    opU = np.triu(U).conj().T
    R = X @ opU - al[0] * B2
    assert np.linalg.norm(R) <= 8 * n2 * eps * (np.linalg.norm(opU) * np.linalg.norm(X) + np.linalg.norm(B2))


def test_complex_on_tracked_operands(gpu):
    """ZGEMM and ZHERK on tracked (HBM-resident) complex arrays: operands used
    in place (hits, no misses), results equal to the untracked calls bit for
    bit, and alpha = 0 on a tracked C scaled on the device (C := beta C)."""
    gb = gpu
    n, k = 160, 96
    A0, B0, C0 = (_cplx(s, (n, n), np.complex128) for s in (100, 101, 102))
    A, B, C = gb.tracked_copy(A0), gb.tracked_copy(B0), gb.tracked_copy(C0)
    Cu = C0.copy(order="F")
    gb.zgemm("C", "N", n, n, n, 0.5 + 0.5j, A0, n, B0, n, 1.5 - 0.5j, Cu, n)
    gb.reset_stats()
    gb.zgemm("C", "N", n, n, n, 0.5 + 0.5j, A, n, B, n, 1.5 - 0.5j, C, n)
    s = gb.stats()
    assert s["hits"] == 3 and s["misses"] == 0 and s["device_calls"] == 1
    assert np.array_equal(np.array(C), Cu)
    # alpha = 0: C := beta C on the device (C is tracked)
    before = gb.stats()["device_calls"]
    gb.zgemm("N", "N", n, n, n, 0, A, n, B, n, 0.25 + 1j, C, n)
    assert gb.stats()["device_calls"] == before + 1
    np.testing.assert_allclose(np.array(C), (0.25 + 1j) * Cu, rtol=4 * np.finfo(float).eps, atol=0)
    # ZHERK on a tracked operand and C
    X = gb.tracked_copy(_cplx(103, (n, k), np.complex128))
    H = gb.tracked_copy(_cplx(104, (n, n), np.complex128))
    H0 = np.array(H)
    lib = gb.library()
    al, be = np.array([0.9]), np.array([0.1])
    lib.zherk_(gb._c("L"), gb._c("N"), gb._i(n), gb._i(k), p(al), p(X), gb._i(n), p(be), p(H),
               gb._i(n))
    Xh = np.array(X)
    want = 0.9 * Xh @ Xh.conj().T + 0.1 * (H0 - 1j * np.diag(np.diag(H0).imag))
    lo = np.tril(np.ones((n, n), bool))
    G = 0.9 * np.abs(Xh) @ np.abs(Xh).T + 0.1 * np.abs(H0)
    assert np.all(np.abs(np.array(H) - want)[lo] <= 16 * (k + 2) * np.finfo(float).eps * G[lo])
    assert np.array_equal(np.array(H)[~lo], H0[~lo])


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("routine", ["syrk", "herk", "syr2k", "her2k", "symm", "hemm", "trmm", "trsm"])
def test_complex_level3_against_oracle(gpu, ora, prec, routine):
    """Every complex Level-3 routine against the oracle's restatement
    (oracle/refblas.c) at shapes past the netlib testers' N <= 9, in two
    forms each, within the ZMMCH-style gauge (TRSM: the oracle's own
    substitution, relative to the size of the solution)."""
    dt, rdt = (np.complex128, np.float64) if prec == "z" else (np.complex64, np.float32)
    eps = np.finfo(rdt).eps
    lib = gpu.library()
    n, k, m = 130, 90, 110
    for form in range(2):
        if routine in ("syrk", "herk", "syr2k", "her2k"):
            herm = routine.startswith("he")
            uplo, tr = ("L", "N") if form == 0 else ("U", "C" if herm else "T")
            shape = (n, k) if tr == "N" else (k, n)
            A, B, C0 = _cplx(110, shape, dt), _cplx(111, shape, dt), _cplx(112, (n, n), dt)
            al = np.array([0.8], rdt) if routine == "herk" else np.array([0.8 - 0.4j], dt)
            be = np.array([0.6], rdt) if herm else np.array([0.6 + 0.2j], dt)
            C, R = C0.copy(order="F"), C0.copy(order="F")
            head = [gpu._c(uplo), gpu._c(tr), gpu._i(n), gpu._i(k), p(al), p(A), gpu._i(shape[0])]
            two = [p(B), gpu._i(shape[0])] if routine.endswith("2k") else []
            getattr(lib, prec + routine + "_")(*head, *two, p(be), p(C), gpu._i(n))
            oargs = [c(uplo), c(tr), n, k, p(al), p(A), shape[0]]
            otwo = [p(B), shape[0]] if routine.endswith("2k") else []
            getattr(ora, f"ora_{prec}{routine}")(*oargs, *otwo, p(be), p(R), n)
            Xa = np.abs(A.astype(np.complex128))
            X = Xa if tr == "N" else Xa.T
            Yb = np.abs(B.astype(np.complex128))
            Y = Yb if tr == "N" else Yb.T
            G = abs(al[0]) * ((X @ Y.T + Y @ X.T) if routine.endswith("2k") else X @ X.T) + abs(be[0]) * np.abs(C0)
            assert np.all(np.abs(C - R) <= 16 * (k + 2) * eps * G), (routine, form)
        elif routine in ("symm", "hemm"):
            side, uplo = ("L", "U") if form == 0 else ("R", "L")
            na = m if side == "L" else n
            A, B, C0 = _cplx(113, (na, na), dt), _cplx(114, (m, n), dt), _cplx(115, (m, n), dt)
            al, be = np.array([0.7 + 0.1j], dt), np.array([-0.3 + 0.9j], dt)
            C, R = C0.copy(order="F"), C0.copy(order="F")
            getattr(lib, prec + routine + "_")(gpu._c(side), gpu._c(uplo), gpu._i(m), gpu._i(n), p(al),
                                               p(A), gpu._i(na), p(B), gpu._i(m), p(be), p(C), gpu._i(m))
            getattr(ora, f"ora_{prec}{routine}")(c(side), c(uplo), m, n, p(al), p(A), na, p(B), m,
                                                  p(be), p(R), m)
            Aa, Ba = np.abs(A.astype(np.complex128)), np.abs(B.astype(np.complex128))
            F = np.maximum(Aa, Aa.T)
            G = abs(al[0]) * (F @ Ba if side == "L" else Ba @ F) + abs(be[0]) * np.abs(C0)
            assert np.all(np.abs(C - R) <= 16 * (na + 2) * eps * G), (routine, form)
        else:
            side, uplo, ta, diag = ("L", "L", "C", "N") if form == 0 else ("R", "U", "T", "U")
            na = m if side == "L" else n
            A = np.asfortranarray(_cplx(116, (na, na), dt) / na)
            A[np.diag_indices(na)] = 1.0 + 0.5j
            B0 = _cplx(117, (m, n), dt)
            al = np.array([0.6 - 0.8j], dt)
            X, R = B0.copy(order="F"), B0.copy(order="F")
            getattr(lib, prec + routine + "_")(gpu._c(side), gpu._c(uplo), gpu._c(ta), gpu._c(diag),
                                               gpu._i(m), gpu._i(n), p(al), p(A), gpu._i(na), p(X),
                                               gpu._i(m))
            getattr(ora, f"ora_{prec}{routine}")(c(side), c(uplo), c(ta), c(diag), m, n, p(al), p(A), na,
                                                  p(R), m)
            scale = np.max(np.abs(R)) if routine == "trsm" else np.max(np.abs(B0)) * 2
            assert np.max(np.abs(X - R)) <= 32 * (na + 2) * eps * scale, (routine, form)
