"""Split-precision DGEMM / DSYRK (GPUBLAS_OPTIONS split=<s>) against the
FP64 oracle.

Bars:
  * integer inputs that fit two slices exactly (|a| <= 255 per row scale) give
    results identical to the exact integer product -- this pins the slice
    arithmetic, the power-of-two scalings and the exactness of the MFMA
    accumulation (FP32 for BF16 / FP16 slices, int32 for INT8 slices: any
    rounding there would show);
  * uniform[-1,1] inputs with the format's full slice count (6 BF16 / FP16
    slices of 8 bits, 7 INT8 slices of 7 bits): normwise relative error
    ||C - Cref||_F / ||Cref||_F <= 1e-13 and max elementwise error within
    1e-13 * max|Cref|, for every trans combination;
  * rows of A and columns of B scaled by 2^U[-20,20] (the wide-range case the
    per-row/column exponents exist for): the same 1e-13 normwise bound;
  * fewer slices give a larger error (monotone in s), so the bound is not met
    by accident;
  * a NaN input falls back to the FP64 kernel (NaN propagates like DGEMM);
  * BF16 slices (pre-scaled by 2^-8p) and FP16 slices (the integers, each
    diagonal's power-of-two weight applied in FP64) give bit-identical
    results: the weights are powers of two and no partial sum is subnormal;
  * INT8 with k beyond one launch's exact int32 range (several launches
    accumulating into C) keeps the 1e-13 bound; at 8192^3 (both input
    distributions of the split-precision workload) INT8 with 7 slices stays
    within 1e-13 of the host BLAS's FP64 product.
Every test runs with all three slice formats.
"""
import numpy as np
import pytest

from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu

FULL = {"bf16": 6, "fp16": 6, "int8": 7}  # slices for <= 1e-13


class _Split:
    def __init__(self, gpu, fmt):
        self.gpu, self.fmt, self.full = gpu, fmt, FULL[fmt]

    def __call__(self, s):
        assert self.gpu.set_options(f"split={s};slice={self.fmt}") == 0


@pytest.fixture(params=["bf16", "fp16", "int8"])
def split(gpu, request):
    """Split mode on, with each slice format in turn; split(s) sets s slices,
    split.full is the format's slice count for FP64-level accuracy."""
    yield _Split(gpu, request.param)
    gpu.set_options("split=0;slice=bf16")


def _ref(ora, ta, tb, m, n, k, alpha, A, B, beta, C0):
    R = C0.copy(order="F")
    ora.ora_dgemm(c(ta), c(tb), m, n, k, alpha, p(A), A.shape[0], p(B), B.shape[0], beta, p(R), m)
    return R


def test_split_integer_inputs_are_exact(gpu, split, ora):
    """Integers that fit the first slice (8-bit BF16 / FP16 slices: |a| <= 255,
    7-bit INT8 slices: |a| <= 127): every later slice is zero, so the two-slice
    product is the exact integer product."""
    rng = np.random.default_rng(11)
    m, n, k = 300, 260, 700
    hi = 127 if split.fmt == "int8" else 255
    A = np.asfortranarray(rng.integers(-hi, hi + 1, size=(m, k)).astype(float))
    B = np.asfortranarray(rng.integers(-hi, hi + 1, size=(k, n)).astype(float))
    C = np.zeros((m, n), order="F")
    split(2)
    gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(float)
    assert np.array_equal(C, exact)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_split_uniform_error_bound(gpu, split, ora, ta, tb):
    m, n, k = 384, 320, 1000
    A = uniform(31, (m, k) if ta == "N" else (k, m))
    B = uniform(32, (k, n) if tb == "N" else (n, k))
    C0 = uniform(33, (m, n))
    C = C0.copy(order="F")
    split(split.full)
    gpu.dgemm(ta, tb, m, n, k, 1.0, A, A.shape[0], B, B.shape[0], 0.5, C, m)
    R = _ref(ora, ta, tb, m, n, k, 1.0, A, B, 0.5, C0)
    assert np.linalg.norm(C - R) / np.linalg.norm(R) <= 1e-13
    assert np.max(np.abs(C - R)) <= 1e-13 * np.max(np.abs(R))


def test_split_wide_range_rows_and_columns(gpu, split, ora):
    m, n, k = 256, 256, 512
    ra = 2.0 ** np.floor(uniform(41, (m,)) * 20.0)  # per-row scale 2^U[-20,20]
    cb = 2.0 ** np.floor(uniform(42, (n,)) * 20.0)  # per-column scale
    A = np.asfortranarray(uniform(43, (m, k)) * ra[:, None])
    B = np.asfortranarray(uniform(44, (k, n)) * cb[None, :])
    C = np.zeros((m, n), order="F")
    split(split.full)
    gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    R = _ref(ora, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"))
    assert np.linalg.norm(C - R) / np.linalg.norm(R) <= 1e-13
    # elementwise relative to each entry's own scale (row scale x column scale)
    scale = np.outer(ra, cb) * np.sqrt(k)
    assert np.max(np.abs(C - R) / scale) <= 1e-13


def test_split_error_decreases_with_slices(gpu, split, ora):
    m = n = k = 256
    A, B = uniform(51, (m, k)), uniform(52, (k, n))
    R = _ref(ora, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"))
    errs = []
    for s in (2, 3, 4, split.full):
        C = np.zeros((m, n), order="F")
        split(s)
        gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        errs.append(np.linalg.norm(C - R) / np.linalg.norm(R))
    assert errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[0] > 1e-6  # two slices really are ~16-bit accurate
    assert errs[3] < 1e-13


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_split_dsyrk(gpu, split, ora, uplo, trans):
    n, k = 300, 500
    A = uniform(61, (n, k) if trans == "N" else (k, n))
    C0 = uniform(62, (n, n))
    C, R = C0.copy(order="F"), C0.copy(order="F")
    split(split.full)
    gpu.dsyrk(uplo, trans, n, k, 0.8, A, A.shape[0], 0.3, C, n)
    ora.ora_dsyrk(c(uplo), c(trans), n, k, 0.8, p(A), A.shape[0], 0.3, p(R), n)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(C[~tri], C0[~tri])
    assert np.linalg.norm((C - R)[tri]) / np.linalg.norm(R[tri]) <= 1e-13


def test_split_nan_falls_back_to_fp64(gpu, split):
    m = n = k = 64
    A, B = uniform(71, (m, k)), uniform(72, (k, n))
    A[3, 5] = np.nan
    C = np.zeros((m, n), order="F")
    split(split.full)
    gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.all(np.isnan(C[3, :])) and np.all(np.isfinite(np.delete(C, 3, axis=0)))


def test_split_fp16_matches_bf16_bitwise(gpu):
    m, n, k = 256, 192, 900
    A = uniform(71, (m, k))
    B = uniform(72, (k, n))
    C0 = uniform(73, (m, n))
    out = {}
    try:
        for fmt in ("bf16", "fp16"):
            assert gpu.set_options(f"split=6;slice={fmt}") == 0
            C = C0.copy(order="F")
            gpu.dgemm("N", "N", m, n, k, 1.3, A, m, B, k, -0.7, C, m)
            out[fmt] = C
    finally:
        gpu.set_options("split=0;slice=bf16")
    assert np.array_equal(out["bf16"], out["fp16"])


def test_split_int8_two_digit_integers_need_the_cross_diagonal(gpu):
    """8-bit integers take two 7-bit INT8 slices (a = 2 q0 + q1 / 64, q1 in
    {0, +-64}); three slices keep the pair (1, 1) of diagonal 2, so the
    product is exact, while two slices drop it (an error of exactly
    sum q1 q1' / 4096 per entry, nonzero here)."""
    rng = np.random.default_rng(12)
    m, n, k = 200, 150, 600
    A = np.asfortranarray(rng.integers(-255, 256, size=(m, k)).astype(float))
    B = np.asfortranarray(rng.integers(-255, 256, size=(k, n)).astype(float))
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(float)
    out = {}
    try:
        for s in (2, 3):
            assert gpu.set_options(f"split={s};slice=int8") == 0
            C = np.zeros((m, n), order="F")
            gpu.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
            out[s] = C
    finally:
        gpu.set_options("split=0;slice=bf16")
    assert np.array_equal(out[3], exact)
    # the dropped pair (1, 1): the scaled digits are q1 = 64 * (a mod 2) * sign(a)
    ea = np.ceil(np.log2(np.max(np.abs(A), axis=1) + 1))[:, None]
    eb = np.ceil(np.log2(np.max(np.abs(B), axis=0) + 1))[None, :]
    assert np.all(ea == 8) and np.all(eb == 8)
    q1a = np.sign(A) * (np.abs(A) % 2)
    q1b = np.sign(B) * (np.abs(B) % 2)
    assert np.array_equal(exact - out[2], q1a @ q1b)


def test_split_int8_long_k_several_launches(gpu, ora):
    """k = 20000 with 8 slices: one launch keeps int32 exact over at most
    2^31 / (8 * 32 * 127^2) = 520 slabs of 32 (16,640 k), so the product runs
    as two launches, the second accumulating into C."""
    m, n, k = 96, 80, 20000
    A, B, C0 = uniform(81, (m, k)), uniform(82, (k, n)), uniform(83, (m, n))
    C = C0.copy(order="F")
    try:
        assert gpu.set_options("split=8;slice=int8") == 0
        gpu.dgemm("N", "N", m, n, k, 0.9, A, m, B, k, -0.4, C, m)
    finally:
        gpu.set_options("split=0;slice=bf16")
    R = _ref(ora, "N", "N", m, n, k, 0.9, A, B, -0.4, C0)
    assert np.linalg.norm(C - R) / np.linalg.norm(R) <= 1e-13


@pytest.mark.parametrize("dist", ["uniform", "wide"])
def test_split_int8_8192_cube(gpu, dist):
    """The split-precision workload at its size, M = N = K = 8192: INT8 with 7
    slices, (a) uniform[-1,1] and (b) sign-random magnitudes 2^U[-20,20] per
    row of A and per column of B; normwise error <= 1e-13 against the FP64
    product of the host BLAS (NumPy's matmul, whose own error is ~1e-16)."""
    n = 8192
    A, B = uniform(91, (n, n)), uniform(92, (n, n))
    if dist == "wide":
        A *= (2.0 ** (uniform(93, (n,)) * 20.0))[:, None]
        B *= (2.0 ** (uniform(94, (n,)) * 20.0))[None, :]
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    C = np.zeros((n, n), order="F")
    try:
        assert gpu.set_options("split=7;slice=int8") == 0
        gpu.dgemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n)
    finally:
        gpu.set_options("split=0;slice=bf16")
    R = A @ B
    err = np.linalg.norm(C - R) / np.linalg.norm(R)
    assert err <= 1e-13, err
