"""Split-precision DGEMM / DSYRK (GPUBLAS_OPTIONS split=<s>) against the
FP64 oracle.

Bars:
  * integer inputs that fit one slice exactly (|a| <= 255 per row scale) give
    results identical to the exact integer product -- this pins the slice
    arithmetic, the power-of-two scalings and the exactness of the FP32
    accumulation inside the BF16 MFMA (any rounding there would show);
  * uniform[-1,1] inputs with s = 6: normwise relative error
    ||C - Cref||_F / ||Cref||_F <= 1e-13 and max elementwise error within
    1e-13 * max|Cref|, for every trans combination;
  * rows of A and columns of B scaled by 2^U[-20,20] (the wide-range case the
    per-row/column exponents exist for): the same 1e-13 normwise bound;
  * fewer slices give a larger error (monotone in s), so the bound is not met
    by accident;
  * a NaN input falls back to the FP64 kernel (NaN propagates like DGEMM);
  * BF16 slices (pre-scaled by 2^-8p) and FP16 slices (the integers, each
    diagonal's power-of-two weight applied in FP64) give bit-identical
    results: the weights are powers of two and no partial sum is subnormal.
Every test runs with both slice formats.
"""
import numpy as np
import pytest

from _oracle import c, p, uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["bf16", "fp16"])
def split(gpu, request):
    """Split mode on, with either slice format (every test runs for both)."""
    def use(s):
        assert gpu.set_options(f"split={s};slice={request.param}") == 0
    yield use
    gpu.set_options("split=0;slice=bf16")


def _ref(ora, ta, tb, m, n, k, alpha, A, B, beta, C0):
    R = C0.copy(order="F")
    ora.ora_dgemm(c(ta), c(tb), m, n, k, alpha, p(A), A.shape[0], p(B), B.shape[0], beta, p(R), m)
    return R


def test_split_integer_inputs_are_exact(gpu, split, ora):
    rng = np.random.default_rng(11)
    m, n, k = 300, 260, 700
    A = np.asfortranarray(rng.integers(-255, 256, size=(m, k)).astype(float))
    B = np.asfortranarray(rng.integers(-255, 256, size=(k, n)).astype(float))
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
    split(6)
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
    split(6)
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
    for s in (2, 3, 4, 6):
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
    split(6)
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
    split(6)
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
