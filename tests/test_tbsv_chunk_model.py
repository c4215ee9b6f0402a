"""The mathematics of the chunked band triangular solve, checked on the CPU.

The GPU DTBSV for k <= 63 (libgpublas_amd/csrc/dev/blas2b.hip) cuts the solve
positions into chunks of L and relies on three facts, restated here in numpy
with dense triangular solves and checked against the oracle's sequential
netlib DTBSV (the column sweep of reference blas_level2/tbsv.cc's host
fall-through):
  * a chunk's unknowns are affine in its incoming state z_c (the last k
    unknowns before it): x_c = y_c + G_c z_c, y_c the solve from a zero state
    and G_c the responses to unit states;
  * the states follow z_{c+1} = h_c + H_c z_c (h_c, H_c the last k rows of
    y_c, G_c), which may be evaluated in two levels: maps composed within
    groups of chunks, one pass over the group maps, then each group filled in;
  * re-solving a chunk with z_c taken off its first k equations gives x_c.
The growth guard (a response outside [-1, 1] declines the chunked result) is
checked to pass a slowly decaying diagonally dominant band and to decline the integer band
with a unit diagonal and -1 below it, whose responses grow without bound.
"""
import math

import numpy as np
import pytest
from scipy.linalg import solve_triangular

import _oracle
from _oracle import c, p, uniform


@pytest.fixture(scope="module")
def L():
    return _oracle.load()


def _band(D, uplo, k):
    """netlib band storage of the triangle of D (lda = k + 1)."""
    n = D.shape[0]
    AB = np.zeros((k + 1, n), order="F")
    for j in range(n):
        rows = range(j, min(n, j + k + 1)) if uplo == "L" else range(max(0, j - k), j + 1)
        for i in rows:
            AB[(i - j) if uplo == "L" else (k + i - j), j] = D[i, j]
    return AB


def _solve_order(T, uplo, trans):
    """op(T) permuted so that the solve runs forward: a lower triangular M and
    the permutation P (M = op(T)[P][:, P])."""
    n = T.shape[0]
    Op = T if trans == "N" else T.T
    fwd = (uplo == "L") == (trans == "N")
    P = np.arange(n) if fwd else np.arange(n)[::-1]
    return Op[np.ix_(P, P)], P


def _chunk_maps(M, b, k, Lc):
    """Per chunk: the zero-state solve y_c, the unit responses G_c, and the
    state map (h_c, H_c) = last k rows of (y_c, G_c)."""
    n = M.shape[0]
    out = []
    for s in range(0, n, Lc):
        e = min(n, s + Lc)
        Mc = M[s:e, s:e]
        y = solve_triangular(Mc, b[s:e], lower=True)
        # column j: the response to a unit in state position j = unknown s - k + j
        B = np.zeros((e - s, k))
        for j in range(k):
            col = s - k + j
            if col >= 0:
                B[:, j] = -M[s:e, col]
        G = solve_triangular(Mc, B, lower=True)
        out.append((s, e, y, G, y[-k:], G[-k:, :]))
    return out


def _states_flat(maps, k):
    z = [np.zeros(k)]
    for (_, _, _, _, h, H) in maps[:-1]:
        z.append(h + H @ z[-1])
    return z


def _states_two_level(maps, k):
    """Groups of g ~ sqrt(C - 1) maps composed, one pass over the groups,
    then each group's states filled in from its incoming state."""
    C = len(maps)
    g = max(1, math.ceil(math.sqrt(C - 1)))
    trans = [(m[4], m[5]) for m in maps[:-1]]  # the C - 1 transitions
    groups = [trans[i:i + g] for i in range(0, len(trans), g)]
    comp = []
    for grp in groups:  # compose: z_out = a + A z_in
        a, A = np.zeros(k), np.eye(k)
        for (h, H) in grp:
            a, A = h + H @ a, H @ A
        comp.append((a, A))
    zg = [np.zeros(k)]
    for (a, A) in comp:
        zg.append(a + A @ zg[-1])
    z = []
    for gi, grp in enumerate(groups):
        zz = zg[gi]
        z.append(zz)
        for (h, H) in grp[:-1]:
            zz = h + H @ zz
            z.append(zz)
    z.append(zg[-1])
    assert len(z) == C
    return z


def _resolve(M, b, k, maps, z):
    x = np.empty_like(b)
    for (s, e, _, _, _, _), zc in zip(maps, z):
        r = b[s:e].copy()
        for j in range(k):
            col = s - k + j
            if col >= 0:
                r -= M[s:e, col] * zc[j]
        x[s:e] = solve_triangular(M[s:e, s:e], r, lower=True)
    return x


def _declined(maps):
    return any(not np.all(np.abs(G) <= 1.0) for (_, _, _, G, _, _) in maps)


def _slow_decay(n, k, uplo, seed):
    """An M-matrix band: off-diagonals in (-1/k, -0.9/k], diagonal in
    [1, 1.02).  Every row and column of |off-diagonals| sums below the
    diagonal, so unit responses stay in [0, 1] (the guard passes), yet they
    decay slowly (about 0.94 a step): the state maps are far from zero and a
    dropped term in them shows."""
    u = (uniform(seed, (n, k + 1)) + 1.0) / 2.0
    T = np.zeros((n, n))
    for d in range(1, k + 1):
        v = -(0.9 + 0.1 * u[:n - d, d]) / k
        if uplo == "L":
            T[np.arange(d, n), np.arange(n - d)] = v
        else:
            T[np.arange(n - d), np.arange(d, n)] = v
    T[np.diag_indices(n)] = 1.0 + 0.02 * u[:, 0]
    return T


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "N"), ("L", "T"), ("U", "T")])
@pytest.mark.parametrize("n,k,Lc", [(700, 3, 64), (3000, 5, 64), (2500, 16, 128), (1000, 63, 64)])
def test_chunked_solve_matches_sequential(L, uplo, trans, n, k, Lc):
    T = _slow_decay(n, k, uplo, 900 + n + k)
    xs = uniform(901 + k, (n,))
    Op = T if trans == "N" else T.T
    b = Op @ xs
    want, AB = b.copy(), _band(T, uplo, k)
    L.ora_dtbsv(c(uplo), c(trans), c("N"), n, k, p(AB), k + 1, p(want), 1)
    np.testing.assert_allclose(want, xs, rtol=1e-12, atol=1e-12)

    M, P = _solve_order(T, uplo, trans)
    bp = b[P]
    maps = _chunk_maps(M, bp, k, Lc)
    assert not _declined(maps)
    # affine in the incoming state: x_c = y_c + G_c z_c with the true state
    xp = want[P]
    for (s, e, y, G, _, _) in maps:
        zc = np.array([xp[s - k + j] if s - k + j >= 0 else 0.0 for j in range(k)])
        np.testing.assert_allclose(y + G @ zc, xp[s:e], rtol=1e-12, atol=1e-12)
    zf = _states_flat(maps, k)
    z2 = _states_two_level(maps, k)
    for a, b2 in zip(zf, z2):
        np.testing.assert_allclose(a, b2, rtol=1e-13, atol=1e-14)
    x = np.empty(n)
    x[P] = _resolve(M, bp, k, maps, z2)
    np.testing.assert_allclose(x, want, rtol=1e-12, atol=1e-12)


def test_planted_wrong_state_is_caught(L):
    """A dropped map term in the propagation (z_{c+1} = h_c instead of
    h_c + H_c z_c) changes the result: the comparison above has teeth."""
    n, k, Lc = 700, 3, 64
    T = _slow_decay(n, k, "L", 77)
    xs = uniform(78, (n,))
    b = T @ xs
    M, P = _solve_order(T, "L", "N")
    maps = _chunk_maps(M, b, k, Lc)
    z = [np.zeros(k)] + [m[4] for m in maps[:-1]]
    x = _resolve(M, b, k, maps, z)
    assert np.max(np.abs(x - xs)) > 1e-6


@pytest.mark.parametrize("n,k", [(640, 1), (640, 7)])
def test_growth_guard_declines_unit_minus_one_band(L, n, k):
    """T with a unit diagonal and -1 on the k subdiagonals: the responses of a
    chunk grow (for k = 1 they are all 1; wider bands reach far past 1), so
    the guard declines for k > 1 and the exact integer solve stays with the
    sequential sweep."""
    T = np.eye(n)
    for d in range(1, k + 1):
        T[np.arange(d, n), np.arange(n - d)] = -1.0
    M, P = _solve_order(T, "L", "N")
    xs = np.zeros(n)
    xs[:8] = [1, -1, 2, 0, 1, -2, 1, 1]
    b = T @ xs
    maps = _chunk_maps(M, b, k, 64)
    assert _declined(maps) == (k > 1)
    want, AB = b.copy(), _band(T, "L", k)
    L.ora_dtbsv(c("L"), c("N"), c("N"), n, k, p(AB), k + 1, p(want), 1)
    if k == 1:
        assert np.array_equal(want, xs)
