"""Level-2 conformance sweep in the manner of the netlib tester DBLAT2.

DBLAT2 (not shipped with the reference; the reference's tests/netlib holds
only DBLAT3) runs every Level-2 routine over a small parameter grid -- its
default data: N in {0, 1, 2, 3, 5, 9}, band widths K in {0, 1, 2, 4},
increments {1, 2, -1, -2}, ALPHA in {0, 1, 0.7}, BETA in {0, 1, 0.9} -- and
accepts a result when, element by element, |computed - expected| <= 16 eps
times the gauge sum |alpha| |A| |x| + |beta| |y| (DMVCH), and when every
argument that must not change is unchanged (its LDE / LDERES checks).  This
restatement runs the same grid on the device (cutoff 0, every call offloaded)
against the oracle, with the increments paired rather than crossed and the
(alpha, beta) pairs that select distinct code paths.  The triangular solves
use diagonally dominant matrices and are compared relatively.
"""
import itertools

import numpy as np
import pytest

from _oracle import c, p

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
NS = [0, 1, 2, 3, 5, 9]
KS = [0, 1, 2, 4]
INCS = [(1, 1), (2, -1), (-2, 2), (-1, -2)]
AB = [(0.0, 1.0), (0.0, 0.9), (1.0, 0.0), (0.7, 0.9), (1.0, 1.0), (0.7, 0.0)]
THRESH = 16.0


def _vec(n, inc, rng, fill=7.5):
    v = np.full(max(1, 1 + (n - 1) * abs(inc)), fill)
    vals = rng.uniform(-1, 1, n)
    if n:
        v[:: abs(inc)][:n] = vals if inc > 0 else vals[::-1]
    return v


def _get(v, n, inc):
    g = v[:: abs(inc)][:n]
    return g if inc > 0 else g[::-1]


def _gaps_same(a, b, n, inc):
    m = np.ones(len(a), bool)
    m[:: abs(inc)][:n] = False
    return np.array_equal(a[m], b[m])


def _check(got, ref, gauge, what):
    bound = THRESH * EPS * gauge + 1e-300
    bad = np.abs(got - ref) > bound
    assert not bad.any(), (what, np.max(np.abs(got - ref) / (EPS * gauge + 1e-300)))


def _band_dense(AB_, m, n, kl, ku):
    D = np.zeros((m, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            D[i, j] = AB_[ku + i - j, j]
    return D


def _pack(A, uplo):
    n = A.shape[0]
    if n == 0:
        return np.zeros(1)
    if uplo == "U":
        return np.concatenate([A[: j + 1, j] for j in range(n)])
    return np.concatenate([A[j:, j] for j in range(n)])


def test_mv_products(gpu, ora):
    """DGEMV, DGBMV, DSYMV, DSBMV, DSPMV over the grid."""
    rng = np.random.default_rng(150)
    gpu.reset_stats()
    calls = 0
    for n, (ix, iy), (al, be) in itertools.product(NS, INCS, AB):
        for m in (n, max(0, n - 1), n + 2):
            for tr in "NT":
                lda = max(1, m) + 1
                A = np.asfortranarray(rng.uniform(-1, 1, (lda, max(n, 1))))
                lx, ly = (n, m) if tr == "N" else (m, n)
                x, y0 = _vec(lx, ix, rng), _vec(ly, iy, rng)
                y, r, A0 = y0.copy(), y0.copy(), A.copy(order="F")
                gpu.dgemv(tr, m, n, al, A, lda, x, ix, be, y, iy)
                ora.ora_dgemv(c(tr), m, n, al, p(A), lda, p(x), ix, be, p(r), iy)
                Op = A[:m, :n] if tr == "N" else A[:m, :n].T
                g = abs(al) * np.abs(Op) @ np.abs(_get(x, lx, ix)) + abs(be) * np.abs(_get(y0, ly, iy))
                _check(_get(y, ly, iy), _get(r, ly, iy), g, ("dgemv", tr, m, n, ix, iy, al, be))
                assert np.array_equal(A, A0) and _gaps_same(y, y0, ly, iy)
                calls += 1
                for kl, ku in itertools.product(KS, KS):
                    ldb = kl + ku + 2
                    Bd = np.asfortranarray(rng.uniform(-1, 1, (ldb, max(n, 1))))
                    y, r = y0.copy(), y0.copy()
                    gpu.dgbmv(tr, m, n, kl, ku, al, Bd, ldb, x, ix, be, y, iy)
                    ora.ora_dgbmv(c(tr), m, n, kl, ku, al, p(Bd), ldb, p(x), ix, be, p(r), iy)
                    D = _band_dense(Bd, m, n, kl, ku)
                    Op = D if tr == "N" else D.T
                    g = abs(al) * np.abs(Op) @ np.abs(_get(x, lx, ix)) + abs(be) * np.abs(_get(y0, ly, iy))
                    _check(_get(y, ly, iy), _get(r, ly, iy), g, ("dgbmv", tr, m, n, kl, ku, ix, iy, al, be))
                    assert _gaps_same(y, y0, ly, iy)
                    calls += 1
        for uplo in "UL":
            lda = max(1, n) + 1
            A = np.asfortranarray(rng.uniform(-1, 1, (lda, max(n, 1))))
            x, y0 = _vec(n, ix, rng), _vec(n, iy, rng)
            T = np.triu(A[:n, :n]) if uplo == "U" else np.tril(A[:n, :n])
            S = T + T.T - np.diag(np.diag(T))
            g = abs(al) * np.abs(S) @ np.abs(_get(x, n, ix)) + abs(be) * np.abs(_get(y0, n, iy))
            y, r = y0.copy(), y0.copy()
            gpu.dsymv(uplo, n, al, A, lda, x, ix, be, y, iy)
            ora.ora_dsymv(c(uplo), n, al, p(A), lda, p(x), ix, be, p(r), iy)
            _check(_get(y, n, iy), _get(r, n, iy), g, ("dsymv", uplo, n, ix, iy, al, be))
            AP = _pack(A[:n, :n], uplo)
            y, r = y0.copy(), y0.copy()
            gpu.dspmv(uplo, n, al, AP, x, ix, be, y, iy)
            ora.ora_dspmv(c(uplo), n, al, p(AP), p(x), ix, be, p(r), iy)
            _check(_get(y, n, iy), _get(r, n, iy), g, ("dspmv", uplo, n, ix, iy, al, be))
            calls += 2
            for k in KS:
                ldb = k + 2
                Bd = np.asfortranarray(rng.uniform(-1, 1, (ldb, max(n, 1))))
                kl, ku = (0, k) if uplo == "U" else (k, 0)
                Tb = _band_dense(Bd, n, n, kl, ku)
                Sb = Tb + Tb.T - np.diag(np.diag(Tb))
                g = abs(al) * np.abs(Sb) @ np.abs(_get(x, n, ix)) + abs(be) * np.abs(_get(y0, n, iy))
                y, r = y0.copy(), y0.copy()
                gpu.dsbmv(uplo, n, k, al, Bd, ldb, x, ix, be, y, iy)
                ora.ora_dsbmv(c(uplo), n, k, al, p(Bd), ldb, p(x), ix, be, p(r), iy)
                _check(_get(y, n, iy), _get(r, n, iy), g, ("dsbmv", uplo, n, k, ix, iy, al, be))
                calls += 1
    assert calls > 5000
    st = gpu.stats()
    assert st["fallthrough"] == 0 and st["device_calls"] > 5000, st


def test_triangular(gpu, ora):
    """DTRMV, DTBMV, DTPMV (gauge |op(T)| |x|) and DTRSV, DTBSV, DTPSV
    (diagonally dominant, relative 1e-12) over the grid."""
    rng = np.random.default_rng(151)
    gpu.reset_stats()
    for n, (ix, _), uplo, tr, dg in itertools.product(NS, INCS, "UL", "NT", "NU"):
        lda = max(1, n) + 1
        A = np.asfortranarray(rng.uniform(-1, 1, (lda, max(n, 1))) * 0.25)
        if n:
            A[np.arange(n), np.arange(n)] = 2.0 + rng.uniform(0, 1, n)
        x0 = _vec(n, ix, rng)
        T = np.triu(A[:n, :n]) if uplo == "U" else np.tril(A[:n, :n])
        if dg == "U" and n:
            T[np.diag_indices(n)] = 1.0
        Op = T if tr == "N" else T.T
        g = np.abs(Op) @ np.abs(_get(x0, n, ix))
        AP = _pack(A[:n, :n], uplo)
        for name, fg, fo, args_g, args_o in (
                ("dtrmv", gpu.dtrmv, ora.ora_dtrmv, (A, lda), (p(A), lda)),
                ("dtpmv", gpu.dtpmv, ora.ora_dtpmv, (AP,), (p(AP),))):
            x, r = x0.copy(), x0.copy()
            fg(uplo, tr, dg, n, *args_g, x, ix)
            fo(c(uplo), c(tr), c(dg), n, *args_o, p(r), ix)
            _check(_get(x, n, ix), _get(r, n, ix), g, (name, uplo, tr, dg, n, ix))
            assert _gaps_same(x, x0, n, ix)
        for name, fg, fo, args_g, args_o in (
                ("dtrsv", gpu.dtrsv, ora.ora_dtrsv, (A, lda), (p(A), lda)),
                ("dtpsv", gpu.dtpsv, ora.ora_dtpsv, (AP,), (p(AP),))):
            x, r = x0.copy(), x0.copy()
            fg(uplo, tr, dg, n, *args_g, x, ix)
            fo(c(uplo), c(tr), c(dg), n, *args_o, p(r), ix)
            np.testing.assert_allclose(_get(x, n, ix), _get(r, n, ix), rtol=1e-12, atol=1e-13,
                                       err_msg=str((name, uplo, tr, dg, n, ix)))
            assert _gaps_same(x, x0, n, ix)
        for k in KS:
            kl, ku = (0, k) if uplo == "U" else (k, 0)
            ldb = k + 2
            Bd = np.asfortranarray(rng.uniform(-1, 1, (ldb, max(n, 1))) * 0.25)
            Bd[ku, :] = 2.0 + rng.uniform(0, 1, max(n, 1))
            Tb = _band_dense(Bd, n, n, kl, ku)
            if dg == "U" and n:
                Tb[np.diag_indices(n)] = 1.0
            Opb = Tb if tr == "N" else Tb.T
            x, r = x0.copy(), x0.copy()
            gpu.dtbmv(uplo, tr, dg, n, k, Bd, ldb, x, ix)
            ora.ora_dtbmv(c(uplo), c(tr), c(dg), n, k, p(Bd), ldb, p(r), ix)
            _check(_get(x, n, ix), _get(r, n, ix), np.abs(Opb) @ np.abs(_get(x0, n, ix)),
                   ("dtbmv", uplo, tr, dg, n, k, ix))
            x, r = x0.copy(), x0.copy()
            gpu.dtbsv(uplo, tr, dg, n, k, Bd, ldb, x, ix)
            ora.ora_dtbsv(c(uplo), c(tr), c(dg), n, k, p(Bd), ldb, p(r), ix)
            np.testing.assert_allclose(_get(x, n, ix), _get(r, n, ix), rtol=1e-12, atol=1e-13,
                                       err_msg=str(("dtbsv", uplo, tr, dg, n, k, ix)))
            assert _gaps_same(x, x0, n, ix)
    st = gpu.stats()
    assert st["fallthrough"] == 0 and st["device_calls"] > 1000, st


def test_rank_updates(gpu, ora):
    """DGER, DSYR, DSPR, DSYR2, DSPR2 over the grid; the gauge is
    |A| + |alpha| |x| |y|' per element; the unreferenced triangle and the
    rows beyond m (leading-dimension padding) must not change."""
    rng = np.random.default_rng(152)
    for n, (ix, iy), al in itertools.product(NS, INCS, [0.0, 1.0, 0.7]):
        for m in (n, max(0, n - 1), n + 2):
            lda = max(1, m) + 1
            A0 = np.asfortranarray(rng.uniform(-1, 1, (lda, max(n, 1))))
            x, y = _vec(m, ix, rng), _vec(n, iy, rng)
            A, R = A0.copy(order="F"), A0.copy(order="F")
            gpu.dger(m, n, al, x, ix, y, iy, A, lda)
            ora.ora_dger(m, n, al, p(x), ix, p(y), iy, p(R), lda)
            g = np.abs(A0[:m, :n]) + abs(al) * np.outer(np.abs(_get(x, m, ix)), np.abs(_get(y, n, iy)))
            _check(A[:m, :n], R[:m, :n], g, ("dger", m, n, ix, iy, al))
            assert np.array_equal(A[m:], A0[m:])
        lda = max(1, n) + 1
        A0 = np.asfortranarray(rng.uniform(-1, 1, (lda, max(n, 1))))
        x, y = _vec(n, ix, rng), _vec(n, iy, rng)
        xs, ys = np.abs(_get(x, n, ix)), np.abs(_get(y, n, iy))
        for uplo in "UL":
            other = np.tril_indices(n, -1) if uplo == "U" else np.triu_indices(n, 1)
            g1 = np.abs(A0[:n, :n]) + abs(al) * np.outer(xs, xs)
            g2 = np.abs(A0[:n, :n]) + abs(al) * (np.outer(xs, ys) + np.outer(ys, xs))
            A, R = A0.copy(order="F"), A0.copy(order="F")
            gpu.dsyr(uplo, n, al, x, ix, A, lda)
            ora.ora_dsyr(c(uplo), n, al, p(x), ix, p(R), lda)
            _check(A[:n, :n], R[:n, :n], g1, ("dsyr", uplo, n, ix, al))
            assert np.array_equal(A[:n, :n][other], A0[:n, :n][other]) and np.array_equal(A[n:], A0[n:])
            A, R = A0.copy(order="F"), A0.copy(order="F")
            gpu.dsyr2(uplo, n, al, x, ix, y, iy, A, lda)
            ora.ora_dsyr2(c(uplo), n, al, p(x), ix, p(y), iy, p(R), lda)
            _check(A[:n, :n], R[:n, :n], g2, ("dsyr2", uplo, n, ix, iy, al))
            assert np.array_equal(A[:n, :n][other], A0[:n, :n][other])
            P0 = _pack(A0[:n, :n], uplo)
            G1, G2 = _pack(g1, uplo), _pack(g2, uplo)
            P, Q = P0.copy(), P0.copy()
            gpu.dspr(uplo, n, al, x, ix, P)
            ora.ora_dspr(c(uplo), n, al, p(x), ix, p(Q))
            _check(P[: len(G1)], Q[: len(G1)], G1 if n else np.zeros(0), ("dspr", uplo, n, ix, al))
            P, Q = P0.copy(), P0.copy()
            gpu.dspr2(uplo, n, al, x, ix, y, iy, P)
            ora.ora_dspr2(c(uplo), n, al, p(x), ix, p(y), iy, p(Q))
            _check(P[: len(G2)], Q[: len(G2)], G2 if n else np.zeros(0), ("dspr2", uplo, n, ix, iy, al))
