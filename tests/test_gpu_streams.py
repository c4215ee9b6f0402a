"""Device C ABI on the caller's own streams.

A caller of the gpublas_dev_* entry points may pass any hipStream_t.  Calls
whose kernels need scratch memory (DGEMV 'N' column-panel partials, DSYMV
tile partials, the Level-1 reduction partials and tickets) get scratch that
belongs to the stream they run on, so two streams of one host thread never
share partials.  Here one thread alternates DGEMV, DSYMV and DDOT calls over
two streams of its own with nothing synchronised in between, and every
result must still equal the oracle's (integer data: bit-exact).
"""
import ctypes

import numpy as np
import pytest

from _oracle import c, p

pytestmark = pytest.mark.gpu

VP, CI, CD, CH = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipStreamCreate.argtypes = [ctypes.POINTER(VP)]
    h.hipStreamSynchronize.argtypes = [VP]
    h.hipStreamDestroy.argtypes = [VP]
    return h


def _dev(gb, a):
    lib = gb.library()
    d = lib.gpublas_dev_alloc(a.nbytes)
    assert d
    assert lib.gpublas_dev_h2d(VP(d), p(a), a.nbytes, None) == 0
    return d


def _get(gb, d, like):
    out = np.empty_like(like)
    lib = gb.library()
    assert lib.gpublas_dev_sync(None) == 0
    assert lib.gpublas_dev_d2h(p(out), VP(d), out.nbytes, None) == 0
    assert lib.gpublas_dev_sync(None) == 0
    return out


def test_two_user_streams_one_thread(gpu, ora, hip):
    gb = gpu
    lib = gb.library()
    lib.gpublas_dev_dgemv.argtypes = [VP, CH, CI, CI, CD, VP, CI, VP, CI, CD, VP, CI]
    lib.gpublas_dev_dsymv.argtypes = [VP, CH, CI, CD, VP, CI, VP, CI, CD, VP, CI]
    lib.gpublas_dev_ddot.argtypes = [VP, CI, VP, CI, VP, CI, VP]
    lib.gpublas_dev_host_alloc.restype = VP
    lib.gpublas_dev_host_alloc.argtypes = [ctypes.c_size_t]
    lib.gpublas_dev_host_free.argtypes = [VP]
    rng = np.random.default_rng(12)
    m, n = 2048, 4096  # 4 row blocks x 128 column panels: the DGEMV 'N' partials are used
    A = np.asfortranarray(rng.integers(-3, 4, size=(m, n)).astype(float))
    S = rng.integers(-3, 4, size=(m, m)).astype(float)
    S = np.asfortranarray(np.tril(S) + np.tril(S, -1).T)
    xs = [rng.integers(-3, 4, size=n).astype(float) for _ in range(2)]
    us = [rng.integers(-3, 4, size=m).astype(float) for _ in range(2)]
    nd = 100000  # longer than one workgroup's share: the reduction uses partials and a ticket
    ds = [rng.integers(-3, 4, size=nd).astype(float) for _ in range(3)]
    dA, dS = _dev(gb, A), _dev(gb, S)
    dx = [_dev(gb, x) for x in xs]
    du = [_dev(gb, u) for u in us]
    dd = [_dev(gb, d) for d in ds]
    dy = [_dev(gb, np.zeros(m)) for _ in range(2)]
    dz = [_dev(gb, np.zeros(m)) for _ in range(2)]
    res = lib.gpublas_dev_host_alloc(64)
    hres = (ctypes.c_double * 8).from_address(res)
    streams = [VP(), VP()]
    for s in streams:
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    lib.gpublas_dev_sync(None)
    try:
        dots = []
        for rep in range(6):
            for i, s in enumerate(streams):
                assert lib.gpublas_dev_dgemv(s, b"N", m, n, 1.0, VP(dA), m, VP(dx[i]), 1, 0.0,
                                             VP(dy[i]), 1) == 0
                assert lib.gpublas_dev_dsymv(s, b"L", m, 2.0, VP(dS), m, VP(du[i]), 1, 0.0,
                                             VP(dz[i]), 1) == 0
                assert lib.gpublas_dev_ddot(s, nd, VP(dd[i]), 1, VP(dd[2]), 1,
                                            VP(res + 8 * (2 * (rep % 2) + i))) == 0
            if rep % 2 == 1:
                for s in streams:
                    assert hip.hipStreamSynchronize(s) == 0
                dots.append([hres[j] for j in range(4)])
        for s in streams:
            assert hip.hipStreamSynchronize(s) == 0
        for i in range(2):
            y = _get(gb, dy[i], np.zeros(m))
            z = _get(gb, dz[i], np.zeros(m))
            yr, zr = np.zeros(m), np.zeros(m)
            ora.ora_dgemv(c("N"), m, n, 1.0, p(A), m, p(xs[i]), 1, 0.0, p(yr), 1)
            ora.ora_dsymv(c("L"), m, 2.0, p(S), m, p(us[i]), 1, 0.0, p(zr), 1)
            assert np.array_equal(y, yr), i
            assert np.array_equal(z, zr), i
        want = [float(ds[0] @ ds[2]), float(ds[1] @ ds[2])] * 2
        for got in dots:
            assert got == want
    finally:
        for s in streams:
            hip.hipStreamSynchronize(s)
            hip.hipStreamDestroy(s)
        lib.gpublas_dev_host_free(VP(res))
        for d in [dA, dS] + dx + du + dd + dy + dz:
            lib.gpublas_dev_release(VP(d))
