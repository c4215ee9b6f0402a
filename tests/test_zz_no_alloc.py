"""The allocation-free DGEMV entry point (gpublas_dev_dgemv_noalloc) against
the ordinary one (gpublas_dev_dgemv): the same kernels on the same inputs,
so the results must agree bit for bit.  Two calls in a row per case, the
second smaller, so the fixed scratch array is reused with stale contents
from the first; the 'N' shapes use the column-panel partials (ncb > 1).
Calls go through the C ABI on a stream the test creates."""
import ctypes

import numpy as np
import pytest

from _oracle import p, uniform

pytestmark = pytest.mark.gpu

VP, CI, CD, CH = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char


def _setup(gb):
    lib = gb.library()
    for f in (lib.gpublas_dev_dgemv, lib.gpublas_dev_dgemv_noalloc):
        f.argtypes = [VP, CH, CI, CI, CD, VP, CI, VP, CI, CD, VP, CI]
    lib.gpublas_dev_alloc.restype = VP
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(VP)]
    hip.hipStreamSynchronize.argtypes = [VP]
    hip.hipStreamDestroy.argtypes = [VP]
    return lib, hip


def _up(lib, a):
    d = lib.gpublas_dev_alloc(a.nbytes)
    assert d
    assert lib.gpublas_dev_h2d(VP(d), p(a), a.nbytes, None) == 0
    return d


def _down(lib, d, n):
    out = np.empty(n)
    assert lib.gpublas_dev_sync(None) == 0
    assert lib.gpublas_dev_d2h(p(out), VP(d), out.nbytes, None) == 0
    assert lib.gpublas_dev_sync(None) == 0
    return out


@pytest.mark.parametrize("trans,shapes", [
    ("N", [(3000, 5000), (700, 900)]),      # 6 x 86 then 2 x 57 panels
    ("N", [(8192, 8192), (4096, 1000)]),    # the chain's shape, then a smaller one
    ("T", [(2000, 3000), (513, 77)]),
])
def test_noalloc_dgemv_matches_dgemv(gpu, trans, shapes):
    lib, hip = _setup(gpu)
    assert lib.gpublas_dev_init() == 0
    s = VP()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    keep = []
    try:
        for call, (m, n) in enumerate(shapes):
            A = uniform(70 + call, (m, n))
            lenx, leny = (n, m) if trans == "N" else (m, n)
            x = uniform(80 + call, (lenx,))
            y0 = uniform(90 + call, (leny,))
            dA, dx = _up(lib, A), _up(lib, x)
            dy1, dy2 = _up(lib, y0), _up(lib, y0)
            keep += [dA, dx, dy1, dy2]
            assert lib.gpublas_dev_dgemv(s, trans.encode(), m, n, 1.1, VP(dA), m, VP(dx), 1, 0.5,
                                         VP(dy1), 1) == 0
            assert lib.gpublas_dev_dgemv_noalloc(s, trans.encode(), m, n, 1.1, VP(dA), m, VP(dx),
                                                 1, 0.5, VP(dy2), 1) == 0
            assert hip.hipStreamSynchronize(s) == 0
            y1, y2 = _down(lib, dy1, leny), _down(lib, dy2, leny)
            assert np.array_equal(y1, y2), (call, m, n)
            ref = 1.1 * (A if trans == "N" else A.T) @ x + 0.5 * y0
            assert np.allclose(y2, ref, rtol=1e-12, atol=1e-10)
    finally:
        hip.hipStreamSynchronize(s)
        hip.hipStreamDestroy(s)
        for d in keep:
            lib.gpublas_dev_release(VP(d))
