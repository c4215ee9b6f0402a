"""libgpublas_amd -- Python handle on the MI355X drop-in BLAS library.

The product is the C-ABI shared library ``lib/libgpublas.so`` (and
``lib/libgpublas_preload.so`` for ``LD_PRELOAD``).  This module loads it with
ctypes and mirrors its Fortran-ABI entry points with the reference's
argument order and meaning (``blas.h:202-314`` of Prince781/libgpublas; the
Level-1/2 entries follow the netlib definitions the reference's CBLAS specs
forward to), so Python callers and the test-suite drive exactly the symbols an
interposed application calls.  Arrays are NumPy arrays in column-major
(Fortran) layout; outputs are updated in place, as in Fortran.

Loading fails loudly when the library has not been built: nothing here falls
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
BIN_DIR = os.path.join(_HERE, "bin")
LIB_PATH = os.path.join(LIB_DIR, "libgpublas.so")
PRELOAD_PATH = os.path.join(LIB_DIR, "libgpublas_preload.so")
CHAIN_BENCH = os.path.join(BIN_DIR, "chain_bench")
SERVE_BENCH = os.path.join(BIN_DIR, "serve_bench")

_lib = None


class GpublasError(RuntimeError):
    pass


def library():
    """The loaded libgpublas.so (ctypes.CDLL); raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpublasError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C libgpublas_amd/csrc)")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(lib)
        _lib = lib
    return _lib


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "calls", "device_calls", "fallthrough", "hits", "misses", "h2d_bytes", "d2h_bytes",
        "tracked_objects", "tracked_bytes", "faults", "uploads", "allocs_seen",
        "async_returns", "evictions", "served", "batches", "max_batch", "deferred",
        "over_budget", "resident_bytes", "serve_wait_ns", "serve_run_ns", "serve_launch_ns",
        "resident", "io_syncs", "handler_reinstalls")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def _declare(lib):
    vp, ci, cd, cf, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_size_t
    lib.gpublas_malloc.restype = vp
    lib.gpublas_malloc.argtypes = [sz]
    lib.gpublas_calloc.restype = vp
    lib.gpublas_calloc.argtypes = [sz, sz]
    lib.gpublas_free.argtypes = [vp]
    lib.gpublas_is_tracked.argtypes = [vp]
    lib.gpublas_device_ptr.restype = vp
    lib.gpublas_device_ptr.argtypes = [vp]
    lib.gpublas_set_options.argtypes = [ctypes.c_char_p]
    lib.gpublas_set_cutoff.argtypes = [ci]
    lib.gpublas_set_crossover.argtypes = [ctypes.c_char_p, cd, cd]
    lib.gpublas_get_crossover.argtypes = [ctypes.c_char_p, ctypes.POINTER(cd), ctypes.POINTER(cd)]
    lib.gpublas_crossover_gpu.argtypes = [ctypes.c_char_p, cd, cd]
    lib.gpublas_tune_crossover.argtypes = [ctypes.c_char_p, ci]
    lib.gpublas_get_stats.argtypes = [ctypes.POINTER(Stats)]
    lib.gpublas_kernel_time.argtypes = [ctypes.c_char_p, ctypes.POINTER(cd)]
    lib.gpublas_last_xerbla.argtypes = [ctypes.c_char_p]
    lib.ddot_.restype = cd
    lib.dnrm2_.restype = cd
    lib.idamax_.restype = ci
    lib.idamin_.restype = ci
    lib.cblas_idamin.restype = ci
    lib.cblas_idamin.argtypes = [ci, vp, ci]
    lib.dasum_.restype = cd
    lib.cblas_dasum.restype = cd
    lib.cblas_dasum.argtypes = [ci, vp, ci]
    lib.cblas_ddot.restype = cd
    lib.cblas_ddot.argtypes = [ci, vp, ci, vp, ci]
    lib.cblas_dnrm2.restype = cd
    lib.cblas_dnrm2.argtypes = [ci, vp, ci]
    lib.cblas_idamax.restype = ci
    lib.cblas_idamax.argtypes = [ci, vp, ci]
    lib.cblas_daxpy.argtypes = [ci, cd, vp, ci, vp, ci]
    lib.cblas_dgemm.argtypes = [ci, ci, ci, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_sgemm.argtypes = [ci, ci, ci, ci, ci, ci, cf, vp, ci, vp, ci, cf, vp, ci]
    lib.cblas_zgemm.argtypes = [ci, ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
    lib.cblas_cgemm.argtypes = [ci, ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
    for p_, r_ in (("z", cd), ("c", cf)):
        getattr(lib, f"cblas_{p_}syrk").argtypes = [ci, ci, ci, ci, ci, vp, vp, ci, vp, vp, ci]
        getattr(lib, f"cblas_{p_}herk").argtypes = [ci, ci, ci, ci, ci, r_, vp, ci, r_, vp, ci]
        getattr(lib, f"cblas_{p_}syr2k").argtypes = [ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
        getattr(lib, f"cblas_{p_}her2k").argtypes = [ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, r_, vp, ci]
        for op in ("symm", "hemm"):
            getattr(lib, f"cblas_{p_}{op}").argtypes = [ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
        for op in ("trmm", "trsm"):
            getattr(lib, f"cblas_{p_}{op}").argtypes = [ci, ci, ci, ci, ci, ci, ci, vp, vp, ci, vp, ci]
    lib.cblas_dgemv.argtypes = [ci, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_dger.argtypes = [ci, ci, ci, cd, vp, ci, vp, ci, vp, ci]
    lib.cblas_dsymv.argtypes = [ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_dsyr.argtypes = [ci, ci, ci, cd, vp, ci, vp, ci]
    lib.cblas_dsyrk.argtypes = [ci, ci, ci, ci, ci, cd, vp, ci, cd, vp, ci]
    lib.cblas_dtrsm.argtypes = [ci, ci, ci, ci, ci, ci, ci, cd, vp, ci, vp, ci]
    lib.cblas_dsymm.argtypes = [ci, ci, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_dtrsv.argtypes = [ci, ci, ci, ci, ci, vp, ci, vp, ci]
    lib.gpublas_dev_alloc.restype = vp
    lib.gpublas_dev_alloc.argtypes = [sz]
    lib.gpublas_dev_release.argtypes = [vp]
    lib.gpublas_dev_h2d.argtypes = [vp, vp, sz, vp]
    lib.gpublas_dev_d2h.argtypes = [vp, vp, sz, vp]
    lib.gpublas_dev_sync.argtypes = [vp]
    lib.gpublas_dev_stream.restype = vp
    lib.gpublas_dev_arch.restype = ctypes.c_char_p
    lib.gpublas_dev_fill_uniform.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64, vp]
    lib.gpublas_dev_batch_create.restype = vp
    lib.gpublas_dev_batch_create.argtypes = [ci]
    for f in ("destroy", "clear", "launch", "done", "wait", "count"):
        getattr(lib, "gpublas_dev_batch_" + f).argtypes = [vp]
    lib.gpublas_dev_batch_force_cfg.argtypes = [vp, ci]
    ch = ctypes.c_char
    lib.gpublas_dev_batch_dgemm.argtypes = [vp, ch, ch, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.gpublas_dev_batch_sgemm.argtypes = [vp, ch, ch, ci, ci, ci, cf, vp, ci, vp, ci, cf, vp, ci]
    lib.gpublas_dev_batch_dgemv.argtypes = [vp, ch, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.gpublas_dev_batch_dtrsm.argtypes = [vp, ch, ch, ch, ch, ci, ci, cd, vp, ci, vp, ci]
    lib.gpublas_dev_tune_gemm.argtypes = [ctypes.c_char_p, ci]
    lib.gpublas_dev_tune_choice.argtypes = [ci, ci, ci, ci]
    lib.gpublas_dev_prof_enable.argtypes = [ci]
    lib.gpublas_dev_prof_query.argtypes = [ctypes.c_char_p, ctypes.POINTER(cd),
                                           ctypes.POINTER(cd), ctypes.POINTER(cd)]


# ---------------------------------------------------------------------------
# Fortran-ABI mirrors: every scalar by reference, arrays by address.
# ---------------------------------------------------------------------------
def _c(ch):
    return ctypes.byref(ctypes.c_char(ch.encode() if isinstance(ch, str) else ch))


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _f(v):
    return ctypes.byref(ctypes.c_float(float(v)))


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return ctypes.c_void_p(a.ctypes.data)


def dgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    library().dgemm_(_c(transa), _c(transb), _i(m), _i(n), _i(k), _d(alpha), _p(a), _i(lda),
                     _p(b), _i(ldb), _d(beta), _p(c), _i(ldc))


def sgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    library().sgemm_(_c(transa), _c(transb), _i(m), _i(n), _i(k), _f(alpha), _p(a), _i(lda),
                     _p(b), _i(ldb), _f(beta), _p(c), _i(ldc))


def _z(v, dt=ctypes.c_double):
    """A complex scalar as the (re, im) pair the Fortran ABI passes by address."""
    v = complex(v)
    return (dt * 2)(v.real, v.imag)


def zgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """ZGEMM on complex128 arrays (leading dimensions in complex elements)."""
    library().zgemm_(_c(transa), _c(transb), _i(m), _i(n), _i(k), _z(alpha), _p(a), _i(lda),
                     _p(b), _i(ldb), _z(beta), _p(c), _i(ldc))


def cgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """CGEMM on complex64 arrays."""
    library().cgemm_(_c(transa), _c(transb), _i(m), _i(n), _i(k), _z(alpha, ctypes.c_float),
                     _p(a), _i(lda), _p(b), _i(ldb), _z(beta, ctypes.c_float), _p(c), _i(ldc))


def dsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    library().dsyrk_(_c(uplo), _c(trans), _i(n), _i(k), _d(alpha), _p(a), _i(lda), _d(beta),
                     _p(c), _i(ldc))


def ssyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    library().ssyrk_(_c(uplo), _c(trans), _i(n), _i(k), _f(alpha), _p(a), _i(lda), _f(beta),
                     _p(c), _i(ldc))


def dtrsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    library().dtrsm_(_c(side), _c(uplo), _c(transa), _c(diag), _i(m), _i(n), _d(alpha), _p(a),
                     _i(lda), _p(b), _i(ldb))


def strsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    library().strsm_(_c(side), _c(uplo), _c(transa), _c(diag), _i(m), _i(n), _f(alpha), _p(a),
                     _i(lda), _p(b), _i(ldb))


def dsyr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    library().dsyr2k_(_c(uplo), _c(trans), _i(n), _i(k), _d(alpha), _p(a), _i(lda), _p(b),
                      _i(ldb), _d(beta), _p(c), _i(ldc))


def ssyr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    library().ssyr2k_(_c(uplo), _c(trans), _i(n), _i(k), _f(alpha), _p(a), _i(lda), _p(b),
                      _i(ldb), _f(beta), _p(c), _i(ldc))


def dtrmm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    library().dtrmm_(_c(side), _c(uplo), _c(transa), _c(diag), _i(m), _i(n), _d(alpha), _p(a),
                     _i(lda), _p(b), _i(ldb))


def strmm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    library().strmm_(_c(side), _c(uplo), _c(transa), _c(diag), _i(m), _i(n), _f(alpha), _p(a),
                     _i(lda), _p(b), _i(ldb))


def dsymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    library().dsymm_(_c(side), _c(uplo), _i(m), _i(n), _d(alpha), _p(a), _i(lda), _p(b),
                     _i(ldb), _d(beta), _p(c), _i(ldc))


def ssymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    library().ssymm_(_c(side), _c(uplo), _i(m), _i(n), _f(alpha), _p(a), _i(lda), _p(b),
                     _i(ldb), _f(beta), _p(c), _i(ldc))


def dgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    library().dgemv_(_c(trans), _i(m), _i(n), _d(alpha), _p(a), _i(lda), _p(x), _i(incx),
                     _d(beta), _p(y), _i(incy))


def dger(m, n, alpha, x, incx, y, incy, a, lda):
    library().dger_(_i(m), _i(n), _d(alpha), _p(x), _i(incx), _p(y), _i(incy), _p(a), _i(lda))


def dtrsv(uplo, trans, diag, n, a, lda, x, incx):
    library().dtrsv_(_c(uplo), _c(trans), _c(diag), _i(n), _p(a), _i(lda), _p(x), _i(incx))


def dsymv(uplo, n, alpha, a, lda, x, incx, beta, y, incy):
    library().dsymv_(_c(uplo), _i(n), _d(alpha), _p(a), _i(lda), _p(x), _i(incx), _d(beta),
                     _p(y), _i(incy))


def dsyr(uplo, n, alpha, x, incx, a, lda):
    library().dsyr_(_c(uplo), _i(n), _d(alpha), _p(x), _i(incx), _p(a), _i(lda))


def dsyr2(uplo, n, alpha, x, incx, y, incy, a, lda):
    library().dsyr2_(_c(uplo), _i(n), _d(alpha), _p(x), _i(incx), _p(y), _i(incy), _p(a),
                     _i(lda))


def dspr(uplo, n, alpha, x, incx, ap):
    library().dspr_(_c(uplo), _i(n), _d(alpha), _p(x), _i(incx), _p(ap))


def dspr2(uplo, n, alpha, x, incx, y, incy, ap):
    library().dspr2_(_c(uplo), _i(n), _d(alpha), _p(x), _i(incx), _p(y), _i(incy), _p(ap))


def dtrmv(uplo, trans, diag, n, a, lda, x, incx):
    library().dtrmv_(_c(uplo), _c(trans), _c(diag), _i(n), _p(a), _i(lda), _p(x), _i(incx))


def dtpmv(uplo, trans, diag, n, ap, x, incx):
    library().dtpmv_(_c(uplo), _c(trans), _c(diag), _i(n), _p(ap), _p(x), _i(incx))


def dtpsv(uplo, trans, diag, n, ap, x, incx):
    library().dtpsv_(_c(uplo), _c(trans), _c(diag), _i(n), _p(ap), _p(x), _i(incx))


def dgbmv(trans, m, n, kl, ku, alpha, a, lda, x, incx, beta, y, incy):
    library().dgbmv_(_c(trans), _i(m), _i(n), _i(kl), _i(ku), _d(alpha), _p(a), _i(lda), _p(x),
                     _i(incx), _d(beta), _p(y), _i(incy))


def dsbmv(uplo, n, k, alpha, a, lda, x, incx, beta, y, incy):
    library().dsbmv_(_c(uplo), _i(n), _i(k), _d(alpha), _p(a), _i(lda), _p(x), _i(incx), _d(beta),
                     _p(y), _i(incy))


def dtbmv(uplo, trans, diag, n, k, a, lda, x, incx):
    library().dtbmv_(_c(uplo), _c(trans), _c(diag), _i(n), _i(k), _p(a), _i(lda), _p(x), _i(incx))


def dtbsv(uplo, trans, diag, n, k, a, lda, x, incx):
    library().dtbsv_(_c(uplo), _c(trans), _c(diag), _i(n), _i(k), _p(a), _i(lda), _p(x), _i(incx))


def dspmv(uplo, n, alpha, ap, x, incx, beta, y, incy):
    library().dspmv_(_c(uplo), _i(n), _d(alpha), _p(ap), _p(x), _i(incx), _d(beta), _p(y),
                     _i(incy))


def daxpy(n, alpha, x, incx, y, incy):
    library().daxpy_(_i(n), _d(alpha), _p(x), _i(incx), _p(y), _i(incy))


def ddot(n, x, incx, y, incy):
    return library().ddot_(_i(n), _p(x), _i(incx), _p(y), _i(incy))


def dscal(n, alpha, x, incx):
    library().dscal_(_i(n), _d(alpha), _p(x), _i(incx))


def dcopy(n, x, incx, y, incy):
    library().dcopy_(_i(n), _p(x), _i(incx), _p(y), _i(incy))


def dswap(n, x, incx, y, incy):
    library().dswap_(_i(n), _p(x), _i(incx), _p(y), _i(incy))


def drot(n, x, incx, y, incy, c, s):
    library().drot_(_i(n), _p(x), _i(incx), _p(y), _i(incy), _d(c), _d(s))


def dasum(n, x, incx):
    return library().dasum_(_i(n), _p(x), _i(incx))


def dnrm2(n, x, incx):
    return library().dnrm2_(_i(n), _p(x), _i(incx))


def idamax(n, x, incx):
    return library().idamax_(_i(n), _p(x), _i(incx))


def idamin(n, x, incx):
    return library().idamin_(_i(n), _p(x), _i(incx))


def drotm(n, x, incx, y, incy, param):
    library().drotm_(_i(n), _p(x), _i(incx), _p(y), _i(incy), _p(param))


# ---------------------------------------------------------------------------
# Extension API
# ---------------------------------------------------------------------------
def set_options(opts: str) -> int:
    return library().gpublas_set_options(opts.encode())


def set_cutoff(cutoff: int) -> None:
    library().gpublas_set_cutoff(int(cutoff))


def get_cutoff() -> int:
    """Current cutoff; CUTOFF_AUTO (-1) under cutoff=auto."""
    return library().gpublas_get_cutoff()


CUTOFF_AUTO = -1
CROSSOVER_ROUTINES = ("dgemm", "sgemm", "dsyrk", "ssyrk", "dsyr2k", "ssyr2k", "dsymm", "ssymm",
                      "dtrsm", "strsm", "dtrmm", "strmm", "dgemv", "dger", "dtrsv", "dsymv",
                      "dsyr", "daxpy", "ddot")


def set_crossover(routine: str, min_flops: float, min_intensity: float) -> None:
    """Set one row of the host/GPU crossover table used under cutoff=auto."""
    if library().gpublas_set_crossover(routine.encode(), float(min_flops),
                                       float(min_intensity)) != 0:
        raise ValueError(f"bad crossover row {routine!r} {min_flops} {min_intensity}")


def get_crossover(routine: str) -> tuple:
    """(min_flops, min_intensity) of one routine's crossover row."""
    f, i = ctypes.c_double(), ctypes.c_double()
    if library().gpublas_get_crossover(routine.encode(), ctypes.byref(f), ctypes.byref(i)) != 0:
        raise ValueError(f"unknown routine {routine!r}")
    return f.value, i.value


def crossover_gpu(routine: str, flops: float, nbytes: float) -> bool:
    """True if the crossover table sends such a call (no resident operand) to the GPU."""
    r = library().gpublas_crossover_gpu(routine.encode(), float(flops), float(nbytes))
    if r < 0:
        raise ValueError(f"unknown routine {routine!r}")
    return bool(r)


def tune_crossover(path=None, max_n: int = 2048) -> None:
    """Measure the crossover table on this machine (needs the GPU), install it
    and write it to `path` if given (load it later with xover=<path>)."""
    r = library().gpublas_tune_crossover(None if path is None else str(path).encode(), int(max_n))
    if r != 0:
        raise RuntimeError(f"gpublas_tune_crossover failed ({r})")


def stats() -> dict:
    s = Stats()
    library().gpublas_get_stats(ctypes.byref(s))
    return s.as_dict()


def reset_stats() -> None:
    library().gpublas_reset_stats()


def last_xerbla():
    buf = ctypes.create_string_buffer(8)
    info = library().gpublas_last_xerbla(buf)
    return buf.value.decode(), info


def device_ready() -> bool:
    return bool(library().gpublas_device_ready())


def kernel_time(name: str):
    ms = ctypes.c_double(0)
    n = library().gpublas_kernel_time(name.encode(), ctypes.byref(ms))
    return n, ms.value


def sync() -> None:
    """Wait for every kernel the calling thread launched."""
    library().gpublas_sync()


def device_ptr(a) -> int:
    """Device address of a tracked array's element 0.  The device copy is
    treated as the newer one from then on (a host access downloads it)."""
    addr = library().gpublas_device_ptr(_p(a))
    if not addr:
        raise GpublasError("not a tracked array")
    return addr


class Batch:
    """One batch slot of the serving loop's device layer (device pointers)."""

    def __init__(self, use_graph=True):
        self._lib = library()
        self._b = self._lib.gpublas_dev_batch_create(1 if use_graph else 0)
        if not self._b:
            raise GpublasError("gpublas_dev_batch_create failed")

    def force_cfg(self, cfg):
        self._lib.gpublas_dev_batch_force_cfg(self._b, int(cfg))

    def dgemm(self, ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc):
        return self._lib.gpublas_dev_batch_dgemm(self._b, ta.encode(), tb.encode(), m, n, k, alpha,
                                                 dA, lda, dB, ldb, beta, dC, ldc)

    def sgemm(self, ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc):
        return self._lib.gpublas_dev_batch_sgemm(self._b, ta.encode(), tb.encode(), m, n, k, alpha,
                                                 dA, lda, dB, ldb, beta, dC, ldc)

    def dgemv(self, tr, m, n, alpha, dA, lda, dx, incx, beta, dy, incy):
        return self._lib.gpublas_dev_batch_dgemv(self._b, tr.encode(), m, n, alpha, dA, lda, dx,
                                                 incx, beta, dy, incy)

    def dtrsm(self, side, uplo, ta, diag, m, n, alpha, dA, lda, dB, ldb):
        return self._lib.gpublas_dev_batch_dtrsm(self._b, side.encode(), uplo.encode(),
                                                 ta.encode(), diag.encode(), m, n, alpha, dA, lda,
                                                 dB, ldb)

    def count(self):
        return self._lib.gpublas_dev_batch_count(self._b)

    def run(self):
        err = self._lib.gpublas_dev_batch_launch(self._b)
        if err:
            raise GpublasError(f"batch launch failed ({err})")
        err = self._lib.gpublas_dev_batch_wait(self._b)
        if err:
            raise GpublasError(f"batch wait failed ({err})")
        self._lib.gpublas_dev_batch_clear(self._b)

    def close(self):
        if self._b:
            self._lib.gpublas_dev_batch_destroy(self._b)
            self._b = None


def tune_gemm(path=None, force=False) -> int:
    return library().gpublas_dev_tune_gemm(path.encode() if path else None, 1 if force else 0)


def tune_choice(single, m, n, k) -> int:
    return library().gpublas_dev_tune_choice(1 if single else 0, m, n, k)


def tracked_empty(shape, dtype=np.float64, order="F"):
    """NumPy array backed by a tracked (HBM-resident) object."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape)) if shape else 1
    nbytes = max(count * dtype.itemsize, 1)
    lib = library()
    addr = lib.gpublas_malloc(nbytes)
    if not addr:
        raise MemoryError("gpublas_malloc failed")
    buf = (ctypes.c_char * nbytes).from_address(addr)
    arr = np.frombuffer(buf, dtype=dtype, count=count).reshape(shape, order=order)
    weakref.finalize(buf, lib.gpublas_free, ctypes.c_void_p(addr))
    return arr


def tracked_copy(a, order="F"):
    t = tracked_empty(a.shape, a.dtype, order=order)
    t[...] = a
    return t


def is_tracked(a) -> bool:
    return bool(library().gpublas_is_tracked(_p(a)))
