"""ctypes bindings of the CPU oracle (oracle/refblas.c) -- test-side only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None
C, I, D, F, VP = ctypes.c_char, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
    L = ctypes.CDLL(ORACLE_SO)
    L.ora_lsame.argtypes = [C, C]
    L.ora_gemm_info.argtypes = [C, C, I, I, I, I, I, I]
    L.ora_syrk_info.argtypes = [C, C, I, I, I, I]
    L.ora_trsm_info.argtypes = [C, C, C, C, I, I, I, I]
    L.ora_symm_info.argtypes = [C, C, I, I, I, I, I]
    L.ora_gemv_info.argtypes = [C, I, I, I, I, I]
    L.ora_ger_info.argtypes = [I, I, I, I, I]
    L.ora_trsv_info.argtypes = [C, C, C, I, I, I]
    L.ora_symv_info.argtypes = [C, I, I, I, I]
    L.ora_syr_info.argtypes = [C, I, I, I]
    L.ora_dsymv.argtypes = [C, I, D, VP, I, VP, I, D, VP, I]
    L.ora_dsyr.argtypes = [C, I, D, VP, I, VP, I]
    L.ora_syr2_info.argtypes = [C, I, I, I, I]
    L.ora_spr_info.argtypes = [C, I, I]
    L.ora_spr2_info.argtypes = [C, I, I, I]
    L.ora_dsyr2.argtypes = [C, I, D, VP, I, VP, I, VP, I]
    L.ora_dspr.argtypes = [C, I, D, VP, I, VP]
    L.ora_dspr2.argtypes = [C, I, D, VP, I, VP, I, VP]
    L.ora_trmv_info.argtypes = [C, C, C, I, I, I]
    L.ora_tpmv_info.argtypes = [C, C, C, I, I]
    L.ora_spmv_info.argtypes = [C, I, I, I]
    L.ora_dtrmv.argtypes = [C, C, C, I, VP, I, VP, I]
    L.ora_dtpmv.argtypes = [C, C, C, I, VP, VP, I]
    L.ora_dtpsv.argtypes = [C, C, C, I, VP, VP, I]
    L.ora_dspmv.argtypes = [C, I, D, VP, VP, I, D, VP, I]
    L.ora_idamin.argtypes = [I, VP, I]
    L.ora_drotm.argtypes = [I, VP, I, VP, I, VP]
    L.ora_tbmv_info.argtypes = [C, C, C, I, I, I, I]
    L.ora_gbmv_info.argtypes = [C, I, I, I, I, I, I, I]
    L.ora_sbmv_info.argtypes = [C, I, I, I, I, I]
    L.ora_dtbmv.argtypes = [C, C, C, I, I, VP, I, VP, I]
    L.ora_dtbsv.argtypes = [C, C, C, I, I, VP, I, VP, I]
    L.ora_dgbmv.argtypes = [C, I, I, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_dsbmv.argtypes = [C, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_dgemm.argtypes = [C, C, I, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_sgemm.argtypes = [C, C, I, I, I, F, VP, I, VP, I, F, VP, I]
    L.ora_zgemm.argtypes = [C, C, I, I, I, VP, VP, I, VP, I, VP, VP, I]
    L.ora_cgemm.argtypes = [C, C, I, I, I, VP, VP, I, VP, I, VP, VP, I]
    for pre in "zc":
        for op in ("syrk", "herk"):
            getattr(L, f"ora_{pre}{op}").argtypes = [C, C, I, I, VP, VP, I, VP, VP, I]
        for op in ("syr2k", "her2k"):
            getattr(L, f"ora_{pre}{op}").argtypes = [C, C, I, I, VP, VP, I, VP, I, VP, VP, I]
        for op in ("symm", "hemm"):
            getattr(L, f"ora_{pre}{op}").argtypes = [C, C, I, I, VP, VP, I, VP, I, VP, VP, I]
        for op in ("trmm", "trsm"):
            getattr(L, f"ora_{pre}{op}").argtypes = [C, C, C, C, I, I, VP, VP, I, VP, I]
    L.ora_dsyrk.argtypes = [C, C, I, I, D, VP, I, D, VP, I]
    L.ora_ssyrk.argtypes = [C, C, I, I, F, VP, I, F, VP, I]
    L.ora_dtrsm.argtypes = [C, C, C, C, I, I, D, VP, I, VP, I]
    L.ora_strsm.argtypes = [C, C, C, C, I, I, F, VP, I, VP, I]
    L.ora_dsymm.argtypes = [C, C, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_syr2k_info.argtypes = [C, C, I, I, I, I, I]
    L.ora_trmm_info.argtypes = [C, C, C, C, I, I, I, I]
    L.ora_dsyr2k.argtypes = [C, C, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_dtrmm.argtypes = [C, C, C, C, I, I, D, VP, I, VP, I]
    L.ora_ssymm.argtypes = [C, C, I, I, F, VP, I, VP, I, F, VP, I]
    L.ora_dgemv.argtypes = [C, I, I, D, VP, I, VP, I, D, VP, I]
    L.ora_dger.argtypes = [I, I, D, VP, I, VP, I, VP, I]
    L.ora_dtrsv.argtypes = [C, C, C, I, VP, I, VP, I]
    L.ora_daxpy.argtypes = [I, D, VP, I, VP, I]
    L.ora_dscal.argtypes = [I, D, VP, I]
    L.ora_dcopy.argtypes = [I, VP, I, VP, I]
    L.ora_dswap.argtypes = [I, VP, I, VP, I]
    L.ora_drot.argtypes = [I, VP, I, VP, I, D, D]
    L.ora_dasum.argtypes = [I, VP, I]
    L.ora_dasum.restype = D
    L.ora_ddot.argtypes = [I, VP, I, VP, I]
    L.ora_ddot.restype = D
    L.ora_dnrm2.argtypes = [I, VP, I]
    L.ora_dnrm2.restype = D
    L.ora_idamax.argtypes = [I, VP, I]
    L.ora_idamax.restype = I
    L.ora_gemm_gauge_ratio.argtypes = [C, C, I, I, I, D, VP, I, VP, I, D, VP, VP, I, D]
    L.ora_gemm_gauge_ratio.restype = D
    L.ora_fill_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, VP]
    L.ora_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.ora_splitmix64_at.restype = ctypes.c_uint64
    _lib = L
    return L


def c(ch):
    return ch.encode() if isinstance(ch, str) else ch


def p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def uniform(seed, shape, dtype=np.float64, order="F"):
    """Deterministic U[-1,1) input of the given shape (SplitMix64 stream)."""
    L = load()
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float64)
    L.ora_fill_uniform(ctypes.c_uint64(seed), n, p(out))
    return np.asarray(out.astype(dtype).reshape(shape, order=order), order=order)
