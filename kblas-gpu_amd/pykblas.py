"""ctypes binding of libkblas_amd.so (the C ABI declared in include/kblas*.h).

This is a thin caller-side binding, the Python counterpart of the ctypes stub a
maintainer would write against the reference library: it loads the in-tree
shared object, declares the argument types of every entry point and offers a
small Handle wrapper.  Device memory is passed as integer addresses (for
example ``tensor.data_ptr()`` of a torch tensor on a ROCm device); nothing here
computes -- every call goes to the HIP kernels in the library, and loading
fails loudly if the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkblas_amd.so")

KBLAS_Success = 1
KBLAS_UnknownError = 0
KBLAS_NotSupported = -1
KBLAS_NotImplemented = -2
KBLAS_WrongConfig = -4
KBLAS_CUDA_Error = -5
KBLAS_InsufficientWorkspace = -6
KBLAS_Error_NotInitialized = -9
KBLAS_Error_WrongInput = -10

c_int, c_long, c_char, c_double, c_float = ctypes.c_int, ctypes.c_long, ctypes.c_char, ctypes.c_double, ctypes.c_float
P = ctypes.c_void_p  # any device or host pointer, passed as an address

# name -> (restype, [argtypes]); T is replaced by c_double / c_float
_SIG = {
    "kblasCreate": (c_int, [P]),
    "kblasDestroy": (c_int, [P]),
    "kblasCreateStreams": (c_int, [P, c_int]),
    "kblasGetStream": (P, [P]),
    "kblasSetStream": (None, [P, P]),
    "kblasGetCublasHandle": (P, [P]),
    "kblasEnableMagma": (c_int, [P]),
    "kblasGetErrorString": (ctypes.c_char_p, [c_int]),
    "kblasAllocateWorkspace": (c_int, [P]),
    "kblasFreeWorkspace": (c_int, [P]),
    "kblasTimerTic": (None, [P]),
    "kblasTimerRecordEnd": (None, [P]),
    "kblasTimerToc": (c_double, [P]),
    # workspace queries
    "kblas_potrf_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_potrf_batch_strided_wsquery": (None, [P, c_int, c_int]),
    "kblas_potrs_batch_wsquery": (None, [P, c_int, c_int, c_int]),
    "kblas_potrs_batch_strided_wsquery": (None, [P, c_int, c_int, c_int]),
    "kblas_posv_batch_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_posv_batch_strided_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_trsm_batch_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_trsm_batch_strided_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_trsm_batch_nonuniform_wsquery": (None, [P]),
    "kblas_syrk_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_syrk_batch_nonuniform_wsquery": (None, [P]),
    "kblas_gemm_batch_strided_wsquery": (None, [P, c_int]),
    "kblas_gemm_batch_nonuniform_wsquery": (None, [P]),
    "kblas_potrf_vbatch_wsquery": (None, [P, c_int, c_int]),
    "kblas_posv_vbatch_wsquery": (None, [P, c_int, c_int]),
    "kblas_sposv_vbatch_wsquery": (None, [P, c_int, c_int, c_int]),
    "kblas_trmm_batch_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_trmm_batch_strided_wsquery": (None, [P, c_char, c_int, c_int, c_int]),
    "kblas_trtri_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_trtri_batch_strided_wsquery": (None, [P, c_int, c_int]),
    "kblas_lauum_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_lauum_batch_strided_wsquery": (None, [P, c_int, c_int]),
    "kblas_potri_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_potri_batch_strided_wsquery": (None, [P, c_int, c_int]),
    "kblas_poti_batch_wsquery": (None, [P, c_int, c_int]),
    "kblas_poti_batch_strided_wsquery": (None, [P, c_int, c_int]),
    # kernel choice per size bucket
    "kblas_tuning_load": (c_int, [ctypes.c_char_p]),
    "kblas_tuning_set": (c_int, [c_char, ctypes.c_char_p, c_int, c_int, ctypes.c_char_p]),
    "kblas_tuning_clear": (None, []),
    "kblas_tuning_query": (c_int, [c_char, ctypes.c_char_p, c_int]),
    "kblas_tuning_variant_name": (ctypes.c_char_p, [c_int]),
}
_TSIG = {
    "potrf_batch": (c_int, [P, c_char, c_int, P, c_int, c_int, P]),
    "potrf_batch_strided": (c_int, [P, c_char, c_int, P, c_int, c_long, c_int, P]),
    "potrs_batch": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, P, c_int, c_int]),
    "potrs_batch_strided": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, c_long, P, c_int, c_long, c_int]),
    "posv_batch": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, P, c_int, c_int, P]),
    "posv_batch_strided": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, c_long, P, c_int, c_long, c_int, P]),
    "trsm_batch": (c_int, [P, c_char, c_char, c_char, c_char, c_int, c_int, "T", P, c_int, P, c_int, c_int]),
    "trsm_batch_strided": (c_int, [P, c_char, c_char, c_char, c_char, c_int, c_int, "T", P, c_int, c_long, P, c_int, c_long, c_int]),
    "syrk_batch": (c_int, [P, c_char, c_char, c_int, c_int, "T", P, c_int, "T", P, c_int, c_int]),
    "syrk_batch_strided": (c_int, [P, c_char, c_char, c_int, c_int, "T", P, c_int, c_long, "T", P, c_int, c_long, c_int]),
    "gemm_batch": (c_int, [P, c_char, c_char, c_int, c_int, c_int, "T", P, c_int, P, c_int, "T", P, c_int, c_int]),
    "gemm_batch_strided": (c_int, [P, c_char, c_char, c_int, c_int, c_int, "T", P, c_int, c_long, P, c_int, c_long, "T", P, c_int, c_long, c_int]),
    "gemm_vbatch": (c_int, [P, c_char, c_char, P, P, P, c_int, c_int, c_int, "T", P, P, P, P, "T", P, P, c_int]),
    "syrk_vbatch": (c_int, [P, c_char, c_char, P, P, c_int, c_int, "T", P, P, "T", P, P, c_int]),
    "trsm_vbatch": (c_int, [P, c_char, c_char, c_char, c_char, P, P, c_int, c_int, "T", P, P, P, P, c_int]),
    "potrf_vbatch": (c_int, [P, c_char, P, c_int, P, P, c_int, P]),
    "posv_vbatch": (c_int, [P, c_char, c_char, P, P, c_int, c_int, P, P, P, P, c_int, P]),
    "trmm_batch": (c_int, [P, c_char, c_char, c_char, c_char, c_int, c_int, "T", P, c_int, P, c_int, c_int]),
    "trmm_batch_strided": (c_int, [P, c_char, c_char, c_char, c_char, c_int, c_int, "T", P, c_int, c_long, P, c_int, c_long, c_int]),
    "trtri_batch": (c_int, [P, c_char, c_char, c_int, P, c_int, c_int, P]),
    "trtri_batch_strided": (c_int, [P, c_char, c_char, c_int, P, c_int, c_long, c_int, P]),
    "lauum_batch": (c_int, [P, c_char, c_int, P, c_int, c_int, P]),
    "lauum_batch_strided": (c_int, [P, c_char, c_int, P, c_int, c_long, c_int, P]),
    "potri_batch": (c_int, [P, c_char, c_int, P, c_int, c_int, P]),
    "potri_batch_strided": (c_int, [P, c_char, c_int, P, c_int, c_long, c_int, P]),
    "poti_batch": (c_int, [P, c_char, c_int, P, c_int, c_int, P]),
    "poti_batch_strided": (c_int, [P, c_char, c_int, P, c_int, c_long, c_int, P]),
}
_DONLY = {
    "kblasDpotrs_vbatch": (c_int, [P, c_char, c_char, P, P, c_int, c_int, P, P, P, P, c_int]),
    "kblasDsposv_batch_strided": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, c_long, P, c_int, c_long,
                                          P, c_int, c_long, c_int, P, P]),
    "kblasDsposv_batch": (c_int, [P, c_char, c_char, c_int, c_int, P, c_int, P, c_int, P, c_int, c_int, P, P]),
    "kblasDsposv_vbatch": (c_int, [P, c_char, c_char, P, P, c_int, c_int, P, P, P, P, P, P, c_int, P, P]),
    "kblasDposv_vbatch_noalloc": (c_int, [P, c_char, c_char, P, P, c_int, c_int, P, P, P, P, c_int, P]),
    # randomized-range-finder SPD helpers (include/batch_ara.h)
    "kblas_sara_fused_potrf_batch": (c_int, [P, P, P, P, P, P, c_int, P, c_int, P]),
    "kblas_dara_fused_potrf_batch": (c_int, [P, P, P, P, P, P, c_int, P, c_int, P]),
    "kblas_sara_mp_syrk_batch": (c_int, [P, P, P, c_int, c_int, P, P, P, P, c_int]),
    "kblas_dara_mp_syrk_batch": (c_int, [P, P, P, c_int, c_int, P, P, P, P, c_int]),
    "kblas_sara_trsm_batch": (c_int, [P, P, P, P, P, P, P, c_int, c_int, c_int]),
    "kblas_dara_trsm_batch": (c_int, [P, P, P, P, P, P, P, c_int, c_int, c_int]),
}
_TSIG["pstrf_batch"] = (c_int, [P, c_int, P, c_int, P, P, c_int])
_TSIG["pstrf_batch_strided"] = (c_int, [P, c_int, P, c_int, c_int, P, c_int, P, c_int])


class CF(ctypes.Structure):
    """hipFloatComplex passed by value"""
    _fields_ = [("x", c_float), ("y", c_float)]


class CD(ctypes.Structure):
    """hipDoubleComplex passed by value"""
    _fields_ = [("x", c_double), ("y", c_double)]


# complex entry points with the real signatures above, T = the complex scalar
_CPLX = ("gemm_batch", "gemm_batch_strided", "syrk_batch", "syrk_batch_strided", "trsm_batch", "trsm_batch_strided",
         "potrf_batch", "potrf_batch_strided", "potrs_batch", "potrs_batch_strided", "posv_batch",
         "posv_batch_strided", "trmm_batch", "trmm_batch_strided", "trtri_batch", "trtri_batch_strided",
         "lauum_batch", "lauum_batch_strided", "potri_batch", "potri_batch_strided", "poti_batch",
         "poti_batch_strided")


def _declare(lib):
    for suffix in _CPLX:
        res, args = _TSIG[suffix]
        for prec, t in (("C", CF), ("Z", CD)):
            f = getattr(lib, "kblas%s%s" % (prec, suffix))
            f.restype = res
            f.argtypes = [t if a == "T" else a for a in args]
    for name, (res, args) in _SIG.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    for suffix, (res, args) in _TSIG.items():
        for prec, t in (("S", c_float), ("D", c_double)):
            f = getattr(lib, "kblas%s%s" % (prec, suffix))
            f.restype = res
            f.argtypes = [t if a == "T" else a for a in args]
    for name, (res, args) in _DONLY.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def load(path=None):
    """Load (once) and return the library; raises OSError if it is not built."""
    global _lib
    if _lib is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise OSError("libkblas_amd.so not built at %s (run __graft_entry__.build())" % p)
        _lib = _declare(ctypes.CDLL(p))
    return _lib


def ch(c):
    """char argument"""
    return c.encode() if isinstance(c, str) else c


class Handle:
    """kblasHandle_t owner; optional HIP stream (an integer hipStream_t)."""

    def __init__(self, stream=None):
        self.lib = load()
        self.h = P()
        st = self.lib.kblasCreate(ctypes.byref(self.h))
        if st != KBLAS_Success:
            raise RuntimeError("kblasCreate failed: %s" % self.lib.kblasGetErrorString(st).decode())
        if stream is not None:
            self.lib.kblasSetStream(self.h, P(stream))

    def set_stream(self, stream):
        self.lib.kblasSetStream(self.h, P(stream))

    def close(self):
        if self.h:
            self.lib.kblasDestroy(ctypes.byref(self.h))
            self.h = P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
