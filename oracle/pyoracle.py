"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: used only by tests/, the smoke entry and the benchmark's
CPU-baseline leg, always as the checker.  The product library never loads it.
All arrays are numpy float64/int32, column-major (Fortran order).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_u64 = ctypes.c_uint64
_c = ctypes.c_char
_p = ctypes.c_void_p

_SIG = {
    "orc_hash": (_u64, [_u64, _u64, _u64]),
    "orc_uniform": (_d, [_u64, _u64, _u64]),
    "orc_normal": (_d, [_u64, _u64, _u64]),
    "orc_exp2": (_d, [_d]),
    "orc_gen_spd_shift": (None, [_u64, _u64, _i, _d, _p, _i]),
    "orc_gen_spd_cond": (_d, [_u64, _u64, _i, _d, _d, _p, _i]),
    "orc_gen_rhs": (None, [_u64, _u64, _i, _i, _p, _i]),
    "orc_make_hpd": (None, [_i, _p, _i]),
    "orc_reg_size": (_i, [_i]),
    "orc_closest_reg_size": (_i, [_i]),
    "orc_potrf": (_i, [_i, _p, _i]),
    "orc_spotrf": (_i, [_i, _p, _i]),
    "orc_trsm": (None, [_c, _c, _i, _i, _d, _p, _i, _p, _i]),
    "orc_syrk": (None, [_c, _i, _i, _d, _p, _i, _d, _p, _i]),
    "orc_gemm": (None, [_c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "orc_potrs": (None, [_c, _i, _i, _p, _i, _p, _i]),
    "orc_trtri": (_i, [_c, _i, _p, _i]),
    "orc_pstrf": (_i, [_i, _p, _i, _p]),
    "orc_ara_potrf": (_i, [_i, _i, _p, _i, _p, _i, _p]),
    "orc_zsqrt": (None, [_d, _d, _p]),
    "orc_zpotrf": (_i, [_i, _i, _p, _i]),
    "orc_zpotrs": (None, [_i, _i, _i, _i, _p, _i, _p, _i]),
    "orc_zgemm": (None, [_c, _c, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "orc_zsyrk": (None, [_i, _c, _i, _i, _p, _p, _i, _p, _p, _i]),
    "orc_ztrsm": (None, [_c, _i, _c, _i, _i, _i, _p, _p, _i, _p, _i]),
    "orc_ztrtri": (_i, [_i, _i, _p, _i]),
    "orc_zlauum": (None, [_i, _p, _i]),
    "orc_zpotri": (_i, [_i, _p, _i]),
    "orc_ztrmm": (None, [_c, _c, _i, _i, _i, _p, _p, _i, _p, _i]),
    "orc_ara_syrk": (None, [_i, _i, _p, _i, _p, _i]),
    "orc_ara_trsm": (None, [_i, _i, _i, _p, _i, _p, _i]),
    "orc_lauum": (None, [_i, _p, _i]),
    "orc_potri": (_i, [_i, _p, _i]),
    "orc_trmm": (None, [_c, _c, _c, _i, _i, _d, _p, _i, _p, _i]),
    "orc_posv": (_i, [_c, _i, _i, _p, _i, _p, _i]),
    "orc_sposv": (_i, [_c, _i, _i, _p, _i, _p, _i, _p, _i, _p]),
    "orc_potrf_berr": (_d, [_i, _p, _i, _p, _i]),
    "orc_solve_berr": (_d, [_c, _i, _i, _p, _i, _p, _i, _p, _i]),
    "orc_flops_potrf": (_d, [_i]),
    "orc_flops_trsm": (_d, [_c, _i, _i]),
    "orc_flops_potrs": (_d, [_c, _i, _i]),
    "orc_flops_posv": (_d, [_c, _i, _i]),
    "orc_flops_syrk": (_d, [_i, _i]),
    "orc_flops_gemm": (_d, [_i, _i, _i]),
    "orc_threads": (_i, []),
    "orc_set_threads": (None, [_i]),
    "orc_potrf_batch": (None, [_i, _p, _i, _l, _i, _p]),
    "orc_posv_batch": (None, [_c, _i, _i, _p, _i, _l, _p, _i, _l, _i, _p]),
    "orc_sposv_batch": (None, [_c, _i, _i, _p, _i, _l, _p, _i, _l, _p, _i, _l, _i, _p, _p]),
    "orc_lapack_available": (_i, []),
    "orc_lapack_name": (ctypes.c_char_p, []),
    "orc_lapack_potrf_batch": (_i, [_i, _p, _i, _l, _i, _p]),
    "orc_lapack_posv_batch": (_i, [_c, _i, _i, _p, _i, _l, _p, _i, _l, _i, _p]),
    "orc_lapack_posv_vbatch": (_i, [_i, _p, _p, _p, _p, _p, _i, _p]),
    "orc_posv_vbatch": (None, [_i, _p, _p, _p, _p, _p, _i, _p]),
    "orc_gen_spd_shift_batch": (None, [_u64, _i, _d, _p, _i, _l, _l, _i]),
    "orc_gen_spd_cond_batch": (None, [_u64, _i, _d, _d, _p, _i, _l, _l, _i, _p]),
    "orc_gen_rhs_batch": (None, [_u64, _i, _i, _p, _i, _l, _l, _i]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("oracle not built: %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ch(c):
    return c.encode() if isinstance(c, str) else c


def fmat(rows, cols, dtype=np.float64):
    return np.zeros((rows, cols), dtype=dtype, order="F")


# ---------------- convenience wrappers (single matrix) ----------------
def potrf(A):
    L = np.array(A, dtype=np.float64, order="F", copy=True)
    info = lib().orc_potrf(L.shape[0], ptr(L), L.shape[0])
    return L, info


def spotrf(A):
    L = np.array(A, dtype=np.float32, order="F", copy=True)
    info = lib().orc_spotrf(L.shape[0], ptr(L), L.shape[0])
    return L, info


def trsm(side, trans, alpha, A, B):
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = X.shape
    lib().orc_trsm(ch(side), ch(trans), m, n, alpha, ptr(A), A.shape[0], ptr(X), m)
    return X


def syrk(trans, alpha, A, beta, B):
    C = np.array(B, dtype=np.float64, order="F", copy=True)
    A = np.asfortranarray(A, dtype=np.float64)
    m = C.shape[0]
    k = A.shape[1] if trans == "N" else A.shape[0]
    lib().orc_syrk(ch(trans), m, k, alpha, ptr(A), A.shape[0], beta, ptr(C), m)
    return C


def gemm(ta, tb, alpha, A, B, beta, C):
    R = np.array(C, dtype=np.float64, order="F", copy=True)
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m, n = R.shape
    k = A.shape[1] if ta == "N" else A.shape[0]
    lib().orc_gemm(ch(ta), ch(tb), m, n, k, alpha, ptr(A), A.shape[0], ptr(B), B.shape[0], beta, ptr(R), m)
    return R


def trtri(diag, L):
    """lower triangular inverse in place of a copy; returns (X, info)"""
    X = np.array(L, dtype=np.float64, order="F", copy=True)
    info = lib().orc_trtri(ch(diag), X.shape[0], ptr(X), X.shape[0])
    return X, info


def pstrf(A):
    """pivoted Cholesky of the upper triangle of A; returns (R, piv (0-based), rank)"""
    R = np.array(A, dtype=np.float64, order="F", copy=True)
    n = R.shape[0]
    piv = np.zeros(max(n, 1), dtype=np.int32)
    rank = lib().orc_pstrf(n, ptr(R), n, ptr(piv))
    return R, piv[:n], rank


def ara_potrf(G, bs, d, dim=None):
    """rank-revealing Cholesky of the lower triangle of G's leading dim block;
    returns (R (dim' x dim', dim' = min(dim, bs)), d (bs entries), rank)"""
    G = np.asfortranarray(G, dtype=np.float64)
    dim = G.shape[0] if dim is None else dim
    dd = np.array(d, dtype=np.float64, copy=True)
    assert dd.shape[0] >= bs
    m = max(min(dim, bs), 1)
    R = np.full((m, m), np.nan, order="F")
    rank = lib().orc_ara_potrf(dim, bs, ptr(G), max(G.shape[0], 1), ptr(R), m, ptr(dd))
    return R[:max(min(dim, bs), 0), :max(min(dim, bs), 0)], dd, rank


def ara_syrk(A):
    """A^T A summed over the rows in order (A given as float64 values)"""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    B = np.zeros((n, n), order="F")
    lib().orc_ara_syrk(m, n, ptr(A), max(m, 1), ptr(B), max(n, 1))
    return B


def ara_trsm(B, R, bs):
    """B(:, 0:min(cols, bs)) R^-1 with R upper"""
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    R = np.asfortranarray(R, dtype=np.float64)
    rows, cols = X.shape
    lib().orc_ara_trsm(rows, cols, bs, ptr(X), max(rows, 1), ptr(R), max(R.shape[0], 1))
    return X


def _zpair(v):
    v = complex(v)
    return np.array([v.real, v.imag], dtype=np.float64)


def zsqrt(v):
    out = np.zeros(2)
    lib().orc_zsqrt(complex(v).real, complex(v).imag, ptr(out))
    return complex(out[0], out[1])


def zpotrf(A, upper=False):
    """complex symmetric L L^T (or U^T U) of the lower (upper) triangle; returns (factor array, info)"""
    X = np.array(A, dtype=np.complex128, order="F", copy=True)
    n = X.shape[0]
    info = lib().orc_zpotrf(int(upper), n, ptr(X), max(n, 1))
    return X, info


def zpotrs(L, B, side="L", upper=False):
    """(Lv Lv^T) X = B: side L -> B is n x nrhs, side R -> B is nrhs x n"""
    L = np.asfortranarray(L, dtype=np.complex128)
    X = np.array(B, dtype=np.complex128, order="F", copy=True)
    n = L.shape[0]
    nrhs = X.shape[1] if side == "L" else X.shape[0]
    lib().orc_zpotrs(int(upper), int(side == "L"), n, nrhs, ptr(L), max(n, 1), ptr(X), max(X.shape[0], 1))
    return X


def zgemm(ta, tb, alpha, A, B, beta, C):
    A = np.asfortranarray(A, dtype=np.complex128)
    B = np.asfortranarray(B, dtype=np.complex128)
    X = np.array(C, dtype=np.complex128, order="F", copy=True)
    m, n = X.shape
    k = A.shape[1] if ta == "N" else A.shape[0]
    lib().orc_zgemm(ch(ta), ch(tb), m, n, k, ptr(_zpair(alpha)), ptr(A), max(A.shape[0], 1), ptr(B),
                    max(B.shape[0], 1), ptr(_zpair(beta)), ptr(X), max(m, 1))
    return X


def zsyrk(upper, trans, alpha, A, beta, C):
    A = np.asfortranarray(A, dtype=np.complex128)
    X = np.array(C, dtype=np.complex128, order="F", copy=True)
    n = X.shape[0]
    k = A.shape[1] if trans == "N" else A.shape[0]
    lib().orc_zsyrk(int(upper), ch(trans), n, k, ptr(_zpair(alpha)), ptr(A), max(A.shape[0], 1),
                    ptr(_zpair(beta)), ptr(X), max(n, 1))
    return X


def ztrsm(side, upper, trans, unit, alpha, A, B):
    A = np.asfortranarray(A, dtype=np.complex128)
    X = np.array(B, dtype=np.complex128, order="F", copy=True)
    m, n = X.shape
    lib().orc_ztrsm(ch(side), int(upper), ch(trans), int(unit), m, n, ptr(_zpair(alpha)), ptr(A),
                    max(A.shape[0], 1), ptr(X), max(m, 1))
    return X


def ztrtri(L, unit=False):
    X = np.array(L, dtype=np.complex128, order="F", copy=True)
    info = lib().orc_ztrtri(int(unit), X.shape[0], ptr(X), max(X.shape[0], 1))
    return X, info


def zlauum(L):
    X = np.array(L, dtype=np.complex128, order="F", copy=True)
    lib().orc_zlauum(X.shape[0], ptr(X), max(X.shape[0], 1))
    return X


def zpotri(L):
    X = np.array(L, dtype=np.complex128, order="F", copy=True)
    info = lib().orc_zpotri(X.shape[0], ptr(X), max(X.shape[0], 1))
    return X, info


def ztrmm(side, trans, unit, alpha, L, B):
    L = np.asfortranarray(L, dtype=np.complex128)
    X = np.array(B, dtype=np.complex128, order="F", copy=True)
    m, n = X.shape
    lib().orc_ztrmm(ch(side), ch(trans), int(unit), m, n, ptr(_zpair(alpha)), ptr(L), max(L.shape[0], 1), ptr(X),
                    max(m, 1))
    return X


def lauum(L):
    X = np.array(L, dtype=np.float64, order="F", copy=True)
    lib().orc_lauum(X.shape[0], ptr(X), X.shape[0])
    return X


def potri(L):
    X = np.array(L, dtype=np.float64, order="F", copy=True)
    info = lib().orc_potri(X.shape[0], ptr(X), X.shape[0])
    return X, info


def trmm(side, trans, diag, alpha, A, B):
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = X.shape
    lib().orc_trmm(ch(side), ch(trans), ch(diag), m, n, alpha, ptr(A), A.shape[0], ptr(X), m)
    return X


def potrs(side, L, B):
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    L = np.asfortranarray(L, dtype=np.float64)
    m, n = X.shape
    lib().orc_potrs(ch(side), m, n, ptr(L), L.shape[0], ptr(X), m)
    return X


def posv(side, A, B):
    L = np.array(A, dtype=np.float64, order="F", copy=True)
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    m, n = X.shape
    info = lib().orc_posv(ch(side), m, n, ptr(L), L.shape[0], ptr(X), m)
    return L, X, info


def sposv(side, A, B):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    B = np.asfortranarray(B, dtype=np.float64)
    X = np.zeros_like(B, order="F")
    m, n = B.shape
    it = ctypes.c_int(0)
    info = lib().orc_sposv(ch(side), m, n, ptr(A), A.shape[0], ptr(B), m, ptr(X), m, ctypes.byref(it))
    return X, info, it.value


def potrf_berr(A, L):
    A = np.asfortranarray(A, dtype=np.float64)
    L = np.asfortranarray(L, dtype=np.float64)
    return lib().orc_potrf_berr(A.shape[0], ptr(A), A.shape[0], ptr(L), L.shape[0])


def solve_berr(side, A, X, B):
    A = np.asfortranarray(A, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m, n = B.shape
    return lib().orc_solve_berr(ch(side), m, n, ptr(A), A.shape[0], ptr(X), m, ptr(B), m)


def gen_spd_shift(seed, mid, n, shift=None):
    A = fmat(n, n)
    lib().orc_gen_spd_shift(seed, mid, n, float(n if shift is None else shift), ptr(A), n)
    return A


def gen_spd_cond(seed, mid, n, lk_lo, lk_hi):
    A = fmat(n, n)
    k = lib().orc_gen_spd_cond(seed, mid, n, lk_lo, lk_hi, ptr(A), n)
    return A, k


def gen_rhs(seed, mid, rows, cols):
    B = fmat(rows, cols)
    lib().orc_gen_rhs(seed, mid, rows, cols, ptr(B), rows)
    return B


# ---------------- batches: arrays of shape (batch, rows, cols), each matrix column-major ----
def batch_array(batch, rows, cols, dtype=np.float64):
    """(batch, cols, rows) C-order storage == batch of column-major rows x cols matrices."""
    return np.zeros((batch, cols, rows), dtype=dtype)


def gen_spd_shift_batch(seed, n, batch, first=0, shift=None):
    A = batch_array(batch, n, n)
    lib().orc_gen_spd_shift_batch(seed, n, float(n if shift is None else shift), ptr(A), n, n * n, first, batch)
    return A


def gen_spd_cond_batch(seed, n, batch, lk_lo, lk_hi, first=0):
    A = batch_array(batch, n, n)
    kap = np.zeros(batch)
    lib().orc_gen_spd_cond_batch(seed, n, lk_lo, lk_hi, ptr(A), n, n * n, first, batch, ptr(kap))
    return A, kap


def gen_rhs_batch(seed, rows, cols, batch, first=0):
    B = batch_array(batch, rows, cols)
    lib().orc_gen_rhs_batch(seed, rows, cols, ptr(B), rows, rows * cols, first, batch)
    return B


def col(Ab, i):
    """matrix i of a batch array as a (rows, cols) Fortran-ordered view"""
    return Ab[i].T
