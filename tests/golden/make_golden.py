"""Writes the golden fixtures in this folder from LAPACK/BLAS (Intel MKL, the
vendor the reference's tests link: testing/Makefile:44, make.inc:30-32).

The reference's own tests hold no golden vectors (they compare against LAPACK
at run time with unseeded rand(), testing/testing_helper.cu:361-367), so the
fixtures are produced the same way the reference's tests produce their
expected values, but seeded and committed:
  - SPD inputs built like testing/testing_helper.cu:361-403 (U(0,1) entries,
    A_ii += n, lower mirrored to upper), from numpy's seeded generator;
  - expected outputs from MKL dpotrf / dtrsm / dsyrk / dgemm / dpotrs, with
    POSV side 'R' computed exactly as testing/batch_triangular/test_Xposv_batch.cpp:313-326
    (dpotrf, then dtrsm(R,L,T,N) and dtrsm(R,L,N,N)).
The inverse family (dtrtri / dlauum / dpotri, and dtrmm) and dpstrf go to a file of
their own, lapack_golden_inv.npz, written by ``--inv`` from its own seeded generator.
Mixed precision (MKL dsposv / spotrf) goes to lapack_golden_mixed.npz, written by ``--mixed``.
The randomized-range-finder helpers (Gram matrices of FP32 samples: MKL dsyrk of
the widened samples, dpotrf of the Gram matrix, dtrsm with its transpose) go to
lapack_golden_ara.npz, written by ``--ara``.
Run in the build container only (needs MKL): python tests/golden/make_golden.py [--inv | --mixed | --ara]
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MKL_PATHS = ["/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.1"]


def load_mkl():
    for p in MKL_PATHS:
        try:
            lib = ctypes.CDLL(p)
            lib.MKL_Get_Version_String.argtypes = [ctypes.c_char_p, ctypes.c_int]
            buf = ctypes.create_string_buffer(256)
            lib.MKL_Get_Version_String(buf, 256)
            return lib, buf.value.decode()
        except OSError:
            continue
    sys.exit("MKL not found")


mkl, MKL_VERSION = load_mkl()
ci = ctypes.c_int
cd = ctypes.c_double
cc = ctypes.c_char


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def byr(x):
    return ctypes.byref(x)


def dpotrf(A):
    L = np.array(A, order="F", copy=True)
    n = ci(L.shape[0]); info = ci(0)
    mkl.dpotrf_(byr(cc(b"L")), byr(n), P(L), byr(n), byr(info))
    return L, info.value


def dtrsm(side, trans, alpha, A, B):
    X = np.array(B, order="F", copy=True)
    m, n = X.shape
    lda = A.shape[0]
    mkl.dtrsm_(byr(cc(side)), byr(cc(b"L")), byr(cc(trans)), byr(cc(b"N")), byr(ci(m)), byr(ci(n)),
               byr(cd(alpha)), P(np.asfortranarray(A)), byr(ci(lda)), P(X), byr(ci(m)))
    return X


def dsyrk(trans, alpha, A, beta, C):
    R = np.array(C, order="F", copy=True)
    n = R.shape[0]
    k = A.shape[1] if trans == b"N" else A.shape[0]
    mkl.dsyrk_(byr(cc(b"L")), byr(cc(trans)), byr(ci(n)), byr(ci(k)), byr(cd(alpha)),
               P(np.asfortranarray(A)), byr(ci(A.shape[0])), byr(cd(beta)), P(R), byr(ci(n)))
    return R


def dgemm(ta, tb, alpha, A, B, beta, C):
    R = np.array(C, order="F", copy=True)
    m, n = R.shape
    k = A.shape[1] if ta == b"N" else A.shape[0]
    mkl.dgemm_(byr(cc(ta)), byr(cc(tb)), byr(ci(m)), byr(ci(n)), byr(ci(k)), byr(cd(alpha)),
               P(np.asfortranarray(A)), byr(ci(A.shape[0])), P(np.asfortranarray(B)), byr(ci(B.shape[0])),
               byr(cd(beta)), P(R), byr(ci(m)))
    return R


def dpotrs(L, B):
    X = np.array(B, order="F", copy=True)
    n, nrhs = X.shape
    info = ci(0)
    mkl.dpotrs_(byr(cc(b"L")), byr(ci(n)), byr(ci(nrhs)), P(np.asfortranarray(L)), byr(ci(n)), P(X),
                byr(ci(n)), byr(info))
    return X


def make_hpd(rng, n):
    """testing/testing_helper.cu:361-367 + 394-403"""
    A = np.asfortranarray(rng.random((n, n)))
    for i in range(n):
        A[i, i] += n
        for j in range(i):
            A[j, i] = A[i, j]
    return A


def main():
    rng = np.random.default_rng(20181114)
    out = {"mkl_version": np.array(MKL_VERSION)}
    sizes = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33, 48, 64, 65, 100]
    for n in sizes:
        A = make_hpd(rng, n)
        L, info = dpotrf(A)
        assert info == 0
        out["potrf_A_%d" % n] = A
        out["potrf_L_%d" % n] = np.tril(L)
        # POSV side R (reference layout): B is m x n, m right-hand sides as rows
        for m in (1, 4):
            B = np.asfortranarray(rng.standard_normal((40, n)))
            X = dtrsm(b"R", b"T", 1.0, L, B)
            X = dtrsm(b"R", b"N", 1.0, L, X)
            out["posvR_B_%d_%d" % (n, m)] = B
            out["posvR_X_%d_%d" % (n, m)] = X
        # POSV side L (LAPACK layout): B is n x nrhs
        B = np.asfortranarray(rng.standard_normal((n, 3)))
        out["posvL_B_%d" % n] = B
        out["posvL_X_%d" % n] = dpotrs(L, B)
    # TRSM: all four lower/non-unit variants with alpha != 1
    for n in (5, 16, 24, 40):
        A = make_hpd(rng, n)
        L, _ = dpotrf(A)
        L = np.tril(L)
        for side in (b"L", b"R"):
            for trans in (b"N", b"T"):
                shape = (n, 6) if side == b"L" else (6, n)
                B = np.asfortranarray(rng.standard_normal(shape))
                key = "trsm_%s%s_%d" % (side.decode(), trans.decode(), n)
                out[key + "_A"] = L
                out[key + "_B"] = B
                out[key + "_X"] = dtrsm(side, trans, 0.75, L, B)
    # SYRK lower, both transposes
    for (m, k) in ((8, 8), (16, 32), (33, 17)):
        for trans in (b"N", b"T"):
            A = np.asfortranarray(rng.standard_normal((m, k) if trans == b"N" else (k, m)))
            C = np.asfortranarray(rng.standard_normal((m, m)))
            key = "syrk_%s_%d_%d" % (trans.decode(), m, k)
            out[key + "_A"] = A
            out[key + "_C"] = C
            out[key + "_R"] = dsyrk(trans, -1.0, A, 0.5, C)
    # GEMM, all transposes
    for (m, n, k) in ((4, 32, 32), (32, 16, 16), (7, 9, 11)):
        for ta in (b"N", b"T"):
            for tb in (b"N", b"T"):
                A = np.asfortranarray(rng.standard_normal((m, k) if ta == b"N" else (k, m)))
                B = np.asfortranarray(rng.standard_normal((k, n) if tb == b"N" else (n, k)))
                C = np.asfortranarray(rng.standard_normal((m, n)))
                key = "gemm_%s%s_%d_%d_%d" % (ta.decode(), tb.decode(), m, n, k)
                out[key + "_A"] = A
                out[key + "_B"] = B
                out[key + "_C"] = C
                out[key + "_R"] = dgemm(ta, tb, -1.0, A, B, 1.0, C)
    # planted non-SPD: first non-positive leading minor at position p (1-based)
    for n, p in ((8, 5), (20, 1), (40, 33)):
        A = make_hpd(rng, n)
        A[p - 1, p - 1] = -float(n)
        L, info = dpotrf(A)
        out["nonspd_A_%d" % n] = A
        out["nonspd_info_%d" % n] = np.array(info)
    np.savez_compressed(os.path.join(HERE, "lapack_golden.npz"), **out)
    print("wrote", len(out), "arrays with", MKL_VERSION)


def dtrtri(diag, L):
    X = np.array(L, order="F", copy=True)
    n = X.shape[0]; info = ci(0)
    mkl.dtrtri_(byr(cc(b"L")), byr(cc(diag)), byr(ci(n)), P(X), byr(ci(n)), byr(info))
    return X, info.value


def dlauum(L):
    X = np.array(L, order="F", copy=True)
    n = X.shape[0]; info = ci(0)
    mkl.dlauum_(byr(cc(b"L")), byr(ci(n)), P(X), byr(ci(n)), byr(info))
    return X


def dpotri(L):
    X = np.array(L, order="F", copy=True)
    n = X.shape[0]; info = ci(0)
    mkl.dpotri_(byr(cc(b"L")), byr(ci(n)), P(X), byr(ci(n)), byr(info))
    return X, info.value


def dtrmm(side, trans, diag, alpha, A, B):
    X = np.array(B, order="F", copy=True)
    m, n = X.shape
    mkl.dtrmm_(byr(cc(side)), byr(cc(b"L")), byr(cc(trans)), byr(cc(diag)), byr(ci(m)), byr(ci(n)),
               byr(cd(alpha)), P(np.asfortranarray(A)), byr(ci(A.shape[0])), P(X), byr(ci(m)))
    return X


def inv_main():
    """Inverse family: factors L of make_hpd matrices (well conditioned) and
    unit-lower triangles with small off-diagonal entries."""
    rng = np.random.default_rng(20190601)
    out = {"mkl_version": np.array(MKL_VERSION)}
    for n in (1, 2, 3, 5, 8, 16, 17, 32, 33, 64, 65):
        A = make_hpd(rng, n)
        L, info = dpotrf(A)
        assert info == 0
        L = np.asfortranarray(np.tril(L))
        out["inv_A_%d" % n] = A
        out["inv_L_%d" % n] = L
        X, info = dtrtri(b"N", L)
        assert info == 0
        out["trtri_X_%d" % n] = np.tril(X)
        out["lauum_X_%d" % n] = np.tril(dlauum(L))
        X, info = dpotri(L)
        assert info == 0
        out["potri_X_%d" % n] = np.tril(X)
        U = np.asfortranarray(np.tril(rng.uniform(-0.5, 0.5, (n, n)) / max(n, 1) ** 0.5, -1) +
                              np.diag(rng.uniform(2.0, 3.0, n)))   # diagonal ignored with diag = U
        X, info = dtrtri(b"U", U)
        assert info == 0
        out["trtriU_A_%d" % n] = U
        out["trtriU_X_%d" % n] = np.tril(X, -1)
    for n, p in ((6, 3), (20, 1), (40, 40)):
        A = make_hpd(rng, n)
        L, _ = dpotrf(A)
        L = np.asfortranarray(np.tril(L))
        L[p - 1, p - 1] = 0.0
        _, info = dtrtri(b"N", L)
        out["trtri_sing_L_%d" % n] = L
        out["trtri_sing_info_%d" % n] = np.array(info)
    for n in (5, 16, 40):
        A = make_hpd(rng, n)
        L, _ = dpotrf(A)
        L = np.asfortranarray(np.tril(L))
        for side in (b"L", b"R"):
            for trans in (b"N", b"T"):
                for diag in (b"N", b"U"):
                    shape = (n, 6) if side == b"L" else (6, n)
                    B = np.asfortranarray(rng.standard_normal(shape))
                    key = "trmm_%s%s%s_%d" % (side.decode(), trans.decode(), diag.decode(), n)
                    out[key + "_A"] = L
                    out[key + "_B"] = B
                    out[key + "_X"] = dtrmm(side, trans, diag, 0.75, L, B)
    # PSTRF (upper): full-rank SPD and low-rank PSD matrices with distinct diagonals
    for n, r in ((5, 5), (16, 16), (32, 32), (33, 33), (48, 48), (64, 64), (20, 7), (40, 13), (64, 30)):
        Bm = rng.standard_normal((n, r))
        A = np.asfortranarray(Bm @ Bm.T + (np.diag(rng.uniform(0.5, 1.5, n)) if r == n else 0.0))
        R = np.array(A, order="F", copy=True)
        piv = np.zeros(n, dtype=np.int32)
        rank = ci(0); info = ci(0)
        work = np.zeros(2 * n)
        mkl.dpstrf_(byr(cc(b"U")), byr(ci(n)), P(R), byr(ci(n)), P(piv), byr(rank), byr(cd(-1.0)), P(work),
                    byr(info))
        out["pstrf_A_%d_%d" % (n, r)] = A
        out["pstrf_R_%d_%d" % (n, r)] = np.triu(R)
        out["pstrf_piv_%d_%d" % (n, r)] = piv - 1
        out["pstrf_rank_%d_%d" % (n, r)] = np.array(rank.value)
    np.savez_compressed(os.path.join(HERE, "lapack_golden_inv.npz"), **out)
    print("wrote", len(out), "arrays with", MKL_VERSION)


def dsposv(A, B):
    """MKL dsposv('L') on copies: returns (X, ITER, INFO); B is n x nrhs."""
    Aw = np.array(A, order="F", copy=True)
    Bw = np.asfortranarray(B, dtype=np.float64)
    n, nrhs = Bw.shape
    X = np.zeros((n, nrhs), order="F")
    work = np.zeros(n * nrhs)
    swork = np.zeros(n * (n + nrhs), dtype=np.float32)
    it = ci(0); info = ci(0)
    mkl.dsposv_(byr(cc(b"L")), byr(ci(n)), byr(ci(nrhs)), P(Aw), byr(ci(n)), P(Bw), byr(ci(n)), P(X), byr(ci(n)),
                P(work), P(swork), byr(it), byr(info))
    return X, it.value, info.value


def spotrf(A32):
    L = np.array(A32, dtype=np.float32, order="F", copy=True)
    n = ci(L.shape[0]); info = ci(0)
    mkl.spotrf_(byr(cc(b"L")), byr(n), P(L), byr(n), byr(info))
    return L, info.value


def mixed_main():
    """Mixed precision (LAPACK dsposv): the inputs are the oracle's own seeded
    Q diag(s) Q^T generator (seed, matrix id, order, log10 cond range), which the
    fixture records together with a checksum of each generated A so the test can
    regenerate it and prove it is the same matrix; the expected outputs are MKL's
    ITER, INFO and X.  spotrf fixtures hold their (small) inputs explicitly."""
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import pyoracle as po
    out = {"mkl_version": np.array(MKL_VERSION)}
    cases = []
    # n = 64 family of the headline workload (cond 1e2..1e6, 4 right-hand sides) and smaller orders
    for n, cnt in ((16, 6), (32, 6), (64, 12)):
        for mid in range(cnt):
            cases.append((n, 4, 2.0, 6.0, 7100 + n, mid))
    # the large non-uniform family (cond 1e2..1e4), fewer right-hand sides to keep the fixture small
    for n, cnt in ((128, 2), (200, 2), (256, 2), (333, 1), (512, 2)):
        for mid in range(cnt):
            cases.append((n, 8, 2.0, 4.0, 7500 + n, mid))
    # beyond FP32: cond 1e7..1e9 (refinement stalls or the FP32 factor fails) and 1e12
    for n, lo, hi, cnt in ((32, 7.0, 9.0, 4), (64, 7.5, 8.5, 4), (48, 12.0, 12.0, 1)):
        for mid in range(cnt):
            cases.append((n, 3, lo, hi, 7900 + n, mid))
    meta = []
    for k, (n, nrhs, lo, hi, seed, mid) in enumerate(cases):
        A, kappa = po.gen_spd_cond(seed, mid, n, lo, hi)
        B = po.gen_rhs(seed + 1, mid, n, nrhs)
        X, it, info = dsposv(A, B)
        meta.append([n, nrhs, lo, hi, seed, mid, kappa, it, info, float(np.sum(A)), float(np.sum(A * A))])
        out["sposv_X_%d" % k] = X
    out["sposv_meta"] = np.array(meta)
    # ITER = -2: an entry of B beyond the FP32 range (LAPACK's DLAG2S check)
    A, _ = po.gen_spd_cond(7990, 0, 16, 2.0, 3.0)
    B = po.gen_rhs(7991, 0, 16, 2)
    B[3, 1] = 1e39
    X, it, info = dsposv(A, B)
    out["sposv_ovf_A"] = A
    out["sposv_ovf_B"] = B
    out["sposv_ovf_X"] = X
    out["sposv_ovf_iter"] = np.array(it)
    # ITER = -2 from A: scaled beyond the FP32 range (B within it)
    As = A * 1e39
    X, it, info = dsposv(As, B / 1e30)
    out["sposv_ovfA_A"] = As
    out["sposv_ovfA_B"] = B / 1e30
    out["sposv_ovfA_X"] = X
    out["sposv_ovfA_iter"] = np.array(it)
    # spotrf on FP32 copies of the same family (orders 16..64)
    for n in (16, 33, 64):
        A, _ = po.gen_spd_cond(7800 + n, 0, n, 2.0, 4.0)
        A32 = np.asfortranarray(A.astype(np.float32))
        L, info = spotrf(A32)
        assert info == 0
        out["spotrf_A_%d" % n] = A32
        out["spotrf_L_%d" % n] = np.tril(L)
    np.savez_compressed(os.path.join(HERE, "lapack_golden_mixed.npz"), **out)
    print("wrote", len(out), "arrays with", MKL_VERSION)
    its = [int(m[7]) for m in meta]
    print("ITER values:", its)


def ara_main():
    """Gram matrices of FP32 sample blocks Y (m x n, entries U(-1,1) rounded to
    FP32), as the range finder forms them: G = Y^T Y from MKL dsyrk('L','T') of
    the FP64-widened samples (lower, mirrored), its MKL dpotrf factor L
    (R = L^T), and X = B R^-1 from MKL dtrsm('R','L','T') with L, for B of
    40 x n N(0,1) entries."""
    rng = np.random.default_rng(20240613)
    out = {"mkl_version": np.array(MKL_VERSION)}
    for m, n in ((40, 1), (64, 7), (100, 16), (96, 17), (250, 29), (300, 32)):
        Y = np.asfortranarray(rng.uniform(-1.0, 1.0, (m, n)).astype(np.float32))
        G = dsyrk(b"T", 1.0, Y.astype(np.float64), 0.0, np.zeros((n, n), order="F"))
        G = np.tril(G) + np.tril(G, -1).T
        L, info = dpotrf(G)
        assert info == 0
        B = np.asfortranarray(rng.standard_normal((40, n)))
        X = dtrsm(b"R", b"T", 1.0, np.tril(L), B)
        key = "ara_%d_%d" % (m, n)
        out[key + "_Y"] = Y
        out[key + "_G"] = G
        out[key + "_L"] = np.tril(L)
        out[key + "_B"] = B
        out[key + "_X"] = X
    np.savez_compressed(os.path.join(HERE, "lapack_golden_ara.npz"), **out)
    print("wrote", len(out), "arrays with", MKL_VERSION)


if __name__ == "__main__":
    if "--ara" in sys.argv[1:]:
        ara_main()
    elif "--inv" in sys.argv[1:]:
        inv_main()
    elif "--mixed" in sys.argv[1:]:
        mixed_main()
    else:
        main()
