"""GPU tests of the complex (C, Z) batched entry points and of the real
upper-triangle and unit-diagonal variants, through the C ABI against the
oracle's complex restatement (orc_zpotrf / orc_zpotrs / orc_ztrsm / orc_zsyrk /
orc_zgemm; real data runs through it with zero imaginary parts).  Strided and
pointer-array batches, padded leading dimensions (the padding and the other
triangle filled with NaN and checked untouched), planted zero pivots."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu

PREC = {  # prefix -> (numpy dtype, eps, ctypes scalar)
    "Z": (np.complex128, np.finfo(np.float64).eps, pykblas.CD),
    "C": (np.complex64, np.finfo(np.float32).eps, pykblas.CF),
    "D": (np.float64, np.finfo(np.float64).eps, None),
    "S": (np.float32, np.finfo(np.float32).eps, None),
}


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def scalar(p, v):
    t = PREC[p][2]
    return t(v.real, v.imag) if t is not None else float(v.real)


def crand(rng, p, *shape):
    x = rng.standard_normal(shape)
    if PREC[p][0] in (np.complex128, np.complex64):
        x = x + 1j * rng.standard_normal(shape)
    return x


def stage(mats, ld, dtype):
    """column-major matrices with leading dimension ld (padding NaN) -> strided device buffer"""
    ncols = max(M.shape[1] for M in mats)
    H = np.full((len(mats), ncols, ld), np.nan, dtype=np.complex128)
    for b, M in enumerate(mats):
        H[b, :M.shape[1], :M.shape[0]] = M.T
    if dtype in (np.float64, np.float32):
        H = H.real
    return kt.to_dev(H.astype(dtype)), ncols * ld


def unstage(buf, b, rows, cols):
    H = kt.to_host(buf)
    return H[b, :cols, :rows].T.astype(np.complex128)


def lfactor(rng, p, n):
    L = np.tril(crand(rng, p, n, n)) * 0.3
    d = rng.uniform(1.0, 2.0, n)
    if PREC[p][0] in (np.complex128, np.complex64):
        d = d + 1j * rng.uniform(-1.0, 1.0, n)
    L[np.diag_indices(n)] = d
    return L


def sym_batch(rng, p, n, batch, upper):
    """A = L L^T (complex symmetric / real SPD) with the unused triangle set to NaN"""
    As, Ls = [], []
    for _ in range(batch):
        L = lfactor(rng, p, n)
        A = (L @ L.T).astype(np.complex128)
        S = A.copy()
        if upper:
            S[np.tril_indices(n, -1)] = np.nan
        else:
            S[np.triu_indices(n, 1)] = np.nan
        As.append(S)
        Ls.append(L)
    return As, Ls


def ptrs(buf, count, stride):
    return kt.ptr_array(buf, count, stride)


@pytest.mark.parametrize("p", ["Z", "C", "D", "S"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [1, 5, 16, 33, 70])
def test_potrf(h, p, upper, n):
    if p in ("D", "S") and not upper:
        pytest.skip("real lower runs the specialised kernels (covered elsewhere)")
    rng = np.random.default_rng(100 + n + 3 * upper)
    dt, eps, _ = PREC[p]
    batch = 6
    As, Ls = sym_batch(rng, p, n, batch, upper)
    As[-1][0, 0] = 0.0                           # planted zero pivot: info 1, matrix left as it was
    ld = n + 3
    buf, stride = stage(As, ld, dt)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    fn = getattr(h.lib, "kblas%spotrf_batch_strided" % p)
    st = fn(h.h, b"U" if upper else b"L", n, kt.P(buf), ld, stride, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    for b in range(batch):
        X = unstage(buf, b, ld, n)
        Xo, io = po.zpotrf(As[b], upper=upper)
        assert inf[b] == io, (b, inf[b], io)
        if io:
            stored = As[b].real if dt in (np.float64, np.float32) else As[b]
            stored = stored.astype(dt).astype(np.complex128)              # as the device holds it
            assert io == 1 and np.array_equal(np.nan_to_num(X[:n], nan=7.0), np.nan_to_num(stored, nan=7.0))
        F, Fo = (np.triu(X[:n]), np.triu(Xo)) if upper else (np.tril(X[:n]), np.tril(Xo))
        if io == 0:
            tol = 64 * n * eps * np.linalg.cond(Ls[b]) * np.max(np.abs(Fo))
            assert np.max(np.abs(F - Fo)) <= tol, (b, np.max(np.abs(F - Fo)), tol)
        other = np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)
        assert np.all(np.isnan(X[:n][other]))                      # the other triangle untouched
        assert np.all(np.isnan(X[n:]))                              # ld padding untouched


@pytest.mark.parametrize("p", ["Z", "C", "D"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("side", ["L", "R"])
def test_posv_and_potrs(h, p, upper, side):
    if p == "D" and not upper:
        pytest.skip("real lower runs the specialised kernels (covered elsewhere)")
    rng = np.random.default_rng(7 + upper + 2 * (side == "L"))
    dt, eps, _ = PREC[p]
    n, nrhs, batch = 29, 3, 5
    As, Ls = sym_batch(rng, p, n, batch, upper)
    Bs = [crand(rng, p, n, nrhs) if side == "L" else crand(rng, p, nrhs, n) for _ in range(batch)]
    if dt in (np.float64, np.float32):
        Bs = [B.real.astype(np.complex128) for B in Bs]
    lda = n + 2
    m_, n_ = (n, nrhs) if side == "L" else (nrhs, n)
    ldb = m_ + 1
    dA, sA = stage(As, lda, dt)
    dB, sB = stage(Bs, ldb, dt)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    pA, pB = ptrs(dA, batch, sA), ptrs(dB, batch, sB)
    st = getattr(h.lib, "kblas%sposv_batch" % p)(h.h, side.encode(), b"U" if upper else b"L", m_, n_, kt.P(pA),
                                                 lda, kt.P(pB), ldb, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    for b in range(batch):
        A = np.where(np.isnan(As[b]), As[b].T, As[b])              # the full symmetric matrix
        X = unstage(dB, b, ldb, n_)[:m_]
        ref = np.linalg.solve(A, Bs[b]) if side == "L" else np.linalg.solve(A.T, Bs[b].T).T
        assert np.max(np.abs(X - ref)) <= 64 * n * eps * np.linalg.cond(A) * np.max(np.abs(ref)), b
        F = unstage(dA, b, lda, n)[:n]
        Xo = po.zpotrs(F, Bs[b], side=side, upper=upper)
        assert np.max(np.abs(X - Xo)) <= 64 * n * eps * np.linalg.cond(A) * np.max(np.abs(Xo)), b
    # POTRS with the factor just computed, strided
    dB2, _ = stage(Bs, ldb, dt)
    st = getattr(h.lib, "kblas%spotrs_batch_strided" % p)(h.h, side.encode(), b"U" if upper else b"L", m_, n_,
                                                          kt.P(dA), lda, sA, kt.P(dB2), ldb, sB, batch)
    assert st == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(dB2), kt.to_host(dB), equal_nan=True)


@pytest.mark.parametrize("p", ["Z", "C", "D"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm(h, p, side, upper, trans, diag):
    if p == "D" and not upper and diag == "N":
        pytest.skip("real lower non-unit: covered by the existing TRSM tests")
    rng = np.random.default_rng(3 + 5 * upper + 11 * (trans == "C"))
    dt, eps, _ = PREC[p]
    m, n, batch = 37, 9, 4
    k = m if side == "L" else n
    Ts, Bs = [], []
    for _ in range(batch):
        T = crand(rng, p, k, k) * 0.1 + np.diag(2.0 + rng.uniform(0, 1, k) + 0j)
        T = (np.triu(T) if upper else np.tril(T)).astype(np.complex128)
        T[np.tril_indices(k, -1) if upper else np.triu_indices(k, 1)] = np.nan
        Ts.append(T)
        Bs.append(crand(rng, p, m, n).astype(np.complex128))
    if dt in (np.float64, np.float32):
        Ts = [T.real.astype(np.complex128) for T in Ts]
        Bs = [B.real.astype(np.complex128) for B in Bs]
    al = 0.75 + 0.25j if dt in (np.complex128, np.complex64) else 0.75 + 0j
    dT, sT = stage(Ts, k + 1, dt)
    dB, sB = stage(Bs, m + 2, dt)
    st = getattr(h.lib, "kblas%strsm_batch_strided" % p)(h.h, side.encode(), b"U" if upper else b"L",
                                                         trans.encode(), diag.encode(), m, n, scalar(p, al), kt.P(dT),
                                                         k + 1, sT, kt.P(dB), m + 2, sB, batch)
    assert st == pykblas.KBLAS_Success
    for b in range(batch):
        X = unstage(dB, b, m + 2, n)
        Xo = po.ztrsm(side, upper, trans, diag == "U", al, np.nan_to_num(Ts[b]), Bs[b])
        assert np.max(np.abs(X[:m] - Xo)) <= 64 * k * eps * max(1.0, np.max(np.abs(Xo))), b
        assert np.all(np.isnan(X[m:]))


@pytest.mark.parametrize("p", ["Z", "C", "D"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_syrk(h, p, upper, trans):
    if p == "D" and not upper:
        pytest.skip("real lower: covered by the existing SYRK tests")
    rng = np.random.default_rng(21 + upper)
    dt, eps, _ = PREC[p]
    n, k, batch = 70, 13, 3
    As = [crand(rng, p, n, k) if trans == "N" else crand(rng, p, k, n) for _ in range(batch)]
    Cs = [crand(rng, p, n, n) for _ in range(batch)]
    if dt in (np.float64, np.float32):
        As = [A.real.astype(np.complex128) for A in As]
        Cs = [C.real.astype(np.complex128) for C in Cs]
    else:
        As = [A.astype(np.complex128) for A in As]
        Cs = [C.astype(np.complex128) for C in Cs]
    al = 1.5 + 0.5j if dt in (np.complex128, np.complex64) else 1.5 + 0j
    be = -0.75 + 0.25j if dt in (np.complex128, np.complex64) else -0.75 + 0j
    ra = As[0].shape[0]
    dA, sA = stage(As, ra, dt)
    dC, sC = stage(Cs, n, dt)
    pA, pC = ptrs(dA, batch, sA), ptrs(dC, batch, sC)   # held: a freed pointer array could be handed out again
    st = getattr(h.lib, "kblas%ssyrk_batch" % p)(h.h, b"U" if upper else b"L", trans.encode(), n, k, scalar(p, al),
                                                 kt.P(pA), ra, scalar(p, be), kt.P(pC), n, batch)
    assert st == pykblas.KBLAS_Success
    for b in range(batch):
        X = unstage(dC, b, n, n)
        Xo = po.zsyrk(upper, trans, al, As[b], be, Cs[b])
        assert np.max(np.abs(X - Xo)) <= 16 * k * eps * max(1.0, np.max(np.abs(Xo))), b
    if p in ("Z", "C"):
        st = getattr(h.lib, "kblas%ssyrk_batch" % p)(h.h, b"L", b"C", n, k, scalar(p, al), kt.P(pA), ra,
                                                     scalar(p, be), kt.P(pC), n, batch)
        assert st == pykblas.KBLAS_Error_WrongInput


@pytest.mark.parametrize("p", ["Z", "C"])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
def test_gemm(h, p, ta, tb):
    rng = np.random.default_rng(31)
    dt, eps, _ = PREC[p]
    m, n, k, batch = 45, 70, 19, 3
    As = [(crand(rng, p, m, k) if ta == "N" else crand(rng, p, k, m)).astype(np.complex128) for _ in range(batch)]
    Bs = [(crand(rng, p, k, n) if tb == "N" else crand(rng, p, n, k)).astype(np.complex128) for _ in range(batch)]
    Cs = [crand(rng, p, m, n).astype(np.complex128) for _ in range(batch)]
    al, be = 0.5 - 1.25j, 2.0 + 0.5j
    dA, sA = stage(As, As[0].shape[0], dt)
    dB, sB = stage(Bs, Bs[0].shape[0], dt)
    dC, sC = stage(Cs, m, dt)
    st = getattr(h.lib, "kblas%sgemm_batch_strided" % p)(h.h, ta.encode(), tb.encode(), m, n, k, scalar(p, al),
                                                         kt.P(dA), As[0].shape[0], sA, kt.P(dB), Bs[0].shape[0], sB,
                                                         scalar(p, be), kt.P(dC), m, sC, batch)
    assert st == pykblas.KBLAS_Success
    for b in range(batch):
        X = unstage(dC, b, m, n)
        Xo = po.zgemm(ta, tb, al, As[b], Bs[b], be, Cs[b])
        assert np.max(np.abs(X - Xo)) <= 16 * k * eps * max(1.0, np.max(np.abs(Xo))), b


@pytest.mark.parametrize("p", ["Z", "C"])
@pytest.mark.parametrize("what", ["trtri", "trtriU", "lauum", "potri", "poti"])
@pytest.mark.parametrize("n", [5, 16, 17, 40, 70])
def test_inverse_family(h, p, what, n):
    rng = np.random.default_rng(500 + n)
    dt, eps, _ = PREC[p]
    batch = 5
    Ls, Ss = [], []
    for _ in range(batch):
        L = lfactor(rng, p, n)
        M = (L @ L.T) if what == "poti" else L
        S = np.tril(M).astype(np.complex128)
        S[np.triu_indices(n, 1)] = np.nan
        Ls.append(L)
        Ss.append(S)
    if what in ("trtri", "potri"):
        Ss[-1][2, 2] = 0.0                     # info 3, matrix left as it was
    ld = n + 1
    buf, stride = stage(Ss, ld, dt)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    if what.startswith("trtri"):
        st = getattr(h.lib, "kblas%strtri_batch_strided" % p)(h.h, b"L", b"U" if what == "trtriU" else b"N", n,
                                                               kt.P(buf), ld, stride, batch, kt.P(info))
    else:
        st = getattr(h.lib, "kblas%s%s_batch_strided" % (p, what))(h.h, b"L", n, kt.P(buf), ld, stride, batch,
                                                                    kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    for b in range(batch):
        X = unstage(buf, b, ld, n)[:n]
        if what == "trtri":
            Xo, io = po.ztrtri(Ss[b])
        elif what == "trtriU":
            Xo, io = po.ztrtri(Ss[b], unit=True)
        elif what == "lauum":
            Xo, io = po.zlauum(Ss[b]), 0
        elif what == "potri":
            Xo, io = po.zpotri(Ss[b])
        else:
            F, io = po.zpotrf(Ss[b])
            Xo, io2 = po.zpotri(F)
            io = io or io2
        assert (inf[b] if what != "lauum" else 0) == io, (b, inf[b], io)
        if io:
            stored = Ss[b].astype(dt).astype(np.complex128)
            assert np.array_equal(np.nan_to_num(X, nan=7.0), np.nan_to_num(stored, nan=7.0))
            continue
        c = np.linalg.cond(Ls[b])
        tol = 64 * n * eps * c * c * max(1.0, np.max(np.abs(np.nan_to_num(Xo))))
        low = np.tril_indices(n, -1) if what == "trtriU" else np.tril_indices(n)
        assert np.max(np.abs(X[low] - Xo[low])) <= tol, (b, np.max(np.abs(X[low] - Xo[low])), tol)
        assert np.all(np.isnan(X[np.triu_indices(n, 1)]))                 # strict upper untouched


@pytest.mark.parametrize("p", ["Z", "C"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("unit", [False, True])
def test_trmm(h, p, side, trans, unit):
    rng = np.random.default_rng(61)
    dt, eps, _ = PREC[p]
    m, n, batch = 21, 34, 3
    k = m if side == "L" else n
    Ls = []
    for _ in range(batch):
        L = np.tril(crand(rng, p, k, k)).astype(np.complex128)
        L[np.triu_indices(k, 1)] = np.nan
        Ls.append(L)
    Bs = [crand(rng, p, m, n).astype(np.complex128) for _ in range(batch)]
    al = 1.5 - 0.5j
    dL, sL = stage(Ls, k, dt)
    dB, sB = stage(Bs, m + 1, dt)
    st = getattr(h.lib, "kblas%strmm_batch_strided" % p)(h.h, side.encode(), b"L", trans.encode(),
                                                         b"U" if unit else b"N", m, n, scalar(p, al), kt.P(dL), k,
                                                         sL, kt.P(dB), m + 1, sB, batch)
    assert st == pykblas.KBLAS_Success
    for b in range(batch):
        X = unstage(dB, b, m + 1, n)
        Xo = po.ztrmm(side, trans, unit, al, np.nan_to_num(Ls[b]), Bs[b])
        assert np.max(np.abs(X[:m] - Xo)) <= 16 * k * eps * max(1.0, np.max(np.abs(Xo))), b
        assert np.all(np.isnan(X[m:]))
