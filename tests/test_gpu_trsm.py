"""GPU parity tests of the batched real TRSM kernels (kb_trsm.hip): the
register kernel for orders up to 64 and one sweep of the blocked solve above
it, against the oracle's TRSM (orc_trsm for lower non-unit triangles,
orc_ztrsm on real data for upper and unit ones).

The shapes cross every order bucket of the register kernel (8, 16, 32, 64,
orders just above and below each), vector counts below, at and above one
wavefront (one to 130), both sides and transpositions, alpha != 1, strided
and pointer-array operands with leading dimensions above the order, NaN in
the unreferenced triangle and in the rows between the order and the leading
dimension (never read, never written), and odd batch counts so that the last
wavefront holds groups with no matrix.  Bar: relative max-norm error within
64 * order * eps of the working precision."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu

PREC = {"D": (np.float64, kt.EPS64), "S": (np.float32, kt.EPS32)}


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def triangle(rng, k, upper, unit):
    T = rng.standard_normal((k, k)) / k + np.diag(1.0 + rng.uniform(0.0, 1.0, k))
    T = np.triu(T) if upper else np.tril(T)
    T[np.tril_indices(k, -1) if upper else np.triu_indices(k, 1)] = np.nan
    if unit:
        T[np.diag_indices(k)] = np.nan     # a unit diagonal is not referenced
    return T


def stage(mats, ld, dt):
    """column-major matrices with leading dimension ld, NaN padding"""
    r, c = mats[0].shape
    buf = np.full((len(mats), c, ld), np.nan)
    for i, M in enumerate(mats):
        buf[i, :, :r] = M.T
    return kt.to_dev(buf.astype(dt)), c * ld


def unstage(t, b, r):
    return kt.to_host(t)[b].T[:r].astype(np.float64)


def run_case(h, p, side, upper, trans, unit, m, n, batch, seed, ptrs):
    dt, eps = PREC[p]
    rng = np.random.default_rng(seed)
    k = m if side == "L" else n
    Ts = [triangle(rng, k, upper, unit) for _ in range(batch)]
    Bs = [rng.standard_normal((m, n)) for _ in range(batch)]
    if dt == np.float32:
        Ts = [T.astype(np.float32).astype(np.float64) for T in Ts]
        Bs = [B.astype(np.float32).astype(np.float64) for B in Bs]
    lda, ldb = k + 3, m + 2
    dT, sT = stage(Ts, lda, dt)
    dB, sB = stage(Bs, ldb, dt)
    alpha = 0.75
    if ptrs:
        pT, pB = kt.ptr_array(dT, batch, sT), kt.ptr_array(dB, batch, sB)
        st = getattr(h.lib, "kblas%strsm_batch" % p)(h.h, side.encode(), b"U" if upper else b"L", trans.encode(),
                                                     b"U" if unit else b"N", m, n, alpha, kt.P(pT), lda, kt.P(pB),
                                                     ldb, batch)
    else:
        st = getattr(h.lib, "kblas%strsm_batch_strided" % p)(h.h, side.encode(), b"U" if upper else b"L",
                                                             trans.encode(), b"U" if unit else b"N", m, n, alpha,
                                                             kt.P(dT), lda, sT, kt.P(dB), ldb, sB, batch)
    assert st == pykblas.KBLAS_Success
    Xall = kt.to_host(dB)
    for b in range(batch):
        X = Xall[b].T.astype(np.float64)
        T = np.nan_to_num(Ts[b])
        if upper or unit:
            Xo = po.ztrsm(side, upper, trans, unit, alpha, T.astype(np.complex128), Bs[b]).real
        else:
            Xo = po.trsm(side, trans, alpha, T, Bs[b])
        err = np.max(np.abs(X[:m] - Xo)) / max(np.max(np.abs(Xo)), 1e-300)
        assert err <= 64 * k * eps, (b, err)
        assert np.all(np.isnan(X[m:])), b


# order -> vector count: below, at and above one wavefront across the buckets
REG_SHAPES = [(1, 5), (3, 64), (8, 8), (9, 65), (16, 3), (17, 16), (31, 1), (32, 130), (33, 2), (48, 32), (63, 7),
              (64, 64)]


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("shape", REG_SHAPES)
def test_register_trsm_lower(h, p, side, trans, shape):
    k, nv = shape
    m, n = (k, nv) if side == "L" else (nv, k)
    run_case(h, p, side, False, trans, False, m, n, batch=11, seed=k * 131 + nv, ptrs=(k % 2 == 1))


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("upper,unit", [(True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(5, 9), (24, 70), (64, 3)])
def test_register_trsm_upper_unit(h, side, trans, upper, unit, shape):
    k, nv = shape
    m, n = (k, nv) if side == "L" else (nv, k)
    run_case(h, "D", side, upper, trans, unit, m, n, batch=5, seed=7 * k + nv + 3 * upper, ptrs=unit)


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("shape", [(65, 1), (96, 33), (130, 64), (257, 100)])
def test_blocked_trsm_lower(h, p, side, trans, shape):
    k, nv = shape
    m, n = (k, nv) if side == "L" else (nv, k)
    run_case(h, p, side, False, trans, False, m, n, batch=3, seed=k + nv, ptrs=(nv % 2 == 1))


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_blocked_trsm_unit_and_upper(h, side, trans):
    # lower unit: the blocked sweep with a unit diagonal; upper above 64: the generic kernel
    run_case(h, "D", side, False, trans, True, *((100, 20) if side == "L" else (20, 100)), batch=3, seed=5, ptrs=False)
    run_case(h, "D", side, True, trans, False, *((90, 12) if side == "L" else (12, 90)), batch=2, seed=6, ptrs=True)


def ptrs(torch, tensors):
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device="cuda")


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("max_order", [12, 50, 150])
def test_nonuniform_trsm_with_bounds(h, p, side, trans, max_order):
    # kblas{D,S}trsm_vbatch with max_m / max_n given: the register (<= 16), matrix-core (<= 64) or blocked
    # kernel chosen by the largest order, each matrix solved at its own sizes; one matrix has a zero size
    dt, eps = PREC[p]
    torch = kt.torch_dev()
    rng = np.random.default_rng(max_order + 7 * (side == "L") + 3 * (trans == "T"))
    batch = 9
    orders = rng.integers(1, max_order + 1, batch)
    orders[0] = max_order
    nvecs = rng.integers(1, 70, batch)
    nvecs[3] = 0
    ms = np.where(side == "L", orders, nvecs).astype(np.int32)
    ns = np.where(side == "L", nvecs, orders).astype(np.int32)
    Ts, Bs, dT, dB = [], [], [], []
    for i in range(batch):
        k = int(orders[i])
        T = triangle(rng, k, False, False)
        B = rng.standard_normal((int(ms[i]), int(ns[i])))
        if dt == np.float32:
            T = T.astype(np.float32).astype(np.float64)
            B = B.astype(np.float32).astype(np.float64)
        Ts.append(T)
        Bs.append(B)
        dT.append(kt.to_dev(np.asfortranarray(T).T.copy(), dt))
        dB.append(kt.to_dev(np.asfortranarray(B).T.copy() if B.size else np.zeros((1, 1)), dt))
    lda = kt.to_dev(orders.astype(np.int32))
    ldb = kt.to_dev(np.maximum(ms, 1).astype(np.int32))
    dm, dn = kt.to_dev(ms), kt.to_dev(ns)
    pT, pB = ptrs(torch, dT), ptrs(torch, dB)   # held until the call has run
    st = getattr(h.lib, "kblas%strsm_vbatch" % p)(h.h, side.encode(), b"L", trans.encode(), b"N", kt.P(dm), kt.P(dn),
                                                  int(ms.max()), int(ns.max()), 0.5, kt.P(pT), kt.P(lda), kt.P(pB),
                                                  kt.P(ldb), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        if ms[i] == 0 or ns[i] == 0:
            continue
        X = kt.to_host(dB[i]).T.astype(np.float64)
        Xo = po.trsm(side, trans, 0.5, np.nan_to_num(Ts[i]), Bs[i])
        err = np.max(np.abs(X - Xo)) / max(np.max(np.abs(Xo)), 1e-300)
        assert err <= 64 * int(orders[i]) * eps, (i, err)


@pytest.mark.parametrize("k", [8, 40, 100])
@pytest.mark.parametrize("side", ["L", "R"])
def test_trsm_alpha_zero_sets_b_without_reading(h, k, side):
    # alpha = 0: B := 0 and B is not read (NaN and Inf in B leave zeros), in each kernel bucket
    rng = np.random.default_rng(k)
    batch, nv = 5, 7
    m, n = (k, nv) if side == "L" else (nv, k)
    T = [triangle(rng, k, False, False) for _ in range(batch)]
    B = np.full((batch, n, m), np.nan)
    B[:, 0, 0] = np.inf
    dT, dB = kt.to_dev(np.stack([t.T for t in T])), kt.to_dev(B)
    st = h.lib.kblasDtrsm_batch_strided(h.h, side.encode(), b"L", b"N", b"N", m, n, 0.0, kt.P(dT), k, k * k,
                                        kt.P(dB), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(dB), np.zeros_like(B))


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("shape", [(72, 64), (129, 65), (200, 96)])
def test_recursive_trsm_lower(h, p, side, trans, shape):
    # orders above 64 with at least 64 vectors: half-order recursion with
    # matrix-core GEMM updates, leaves of order 33..64; pointer arrays for
    # odd orders (sub-block offsets on top of each matrix's pointer)
    k, nv = shape
    m, n = (k, nv) if side == "L" else (nv, k)
    run_case(h, p, side, False, trans, False, m, n, batch=5, seed=3 * k + nv, ptrs=(k % 2 == 1))


@pytest.mark.parametrize("side", ["L", "R"])
def test_recursive_trsm_unit_diagonal(h, side):
    run_case(h, "D", side, False, "T", True, *((150, 70) if side == "L" else (70, 150)), batch=3, seed=11, ptrs=True)
