"""GPU tests of the non-uniform (vbatch) Cholesky family: one persistent
launch over a device-sorted queue, pointer arrays and device size arrays,
orders on both sides of the register/blocked boundary, against the oracle.
Also the workspace protocol (query -> allocate -> call)."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu
EPS = kt.EPS64


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def make_batch(ns, nrhs, side, seed, cond=None):
    torch = kt.torch_dev()
    As, Bs = [], []
    for i, n in enumerate(ns):
        n = int(n)
        if cond is None:
            A = po.gen_spd_shift(seed, i, n)
        else:
            A, _ = po.gen_spd_cond(seed, i, n, cond[0], cond[1])
        B = po.gen_rhs(seed + 1, i, n, nrhs) if side == "L" else po.gen_rhs(seed + 1, i, nrhs, n)
        As.append(A)
        Bs.append(B)
    dA = [kt.to_dev(np.asfortranarray(A).T.copy()) for A in As]
    dB = [kt.to_dev(np.asfortranarray(B).T.copy()) for B in Bs]
    pA = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
    pB = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
    return As, Bs, dA, dB, pA, pB


def alloc_ws(h, maxn, batch):
    h.lib.kblas_posv_vbatch_wsquery(h.h, maxn, batch)
    assert h.lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success


@pytest.mark.parametrize("side", ["L", "R"])
def test_dposv_vbatch_mixed_orders(h, side):
    rng = np.random.default_rng(33)
    batch = 60
    ns = rng.integers(1, 200, batch).astype(np.int32)
    ns[:4] = [1, 64, 65, 8]
    nrhs = 4
    As, Bs, dA, dB, pA, pB = make_batch(ns, nrhs, side, 71)
    alloc_ws(h, 200, batch)
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    if side == "L":
        dm, dnn = dn, dr          # B is n x nrhs
        ldb = dn
    else:
        dm, dnn = dr, dn          # B is nrhs x n
        ldb = dr
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_vbatch(h.h, side.encode(), b"L", kt.P(dm), kt.P(dnn), 0, 0, kt.P(pA), kt.P(dn),
                                 kt.P(pB), kt.P(ldb), batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        n = int(ns[i])
        L = kt.to_host(dA[i]).T
        X = kt.to_host(dB[i]).T
        assert po.potrf_berr(As[i], L) <= 4 * n * EPS, (i, n)
        _, Xo, _ = po.posv(side, As[i], Bs[i])
        assert np.max(np.abs(X - Xo)) <= 64 * n * EPS * max(1.0, np.max(np.abs(Xo))), (i, n)


def test_dpotrf_vbatch_and_info(h):
    rng = np.random.default_rng(5)
    batch = 40
    ns = rng.integers(2, 150, batch).astype(np.int32)
    As, _, dA, _, pA, _ = make_batch(ns, 1, "L", 91)
    # plant a non-positive pivot in one small and one large matrix
    for i, p in ((3, 2), (7, 1)):
        ns_i = int(ns[i])
        As[i][p - 1, p - 1] = -5.0
        dA[i].copy_(kt.to_dev(np.asfortranarray(As[i]).T.copy()))
    alloc_ws(h, 150, batch)
    dn = kt.to_dev(ns)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDpotrf_vbatch(h.h, b"L", kt.P(dn), 0, kt.P(pA), kt.P(dn), batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    for i in range(batch):
        _, io = po.potrf(As[i])
        assert inf[i] == io, (i, int(ns[i]), inf[i], io)
        if io == 0:
            assert po.potrf_berr(As[i], kt.to_host(dA[i]).T) <= 4 * int(ns[i]) * EPS


def test_dsposv_vbatch(h):
    rng = np.random.default_rng(8)
    batch = 30
    ns = rng.integers(8, 100, batch).astype(np.int32)
    nrhs = 4
    As, Bs, dA, dB, pA, pB = make_batch(ns, nrhs, "R", 17, cond=(2.0, 6.0))
    torch = kt.torch_dev()
    dX = [torch.zeros_like(t) for t in dB]
    pX = torch.tensor([t.data_ptr() for t in dX], dtype=torch.int64, device="cuda")
    alloc_ws(h, 100, batch)
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    it = kt.to_dev(np.full(batch, -99, np.int32))
    st = h.lib.kblasDsposv_vbatch(h.h, b"R", b"L", kt.P(dr), kt.P(dn), 0, 0, kt.P(pA), kt.P(dn), kt.P(pB),
                                  kt.P(dr), kt.P(pX), kt.P(dr), batch, kt.P(info), kt.P(it))
    assert st == pykblas.KBLAS_Success
    itr = kt.to_host(it)
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        X = kt.to_host(dX[i]).T
        assert po.solve_berr("R", As[i], X, Bs[i]) <= 1e-13, (i, int(ns[i]), itr[i])
        if ns[i] > 64:
            assert itr[i] == -1


def test_vbatch_requires_workspace():
    hh = kt.handle()
    try:
        dn = kt.to_dev(np.array([4, 5], np.int32))
        pA = kt.to_dev(np.zeros(2, np.int64))
        st = hh.lib.kblasDpotrf_vbatch(hh.h, b"L", kt.P(dn), 0, kt.P(pA), kt.P(dn), 2, None)
        assert st == pykblas.KBLAS_InsufficientWorkspace
    finally:
        hh.close()


def test_sposv_vbatch_single_precision(h):
    # FP32 non-uniform: both phases (orders above and below 64)
    rng = np.random.default_rng(41)
    batch = 24
    ns = rng.integers(2, 180, batch).astype(np.int32)
    ns[:3] = [65, 64, 1]
    nrhs = 2
    torch = kt.torch_dev()
    As, Bs = [], []
    for i, n in enumerate(ns):
        As.append(po.gen_spd_shift(51, i, int(n)))
        Bs.append(po.gen_rhs(52, i, int(n), nrhs))
    dA = [kt.to_dev(np.asfortranarray(A).T.copy(), np.float32) for A in As]
    dB = [kt.to_dev(np.asfortranarray(B).T.copy(), np.float32) for B in Bs]
    pA = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
    pB = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
    alloc_ws(h, 180, batch)
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasSposv_vbatch(h.h, b"L", b"L", kt.P(dn), kt.P(dr), 0, 0, kt.P(pA), kt.P(dn), kt.P(pB), kt.P(dn),
                                 batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        n = int(ns[i])
        X = kt.to_host(dB[i]).T.astype(np.float64)
        assert po.solve_berr("L", As[i], X, Bs[i]) <= 8 * n * kt.EPS32, (i, n)
