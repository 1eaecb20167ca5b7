"""GPU tests at the benchmark workloads' size ranges and under HIP-graph
capture, against the oracle:
  * non-uniform mixed POSV with orders across [128, 512] (past 257) and 64
    right-hand sides, cond(A) in [1e2, 1e4] (the large-tile mixed workload);
  * non-uniform FP64 POSV with orders across [8, 256] (its whole range);
  * a captured non-uniform POSV step and a captured mixed POSV step, each
    replayed several times with eager work in between, every replay bitwise
    equal to the uncaptured call.
"""
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


def dev_list(mats, dtype=None):
    return [kt.to_dev(np.asfortranarray(M).T.copy(), dtype) for M in mats]


def ptrs(ts):
    torch = kt.torch_dev()
    return torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")


def alloc_ws(h, maxn, batch):
    h.lib.kblas_posv_vbatch_wsquery(h.h, maxn, batch)
    assert h.lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success


def test_dsposv_vbatch_large_orders_64_rhs(h):
    """orders 128..512 (and a few small ones so both phases run), 64
    right-hand sides in the reference's Right layout, cond in [1e2, 1e4]"""
    ns = np.array([128, 512, 258, 300, 129, 257, 384, 511, 450, 200, 40, 64], np.int32)
    nrhs = 64
    batch = len(ns)
    As, Bs, kap = [], [], []
    for i, n in enumerate(ns):
        A, k = po.gen_spd_cond(505, i, int(n), 2.0, 4.0)
        As.append(A)
        kap.append(k)
        Bs.append(po.gen_rhs(506, i, nrhs, int(n)))          # nrhs x n
    dA, dB = dev_list(As), dev_list(Bs)
    torch = kt.torch_dev()
    dX = [torch.zeros_like(t) for t in dB]
    # workspace with the FP32 slots: the orders above 64 refine too
    h.lib.kblas_sposv_vbatch_wsquery(h.h, nrhs, int(ns.max()), batch)
    assert h.lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    it = kt.to_dev(np.full(batch, -99, np.int32))
    pA, pB, pX = ptrs(dA), ptrs(dB), ptrs(dX)      # kept alive until the call has run
    st = h.lib.kblasDsposv_vbatch(h.h, b"R", b"L", kt.P(dr), kt.P(dn), nrhs, int(ns.max()), kt.P(pA), kt.P(dn),
                                  kt.P(pB), kt.P(dr), kt.P(pX), kt.P(dr), batch, kt.P(info), kt.P(it))
    assert st == pykblas.KBLAS_Success
    inf, itr = kt.to_host(info), kt.to_host(it)
    assert np.all(inf == 0)
    for i in range(batch):
        n = int(ns[i])
        X = kt.to_host(dX[i]).T
        be = po.solve_berr("R", As[i], X, Bs[i])
        assert be <= 1e-13, (i, n, itr[i], be)
        _, Xo, _ = po.posv("R", As[i], Bs[i])
        fe = np.max(np.abs(X - Xo)) / np.max(np.abs(Xo))
        assert fe <= kap[i] * 1e-13, (i, n, fe)
        # cond <= 1e4: refinement serves every order, within 2 of LAPACK dsposv's count
        _, _, it_r = po.sposv("R", As[i], Bs[i])
        assert itr[i] >= 0 and it_r >= 0, (i, n, itr[i], it_r)
        assert abs(int(itr[i]) - it_r) <= 2, (i, n, itr[i], it_r)
        assert np.array_equal(kt.to_host(dB[i]).T, Bs[i])    # B read-only
        assert np.array_equal(kt.to_host(dA[i]).T, As[i])    # A read-only when refinement served


def test_dposv_vbatch_whole_range(h):
    """orders across [8, 256], including the top of the range"""
    rng = np.random.default_rng(77)
    ns = rng.integers(8, 257, 36).astype(np.int32)
    ns[:6] = [256, 255, 250, 233, 201, 8]
    nrhs = 4
    batch = len(ns)
    As = [po.gen_spd_shift(78, i, int(n)) for i, n in enumerate(ns)]
    Bs = [po.gen_rhs(79, i, nrhs, int(n)) for i, n in enumerate(ns)]
    dA, dB = dev_list(As), dev_list(Bs)
    alloc_ws(h, 256, batch)
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    pA, pB = ptrs(dA), ptrs(dB)                    # kept alive until the call has run
    st = h.lib.kblasDposv_vbatch(h.h, b"R", b"L", kt.P(dr), kt.P(dn), nrhs, 256, kt.P(pA), kt.P(dn), kt.P(pB),
                                 kt.P(dr), batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        n = int(ns[i])
        L = kt.to_host(dA[i]).T
        X = kt.to_host(dB[i]).T
        assert po.potrf_berr(As[i], L) <= 4 * n * EPS, (i, n)
        _, Xo, _ = po.posv("R", As[i], Bs[i])
        assert np.max(np.abs(X - Xo)) <= 64 * n * EPS * max(1.0, np.max(np.abs(Xo))), (i, n)


def capture(torch, h, fn):
    """fn() issues library calls on the handle's stream; returns the graph"""
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        h.set_stream(torch.cuda.current_stream().cuda_stream)
        fn()
    h.set_stream(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return g


def test_graph_replay_vbatch_posv(h):
    torch = kt.torch_dev()
    rng = np.random.default_rng(90)
    ns = rng.integers(8, 200, 40).astype(np.int32)
    ns[:2] = [199, 65]
    nrhs = 3
    batch = len(ns)
    As = [po.gen_spd_shift(91, i, int(n)) for i, n in enumerate(ns)]
    Bs = [po.gen_rhs(92, i, int(n), nrhs) for i, n in enumerate(ns)]        # side L: n x nrhs
    A0, B0 = dev_list(As), dev_list(Bs)
    dA = [t.clone() for t in A0]
    dB = [t.clone() for t in B0]
    pA, pB = ptrs(dA), ptrs(dB)
    alloc_ws(h, 200, batch)
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    info = kt.to_dev(np.full(batch, -7, np.int32))

    def step():
        assert h.lib.kblasDposv_vbatch(h.h, b"L", b"L", kt.P(dn), kt.P(dr), 0, 0, kt.P(pA), kt.P(dn), kt.P(pB),
                                       kt.P(dn), batch, kt.P(info)) == pykblas.KBLAS_Success

    def restore():
        for a, a0 in zip(dA, A0):
            a.copy_(a0)
        for b, b0 in zip(dB, B0):
            b.copy_(b0)
        info.fill_(-7)

    step()                                   # eager reference
    refA = [kt.to_host(t) for t in dA]
    refB = [kt.to_host(t) for t in dB]
    assert np.all(kt.to_host(info) == 0)
    restore()
    g = capture(torch, h, step)
    for rep in range(3):
        restore()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.all(kt.to_host(info) == 0), rep
        for i in range(batch):
            assert np.array_equal(kt.to_host(dA[i]), refA[i]), (rep, i)
            assert np.array_equal(kt.to_host(dB[i]), refB[i]), (rep, i)


def test_graph_replay_mixed_posv(h):
    torch = kt.torch_dev()
    n, nrhs, batch = 64, 4, 300
    A, kap = po.gen_spd_cond_batch(41, n, batch, 2.0, 6.0)
    B = po.gen_rhs_batch(42, nrhs, n, batch)                                 # side R: nrhs x n
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    dX = torch.zeros_like(dB)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    it = kt.to_dev(np.full(batch, -99, np.int32))

    def step():
        assert h.lib.kblasDsposv_batch_strided(h.h, b"R", b"L", nrhs, n, kt.P(dA), n, n * n, kt.P(dB), nrhs,
                                               nrhs * n, kt.P(dX), nrhs, nrhs * n, batch, kt.P(info),
                                               kt.P(it)) == pykblas.KBLAS_Success

    step()
    refX, refI, refIt = kt.to_host(dX), kt.to_host(info), kt.to_host(it)
    assert np.all(refI == 0) and np.all(refIt >= 0)
    g = capture(torch, h, step)
    for rep in range(3):
        dX.fill_(float("nan"))
        info.fill_(-7)
        it.fill_(-99)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(kt.to_host(dX), refX), rep
        assert np.array_equal(kt.to_host(info), refI), rep
        assert np.array_equal(kt.to_host(it), refIt), rep
    assert np.array_equal(kt.to_host(dA), A)                                 # A read-only


def test_dsposv_vbatch_large_orders_fallbacks(h):
    """orders above 64 on the mixed path: an FP32 factorization that fails
    (cond 1e12) gives ITER = -3, a right-hand side beyond FP32 range -2, both
    solved in FP64 with A := L; the others refine; a workspace without the
    FP32 slots sends every order above 64 to the FP64 path (ITER = -1)"""
    ns = np.array([100, 130, 90, 70, 40], np.int32)
    nrhs = 8
    batch = len(ns)
    As, Bs = [], []
    for i, n in enumerate(ns):
        lk = 12.0 if i == 0 else 3.0
        A, _ = po.gen_spd_cond(707, i, int(n), lk, lk)
        As.append(A)
        B = po.gen_rhs(708, i, int(n), nrhs)                  # side L: n x nrhs
        if i == 1:
            B[:, 3] *= 1e39
        Bs.append(B)
    torch = kt.torch_dev()
    dn = kt.to_dev(ns)
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    for slots in (True, False):
        dA, dB = dev_list(As), dev_list(Bs)
        dX = [torch.zeros_like(t) for t in dB]
        hh = h if slots else kt.handle()        # a fresh handle: only the queue
        if slots:
            hh.lib.kblas_sposv_vbatch_wsquery(hh.h, nrhs, int(ns.max()), batch)
        else:
            hh.lib.kblas_posv_vbatch_wsquery(hh.h, int(ns.max()), batch)
        assert hh.lib.kblasAllocateWorkspace(hh.h) == pykblas.KBLAS_Success
        info = kt.to_dev(np.full(batch, -7, np.int32))
        it = kt.to_dev(np.full(batch, -99, np.int32))
        pA, pB, pX = ptrs(dA), ptrs(dB), ptrs(dX)
        st = hh.lib.kblasDsposv_vbatch(hh.h, b"L", b"L", kt.P(dn), kt.P(dr), int(ns.max()), nrhs, kt.P(pA),
                                       kt.P(dn), kt.P(pB), kt.P(dn), kt.P(pX), kt.P(dn), batch, kt.P(info), kt.P(it))
        assert st == pykblas.KBLAS_Success
        inf, itr = kt.to_host(info), kt.to_host(it)
        if not slots:
            hh.close()
        assert np.all(inf == 0), inf
        want = [-3, -2, None, None, None] if slots else [-1, -1, -1, -1, None]
        for i in range(batch):
            n = int(ns[i])
            X = kt.to_host(dX[i]).T
            assert po.solve_berr("L", As[i], X, Bs[i]) <= 1e-13, (slots, i, itr[i])
            if want[i] is None:
                assert itr[i] >= 0, (slots, i, itr[i])
                assert np.array_equal(kt.to_host(dA[i]).T, As[i])
            else:
                assert itr[i] == want[i], (slots, i, itr[i])
                assert po.potrf_berr(As[i], kt.to_host(dA[i]).T) <= 4 * n * EPS   # A holds L
        if slots:
            for i in (0, 1):
                _, _, it_r = po.sposv("L", As[i], Bs[i])
                assert it_r == want[i], (i, it_r)
