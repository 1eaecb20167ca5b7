"""GPU test of kblasDposv_vbatch_noalloc (the non-uniform FP64 POSV with its
queue in the library's fixed scratch area): its results must be bitwise equal
to kblasDposv_vbatch's (same kernels, workspace from the handle) on two calls
in a row, the second with a smaller batch, and batches above the scratch
capacity are rejected before any launch."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu
NOALLOC_MAX_BATCH = 1 << 20


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def batch_inputs(ns, nrhs, seed):
    torch = kt.torch_dev()
    As, Bs = [], []
    for i, n in enumerate(ns):
        As.append(po.gen_spd_shift(seed, i, int(n)))
        Bs.append(po.gen_rhs(seed + 1, i, int(n), nrhs))      # side L: n x nrhs

    def copy():
        dA = [kt.to_dev(np.asfortranarray(A).T.copy()) for A in As]
        dB = [kt.to_dev(np.asfortranarray(B).T.copy()) for B in Bs]
        pA = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
        pB = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
        return dA, dB, pA, pB
    return As, Bs, copy


def run_pair(h, ns, nrhs, seed):
    batch = len(ns)
    As, Bs, copy = batch_inputs(ns, nrhs, seed)
    dn = kt.to_dev(np.asarray(ns, np.int32))
    dr = kt.to_dev(np.full(batch, nrhs, np.int32))
    # reference call: workspace from the handle
    h.lib.kblas_posv_vbatch_wsquery(h.h, int(max(ns)), batch)
    assert h.lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
    dA1, dB1, pA1, pB1 = copy()
    info1 = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_vbatch(h.h, b"L", b"L", kt.P(dn), kt.P(dr), 0, 0, kt.P(pA1), kt.P(dn), kt.P(pB1),
                                 kt.P(dn), batch, kt.P(info1))
    assert st == pykblas.KBLAS_Success
    # the same batch through the scratch-area entry point
    dA2, dB2, pA2, pB2 = copy()
    info2 = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_vbatch_noalloc(h.h, b"L", b"L", kt.P(dn), kt.P(dr), 0, 0, kt.P(pA2), kt.P(dn),
                                         kt.P(pB2), kt.P(dn), batch, kt.P(info2))
    assert st == pykblas.KBLAS_Success
    i1, i2 = kt.to_host(info1), kt.to_host(info2)
    assert np.array_equal(i1, i2)
    assert np.all(i2 == 0)
    for i in range(batch):
        assert np.array_equal(kt.to_host(dA1[i]), kt.to_host(dA2[i])), i
        assert np.array_equal(kt.to_host(dB1[i]), kt.to_host(dB2[i])), i
    # and the answer is the oracle's (same bar as the vbatch tests)
    for i in range(0, batch, 7):
        n = int(ns[i])
        _, Xo, _ = po.posv("L", As[i], Bs[i])
        X = kt.to_host(dB2[i]).T
        assert np.max(np.abs(X - Xo)) <= 64 * n * kt.EPS64 * max(1.0, np.max(np.abs(Xo))), (i, n)


def test_noalloc_matches_workspace_call_twice(h):
    rng = np.random.default_rng(2024)
    ns = rng.integers(1, 140, 70).astype(np.int32)
    ns[:3] = [1, 64, 65]
    run_pair(h, ns, 3, 501)
    # second call, smaller batch: the scratch area keeps the first call's queue
    ns2 = rng.integers(1, 100, 23).astype(np.int32)
    ns2[0] = 97
    run_pair(h, ns2, 2, 502)


def test_noalloc_rejects_batch_above_capacity(h):
    torch = kt.torch_dev()
    dn = kt.to_dev(np.ones(4, np.int32))
    dummy = torch.zeros(4, dtype=torch.int64, device="cuda")
    info = kt.to_dev(np.full(4, -7, np.int32))
    st = h.lib.kblasDposv_vbatch_noalloc(h.h, b"L", b"L", kt.P(dn), kt.P(dn), 0, 0, kt.P(dummy), kt.P(dn),
                                         kt.P(dummy), kt.P(dn), NOALLOC_MAX_BATCH + 1, kt.P(info))
    assert st == pykblas.KBLAS_InsufficientWorkspace
    assert np.all(kt.to_host(info) == -7)     # nothing ran
