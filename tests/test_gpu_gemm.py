"""GPU parity tests of the matrix-core batched GEMM / SYRK kernel (kb_gemm.hip)
against the oracle (orc_gemm, orc_syrk; orc_zsyrk on real data for the upper
triangle).

Shapes cross both tile sizes (16 and 32) and their edges, inner dimensions
around the K chunk (1, 3, 31, 32, 33, 65, 130), every transposition pair,
alpha / beta combinations including beta = 0 with NaN in C (C must not be
read), batches that are not a multiple of 8 (the launch numbers matrices in
groups of 8), leading dimensions above the rows with NaN padding (never read,
never written) and pointer arrays.  SYRK leaves the other triangle of C
bit-unchanged.  Bar: relative max-norm error within 16 * k * eps."""
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


def stage(mats, ld, dt):
    r, c = mats[0].shape
    buf = np.full((len(mats), c, ld), np.nan)
    for i, M in enumerate(mats):
        buf[i, :, :r] = M.T
    return kt.to_dev(buf.astype(dt)), c * ld


def rnd(rng, shape, dt):
    return rng.standard_normal(shape).astype(dt).astype(np.float64)


# the last four take the 64 x 64 workgroup tiles (m, n >= 64): K chunks of 16 with a partial last chunk,
# edge tiles in both directions
GEMM_SHAPES = [(8, 8, 8), (16, 16, 1), (5, 17, 3), (17, 5, 31), (32, 32, 32), (33, 40, 33), (70, 31, 130),
               (64, 64, 65), (100, 96, 16), (130, 65, 7), (129, 200, 33)]


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("mnk", GEMM_SHAPES)
def test_gemm(h, p, ta, tb, mnk):
    dt, eps = PREC[p]
    m, n, k = mnk
    batch = 13
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    sa = (m, k) if ta == "N" else (k, m)
    sb = (k, n) if tb == "N" else (n, k)
    As = [rnd(rng, sa, dt) for _ in range(batch)]
    Bs = [rnd(rng, sb, dt) for _ in range(batch)]
    Cs = [rnd(rng, (m, n), dt) for _ in range(batch)]
    beta = 0.0 if k % 2 else 0.5
    if beta == 0.0:
        Cs = [np.full((m, n), np.nan) for _ in range(batch)]
    lda, ldb, ldc = sa[0] + 1, sb[0] + 3, m + 2
    dA, sA = stage(As, lda, dt)
    dB, sB = stage(Bs, ldb, dt)
    dC, sC = stage(Cs, ldc, dt)
    alpha = -1.25
    if m % 2:
        pA, pB, pC = kt.ptr_array(dA, batch, sA), kt.ptr_array(dB, batch, sB), kt.ptr_array(dC, batch, sC)
        st = getattr(h.lib, "kblas%sgemm_batch" % p)(h.h, ta.encode(), tb.encode(), m, n, k, alpha, kt.P(pA), lda,
                                                     kt.P(pB), ldb, beta, kt.P(pC), ldc, batch)
    else:
        st = getattr(h.lib, "kblas%sgemm_batch_strided" % p)(h.h, ta.encode(), tb.encode(), m, n, k, alpha, kt.P(dA),
                                                             lda, sA, kt.P(dB), ldb, sB, beta, kt.P(dC), ldc, sC,
                                                             batch)
    assert st == pykblas.KBLAS_Success
    R = kt.to_host(dC)
    for b in range(batch):
        X = R[b].T.astype(np.float64)
        Co = po.gemm(ta, tb, alpha, As[b], Bs[b], beta, np.nan_to_num(Cs[b]))
        err = np.max(np.abs(X[:m] - Co)) / max(np.max(np.abs(Co)), 1e-300)
        assert err <= 16 * max(k, 1) * eps, (b, err)
        assert np.all(np.isnan(X[m:])), b


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("mk", [(8, 8), (16, 3), (17, 33), (32, 32), (45, 64), (96, 130)])
def test_syrk(h, p, trans, upper, mk):
    dt, eps = PREC[p]
    m, k = mk
    batch = 11
    rng = np.random.default_rng(m * 5 + k + 100 * upper)
    sa = (m, k) if trans == "N" else (k, m)
    As = [rnd(rng, sa, dt) for _ in range(batch)]
    Cs = [rnd(rng, (m, m), dt) for _ in range(batch)]
    other = np.tril_indices(m, -1) if upper else np.triu_indices(m, 1)
    for C in Cs:
        C[other] = np.nan             # the other triangle is neither read nor written
    lda, ldc = sa[0] + 2, m + 1
    dA, sA = stage(As, lda, dt)
    dC, sC = stage(Cs, ldc, dt)
    alpha, beta = 0.5, -1.0
    st = getattr(h.lib, "kblas%ssyrk_batch_strided" % p)(h.h, b"U" if upper else b"L", trans.encode(), m, k, alpha,
                                                         kt.P(dA), lda, sA, beta, kt.P(dC), ldc, sC, batch)
    assert st == pykblas.KBLAS_Success
    R = kt.to_host(dC)
    tri = np.triu_indices(m) if upper else np.tril_indices(m)
    for b in range(batch):
        X = R[b].T.astype(np.float64)[:m]
        C0 = np.nan_to_num(Cs[b])
        if upper:
            Co = po.zsyrk(True, trans, alpha, As[b].astype(np.complex128), beta, C0.astype(np.complex128)).real
        else:
            Co = po.syrk(trans, alpha, As[b], beta, C0)
        err = np.max(np.abs(X[tri] - Co[tri])) / max(np.max(np.abs(Co[tri])), 1e-300)
        assert err <= 16 * k * eps, (b, err)
        assert np.all(np.isnan(X[other])), b


def test_gemm_alpha_zero_reads_no_operands(h):
    # alpha = 0: C := beta C, A and B not read (NaN operands leave C finite)
    m, n, k, batch = 20, 9, 7, 3
    A = np.full((batch, k, m), np.nan)
    B = np.full((batch, n, k), np.nan)
    C = np.random.default_rng(1).standard_normal((batch, n, m))
    dA, dB, dC = kt.to_dev(A), kt.to_dev(B), kt.to_dev(C)
    st = h.lib.kblasDgemm_batch_strided(h.h, b"N", b"N", m, n, k, 0.0, kt.P(dA), m, m * k, kt.P(dB), k, k * n,
                                        2.0, kt.P(dC), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(dC), 2.0 * C)


def dev_list(mats, dt):
    return [kt.to_dev(np.asfortranarray(M).T.copy() if M.size else np.zeros((1, 1)), dt) for M in mats]


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("maxs", [(20, 14, 9), (150, 90, 70)])
def test_nonuniform_gemm_with_bounds(h, p, ta, tb, maxs):
    # kblas{D,S}gemm_vbatch with max_m / max_n / max_k: the matrix-core kernels sized by the maxima (the
    # workgroup tiles when both reach 64), each matrix at its own sizes, one with m = 0 skipped
    dt, eps = PREC[p]
    torch = kt.torch_dev()
    rng = np.random.default_rng(sum(maxs) + 2 * (ta == "T") + (tb == "T"))
    batch = 11
    ms = rng.integers(1, maxs[0] + 1, batch).astype(np.int32)
    ns = rng.integers(1, maxs[1] + 1, batch).astype(np.int32)
    ks = rng.integers(1, maxs[2] + 1, batch).astype(np.int32)
    ms[0], ns[0], ks[0] = maxs
    ms[4] = 0
    As = [rnd(rng, (m, k) if ta == "N" else (k, m), dt) for m, k in zip(ms, ks)]
    Bs = [rnd(rng, (k, n) if tb == "N" else (n, k), dt) for n, k in zip(ns, ks)]
    Cs = [rnd(rng, (m, n), dt) for m, n in zip(ms, ns)]
    dA, dB, dC = dev_list(As, dt), dev_list(Bs, dt), dev_list(Cs, dt)
    lda = kt.to_dev(np.array([max(A.shape[0], 1) for A in As], np.int32))
    ldb = kt.to_dev(np.array([max(B.shape[0], 1) for B in Bs], np.int32))
    ldc = kt.to_dev(np.maximum(ms, 1).astype(np.int32))
    P = lambda L: torch.tensor([t.data_ptr() for t in L], dtype=torch.int64, device="cuda")
    pA, pB, pC = P(dA), P(dB), P(dC)                            # held until the call has run
    dm, dn, dk = kt.to_dev(ms), kt.to_dev(ns), kt.to_dev(ks)
    st = getattr(h.lib, "kblas%sgemm_vbatch" % p)(h.h, ta.encode(), tb.encode(), kt.P(dm), kt.P(dn), kt.P(dk),
                                                  maxs[0], maxs[1], maxs[2], -0.5, kt.P(pA), kt.P(lda), kt.P(pB),
                                                  kt.P(ldb), 2.0, kt.P(pC), kt.P(ldc), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        if ms[i] == 0:
            continue
        X = kt.to_host(dC[i]).T.astype(np.float64)
        Co = po.gemm(ta, tb, -0.5, As[i], Bs[i], 2.0, Cs[i])
        err = np.max(np.abs(X - Co)) / max(np.max(np.abs(Co)), 1e-300)
        assert err <= 16 * int(ks[i]) * eps, (i, err)


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("maxs", [(24, 10), (130, 70)])
def test_nonuniform_syrk_with_bounds(h, p, trans, maxs):
    dt, eps = PREC[p]
    torch = kt.torch_dev()
    rng = np.random.default_rng(sum(maxs) + (trans == "T"))
    batch = 10
    ms = rng.integers(1, maxs[0] + 1, batch).astype(np.int32)
    ks = rng.integers(1, maxs[1] + 1, batch).astype(np.int32)
    ms[0], ks[0] = maxs
    As = [rnd(rng, (m, k) if trans == "N" else (k, m), dt) for m, k in zip(ms, ks)]
    Cs = [rnd(rng, (m, m), dt) for m in ms]
    for C in Cs:
        C[np.triu_indices(C.shape[0], 1)] = np.nan
    dA, dC = dev_list(As, dt), dev_list(Cs, dt)
    lda = kt.to_dev(np.array([A.shape[0] for A in As], np.int32))
    ldc = kt.to_dev(ms.copy())
    P = lambda L: torch.tensor([t.data_ptr() for t in L], dtype=torch.int64, device="cuda")
    pA, pC = P(dA), P(dC)                                       # held until the call has run
    dm, dk = kt.to_dev(ms), kt.to_dev(ks)
    st = getattr(h.lib, "kblas%ssyrk_vbatch" % p)(h.h, b"L", trans.encode(), kt.P(dm), kt.P(dk), maxs[0], maxs[1],
                                                  0.75, kt.P(pA), kt.P(lda), -1.0, kt.P(pC), kt.P(ldc), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        m = int(ms[i])
        X = kt.to_host(dC[i]).T.astype(np.float64)
        Co = po.syrk(trans, 0.75, As[i], -1.0, np.nan_to_num(Cs[i]))
        tri = np.tril_indices(m)
        err = np.max(np.abs(X[tri] - Co[tri])) / max(np.max(np.abs(Co[tri])), 1e-300)
        assert err <= 16 * int(ks[i]) * eps, (i, err)
        assert np.all(np.isnan(X[np.triu_indices(m, 1)])), i


def test_gemm_workgroup_tiles_alpha_zero_reads_no_operands(h):
    # the 64 x 64 workgroup-tile kernel: alpha = 0 leaves C := beta C without reading A or B
    m, n, k, batch = 70, 66, 20, 3
    A = np.full((batch, k, m), np.nan)
    B = np.full((batch, n, k), np.nan)
    C = np.random.default_rng(2).standard_normal((batch, n, m))
    dA, dB, dC = kt.to_dev(A), kt.to_dev(B), kt.to_dev(C)
    st = h.lib.kblasDgemm_batch_strided(h.h, b"N", b"N", m, n, k, 0.0, kt.P(dA), m, m * k, kt.P(dB), k, k * n,
                                        -1.5, kt.P(dC), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(dC), -1.5 * C)
