"""GPU tests of the randomized-range-finder SPD helpers (include/batch_ara.h)
through the C ABI against the oracle (orc_ara_potrf / orc_ara_syrk /
orc_ara_trsm): non-uniform batches through device pointer and size arrays,
the rank-revealing stop, FP32 and FP64 outputs, padding left untouched."""
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


def dev_list(mats, dtype):
    """column-major matrices -> device buffers (kept alive) + device pointer array"""
    torch = kt.torch_dev()
    bufs = [kt.to_dev(np.asfortranarray(M).T.copy(), dtype) for M in mats]
    ptrs = torch.tensor([b.data_ptr() for b in bufs], dtype=torch.int64, device="cuda")
    return bufs, ptrs


def back(buf):
    return kt.to_host(buf).T


def gram_batch(seed, dims, ld_pad=3):
    rng = np.random.default_rng(seed)
    Gs = []
    for i, d in enumerate(dims):
        n = max(int(d), 1)
        Y = rng.uniform(-1.0, 1.0, (n + 20, n)).astype(np.float32).astype(np.float64)
        G = np.full((n + ld_pad, n), np.nan)
        G[:n, :n] = Y.T @ Y
        Gs.append(G)
    # rank-deficient members: an exact rank-one integer Gram matrix and a planted zero pivot
    v = np.array([2.0, 4.0, 6.0, -8.0, 10.0, 1.0, 3.0])
    Gs[1] = np.full((7 + ld_pad, 7), np.nan)
    Gs[1][:7, :7] = np.outer(v, v)
    Gs[2] = np.full((9 + ld_pad, 9), np.nan)
    Gs[2][:9, :9] = np.diag([4.0, 9.0, 1.0, 0.0, 2.0, 3.0, 5.0, 6.0, 7.0])
    return Gs


@pytest.mark.parametrize("bs", [16, 32])
@pytest.mark.parametrize("rprec", ["d", "s"])
def test_fused_potrf_matches_oracle(h, bs, rprec):
    dims = np.array([32, 7, 9, 1, 5, 16, 17, 31, 0, 40, 12, 24, 32, 3, 29, 20, 8, 30, 2], np.int32)
    dims[1], dims[2] = 7, 9
    Gs = gram_batch(5 + bs, dims)
    nb = len(dims)
    torch = kt.torch_dev()
    dG, pG = dev_list(Gs, np.float64)
    ldg = kt.to_dev(np.array([G.shape[0] for G in Gs], np.int32))
    rdt = np.float64 if rprec == "d" else np.float32
    Rs = [np.full((min(max(int(d), 1), bs) + 2, min(max(int(d), 1), bs)), np.nan) for d in dims]
    dR, pR = dev_list(Rs, rdt)
    ldr = kt.to_dev(np.array([R.shape[0] for R in Rs], np.int32))
    d0 = np.random.default_rng(9).uniform(0.5, 2.0, nb * bs)
    dd = kt.to_dev(d0)
    ranks = kt.to_dev(np.full(nb, -7, np.int32))
    fn = h.lib.kblas_dara_fused_potrf_batch if rprec == "d" else h.lib.kblas_sara_fused_potrf_batch
    st = fn(kt.P(kt.to_dev(dims)), kt.P(pG), kt.P(ldg), kt.P(pR), kt.P(ldr), kt.P(dd), bs, kt.P(ranks), nb,
            kt.P(torch.cuda.current_stream().cuda_stream))
    assert st == pykblas.KBLAS_Success
    rk, dh = kt.to_host(ranks), kt.to_host(dd)
    for i, dim in enumerate(dims):
        Ro, do, ro = po.ara_potrf(Gs[i][:max(int(dim), 1)], bs, d0[i * bs:(i + 1) * bs], dim=int(dim))
        assert rk[i] == ro, (i, dim, rk[i], ro)
        R = back(dR[i])
        c = min(int(dim), bs)
        if c <= 0:
            assert np.all(np.isnan(R))                                   # nothing written
            assert np.array_equal(dh[i * bs:(i + 1) * bs], d0[i * bs:(i + 1) * bs])
            continue
        tol = 4 * c * EPS * max(1.0, np.max(np.abs(Ro))) if rprec == "d" else 0.0
        Rg = R[:c, :c]
        if rprec == "d":
            assert np.max(np.abs(Rg - Ro)) <= tol, (i, dim)
        else:
            assert np.max(np.abs(Rg - Ro.astype(np.float32)) / np.maximum(np.abs(Ro), 1e-30)) <= 2 * kt.EPS32 + 8 * c * EPS, (i, dim)
        assert np.all(np.isnan(R[c:, :]))                                # ldr padding untouched
        assert np.allclose(dh[i * bs:(i + 1) * bs], do, rtol=8 * c * EPS, atol=0), (i, dim)
    assert rk[1] == 1 and rk[2] == 3                                     # the planted rank deficiencies


def test_fused_potrf_rejects_bad_bs(h):
    torch = kt.torch_dev()
    z = kt.P(0)
    st = h.lib.kblas_dara_fused_potrf_batch(z, z, z, z, z, z, 24, z, 4, kt.P(torch.cuda.current_stream().cuda_stream))
    assert st == pykblas.KBLAS_NotImplemented


@pytest.mark.parametrize("integer", [False, True])
def test_sara_mp_syrk_matches_oracle(h, integer):
    rng = np.random.default_rng(11 + integer)
    shapes = [(100, 32), (15, 7), (16, 16), (17, 17), (257, 29), (1, 1), (0, 5), (64, 33), (40, 20), (300, 1)]
    As = []
    for m, n in shapes:
        mm = max(m, 1)
        A = rng.integers(-7, 8, (mm + 5, n)) if integer else rng.uniform(-1.0, 1.0, (mm + 5, n))
        As.append(A.astype(np.float32).astype(np.float64))
    dA, pA = dev_list(As, np.float32)
    lda = kt.to_dev(np.array([A.shape[0] for A in As], np.int32))
    Bs = [np.full((min(n, 32) + 4, min(n, 32)), np.nan) for _, n in shapes]
    dB, pB = dev_list(Bs, np.float64)
    ldb = kt.to_dev(np.array([B.shape[0] for B in Bs], np.int32))
    ms = kt.to_dev(np.array([m for m, _ in shapes], np.int32))
    ns = kt.to_dev(np.array([n for _, n in shapes], np.int32))
    st = h.lib.kblas_sara_mp_syrk_batch(h.h, kt.P(ms), kt.P(ns), 300, 32, kt.P(pA), kt.P(lda), kt.P(pB), kt.P(ldb),
                                        len(shapes))
    assert st == pykblas.KBLAS_Success
    for i, (m, n) in enumerate(shapes):
        B = back(dB[i])
        if m == 0:
            assert np.all(np.isnan(B))
            continue
        c = min(n, 32)
        Bo = po.ara_syrk(As[i][:m, :c])
        Bg = B[:c, :c]
        assert np.array_equal(Bg, Bg.T), i                              # exactly symmetric
        if integer:
            assert np.array_equal(Bg, Bo), i
        else:
            bound = 2 * m * EPS * (np.abs(As[i][:m, :c]).T @ np.abs(As[i][:m, :c]))
            assert np.all(np.abs(Bg - Bo) <= bound), i
        assert np.all(np.isnan(B[c:, :])), i                             # ldb padding untouched
    st = h.lib.kblas_sara_mp_syrk_batch(h.h, kt.P(ms), kt.P(ns), 300, 33, kt.P(pA), kt.P(lda), kt.P(pB), kt.P(ldb),
                                        len(shapes))
    assert st == pykblas.KBLAS_NotImplemented


def test_dara_mp_syrk_matches_oracle(h):
    rng = np.random.default_rng(13)
    shapes = [(100, 32), (15, 7), (257, 40), (1, 1), (64, 64)]
    As = [rng.standard_normal((m + 3, n)) for m, n in shapes]
    dA, pA = dev_list(As, np.float64)
    lda = kt.to_dev(np.array([A.shape[0] for A in As], np.int32))
    Bs = [np.full((n + 2, n), np.nan) for _, n in shapes]
    dB, pB = dev_list(Bs, np.float64)
    ldb = kt.to_dev(np.array([B.shape[0] for B in Bs], np.int32))
    ms = kt.to_dev(np.array([m for m, _ in shapes], np.int32))
    ns = kt.to_dev(np.array([n for _, n in shapes], np.int32))
    st = h.lib.kblas_dara_mp_syrk_batch(h.h, kt.P(ms), kt.P(ns), 257, 64, kt.P(pA), kt.P(lda), kt.P(pB), kt.P(ldb),
                                        len(shapes))
    assert st == pykblas.KBLAS_Success
    for i, (m, n) in enumerate(shapes):
        B = back(dB[i])
        Bo = po.ara_syrk(As[i][:m, :])
        bound = 2 * m * EPS * (np.abs(As[i][:m]).T @ np.abs(As[i][:m]))
        assert np.all(np.abs(B[:n, :n] - Bo) <= bound), i
        assert np.all(np.isnan(B[n:, :])), i


@pytest.mark.parametrize("bs", [16, 32])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_ara_trsm_matches_oracle(h, bs, prec):
    rng = np.random.default_rng(17 + bs)
    shapes = [(100, 32), (1, 1), (300, 17), (257, 16), (40, 7), (600, 29), (5, 40), (0, 3), (64, 24)]
    dt = np.float64 if prec == "d" else np.float32
    Rs, Bs = [], []
    for rows, cols in shapes:
        Y = rng.uniform(-1.0, 1.0, (cols + 24, cols))
        R = np.linalg.cholesky(Y.T @ Y).T.astype(dt).astype(np.float64)
        R[np.tril_indices(cols, -1)] = np.nan                            # only the upper triangle may be read
        Rp = np.full((cols + 2, cols), np.nan)
        Rp[:cols] = R
        Rs.append(Rp)
        Bs.append(rng.standard_normal((max(rows, 1) + 3, cols)).astype(dt).astype(np.float64))
    dR, pR = dev_list(Rs, dt)
    dB, pB = dev_list(Bs, dt)
    ldr = kt.to_dev(np.array([R.shape[0] for R in Rs], np.int32))
    ldb = kt.to_dev(np.array([B.shape[0] for B in Bs], np.int32))
    rows = kt.to_dev(np.array([r for r, _ in shapes], np.int32))
    cols = kt.to_dev(np.array([c for _, c in shapes], np.int32))
    fn = h.lib.kblas_dara_trsm_batch if prec == "d" else h.lib.kblas_sara_trsm_batch
    st = fn(h.h, kt.P(pB), kt.P(ldb), kt.P(pR), kt.P(ldr), kt.P(rows), kt.P(cols), len(shapes), 600, bs)
    assert st == pykblas.KBLAS_Success
    eps = EPS if prec == "d" else kt.EPS32
    for i, (r, c) in enumerate(shapes):
        X = back(dB[i])
        k = min(c, bs)
        Xo = po.ara_trsm(Bs[i][:r], Rs[i][:c], bs)
        Rk = np.nan_to_num(Rs[i][:k, :k])
        scale = np.abs(Xo[:, :k]) @ np.abs(Rk) + np.abs(Bs[i][:r, :k]) if r else np.zeros((0, k))
        assert np.all(np.abs(X[:r, :k] @ Rk - Bs[i][:r, :k]) <= 4 * k * eps * scale + 1e-300), (i, r, c)
        if prec == "d":
            cond = np.linalg.cond(Rk)
            assert np.max(np.abs(X[:r, :k] - Xo[:, :k]), initial=0.0) <= 8 * k * EPS * cond * max(1.0, np.max(np.abs(Xo), initial=0.0)), i
        assert np.array_equal(X[:r, k:], Bs[i][:r, k:]), i               # columns beyond bs untouched
        assert np.array_equal(X[r:, :], Bs[i][r:, :]), i                 # rows beyond rows untouched
