"""Seeded randomized parity of the batched TRSM / GEMM / SYRK entry points
(uniform strided, pointer arrays, non-uniform with and without size bounds)
against the oracle: 40 draws of precision, operation, shape (1..200 in each
dimension, so every kernel bucket is hit), transpositions, side, alpha / beta
and leading-dimension padding.  Bar: 64 * k * eps of the working precision
(TRSM, k = the order; GEMM / SYRK, k = the inner dimension)."""
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


def rnd(rng, shape, dt):
    return rng.standard_normal(shape).astype(dt).astype(np.float64)


def tri(rng, k, dt):
    T = np.tril(rng.standard_normal((k, k)) / k) + np.diag(1.0 + rng.uniform(0.0, 1.0, k))
    return T.astype(dt).astype(np.float64)


def dev_mats(mats, lds, dt):
    """one device buffer per matrix (column-major, leading dimension lds[i], NaN padding)"""
    out = []
    for M, ld in zip(mats, lds):
        r, c = M.shape
        buf = np.full((max(c, 1), ld), np.nan)
        buf[:c, :r] = M.T
        out.append(kt.to_dev(buf, dt))
    return out


def ptrs(torch, L):
    return torch.tensor([t.data_ptr() for t in L], dtype=torch.int64, device="cuda")


def host(t, r, c):
    return kt.to_host(t).astype(np.float64)[:c, :r].T


@pytest.mark.parametrize("draw", range(40))
def test_random_blas3(h, draw):
    rng = np.random.default_rng(9000 + draw)
    torch = kt.torch_dev()
    p = "D" if draw % 3 else "S"
    dt, eps = PREC[p]
    op = ["trsm", "gemm", "syrk"][draw % 3]
    mode = ["ptr", "vb", "vb0"][(draw // 3) % 3]       # pointer arrays; non-uniform bounded / unbounded
    batch = int(rng.integers(1, 9))
    big = rng.random() < 0.5
    hi = 200 if big else 40
    alpha = float(rng.choice([1.0, -0.5, 2.0]))
    if op == "trsm":
        side = "L" if rng.random() < 0.5 else "R"
        trans = "N" if rng.random() < 0.5 else "T"
        if mode == "ptr":
            ms = np.full(batch, int(rng.integers(1, hi + 1)), np.int32)
            ns = np.full(batch, int(rng.integers(1, 80)), np.int32)
        else:
            ms = rng.integers(1, hi + 1, batch).astype(np.int32)
            ns = rng.integers(1, 80, batch).astype(np.int32)
        if side == "R":
            ms, ns = ns, ms
        ks = ms if side == "L" else ns
        Ts = [tri(rng, int(k), dt) for k in ks]
        Bs = [rnd(rng, (int(m), int(n)), dt) for m, n in zip(ms, ns)]
        lda = [int(k) + int(rng.integers(0, 3)) for k in ks]
        ldb = [int(m) + int(rng.integers(0, 3)) for m in ms]
        dT, dB = dev_mats(Ts, lda, dt), dev_mats(Bs, ldb, dt)
        pT, pB = ptrs(torch, dT), ptrs(torch, dB)
        if mode == "ptr":
            st = getattr(h.lib, "kblas%strsm_batch" % p)(h.h, side.encode(), b"L", trans.encode(), b"N",
                                                         int(ms[0]), int(ns[0]), alpha, kt.P(pT), lda[0], kt.P(pB),
                                                         ldb[0], batch) if len(set(lda)) == 1 and len(set(ldb)) == 1 \
                else None
            if st is None:   # uniform pointer arrays need one leading dimension
                lda = [max(lda)] * batch
                ldb = [max(ldb)] * batch
                dT, dB = dev_mats(Ts, lda, dt), dev_mats(Bs, ldb, dt)
                pT, pB = ptrs(torch, dT), ptrs(torch, dB)
                st = getattr(h.lib, "kblas%strsm_batch" % p)(h.h, side.encode(), b"L", trans.encode(), b"N",
                                                             int(ms[0]), int(ns[0]), alpha, kt.P(pT), lda[0],
                                                             kt.P(pB), ldb[0], batch)
        else:
            dm, dn = kt.to_dev(ms), kt.to_dev(ns)
            dla, dlb = kt.to_dev(np.array(lda, np.int32)), kt.to_dev(np.array(ldb, np.int32))
            mx = (int(ms.max()), int(ns.max())) if mode == "vb" else (0, 0)
            st = getattr(h.lib, "kblas%strsm_vbatch" % p)(h.h, side.encode(), b"L", trans.encode(), b"N", kt.P(dm),
                                                          kt.P(dn), mx[0], mx[1], alpha, kt.P(pT), kt.P(dla),
                                                          kt.P(pB), kt.P(dlb), batch)
        assert st == pykblas.KBLAS_Success
        for i in range(batch):
            X = host(dB[i], int(ms[i]), int(ns[i]))
            Xo = po.trsm(side, trans, alpha, Ts[i], Bs[i])
            err = np.max(np.abs(X - Xo)) / max(np.max(np.abs(Xo)), 1e-300)
            assert err <= 64 * int(ks[i]) * eps, (i, err)
    else:
        ta = "N" if rng.random() < 0.5 else "T"
        tb = "N" if rng.random() < 0.5 else "T"
        beta = float(rng.choice([0.0, 1.0, -0.75]))
        if mode == "ptr":
            ms = np.full(batch, int(rng.integers(1, hi + 1)), np.int32)
            ns = np.full(batch, int(rng.integers(1, hi + 1)), np.int32)
            ks = np.full(batch, int(rng.integers(1, hi + 1)), np.int32)
        else:
            ms = rng.integers(1, hi + 1, batch).astype(np.int32)
            ns = rng.integers(1, hi + 1, batch).astype(np.int32)
            ks = rng.integers(1, hi + 1, batch).astype(np.int32)
        if op == "syrk":
            ns = ms
        As = [rnd(rng, (int(m), int(k)) if ta == "N" else (int(k), int(m)), dt) for m, k in zip(ms, ks)]
        Bs = [rnd(rng, (int(k), int(n)) if tb == "N" else (int(n), int(k)), dt) for n, k in zip(ns, ks)]
        Cs = [rnd(rng, (int(m), int(n)), dt) for m, n in zip(ms, ns)]
        if op == "syrk":
            for C in Cs:
                C[np.triu_indices(C.shape[0], 1)] = np.nan
        pad = int(rng.integers(0, 3))
        lda = [A.shape[0] + pad for A in As]
        ldb = [B.shape[0] + pad for B in Bs]
        ldc = [int(m) + pad for m in ms]
        if mode == "ptr":
            lda, ldb, ldc = [max(lda)] * batch, [max(ldb)] * batch, [max(ldc)] * batch
        dA, dB, dC = dev_mats(As, lda, dt), dev_mats(Bs, ldb, dt), dev_mats(Cs, ldc, dt)
        pA, pB, pC = ptrs(torch, dA), ptrs(torch, dB), ptrs(torch, dC)
        if op == "gemm":
            if mode == "ptr":
                st = getattr(h.lib, "kblas%sgemm_batch" % p)(h.h, ta.encode(), tb.encode(), int(ms[0]), int(ns[0]),
                                                             int(ks[0]), alpha, kt.P(pA), lda[0], kt.P(pB), ldb[0],
                                                             beta, kt.P(pC), ldc[0], batch)
            else:
                dm, dn, dk = kt.to_dev(ms), kt.to_dev(ns), kt.to_dev(ks)
                dla, dlb, dlc = (kt.to_dev(np.array(x, np.int32)) for x in (lda, ldb, ldc))
                mx = (int(ms.max()), int(ns.max()), int(ks.max())) if mode == "vb" else (0, 0, 0)
                st = getattr(h.lib, "kblas%sgemm_vbatch" % p)(h.h, ta.encode(), tb.encode(), kt.P(dm), kt.P(dn),
                                                              kt.P(dk), mx[0], mx[1], mx[2], alpha, kt.P(pA),
                                                              kt.P(dla), kt.P(pB), kt.P(dlb), beta, kt.P(pC),
                                                              kt.P(dlc), batch)
        else:
            if mode == "ptr":
                st = getattr(h.lib, "kblas%ssyrk_batch" % p)(h.h, b"L", ta.encode(), int(ms[0]), int(ks[0]), alpha,
                                                             kt.P(pA), lda[0], beta, kt.P(pC), ldc[0], batch)
            else:
                dm, dk = kt.to_dev(ms), kt.to_dev(ks)
                dla, dlc = kt.to_dev(np.array(lda, np.int32)), kt.to_dev(np.array(ldc, np.int32))
                mx = (int(ms.max()), int(ks.max())) if mode == "vb" else (0, 0)
                st = getattr(h.lib, "kblas%ssyrk_vbatch" % p)(h.h, b"L", ta.encode(), kt.P(dm), kt.P(dk), mx[0],
                                                              mx[1], alpha, kt.P(pA), kt.P(dla), beta, kt.P(pC),
                                                              kt.P(dlc), batch)
        assert st == pykblas.KBLAS_Success
        for i in range(batch):
            X = host(dC[i], int(ms[i]), int(ns[i]))
            C0 = np.nan_to_num(Cs[i])
            if op == "gemm":
                Co = po.gemm(ta, tb, alpha, As[i], Bs[i], beta, C0)
                sel = np.ones_like(X, dtype=bool)
            else:
                Co = po.syrk(ta, alpha, As[i], beta, C0)
                sel = np.tril(np.ones_like(X, dtype=bool))
                assert np.all(np.isnan(X[~sel])), i
            err = np.max(np.abs(X[sel] - Co[sel])) / max(np.max(np.abs(Co[sel])), 1e-300)
            assert err <= 64 * int(ks[i]) * eps, (i, err)
