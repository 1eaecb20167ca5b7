"""GPU tests of the mixed-precision POSV (kblasDsposv_batch*): FP32 factor +
FP64 refinement, checked against the FP64 oracle.

Bar: normwise backward error ||b - A x||inf / (||A||inf ||x||inf
+ ||b||inf) <= 1e-13 for every solve; forward error within kappa * 1e-13 of
the FP64 oracle's solution; iteration counts within 2 of the oracle's dsposv
restatement; fallback (iter < 0) exactly when the FP32 factor cannot serve.
"""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def run_sposv(h, side, A, B, nrhs_rows, order):
    batch = A.shape[0]
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    dX = kt.to_dev(np.zeros_like(B))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    it = kt.to_dev(np.full(batch, -99, np.int32))
    rows = B.shape[2]
    cols = B.shape[1]
    st = h.lib.kblasDsposv_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dA), order, order * order,
                                         kt.P(dB), rows, rows * cols, kt.P(dX), rows, rows * cols,
                                         batch, kt.P(info), kt.P(it))
    assert st == pykblas.KBLAS_Success
    return kt.to_host(dX), kt.to_host(info), kt.to_host(it), kt.to_host(dA)


@pytest.mark.parametrize("n", [16, 32, 40, 57, 64])
@pytest.mark.parametrize("side", ["R", "L"])
def test_sposv_conditioned_batch(h, n, side):
    batch = 40
    A, kap = po.gen_spd_cond_batch(4, n, batch, 2.0, 6.0)
    B = po.gen_rhs_batch(44, 4, n, batch) if side == "R" else po.gen_rhs_batch(44, n, 4, batch)
    X, info, it, Aout = run_sposv(h, side, A, B, 4, n)
    assert np.all(info == 0)
    for i in range(batch):
        Ai, Bi, Xi = po.col(A, i), po.col(B, i), po.col(X, i)
        assert po.solve_berr(side, Ai, Xi, Bi) <= 1e-13, (i, kap[i], it[i])
        _, Xo, _ = po.posv(side, Ai, Bi)
        fe = np.max(np.abs(Xi - Xo)) / np.max(np.abs(Xo))
        assert fe <= kap[i] * 1e-13
        Xr, inf_r, it_r = po.sposv(side, Ai, Bi)
        # cond <= 1e6: refinement must serve on both sides (a spurious FP64
        # fallback would still give a fine answer, so it is checked here)
        assert it[i] >= 0 and it_r >= 0, (i, kap[i], it[i], it_r)
        assert abs(int(it[i]) - it_r) <= 2, (i, kap[i], it[i], it_r)
        assert np.array_equal(Aout[i], A[i])     # A untouched when refinement served


@pytest.mark.parametrize("n", [32, 48])
def test_sposv_fallback_on_fp32_failure(h, n):
    batch = 6
    A, kap = po.gen_spd_cond_batch(9, n, batch, 12.0, 12.0)   # kappa 1e12: FP32 factor fails
    easy, _ = po.gen_spd_cond_batch(9, n, 1, 2.0, 2.0, first=1)
    A[1] = easy[0]                                             # one easy matrix
    B = po.gen_rhs_batch(9, n, 2, batch)
    X, info, it, Aout = run_sposv(h, "L", A, B, 2, n)
    assert np.all(info == 0)
    assert it[1] >= 0
    for i in range(batch):
        assert po.solve_berr("L", po.col(A, i), po.col(X, i), po.col(B, i)) <= 1e-13
        if i != 1:
            assert it[i] < 0
            L = po.col(Aout, i)
            assert po.potrf_berr(po.col(A, i), L) <= 4 * n * kt.EPS64   # A holds L after fallback


@pytest.mark.parametrize("n", [65, 130])
@pytest.mark.parametrize("side", ["R", "L"])
def test_sposv_large_order_full_precision(h, n, side):
    # above the register range the mixed entry takes dsposv's full-precision
    # path: ITER = -1, A := L, X := the FP64 solution
    batch = 5
    A, kap = po.gen_spd_cond_batch(12, n, batch, 2.0, 5.0)
    B = po.gen_rhs_batch(13, 3, n, batch) if side == "R" else po.gen_rhs_batch(13, n, 3, batch)
    X, info, it, Aout = run_sposv(h, side, A, B, 3, n)
    assert np.all(info == 0)
    assert np.all(it == -1)
    for i in range(batch):
        assert po.solve_berr(side, po.col(A, i), po.col(X, i), po.col(B, i)) <= 1e-13
        assert po.potrf_berr(po.col(A, i), po.col(Aout, i)) <= 4 * n * kt.EPS64


@pytest.mark.parametrize("n", [24, 50, 64])
def test_sposv_fp32_overflow_iter_minus_2(h, n):
    """ITER = -2 (LAPACK's DLAG2S / DLAT2S overflow checks): a column of B
    beyond FP32 range in matrix 0, A scaled beyond FP32 range in matrix 2;
    both are solved in FP64 (X equal to the FP64 oracle's within the solve
    bar), the other matrices refine as usual and A stays untouched for them."""
    batch = 4
    A, kap = po.gen_spd_cond_batch(61, n, batch, 2.0, 3.0)
    B = po.gen_rhs_batch(62, n, 3, batch)                 # side L: n x 3
    B[0][1, :] *= 1e39                                    # right-hand side 1 of matrix 0
    A[2] *= 1e39 / np.max(np.abs(A[2]))                   # |A| up to 1e39 > FLT_MAX
    X, info, it, Aout = run_sposv(h, "L", A, B, 3, n)
    assert np.all(info == 0)
    assert it[0] == -2 and it[2] == -2, it
    assert it[1] >= 0 and it[3] >= 0, it
    for i in range(batch):
        Ai, Bi, Xi = po.col(A, i), po.col(B, i), po.col(X, i)
        _, _, it_r = po.sposv("L", Ai, Bi)
        if i in (0, 2):
            assert it_r == -2
        assert po.solve_berr("L", Ai, Xi, Bi) <= 1e-13, (i, it[i])
        if i in (0, 2):
            _, Xo, _ = po.posv("L", Ai, Bi)
            assert np.max(np.abs(Xi - Xo)) <= 64 * n * kt.EPS64 * np.max(np.abs(Xo)) * kap[i], i
            assert po.potrf_berr(Ai, po.col(Aout, i)) <= 4 * n * kt.EPS64    # A holds L
        else:
            assert np.array_equal(Aout[i], A[i])
