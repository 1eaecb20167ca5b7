"""GPU parity tests of the batched Cholesky family through the C ABI
(kblas{D,S}potrf/potrs/posv_batch[_strided]) against the CPU oracle.

Bar (FP64): backward error ||A - L L^T||_F/||A||_F <= 4 n eps64 (per matrix),
factor within 64 n eps64 (max-norm, relative) of the oracle's, solves within
the same order of the oracle's with normwise residual <= 4 n eps64; info
bit-exact with LAPACK's; strict upper triangle bit-exact unchanged (it is
filled with NaN so any read of it would poison the factor).
FP32: same checks against the FP64 oracle with eps32.
"""
import ctypes

import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu

EPS = kt.EPS64
REG_NS = [1, 2, 3, 5, 8, 13, 16, 17, 24, 31, 32, 33, 48, 63, 64]


def nan_upper(Ab):
    """Ab: (batch, n, ld) batch array; set the strict upper triangle of each matrix to NaN."""
    B = Ab.copy()
    n = B.shape[1]
    for c in range(n):
        B[:, c, :c] = np.nan
    return B


def lower(M):
    return np.tril(M)


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def check_factor(A, L, n, eps, tag=""):
    Lo, info = po.potrf(A)
    assert info == 0
    scale = np.max(np.abs(np.tril(Lo)))
    err = np.max(np.abs(np.tril(L) - np.tril(Lo))) / scale
    assert err <= 64 * n * eps, (tag, err)
    assert po.potrf_berr(A, L) <= 4 * n * eps, tag


@pytest.mark.parametrize("n", REG_NS)
def test_dpotrf_strided(h, n):
    batch = 37
    A = po.gen_spd_shift_batch(1, n, batch)
    Ain = nan_upper(A)
    dA = kt.to_dev(Ain)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDpotrf_batch_strided(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    L = kt.to_host(dA)
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        Li, Ai = po.col(L, i), po.col(A, i)
        iu = np.triu_indices(n, 1)
        assert np.all(np.isnan(Li[iu]))          # strict upper untouched
        check_factor(Ai, Li, n, EPS, "mat %d" % i)


@pytest.mark.parametrize("n", [7, 16, 29, 64])
def test_dpotrf_pointer_array_padded_lda_bitwise_equal_to_strided(h, n):
    batch = 21
    lda = n + 3
    A = po.gen_spd_shift_batch(2, n, batch)
    Ap = np.full((batch, n, lda), np.nan)
    Ap[:, :, :n] = A
    dAp = kt.to_dev(Ap)
    ptrs = kt.ptr_array(dAp, batch, n * lda)
    st = h.lib.kblasDpotrf_batch(h.h, b"L", n, kt.P(ptrs), lda, batch, None)
    assert st == pykblas.KBLAS_Success
    dAs = kt.to_dev(A)
    st = h.lib.kblasDpotrf_batch_strided(h.h, b"L", n, kt.P(dAs), n, n * n, batch, None)
    assert st == pykblas.KBLAS_Success
    Lp, Ls = kt.to_host(dAp), kt.to_host(dAs)
    for i in range(batch):
        assert np.array_equal(np.tril(Lp[i, :, :n].T), np.tril(Ls[i].T))
        assert np.all(np.isnan(Lp[i, :, n:]))    # padding rows never written


@pytest.mark.parametrize("n", [8, 20, 40])
def test_dpotrf_info_matches_lapack(h, n):
    import os
    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_golden.npz"))
    bad = np.asfortranarray(G["nonspd_A_%d" % n])
    batch = 9
    A = po.gen_spd_shift_batch(3, n, batch)
    A[4] = bad.T          # matrix 4 is not positive definite
    dA = kt.to_dev(A)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDpotrf_batch_strided(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    expect = np.zeros(batch, np.int32)
    expect[4] = int(G["nonspd_info_%d" % n])
    assert np.array_equal(inf, expect)
    L = kt.to_host(dA)
    p = expect[4]
    Lo, _ = po.potrf(bad)
    # LAPACK semantics: columns before the failing one hold the factor
    assert np.allclose(np.tril(po.col(L, 4))[:, :p - 1], np.tril(Lo)[:, :p - 1], rtol=1e-12, atol=1e-12)
    for i in (0, 8):
        check_factor(po.col(A, i), po.col(L, i), n, EPS)


@pytest.mark.parametrize("n", REG_NS)
@pytest.mark.parametrize("m", [1, 4, 5])
def test_dposv_right_strided(h, n, m):
    batch = 23
    A = po.gen_spd_shift_batch(4, n, batch)
    B = po.gen_rhs_batch(5, m, n, batch)         # m x n, ldb = m (reference side 'R')
    dA, dB = kt.to_dev(nan_upper(A)), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_batch_strided(h.h, b"R", b"L", m, n, kt.P(dA), n, n * n, kt.P(dB), m, m * n,
                                        batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    X, L = kt.to_host(dB), kt.to_host(dA)
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        Ai, Bi = po.col(A, i), po.col(B, i)
        _, Xo, _ = po.posv("R", Ai, Bi)
        Xi = po.col(X, i)
        assert np.max(np.abs(Xi - Xo)) <= 64 * n * EPS * max(np.max(np.abs(Xo)), 1)
        assert po.solve_berr("R", Ai, Xi, Bi) <= 4 * n * EPS
        check_factor(Ai, po.col(L, i), n, EPS)


@pytest.mark.parametrize("n", [1, 9, 16, 32, 45, 64])
@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_dposv_left_pointer_array(h, n, nrhs):
    batch = 19
    A = po.gen_spd_shift_batch(6, n, batch)
    B = po.gen_rhs_batch(7, n, nrhs, batch)      # n x nrhs (LAPACK layout)
    dA, dB = kt.to_dev(nan_upper(A)), kt.to_dev(B)
    pA, pB = kt.ptr_array(dA, batch, n * n), kt.ptr_array(dB, batch, n * nrhs)
    st = h.lib.kblasDposv_batch(h.h, b"L", b"L", n, nrhs, kt.P(pA), n, kt.P(pB), n, batch, None)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Ai, Bi = po.col(A, i), po.col(B, i)
        _, Xo, _ = po.posv("L", Ai, Bi)
        assert np.max(np.abs(po.col(X, i) - Xo)) <= 64 * n * EPS * max(np.max(np.abs(Xo)), 1)
        assert po.solve_berr("L", Ai, po.col(X, i), Bi) <= 4 * n * EPS


@pytest.mark.parametrize("n", [4, 16, 31, 64])
@pytest.mark.parametrize("side", ["L", "R"])
def test_dpotrs_uses_given_factor(h, n, side):
    batch = 11
    A = po.gen_spd_shift_batch(8, n, batch)
    Lh = np.zeros_like(A)
    for i in range(batch):
        Lh[i] = np.tril(po.potrf(po.col(A, i))[0]).T
    rows, cols = (n, 2) if side == "L" else (3, n)
    B = po.gen_rhs_batch(9, rows, cols, batch)
    dL, dB = kt.to_dev(nan_upper(Lh)), kt.to_dev(B)
    st = h.lib.kblasDpotrs_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dL), n, n * n,
                                         kt.P(dB), rows, rows * cols, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.potrs(side, po.col(Lh, i), po.col(B, i))
        assert np.max(np.abs(po.col(X, i) - Xo)) <= 64 * n * EPS * max(np.max(np.abs(Xo)), 1)


@pytest.mark.parametrize("n", [3, 16, 32, 40, 50, 64])
def test_spotrf_and_sposv_single_precision(h, n):
    batch = 17
    A = po.gen_spd_shift_batch(10, n, batch)
    B = po.gen_rhs_batch(11, 2, n, batch)
    dA = kt.to_dev(nan_upper(A), np.float32)
    dB = kt.to_dev(B, np.float32)
    st = h.lib.kblasSposv_batch_strided(h.h, b"R", b"L", 2, n, kt.P(dA), n, n * n, kt.P(dB), 2, 2 * n, batch, None)
    assert st == pykblas.KBLAS_Success
    L, X = kt.to_host(dA).astype(np.float64), kt.to_host(dB).astype(np.float64)
    for i in range(batch):
        Ai = po.col(A, i)
        assert po.potrf_berr(Ai, po.col(L, i)) <= 4 * n * kt.EPS32
        assert po.solve_berr("R", Ai, po.col(X, i), po.col(B, i)) <= 8 * n * kt.EPS32


def test_argument_errors(h):
    lib = h.lib
    P = ctypes.c_void_p
    dA = kt.to_dev(np.zeros((4, 8, 8)))
    # Upper runs the generic kernels (NotImplemented in the reference); here a zero matrix stops at its first pivot
    assert lib.kblasDpotrf_batch_strided(h.h, b"U", 8, kt.P(dA), 8, 64, 4, None) == pykblas.KBLAS_Success
    assert lib.kblasDpotrf_batch_strided(h.h, b"L", 8, kt.P(dA), 7, 64, 4, None) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblasDpotrf_batch_strided(h.h, b"L", -1, kt.P(dA), 8, 64, 4, None) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblasDpotrf_batch_strided(h.h, b"L", 8, kt.P(dA), 8, 64, 0, None) == pykblas.KBLAS_Success
    assert lib.kblasDpotrf_batch_strided(h.h, b"X", 8, kt.P(dA), 8, 64, 4, None) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblasDposv_batch_strided(h.h, b"R", b"U", 1, 8, kt.P(dA), 8, 64, kt.P(dA), 1, 8, 4, None) == pykblas.KBLAS_Success
    assert lib.kblasDposv_batch_strided(h.h, b"R", b"L", 2, 8, kt.P(dA), 8, 64, kt.P(dA), 1, 16, 4, None) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblasDpotrf_batch(h.h, b"L", 8, P(0), 8, 4, None) == pykblas.KBLAS_Error_WrongInput


@pytest.mark.parametrize("n", [33, 40, 48])
def test_spotrf_blocked_rows_single_precision(h, n):
    # 32 < n <= 48 POTRF runs the blocked DPP-row kernel (FP32 instance)
    batch = 19
    A = po.gen_spd_shift_batch(14, n, batch)
    dA = kt.to_dev(nan_upper(A), np.float32)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasSpotrf_batch_strided(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    L = kt.to_host(dA).astype(np.float64)
    for i in range(batch):
        assert po.potrf_berr(po.col(A, i), po.col(L, i)) <= 4 * n * kt.EPS32
        assert np.all(np.isnan(po.col(L, i)[np.triu_indices(n, 1)]))


@pytest.mark.parametrize("n", [70, 129, 200])
def test_sposv_blocked_single_precision(h, n):
    # n > 64: the blocked workgroup kernel (FP32 instance): register-tile
    # diagonal blocks, paired panels, fused forward sweep
    batch = 5
    A = po.gen_spd_shift_batch(16, n, batch)
    B = po.gen_rhs_batch(17, 3, n, batch)
    dA = kt.to_dev(nan_upper(A), np.float32)
    dB = kt.to_dev(B, np.float32)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasSposv_batch_strided(h.h, b"R", b"L", 3, n, kt.P(dA), n, n * n, kt.P(dB), 3, 3 * n, batch,
                                        kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    L, X = kt.to_host(dA).astype(np.float64), kt.to_host(dB).astype(np.float64)
    for i in range(batch):
        Ai = po.col(A, i)
        assert po.potrf_berr(Ai, po.col(L, i)) <= 4 * n * kt.EPS32
        assert po.solve_berr("R", Ai, po.col(X, i), po.col(B, i)) <= 8 * n * kt.EPS32
        assert np.all(np.isnan(po.col(L, i)[np.triu_indices(n, 1)]))
