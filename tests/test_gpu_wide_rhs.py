"""GPU parity tests of POSV / POTRS with many right-hand sides at orders
32 < n <= 64, which run as a factorization followed by the two sweeps of the
matrix-core TRSM (kb_launch_trsm_solve), against the oracle (orc_posv,
orc_potrs).

Bar (FP64, the same as the other Cholesky tests): solutions within
64 n eps64 of the oracle's (max-norm, relative to max(|X|, 1)) and normwise
backward error <= 4 n eps64; FP32 against the FP64 oracle with eps32.  A
matrix that is not positive definite gets LAPACK's info and keeps its
right-hand sides unchanged; the strict upper triangle holds NaN (never read).
Both sides, strided and pointer-array operands, right-hand-side counts at,
between and above the 32-wide wavefront chunks."""
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


def nan_upper(Ab):
    B = Ab.copy()
    n = B.shape[1]
    for c in range(n):
        B[:, c, :c] = np.nan
    return B


@pytest.mark.parametrize("p", ["D", "S"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("n", [33, 40, 48, 64])
@pytest.mark.parametrize("nrhs", [32, 70])
def test_posv_wide(h, p, side, n, nrhs):
    dt, eps = (np.float64, kt.EPS64) if p == "D" else (np.float32, kt.EPS32)
    batch = 9
    A = po.gen_spd_shift_batch(20 + n, n, batch)
    bad = 5
    A[bad][n // 2, n // 2] = -1.0             # not positive definite: info = n/2 + 1 (or earlier)
    rows, cols = (n, nrhs) if side == "L" else (nrhs, n)
    B = po.gen_rhs_batch(30 + nrhs, rows, cols, batch)
    if dt == np.float32:
        A = A.astype(np.float32).astype(np.float64)
        B = B.astype(np.float32).astype(np.float64)
    dA, dB = kt.to_dev(nan_upper(A), dt), kt.to_dev(B, dt)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    if n % 2:
        pA, pB = kt.ptr_array(dA, batch, n * n), kt.ptr_array(dB, batch, rows * cols)
        st = getattr(h.lib, "kblas%sposv_batch" % p)(h.h, side.encode(), b"L", rows, cols, kt.P(pA), n, kt.P(pB),
                                                     rows, batch, kt.P(info))
    else:
        st = getattr(h.lib, "kblas%sposv_batch_strided" % p)(h.h, side.encode(), b"L", rows, cols, kt.P(dA), n,
                                                             n * n, kt.P(dB), rows, rows * cols, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB).astype(np.float64)
    inf = kt.to_host(info)
    for i in range(batch):
        Ai, Bi = po.col(A, i), po.col(B, i)
        Lo, Xo, io = po.posv(side, Ai, Bi)
        assert inf[i] == io, i
        if i == bad:
            assert io > 0
            assert np.array_equal(po.col(X, i), Bi), "a failed factorization leaves B as it was"
            continue
        Xi = po.col(X, i)
        assert np.max(np.abs(Xi - Xo)) <= 64 * n * eps * max(np.max(np.abs(Xo)), 1), i
        assert po.solve_berr(side, Ai, Xi, Bi) <= 4 * n * eps, i


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("n", [35, 64])
def test_potrs_wide_uses_given_factor(h, side, n):
    batch = 7
    nrhs = 64
    A = po.gen_spd_shift_batch(40 + n, n, batch)
    Lh = np.zeros_like(A)
    for i in range(batch):
        Lh[i] = np.tril(po.potrf(po.col(A, i))[0]).T
    rows, cols = (n, nrhs) if side == "L" else (nrhs, n)
    B = po.gen_rhs_batch(50, rows, cols, batch)
    dL, dB = kt.to_dev(nan_upper(Lh)), kt.to_dev(B)
    st = h.lib.kblasDpotrs_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dL), n, n * n, kt.P(dB), rows,
                                         rows * cols, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.potrs(side, po.col(Lh, i), po.col(B, i))
        assert np.max(np.abs(po.col(X, i) - Xo)) <= 64 * n * kt.EPS64 * max(np.max(np.abs(Xo)), 1)
