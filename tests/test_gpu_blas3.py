"""GPU parity tests of the batched TRSM / SYRK / GEMM entry points (uniform
strided, pointer-array and non-uniform) and of the blocked (n > 64) Cholesky
family, against the CPU oracle.  FP64 bar: relative max-norm error within
c * k * eps64 of the oracle (k = inner dimension / order)."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu
EPS = kt.EPS64
RNG = np.random.default_rng(2024)


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def lower_factor_batch(n, batch, seed):
    A = po.gen_spd_shift_batch(seed, n, batch)
    L = np.zeros_like(A)
    for i in range(batch):
        L[i] = np.tril(po.potrf(po.col(A, i))[0]).T
    return L


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("mn", [(16, 16), (4, 64), (70, 33), (5, 100)])
def test_dtrsm_strided(h, side, trans, mn):
    m, n = mn
    batch = 9
    order = m if side == "L" else n
    L = lower_factor_batch(order, batch, 3)
    B = RNG.standard_normal((batch, n, m))
    dL, dB = kt.to_dev(L), kt.to_dev(B)
    st = h.lib.kblasDtrsm_batch_strided(h.h, side.encode(), b"L", trans.encode(), b"N", m, n, 0.5,
                                        kt.P(dL), order, order * order, kt.P(dB), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.trsm(side, trans, 0.5, po.col(L, i), po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 64 * order * EPS


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("mk", [(16, 16), (32, 32), (70, 20), (8, 130)])
def test_dsyrk_pointer_array(h, trans, mk):
    m, k = mk
    batch = 7
    shapeA = (m, k) if trans == "N" else (k, m)
    A = RNG.standard_normal((batch, shapeA[1], shapeA[0]))
    C = RNG.standard_normal((batch, m, m))
    dA, dC = kt.to_dev(A), kt.to_dev(C)
    pA, pC = kt.ptr_array(dA, batch, m * k), kt.ptr_array(dC, batch, m * m)
    st = h.lib.kblasDsyrk_batch(h.h, b"L", trans.encode(), m, k, -1.0, kt.P(pA), shapeA[0], 0.5, kt.P(pC), m, batch)
    assert st == pykblas.KBLAS_Success
    R = kt.to_host(dC)
    il, iu = np.tril_indices(m), np.triu_indices(m, 1)
    for i in range(batch):
        Ro = po.syrk(trans, -1.0, po.col(A, i), 0.5, po.col(C, i))
        assert rel(po.col(R, i)[il], Ro[il]) <= 16 * k * EPS
        assert np.array_equal(po.col(R, i)[iu], po.col(C, i)[iu])


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("mnk", [(4, 32, 32), (32, 16, 16), (100, 70, 45), (1, 1, 1)])
def test_dgemm_strided(h, ta, tb, mnk):
    m, n, k = mnk
    batch = 6
    sa = (m, k) if ta == "N" else (k, m)
    sb = (k, n) if tb == "N" else (n, k)
    A = RNG.standard_normal((batch, sa[1], sa[0]))
    B = RNG.standard_normal((batch, sb[1], sb[0]))
    C = RNG.standard_normal((batch, n, m))
    dA, dB, dC = kt.to_dev(A), kt.to_dev(B), kt.to_dev(C)
    st = h.lib.kblasDgemm_batch_strided(h.h, ta.encode(), tb.encode(), m, n, k, -1.0, kt.P(dA), sa[0], sa[0] * sa[1],
                                        kt.P(dB), sb[0], sb[0] * sb[1], 1.0, kt.P(dC), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    R = kt.to_host(dC)
    for i in range(batch):
        Ro = po.gemm(ta, tb, -1.0, po.col(A, i), po.col(B, i), 1.0, po.col(C, i))
        assert rel(po.col(R, i), Ro) <= 16 * k * EPS


def test_nonuniform_trsm_syrk_gemm(h):
    batch = 12
    ms = RNG.integers(1, 60, batch).astype(np.int32)
    ns = RNG.integers(1, 60, batch).astype(np.int32)
    # TRSM side R, trans T: B_i is m_i x n_i, L_i of order n_i
    Ls = [np.tril(po.potrf(po.gen_spd_shift(5, i, int(ns[i])))[0]) for i in range(batch)]
    Bs = [RNG.standard_normal((int(ms[i]), int(ns[i]))) for i in range(batch)]
    dLs = [kt.to_dev(np.asfortranarray(L).T.copy()) for L in Ls]   # column-major storage
    dBs = [kt.to_dev(np.asfortranarray(B).T.copy()) for B in Bs]
    torch = kt.torch_dev()
    pL = torch.tensor([t.data_ptr() for t in dLs], dtype=torch.int64, device="cuda")
    pB = torch.tensor([t.data_ptr() for t in dBs], dtype=torch.int64, device="cuda")
    dm, dn = kt.to_dev(ms), kt.to_dev(ns)
    st = h.lib.kblasDtrsm_vbatch(h.h, b"R", b"L", b"T", b"N", kt.P(dm), kt.P(dn), 0, 0, 1.0, kt.P(pL), kt.P(dn),
                                 kt.P(pB), kt.P(dm), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        X = kt.to_host(dBs[i]).T
        Xo = po.trsm("R", "T", 1.0, Ls[i], Bs[i])
        assert rel(X, Xo) <= 64 * int(ns[i]) * EPS
    # SYRK non-uniform: C_i (m_i x m_i) -= A_i A_i^T, A_i m_i x n_i
    As = [RNG.standard_normal((int(ms[i]), int(ns[i]))) for i in range(batch)]
    Cs = [RNG.standard_normal((int(ms[i]), int(ms[i]))) for i in range(batch)]
    dAs = [kt.to_dev(A.T.copy()) for A in As]
    dCs = [kt.to_dev(C.T.copy()) for C in Cs]
    pA = torch.tensor([t.data_ptr() for t in dAs], dtype=torch.int64, device="cuda")
    pC = torch.tensor([t.data_ptr() for t in dCs], dtype=torch.int64, device="cuda")
    st = h.lib.kblasDsyrk_vbatch(h.h, b"L", b"N", kt.P(dm), kt.P(dn), 0, 0, -1.0, kt.P(pA), kt.P(dm), 1.0,
                                 kt.P(pC), kt.P(dm), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        R = kt.to_host(dCs[i]).T
        Ro = po.syrk("N", -1.0, As[i], 1.0, Cs[i])
        il = np.tril_indices(int(ms[i]))
        assert rel(R[il], Ro[il]) <= 16 * int(ns[i]) * EPS
    # GEMM non-uniform: C_i (m_i x n_i) = A_i (m_i x k_i) B_i (k_i x n_i)
    ks = RNG.integers(1, 40, batch).astype(np.int32)
    As = [RNG.standard_normal((int(ms[i]), int(ks[i]))) for i in range(batch)]
    Bs = [RNG.standard_normal((int(ks[i]), int(ns[i]))) for i in range(batch)]
    Cs = [RNG.standard_normal((int(ms[i]), int(ns[i]))) for i in range(batch)]
    dAs = [kt.to_dev(A.T.copy()) for A in As]
    dBs = [kt.to_dev(B.T.copy()) for B in Bs]
    dCs = [kt.to_dev(C.T.copy()) for C in Cs]
    pA = torch.tensor([t.data_ptr() for t in dAs], dtype=torch.int64, device="cuda")
    pB = torch.tensor([t.data_ptr() for t in dBs], dtype=torch.int64, device="cuda")
    pC = torch.tensor([t.data_ptr() for t in dCs], dtype=torch.int64, device="cuda")
    dk = kt.to_dev(ks)
    st = h.lib.kblasDgemm_vbatch(h.h, b"N", b"N", kt.P(dm), kt.P(dn), kt.P(dk), 0, 0, 0, 2.0, kt.P(pA), kt.P(dm),
                                 kt.P(pB), kt.P(dk), 0.0, kt.P(pC), kt.P(dm), batch)
    assert st == pykblas.KBLAS_Success
    for i in range(batch):
        R = kt.to_host(dCs[i]).T
        assert rel(R, 2.0 * As[i] @ Bs[i]) <= 16 * int(ks[i]) * EPS


@pytest.mark.parametrize("n", [65, 100, 128, 200, 257])
def test_dposv_blocked_orders(h, n):
    batch = 5
    A = po.gen_spd_shift_batch(12, n, batch)
    B = po.gen_rhs_batch(13, n, 4, batch)
    Ain = A.copy()
    for c in range(n):
        Ain[:, c, :c] = np.nan
    dA, dB = kt.to_dev(Ain), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_batch_strided(h.h, b"L", b"L", n, 4, kt.P(dA), n, n * n, kt.P(dB), n, 4 * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    L, X = kt.to_host(dA), kt.to_host(dB)
    assert np.all(kt.to_host(info) == 0)
    for i in range(batch):
        Ai = po.col(A, i)
        assert po.potrf_berr(Ai, po.col(L, i)) <= 4 * n * EPS
        _, Xo, _ = po.posv("L", Ai, po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 64 * n * EPS
        assert np.all(np.isnan(po.col(L, i)[np.triu_indices(n, 1)]))


@pytest.mark.parametrize("n", [70, 161])
@pytest.mark.parametrize("side", ["L", "R"])
def test_dposv_blocked_many_rhs(h, n, side):
    # 20 right-hand sides: two passes (16 + 4) of the blocked solve
    batch = 3
    nrhs = 20
    A = po.gen_spd_shift_batch(15, n, batch)
    B = po.gen_rhs_batch(16, n, nrhs, batch) if side == "L" else po.gen_rhs_batch(16, nrhs, n, batch)
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    rows, cols = B.shape[2], B.shape[1]
    st = h.lib.kblasDposv_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dA), n, n * n, kt.P(dB), rows,
                                        rows * cols, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    X = kt.to_host(dB)
    for i in range(batch):
        Ai = po.col(A, i)
        _, Xo, _ = po.posv(side, Ai, po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 64 * n * EPS
        assert po.solve_berr(side, Ai, po.col(X, i), po.col(B, i)) <= 4 * n * EPS


@pytest.mark.parametrize("n", [80, 150])
def test_dpotrf_blocked_info_and_right_side_solve(h, n):
    batch = 4
    A = po.gen_spd_shift_batch(14, n, batch)
    A[2, 70, 70] = -1.0   # leading minor 71 not positive definite
    B = po.gen_rhs_batch(15, 3, n, batch)
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDposv_batch_strided(h.h, b"R", b"L", 3, n, kt.P(dA), n, n * n, kt.P(dB), 3, 3 * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    _, info_o = po.potrf(po.col(A, 2))
    assert inf[2] == info_o and inf[2] > 0
    X = kt.to_host(dB)
    for i in (0, 1, 3):
        assert inf[i] == 0
        _, Xo, _ = po.posv("R", po.col(A, i), po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 64 * n * EPS


@pytest.mark.parametrize("n,nrhs", [(96, 40), (129, 64), (193, 70)])
@pytest.mark.parametrize("side", ["L", "R"])
def test_dposv_blocked_wide_rhs(h, n, nrhs, side):
    # up to 64 right-hand sides per pass of the blocked solve: three and four
    # column blocks of 16, and two passes (64 + 6); n = 129 ends with a
    # one-column panel, n = 96 with a lone panel after a pair
    batch = 2
    A, _ = po.gen_spd_cond_batch(21, n, batch, 2.0, 5.0)
    B = po.gen_rhs_batch(22, n, nrhs, batch) if side == "L" else po.gen_rhs_batch(22, nrhs, n, batch)
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    rows, cols = B.shape[2], B.shape[1]
    st = h.lib.kblasDposv_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dA), n, n * n, kt.P(dB), rows,
                                        rows * cols, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    L, X = kt.to_host(dA), kt.to_host(dB)
    for i in range(batch):
        Ai = po.col(A, i)
        assert po.potrf_berr(Ai, po.col(L, i)) <= 4 * n * EPS
        assert po.solve_berr(side, Ai, po.col(X, i), po.col(B, i)) <= 4 * n * EPS


@pytest.mark.parametrize("n,nrhs", [(100, 3), (160, 50)])
def test_dpotrs_blocked_uses_given_factor(h, n, nrhs):
    batch = 3
    A = po.gen_spd_shift_batch(23, n, batch)
    Lh = np.zeros_like(A)
    for i in range(batch):
        Lh[i] = np.tril(po.potrf(po.col(A, i))[0]).T
    B = po.gen_rhs_batch(24, n, nrhs, batch)
    Lin = Lh.copy()
    for c in range(n):
        Lin[:, c, :c] = np.nan
    dL, dB = kt.to_dev(Lin), kt.to_dev(B)
    st = h.lib.kblasDpotrs_batch_strided(h.h, b"L", b"L", n, nrhs, kt.P(dL), n, n * n, kt.P(dB), n, n * nrhs,
                                         batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.potrs("L", po.col(Lh, i), po.col(B, i))
        assert np.max(np.abs(po.col(X, i) - Xo)) <= 64 * n * EPS * max(np.max(np.abs(Xo)), 1)


@pytest.mark.parametrize("fail_col", [40, 100, 130])
def test_dpotrf_blocked_info_in_second_panel(h, fail_col):
    # panels go in pairs: columns 32..63 and 96..127 are the second panel of a
    # pair, 128.. the first of the third pair
    n, batch = 150, 3
    A = po.gen_spd_shift_batch(25, n, batch)
    A[1, fail_col, fail_col] = -2.0
    dA = kt.to_dev(A)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    st = h.lib.kblasDpotrf_batch_strided(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf = kt.to_host(info)
    L = kt.to_host(dA)
    _, io = po.potrf(po.col(A, 1))
    assert inf[1] == io and 0 < io <= fail_col + 1
    for i in (0, 2):
        assert inf[i] == 0
        assert po.potrf_berr(po.col(A, i), po.col(L, i)) <= 4 * n * EPS


@pytest.mark.parametrize("n,nrhs,side", [(100, 64, "L"), (150, 80, "R"), (256, 64, "L")])
def test_dposv_recursive_solve_failed_matrix_keeps_b(h, n, nrhs, side):
    # orders above 64 with 64 or more right-hand sides: blocked factorization,
    # then the recursive solves; the matrix whose factorization fails keeps
    # its B bit for bit, the others match LAPACK
    batch = 4
    A = po.gen_spd_shift_batch(31, n, batch)
    A[2, n // 2, n // 2] = -1.0
    B = po.gen_rhs_batch(32, n, nrhs, batch) if side == "L" else po.gen_rhs_batch(32, nrhs, n, batch)
    dA, dB = kt.to_dev(A), kt.to_dev(B)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    rows, cols = B.shape[2], B.shape[1]
    st = h.lib.kblasDposv_batch_strided(h.h, side.encode(), b"L", rows, cols, kt.P(dA), n, n * n, kt.P(dB), rows,
                                        rows * cols, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    inf, X = kt.to_host(info), kt.to_host(dB)
    _, info_o = po.potrf(po.col(A, 2))
    assert inf[2] == info_o and inf[2] > 0
    assert np.array_equal(X[2], B[2])
    for i in (0, 1, 3):
        assert inf[i] == 0
        Ai = po.col(A, i)
        assert po.solve_berr(side, Ai, po.col(X, i), po.col(B, i)) <= 4 * n * EPS
        _, Xo, _ = po.posv(side, Ai, po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 64 * n * EPS


def test_dpotrs_recursive_pointer_array(h):
    n, nrhs, batch = 130, 70, 3
    A = po.gen_spd_shift_batch(33, n, batch)
    Lh = np.zeros_like(A)
    for i in range(batch):
        Lh[i] = np.tril(po.potrf(po.col(A, i))[0]).T
    B = po.gen_rhs_batch(34, nrhs, n, batch)     # right side: nrhs x n
    torch = kt.torch_dev()
    dL, dB = kt.to_dev(Lh), kt.to_dev(B)
    pL = torch.tensor([dL[i].data_ptr() for i in range(batch)], dtype=torch.int64, device="cuda")
    pB = torch.tensor([dB[i].data_ptr() for i in range(batch)], dtype=torch.int64, device="cuda")
    st = h.lib.kblasDpotrs_batch(h.h, b"R", b"L", nrhs, n, kt.P(pL), n, kt.P(pB), nrhs, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.potrs("R", po.col(Lh, i), po.col(B, i))
        assert np.max(np.abs(po.col(X, i) - Xo)) <= 64 * n * EPS * max(np.max(np.abs(Xo)), 1)


@pytest.mark.parametrize("p,n,nrhs,side", [("D", 300, 70, "L"), ("S", 130, 64, "R")])
def test_posv_recursive_solve_orders(h, p, n, nrhs, side):
    # the recursive solves above order 256 (FP64) and in FP32: backward error
    # of every system within the working precision's bar
    dt, eps = (np.float64, kt.EPS64) if p == "D" else (np.float32, kt.EPS32)
    batch = 3
    A = po.gen_spd_shift_batch(35, n, batch)
    B = po.gen_rhs_batch(36, n, nrhs, batch) if side == "L" else po.gen_rhs_batch(36, nrhs, n, batch)
    A, B = A.astype(dt).astype(np.float64), B.astype(dt).astype(np.float64)
    dA, dB = kt.to_dev(A.astype(dt)), kt.to_dev(B.astype(dt))
    info = kt.to_dev(np.full(batch, -7, np.int32))
    rows, cols = B.shape[2], B.shape[1]
    st = getattr(h.lib, "kblas%sposv_batch_strided" % p)(h.h, side.encode(), b"L", rows, cols, kt.P(dA), n, n * n,
                                                         kt.P(dB), rows, rows * cols, batch, kt.P(info))
    assert st == pykblas.KBLAS_Success
    assert np.all(kt.to_host(info) == 0)
    X = kt.to_host(dB).astype(np.float64)
    for i in range(batch):
        assert po.solve_berr(side, po.col(A, i), po.col(X, i), po.col(B, i)) <= 4 * n * eps
