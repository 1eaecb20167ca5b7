"""GPU parity tests of the batched inverse family (TRTRI, LAUUM, POTRI, POTI)
and TRMM against the CPU oracle (pinned in test_oracle_inv.py).

Bars: FP64 relative max-norm error within c * n * eps64 of the oracle (the
kernels follow the oracle's recurrences; for n <= 32 TRTRI's sums run in the
other direction and use a refined reciprocal of the pivot), FP32 within
c * n * eps32 of the FP64 oracle on the same FP32-rounded inputs; integer info
and untouched memory (strict upper triangle, leading-dimension padding, a
singular matrix, POTI after a failed factorization) bit-exact; strided and
pointer-array calls bit-identical.  Orders cover the one-wave, multi-wave,
LDS-staged and in-place (not staged) kernel paths."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu
EPS = kt.EPS64
RNG = np.random.default_rng(77)
# register form up to 64 (16, 32 or 64 lanes per matrix); above: one workgroup per matrix,
# staged in LDS up to n = 87 (FP64) / 123 (FP32), in place in HBM beyond
SIZES = [1, 2, 5, 16, 31, 32, 33, 64, 65, 88, 100, 130]


@pytest.fixture(scope="module")
def h():
    hh = kt.handle()
    yield hh
    hh.close()


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def factors(n, batch, seed, lda=None):
    """batch of lower Cholesky factors (upper triangle = the SPD input's), ld lda"""
    lda = lda or n
    A = po.gen_spd_shift_batch(seed, n, batch)
    out = po.batch_array(batch, lda, n)
    out[:] = -7.0                                   # padding sentinel
    for i in range(batch):
        L, info = po.potrf(po.col(A, i))
        assert info == 0
        out[i][:, :n] = L.T
    return A, out


def call_inv(h, name, prec, n, dA, lda, batch, info, diag=None, strided=True):
    f = getattr(h.lib, "kblas%s%s_batch%s" % (prec, name, "_strided" if strided else ""))
    if name == "trtri" and diag is None:
        diag = "N"
    args = [h.h, b"L"] + ([diag.encode()] if diag else []) + [n]
    if strided:
        args += [kt.P(dA), lda, lda * n, batch, kt.P(info)]
    else:
        args += [kt.P(kt.ptr_array(dA, batch, lda * n)), lda, batch, kt.P(info)]
    return f(*args)


def oracle_op(name, L, diag="N"):
    if name == "trtri":
        return po.trtri(diag, L)
    if name == "lauum":
        return po.lauum(L), 0
    return po.potri(L)


@pytest.mark.parametrize("name", ["trtri", "lauum", "potri"])
@pytest.mark.parametrize("n", SIZES)
def test_dinv_strided(h, name, n):
    batch = 5
    torch = kt.torch_dev()
    _, L = factors(n, batch, 11 + n)
    dL = kt.to_dev(L)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    assert call_inv(h, name, "D", n, dL, n, batch, info) == pykblas.KBLAS_Success
    X = kt.to_host(dL)
    assert np.array_equal(kt.to_host(info), np.zeros(batch, np.int32))
    iu = np.triu_indices(n, 1)
    for i in range(batch):
        Xo, inf = oracle_op(name, po.col(L, i))
        assert inf == 0
        Xg = po.col(X, i)
        assert rel(np.tril(Xg), np.tril(Xo)) <= 16 * n * EPS, (name, n, i)
        assert np.array_equal(Xg[iu], po.col(L, i)[iu])


@pytest.mark.parametrize("n", [7, 33, 100])
def test_dtrtri_unit(h, n):
    batch = 4
    torch = kt.torch_dev()
    U = po.batch_array(batch, n, n)
    for i in range(batch):
        T = np.tril(RNG.uniform(-0.5, 0.5, (n, n)) / np.sqrt(n), -1) + np.diag(RNG.uniform(2, 3, n))
        T[np.triu_indices(n, 1)] = 5.0
        U[i] = T.T
    dU = kt.to_dev(U)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    assert call_inv(h, "trtri", "D", n, dU, n, batch, info, diag="U") == pykblas.KBLAS_Success
    X = kt.to_host(dU)
    assert np.array_equal(kt.to_host(info), np.zeros(batch, np.int32))
    for i in range(batch):
        Xo, _ = po.trtri("U", po.col(U, i))
        Xg = po.col(X, i)
        assert rel(np.tril(Xg, -1), np.tril(Xo, -1)) <= 16 * n * EPS
        assert np.array_equal(np.diag(Xg), np.diag(po.col(U, i)))    # unit diagonal not written
        assert np.array_equal(np.triu(Xg, 1), np.triu(po.col(U, i), 1))


@pytest.mark.parametrize("name", ["trtri", "lauum", "potri", "poti"])
@pytest.mark.parametrize("n", [16, 40, 100])
def test_dinv_pointer_array_equals_strided_with_padding(h, name, n):
    batch, lda = 6, n + 3
    torch = kt.torch_dev()
    A, L = factors(n, batch, 5 + n, lda)
    if name == "poti":   # POTI takes the SPD matrix itself
        for i in range(batch):
            L[i][:, :n] = A[i]
    d1, d2 = kt.to_dev(L), kt.to_dev(L)
    i1 = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    i2 = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    assert call_inv(h, name, "D", n, d1, lda, batch, i1, strided=True) == pykblas.KBLAS_Success
    assert call_inv(h, name, "D", n, d2, lda, batch, i2, strided=False) == pykblas.KBLAS_Success
    X1, X2 = kt.to_host(d1), kt.to_host(d2)
    assert np.array_equal(X1, X2)
    assert np.array_equal(kt.to_host(i1), np.zeros(batch, np.int32))
    assert np.all(X1[:, :, n:] == -7.0)             # leading-dimension padding untouched
    if name == "poti":
        for i in range(batch):
            Ai = A[i].T
            Xi = np.tril(X1[i][:, :n].T)
            Xf = Xi + np.tril(Xi, -1).T
            assert rel(Xf @ Ai, np.eye(n)) <= 64 * n * EPS


@pytest.mark.parametrize("n,p", [(6, 3), (40, 1), (100, 100)])
def test_dtrtri_potri_zero_diagonal(h, n, p):
    batch = 3
    torch = kt.torch_dev()
    _, L = factors(n, batch, 99)
    L[1][p - 1, p - 1] = 0.0                        # matrix 1 singular at p
    for name in ("trtri", "potri"):
        dL = kt.to_dev(L)
        info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
        assert call_inv(h, name, "D", n, dL, n, batch, info) == pykblas.KBLAS_Success
        X = kt.to_host(dL)
        assert list(kt.to_host(info)) == [0, p, 0]
        assert np.array_equal(X[1], L[1])           # left unchanged, as LAPACK


def test_dpoti_failed_factorization_keeps_potrf_result(h):
    n, batch = 24, 3
    torch = kt.torch_dev()
    A = po.gen_spd_shift_batch(3, n, batch)
    A[2][7, 7] = -50.0                              # first non-positive minor at 8
    dA, dP = kt.to_dev(A), kt.to_dev(A)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    ip = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    assert h.lib.kblasDpoti_batch_strided(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info)) == pykblas.KBLAS_Success
    assert h.lib.kblasDpotrf_batch_strided(h.h, b"L", n, kt.P(dP), n, n * n, batch, kt.P(ip)) == pykblas.KBLAS_Success
    X, Pf = kt.to_host(dA), kt.to_host(dP)
    assert list(kt.to_host(info)) == [0, 0, 8] and list(kt.to_host(ip)) == [0, 0, 8]
    assert np.array_equal(X[2], Pf[2], equal_nan=True)   # POTRF leaves NaN past the failed pivot
    for i in range(2):
        Xi = np.tril(X[i].T)
        assert rel((Xi + np.tril(Xi, -1).T) @ A[i].T, np.eye(n)) <= 64 * n * EPS


@pytest.mark.parametrize("name", ["trtri", "lauum", "potri"])
@pytest.mark.parametrize("n", [8, 32, 64, 100, 130])
def test_sinv_strided(h, name, n):
    batch = 4
    torch = kt.torch_dev()
    _, L = factors(n, batch, 21 + n)
    L32 = L.astype(np.float32)
    dL = kt.to_dev(L32)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    assert call_inv(h, name, "S", n, dL, n, batch, info) == pykblas.KBLAS_Success
    X = kt.to_host(dL)
    assert np.array_equal(kt.to_host(info), np.zeros(batch, np.int32))
    for i in range(batch):
        Xo, _ = oracle_op(name, po.col(L32, i).astype(np.float64))
        assert rel(np.tril(po.col(X, i)).astype(np.float64), np.tril(Xo)) <= 16 * n * kt.EPS32


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("mn", [(16, 4), (5, 64), (70, 33), (200, 3), (3, 300)])
def test_dtrmm_strided(h, side, trans, diag, mn):
    m, n = mn
    batch = 5
    order = m if side == "L" else n
    _, L = factors(order, batch, 31 + order)
    B = RNG.standard_normal((batch, n, m))
    dL, dB = kt.to_dev(L), kt.to_dev(B)
    st = h.lib.kblasDtrmm_batch_strided(h.h, side.encode(), b"L", trans.encode(), diag.encode(), m, n, -0.5,
                                        kt.P(dL), order, order * order, kt.P(dB), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.trmm(side, trans, diag, -0.5, po.col(L, i), po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 16 * order * EPS


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtrmm_panel_loop(h, side, trans):
    """an order too large for the triangle to be staged in LDS, so B goes
    through LDS in panels of whole lines (1100 x 40 / 40 x 1100: six
    seven-line panels), for both transposes"""
    m, n = (1100, 40) if side == "L" else (40, 1100)
    batch = 2
    order = m if side == "L" else n
    _, L = factors(order, batch, 77)
    B = RNG.standard_normal((batch, n, m))
    dL, dB = kt.to_dev(L), kt.to_dev(B)
    st = h.lib.kblasDtrmm_batch_strided(h.h, side.encode(), b"L", trans.encode(), b"N", m, n, 1.5,
                                        kt.P(dL), order, order * order, kt.P(dB), m, m * n, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.trmm(side, trans, "N", 1.5, po.col(L, i), po.col(B, i))
        assert rel(po.col(X, i), Xo) <= 16 * order * EPS


@pytest.mark.parametrize("side", ["L", "R"])
def test_strmm_pointer_array_and_alpha_zero(h, side):
    m, n, batch = 20, 12, 4
    order = m if side == "L" else n
    _, L = factors(order, batch, 41)
    L32 = L.astype(np.float32)
    B = RNG.standard_normal((batch, n, m)).astype(np.float32)
    dL, dB = kt.to_dev(L32), kt.to_dev(B)
    pL, pB = kt.ptr_array(dL, batch, order * order), kt.ptr_array(dB, batch, m * n)
    st = h.lib.kblasStrmm_batch(h.h, side.encode(), b"L", b"N", b"N", m, n, 2.0, kt.P(pL), order, kt.P(pB), m, batch)
    assert st == pykblas.KBLAS_Success
    X = kt.to_host(dB)
    for i in range(batch):
        Xo = po.trmm(side, "N", "N", 2.0, po.col(L32, i).astype(np.float64), po.col(B, i).astype(np.float64))
        assert rel(po.col(X, i).astype(np.float64), Xo) <= 16 * order * kt.EPS32
    dB.fill_(float("nan"))
    st = h.lib.kblasStrmm_batch(h.h, side.encode(), b"L", b"N", b"N", m, n, 0.0, kt.P(pL), order, kt.P(pB), m, batch)
    assert st == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(dB), np.zeros_like(B))


def test_inv_edge_cases(h):
    torch = kt.torch_dev()
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    info = torch.full((3,), -1, dtype=torch.int32, device="cuda")
    # n = 0 zeroes info; batchCount = 0 is a no-op; Upper is not implemented
    assert h.lib.kblasDpotri_batch_strided(h.h, b"L", 0, kt.P(d), 1, 0, 3, kt.P(info)) == pykblas.KBLAS_Success
    assert list(kt.to_host(info)) == [0, 0, 0]
    assert h.lib.kblasDtrtri_batch_strided(h.h, b"L", b"N", 4, kt.P(d), 4, 16, 0, kt.P(info)) == pykblas.KBLAS_Success
    assert h.lib.kblasDlauum_batch_strided(h.h, b"U", 4, kt.P(d), 4, 16, 1, kt.P(info)) == pykblas.KBLAS_NotImplemented
    assert h.lib.kblasDtrmm_batch_strided(h.h, b"L", b"U", b"N", b"N", 4, 4, 1.0, kt.P(d), 4, 16, kt.P(d), 4, 16,
                                          1) == pykblas.KBLAS_NotImplemented
    # too-small stride / leading dimension; orders above 1024 are not served
    assert h.lib.kblasDpotri_batch_strided(h.h, b"L", 4, kt.P(d), 4, 15, 2, kt.P(info)) == pykblas.KBLAS_Error_WrongInput
    assert h.lib.kblasDpotri_batch_strided(h.h, b"L", 4, kt.P(d), 3, 16, 2, kt.P(info)) == pykblas.KBLAS_Error_WrongInput
    assert h.lib.kblasDpotri_batch_strided(h.h, b"L", 1025, kt.P(d), 1025, 1025 * 1025, 1,
                                           kt.P(info)) == pykblas.KBLAS_NotImplemented


# ------------------------------------------------------------------ PSTRF ----
def pstrf_inputs(n, r, batch, seed):
    g = np.random.default_rng(seed)
    A = po.batch_array(batch, n, n)
    for i in range(batch):
        Bm = g.standard_normal((n, r))
        M = Bm @ Bm.T + (np.diag(g.uniform(0.5, 1.5, n)) if r == n else 0.0)
        M[np.tril_indices(n, -1)] = 1e30          # only the upper triangle is read
        A[i] = M.T
    return A


@pytest.mark.parametrize("n,r", [(1, 1), (5, 5), (16, 16), (31, 9), (32, 32), (33, 33), (48, 20), (64, 64), (64, 30)])
def test_dpstrf_strided_matches_oracle(h, n, r):
    # n <= 32 pairs two matrices per wave: an odd batch leaves a live matrix
    # beside an idle half in the last wave
    batch = 7 if n <= 32 else 6
    torch = kt.torch_dev()
    A = pstrf_inputs(n, r, batch, 100 + n + r)
    dA = kt.to_dev(A)
    piv = torch.full((batch, n), -1, dtype=torch.int32, device="cuda")
    ranks = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    st = h.lib.kblasDpstrf_batch_strided(h.h, n, kt.P(dA), n, n * n, kt.P(piv), n, kt.P(ranks), batch)
    assert st == pykblas.KBLAS_Success
    R, pv, rk = kt.to_host(dA), kt.to_host(piv), kt.to_host(ranks)
    for i in range(batch):
        Ro, po_piv, ro = po.pstrf(po.col(A, i))
        assert rk[i] == ro == r
        assert np.array_equal(pv[i], po_piv)                      # integer output bit-exact
        Rg = po.col(R, i)
        assert np.array_equal(np.tril(Rg, -1), np.zeros((n, n)))
        assert rel(np.triu(Rg)[:r], np.triu(Ro)[:r]) <= 16 * n * EPS


def test_dpstrf_ties_and_pointer_array(h):
    """Identity matrices (all pivots tie): pivots follow the reference's reduction, as the
    oracle; pointer-array and strided calls give identical results."""
    torch = kt.torch_dev()
    n, batch = 40, 3
    A = po.batch_array(batch, n, n)
    for i in range(batch):
        A[i] = np.eye(n) * (i + 1)
    d1, d2 = kt.to_dev(A), kt.to_dev(A)
    p1 = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    p2 = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    r1 = torch.zeros(batch, dtype=torch.int32, device="cuda")
    r2 = torch.zeros(batch, dtype=torch.int32, device="cuda")
    assert h.lib.kblasDpstrf_batch_strided(h.h, n, kt.P(d1), n, n * n, kt.P(p1), n, kt.P(r1), batch) == pykblas.KBLAS_Success
    pA = kt.ptr_array(d2, batch, n * n)
    pP = kt.ptr_array(p2, batch, n)
    assert h.lib.kblasDpstrf_batch(h.h, n, kt.P(pA), n, kt.P(pP), kt.P(r2), batch) == pykblas.KBLAS_Success
    assert np.array_equal(kt.to_host(d1), kt.to_host(d2))
    assert np.array_equal(kt.to_host(p1), kt.to_host(p2)) and np.array_equal(kt.to_host(r1), kt.to_host(r2))
    _, po_piv, _ = po.pstrf(np.eye(n))
    assert all(np.array_equal(row, po_piv) for row in kt.to_host(p1))


@pytest.mark.parametrize("n", [16, 64])
def test_spstrf_strided(h, n):
    batch = 4
    torch = kt.torch_dev()
    A = pstrf_inputs(n, n, batch, 7 + n)
    A32 = A.astype(np.float32)
    dA = kt.to_dev(A32)
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    ranks = torch.zeros(batch, dtype=torch.int32, device="cuda")
    assert h.lib.kblasSpstrf_batch_strided(h.h, n, kt.P(dA), n, n * n, kt.P(piv), n, kt.P(ranks), batch) == pykblas.KBLAS_Success
    R, pv = kt.to_host(dA), kt.to_host(piv)
    for i in range(batch):
        M = np.triu(po.col(A32, i).astype(np.float64))
        M = M + np.triu(M, 1).T
        Pm = np.eye(n)[:, pv[i]]
        Rg = np.triu(po.col(R, i)).astype(np.float64)
        assert rel(Rg.T @ Rg, Pm.T @ M @ Pm) <= 16 * n * kt.EPS32
        assert sorted(pv[i].tolist()) == list(range(n))


def test_pstrf_edge_cases(h):
    torch = kt.torch_dev()
    d = torch.zeros(65 * 65, dtype=torch.float64, device="cuda")
    piv = torch.zeros(65, dtype=torch.int32, device="cuda")
    ranks = torch.full((2,), -1, dtype=torch.int32, device="cuda")
    assert h.lib.kblasDpstrf_batch_strided(h.h, 65, kt.P(d), 65, 65 * 65, kt.P(piv), 65, kt.P(ranks), 1) == pykblas.KBLAS_Error_WrongInput
    assert h.lib.kblasDpstrf_batch_strided(h.h, 0, kt.P(d), 1, 0, kt.P(piv), 0, kt.P(ranks), 2) == pykblas.KBLAS_Success
    assert list(kt.to_host(ranks)) == [0, 0]
    # zero matrix: rank 0, nothing factored
    z = torch.zeros(16, dtype=torch.float64, device="cuda")
    assert h.lib.kblasDpstrf_batch_strided(h.h, 4, kt.P(z), 4, 16, kt.P(piv), 4, kt.P(ranks), 1) == pykblas.KBLAS_Success
    assert kt.to_host(ranks)[0] == 0 and np.array_equal(kt.to_host(piv)[:4], np.arange(4))
