"""CPU tests that pin the oracle's inverse family and TRMM (oracle/kblas_oracle.c:
orc_trtri, orc_lauum, orc_potri, orc_trmm) before the GPU tests trust it.

Pins, per restated function:
  - golden vectors from MKL dtrtri / dlauum / dpotri / dtrmm
    (tests/golden/lapack_golden_inv.npz, written by make_golden.py --inv);
  - identities that fix the result independently of any library:
    L X = I for TRTRI (non-unit and unit), LAUUM = L^T L by a plain matrix
    product, POTRI(L) A = I for A = L L^T, TRMM against the dense product of
    the triangle, closed forms (diagonal, n = 1, unit bidiagonal), and
    LAPACK's info on planted zero diagonals (matrix left unchanged).
"""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lapack_golden_inv.npz"))
EPS = np.finfo(np.float64).eps
SIZES = [1, 2, 3, 5, 8, 16, 17, 32, 33, 64, 65]


def relmax(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    return d / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("n", SIZES)
def test_trtri_matches_lapack(n):
    L = G["inv_L_%d" % n]
    X, info = po.trtri("N", L)
    assert info == 0
    assert relmax(np.tril(X), G["trtri_X_%d" % n]) <= 8 * n * EPS
    iu = np.triu_indices(n, 1)
    assert np.array_equal(X[iu], L[iu])           # strict upper untouched
    assert relmax(np.tril(L) @ np.tril(X), np.eye(n)) <= 8 * n * EPS


@pytest.mark.parametrize("n", SIZES)
def test_trtri_unit_matches_lapack(n):
    U = G["trtriU_A_%d" % n]
    X, info = po.trtri("U", U)
    assert info == 0
    if n > 1:
        assert relmax(np.tril(X, -1), G["trtriU_X_%d" % n]) <= 8 * n * EPS
    assert np.array_equal(np.diag(X), np.diag(U))  # unit diagonal neither read nor written
    Lu = np.tril(U, -1) + np.eye(n)
    Xu = np.tril(X, -1) + np.eye(n)
    assert relmax(Lu @ Xu, np.eye(n)) <= 8 * n * EPS


@pytest.mark.parametrize("n", SIZES)
def test_lauum_matches_lapack_and_product(n):
    L = G["inv_L_%d" % n]
    X = po.lauum(L)
    assert relmax(np.tril(X), G["lauum_X_%d" % n]) <= 8 * n * EPS
    Lt = np.tril(L)
    assert relmax(np.tril(X), np.tril(Lt.T @ Lt)) <= 8 * n * EPS


@pytest.mark.parametrize("n", SIZES)
def test_potri_matches_lapack_and_inverts(n):
    L = G["inv_L_%d" % n]
    A = G["inv_A_%d" % n]
    X, info = po.potri(L)
    assert info == 0
    assert relmax(np.tril(X), G["potri_X_%d" % n]) <= 16 * n * EPS
    Xf = np.tril(X) + np.tril(X, -1).T
    assert relmax(Xf @ A, np.eye(n)) <= 64 * n * EPS


@pytest.mark.parametrize("n,p", [(6, 3), (20, 1), (40, 40)])
def test_trtri_zero_diagonal_info(n, p):
    L = G["trtri_sing_L_%d" % n]
    assert int(G["trtri_sing_info_%d" % n]) == p
    X, info = po.trtri("N", L)
    assert info == p
    assert np.array_equal(X, L)
    X, info = po.potri(L)
    assert info == p and np.array_equal(X, L)


def test_diagonal_closed_form():
    d = np.array([4.0, 0.5, 2.0, 8.0, 1.0])
    X, info = po.trtri("N", np.diag(d))
    assert info == 0 and np.array_equal(np.diag(X), 1.0 / d)
    assert np.array_equal(np.diag(po.lauum(np.diag(d))), d * d)
    X, _ = po.potri(np.diag(np.sqrt(d)))
    assert np.allclose(np.diag(X), 1.0 / d, rtol=4 * EPS, atol=0)


def test_n1():
    X, info = po.trtri("N", np.array([[4.0]]))
    assert info == 0 and X[0, 0] == 0.25
    assert po.lauum(np.array([[3.0]]))[0, 0] == 9.0
    X, _ = po.potri(np.array([[2.0]]))
    assert X[0, 0] == 0.25


def test_unit_bidiagonal_closed_form():
    # L = I + c * subdiagonal  =>  L^-1(i,j) = (-c)^(i-j) for i >= j
    n, c = 9, 0.5
    L = np.eye(n) + c * np.eye(n, k=-1)
    X, _ = po.trtri("N", L)
    i, j = np.tril_indices(n)
    assert np.array_equal(X[i, j], (-c) ** (i - j).astype(float))


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n", [5, 16, 40])
def test_trmm_matches_blas_and_product(side, trans, diag, n):
    key = "trmm_%s%s%s_%d" % (side, trans, diag, n)
    L, B = G[key + "_A"], G[key + "_B"]
    X = po.trmm(side, trans, diag, 0.75, L, B)
    assert relmax(X, G[key + "_X"]) <= 8 * n * EPS
    T = np.tril(L, -1) + (np.eye(n) if diag == "U" else np.diag(np.diag(L)))
    opT = T.T if trans == "T" else T
    ref = 0.75 * (opT @ B if side == "L" else B @ opT)
    assert relmax(X, ref) <= 8 * n * EPS


def test_trmm_alpha_zero_clears_nan():
    B = np.full((3, 2), np.nan)
    X = po.trmm("L", "N", "N", 0.0, np.eye(3), B)
    assert np.array_equal(X, np.zeros((3, 2)))


def test_planted_mistake_is_caught():
    """A dropped diagonal term in the TRTRI sums (a plausible slip) breaks L X = I."""
    L = G["inv_L_16"]
    X, _ = po.trtri("N", L)
    Xbad = np.tril(X).copy()
    Xbad[5:, 3] *= 1.0 + 1e-9
    assert relmax(np.tril(L) @ Xbad, np.eye(16)) > 8 * 16 * EPS


PSTRF_CASES = [(5, 5), (16, 16), (32, 32), (33, 33), (48, 48), (64, 64), (20, 7), (40, 13), (64, 30)]


@pytest.mark.parametrize("n,r", PSTRF_CASES)
def test_pstrf_matches_lapack(n, r):
    """MKL dpstrf('U', tol < 0): same pivots and rank, R within c n eps (distinct diagonals, no ties)."""
    A = G["pstrf_A_%d_%d" % (n, r)]
    R, piv, rank = po.pstrf(A)
    assert rank == int(G["pstrf_rank_%d_%d" % (n, r)]) == r
    assert np.array_equal(piv[:rank], G["pstrf_piv_%d_%d" % (n, r)][:rank])
    assert relmax(np.triu(R)[:rank], G["pstrf_R_%d_%d" % (n, r)][:rank]) <= 8 * n * EPS
    assert np.array_equal(np.tril(R, -1), np.zeros((n, n)))
    # P^T A P = R^T R on the leading rank rows, the diagonal of R non-increasing
    Pm = np.eye(n)[:, piv]
    Rr = np.triu(R)[:rank]
    assert relmax((Pm.T @ A @ Pm)[:rank, :rank], (Rr.T @ Rr)[:rank, :rank]) <= 16 * n * EPS
    d = np.diag(R)[:rank]
    assert np.all(d[1:] <= d[:-1] * (1 + 1e-12))
    assert sorted(piv.tolist()) == list(range(n))


def test_pstrf_tie_rule_on_identity():
    """Equal diagonals: the pivot is the reference's 32-slot butterfly winner, not the first
    maximum (LAPACK's choice); for I_4 step 1 picks index 2 (slot 0 keeps the even-side
    winner 2 against the odd-side 1 on equal values)."""
    R, piv, rank = po.pstrf(np.eye(4))
    assert rank == 4 and list(piv) == [0, 2, 1, 3]
    assert np.array_equal(R, np.eye(4))


def test_pstrf_diagonal_closed_form():
    d = np.array([3.0, 9.0, 1.0, 4.0, 0.0])
    R, piv, rank = po.pstrf(np.diag(d))
    assert rank == 4 and list(piv[:4]) == [1, 3, 0, 2]
    assert np.array_equal(np.diag(R)[:4], np.sqrt(d[[1, 3, 0, 2]]))


def test_pstrf_reads_upper_triangle_only():
    A = G["pstrf_A_16_16"].copy()
    B = A.copy()
    B[np.tril_indices(16, -1)] = 1e30
    R1, p1, r1 = po.pstrf(A)
    R2, p2, r2 = po.pstrf(B)
    assert r1 == r2 and np.array_equal(p1, p2) and np.array_equal(R1, R2)
