"""CPU tests that pin the oracle's randomized-range-finder helpers
(oracle/kblas_oracle.c: orc_ara_potrf, orc_ara_syrk, orc_ara_trsm) before the
GPU tests trust them.

Pins, per restated function:
  - golden vectors from MKL (tests/golden/lapack_golden_ara.npz, written by
    make_golden.py --ara): dsyrk of FP32 sample blocks widened to FP64, dpotrf
    of the resulting Gram matrices (R = L^T, rank = order), dtrsm with L^T;
  - closed forms that fix the rank-revealing rule independently of any
    library: a rank-one integer Gram matrix (exact zero Schur complement),
    a diagonal matrix with a zero (stop, the rest left unfactored), the
    tolerance boundary max(diag) * 2^-52, dim clamped to bs, dim <= 0;
  - exact integer Gram products (every partial sum representable), only the
    upper triangle of R read by the solve, columns beyond bs untouched.
"""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lapack_golden_ara.npz"))
EPS = np.finfo(np.float64).eps
CASES = [(40, 1), (64, 7), (100, 16), (96, 17), (250, 29), (300, 32)]


def relmax(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    return d / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("m,n", CASES)
def test_gram_product_matches_mkl(m, n):
    k = "ara_%d_%d" % (m, n)
    Y = G[k + "_Y"].astype(np.float64)
    B = po.ara_syrk(Y)
    assert np.array_equal(B, B.T)                        # both triangles, exactly symmetric
    bound = m * EPS * (np.abs(Y).T @ np.abs(Y))
    assert np.all(np.abs(B - G[k + "_G"]) <= bound)


def test_gram_product_integers_exact():
    rng = np.random.default_rng(3)
    Y = rng.integers(-9, 10, (37, 11)).astype(np.float64)
    assert np.array_equal(po.ara_syrk(Y), Y.T @ Y)
    assert np.array_equal(po.ara_syrk(Y[:, :1]), Y[:, :1].T @ Y[:, :1])


@pytest.mark.parametrize("m,n", CASES)
@pytest.mark.parametrize("bs", [16, 32])
def test_rank_revealing_potrf_full_rank_matches_mkl(m, n, bs):
    k = "ara_%d_%d" % (m, n)
    Gm, L = G[k + "_G"], G[k + "_L"]
    d = np.linspace(0.5, 2.0, bs)
    R, dd, rank = po.ara_potrf(Gm, bs, d)
    c = min(n, bs)
    assert rank == c
    Lc = L[:c, :c]                                       # leading block of the MKL factor
    assert relmax(R, Lc.T) <= 16 * c * EPS
    assert np.all(np.tril(R, -1) == 0.0)
    assert np.allclose(dd[:c], d[:c] * np.diag(Lc), rtol=16 * c * EPS, atol=0)
    assert np.all(dd[c:] == 0.0)


def test_rank_revealing_potrf_reads_only_the_lower_triangle():
    k = "ara_96_17"
    Gm = G[k + "_G"].copy()
    Gm[np.triu_indices(17, 1)] = np.nan
    R, _, rank = po.ara_potrf(Gm, 32, np.ones(32))
    assert rank == 17
    assert relmax(R, G[k + "_L"].T) <= 16 * 17 * EPS


def test_rank_one_integer_gram_stops_after_one_column():
    v = np.array([2.0, 4.0, 6.0, -8.0, 10.0])
    R, d, rank = po.ara_potrf(np.outer(v, v), 16, np.full(16, 3.0))
    assert rank == 1
    assert np.array_equal(R[0], v)                       # L(:,0) = G(:,0) / 2 = v exactly
    assert np.all(R[1:] == 0.0)                          # exact zero Schur complement, zero lower part
    assert d[0] == 6.0 and np.all(d[1:] == 0.0)


def test_zero_pivot_leaves_the_rest_unfactored():
    Gm = np.diag([4.0, 9.0, 0.0, 16.0])
    R, d, rank = po.ara_potrf(Gm, 16, np.ones(16))
    assert rank == 2
    assert np.array_equal(R, np.diag([2.0, 3.0, 0.0, 16.0]))   # 16 is the unfactored Schur complement
    assert np.array_equal(d[:4], [2.0, 3.0, 0.0, 0.0])


@pytest.mark.parametrize("tiny,rank", [(1e-17, 1), (2.2204460492503131e-16, 1), (1e-15, 2)])
def test_tolerance_is_max_diag_times_eps(tiny, rank):
    _, _, r = po.ara_potrf(np.diag([1.0, tiny]), 16, np.ones(16))
    assert r == rank


def test_dim_clamped_to_bs_and_empty():
    k = "ara_250_29"
    Gm, L = G[k + "_G"], G[k + "_L"]
    R, d, rank = po.ara_potrf(Gm, 16, np.ones(16))       # dim 29 > bs 16: leading 16 x 16 block
    assert rank == 16 and R.shape == (16, 16)
    assert relmax(R, L[:16, :16].T) <= 16 * 16 * EPS
    R, d, rank = po.ara_potrf(Gm, 16, np.full(16, 5.0), dim=0)
    assert rank == 0 and np.all(d == 5.0)                # nothing else written


@pytest.mark.parametrize("m,n", CASES)
@pytest.mark.parametrize("bs", [16, 32])
def test_right_trsm_matches_mkl(m, n, bs):
    k = "ara_%d_%d" % (m, n)
    R = G[k + "_L"].T.copy()
    R[np.tril_indices(n, -1)] = np.nan                   # only the upper triangle may be read
    B, Xm = G[k + "_B"], G[k + "_X"]
    X = po.ara_trsm(B, R, bs)
    c = min(n, bs)
    assert relmax(X[:, :c], Xm[:, :c]) <= 64 * c * EPS * np.linalg.cond(G[k + "_L"][:c, :c])
    assert np.array_equal(X[:, c:], B[:, c:])            # columns beyond bs untouched
    Rc = np.triu(G[k + "_L"].T[:c, :c])
    assert relmax(X[:, :c] @ Rc, B[:, :c]) <= 64 * c * EPS


def test_right_trsm_closed_form():
    R = np.array([[2.0, 1.0], [0.0, 4.0]])
    B = np.array([[2.0, 9.0], [4.0, 2.0]])
    # x0 = b0 / 2, x1 = (b1 - 1 * x0) / 4
    assert np.array_equal(po.ara_trsm(B, R, 16), [[1.0, 2.0], [2.0, 0.0]])
