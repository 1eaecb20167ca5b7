"""CPU tests that pin the oracle's mixed-precision restatement (orc_sposv, the
LAPACK dsposv algorithm, and orc_spotrf) against Intel MKL's own dsposv /
spotrf, from tests/golden/lapack_golden_mixed.npz (make_golden.py --mixed).

The inputs are the oracle's seeded Q diag(s) Q^T generator; the fixture keeps
the generator arguments and a checksum of every generated matrix, so each test
first proves it regenerated the same A.  Pins:
  - ITER: for cond(A) <= 1e6 (where FP32 refinement converges) ITER >= 0 on
    both sides and |ITER_oracle - ITER_mkl| <= 1 (the FP32 factors differ in
    rounding, which can move the last step across the threshold); a loop that
    never converged and always fell back (ITER = -31 with a fine FP64 answer),
    or one that skipped refinement, fails here;
  - beyond FP32 (cond 1e7 .. 1e12) both sides must return a solution with
    backward error <= 1e-13 and the same X to cond * eps accuracy; the ITER
    values there are not compared (both sit at kappa * u32 ~ 1, where the
    blocked and unblocked FP32 factors legitimately take different paths);
  - ITER = -2 (an entry of B or of A beyond FP32 range, LAPACK's DLAG2S /
    DLAT2S check), with X equal to the FP64 solution;
  - X against MKL's X within cond(A) * n * eps;
  - orc_spotrf against MKL spotrf within 8 n u32 (relative, max norm).
"""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lapack_golden_mixed.npz"))
META = G["sposv_meta"]
EPS = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def case(k):
    n, nrhs, lo, hi, seed, mid, kappa, it, info, s1, s2 = META[k]
    n, nrhs, seed, mid = int(n), int(nrhs), int(seed), int(mid)
    A, kap = po.gen_spd_cond(seed, mid, n, lo, hi)
    assert np.sum(A) == s1 and np.sum(A * A) == s2, "generator no longer reproduces the fixture's matrix"
    assert kap == kappa
    B = po.gen_rhs(seed + 1, mid, n, nrhs)
    return A, B, kappa, int(it), int(info)


@pytest.mark.parametrize("k", range(len(META)))
def test_sposv_matches_mkl_dsposv(k):
    A, B, kappa, it_mkl, info_mkl = case(k)
    n = A.shape[0]
    X, info, it = po.sposv("L", A, B)
    assert info == info_mkl == 0
    Xm = G["sposv_X_%d" % k]
    assert po.solve_berr("L", A, X, B) <= 1e-13
    rel = np.max(np.abs(X - Xm)) / np.max(np.abs(Xm))
    assert rel <= 4 * n * kappa * EPS, rel
    if kappa <= 1e6 * 1.001:
        assert it >= 0 and it_mkl >= 0, (it, it_mkl)
        assert abs(it - it_mkl) <= 1, (it, it_mkl)
    else:
        assert it != 0 and it_mkl != 0   # far beyond FP32: never converged without refinement


def test_sposv_iter_distribution_matches_mkl():
    """Over the whole cond <= 1e6 family, the iteration counts agree exactly
    on most systems (a systematic shift would be a restatement error)."""
    same = tot = 0
    for k in range(len(META)):
        A, B, kappa, it_mkl, _ = case(k)
        if kappa > 1e6 * 1.001:
            continue
        _, _, it = po.sposv("L", A, B)
        tot += 1
        same += it == it_mkl
    assert tot >= 30 and same >= 0.85 * tot, (same, tot)


@pytest.mark.parametrize("key", ["ovf", "ovfA"])
def test_sposv_fp32_overflow_gives_iter_minus_2(key):
    A = G["sposv_ovfA_A"] if key == "ovfA" else G["sposv_ovf_A"]
    B = G["sposv_ovfA_B"] if key == "ovfA" else G["sposv_ovf_B"]
    X, info, it = po.sposv("L", A, B)
    assert int(G["sposv_%s_iter" % key]) == -2
    assert it == -2 and info == 0
    Xm = G["sposv_%s_X" % key]
    assert np.max(np.abs(X - Xm)) / np.max(np.abs(Xm)) <= 64 * A.shape[0] * EPS * 1e3
    # the answer is the FP64 path's
    _, X64, _ = po.posv("L", A, B)
    assert np.array_equal(X, X64)


def test_sposv_right_side_is_transpose_of_left():
    A, B, kappa, it_mkl, _ = case(20)
    XL, info, itL = po.sposv("L", A, B)
    XR, info, itR = po.sposv("R", A, np.asfortranarray(B.T))
    assert itL == itR
    assert np.array_equal(XR, XL.T)


@pytest.mark.parametrize("n", [16, 33, 64])
def test_spotrf_matches_mkl(n):
    A32 = G["spotrf_A_%d" % n]
    L, info = po.spotrf(A32)
    assert info == 0
    Lm = G["spotrf_L_%d" % n]
    assert np.max(np.abs(np.tril(L) - Lm)) / np.max(np.abs(Lm)) <= 8 * n * EPS32
    iu = np.triu_indices(n, 1)
    assert np.array_equal(L[iu], A32[iu])
