"""CPU tests that pin the oracle (oracle/kblas_oracle.c) before it is trusted.

Pins, per restated function:
  - golden vectors from MKL LAPACK/BLAS (tests/golden/lapack_golden.npz):
    POTRF, POSV (side R exactly as the reference test's CPU path, side L as
    dpotrs), all four lower TRSM variants, SYRK N/T, GEMM NN/NT/TN/TT, and
    LAPACK's info on planted non-SPD matrices;
  - closed forms: diagonal, tridiag(-1,2,-1), n = 1, A = L0 L0^T, planted
    solutions;
  - invariants: backward error bound, symmetric/SPD generator outputs with
    the requested condition number (numpy eigvalsh), FLOP model values.
"""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lapack_golden.npz"))
EPS = np.finfo(np.float64).eps
SIZES = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33, 48, 64, 65, 100]


def relmax(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = max(np.max(np.abs(b)), 1e-300)
    return d / s


@pytest.mark.parametrize("n", SIZES)
def test_potrf_matches_lapack(n):
    A = G["potrf_A_%d" % n]
    L, info = po.potrf(A)
    assert info == 0
    assert relmax(np.tril(L), G["potrf_L_%d" % n]) <= 8 * n * EPS
    # strict upper triangle untouched
    iu = np.triu_indices(n, 1)
    assert np.array_equal(L[iu], A[iu])
    assert po.potrf_berr(A, L) <= 4 * n * EPS


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("m", [1, 4])
def test_posv_right_matches_reference_cpu_path(n, m):
    A = G["potrf_A_%d" % n]
    B = G["posvR_B_%d_%d" % (n, m)]
    L, X, info = po.posv("R", A, B)
    assert info == 0
    assert relmax(X, G["posvR_X_%d_%d" % (n, m)]) <= 16 * n * EPS
    assert po.solve_berr("R", A, X, B) <= 4 * n * EPS


@pytest.mark.parametrize("n", SIZES)
def test_posv_left_matches_dpotrs(n):
    A = G["potrf_A_%d" % n]
    B = G["posvL_B_%d" % n]
    L, X, info = po.posv("L", A, B)
    assert info == 0
    assert relmax(X, G["posvL_X_%d" % n]) <= 16 * n * EPS


@pytest.mark.parametrize("n", [5, 16, 24, 40])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_trsm_matches_blas(n, side, trans):
    key = "trsm_%s%s_%d" % (side, trans, n)
    X = po.trsm(side, trans, 0.75, G[key + "_A"], G[key + "_B"])
    assert relmax(X, G[key + "_X"]) <= 16 * n * EPS


@pytest.mark.parametrize("mk", [(8, 8), (16, 32), (33, 17)])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_syrk_matches_blas(mk, trans):
    m, k = mk
    key = "syrk_%s_%d_%d" % (trans, m, k)
    R = po.syrk(trans, -1.0, G[key + "_A"], 0.5, G[key + "_C"])
    il = np.tril_indices(m)
    assert relmax(R[il], G[key + "_R"][il]) <= 4 * k * EPS
    iu = np.triu_indices(m, 1)
    assert np.array_equal(R[iu], G[key + "_C"][iu])   # upper untouched


@pytest.mark.parametrize("mnk", [(4, 32, 32), (32, 16, 16), (7, 9, 11)])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_matches_blas(mnk, ta, tb):
    m, n, k = mnk
    key = "gemm_%s%s_%d_%d_%d" % (ta, tb, m, n, k)
    R = po.gemm(ta, tb, -1.0, G[key + "_A"], G[key + "_B"], 1.0, G[key + "_C"])
    assert relmax(R, G[key + "_R"]) <= 4 * k * EPS


@pytest.mark.parametrize("n", [8, 20, 40])
def test_info_on_non_spd_matches_lapack(n):
    A = G["nonspd_A_%d" % n]
    _, info = po.potrf(A)
    assert info == int(G["nonspd_info_%d" % n])


# ---------------- closed forms ----------------
def test_diagonal_closed_form():
    d = np.array([4.0, 9.0, 2.0, 1e-3, 7.5] * 7)
    L, info = po.potrf(np.diag(d))
    assert info == 0
    assert np.allclose(np.diag(L), np.sqrt(d), rtol=2 * EPS, atol=0)
    assert np.array_equal(np.tril(L, -1), np.zeros_like(L))


@pytest.mark.parametrize("n", [5, 17, 40])
def test_tridiagonal_closed_form(n):
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    L, info = po.potrf(np.asfortranarray(A))
    assert info == 0
    i = np.arange(1, n + 1, dtype=float)
    assert np.allclose(np.diag(L), np.sqrt((i + 1) / i), rtol=8 * EPS)
    assert np.allclose(np.diag(L, -1), -np.sqrt(i[:-1] / (i[:-1] + 1)), rtol=8 * EPS)


def test_n1():
    L, info = po.potrf(np.array([[6.25]]))
    assert info == 0 and L[0, 0] == 2.5
    _, X, info = po.posv("R", np.array([[4.0]]), np.array([[2.0], [8.0]]))
    assert info == 0 and np.array_equal(X[:, 0], [0.5, 2.0])


@pytest.mark.parametrize("n", [12, 33])
def test_known_factor_recovered(n):
    rng = np.random.default_rng(n)
    L0 = np.tril(rng.uniform(-1, 1, (n, n)))
    L0[np.diag_indices(n)] = rng.uniform(1, 2, n)
    A = np.asfortranarray(L0 @ L0.T)
    L, info = po.potrf(A)
    assert info == 0
    assert relmax(np.tril(L), L0) <= 1e-11


@pytest.mark.parametrize("side", ["L", "R"])
def test_planted_solution(side):
    A = po.gen_spd_shift(7, 0, 40)
    rng = np.random.default_rng(3)
    X0 = rng.standard_normal((40, 3)) if side == "L" else rng.standard_normal((3, 40))
    B = A @ X0 if side == "L" else X0 @ A
    _, X, info = po.posv(side, A, B)
    assert info == 0
    assert relmax(X, X0) <= 1e3 * EPS * np.linalg.cond(A)


# ---------------- generators ----------------
def test_gen_spd_shift_is_BBt_plus_shift():
    n = 24
    A = po.gen_spd_shift(11, 5, n)
    B = np.array([[po.lib().orc_uniform(11, 5, i + k * n) for k in range(n)] for i in range(n)])
    assert np.allclose(A, B @ B.T + n * np.eye(n), rtol=0, atol=64 * n * EPS)
    assert np.array_equal(A, A.T)
    assert np.all(np.abs(B) < 1)


@pytest.mark.parametrize("n", [16, 64])
def test_gen_spd_cond_spectrum(n):
    for mid in range(4):
        A, kappa = po.gen_spd_cond(4, mid, n, 2.0, 6.0)
        assert np.array_equal(A, A.T)
        w = np.linalg.eigvalsh(A)
        assert w[0] > 0
        assert 1e2 * 0.999 <= kappa <= 1e6 * 1.001
        assert abs(w[-1] / w[0] / kappa - 1) < 1e-6 * kappa
        assert abs(w[-1] - 1.0) < 1e-12


def test_uniform_and_normal_moments():
    u = np.array([po.lib().orc_uniform(1, 2, k) for k in range(20000)])
    z = np.array([po.lib().orc_normal(1, 2, k) for k in range(20000)])
    assert np.all(np.abs(u) < 1)
    assert abs(u.mean()) < 0.02 and abs(u.var() - 1 / 3) < 0.01
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1) < 0.05
    assert abs(np.mean(z ** 4) - 3) < 0.25


def test_exp2_matches_libm():
    for x in np.linspace(-20, 0, 101):
        assert abs(po.lib().orc_exp2(x) / 2.0 ** x - 1) < 4 * EPS


# ---------------- mixed precision restatement ----------------
def test_sposv_converges_to_fp64_accuracy():
    for mid in range(6):
        A, kappa = po.gen_spd_cond(4, mid, 32, 2.0, 6.0)
        B = po.gen_rhs(44, mid, 32, 4)
        X, info, it = po.sposv("L", A, B)
        assert info == 0
        assert po.solve_berr("L", A, X, B) <= 1e-13
        if it >= 0:
            assert it <= 30


def test_sposv_falls_back_when_fp32_fails():
    # kappa = 1e12: the FP32 factor cannot represent it; FP64 path must answer
    A, _ = po.gen_spd_cond(9, 0, 32, 12.0, 12.0)
    B = po.gen_rhs(9, 0, 32, 2)
    X, info, it = po.sposv("L", A, B)
    assert info == 0 and it < 0
    assert po.solve_berr("L", A, X, B) <= 1e-13


# ---------------- FLOP model (testing/flops.h) ----------------
def test_flop_model_values():
    L = po.lib()
    assert round(L.orc_flops_potrf(32)) == 11440
    assert round(L.orc_flops_posv(b"R", 1, 16)) == 2008
    assert round(L.orc_flops_posv(b"R", 4, 64)) == 122208
    assert round(L.orc_flops_potrs(b"R", 4, 64)) == 2 * 4 * 64 * 64
    assert round(L.orc_flops_syrk(16, 16)) == 16 * 16 * 17


def test_split_rule():
    L = po.lib()
    assert [L.orc_reg_size(n) for n in (1, 2, 3, 16, 24, 32)] == [1, 1, 0, 1, 0, 1]
    assert [L.orc_closest_reg_size(n) for n in (17, 24, 33, 63, 100)] == [16, 16, 32, 32, 64]
