"""CPU tests that pin the oracle's complex restatement (oracle/kblas_oracle.c:
orc_zsqrt, orc_zpotrf, orc_zpotrs, orc_zgemm, orc_zsyrk, orc_ztrsm) before the
GPU tests of the C / Z entry points trust it.

The reference's complex instantiations compute complex SYMMETRIC arithmetic
(include/kblas_operators.h: no conjugation, principal square root), so the
pins are:
  - the square root against Python's cmath.sqrt (principal branch) on every
    quadrant and the negative real axis;
  - A = L0 L0^T built from a lower L0 whose diagonal has positive real part is
    factored back to L0 (the factorization is unique once the square roots are
    principal); upper storage gives U = L0^T;
  - real data reduces to numpy's real Cholesky (LAPACK);
  - solves against numpy.linalg.solve, GEMM / SYRK against numpy products
    (with conj for 'C'), TRSM against scipy's triangular solver (LAPACK
    ztrtrs semantics, including conjugate transposes and unit diagonals);
  - a zero pivot sets info and stops.
"""
import cmath

import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as po

EPS = np.finfo(np.float64).eps


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("v", [4, -4, 3 + 4j, 3 - 4j, -3 + 4j, -3 - 4j, 1j, -1j, -2 + 0j, complex(-2, -0.0),
                               1e-300 + 1e-300j, 7.25 - 1e-9j])
def test_principal_square_root(v):
    got = po.zsqrt(v)
    ref = cmath.sqrt(v)
    # the reference's formula takes Im from sqrt(|v| (1 - cos arg) / 2): accurate to
    # sqrt(eps |v|) when the argument is tiny, to a few ulps elsewhere
    assert abs(got - ref) <= 4 * EPS * abs(ref) + np.sqrt(4 * EPS * abs(v))
    assert got.real >= 0
    assert np.signbit(got.imag) == np.signbit(complex(v).imag)


def test_square_root_formula_of_the_reference():
    # tiny argument: cos(arg) rounds to 1, the imaginary part comes out as -0
    got = po.zsqrt(7.25 - 1e-9j)
    assert got.real == np.sqrt(7.25) and got.imag == 0.0 and np.signbit(got.imag)
    assert po.zsqrt(-4 + 0j) == 2j and po.zsqrt(complex(-4, -0.0)) == -2j


def l0(rng, n):
    L = np.tril(crand(rng, n, n)) * 0.3
    d = rng.uniform(1.0, 2.0, n) + 1j * rng.uniform(-1.0, 1.0, n)   # positive real parts
    L[np.diag_indices(n)] = d
    return L


@pytest.mark.parametrize("n", [1, 2, 5, 16, 33, 70])
@pytest.mark.parametrize("upper", [False, True])
def test_zpotrf_recovers_the_symmetric_factor(n, upper):
    rng = np.random.default_rng(n + 7 * upper)
    L = l0(rng, n)
    A = L @ L.T                          # complex symmetric, no conjugation
    assert np.allclose(A, A.T)
    S = A.copy()
    if upper:
        S[np.tril_indices(n, -1)] = np.nan   # only the upper triangle may be read
    else:
        S[np.triu_indices(n, 1)] = np.nan
    X, info = po.zpotrf(S, upper=upper)
    assert info == 0
    F = np.triu(X).T if upper else np.tril(X)
    assert rel(F, L) <= 64 * n * EPS * np.linalg.cond(L)
    assert rel(F @ F.T, A) <= 16 * n * EPS


def test_zpotrf_real_data_is_the_real_cholesky():
    rng = np.random.default_rng(3)
    B = rng.standard_normal((12, 12))
    A = B @ B.T + 12 * np.eye(12)
    X, info = po.zpotrf(A.astype(np.complex128))
    assert info == 0
    assert np.all(np.tril(X).imag == 0.0)
    assert rel(np.tril(X).real, np.linalg.cholesky(A)) <= 16 * 12 * EPS


def test_zpotrf_zero_pivot_sets_info():
    A = np.diag([4.0 + 0j, 9.0, 0.0, 1.0]).astype(np.complex128)
    X, info = po.zpotrf(A)
    assert info == 3
    assert X[0, 0] == 2 and X[1, 1] == 3 and X[3, 3] == 1   # stopped before column 3


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("upper", [False, True])
def test_zpotrs_solves_the_symmetric_system(side, upper):
    rng = np.random.default_rng(11)
    n, nrhs = 23, 3
    L = l0(rng, n)
    A = L @ L.T
    F = L.T if upper else L
    B = crand(rng, n, nrhs) if side == "L" else crand(rng, nrhs, n)
    X = po.zpotrs(F, B, side=side, upper=upper)
    ref = np.linalg.solve(A, B) if side == "L" else np.linalg.solve(A.T, B.T).T
    assert rel(X, ref) <= 64 * n * EPS * np.linalg.cond(A)


@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
def test_zgemm_matches_numpy(ta, tb):
    rng = np.random.default_rng(5)
    m, n, k = 7, 9, 11
    op = {"N": lambda X: X, "T": lambda X: X.T, "C": lambda X: X.conj().T}
    A = crand(rng, m, k) if ta == "N" else crand(rng, k, m)
    B = crand(rng, k, n) if tb == "N" else crand(rng, n, k)
    C = crand(rng, m, n)
    al, be = 0.5 - 1.25j, 2.0 + 0.5j
    got = po.zgemm(ta, tb, al, A, B, be, C)
    ref = al * (op[ta](A) @ op[tb](B)) + be * C
    assert rel(got, ref) <= 16 * k * EPS
    got0 = po.zgemm(ta, tb, al, A, B, 0.0, np.full((m, n), np.nan + 0j))   # beta 0: C not read
    assert rel(got0, al * (op[ta](A) @ op[tb](B))) <= 16 * k * EPS


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_zsyrk_matches_numpy(upper, trans):
    rng = np.random.default_rng(6)
    n, k = 10, 6
    A = crand(rng, n, k) if trans == "N" else crand(rng, k, n)
    C = crand(rng, n, n)
    al, be = 1.5 + 0.5j, -0.75 + 0.0j
    got = po.zsyrk(upper, trans, al, A, be, C)
    P = A @ A.T if trans == "N" else A.T @ A          # symmetric, no conjugation
    ref = al * P + be * C
    tri = np.triu_indices(n) if upper else np.tril_indices(n)
    other = np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)
    assert rel(got[tri], ref[tri]) <= 16 * k * EPS
    assert np.array_equal(got[other], C[other])        # the other triangle untouched


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("unit", [False, True])
def test_ztrsm_matches_scipy(side, upper, trans, unit):
    rng = np.random.default_rng(9)
    m, n = 13, 6
    k = m if side == "L" else n
    T = crand(rng, k, k) * 0.2 + np.diag(2.0 + rng.uniform(0, 1, k) + 1j * rng.uniform(-1, 1, k))
    T = np.triu(T) if upper else np.tril(T)
    B = crand(rng, m, n)
    al = 0.75 + 0.25j
    got = po.ztrsm(side, upper, trans, unit, al, T, B)
    tflag = {"N": 0, "T": 1, "C": 2}[trans]
    if side == "L":
        ref = sla.solve_triangular(T, al * B, trans=tflag, lower=not upper, unit_diagonal=unit)
    else:
        # X op(T) = al B  <=>  op(T)^T X^T = al B^T
        Tu = T.copy()
        if unit:
            Tu[np.diag_indices(k)] = 1.0
        opT = {"N": Tu, "T": Tu.T, "C": Tu.conj().T}[trans]
        ref = np.linalg.solve(opT.T, (al * B).T).T
    assert rel(got, ref) <= 64 * k * EPS


@pytest.mark.parametrize("n", [1, 4, 17, 40])
@pytest.mark.parametrize("unit", [False, True])
def test_ztrtri_lauum_potri_match_numpy(n, unit):
    rng = np.random.default_rng(40 + n)
    L = l0(rng, n)
    Lu = L.copy()
    if unit:
        Lu[np.diag_indices(n)] = 1.0
    S = L.copy()
    S[np.triu_indices(n, 1)] = np.nan                  # the strict upper triangle is neither read nor written
    X, info = po.ztrtri(S, unit=unit)
    assert info == 0
    assert rel(np.tril(X, -1) if unit else np.tril(X), np.tril(np.linalg.inv(Lu), -1 if unit else 0)) <= 64 * n * EPS
    if unit:
        assert np.array_equal(np.diag(X), np.diag(L))  # unit diagonal left alone
    assert np.all(np.isnan(X[np.triu_indices(n, 1)]))
    Y = po.zlauum(S)
    assert rel(np.tril(Y), np.tril(L.T @ L)) <= 16 * n * EPS       # no conjugation
    Z, info = po.zpotri(S)
    assert info == 0
    assert rel(np.tril(Z), np.tril(np.linalg.inv(L @ L.T))) <= 64 * n * EPS * np.linalg.cond(L) ** 2


def test_ztrtri_zero_diagonal_info():
    L = np.tril(np.ones((5, 5), dtype=np.complex128))
    L[3, 3] = 0.0
    X, info = po.ztrtri(L)
    assert info == 4 and np.array_equal(X, L)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("unit", [False, True])
def test_ztrmm_matches_numpy(side, trans, unit):
    rng = np.random.default_rng(77)
    m, n = 9, 6
    k = m if side == "L" else n
    L = np.tril(crand(rng, k, k))
    Lu = L.copy()
    if unit:
        Lu[np.diag_indices(k)] = 1.0
    opL = {"N": Lu, "T": Lu.T, "C": Lu.conj().T}[trans]
    B = crand(rng, m, n)
    al = 1.5 - 0.5j
    got = po.ztrmm(side, trans, unit, al, L, B)
    ref = al * (opL @ B) if side == "L" else al * (B @ opL)
    assert rel(got, ref) <= 16 * k * EPS
