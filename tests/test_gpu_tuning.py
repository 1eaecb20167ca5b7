"""GPU tests of the per-size-bucket kernel choice: every variant the table
may name, forced for the orders it can serve, gives the oracle's results (the
tuner may pick any of them), and the uniform entry points follow the table
(the same inputs give each forced variant's own bitwise result)."""
import numpy as np
import pytest

import kbtest as kt
import pykblas
import pyoracle as po

pytestmark = pytest.mark.gpu
ORDERS = (1, 5, 16, 17, 31, 32, 33, 40, 48, 57, 64)


def candidates(op, n):
    if op == "sposv":
        return (["reg"] if n <= 32 else []) + ["tile"]
    c = []
    if n <= 16:
        c.append("rg16")
    if n <= 32:
        c.append("rg32")
    if op == "potrf" and 32 < n <= 64:
        c.append("rg4")
    return c + ["tile", "blk"]


@pytest.fixture()
def h():
    hh = kt.handle()
    hh.lib.kblas_tuning_clear()
    yield hh
    hh.lib.kblas_tuning_clear()
    hh.close()


def force(h, prec, op, n, v):
    h.lib.kblas_tuning_clear()
    assert h.lib.kblas_tuning_set(prec.encode(), op.encode(), n, n, v.encode()) == pykblas.KBLAS_Success


def run(h, prec, op, n, A, B):
    """one uniform call; returns (A out, B/X out, info, iter)"""
    dt = np.float64 if prec == "d" else np.float32
    batch = A.shape[0]
    dA = kt.to_dev(A, dt)
    dB = kt.to_dev(B, dt)
    info = kt.to_dev(np.full(batch, -7, np.int32))
    it = kt.to_dev(np.full(batch, -99, np.int32))
    lib = h.lib
    nrhs = B.shape[1]
    P = prec.upper()
    if op == "potrf":
        st = getattr(lib, "kblas%spotrf_batch_strided" % P)(h.h, b"L", n, kt.P(dA), n, n * n, batch, kt.P(info))
    elif op == "posv":
        st = getattr(lib, "kblas%sposv_batch_strided" % P)(h.h, b"L", b"L", n, nrhs, kt.P(dA), n, n * n, kt.P(dB), n,
                                                           n * nrhs, batch, kt.P(info))
    elif op == "potrs":
        st = getattr(lib, "kblas%spotrs_batch_strided" % P)(h.h, b"L", b"L", n, nrhs, kt.P(dA), n, n * n, kt.P(dB), n,
                                                            n * nrhs, batch)
    else:
        dX = kt.to_dev(np.zeros_like(B))
        st = lib.kblasDsposv_batch_strided(h.h, b"L", b"L", n, nrhs, kt.P(dA), n, n * n, kt.P(dB), n, n * nrhs,
                                           kt.P(dX), n, n * nrhs, batch, kt.P(info), kt.P(it))
        dB = dX
    assert st == pykblas.KBLAS_Success, (op, n, st)
    return kt.to_host(dA), kt.to_host(dB), kt.to_host(info), kt.to_host(it)


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("op", ["potrf", "posv", "potrs", "sposv"])
def test_every_variant_matches_oracle(h, prec, op):
    if op == "sposv" and prec == "s":
        pytest.skip("mixed precision is FP64 in / FP64 out")
    eps = kt.EPS64 if prec == "d" else kt.EPS32
    batch = 7
    for n in ORDERS:
        if op == "sposv":
            A, _ = po.gen_spd_cond_batch(300 + n, n, batch, 2.0, 5.0)
        else:
            A = po.gen_spd_shift_batch(300 + n, n, batch)
        B = po.gen_rhs_batch(400 + n, n, 3, batch)              # side L: n x 3
        if op == "potrs":     # the factor (lower triangle; POTRS reads nothing else)
            A_in = np.stack([po.potrf(po.col(A, i))[0].T for i in range(batch)])
        else:
            A_in = A
        for v in candidates(op, n):
            force(h, prec, op, n, v)
            Ao, Bo, info, it = run(h, prec, op, n, A_in, B)
            assert np.all(info[:] == 0) or op == "potrs", (op, n, v, info)
            for i in range(batch):
                Ai = po.col(A, i)
                if op == "potrf":
                    assert po.potrf_berr(Ai, po.col(Ao, i).astype(np.float64)) <= 4 * n * eps, (n, v, i)
                elif op == "sposv":
                    assert it[i] >= 0, (n, v, i, it[i])
                    assert po.solve_berr("L", Ai, po.col(Bo, i), po.col(B, i)) <= 1e-13, (n, v, i)
                else:
                    X = po.col(Bo, i).astype(np.float64)
                    assert po.solve_berr("L", Ai, X, po.col(B, i)) <= 8 * n * eps, (op, n, v, i)


def test_dispatch_follows_the_table(h):
    """the uniform entry point runs the table's variant: results equal that
    variant's own, bit for bit, and differ between variants that round
    differently"""
    n, batch = 30, 64
    A = po.gen_spd_shift_batch(555, n, batch)
    B = po.gen_rhs_batch(556, n, 2, batch)
    outs = {}
    for v in ("rg32", "tile", "blk"):
        force(h, "d", "posv", n, v)
        outs[v] = run(h, "d", "posv", n, A, B)[1]
        again = run(h, "d", "posv", n, A, B)[1]
        assert np.array_equal(outs[v], again), v             # deterministic per variant
    h.lib.kblas_tuning_clear()
    # a table naming each variant in turn for a range around n
    for v in ("rg32", "tile", "blk"):
        h.lib.kblas_tuning_clear()
        assert h.lib.kblas_tuning_set(b"d", b"posv", 17, 32, v.encode()) == pykblas.KBLAS_Success
        got = run(h, "d", "posv", n, A, B)[1]
        assert np.array_equal(got, outs[v]), v
    differ = [(a, b) for a in outs for b in outs if a < b and not np.array_equal(outs[a], outs[b])]
    assert differ, "the variants should not all round alike"
