#!/usr/bin/env python3
"""Autotuner of the batched Cholesky family: measures every kernel variant
that can serve each (precision, operation, order) on the current GPU and
writes the per-size-bucket table the library dispatches from
(kblas-gpu_amd/tuning.json by default; the library loads it on first use,
or $KBLAS_TUNING_FILE).

Variants are the ways one matrix is mapped onto the device (see
include/kblas_batch_ext.h): four matrices per wavefront in registers
(rg16 / rg32), one matrix per 16-lane row in 16-column blocks (rg4, POTRF),
one matrix per wavefront as matrix-core tiles (tile), one workgroup per
matrix with LDS-staged panels (blk); for the mixed POSV the small-order
register kernel (reg) or the tile kernel.  For non-uniform batches the
register/workgroup split ("vsplit") is chosen the same way.

Each candidate is forced with kblas_tuning_set, run on a synthetic batch
(inputs generated on the GPU, A = B B^T + n I), timed with events (median of
the repetitions after a warm-up), and the fastest kept; adjacent orders
with the same winner are merged into one bucket.  A variant replaces the
library's built-in choice only when it is at least 3 % faster.

usage: python kblas-gpu_amd/tune.py [--out PATH] [--batch N] [--reps R] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pykblas  # noqa: E402

OPS = ("potrf", "posv", "potrs", "sposv")
NRHS = {"potrf": 0, "posv": 1, "potrs": 4, "sposv": 4}


def candidates(op, n):
    """variants able to serve order n (the library rejects the others)"""
    if op == "sposv":
        c = ["tile"]
        if n <= 32:
            c.insert(0, "reg")
        return c
    c = []
    if n <= 16:
        c.append("rg16")
    if n <= 32:
        c.append("rg32")
    if op == "potrf" and 32 < n <= 64:
        c.append("rg4")
    c += ["tile", "blk"]
    return c


# a measured variant replaces the library's built-in choice only when it is at
# least this much faster (run-to-run noise is a few per cent)
MARGIN = 0.97


def builtin(op, n):
    """the library's built-in choice (kb_reg.hip) for a uniform batch"""
    if op == "sposv":
        return "reg" if n <= 32 else "tile"
    if n <= 16:
        return "rg16"
    if n <= 32:
        return "rg32"
    if op == "potrf" and n <= 48:
        return "rg4"
    return "tile"


def pick(times, default):
    best = min(times, key=times.get)
    return best if times[best] < MARGIN * times[default] else default


def gen_spd(torch, g, bc, n, dtype):
    X = torch.rand((bc, n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    A = torch.bmm(X, X.transpose(1, 2).contiguous())
    A += n * torch.eye(n, device="cuda", dtype=torch.float64)
    return A.to(dtype).contiguous()


def time_call(torch, fn, restore, reps):
    ts = []
    for r in range(reps + 1):
        restore()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(200000)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        if r:
            ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def tune_uniform(torch, h, prec, op, ns, batch, reps, log):
    lib = h.lib
    dtype = torch.float64 if prec == "d" or op == "sposv" else torch.float32
    g = torch.Generator(device="cuda")
    rows = []
    for n in ns:
        g.manual_seed(1000 + n)
        A0 = gen_spd(torch, g, batch, n, dtype)
        A = A0.clone()
        nrhs = NRHS[op]
        B0 = torch.randn((batch, n, max(nrhs, 1)), generator=g, device="cuda", dtype=torch.float64).to(dtype)
        B = B0.clone()
        X = torch.empty_like(B0)
        info = torch.zeros(batch, dtype=torch.int32, device="cuda")
        it = torch.zeros(batch, dtype=torch.int32, device="cuda")
        if op == "potrs":
            # a factor to solve with
            lib.kblas_tuning_clear()
            call = lib.kblasDpotrf_batch_strided if prec == "d" else lib.kblasSpotrf_batch_strided
            assert call(h.h, b"L", n, A0.data_ptr(), n, n * n, batch, info.data_ptr()) == pykblas.KBLAS_Success
            torch.cuda.synchronize()

        def restore():
            A.copy_(A0)
            B.copy_(B0)

        def run():
            if op == "potrf":
                f = lib.kblasDpotrf_batch_strided if prec == "d" else lib.kblasSpotrf_batch_strided
                st = f(h.h, b"L", n, A.data_ptr(), n, n * n, batch, info.data_ptr())
            elif op == "posv":
                f = lib.kblasDposv_batch_strided if prec == "d" else lib.kblasSposv_batch_strided
                st = f(h.h, b"L", b"L", n, nrhs, A.data_ptr(), n, n * n, B.data_ptr(), n, n * nrhs, batch,
                       info.data_ptr())
            elif op == "potrs":
                f = lib.kblasDpotrs_batch_strided if prec == "d" else lib.kblasSpotrs_batch_strided
                st = f(h.h, b"L", b"L", n, nrhs, A.data_ptr(), n, n * n, B.data_ptr(), n, n * nrhs, batch)
            else:
                st = lib.kblasDsposv_batch_strided(h.h, b"L", b"L", n, nrhs, A.data_ptr(), n, n * n, B.data_ptr(),
                                                   n, n * nrhs, X.data_ptr(), n, n * nrhs, batch, info.data_ptr(),
                                                   it.data_ptr())
            assert st == pykblas.KBLAS_Success, st

        times = {}
        for v in candidates(op, n):
            lib.kblas_tuning_clear()
            assert lib.kblas_tuning_set(prec.encode(), op.encode(), n, n, v.encode()) == pykblas.KBLAS_Success
            times[v] = time_call(torch, run, restore, reps)
        lib.kblas_tuning_clear()
        best = pick(times, builtin(op, n))
        rows.append((n, best, times))
        log("%s %-5s n=%3d  %s  -> %s" % (prec, op, n, "  ".join("%s %.4f" % kv for kv in times.items()), best))
        del A, A0, B, B0, X
    return rows


def tune_vsplit(torch, h, prec, batch, reps, log):
    """non-uniform POSV over orders U{8..256}, 4 right-hand sides: the order
    at and below which a matrix is served by one wavefront"""
    lib = h.lib
    dtype = torch.float64 if prec == "d" else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    ns = torch.randint(8, 257, (batch,), generator=g, device="cuda", dtype=torch.int32)
    ns_h = ns.cpu().numpy()
    off = np.zeros(batch + 1, np.int64)
    off[1:] = np.cumsum(ns_h.astype(np.int64) ** 2)
    offb = np.zeros(batch + 1, np.int64)
    offb[1:] = np.cumsum(ns_h.astype(np.int64) * 4)
    A0 = torch.empty(int(off[-1]), device="cuda", dtype=dtype)
    for n in np.unique(ns_h):
        idx = np.nonzero(ns_h == n)[0]
        blk = gen_spd(torch, g, len(idx), int(n), dtype)
        pos = torch.from_numpy((off[idx][:, None] + np.arange(n * n)[None, :]).reshape(-1)).cuda()
        A0.index_copy_(0, pos, blk.reshape(-1))
    B0 = torch.randn(int(offb[-1]), generator=g, device="cuda", dtype=torch.float64).to(dtype)
    A, B = A0.clone(), B0.clone()
    esz = A.element_size()
    pA = torch.from_numpy(off[:-1] * esz).cuda() + A.data_ptr()
    pB = torch.from_numpy(offb[:-1] * esz).cuda() + B.data_ptr()
    r4 = torch.full((batch,), 4, dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    lib.kblas_posv_vbatch_wsquery(h.h, 256, batch)
    assert lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
    f = lib.kblasDposv_vbatch if prec == "d" else lib.kblasSposv_vbatch

    def restore():
        A.copy_(A0)
        B.copy_(B0)

    def run():
        assert f(h.h, b"R", b"L", r4.data_ptr(), ns.data_ptr(), 4, 256, pA.data_ptr(), ns.data_ptr(), pB.data_ptr(),
                 r4.data_ptr(), batch, info.data_ptr()) == pykblas.KBLAS_Success

    times = {}
    for split in (0, 16, 32, 48, 64):
        lib.kblas_tuning_clear()
        assert lib.kblas_tuning_set(prec.encode(), b"vsplit", 0, 1 << 30, str(split).encode()) == pykblas.KBLAS_Success
        times[split] = time_call(torch, run, restore, reps)
    lib.kblas_tuning_clear()
    best = pick(times, 64)
    log("%s vsplit  %s  -> %d" % (prec, "  ".join("%d %.4f" % kv for kv in times.items()), best))
    return best, times


def buckets(rows):
    """merge adjacent orders with the same winner"""
    out = []
    for n, v, t in rows:
        if out and out[-1]["variant"] == v and out[-1]["n_hi"] + 1 >= n:
            out[-1]["n_hi"] = n
            out[-1]["ms"][str(n)] = t
        else:
            out.append({"n_lo": n, "n_hi": n, "variant": v, "ms": {str(n): t}})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--batch", type=int, default=20000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="orders 4, 8, ..., 64 only (filled in between)")
    args = ap.parse_args()
    import torch
    assert torch.cuda.is_available(), "the tuner measures on the GPU"
    log = lambda m: print("tune: " + m, file=sys.stderr, flush=True)  # noqa: E731
    h = pykblas.Handle(torch.cuda.current_stream().cuda_stream)
    t0 = time.time()
    ns = list(range(4, 65, 4)) if args.quick else list(range(1, 65))
    entries = []
    for prec in ("d", "s"):
        for op in OPS:
            if op == "sposv" and prec == "s":
                continue
            rows = tune_uniform(torch, h, prec, op, ns, args.batch, args.reps, log)
            if args.quick:   # each measured order stands for the orders up to it
                filled = []
                lo = 1
                for n, v, t in rows:
                    for m in range(lo, n + 1):
                        if v in candidates(op, m):
                            filled.append((m, v, t))
                        else:
                            filled.append((m, candidates(op, m)[0], t))
                    lo = n + 1
                rows = filled
            for b in buckets(rows):
                entries.append(dict(prec=prec, op=op, **b))
        best, times = tune_vsplit(torch, h, prec, args.batch, args.reps, log)
        entries.append(dict(prec=prec, op="vsplit", n_lo=0, n_hi=1 << 30, variant=best,
                            ms={str(k): v for k, v in times.items()}))
    h.close()
    props = torch.cuda.get_device_properties(0)
    table = {"device": props.name, "gcn_arch": getattr(props, "gcnArchName", ""),
             "generated_by": "kblas-gpu_amd/tune.py", "batch": args.batch,
             "seconds": round(time.time() - t0, 1), "entries": entries}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log("wrote %s (%d entries)" % (args.out, len(entries)))


if __name__ == "__main__":
    main()
