#!/usr/bin/env python3
"""Benchmark of the batched SPD path on one MI355X.

Default workload (the headline metric): 50,000 SPD 64x64 systems,
A = Q diag(s) Q^T with Q Haar-orthogonal (Stewart: product of Householder
reflectors from Gaussian vectors) and s log-spaced
so cond(A) = kappa, log10(kappa) ~ U[2, 6]; 4 right-hand sides ~ N(0,1) in the
reference's Right-side layout (B is 4 x 64 per system).  One step = one
mixed-precision batched POSV (FP32 factor + FP64 refinement, FP64 fallback)
over the whole batch; the plain FP64 batched POSV on the same inputs is timed
the same way.  value = mixed solves/s.

Inputs are generated on the GPU with torch (plumbing only) and staged W+K
times in HBM, so the timed loop holds exactly K library calls back to back,
bracketed by synchronisation, timed with HIP events on the library's stream.
The CPU baseline (the oracle's OpenMP FP64 POSV restatement, and LAPACK when a
library is loadable) runs on a bounded sample of the same inputs, which also
yields the per-solve backward errors of the GPU results.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kblas-gpu_amd"))

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "c1": dict(op="potrf", n=32, batch=10000, nrhs=0, gen="shift",
               workload="dpotrf_batch_strided n=32 x 10000, A=B*B^T+32I"),
    "c2": dict(op="posv", n=16, batch=100000, nrhs=1, gen="shift",
               workload="dposv_batch_strided n=16 x 100000, 1 RHS, A=B*B^T+16I"),
    "c3": dict(op="vposv", n=256, nmin=8, batch=20000, nrhs=4, gen="shift",
               workload="dposv_vbatch n~U{8..256} x 20000, 4 RHS, pointer arrays, one persistent launch "
                        "over a device size-sorted queue, each step a captured HIP graph"),
    "c5": dict(op="vsposv", n=512, nmin=128, batch=2000, nrhs=64, gen="cond", lkmin=2.0, lkmax=4.0,
               workload="dsposv_vbatch n~U{128..512} x 2000, 64 RHS, cond in [1e2,1e4], non-uniform mixed POSV"),
    "c4": dict(op="sposv", n=64, batch=50000, nrhs=4, gen="cond",
               workload="dsposv_batch_strided n=64 x 50000, 4 RHS, cond in [1e2,1e6], mixed FP32+IR vs FP64"),
    # inverse family (SPD inverse), sized like c1; not a headline configuration
    "poti": dict(op="poti", n=32, batch=10000, nrhs=0, gen="shift", seed=1240,
                 workload="dpoti_batch_strided n=32 x 10000 (A := A^-1, POTRF then POTRI), A=B*B^T+32I"),
}


def gen_inputs(torch, cfg, seed):
    n, bc, nrhs = cfg["n"], cfg["batch"], cfg["nrhs"]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dev = "cuda"
    if cfg["gen"] == "shift":
        Bm = torch.rand((bc, n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1
        A = torch.bmm(Bm, Bm.transpose(1, 2).contiguous())
        A = 0.5 * (A + A.transpose(1, 2))
        A += n * torch.eye(n, device=dev, dtype=torch.float64)
        kappa = None
    else:
        # A = Q diag(s) Q^T, Q = H_0 ... H_{n-2} with H_p a Householder reflector
        # from a Gaussian vector supported on entries p..n-1 (Stewart's
        # construction of a Haar orthogonal matrix), applied two-sidedly to
        # diag(s) in a few hundred batched kernels.
        lk = 2.0 + 4.0 * torch.rand((bc, 1), generator=g, device=dev, dtype=torch.float64)
        t = torch.rand((bc, n), generator=g, device=dev, dtype=torch.float64)
        t[:, 0] = 0.0
        t[:, -1] = 1.0
        s = torch.pow(10.0, -lk * t)
        A = torch.diag_embed(s)
        for p in range(n - 2, -1, -1):
            v = torch.randn((bc, n, 1), generator=g, device=dev, dtype=torch.float64)
            v[:, :p, :] = 0.0
            tau = 2.0 / (v * v).sum(dim=1, keepdim=True)          # (bc,1,1)
            w = tau * torch.bmm(A, v)                              # (bc,n,1)
            alpha = 0.5 * tau * (v * w).sum(dim=1, keepdim=True)
            w = w - alpha * v
            A = A - torch.bmm(v, w.transpose(1, 2).contiguous()) - torch.bmm(w, v.transpose(1, 2).contiguous())
        A = 0.5 * (A + A.transpose(1, 2))
        kappa = torch.pow(10.0, lk[:, 0])
    A = A.contiguous()                       # symmetric: row- and column-major agree
    B = None
    if nrhs:
        # Right side (reference layout): B is nrhs x n, ldb = nrhs, column-major
        B = torch.randn((bc, n, nrhs), generator=g, device=dev, dtype=torch.float64).contiguous()
    return A, B, kappa


def gen_vbatch(torch, cfg, seed):
    """Non-uniform batch: sizes n_i ~ U{nmin..n}, A_i = B_i B_i^T + n_i I packed one
    after another (column-major, lda = n_i), B_i (right side: nrhs x n_i, ldb = nrhs)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    bc, nmax, nmin, nrhs = cfg["batch"], cfg["n"], cfg["nmin"], cfg["nrhs"]
    ns = torch.randint(nmin, nmax + 1, (bc,), generator=g, device="cuda", dtype=torch.int32)
    ns_h = ns.cpu().numpy()
    offA = np.zeros(bc + 1, dtype=np.int64)
    offA[1:] = np.cumsum(ns_h.astype(np.int64) ** 2)
    offB = np.zeros(bc + 1, dtype=np.int64)
    offB[1:] = np.cumsum(ns_h.astype(np.int64) * nrhs)
    Abuf = torch.empty(int(offA[-1]), device="cuda", dtype=torch.float64)
    Bbuf = torch.randn(int(offB[-1]), generator=g, device="cuda", dtype=torch.float64)
    for n in np.unique(ns_h):
        idx = np.nonzero(ns_h == n)[0]
        k = len(idx)
        if cfg["gen"] == "cond":
            # A = Q diag(s) Q^T, Q from the QR factorization of a Gaussian matrix,
            # s log-spaced with endpoints 1 and 1/kappa, log10(kappa) ~ U[lkmin, lkmax]
            G = torch.randn((k, n, n), generator=g, device="cuda", dtype=torch.float64)
            Q, _ = torch.linalg.qr(G)
            lk = cfg["lkmin"] + (cfg["lkmax"] - cfg["lkmin"]) * torch.rand((k, 1), generator=g, device="cuda",
                                                                           dtype=torch.float64)
            t = torch.linspace(0.0, 1.0, int(n), device="cuda", dtype=torch.float64)[None, :]
            sv = torch.pow(10.0, -lk * t)
            A = torch.bmm(Q * sv[:, None, :], Q.transpose(1, 2).contiguous())
            A = 0.5 * (A + A.transpose(1, 2))
        else:
            Bm = torch.rand((k, n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
            A = torch.bmm(Bm, Bm.transpose(1, 2).contiguous())
            A = 0.5 * (A + A.transpose(1, 2)) + n * torch.eye(int(n), device="cuda", dtype=torch.float64)
        pos = torch.from_numpy((offA[idx][:, None] + np.arange(n * n)[None, :]).reshape(-1)).cuda()
        Abuf.index_copy_(0, pos, A.reshape(-1))
    return ns, ns_h, offA, offB, Abuf, Bbuf


def measured_traffic(config):
    """HBM bytes per launch of the workload's library kernel, from the committed
    rocprofv3 counter summary profiles/traffic_<config>.md (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters read for a plain copy of known
    size).  None when no summary exists for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.md" % config)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return {"bytes_per_launch": float(ln.split(":", 1)[1].split()[0]),
                    "source": "profiles/traffic_%s.md" % config}
    return None


def time_steps(torch, call, K, W):
    for i in range(W):
        call(i)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for i in range(K):
        call(W + i)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1000.0, wall


def cpu_baseline(cfg, A_s, B_s, X_gpu, cores):
    """Oracle (FP64 POSV restatement, OpenMP over the sample) on host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    L = po.lib()
    L.orc_set_threads(cores)
    n, nrhs = cfg["n"], cfg["nrhs"]
    bs = A_s.shape[0]
    res = {}
    # accuracy of the GPU results on the sample, against the FP64 oracle metrics
    if X_gpu is not None:
        be = np.array([po.solve_berr("R", A_s[i].T, X_gpu[i].T, B_s[i].T) for i in range(bs)])
        res["berr_max"] = float(be.max())
        res["berr_median"] = float(np.median(be))
    A = np.ascontiguousarray(A_s)
    reps, t_best = 0, None
    t_end = time.perf_counter() + 12.0
    while reps < 5 and (reps == 0 or time.perf_counter() < t_end):
        Aw = A.copy()
        if cfg["op"] == "potrf":
            t0 = time.perf_counter()
            L.orc_potrf_batch(n, po.ptr(Aw), n, n * n, bs, None)
        else:
            Bw = np.ascontiguousarray(B_s).copy()
            t0 = time.perf_counter()
            L.orc_posv_batch(b"R", nrhs, n, po.ptr(Aw), n, n * n, po.ptr(Bw), nrhs, nrhs * n, bs, None)
        dt = time.perf_counter() - t0
        t_best = dt if t_best is None else min(t_best, dt)
        reps += 1
    res["value"] = bs / t_best
    lap = None
    if L.orc_lapack_available():
        Aw = A.copy()
        t0 = time.perf_counter()
        if cfg["op"] == "potrf":
            L.orc_lapack_potrf_batch(n, po.ptr(Aw), n, n * n, bs, None)
        else:
            Bw = np.ascontiguousarray(B_s).copy()
            L.orc_lapack_posv_batch(b"R", nrhs, n, po.ptr(Aw), n, n * n, po.ptr(Bw), nrhs, nrhs * n, bs, None)
        lap = {"value": bs / (time.perf_counter() - t0), "library": L.orc_lapack_name().decode()}
    res["lapack"] = lap
    return res


def cpu_baseline_vbatch(cfg, vb, K, unit):
    """Oracle FP64 POSV (OpenMP over matrices) on the first matrices of the same batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    import ctypes
    L = po.lib()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    L.orc_set_threads(cores)
    ns, ns_h, offA, offB, Abuf, Bbuf = vb
    nrhs = cfg["nrhs"]
    bs = min(2000 if cfg["op"] == "vposv" else 40, cfg["batch"])
    A = Abuf[: int(offA[bs])].cpu().numpy()
    B = Bbuf[: int(offB[bs])].cpu().numpy()
    best = None
    for rep in range(3):
        Aw, Bw = A.copy(), B.copy()
        t0 = time.perf_counter()
        # one OpenMP batch per distinct order keeps the loop inside C
        for n in np.unique(ns_h[:bs]):
            idx = np.nonzero(ns_h[:bs] == n)[0]
            for i in idx:
                L.orc_posv(b"R", nrhs, int(n), Aw[offA[i]:].ctypes.data_as(ctypes.c_void_p), int(n),
                           Bw[offB[i]:].ctypes.data_as(ctypes.c_void_p), nrhs)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {"value": bs / best, "unit": unit, "cores": 1, "kind": "port",
            "sample": "first %d systems of the non-uniform batch, oracle FP64 POSV, one thread" % bs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-sample", type=int, default=0, help="matrices in the CPU-baseline sample (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed region, run a plain device copy of MIB MiB three times "
                         "(the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the batched SPD path shards by matrix, each GPU a contiguous "
              "range of the batch, with no exchange step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import pykblas

    cfg = CONFIGS[args.config]
    n, bc, nrhs = cfg["n"], cfg["batch"], cfg["nrhs"]
    K, W = args.steps, args.warmup
    print("bench: config %s (%s), generating inputs" % (args.config, cfg["workload"]), file=sys.stderr, flush=True)
    vb = None
    if cfg["op"] in ("vposv", "vsposv"):
        vb = gen_vbatch(torch, cfg, cfg.get("seed", 1234 + int(args.config[1])))
        A0, B0, kappa = None, None, None
    else:
        A0, B0, kappa = gen_inputs(torch, cfg, cfg.get("seed") or 1234 + int(args.config[1]))
    torch.cuda.synchronize()
    h = pykblas.Handle(torch.cuda.current_stream().cuda_stream)
    lib = h.lib
    info = torch.zeros(bc, dtype=torch.int32, device="cuda")
    iters = torch.zeros(bc, dtype=torch.int32, device="cuda")
    NA = n * n
    out = {}

    if cfg["op"] == "vsposv":
        # non-uniform mixed POSV: A, B read-only, X written; every step restores A
        # (the full-precision path overwrites it with L), outside the timed region
        ns, ns_h, offA, offB, Abuf0, Bbuf0 = vb
        nrhs_arr = torch.full((bc,), nrhs, dtype=torch.int32, device="cuda")
        lib.kblas_posv_vbatch_wsquery(h.h, n, bc)
        assert lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
        offA_d = torch.from_numpy(offA[:-1] * 8).cuda()
        offB_d = torch.from_numpy(offB[:-1] * 8).cuda()
        As = [Abuf0.clone() for _ in range(K + W)]
        Xbuf = torch.empty_like(Bbuf0)
        pA = [offA_d + a.data_ptr() for a in As]
        pB = offB_d + Bbuf0.data_ptr()
        pX = offB_d + Xbuf.data_ptr()
        torch.cuda.synchronize()

        def call(i):
            st = lib.kblasDsposv_vbatch(h.h, b"R", b"L", nrhs_arr.data_ptr(), ns.data_ptr(), nrhs, n,
                                        pA[i].data_ptr(), ns.data_ptr(), pB.data_ptr(), nrhs_arr.data_ptr(),
                                        pX.data_ptr(), nrhs_arr.data_ptr(), bc, info.data_ptr(), iters.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        t, wall = time_steps(torch, call, K, W)
        nn = ns_h.astype(np.float64)
        per_unit_bytes = float(np.mean(4 * nn * (nn + 1) + 16 * nn * nrhs))
        flops = float(np.sum(nn ** 3 / 3 + nn ** 2 / 2 + nn / 6 + 2 * nrhs * nn ** 2))
        value = bc * K / t
        metric, unit = "solves/s (non-uniform mixed DSPOSV n~U{128..512}, 64 RHS)", "solves/s"
        dtype = "f32+f64"
        X_sample = None
        out["gflops"] = flops * K / t / 1e9
        out["fp64_peak_frac"] = out["gflops"] / 78600.0
        it_h = iters.cpu().numpy()
        hist = {}
        for v in it_h.tolist():
            hist[str(v)] = hist.get(str(v), 0) + 1
        out["iter_hist"] = hist
        out["info_nonzero"] = int((info != 0).sum().item())
        del As
    elif cfg["op"] == "vposv":
        ns, ns_h, offA, offB, Abuf0, Bbuf0 = vb
        nrhs_arr = torch.full((bc,), nrhs, dtype=torch.int32, device="cuda")
        lib.kblas_posv_vbatch_wsquery(h.h, n, bc)
        assert lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
        offA_d = torch.from_numpy(offA[:-1] * 8).cuda()
        offB_d = torch.from_numpy(offB[:-1] * 8).cuda()
        As = [Abuf0.clone() for _ in range(K + W)]
        Bs = [Bbuf0.clone() for _ in range(K + W)]
        pA = [offA_d + a.data_ptr() for a in As]
        pB = [offB_d + b.data_ptr() for b in Bs]
        torch.cuda.synchronize()

        def issue(i):
            st = lib.kblasDposv_vbatch(h.h, b"R", b"L", nrhs_arr.data_ptr(), ns.data_ptr(), nrhs, n,
                                       pA[i].data_ptr(), ns.data_ptr(), pB[i].data_ptr(), nrhs_arr.data_ptr(),
                                       bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        # one HIP graph per step (sort kernels + persistent kernel), captured on torch's stream
        graphs = []
        cap = torch.cuda.Stream()
        for i in range(K + W):
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr, stream=cap):
                h.set_stream(torch.cuda.current_stream().cuda_stream)
                issue(i)
            graphs.append(gr)
        h.set_stream(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()

        def call(i):
            graphs[i].replay()
        t, wall = time_steps(torch, call, K, W)
        nn = ns_h.astype(np.float64)
        per_unit_bytes = float(np.mean(8 * nn * (nn + 1) + 16 * nn * nrhs))
        flops = float(np.sum(nn ** 3 / 3 + nn ** 2 / 2 + nn / 6 + 2 * nrhs * nn ** 2))
        value = bc * K / t
        metric, unit = "solves/s (non-uniform batched DPOSV n~U{8..256}, 4 RHS)", "solves/s"
        dtype = "f64"
        X_sample = None
        out["gflops"] = flops * K / t / 1e9
        out["fp64_peak_frac"] = out["gflops"] / 78600.0
        out["info_nonzero"] = int((info != 0).sum().item())
        del As
    elif cfg["op"] == "potrf":
        As = [A0.clone() for _ in range(K + W)]

        def call(i):
            st = lib.kblasDpotrf_batch_strided(h.h, b"L", n, As[i].data_ptr(), n, NA, bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        t, wall = time_steps(torch, call, K, W)
        per_unit_bytes = 8 * n * (n + 1)
        value = bc * K / t
        metric, unit = "factorizations/s (batched DPOTRF n=%d)" % n, "factorizations/s"
        dtype = "f64"
        X_sample = None
        del As
    elif cfg["op"] == "poti":
        As = [A0.clone() for _ in range(K + W)]

        def call(i):
            st = lib.kblasDpoti_batch_strided(h.h, b"L", n, As[i].data_ptr(), n, NA, bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        t, wall = time_steps(torch, call, K, W)
        assert int((info != 0).sum().item()) == 0
        X_sample = As[0]
        # POTRI alone, on the factors
        Ls = [A0.clone() for _ in range(K + W)]
        for i in range(K + W):
            assert lib.kblasDpotrf_batch_strided(h.h, b"L", n, Ls[i].data_ptr(), n, NA, bc,
                                                 info.data_ptr()) == pykblas.KBLAS_Success

        def call_potri(i):
            st = lib.kblasDpotri_batch_strided(h.h, b"L", n, Ls[i].data_ptr(), n, NA, bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        tp, _ = time_steps(torch, call_potri, K, W)
        out["potri"] = {"value": bc * K / tp, "ms_per_step": 1000 * tp / K,
                        "achieved_GBs": 8 * n * (n + 1) * bc * K / tp / 1e9}
        # TRMM B := L^T B (left, lower, transposed) with the factors and n right-hand sides
        Bt = [torch.randn((bc, n, n), device="cuda", dtype=torch.float64) for _ in range(K + W)]

        def call_trmm(i):
            st = lib.kblasDtrmm_batch_strided(h.h, b"L", b"L", b"T", b"N", n, n, 1.0, Ls[i].data_ptr(), n, NA,
                                              Bt[i].data_ptr(), n, NA, bc)
            assert st == pykblas.KBLAS_Success, st
        tt, _ = time_steps(torch, call_trmm, K, W)
        out["trmm_LLTN"] = {"value": bc * K / tt, "unit": "products/s", "ms_per_step": 1000 * tt / K,
                            "achieved_GBs": (4 * n * (n + 1) + 16 * n * n) * bc * K / tt / 1e9}
        del Bt
        # pivoted Cholesky (PSTRF) of the same SPD inputs
        Ps = [A0.clone() for _ in range(K + W)]
        pv = torch.zeros((bc, n), dtype=torch.int32, device="cuda")
        rk = torch.zeros(bc, dtype=torch.int32, device="cuda")

        def call_pstrf(i):
            st = lib.kblasDpstrf_batch_strided(h.h, n, Ps[i].data_ptr(), n, NA, pv.data_ptr(), n, rk.data_ptr(), bc)
            assert st == pykblas.KBLAS_Success, st
        tq, _ = time_steps(torch, call_pstrf, K, W)
        assert int((rk != n).sum().item()) == 0
        out["pstrf"] = {"value": bc * K / tq, "unit": "factorizations/s", "ms_per_step": 1000 * tq / K,
                        "achieved_GBs": (8 * n * (n + 1) + 4 * n) * bc * K / tq / 1e9}
        del Ps
        per_unit_bytes = 8 * n * (n + 1)
        value = bc * K / t
        metric, unit = "inversions/s (batched DPOTI n=%d)" % n, "inversions/s"
        dtype = "f64"
        del Ls
    elif cfg["op"] == "posv":
        As = [A0.clone() for _ in range(K + W)]
        Bs = [B0.clone() for _ in range(K + W)]

        def call(i):
            st = lib.kblasDposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[i].data_ptr(), n, NA,
                                              Bs[i].data_ptr(), nrhs, nrhs * n, bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        t, wall = time_steps(torch, call, K, W)
        per_unit_bytes = 8 * n * (n + 1) + 16 * n * nrhs
        value = bc * K / t
        metric, unit = "solves/s (batched DPOSV n=%d, %d RHS)" % (n, nrhs), "solves/s"
        dtype = "f64"
        X_sample = Bs[0]
        del As
    else:
        # mixed precision: A and B read-only, X written (A only on fallback)
        As = [A0.clone() for _ in range(K + W)]
        Xs = torch.empty_like(B0)

        def call_mixed(i):
            st = lib.kblasDsposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[i].data_ptr(), n, NA,
                                               B0.data_ptr(), nrhs, nrhs * n, Xs.data_ptr(), nrhs, nrhs * n,
                                               bc, info.data_ptr(), iters.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        print("bench: timing mixed FP32+IR POSV", file=sys.stderr, flush=True)
        t, wall = time_steps(torch, call_mixed, K, W)
        it_h = iters.cpu().numpy()
        X_sample = Xs
        del As
        # plain FP64 POSV on the same inputs
        As = [A0.clone() for _ in range(K + W)]
        Bs = [B0.clone() for _ in range(K + W)]

        def call_fp64(i):
            st = lib.kblasDposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[i].data_ptr(), n, NA,
                                              Bs[i].data_ptr(), nrhs, nrhs * n, bc, info.data_ptr())
            assert st == pykblas.KBLAS_Success, st
        print("bench: timing FP64 POSV", file=sys.stderr, flush=True)
        t64, _ = time_steps(torch, call_fp64, K, W)
        X64 = Bs[0]
        del As
        per_unit_bytes = 4 * n * (n + 1) + 16 * n * nrhs     # A lower read once (FP64), B in, X out
        value = bc * K / t
        metric = "solves/s (batched POSV n=64 mixed FP32+IR, bwd err<=1e-13)"
        unit = "solves/s"
        dtype = "f32+f64"
        out["fp64_posv"] = {"value": bc * K / t64, "ms_per_step": 1000 * t64 / K,
                            "achieved_GBs": (8 * n * (n + 1) + 16 * n * nrhs) * bc * K / t64 / 1e9}
        out["mixed_over_fp64"] = (bc * K / t) / (bc * K / t64)
        hist = {}
        for v in it_h.tolist():
            hist[str(v)] = hist.get(str(v), 0) + 1
        out["iter_hist"] = hist
        out["fallbacks"] = int((it_h < 0).sum())

    ms = 1000.0 * t / K
    achieved = per_unit_bytes * bc / (t / K) / 1e9
    line = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": cfg["workload"], "batch": bc, "n": n, "nrhs": nrhs},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_unit": per_unit_bytes},
    }
    line.update(out)
    tr = measured_traffic(args.config)
    if tr is not None:
        line["roofline"]["traffic"] = tr["bytes_per_launch"]
        line["roofline"]["traffic_source"] = tr["source"]
    torch.cuda.synchronize()
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu and cfg["op"] == "poti":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        bs = min(args.cpu_sample or 4000, bc)
        A_s = A0[:bs].cpu().numpy()
        Xg = X_sample[:bs].cpu().numpy()
        print("bench: CPU baseline on %d matrices, 1 thread" % bs, file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        worst = 0.0
        for i in range(bs):
            L, inf = po.potrf(A_s[i])
            Xo, inf2 = po.potri(L)
            if i < 200:   # accuracy of the GPU inverse on a sub-sample: |X A - I| / (|X| |A|)
                Xi = np.tril(Xg[i].T)
                Xf = Xi + np.tril(Xi, -1).T
                worst = max(worst, float(np.abs(Xf @ A_s[i] - np.eye(n)).max()
                                         / (np.abs(Xf).max() * np.abs(A_s[i]).max() * n)))
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": bs / tc, "unit": unit, "cores": 1, "kind": "port",
                                "sample": "%d of the %d matrices, FP64 oracle POTRF + POTRI restatement, "
                                          "one thread (includes the residual check of 200)" % (bs, bc)}
        line["inverse_residual_max_sample"] = worst
    elif not args.no_cpu and cfg["op"] in ("vposv", "vsposv"):
        line["cpu_baseline"] = cpu_baseline_vbatch(cfg, vb, K, unit)
    elif not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        bs = args.cpu_sample or {"c1": 20000, "c2": 50000, "c4": 2000}[args.config]
        bs = min(bs, bc)
        A_s = A0[:bs].cpu().numpy()
        B_s = B0[:bs].cpu().numpy() if B0 is not None else None
        Xg = X_sample[:bs].cpu().numpy() if X_sample is not None else None
        print("bench: CPU baseline on %d matrices, %d threads" % (bs, cores), file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, A_s, B_s, Xg, cores)
        line["cpu_baseline"] = {"value": cb["value"], "unit": unit, "cores": cores, "kind": "port",
                                "sample": "%d of the %d systems, FP64 oracle POSV/POTRF restatement, OpenMP"
                                          % (bs, bc),
                                "lapack": cb.get("lapack")}
        if "berr_max" in cb:
            line["berr_max_sample"] = cb["berr_max"]
            line["berr_median_sample"] = cb["berr_median"]
    h.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
