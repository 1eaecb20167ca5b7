#!/usr/bin/env python3
"""Benchmark of the batched SPD (Cholesky) path on one MI355X.

One run times every workload of the path and prints ONE JSON line.  The
headline (the line's top-level metric/value) is the mixed-precision batched
POSV: 50,000 SPD 64x64 systems A = Q diag(s) Q^T (Q Haar-orthogonal, built as
a product of Householder reflectors from Gaussian vectors, s log-spaced with
pinned endpoints so cond(A) = kappa, log10(kappa) ~ U[2, 6]), 4 right-hand
sides ~ N(0,1) in the reference's Right-side layout (B is 4 x 64 per system);
one step = one kblasDsposv_batch_strided call over the whole batch (FP32
factor + FP64 refinement, FP64 fallback), the plain FP64 kblasDposv on the
same inputs timed beside it.  The other workloads go under "configs":
  c1  DPOTRF n=32 x 10,000 (A = B B^T + 32 I)
  c2  DPOSV  n=16 x 100,000, 1 RHS
  c3  DPOSV  non-uniform n ~ U{8..256} x 20,000, 4 RHS, pointer arrays
  c5  DSPOSV non-uniform n ~ U{128..512} x 2,000, 64 RHS, cond in [1e2, 1e4]
  sweep  the reference's published chart shapes: batch 10,240, n = m = 8..256,
         DPOSV (nrhs = n), DTRSM, DSYRK, DGEMM in GFLOP/s (testing/flops.h)

Every step is one library call captured as a HIP graph (torch.cuda.CUDAGraph
on a side stream) and replayed; inputs are generated on the GPU with torch
(plumbing only) and are resident in HBM before timing starts.  c1/c2/c4 stage
W+2K copies of their inputs; c3/c5 restore theirs with a device copy before
each replay.  Each timed replay is queued behind a short spin kernel and
bracketed by HIP events on the replay stream, so the events hold the
library's device time and not the host's graph-launch latency (median_ms,
min_ms, gated_mean_ms).  c1/c2/c4 also capture K further steps back to back
in ONE graph (fresh input slots) and replay it once behind a spin kernel; its
span / K is ms_per_step there (whole-job throughput of a stream of calls,
without the per-replay launch gap of a one-call graph, about 8 us here);
c3/c5 use the gated mean.  roofline.achieved = algorithmic bytes (or flops)
per step / ms_per_step.

cpu_baseline is the reference test's CPU comparator: looped LAPACK (MKL,
sequential per thread) dpotrf + dtrsm(R,L,T) + dtrsm(R,L,N), as in
testing/batch_triangular/test_Xposv_batch.cpp:313-326 (dpotrf alone for
POTRF, test_Xpotrf_batch.cpp:293-305), over a bounded sample of the same
inputs, OpenMP over the batch on every host core, warmed up, median of 5;
a single-thread figure and the oracle's own C restatement are reported
beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config all|c1|c2|c3|c4|c5|poti|sweep]
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
DEV = "cuda"
FP64_PEAK_GFS = 78600.0
FP32_PEAK_GFS = 157300.0
# spin-kernel length (GPU clock cycles, ~0.4 ms) queued before every timed step
GATE_CYCLES = 1000000

CONFIGS = {
    "c1": dict(op="potrf", n=32, batch=10000, nrhs=0, gen="shift", seed=1,
               workload="dpotrf_batch_strided n=32 x 10000, A=B*B^T+32I"),
    "c2": dict(op="posv", n=16, batch=100000, nrhs=1, gen="shift", seed=2,
               workload="dposv_batch_strided n=16 x 100000, 1 RHS, A=B*B^T+16I"),
    "c3": dict(op="vposv", n=256, nmin=8, batch=20000, nrhs=4, gen="shift", seed=3,
               workload="dposv_vbatch n~U{8..256} x 20000, 4 RHS, pointer arrays, one persistent launch "
                        "over a device size-sorted queue"),
    "c4": dict(op="sposv", n=64, batch=50000, nrhs=4, gen="cond", lkmin=2.0, lkmax=6.0, seed=4,
               workload="dsposv_batch_strided n=64 x 50000, 4 RHS, cond in [1e2,1e6], mixed FP32+IR vs FP64"),
    "c5": dict(op="vsposv", n=512, nmin=128, batch=2000, nrhs=64, gen="cond", lkmin=2.0, lkmax=4.0, seed=5,
               workload="dsposv_vbatch n~U{128..512} x 2000, 64 RHS, cond in [1e2,1e4], non-uniform mixed POSV"),
    # inverse family (SPD inverse), sized like c1; not part of the default run
    "poti": dict(op="poti", n=32, batch=10000, nrhs=0, gen="shift", seed=1240,
                 workload="dpoti_batch_strided n=32 x 10000 (A := A^-1, POTRF then POTRI), A=B*B^T+32I"),
    # the shapes of the reference's published batch charts (batch 10240, m = n)
    "sweep": dict(op="sweep", batch=10240, sizes=[8, 16, 32, 64, 128, 256], seed=6,
                  workload="batch 10240, n = m = 8..256: dposv_batch_strided (R, nrhs = n), dtrsm_batch_strided "
                           "(L, L, N, N), dsyrk_batch_strided (L, N), dgemm_batch_strided (N, N)"),
}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# input generation (torch on the GPU: plumbing only)
# ---------------------------------------------------------------------------
def gen_shift(torch, g, bc, n):
    Bm = torch.rand((bc, n, n), generator=g, device=DEV, dtype=torch.float64) * 2 - 1
    A = torch.bmm(Bm, Bm.transpose(1, 2).contiguous())
    A = 0.5 * (A + A.transpose(1, 2))
    A += n * torch.eye(n, device=DEV, dtype=torch.float64)
    return A.contiguous()


def gen_cond(torch, g, ns, nmax, lkmin, lkmax):
    """A_b = Q_b diag(s_b) Q_b^T of order ns[b] in the leading block of an
    nmax x nmax slot: Q_b = H_0 ... H_{n-2}, H_p the Householder reflector of a
    Gaussian vector supported on p..n_b-1 (Stewart's construction of a Haar
    orthogonal matrix, the distribution of the Q factor of a Gaussian matrix),
    s_i = kappa^(-t_i) with t_0 = 0, t_{n-1} = 1 and the rest U(0,1), so
    cond(A_b) = kappa_b, log10(kappa_b) ~ U[lkmin, lkmax].  Slots must be
    sorted by order (ascending) so the matrices still being built at step p
    are a suffix of the batch."""
    bc = ns.shape[0]
    dev = DEV
    lk = lkmin + (lkmax - lkmin) * torch.rand((bc, 1), generator=g, device=dev, dtype=torch.float64)
    idx = torch.arange(nmax, device=dev)
    t = torch.rand((bc, nmax), generator=g, device=dev, dtype=torch.float64)
    t[:, 0] = 0.0
    t = torch.where(idx[None, :] == (ns[:, None] - 1), torch.ones_like(t), t)
    s = torch.pow(10.0, -lk * t)
    s = torch.where(idx[None, :] < ns[:, None], s, torch.zeros_like(s))
    A = torch.diag_embed(s)
    ns_h = ns.cpu().numpy()
    for p in range(nmax - 2, -1, -1):
        first = int(np.searchsorted(ns_h, p + 2))      # slots with n_b > p + 1
        if first >= bc:
            continue
        As = A[first:]
        nb = ns[first:]
        v = torch.randn((bc - first, nmax, 1), generator=g, device=dev, dtype=torch.float64)
        keep = (idx[None, :, None] >= p) & (idx[None, :, None] < nb[:, None, None])
        v = torch.where(keep, v, torch.zeros_like(v))
        vv = (v * v).sum(dim=1, keepdim=True)
        tau = torch.where(vv > 0, 2.0 / torch.clamp(vv, min=1e-300), torch.zeros_like(vv))
        w = tau * torch.bmm(As, v)
        alpha = 0.5 * tau * (v * w).sum(dim=1, keepdim=True)
        w = w - alpha * v
        left = torch.cat((v, w), dim=2).contiguous()                  # (k, nmax, 2)
        right = torch.cat((w, v), dim=2).transpose(1, 2).contiguous()  # (k, 2, nmax)
        As.baddbmm_(left, right, alpha=-1.0)
    A = 0.5 * (A + A.transpose(1, 2))
    return A.contiguous(), torch.pow(10.0, lk[:, 0])


def gen_uniform(torch, cfg):
    g = torch.Generator(device=DEV)
    g.manual_seed(1234 + cfg["seed"])
    n, bc, nrhs = cfg["n"], cfg["batch"], cfg["nrhs"]
    kappa = None
    if cfg["gen"] == "shift":
        A = gen_shift(torch, g, bc, n)
    else:
        ns = torch.full((bc,), n, dtype=torch.int64, device=DEV)
        A, kappa = gen_cond(torch, g, ns, n, cfg["lkmin"], cfg["lkmax"])
    B = None
    if nrhs:
        # right side (reference layout): B is nrhs x n, ldb = nrhs, column-major
        B = torch.randn((bc, n, nrhs), generator=g, device=DEV, dtype=torch.float64).contiguous()
    return A, B, kappa


def gen_vbatch(torch, cfg):
    """Non-uniform batch: n_b ~ U{nmin..n}; A_b packed one after another
    (column-major, lda = n_b), B_b (right side: nrhs x n_b, ldb = nrhs)."""
    g = torch.Generator(device=DEV)
    g.manual_seed(1234 + cfg["seed"])
    bc, nmax, nmin, nrhs = cfg["batch"], cfg["n"], cfg["nmin"], cfg["nrhs"]
    ns = torch.randint(nmin, nmax + 1, (bc,), generator=g, device=DEV, dtype=torch.int32)
    ns_h = ns.cpu().numpy()
    offA = np.zeros(bc + 1, dtype=np.int64)
    offA[1:] = np.cumsum(ns_h.astype(np.int64) ** 2)
    offB = np.zeros(bc + 1, dtype=np.int64)
    offB[1:] = np.cumsum(ns_h.astype(np.int64) * nrhs)
    Abuf = torch.empty(int(offA[-1]), device=DEV, dtype=torch.float64)
    Bbuf = torch.randn(int(offB[-1]), generator=g, device=DEV, dtype=torch.float64)
    kappa = None
    if cfg["gen"] == "cond":
        order = np.argsort(ns_h, kind="stable")
        ns_sorted = torch.from_numpy(ns_h[order].astype(np.int64)).to(DEV)
        Aslots, kap_sorted = gen_cond(torch, g, ns_sorted, nmax, cfg["lkmin"], cfg["lkmax"])
        kappa = np.empty(bc)
        kappa[order] = kap_sorted.cpu().numpy()
        for n in np.unique(ns_h):
            sel = np.nonzero(ns_h[order] == n)[0]          # slots of this order
            dst = order[sel]                                # their batch positions
            blk = Aslots[torch.from_numpy(sel).to(DEV), :n, :n]
            # symmetric: the slot's row-major n x n block is also its column-major one
            pos = torch.from_numpy((offA[dst][:, None] + np.arange(n * n)[None, :]).reshape(-1)).to(DEV)
            Abuf.index_copy_(0, pos, blk.reshape(-1))
        del Aslots
    else:
        for n in np.unique(ns_h):
            idx = np.nonzero(ns_h == n)[0]
            A = gen_shift(torch, g, len(idx), int(n))
            pos = torch.from_numpy((offA[idx][:, None] + np.arange(n * n)[None, :]).reshape(-1)).to(DEV)
            Abuf.index_copy_(0, pos, A.reshape(-1))
    return dict(ns=ns, ns_h=ns_h, offA=offA, offB=offB, A=Abuf, B=Bbuf, kappa=kappa)


_CEILING = {}


def copy_ceiling(torch, lib, total_bytes):
    """Achievable HBM rate for a job that moves total_bytes (read + write):
    the library's streaming copy kernel (16-byte vector accesses) copying
    total_bytes / 2, gated and timed like the workloads, the best grid of a
    few, median of five runs.  GB/s."""
    half = (int(total_bytes) // 2 + 4095) // 4096 * 4096
    if half in _CEILING:
        return _CEILING[half]
    import ctypes
    f = lib.kb_stream_copy
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_void_p]
    src = torch.ones(half // 8, dtype=torch.float64, device=DEV)
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    best = 0.0
    for blocks in (1024, 2048, 4096, 8192):
        ts = []
        for r in range(6):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            torch.cuda._sleep(GATE_CYCLES)
            e0.record(st)
            assert f(dst.data_ptr(), src.data_ptr(), half, blocks, st.cuda_stream) == 0
            e1.record(st)
            torch.cuda.synchronize()
            if r:
                ts.append(e0.elapsed_time(e1) / 1e3)
        best = max(best, 2 * half / float(np.median(ts)) / 1e9)
    assert torch.equal(src[:1024], dst[:1024])
    del src, dst
    _CEILING[half] = best
    return best


def measured_traffic(config):
    """HBM bytes per launch of the workload's library kernel, from the committed
    rocprofv3 counter summary profiles/traffic_<config>.md (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters read for a plain copy of known
    size).  None when no summary exists for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.md" % config)
    if not os.path.exists(path):
        return None, None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":", 1)[1].split()[0]), "profiles/traffic_%s.md" % config
    return None, None


# ---------------------------------------------------------------------------
# timing: one HIP graph per step, HIP events around every replay
# ---------------------------------------------------------------------------
class Stepper:
    """Captures issue(slot) once per input slot as a HIP graph on a side stream
    (the library handle follows the capturing stream) and replays it."""

    def __init__(self, torch, h, issue, slots, chain=None):
        """slots: one graph per input slot 0..slots-1; chain: further slots
        captured back to back into ONE graph (a stream of calls)."""
        self.torch = torch
        self.graphs = []
        self.chain = None
        self.stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        for s in range(slots):
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr, stream=self.stream):
                h.set_stream(torch.cuda.current_stream().cuda_stream)
                issue(s)
            self.graphs.append(gr)
        if chain:
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr, stream=self.stream):
                h.set_stream(torch.cuda.current_stream().cuda_stream)
                for s in chain:
                    issue(s)
            self.chain = (gr, len(chain))
        h.set_stream(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()

    def run(self, K, W, slot_of, restore=None):
        """W warm-up replays, then K timed ones.  Each timed replay is queued
        behind a short spin kernel (torch.cuda._sleep) so the host has queued
        the start event, the graph and the end event before the GPU reaches
        them: the events then bracket the library's own device time, not the
        host's graph-launch latency.  Where the stepper holds a chain (K
        further steps captured back to back in one graph, on fresh input
        slots), it is replayed once behind a spin kernel and its span / K is
        reported beside: the device time per call of a stream of calls,
        without a graph launch per call."""
        torch = self.torch
        st = self.stream
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
        e_all0 = torch.cuda.Event(enable_timing=True)
        e_all1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(st):
            for i in range(W):
                if restore:
                    restore(i)
                self.graphs[slot_of(i)].replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.cuda.stream(st):
            for i in range(K):
                if restore:
                    restore(W + i)
                torch.cuda._sleep(GATE_CYCLES)
                ev[i][0].record(st)
                self.graphs[slot_of(W + i)].replay()
                ev[i][1].record(st)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = np.array([a.elapsed_time(b) for a, b in ev]) / 1000.0
        span = None
        if self.chain is not None:
            gr, kc = self.chain
            with torch.cuda.stream(st):
                torch.cuda._sleep(GATE_CYCLES)
                e_all0.record(st)
                gr.replay()
                e_all1.record(st)
            torch.cuda.synchronize()
            span = e_all0.elapsed_time(e_all1) / 1000.0 / kc * K
        return dict(total_s=float(per.sum()), wall_s=wall,
                    back_to_back_ms_per_step=None if span is None else 1e3 * span / K,
                    median_ms=float(np.median(per) * 1e3), min_ms=float(per.min() * 1e3),
                    mean_ms=float(per.mean() * 1e3))


# ---------------------------------------------------------------------------
# CPU baseline (the reference test's comparator) and the oracle port
# ---------------------------------------------------------------------------
def host_info():
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model


def host_cores():
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return env if env > 0 else aff


def median_rate(fn, units, reps=5):
    fn()                                     # warm-up (MKL initialisation, page faults)
    ts = []
    for _ in range(reps):
        ts.append(fn())
    return units / float(np.median(ts)), [units / t for t in ts]


def cpu_uniform(cfg, A_s, B_s, cores):
    """Looped LAPACK over the sample, all cores and one thread; the oracle port beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    L = po.lib()
    n, nrhs = cfg["n"], cfg["nrhs"]
    bs = A_s.shape[0]
    A = np.ascontiguousarray(A_s)
    B = np.ascontiguousarray(B_s) if B_s is not None else None

    def lap(th, count):
        def f():
            L.orc_set_threads(th)
            Aw = A[:count].copy()
            Bw = B[:count].copy() if B is not None else None
            t0 = time.perf_counter()
            if cfg["op"] == "potrf":
                L.orc_lapack_potrf_batch(n, po.ptr(Aw), n, n * n, count, None)
            else:
                L.orc_lapack_posv_batch(b"R", nrhs, n, po.ptr(Aw), n, n * n, po.ptr(Bw), nrhs, nrhs * n, count,
                                        None)
            return time.perf_counter() - t0
        return f

    def port(th, count):
        def f():
            L.orc_set_threads(th)
            Aw = A[:count].copy()
            Bw = B[:count].copy() if B is not None else None
            t0 = time.perf_counter()
            if cfg["op"] == "potrf":
                L.orc_potrf_batch(n, po.ptr(Aw), n, n * n, count, None)
            else:
                L.orc_posv_batch(b"R", nrhs, n, po.ptr(Aw), n, n * n, po.ptr(Bw), nrhs, nrhs * n, count, None)
            return time.perf_counter() - t0
        return f
    res = {}
    if L.orc_lapack_available():
        v, runs = median_rate(lap(cores, bs), bs)
        one = max(1, bs // max(cores, 1))
        v1, _ = median_rate(lap(1, one), one)
        res = {"value": v, "cores": cores, "kind": "port",
               "what": "looped LAPACK dpotrf%s (%s, sequential per thread), OpenMP over the batch" %
                       ("" if cfg["op"] == "potrf" else " + dtrsm(R,L,T) + dtrsm(R,L,N)",
                        L.orc_lapack_name().decode()),
               "runs": runs, "single_thread": v1, "single_thread_sample": one}
    pv, _ = median_rate(port(cores, min(bs, 4000)), min(bs, 4000), reps=3)
    res["oracle_port"] = {"value": pv, "cores": cores, "sample": min(bs, 4000)}
    if "value" not in res:   # no LAPACK on this host: the port is the baseline
        res.update({"value": pv, "cores": cores, "kind": "port", "what": "oracle C restatement (no LAPACK found)"})
    res["sample"] = "%d of the %d matrices (the first ones of the same seeded batch)" % (bs, cfg["batch"])
    res["cpu_model"] = host_info()
    return res


def cpu_vbatch(cfg, vb, cores, count):
    """Looped LAPACK over the first `count` matrices of a non-uniform batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    L = po.lib()
    nrhs = cfg["nrhs"]
    ns = np.ascontiguousarray(vb["ns_h"][:count].astype(np.int32))
    offA = np.ascontiguousarray(vb["offA"][:count].astype(np.int64))
    offB = np.ascontiguousarray(vb["offB"][:count].astype(np.int64))
    A = vb["A"][: int(vb["offA"][count])].cpu().numpy()
    B = vb["B"][: int(vb["offB"][count])].cpu().numpy()
    # largest first, so the dynamic schedule balances
    order = np.argsort(-ns, kind="stable")
    ns_o, offA_o, offB_o = ns[order].copy(), offA[order].copy(), offB[order].copy()

    def mk(fn, th):
        def f():
            L.orc_set_threads(th)
            Aw, Bw = A.copy(), B.copy()
            t0 = time.perf_counter()
            fn(nrhs, po.ptr(ns_o), po.ptr(Aw), po.ptr(offA_o), po.ptr(Bw), po.ptr(offB_o), count, None)
            return time.perf_counter() - t0
        return f
    res = {}
    if L.orc_lapack_available():
        v, runs = median_rate(mk(L.orc_lapack_posv_vbatch, cores), count, reps=5)
        v1, _ = median_rate(mk(L.orc_lapack_posv_vbatch, 1), count, reps=3)
        res = {"value": v, "cores": cores, "kind": "port",
               "what": "looped LAPACK dpotrf + dtrsm(R,L,T) + dtrsm(R,L,N) (%s, sequential per thread), "
                       "OpenMP dynamic over the matrices" % L.orc_lapack_name().decode(),
               "runs": runs, "single_thread": v1}
    pv, _ = median_rate(mk(L.orc_posv_vbatch, cores), count, reps=3)
    res["oracle_port"] = {"value": pv, "cores": cores}
    if "value" not in res:
        res.update({"value": pv, "cores": cores, "kind": "port", "what": "oracle C restatement (no LAPACK found)"})
    res["sample"] = "first %d of the %d matrices of the same seeded non-uniform batch" % (count, cfg["batch"])
    res["cpu_model"] = host_info()
    return res


def berr_stats(side, A_s, X_s, B_s):
    """normwise backward errors ||b - A x|| / (||A|| ||x|| + ||b||) of a sample (oracle metric)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    be = np.array([po.solve_berr(side, A_s[i].T, X_s[i].T, B_s[i].T) for i in range(A_s.shape[0])])
    return {"max": float(be.max()), "median": float(np.median(be)), "sample": int(A_s.shape[0])}


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def hist_of(a):
    h = {}
    for v in a.tolist():
        h[str(v)] = h.get(str(v), 0) + 1
    return dict(sorted(h.items(), key=lambda kv: int(kv[0])))


def step_seconds(r, K):
    """Seconds for the K timed steps: the span of the K back-to-back replays
    where the workload has one (no restore step between replays), else the
    sum of the gated per-step event times."""
    if r.get("back_to_back_ms_per_step") is not None:
        return r["back_to_back_ms_per_step"] * K / 1e3
    return r["total_s"]


def run_uniform(torch, pykblas, h, name, cfg, K, W, with_cpu, cpu_sample):
    n, bc, nrhs = cfg["n"], cfg["batch"], cfg["nrhs"]
    NA = n * n
    lib = h.lib
    log("%s: %s, generating inputs" % (name, cfg["workload"]))
    A0, B0, kappa = gen_uniform(torch, cfg)
    info = torch.zeros(bc, dtype=torch.int32, device=DEV)
    iters = torch.zeros(bc, dtype=torch.int32, device=DEV)
    S = W + 2 * K          # warm-up, gated and back-to-back slots
    CH = list(range(W + K, S))
    out = {"workload": cfg["workload"], "batch": bc, "n": n, "nrhs": nrhs}
    if cfg["op"] == "potrf":
        As = [A0.clone() for _ in range(S)]

        def issue(s):
            assert lib.kblasDpotrf_batch_strided(h.h, b"L", n, As[s].data_ptr(), n, NA, bc,
                                                 info.data_ptr()) == pykblas.KBLAS_Success
        r = Stepper(torch, h, issue, W + K, chain=CH).run(K, W, lambda i: i)
        assert int((info != 0).sum().item()) == 0
        bytes_alg, bytes_full = 8 * n * (n + 1), 16 * n * n
        unit = "factorizations/s"
        del As
    elif cfg["op"] == "posv":
        As = [A0.clone() for _ in range(S)]
        Bs = [B0.clone() for _ in range(S)]

        def issue(s):
            assert lib.kblasDposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[s].data_ptr(), n, NA,
                                                Bs[s].data_ptr(), nrhs, nrhs * n, bc,
                                                info.data_ptr()) == pykblas.KBLAS_Success
        r = Stepper(torch, h, issue, W + K, chain=CH).run(K, W, lambda i: i)
        assert int((info != 0).sum().item()) == 0
        bytes_alg, bytes_full = 8 * n * (n + 1) + 16 * n * nrhs, 16 * n * n + 16 * n * nrhs
        unit = "solves/s"
        X_sample = Bs[0]
        del As
    else:
        # mixed precision: A and B read-only, X written (A only on fallback)
        As = [A0.clone() for _ in range(S)]
        Xs = torch.empty_like(B0)

        def issue(s):
            assert lib.kblasDsposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[s].data_ptr(), n, NA,
                                                 B0.data_ptr(), nrhs, nrhs * n, Xs.data_ptr(), nrhs, nrhs * n,
                                                 bc, info.data_ptr(), iters.data_ptr()) == pykblas.KBLAS_Success
        log("%s: timing mixed FP32+IR POSV" % name)
        r = Stepper(torch, h, issue, W + K, chain=CH).run(K, W, lambda i: i)
        it_h = iters.cpu().numpy()
        X_mixed = Xs
        del As
        # the plain FP64 POSV on the same inputs
        As = [A0.clone() for _ in range(S)]
        Bs = [B0.clone() for _ in range(S)]

        def issue64(s):
            assert lib.kblasDposv_batch_strided(h.h, b"R", b"L", nrhs, n, As[s].data_ptr(), n, NA,
                                                Bs[s].data_ptr(), nrhs, nrhs * n, bc,
                                                info.data_ptr()) == pykblas.KBLAS_Success
        log("%s: timing FP64 POSV" % name)
        r64 = Stepper(torch, h, issue64, W + K, chain=CH).run(K, W, lambda i: i)
        X64 = Bs[0]
        del As, Bs
        bytes_alg = 4 * n * (n + 1) + 16 * n * nrhs       # A lower read once (FP64), B in, X out
        bytes64 = 8 * n * (n + 1) + 16 * n * nrhs
        unit = "solves/s"
        t64 = step_seconds(r64, K)
        out["fp64_posv"] = {"value": bc * K / t64, "ms_per_step": 1e3 * t64 / K,
                            "median_ms": r64["median_ms"], "min_ms": r64["min_ms"], "gated_mean_ms": r64["mean_ms"],
                            "achieved_GBs": bytes64 * bc * K / t64 / 1e9,
                            "bytes_per_unit": bytes64}
        out["mixed_over_fp64"] = t64 / step_seconds(r, K)
        out["iter_hist"] = hist_of(it_h)
        out["fallbacks"] = int((it_h < 0).sum())
        bytes_full = None
    t = step_seconds(r, K)
    out.update({"value": bc * K / t, "unit": unit, "ms_per_step": 1e3 * t / K, "median_ms": r["median_ms"],
                "min_ms": r["min_ms"], "gated_mean_ms": r["mean_ms"], "graph_replay": True})
    ach = bytes_alg * bc * K / t / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "bytes_per_unit": bytes_alg, "units_per_launch": bc}
    if bytes_full:
        roof["achieved_full_tile"] = bytes_full * bc * K / t / 1e9
        roof["frac_full_tile"] = roof["achieved_full_tile"] / HBM_PEAK_GBS
        roof["bytes_per_unit_full_tile"] = bytes_full
    # what a streaming copy moving the same bytes per step reaches on this GPU
    cc = copy_ceiling(torch, lib, (bytes_full or bytes_alg) * bc)
    roof["copy_ceiling_GBs"] = cc
    roof["frac_of_copy_ceiling"] = (roof.get("achieved_full_tile") or ach) / cc
    tr, src = measured_traffic(name)
    if tr is not None:
        roof["traffic"], roof["traffic_source"] = tr, src
    out["roofline"] = roof
    if cfg["op"] == "sposv":
        # accuracy of both GPU paths on a sample, oracle metrics (FP64)
        ss = min(2000, bc)
        A_s, B_s = A0[:ss].cpu().numpy(), B0[:ss].cpu().numpy()
        out["berr_mixed"] = berr_stats("R", A_s, X_mixed[:ss].cpu().numpy(), B_s)
        out["berr_fp64"] = berr_stats("R", A_s, X64[:ss].cpu().numpy(), B_s)
        out["kappa_range"] = [float(kappa.min()), float(kappa.max())]
    elif cfg["op"] == "posv":
        ss = min(2000, bc)
        out["berr_fp64"] = berr_stats("R", A0[:ss].cpu().numpy(), X_sample[:ss].cpu().numpy(),
                                      B0[:ss].cpu().numpy())
    if with_cpu:
        cores = host_cores()
        bs = min(cpu_sample or {"potrf": 20000, "posv": 50000, "sposv": 10000}[cfg["op"]], bc)
        log("%s: CPU baseline on %d matrices, %d threads" % (name, bs, cores))
        A_s = A0[:bs].cpu().numpy()
        B_s = B0[:bs].cpu().numpy() if B0 is not None else None
        cb = cpu_uniform(cfg, A_s, B_s, cores)
        cb["unit"] = unit
        out["cpu_baseline"] = cb
    return out


def run_vbatch(torch, pykblas, h, name, cfg, K, W, with_cpu):
    bc, nmax, nrhs = cfg["batch"], cfg["n"], cfg["nrhs"]
    lib = h.lib
    log("%s: %s, generating inputs" % (name, cfg["workload"]))
    vb = gen_vbatch(torch, cfg)
    ns, ns_h, offA, offB = vb["ns"], vb["ns_h"], vb["offA"], vb["offB"]
    A0, B0 = vb["A"], vb["B"]
    torch.cuda.synchronize()
    info = torch.zeros(bc, dtype=torch.int32, device=DEV)
    iters = torch.zeros(bc, dtype=torch.int32, device=DEV)
    nrhs_arr = torch.full((bc,), nrhs, dtype=torch.int32, device=DEV)
    mixed = cfg["op"] == "vsposv"
    if mixed:   # the queue and the FP32 slots of the mixed phase for orders above 64
        lib.kblas_sposv_vbatch_wsquery(h.h, nrhs, nmax, bc)
    else:
        lib.kblas_posv_vbatch_wsquery(h.h, nmax, bc)
    assert lib.kblasAllocateWorkspace(h.h) == pykblas.KBLAS_Success
    offA_d = torch.from_numpy(offA[:-1] * 8).to(DEV)
    offB_d = torch.from_numpy(offB[:-1] * 8).to(DEV)
    Aw = A0.clone()
    Bw = B0.clone()
    pA = offA_d + Aw.data_ptr()
    pB = offB_d + Bw.data_ptr()
    if mixed:
        Xw = torch.empty_like(B0)
        pX = offB_d + Xw.data_ptr()

        def issue(s):
            assert lib.kblasDsposv_vbatch(h.h, b"R", b"L", nrhs_arr.data_ptr(), ns.data_ptr(), nrhs, nmax,
                                          pA.data_ptr(), ns.data_ptr(), pB.data_ptr(), nrhs_arr.data_ptr(),
                                          pX.data_ptr(), nrhs_arr.data_ptr(), bc, info.data_ptr(),
                                          iters.data_ptr()) == pykblas.KBLAS_Success

        def restore(i):
            Aw.copy_(A0)          # a fallback overwrites A with its FP64 factor
    else:
        def issue(s):
            assert lib.kblasDposv_vbatch(h.h, b"R", b"L", nrhs_arr.data_ptr(), ns.data_ptr(), nrhs, nmax,
                                         pA.data_ptr(), ns.data_ptr(), pB.data_ptr(), nrhs_arr.data_ptr(),
                                         bc, info.data_ptr()) == pykblas.KBLAS_Success

        def restore(i):
            Aw.copy_(A0)
            Bw.copy_(B0)
    log("%s: capturing and timing" % name)
    r = Stepper(torch, h, issue, 1).run(K, W, lambda i: 0, restore=restore)
    t = r["total_s"]
    r64 = None
    if mixed:
        it_h = iters.cpu().numpy()
        X_mixed = Xw.clone()
        # the FP64 non-uniform POSV on the same inputs, for comparison
        def issue64(s):
            assert lib.kblasDposv_vbatch(h.h, b"R", b"L", nrhs_arr.data_ptr(), ns.data_ptr(), nrhs, nmax,
                                         pA.data_ptr(), ns.data_ptr(), pB.data_ptr(), nrhs_arr.data_ptr(),
                                         bc, info.data_ptr()) == pykblas.KBLAS_Success

        def restore64(i):
            Aw.copy_(A0)
            Bw.copy_(B0)
        log("%s: timing the FP64 POSV beside it" % name)
        r64 = Stepper(torch, h, issue64, 1).run(K, W, lambda i: 0, restore=restore64)
        X64 = Bw.clone()
    nn = ns_h.astype(np.float64)
    flops = float(np.sum(nn ** 3 / 3 + nn ** 2 / 2 + nn / 6 + 2 * nrhs * nn ** 2))
    if mixed:
        bytes_alg = float(np.sum(4 * nn * (nn + 1) + 16 * nn * nrhs))
    else:
        bytes_alg = float(np.sum(8 * nn * (nn + 1) + 16 * nn * nrhs))
    out = {"workload": cfg["workload"], "batch": bc, "n_range": [int(ns_h.min()), int(ns_h.max())], "nrhs": nrhs,
           "value": bc * K / t, "unit": "solves/s", "ms_per_step": 1e3 * t / K, "median_ms": r["median_ms"],
           "min_ms": r["min_ms"], "graph_replay": True,
           "gflops": flops * K / t / 1e9}
    out["fp64_peak_frac"] = out["gflops"] / FP64_PEAK_GFS
    ach = bytes_alg * K / t / 1e9
    roof = {"bound": "fp64" if not mixed else "mixed", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None, "bytes_per_step": bytes_alg,
            "gflops": out["gflops"], "fp64_frac": out["fp64_peak_frac"]}
    tr, src = measured_traffic(name)
    if tr is not None:
        roof["traffic"], roof["traffic_source"] = tr, src
    out["roofline"] = roof
    out["info_nonzero"] = int((info != 0).sum().item())
    if mixed:
        out["iter_hist"] = hist_of(it_h)
        out["fallbacks"] = int((it_h < 0).sum())
        t64 = r64["total_s"]
        out["fp64_posv"] = {"value": bc * K / t64, "ms_per_step": 1e3 * t64 / K, "median_ms": r64["median_ms"],
                            "min_ms": r64["min_ms"], "gflops": flops * K / t64 / 1e9}
        out["mixed_over_fp64"] = t64 / t
        # accuracy per matrix on a sample against the FP64 oracle solution
        X = X_mixed
        ss = min(cfg.get("acc_sample", 64), bc)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        Ah = A0[: int(offA[ss])].cpu().numpy()
        Bh = B0[: int(offB[ss])].cpu().numpy()
        Xh = X[: int(offB[ss])].cpu().numpy()
        be, fe = [], []
        for i in range(ss):
            n = int(ns_h[i])
            Ai = Ah[offA[i]:offA[i + 1]].reshape(n, n).T
            Bi = Bh[offB[i]:offB[i + 1]].reshape(n, nrhs).T
            Xi = Xh[offB[i]:offB[i + 1]].reshape(n, nrhs).T
            be.append(po.solve_berr("R", Ai, Xi, Bi))
            _, Xo, _ = po.posv("R", Ai, Bi)
            fe.append(float(np.max(np.abs(Xi - Xo)) / np.max(np.abs(Xo))))
        out["berr_mixed"] = {"max": float(max(be)), "median": float(np.median(be)), "sample": ss}
        out["fwd_err_vs_fp64_oracle"] = {"max": float(max(fe)), "median": float(np.median(fe)), "sample": ss,
                                         "per_matrix": [float("%.3g" % v) for v in fe]}
        X64h = X64[: int(offB[ss])].cpu().numpy()
        be64 = []
        for i in range(ss):
            n = int(ns_h[i])
            Ai = Ah[offA[i]:offA[i + 1]].reshape(n, n).T
            Bi = Bh[offB[i]:offB[i + 1]].reshape(n, nrhs).T
            be64.append(po.solve_berr("R", Ai, X64h[offB[i]:offB[i + 1]].reshape(n, nrhs).T, Bi))
        out["berr_fp64"] = {"max": float(max(be64)), "median": float(np.median(be64)), "sample": ss}
    if with_cpu:
        cores = host_cores()
        count = min(cfg.get("cpu_count", 2000 if not mixed else 500), bc)
        log("%s: CPU baseline on %d matrices, %d threads" % (name, count, cores))
        cb = cpu_vbatch(cfg, vb, cores, count)
        cb["unit"] = "solves/s"
        out["cpu_baseline"] = cb
    del Aw, Bw, vb
    return out


def run_poti(torch, pykblas, h, cfg, K, W, with_cpu):
    """inverse family, sized like c1 (not part of the default run)"""
    n, bc = cfg["n"], cfg["batch"]
    NA = n * n
    lib = h.lib
    A0, _, _ = gen_uniform(torch, cfg)
    info = torch.zeros(bc, dtype=torch.int32, device=DEV)
    S = W + 2 * K          # warm-up, gated and back-to-back slots
    CH = list(range(W + K, S))
    As = [A0.clone() for _ in range(S)]

    def issue(s):
        assert lib.kblasDpoti_batch_strided(h.h, b"L", n, As[s].data_ptr(), n, NA, bc,
                                            info.data_ptr()) == pykblas.KBLAS_Success
    r = Stepper(torch, h, issue, W + K, chain=CH).run(K, W, lambda i: i)
    assert int((info != 0).sum().item()) == 0
    X_sample = As[0]
    Ls = [A0.clone() for _ in range(S)]
    for i in range(S):
        assert lib.kblasDpotrf_batch_strided(h.h, b"L", n, Ls[i].data_ptr(), n, NA, bc,
                                             info.data_ptr()) == pykblas.KBLAS_Success

    def issue_potri(s):
        assert lib.kblasDpotri_batch_strided(h.h, b"L", n, Ls[s].data_ptr(), n, NA, bc,
                                             info.data_ptr()) == pykblas.KBLAS_Success
    rp = Stepper(torch, h, issue_potri, S).run(K, W, lambda i: i)
    Bt = [torch.randn((bc, n, n), device=DEV, dtype=torch.float64) for _ in range(S)]

    def issue_trmm(s):
        assert lib.kblasDtrmm_batch_strided(h.h, b"L", b"L", b"T", b"N", n, n, 1.0, Ls[s].data_ptr(), n, NA,
                                            Bt[s].data_ptr(), n, NA, bc) == pykblas.KBLAS_Success
    rt = Stepper(torch, h, issue_trmm, S).run(K, W, lambda i: i)
    del Bt
    Ps = [A0.clone() for _ in range(S)]
    pv = torch.zeros((bc, n), dtype=torch.int32, device=DEV)
    rk = torch.zeros(bc, dtype=torch.int32, device=DEV)

    def issue_pstrf(s):
        assert lib.kblasDpstrf_batch_strided(h.h, n, Ps[s].data_ptr(), n, NA, pv.data_ptr(), n, rk.data_ptr(),
                                             bc) == pykblas.KBLAS_Success
    rq = Stepper(torch, h, issue_pstrf, S).run(K, W, lambda i: i)
    assert int((rk != n).sum().item()) == 0
    del Ps, Ls
    t = r["total_s"]
    bytes_alg = 8 * n * (n + 1)
    ach = bytes_alg * bc * K / t / 1e9
    out = {"workload": cfg["workload"], "batch": bc, "n": n, "value": bc * K / t, "unit": "inversions/s",
           "ms_per_step": 1e3 * t / K, "median_ms": r["median_ms"], "min_ms": r["min_ms"],
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None, "bytes_per_unit": bytes_alg},
           "potri": {"value": bc * K / rp["total_s"], "ms_per_step": 1e3 * rp["total_s"] / K,
                     "achieved_GBs": bytes_alg * bc * K / rp["total_s"] / 1e9},
           "trmm_LLTN": {"value": bc * K / rt["total_s"], "unit": "products/s", "ms_per_step": 1e3 * rt["total_s"] / K,
                         "achieved_GBs": (4 * n * (n + 1) + 16 * n * n) * bc * K / rt["total_s"] / 1e9},
           # PSTRF reads and writes the full n x n matrix, plus the pivot vector
           "pstrf": {"value": bc * K / rq["total_s"], "unit": "factorizations/s",
                     "ms_per_step": 1e3 * rq["total_s"] / K,
                     "achieved_GBs": (16 * n * n + 4 * n) * bc * K / rq["total_s"] / 1e9}}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        bs = min(4000, bc)
        A_s = A0[:bs].cpu().numpy()
        Xg = X_sample[:bs].cpu().numpy()
        t0 = time.perf_counter()
        worst = 0.0
        for i in range(bs):
            L, inf = po.potrf(A_s[i])
            Xo, inf2 = po.potri(L)
            if i < 200:
                Xi = np.tril(Xg[i].T)
                Xf = Xi + np.tril(Xi, -1).T
                worst = max(worst, float(np.abs(Xf @ A_s[i] - np.eye(n)).max()
                                         / (np.abs(Xf).max() * np.abs(A_s[i]).max() * n)))
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": bs / tc, "unit": "inversions/s", "cores": 1, "kind": "port",
                               "sample": "%d of the %d matrices, FP64 oracle POTRF + POTRI restatement, one thread "
                                         "(includes the residual check of 200)" % (bs, bc)}
        out["inverse_residual_max_sample"] = worst
    return out


# GFLOP/s read off the reference's published "BATCH DPOSV" / "BATCH DTRSM" charts (docs/KBLAS_handout.png,
# batch 10240, m = n, unlabelled NVIDIA GPU, +-10 %): context for the sweep, not a target
POSTER_GFS = {"dposv": {8: 54, 16: 93, 32: 113, 64: 142, 128: 204, 256: 267},
              "dtrsm": {8: 44, 16: 85, 32: 110, 64: 140, 128: 199, 256: 259}}


def flops_of(op, n):
    """testing/flops.h: POTRF n^3/3 + n^2/2 + n/6; POTRS 2 nrhs n^2 (here nrhs = n);
    TRSM (left, m = n) n m^2; SYRK k n (n + 1); GEMM 2 m n k."""
    potrf = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
    return {"dposv": potrf + 2.0 * n ** 3, "dtrsm": float(n ** 3), "dsyrk": float(n * n * (n + 1)),
            "dgemm": 2.0 * n ** 3}[op]


def run_sweep(torch, pykblas, h, cfg, K, W):
    """The reference's own benchmark shapes: batch 10240, m = n = 8..256,
    DPOSV (side R, nrhs = n), DTRSM (L, L, N, N), DSYRK (L, N) and DGEMM (N, N),
    each step a graph replay behind a spin gate, inputs restored between steps
    (outside the timed region) where the call overwrites them."""
    bc = cfg["batch"]
    lib = h.lib
    g = torch.Generator(device=DEV)
    g.manual_seed(1234 + cfg["seed"])
    info = torch.zeros(bc, dtype=torch.int32, device=DEV)
    res = {}
    for n in cfg["sizes"]:
        NA = n * n
        A0 = gen_shift(torch, g, bc, n)
        B0 = torch.randn((bc, n, n), generator=g, device=DEV, dtype=torch.float64)
        A, B = A0.clone(), B0.clone()
        L = A0.clone()
        assert lib.kblasDpotrf_batch_strided(h.h, b"L", n, L.data_ptr(), n, NA, bc,
                                             info.data_ptr()) == pykblas.KBLAS_Success
        torch.cuda.synchronize()
        row = {}

        def restore_ab(i):
            A.copy_(A0)
            B.copy_(B0)

        def restore_b(i):
            B.copy_(B0)

        legs = {
            "dposv": (lambda s: lib.kblasDposv_batch_strided(h.h, b"R", b"L", n, n, A.data_ptr(), n, NA,
                                                             B.data_ptr(), n, NA, bc, info.data_ptr()),
                      restore_ab),
            "dtrsm": (lambda s: lib.kblasDtrsm_batch_strided(h.h, b"L", b"L", b"N", b"N", n, n, 1.0,
                                                             L.data_ptr(), n, NA, B.data_ptr(), n, NA, bc),
                      restore_b),
            "dsyrk": (lambda s: lib.kblasDsyrk_batch_strided(h.h, b"L", b"N", n, n, -1.0, A0.data_ptr(), n, NA,
                                                             1.0, B.data_ptr(), n, NA, bc), restore_b),
            "dgemm": (lambda s: lib.kblasDgemm_batch_strided(h.h, b"N", b"N", n, n, n, -1.0, A0.data_ptr(), n,
                                                             NA, L.data_ptr(), n, NA, 1.0, B.data_ptr(), n, NA,
                                                             bc), restore_b),
        }
        for op, (call, restore) in legs.items():
            def issue(s, call=call):
                assert call(s) == pykblas.KBLAS_Success
            r = Stepper(torch, h, issue, 1).run(K, W, lambda i: 0, restore=restore)
            if op == "dposv":
                assert int((info != 0).sum().item()) == 0
            t = r["total_s"] / K
            e = {"ms_per_step": 1e3 * t, "median_ms": r["median_ms"], "min_ms": r["min_ms"],
                 "gflops": flops_of(op, n) * bc / t / 1e9}
            if op in POSTER_GFS:
                e["reference_published_gflops"] = POSTER_GFS[op][n]
            row[op] = e
        res[str(n)] = row
        log("sweep n=%d: %s" % (n, ", ".join("%s %.0f GF/s" % (k, v["gflops"]) for k, v in row.items())))
        del A0, B0, A, B, L
        torch.cuda.empty_cache()
    top = res[str(cfg["sizes"][-1])]["dposv"]
    return {"workload": cfg["workload"], "batch": bc, "n": cfg["sizes"][-1], "nrhs": cfg["sizes"][-1],
            "value": top["gflops"], "unit": "GFLOP/s", "ms_per_step": top["ms_per_step"],
            "fp64_peak_frac": top["gflops"] / FP64_PEAK_GFS, "sizes": res,
            "flop_model": "testing/flops.h (POSV = POTRF + 2 TRSM with nrhs = n; TRSM left m^2 n; SYRK k n (n+1); "
                          "GEMM 2mnk)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="all", choices=["all"] + sorted(CONFIGS))
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

    K, W = args.steps, args.warmup
    h = pykblas.Handle(torch.cuda.current_stream().cuda_stream)
    names = ["c4", "c1", "c2", "c3", "c5", "sweep"] if args.config == "all" else [args.config]
    results = {}
    t_start = time.perf_counter()
    for name in names:
        cfg = CONFIGS[name]
        if cfg["op"] in ("vposv", "vsposv"):
            results[name] = run_vbatch(torch, pykblas, h, name, cfg, K, W, not args.no_cpu)
        elif cfg["op"] == "poti":
            results[name] = run_poti(torch, pykblas, h, cfg, K, W, not args.no_cpu)
        elif cfg["op"] == "sweep":
            results[name] = run_sweep(torch, pykblas, h, cfg, K, W)
        else:
            results[name] = run_uniform(torch, pykblas, h, name, cfg, K, W, not args.no_cpu, args.cpu_sample)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log("%s: %.4g %s, %.4f ms/step" % (name, results[name]["value"], results[name]["unit"],
                                            results[name]["ms_per_step"]))
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=DEV)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    h.close()

    head_name = names[0]
    head = results[head_name]
    cfg = CONFIGS[head_name]
    metric = {
        "sposv": "solves/s (batched POSV n=64 mixed FP32+IR, bwd err<=1e-13)",
        "potrf": "factorizations/s (batched DPOTRF n=%d)" % cfg.get("n", 0),
        "posv": "solves/s (batched DPOSV n=%d, %d RHS)" % (cfg.get("n", 0), cfg.get("nrhs", 0)),
        "vposv": "solves/s (non-uniform batched DPOSV n~U{8..256}, 4 RHS)",
        "vsposv": "solves/s (non-uniform mixed DSPOSV n~U{128..512}, 64 RHS)",
        "poti": "inversions/s (batched DPOTI n=%d)" % cfg.get("n", 0),
        "sweep": "GFLOP/s (batched DPOSV n=nrhs=256 x 10240, with the DPOSV/DTRSM/DSYRK/DGEMM sweep n=8..256)",
    }[cfg["op"]]
    line = {
        "metric": metric, "value": head["value"], "unit": head["unit"], "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if cfg["op"] in ("sposv", "vsposv") else "f64",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": cfg["workload"], "batch": cfg["batch"], "n": head.get("n", cfg.get("n")),
                   "nrhs": head.get("nrhs", cfg.get("nrhs", 0))},
    }
    for k, v in head.items():
        if k not in ("workload", "batch", "n", "nrhs", "value", "unit", "ms_per_step"):
            line[k] = v
    if len(names) > 1:
        line["configs"] = {k: results[k] for k in names[1:]}
    line["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line))


if __name__ == "__main__":
    main()
