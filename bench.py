#!/usr/bin/env python3
"""Benchmark: unbalanced variable-size FP32 batched GEMM on one MI355X.

Workload (one "step" = one batched launch over the whole batch, inputs resident
in HBM): B = 2048 column-major GEMMs, 90% with M,N,K ~ U{16..128} and 10% with
M,N,K ~ U{512..2048}; A, B ~ N(0,1) (seed 0); C ~ U[-1,1]; alpha = 1.5,
beta = 0.5.  Metric: 2*sum(MNK) / t in GFLOP/s.

Prints ONE JSON line (rank 0).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PEAK = {"fp32": 157.3e12, "bf16x3": 2.5e15 / 3, "fp16": 2.5e15, "fp8": 5.0e15, "bf16": 2.5e15}
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps  # ms


def time_synced(fn, steps, warmup):
    """Mean wall ms of fn() followed by a host synchronisation, per call: the
    reference's mode of one launch then hipDeviceSynchronize
    (ROCm/framework/gemm.cuh:246-249), launch and sync latency included."""
    import torch
    for _ in range(warmup):
        fn()
        torch.cuda.synchronize()
    t = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        t.append(time.perf_counter() - t0)
    return 1e3 * float(np.mean(t))


KERNEL_TEMPLATE = {1: "Fp32MfmaTile", 2: "SeqTile", 3: "SplitTile", 4: "Bf16Tile", 7: "SplitTile"}


def kernel_name(bg, prec, kind, tile):
    """Readable name of a plan kernel: family template and tile shape (the
    symbols rocprof lists are vbg_gemm_kernel / vbg_split_kernel /
    vbg_bf16_kernel instantiated on these templates)."""
    if kind == 0:
        return "vbg_amax_kernel+vbg_scales_kernel"
    if kind == 5:
        return "vbg_splitk_reduce"
    if kind == 6:
        return f"vbg_pack_kernel[{prec}]"
    tag = {3: f"[{prec}]", 7: f"[{prec},packed]"}.get(kind, "")
    # packed FP16 / FP8 and the BF16 family run the shared LDS-DMA tiles
    dma = kind == 4 or (kind == 7 and prec in ("fp16", "fp8"))
    if tile == bg._lib.TILE_MIXED:  # one launch over the 64x64 .. 128x128 shapes
        if dma:
            return f"DmaTileMixed{tag}"
        return {1: "Fp32Mixed", 3: f"SplitMixed{tag}", 7: f"SplitMixed{tag}"}.get(
            kind, f"{KERNEL_TEMPLATE[kind]}<mixed>")
    bm, bn = bg.TILE_SHAPES[tile]
    if kind == 1 and bm == bn and bm <= 32:
        return f"Fp32SmallTile<{bm}>"
    return f"{'DmaTile' if dma else KERNEL_TEMPLATE[kind]}<{bm},{bn}>" + tag


TRAFFIC_FILE = os.path.join("profiles", "traffic_unbalanced_b2048.txt")


def source_digest():
    """sha256 over the library's sources (batchgemm_amd/csrc/*, its Makefile
    and include/vbgemm.h, in name order): names the code a profile was
    measured on, so a figure measured on other code is not reported as this
    code's."""
    import hashlib
    root = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(root, "batchgemm_amd", "csrc")
    files = [os.path.join(csrc, n) for n in sorted(os.listdir(csrc))
             if n.endswith((".hip", ".cpp", ".h"))]
    files += [os.path.join(root, "batchgemm_amd", "Makefile"), os.path.join(root, "include", "vbgemm.h")]
    h = hashlib.sha256()
    for path in files:
        h.update(os.path.basename(path).encode() + b"\0")
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE, calibrated on a copy kernel), from the lines
    `<kernel> <bytes>` of TRAFFIC_FILE, whose header names the commit and the
    source digest it was measured on.  Returns (bytes or None, commit or
    None): bytes only when that kernel was measured on these sources."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_FILE)
    digest = commit = None
    found = None
    try:
        with open(path) as f:
            for ln in f:
                parts = ln.split()
                if ln.startswith("# source-digest:") and len(parts) >= 3:
                    digest = parts[2]
                elif ln.startswith("# commit:") and len(parts) >= 3:
                    commit = parts[2]
                elif len(parts) == 2 and parts[0] == kernel and not ln.startswith("#"):
                    found = int(float(parts[1]))
    except OSError:
        return None, None
    if found is None or digest != source_digest():
        return None, commit
    return found, commit


def per_matrix_library(vb, alpha, beta, steps, warmup):
    """The per-matrix-launch comparator: one torch.addmm (rocBLAS/hipBLASLt
    FP32 GEMM) per matrix, as the reference's hipBLAS loop
    (ROCm/default/gemm.cuh:64-85), on the same column-major buffers."""
    from batchgemm_amd.baseline import library_loop
    return time_launches(library_loop(vb, alpha, beta), steps, warmup)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_sample_rate(orc, M, N, K, alpha, beta, budget_s, nthreads, seed):
    """Host per-matrix SGEMM (the oracle's blocked loop) over matrices of the
    workload in random order (keeping its size mix), until about `budget_s`
    seconds of CPU work: (GFLOP/s, matrices, GFLOP, seconds)."""
    rng = np.random.default_rng(seed)
    order = np.concatenate([rng.permutation(len(M)) for _ in range(4)])  # at most 4 passes
    done_flops, done_t, used = 0, 0.0, 0
    chunk = 64 if nthreads != 1 else 8
    while done_t < budget_s and used < len(order):
        idx = order[used: used + chunk]
        used += len(idx)
        As = [rng.standard_normal(int(M[i]) * int(K[i])).astype(np.float32) for i in idx]
        Bs = [rng.standard_normal(int(K[i]) * int(N[i])).astype(np.float32) for i in idx]
        Cs = [rng.uniform(-1, 1, int(M[i]) * int(N[i])).astype(np.float32) for i in idx]
        m, n, k = M[idx], N[idx], K[idx]
        t0 = time.perf_counter()
        orc.sgemm_vbatched("N", "N", m, n, k, alpha, As, m, Bs, k, beta, Cs, m, nthreads=nthreads)
        done_t += time.perf_counter() - t0
        done_flops += int(2 * np.sum(m.astype(np.int64) * n * k))
    return done_flops / done_t / 1e9, used, done_flops / 1e9, done_t


def cpu_baseline(M, N, K, alpha, beta, budget_s=12.0):
    """Host per-matrix SGEMM loop (oracle/liboracle.so, OpenMP across matrices)
    on a bounded sample of the same workload: matrices in random order (keeping
    the 90/10 mix) until about `budget_s` seconds of CPU work on every thread,
    then a quarter of that on one thread."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc
    rate, used, gf, secs = cpu_sample_rate(orc, M, N, K, alpha, beta, budget_s, 0, 123)
    rate1, used1, gf1, secs1 = cpu_sample_rate(orc, M, N, K, alpha, beta, budget_s / 4, 1, 124)
    return {
        "value": round(rate, 2), "unit": "GFLOP/s",
        "cores": orc.num_threads(), "kind": "port", "cpu_model": cpu_model(),
        "single_thread_value": round(rate1, 2),
        "sample": f"{used} matrices ({used / len(M):.2f} passes over the {len(M)}) of the same "
                  f"workload in random order ({gf:.1f} GFLOP, {secs:.1f} s): host per-matrix "
                  f"FP32 SGEMM (the oracle's cache-blocked loop), OpenMP across matrices; "
                  f"single thread: {used1} matrices, {gf1:.1f} GFLOP, {secs1:.1f} s",
    }


def rel_err_sample(w, vb, M, N, K, alpha, beta, C0, idx):
    """Relative Frobenius error of C against FP64 (torch float64 on the GPU)
    over sampled matrices, pooled."""
    import torch
    num = den = 0.0
    for b in idx:
        A = vb.a(b).double(); B = vb.b(b).double()
        m, n = int(M[b]), int(N[b])
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + m * n].view(n, m).t().double()
        ref = alpha * (A @ B) + beta * c0
        num += float(torch.linalg.norm(vb.c(b).double() - ref) ** 2)
        den += float(torch.linalg.norm(ref) ** 2)
    return (num / den) ** 0.5


def extras(bg, w, vb, M, N, K, alpha, beta, flops, fp32_value, args):
    """bf16x3 on the same batch (GFLOP/s, ratio to FP32, error vs FP64 on a
    sample); all four precision paths on outlier-stressed inputs of the same
    shapes; the per-matrix library loop; the grouped BF16 forward+backward
    (G=128, K=2048, N=1408, Zipf(1.2) group sizes summing to 32768)."""
    import torch
    out = {}
    rng = np.random.default_rng(1)
    big = np.where(M >= 512)[0]
    idx = list(rng.choice(big, 4, replace=False)) + list(rng.choice(np.where(M < 512)[0], 28, replace=False))
    for prec in ("bf16x3", "fp32"):
        p = bg.Plan(M, N, K, prec=prec)
        ms = time_launches(lambda: p.run(vb.dA, vb.dB, vb.dC, alpha, beta), args.steps, args.warmup)
        C0 = vb.C.clone()
        p.run(vb.dA, vb.dB, vb.dC, alpha, beta)
        torch.cuda.synchronize()
        err = rel_err_sample(w, vb, M, N, K, alpha, beta, C0, idx)
        out[prec] = {"gflops": round(flops / (ms * 1e-3) / 1e9, 1), "ms": round(ms, 4),
                     "rel_err_vs_fp64": float(f"{err:.3g}")}
        # one batch at a time, host-synchronised after each (the reference's mode)
        ms_sync = time_synced(lambda: p.run(vb.dA, vb.dB, vb.dC, alpha, beta), args.steps, args.warmup)
        out[prec]["sync_per_batch_ms"] = round(ms_sync, 4)
        out[prec]["sync_per_batch_gflops"] = round(flops / (ms_sync * 1e-3) / 1e9, 1)
        log(f"[bench] {prec}: {ms:.3f} ms, {out[prec]['gflops']} GFLOP/s, rel err {err:.3g}; "
            f"synchronised per batch {ms_sync:.3f} ms")
    out["bf16x3"]["speedup_vs_fp32"] = round(out["bf16x3"]["gflops"] / out["fp32"]["gflops"], 3)
    # Precision paths on the same shapes with N(0,1) inputs and 1% of the
    # entries scaled by 1e3 (stresses the per-matrix scaling); error of the
    # sampled matrices against FP64.
    vb.fill("normal", seed=args.seed + 1, outlier_frac=0.01, outlier_scale=1e3, c_dist="uniform")
    # The split precisions run both ways: operands quantised once per run into
    # a packed workspace ("packed", FP16 / FP8 default on this batch), and
    # quantised on the HBM -> LDS path ("in_loop"); the two give the same bits.
    prec_out = {}
    for prec in ("fp32", "bf16x3", "fp16", "fp8"):
        variants = [(prec, {})] if prec == "fp32" else [
            (f"{prec}_packed", {"pack_operands": 1}), (f"{prec}_in_loop", {"pack_operands": -1})]
        for key, o in variants:
            p = bg.Plan(M, N, K, prec=prec, **o)
            ms = time_launches(lambda: p.run(vb.dA, vb.dB, vb.dC, alpha, beta), max(3, args.steps // 2),
                               args.warmup)
            C0 = vb.C.clone()
            p.run(vb.dA, vb.dB, vb.dC, alpha, beta)
            torch.cuda.synchronize()
            err = rel_err_sample(w, vb, M, N, K, alpha, beta, C0, idx)
            prec_out[key] = {"gflops": round(flops / (ms * 1e-3) / 1e9, 1), "ms": round(ms, 4),
                             "rel_err_vs_fp64": float(f"{err:.3g}")}
            log(f"[bench] outlier inputs {key}: {ms:.3f} ms, {prec_out[key]['gflops']} GFLOP/s, "
                f"rel err {err:.3g}")
            del p
        if prec != "fp32":
            packed_default = bg.Plan(M, N, K, prec=prec, host_only=True).packed_bytes > 0
            prec_out[f"{prec}_default"] = "packed" if packed_default else "in_loop"
    out["precision_outliers"] = {"inputs": "N(0,1), 1% of entries x1e3, C ~ U[-1,1]", **prec_out}
    lib_ms = per_matrix_library(vb, alpha, beta, max(3, args.steps // 4), 1)
    out["per_matrix_library_fp32"] = {
        "what": "torch.addmm per matrix (rocBLAS/hipBLASLt), 2048 launches",
        "ms": round(lib_ms, 3), "gflops": round(flops / (lib_ms * 1e-3) / 1e9, 1)}
    log(f"[bench] per-matrix library loop: {lib_ms:.3f} ms")
    # grouped forward + backward
    from batchgemm_amd.grouped import bench_config4
    out["grouped_bf16"] = bench_config4(args.steps, args.warmup, args.seed)
    f_ms, b_ms = out["grouped_bf16"]["fwd_ms"], out["grouped_bf16"]["bwd_ms"]
    log(f"[bench] grouped fwd {f_ms:.3f} ms bwd {b_ms:.3f} ms")
    out["small_batch"] = small_batch_latency(bg, w, args)
    return out


def small_batch_latency(bg, w, args):
    """The reference's small random batch (B = 64, M, N, K ~ U{16..256},
    alpha 1, beta 0, U[-1, 1] inputs): a latency case, one launch of a few
    hundred items.  Microseconds per batch back to back and with a host
    synchronisation after each."""
    M, N, K = w.shapes_uniform(64, 16, 256, seed=0)
    vb = w.VBatch(M, N, K).fill("uniform", seed=0)
    flops = 2.0 * float(np.sum(M.astype(np.float64) * N * K))
    res = {"batch": "B=64, M,N,K ~ U{16..256}, seed 0", "gflop": round(flops / 1e9, 4)}
    for prec in ("fp32", "bf16x3"):
        p = bg.Plan(M, N, K, prec=prec)
        run = lambda: p.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
        us = 1e3 * time_launches(run, 200, 20)
        us_sync = 1e3 * time_synced(run, 100, 10)
        res[prec] = {"us": round(us, 1), "sync_us": round(us_sync, 1),
                     "gflops": round(flops / (us * 1e-6) / 1e9, 1)}
        log(f"[bench] small batch {prec}: {us:.1f} us back to back, {us_sync:.1f} us synchronised")
    return res


def traffic_probe(bg, vb, M, N, K, alpha, beta, args):
    """Kernels for HBM-traffic counters: a plain copy of 1 GiB (reads 1 GiB,
    writes 1 GiB: the scale against which FETCH_SIZE / WRITE_SIZE are read),
    then the workload's plan, each dispatch on its own."""
    import torch
    n = (1 << 30) // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    kernels = {}
    for prec in args.prec.split(","):
        opts = {"policy": "fixed", "fixed_tile": "128x128", "mixed_tiles": -1} \
            if args.workload == "uniform" else {"policy": args.policy}
        plan = bg.Plan(M, N, K, prec=prec, **opts)
        for _ in range(args.steps):
            plan.run(vb.dA, vb.dB, vb.dC, alpha, beta)
            torch.cuda.synchronize()
        kernels[prec] = [kernel_name(bg, prec, int(k), int(t)) for k, t, *_ in plan.kernels()]
        del plan
    print(json.dumps({"traffic_probe": True, "copy_bytes": n * 4, "workload": args.workload,
                      "steps": args.steps, "kernels": kernels}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="fp32",
                    help="fp32 | bf16x3 | fp16 | fp8 (--traffic-probe: a comma-separated list)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--policy", default="auto")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the bf16x3 and grouped fwd+bwd measurements")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: one 1 GiB device copy (the calibration "
                         "kernel of known traffic), then --steps runs of the workload's plan "
                         "for each --prec; no timing, no extras")
    ap.add_argument("--workload", default="unbalanced", choices=["unbalanced", "uniform", "batching"],
                    help="--traffic-probe only: the unbalanced B=2048 batch (default), a "
                         "uniform 64 x 2048^3 batch on 128x128 tiles, or the first batch of the "
                         "batching workload (B=7007, M,N,K ~ log-U[8,1024], alpha 1, beta 0)")
    ap.add_argument("--small-batch", action="store_true",
                    help="only the small random batch (B = 64, M,N,K ~ U{16..256}): FP32 and "
                         "bf16x3 microseconds per batch, as one JSON object (for kernel traces)")
    args = ap.parse_args()
    precs = args.prec.split(",")
    for pr in precs:
        if pr not in PEAK or pr == "bf16":
            ap.error(f"unknown --prec {pr}")
    if len(precs) > 1 and not args.traffic_probe:
        ap.error("several --prec values only with --traffic-probe")

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batched GEMM path is single-GPU in this "
              "build (multi-GPU sharding is a design note, not implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import batchgemm_amd as bg
    from batchgemm_amd import workloads as w

    if args.small_batch:
        print(json.dumps(small_batch_latency(bg, w, args)))
        return
    alpha, beta = 1.5, 0.5
    if args.traffic_probe and args.workload == "uniform":
        M = N = K = np.full(64, 2048, np.int32)
    elif args.traffic_probe and args.workload == "batching":
        from batchgemm_amd.autotune import batching_workload
        M, N, K = (np.asarray(x, np.int32) for x in batching_workload(50, args.seed)[0])
        alpha, beta = 1.0, 0.0
    else:
        M, N, K = w.shapes_unbalanced(2048, seed=args.seed)
    flops = w.flops(M, N, K)
    log(f"[bench] workload: {len(M)} matrices, {flops / 1e9:.1f} GFLOP, prec={args.prec}")
    vb = w.VBatch(M, N, K).fill("normal", seed=args.seed, c_dist="uniform")
    torch.cuda.synchronize()
    if args.traffic_probe:
        traffic_probe(bg, vb, M, N, K, alpha, beta, args)
        return
    t0 = time.perf_counter()
    plan = bg.Plan(M, N, K, prec=args.prec, policy=args.policy)
    plan_ms = (time.perf_counter() - t0) * 1e3
    log(f"[bench] plan: {plan.n_items} items, {plan.n_launches} launches, grid {plan.grid}, "
        f"{plan_ms:.1f} ms")

    run = lambda: plan.run(vb.dA, vb.dB, vb.dC, alpha, beta)
    ms = time_launches(run, args.steps, args.warmup)
    value = flops / (ms * 1e-3) / 1e9
    log(f"[bench] {ms:.3f} ms/step  {value:.1f} GFLOP/s")

    # Per-kernel durations (HIP events recorded by a timing plan of the same
    # batch around each of its kernels, on the launch stream; a timing plan
    # runs its launches one after another) and each kernel's algorithmic flops
    # (2*rows*cols*k of its items, padding excluded).
    tplan = bg.Plan(M, N, K, prec=args.prec, policy=args.policy, timing_slots=args.steps)
    time_launches(lambda: tplan.run(vb.dA, vb.dB, vb.dC, alpha, beta), args.steps, args.warmup)
    kt = tplan.kernels()
    kms = tplan.kernel_times()
    iflops = tplan.item_flops()
    rows = []
    for (kind, tile, first, cnt, grid), t in zip(kt, kms):
        # GEMM launches (FP32, sequential, split, BF16, split on packed
        # operands) carry flops; amax, split-K reduction and pack passes none
        f = float(iflops[first:first + cnt].sum()) if kind in (1, 2, 3, 4, 7) else 0.0
        rows.append({"kernel": kernel_name(bg, args.prec, int(kind), int(tile)), "items": int(cnt),
                     "grid": int(grid), "ms": round(float(t), 4), "gflop": round(f / 1e9, 3),
                     "tflops": round(f / (t * 1e-3) / 1e12, 1) if f and t > 0 else None})
    dom = max(range(len(rows)), key=lambda r: rows[r]["gflop"])
    dom_ms = float(kms[dom]); dom_flops = rows[dom]["gflop"] * 1e9
    achieved = dom_flops / (dom_ms * 1e-3)
    peak = PEAK[args.prec]
    log(f"[bench] kernels: " + "; ".join(f"{r['kernel']} {r['ms']} ms {r['gflop']} GFLOP" for r in rows))
    log(f"[bench] dominant kernel {rows[dom]['kernel']}: {dom_ms:.3f} ms, "
        f"{achieved / 1e12:.1f} TFLOP/s")
    traffic, traffic_commit = read_traffic(rows[dom]["kernel"])

    line = {
        "metric": "batched GFLOP/s (2*sum(MNK)/t), unbalanced variable-size FP32 batched GEMM",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16x3": "bf16x3 (f32 accumulate)", "fp16": "f16 (f32 accumulate)",
                  "fp8": "e4m3 (f32 accumulate)"}[args.prec],
        "data": "synthetic (seeded N(0,1) A,B; U[-1,1] C)",
        "config": {"workload": "unbalanced vbatched GEMM B=2048: 90% M,N,K~U{16..128}, "
                               "10% M,N,K~U{512..2048}, alpha=1.5, beta=0.5, column-major",
                   "batch": 2048, "gflop_per_step": round(flops / 1e9, 3), "prec": args.prec,
                   "policy": args.policy, "plan_ms": round(plan_ms, 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved / 1e12, 2),
                     "peak": round(peak / 1e12, 1), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "traffic_measured_on": {"commit": traffic_commit, "file": TRAFFIC_FILE,
                                             "source_digest": source_digest(),
                                             "current": traffic is not None},
                     "kernel": rows[dom]["kernel"], "launch_ms": round(dom_ms, 4),
                     "launch_gflop": rows[dom]["gflop"], "kernels": rows},
        "cpu_baseline": None,
    }
    if args.prec == "fp32" and not args.no_extras:
        line["extras"] = extras(bg, w, vb, M, N, K, alpha, beta, flops, value, args)
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(M, N, K, alpha, beta, args.cpu_seconds)
        log(f"[bench] CPU baseline {line['cpu_baseline']['value']} GFLOP/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
