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


def cpu_baseline(M, N, K, alpha, beta, budget_s=12.0):
    """Host per-matrix SGEMM loop (oracle/liboracle.so, OpenMP across matrices)
    on a bounded sample of the same workload: matrices in random order (keeping
    the 90/10 mix) until about `budget_s` seconds of CPU work."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc
    rng = np.random.default_rng(123)
    order = rng.permutation(len(M))
    done_flops, done_t, used = 0, 0.0, 0
    chunk = 64
    while done_t < budget_s and used < len(order):
        idx = order[used: used + chunk]
        used += len(idx)
        As = [rng.standard_normal(int(M[i]) * int(K[i])).astype(np.float32) for i in idx]
        Bs = [rng.standard_normal(int(K[i]) * int(N[i])).astype(np.float32) for i in idx]
        Cs = [rng.uniform(-1, 1, int(M[i]) * int(N[i])).astype(np.float32) for i in idx]
        m, n, k = M[idx], N[idx], K[idx]
        t0 = time.perf_counter()
        orc.sgemm_vbatched("N", "N", m, n, k, alpha, As, m, Bs, k, beta, Cs, m)
        done_t += time.perf_counter() - t0
        done_flops += int(2 * np.sum(m.astype(np.int64) * n * k))
    return {
        "value": round(done_flops / done_t / 1e9, 2), "unit": "GFLOP/s",
        "cores": orc.num_threads(), "kind": "port",
        "sample": f"{used} of {len(M)} matrices of the same workload in random order "
                  f"({done_flops / 1e9:.1f} GFLOP, {done_t:.1f} s): host per-matrix FP32 SGEMM, "
                  f"OpenMP across matrices",
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
    sample), and the grouped BF16 forward+backward (G=128, K=2048, N=1408,
    Zipf(1.2) group sizes summing to 32768)."""
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
        log(f"[bench] {prec}: {ms:.3f} ms, {out[prec]['gflops']} GFLOP/s, rel err {err:.3g}")
    out["bf16x3"]["speedup_vs_fp32"] = round(out["bf16x3"]["gflops"] / out["fp32"]["gflops"], 3)
    # grouped forward + backward
    from batchgemm_amd.grouped import GroupedPlan
    m_sizes = w.zipf_groups(128, 32768, 1.2, seed=args.seed)
    Kg, Ng = 2048, 1408
    g = torch.Generator(device="cuda").manual_seed(args.seed)
    A = (torch.randn(32768, Kg, generator=g, device="cuda") * 0.02).to(torch.bfloat16)
    B = (torch.randn(128, Kg, Ng, generator=g, device="cuda") * 0.02).to(torch.bfloat16)
    dC = torch.randn(32768, Ng, generator=g, device="cuda").to(torch.bfloat16)
    gp = GroupedPlan(m_sizes, Kg, Ng)
    C = torch.empty(32768, Ng, dtype=torch.bfloat16, device="cuda")
    dA = torch.empty_like(A); dB = torch.empty_like(B)
    L = gp._p
    lib = bg.lib()
    from batchgemm_amd.api import _stream_ptr
    def fwd():
        bg._lib.check(lib.vbg_grouped_gemm_fwd(L, A.data_ptr(), B.data_ptr(), C.data_ptr(), 1,
                                               _stream_ptr(None)), "grouped fwd")

    def bwd():
        bg._lib.check(lib.vbg_grouped_gemm_bwd(L, A.data_ptr(), B.data_ptr(), dC.data_ptr(), 1,
                                               dA.data_ptr(), dB.data_ptr(), 1, _stream_ptr(None)),
                      "grouped bwd")
    f_ms = time_launches(fwd, args.steps, args.warmup)
    b_ms = time_launches(bwd, args.steps, args.warmup)
    gfl = 2.0 * 32768 * Kg * Ng
    out["grouped_bf16"] = {
        "workload": "G=128, K=2048, N=1408, Zipf(1.2) rows summing to 32768, BF16 in/out, FP32 accumulate",
        "fwd_ms": round(f_ms, 4), "bwd_ms": round(b_ms, 4), "fwd_bwd_ms": round(f_ms + b_ms, 4),
        "fwd_tflops": round(gfl / (f_ms * 1e-3) / 1e12, 1),
        "bwd_tflops": round(2 * gfl / (b_ms * 1e-3) / 1e12, 1)}
    log(f"[bench] grouped fwd {f_ms:.3f} ms bwd {b_ms:.3f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="fp32", choices=["fp32", "bf16x3", "fp16", "fp8"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--policy", default="auto")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the bf16x3 and grouped fwd+bwd measurements")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batched GEMM path is single-GPU in this "
              "build (multi-GPU sharding is a design note, not implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import batchgemm_amd as bg
    from batchgemm_amd import workloads as w

    alpha, beta = 1.5, 0.5
    M, N, K = w.shapes_unbalanced(2048, seed=args.seed)
    flops = w.flops(M, N, K)
    log(f"[bench] workload: 2048 matrices, {flops / 1e9:.1f} GFLOP, prec={args.prec}")
    vb = w.VBatch(M, N, K).fill("normal", seed=args.seed, c_dist="uniform")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = bg.Plan(M, N, K, prec=args.prec, policy=args.policy)
    plan_ms = (time.perf_counter() - t0) * 1e3
    log(f"[bench] plan: {plan.n_items} items, {plan.n_launches} launches, grid {plan.grid}, "
        f"{plan_ms:.1f} ms")

    run = lambda: plan.run(vb.dA, vb.dB, vb.dC, alpha, beta)
    ms = time_launches(run, args.steps, args.warmup)
    value = flops / (ms * 1e-3) / 1e9
    log(f"[bench] {ms:.3f} ms/step  {value:.1f} GFLOP/s")

    # Dominant kernel: the launch holding the 128x128 FP32 MFMA tiles (the
    # large matrices).  Its matrices are planned alone with the same tile and
    # timed with HIP events on the launch stream.
    tiles = plan.tiles()
    dom_tile = int(np.bincount(tiles[tiles >= 0], weights=(2.0 * M * N * K)[tiles >= 0]).argmax())
    sel = np.where(tiles == dom_tile)[0]
    sub = bg.Plan(M[sel], N[sel], K[sel], prec=args.prec, policy="fixed", fixed_tile=dom_tile)
    dA, dB, dC = vb.dA[sel], vb.dB[sel], vb.dC[sel]
    sub_ms = time_launches(lambda: sub.run(dA, dB, dC, alpha, beta), args.steps, args.warmup)
    sub_flops = w.flops(M[sel], N[sel], K[sel])
    achieved = sub_flops / (sub_ms * 1e-3)
    peak = PEAK[args.prec]
    log(f"[bench] dominant launch: tile {bg.TILE_SHAPES[dom_tile]}, {len(sel)} matrices, "
        f"{sub_ms:.3f} ms, {achieved / 1e12:.1f} TFLOP/s")

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
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": f"vbg_gemm_kernel<{bg.TILE_SHAPES[dom_tile]}>",
                     "launch_ms": round(sub_ms, 4), "launch_gflop": round(sub_flops / 1e9, 3)},
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
