"""Benchmark driver with the reference's protocol.

    python -m batchgemm_amd.driver BATCH TLP_thres wavefront_thres [--data FILE]
           [--policy reference|auto|wang|li|fixed] [--prec fp32|fp32_seq|bf16x3|...] [--verify]

Follows ROCm/framework/gemm.cuh and its runner ROCm/framework/run.sh:2-28:
- reads the first BATCH "M N K" lines of the dataset (gemm.cuh:34-48; default
  the shipped data_MN_K_512_128);
- fills every element of A_i, B_i, C_i with i+1, i+2, i+3 (gemm.cuh:58-80);
- one warm-up launch and a device synchronisation, then N_RUNS = 10 launches,
  each followed by a device synchronisation, between two events (gemm.cuh:232-253);
- prints the reference's two lines, the mean time in seconds and
  sum(2MNK + 2MN)/1e9 GFLOP with its rate (gemm.cuh:112-116, 255-262).

The launches accumulate, C_i <- C_i + A_i B_i (alpha = beta = 1), since the
reference's tiles seed their accumulators from C.  --verify checks every
element against the closed form C_i = (i+3) + 11 K_i (i+1)(i+2), the
known-answer check of ROCm/tiling/log-verify, to 2e-6 relative; with
--prec fp32_seq (one FP32 fma per k, the reference tiles' arithmetic) it
checks bit-exact equality with that fma sequence replayed on the host, which
leaves the closed form once the sums pass 2^24.

The default policy "reference" takes the reference planner's tile choice
under wavefront_thres (gemm.cuh:141-201); TLP_thres is the threshold of the
Wang and Li policies.  Layout is column-major (immaterial for constant data).
Details go to stderr; stdout carries the reference's lines only.
--per-matrix-library times the same protocol on one library GEMM per matrix
(torch.addmm, i.e. rocBLAS / hipBLASLt), the reference's own comparison
program ROCm/default/gemm.cuh.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

N_RUNS = 10


def reference_gflop(M, N, K) -> float:
    """sum(2MNK + 2MN) / 1e9, the reference's count (gemm.cuh:112-116)."""
    M = np.asarray(M, np.int64); N = np.asarray(N, np.int64); K = np.asarray(K, np.int64)
    return float(np.sum(2 * M * N * K + 2 * M * N)) / 1.0e9


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("batch", type=int)
    ap.add_argument("tlp_thres", type=int)
    ap.add_argument("wavefront_thres", type=int)
    ap.add_argument("--data", default=None, help="dataset file (M N K per line)")
    ap.add_argument("--policy", default="reference", choices=["reference", "auto", "wang", "li", "fixed"])
    ap.add_argument("--prec", default="fp32")
    ap.add_argument("--verify", action="store_true")
    ap.add_argument("--per-matrix-library", action="store_true",
                    help="time one library GEMM per matrix instead (torch.addmm: rocBLAS / "
                         "hipBLASLt), the reference's baseline program ROCm/default/gemm.cuh:64-85")
    args = ap.parse_args(argv)

    import torch
    from . import workloads as w
    from .api import Plan

    M, N, K = w.read_dataset(args.data or w.dataset_path(), args.batch)
    B = len(M)
    vb = w.VBatch(M, N, K)
    for i in range(B):
        vb.a(i).fill_(i + 1)
        vb.b(i).fill_(i + 2)
        vb.c(i).fill_(i + 3)
    if args.per_matrix_library:
        from .baseline import library_loop
        run = library_loop(vb, 1.0, 1.0)  # C_i <- C_i + A_i B_i, one library call per matrix
        plan = None
    else:
        plan = Plan(M, N, K, prec=args.prec, policy=args.policy,
                    wavefront_thres=args.wavefront_thres, tlp_thres=args.tlp_thres)

        def run():
            plan.run(vb.dA, vb.dB, vb.dC, 1.0, 1.0)
    torch.cuda.synchronize()

    run()  # warm-up
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(N_RUNS):
        run()
        torch.cuda.synchronize()
    stop.record()
    stop.synchronize()
    t = start.elapsed_time(stop) / N_RUNS / 1.0e3  # seconds
    gflop = reference_gflop(M, N, K)
    print(f"time======{t:f}========N_RUNS====%d=====" % N_RUNS)
    print(f"gflops=={gflop:f}         {gflop / t:f}")
    sys.stdout.flush()
    what = (f"per-matrix library, {B} launches" if plan is None else
            f"policy={args.policy} prec={args.prec} items={plan.n_items} launches={plan.n_launches}")
    print(f"[driver] B={B} {what}; 2*sum(MNK) rate {w.flops(M, N, K) / t / 1e9:.1f} GFLOP/s",
          file=sys.stderr)

    if args.verify:
        seq = args.prec == "fp32_seq" and plan is not None
        want = sequential_fp32(K, N_RUNS + 1) if seq else None
        bad = 0
        for i in range(B):
            got = vb.c(i)
            if seq:  # the reference's arithmetic: bit for bit
                ok = bool((got == float(want[i])).all())
                err = float((got.double() - float(want[i])).abs().max())
            else:
                exact = (i + 3) + (N_RUNS + 1) * float(K[i]) * (i + 1) * (i + 2)
                err = float((got.double() - exact).abs().max()) / exact
                ok = err <= 2e-6
            if not ok:
                bad += 1
                print(f"[driver] matrix {i} ({M[i]}x{N[i]}x{K[i]}): max err {err:.3g}",
                      file=sys.stderr)
        what = "the sequential FP32 fma chain" if seq else "the closed form"
        print(f"[driver] verify: {B - bad} of {B} matrices match {what}", file=sys.stderr)
        return 1 if bad else 0
    return 0


def sequential_fp32(K, runs: int) -> np.ndarray:
    """Value of every element of C_i after `runs` launches of the reference's
    correct tiles on the driver's constant data: one FP32 fma per k,
    c <- fl(c + (i+1)(i+2)) (the product is exact), seeded with c = i+3.
    Equals (i+3) + runs K_i (i+1)(i+2) only while the sums stay below 2^24."""
    K = np.asarray(K, np.int64)
    i = np.arange(len(K), dtype=np.float64)
    ab = ((i + 1) * (i + 2)).astype(np.float32)
    c = (i + 3).astype(np.float32)
    for _ in range(runs):
        for k in range(int(K.max()) if len(K) else 0):
            c = np.where(k < K, (c + ab).astype(np.float32), c)
    return c


if __name__ == "__main__":
    sys.exit(main())
