"""Autotuner: picks the planner's tile-size set and thresholds per shape
distribution, keeping the fixed heuristic unless a candidate beats it.

The reference fixes its planner by hand: one tile table and two thresholds
passed on the command line (ROCm/framework/gemm.cuh:121-201, run.sh:2-4), with
the alternative heuristics of Wang (ROCm/tiling/gemm_wang.cuh:104-182) and
Li (CUDA/batching/gemm.cu:99-157) in other trees.  Here those policies, and
variants of this library's own cost model, are candidates:

- ``auto``                   the default heuristic (cost model, all tiles)
- ``auto`` + tile set        ``tile_mask``: no 128x128 / only >= 64x64 / no 16x16
- ``auto`` + thresholds      ``big_tile_min_mn`` 256 or 512; split-K off;
                             edge tiles off / on; tile grouping off
- ``auto`` + launch shape    one mixed launch for the 64..128 tiles, or one
                             launch per tile shape
- ``auto`` + cost model      the cost model's constants on a grid: the fixed
                             cost per work item (1, 3, 6, 12 units) times the
                             modelled efficiency of the tiles below 128x128
                             (x0.70 .. x1.50)
- ``reference``              the reference planner, wavefront_thres 2500 / 6000 / 20000
- ``wang`` / ``li``          TLP thresholds 65536 / 262144

A batch's distribution key is coarse -- precision, transposes, log2 of the
batch size and of the median M, N, K (half-octave steps), and the fraction of
matrices with min(M, N) >= 512 (tenths) -- so that batches drawn from one
distribution share a choice.  Tuning times every candidate on synthetic data
of the batch's own shapes (HIP events, median of the runs) and keeps the
default unless a candidate is more than ``min_gain`` faster.  Choices are
cached as JSON (``VBG_AUTOTUNE_CACHE`` or ~/.cache/vbgemm/autotune.json).

    python -m batchgemm_amd.autotune --bench [--batches 50] [--prec fp32,bf16x3]

runs the batching workload: 50 successive batches, batch size ~ U{256..8192},
M, N, K ~ log-uniform [8, 1024] independently, entries ~ U[-1, 1]; for each,
the fixed heuristic, the autotuned plan and the autotuned plan replayed as a
HIP graph, reporting per-batch GFLOP/s and p50 / p99 batch latency as one JSON
line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib

CANDIDATES = [
    ("auto", {}),
    ("auto:no128", {"tile_mask": 0b011111}),
    ("auto:ge64", {"tile_mask": 0b111100}),
    ("auto:no16", {"tile_mask": 0b111110}),
    ("auto:big256", {"big_tile_min_mn": 256}),
    ("auto:big512", {"big_tile_min_mn": 512}),
    ("auto:nosplitk", {"split_k_min": -1}),
    ("auto:noedges", {"edge_tiles": -1}),
    ("auto:edges", {"edge_tiles": 1}),
    ("auto:mixed", {"mixed_tiles": 1}),
    ("auto:separate", {"mixed_tiles": -1}),
    ("auto:group1", {"tile_group": 1}),
    ("auto:noconsolidate", {"consolidate": -1}),
    ("auto:only128", {"tile_mask": 0b100000}),
    ("reference:2500", {"policy": "reference", "wavefront_thres": 2500}),
    ("reference:6000", {"policy": "reference", "wavefront_thres": 6000}),
    ("reference:20000", {"policy": "reference", "wavefront_thres": 20000}),
    ("wang:65536", {"policy": "wang", "tlp_thres": 65536}),
    ("wang:262144", {"policy": "wang", "tlp_thres": 262144}),
    ("li:65536", {"policy": "li", "tlp_thres": 65536}),
    ("li:262144", {"policy": "li", "tlp_thres": 262144}),
]
# The AUTO cost model's constants (the per-item fixed cost and the modelled
# MFMA efficiency of the tiles below 128x128, which together set where the
# planner trades large tiles for more, smaller items): a grid around the
# defaults (3.0 units, x1.00).
COST_GRID = [(f, e) for f in (10, 30, 60, 120) for e in (70, 85, 100, 120, 150) if (f, e) != (30, 100)]
CANDIDATES += [(f"auto:cost:f{f}:e{e}", {"item_fixed_x10": f, "small_tile_eff_pct": e}) for f, e in COST_GRID]
DEFAULT = "auto"


def distribution_key(M, N, K, prec="fp32", transA="N", transB="N") -> str:
    M = np.asarray(M, np.int64); N = np.asarray(N, np.int64); K = np.asarray(K, np.int64)
    if len(M) == 0:
        return f"{prec}|{transA}{transB}|empty"

    def oct2(x):
        return int(round(2 * math.log2(max(float(x), 1.0))))
    big = float(np.mean(np.minimum(M, N) >= 512))
    return (f"{prec}|{transA.upper()}{transB.upper()}|B{int(round(math.log2(len(M))))}"
            f"|m{oct2(np.median(M))}n{oct2(np.median(N))}k{oct2(np.median(K))}|big{int(round(10 * big))}")


def default_cache_path() -> str:
    return os.environ.get("VBG_AUTOTUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "vbgemm", "autotune.json")


class Autotuner:
    """Per-distribution planner choice with a JSON cache.

    ``plan(M, N, K, ...)`` returns a Plan built with the cached choice for the
    batch's key, tuning first (on synthetic data) when the key is new and
    ``tune_missing`` is set; ``tune`` forces a measurement."""

    def __init__(self, cache_path: str | None = None, min_gain: float = 0.03, runs: int = 5,
                 tune_missing: bool = True, candidates=None):
        self.cache_path = cache_path or default_cache_path()
        self.min_gain = min_gain
        self.runs = runs
        self.tune_missing = tune_missing
        self.candidates = list(candidates or CANDIDATES)
        self.table: dict[str, dict] = {}
        if os.path.exists(self.cache_path):
            with open(self.cache_path) as f:
                self.table = json.load(f)

    def save(self) -> None:
        d = os.path.dirname(self.cache_path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    def choice(self, key: str) -> tuple[str, dict]:
        name = self.table.get(key, {}).get("choice", DEFAULT)
        return name, dict(self.candidates)[name] if name in dict(self.candidates) else {}

    def tune(self, M, N, K, prec="fp32", transA="N", transB="N", seed=0) -> dict:
        """Time every candidate on synthetic operands of these shapes; record
        and return {"choice", "ms": {candidate: median ms}, "gain"}."""
        import torch
        from .api import Plan
        from .workloads import VBatch
        key = distribution_key(M, N, K, prec, transA, transB)
        vb = VBatch(M, N, K, transA, transB).fill("uniform", seed=seed)
        times = {}
        for name, opts in self.candidates:
            try:
                p = Plan(vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc, transA, transB, prec, **opts)
            except _lib.VbgError:
                continue  # e.g. the reference planner is undefined for this batch
            times[name] = median_ms(lambda: p.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0), self.runs)
            p.close()
        del vb
        torch.cuda.empty_cache()
        base = times.get(DEFAULT, float("inf"))
        best = min(times, key=times.get)
        choice = best if times[best] < base * (1.0 - self.min_gain) else DEFAULT
        rec = {"choice": choice, "ms": {k: round(v, 4) for k, v in times.items()},
               "gain": round(base / times[choice], 4)}
        self.table[key] = rec
        return rec

    def plan(self, M, N, K, lda=None, ldb=None, ldc=None, transA="N", transB="N", prec="fp32",
             **extra):
        from .api import Plan
        key = distribution_key(M, N, K, prec, transA, transB)
        if key not in self.table and self.tune_missing:
            self.tune(M, N, K, prec, transA, transB)
            self.save()
        name, opts = self.choice(key)
        return Plan(M, N, K, lda, ldb, ldc, transA, transB, prec, **{**opts, **extra})


def median_ms(fn, runs: int, warmup: int = 2) -> float:
    """Median over `runs` of one call bracketed by HIP events on the current
    stream (each call synchronised)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(runs):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        ts.append(s.elapsed_time(e))
    return float(np.median(ts))


def batching_workload(n_batches: int, seed: int):
    """Shapes of the batching workload: B ~ U{256..8192}; M, N, K ~
    round(exp(U[ln 8, ln 1024])) independently."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = math.log(8), math.log(1024)
    for _ in range(n_batches):
        B = int(rng.integers(256, 8193))
        dims = np.rint(np.exp(rng.uniform(lo, hi, size=(3, B)))).astype(np.int32)
        out.append((dims[0], dims[1], dims[2]))
    return out


def bench(n_batches=50, precs=("fp32", "bf16x3"), seed=0, runs=5, cache_path=None):
    import torch
    from .api import Plan
    from .graph import GraphedBatch
    from .workloads import VBatch, flops
    batches = batching_workload(n_batches, seed)
    result = {"workload": f"{n_batches} batches, B ~ U{{256..8192}}, M,N,K ~ log-U[8,1024], "
                          f"U[-1,1] entries, alpha 1, beta 0", "seed": seed}
    for prec in precs:
        tuner = Autotuner(cache_path=cache_path, runs=runs)
        t0 = time.perf_counter()
        for M, N, K in batches[:5]:  # tuning set: the first batches of the sequence
            key = distribution_key(M, N, K, prec)
            if key not in tuner.table:
                tuner.tune(M, N, K, prec)
        tune_s = time.perf_counter() - t0
        rows = []
        for i, (M, N, K) in enumerate(batches):
            F = flops(M, N, K)
            vb = VBatch(M, N, K).fill("uniform", seed=seed + 1000 + i)
            t1 = time.perf_counter()
            pd = Plan(M, N, K, prec=prec)
            plan_ms = (time.perf_counter() - t1) * 1e3
            d_ms = median_ms(lambda: pd.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0), runs)
            key = distribution_key(M, N, K, prec)
            tuned_now = key not in tuner.table
            pt = tuner.plan(M, N, K, prec=prec)
            name = tuner.choice(key)[0]
            t_ms = median_ms(lambda: pt.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0), runs)
            g = GraphedBatch(pt, vb.dA, vb.dB, vb.dC, 1.0, 0.0)
            g_ms = median_ms(g.replay, runs)
            rows.append({"B": len(M), "gflop": round(F / 1e9, 3), "plan_ms": round(plan_ms, 2),
                         "fixed_ms": round(d_ms, 4), "tuned_ms": round(t_ms, 4),
                         "graph_ms": round(g_ms, 4), "choice": name, "tuned_on_demand": tuned_now})
            print(f"[autotune] {prec} batch {i}: B={len(M)} {F / 1e9:.1f} GFLOP fixed {d_ms:.3f} ms "
                  f"tuned({name}) {t_ms:.3f} ms graph {g_ms:.3f} ms", file=sys.stderr, flush=True)
            del g, pd, pt, vb
            torch.cuda.empty_cache()
        tuner.save()

        def lat(col):
            v = np.array([r[col] for r in rows])
            return {"p50_ms": round(float(np.percentile(v, 50)), 4),
                    "p99_ms": round(float(np.percentile(v, 99)), 4)}
        gf = np.array([r["gflop"] for r in rows])
        summary = {}
        for col in ("fixed_ms", "tuned_ms", "graph_ms"):
            ms = np.array([r[col] for r in rows])
            summary[col.replace("_ms", "")] = {
                **lat(col), "total_gflops": round(float(gf.sum() / ms.sum() * 1e3), 1),
                "per_batch_gflops_min": round(float((gf / ms * 1e3).min()), 1),
                "per_batch_gflops_median": round(float(np.median(gf / ms * 1e3)), 1)}
        summary["plan_ms"] = lat("plan_ms")
        summary["tuning_s"] = round(tune_s, 2)
        summary["choices"] = {k: v["choice"] for k, v in tuner.table.items() if k.startswith(prec)}
        summary["per_batch"] = [{k: r[k] for k in ("B", "gflop", "fixed_ms", "tuned_ms", "graph_ms")}
                                for r in rows]
        result[prec] = summary
    return result


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="autotuner / batching benchmark")
    ap.add_argument("--bench", action="store_true", help="run the 50-batch batching workload")
    ap.add_argument("--batches", type=int, default=50)
    ap.add_argument("--prec", default="fp32,bf16x3")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--cache", default=None, help="autotune cache file (JSON)")
    args = ap.parse_args(argv)
    if not args.bench:
        ap.print_help()
        return 0
    res = bench(args.batches, tuple(args.prec.split(",")), args.seed, args.runs, args.cache)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
