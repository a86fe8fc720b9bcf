"""Autotuner candidates on the GPU: every candidate planner computes the same
batched GEMM (checked against the FP64 oracle), and tuning records a choice
(needs an MI355X)."""
import sys

import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_every_candidate_is_correct(bg, orc, gpu):
    import torch
    from batchgemm_amd import autotune as at
    from batchgemm_amd import workloads as w
    from test_gpu_fp32 import oracle_check
    M, N, K = w.shapes_loguniform(120, seed=7)
    for name, opts in at.CANDIDATES:
        say(f"[autotune] candidate {name}")
        vb = w.VBatch(M, N, K).fill("uniform", seed=7, c_dist="uniform")
        C0 = vb.C.cpu().numpy()
        try:
            p = bg.Plan(M, N, K, **opts)
        except bg._lib.VbgError as e:
            assert e.code == -6 and opts.get("policy") == "reference"  # undefined for these shapes
            continue
        p.run(vb.dA, vb.dB, vb.dC, 1.25, -0.5)
        torch.cuda.synchronize()
        oracle_check(orc, vb, 1.25, -0.5, C0)


def test_tune_records_a_choice(bg, gpu, tmp_path):
    import numpy as np
    from batchgemm_amd import autotune as at
    say("[autotune] tune")
    M, N, K = at.batching_workload(1, seed=11)[0]
    M, N, K = M[:300], N[:300], K[:300]
    t = at.Autotuner(cache_path=str(tmp_path / "t.json"), runs=3)
    rec = t.tune(M, N, K, "fp32")
    assert rec["choice"] in dict(at.CANDIDATES) and "auto" in rec["ms"]
    assert rec["ms"][rec["choice"]] <= rec["ms"]["auto"]
    p = t.plan(M, N, K, prec="fp32")
    assert p.n_items > 0 and np.isfinite(rec["gain"])
