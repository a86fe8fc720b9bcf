"""HIP-graph capture and replay of planned batches (needs an MI355X).

Every replay is checked against the eager run of the same plan on the same
inputs (bit for bit: the kernels are deterministic, split-K included), with
the program's own eager work and synchronisations between replays, and with
the operands rebound between replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec,ksplit", [("fp32", 0), ("fp32", 3), ("bf16x3", 0), ("fp16", 0),
                                         ("fp8", 2)])
def test_graph_replays_match_eager(bg, gpu, prec, ksplit):
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.graph import GraphedBatch
    say(f"[graph] prec={prec} ksplit={ksplit}")
    M, N, K = w.shapes_unbalanced(160, seed=21)
    alpha, beta = 1.5, 0.5
    vb1 = w.VBatch(M, N, K).fill("normal", seed=21, c_dist="uniform")
    vb2 = w.VBatch(M, N, K).fill("normal", seed=22, outlier_frac=0.01, c_dist="uniform")
    C1, C2 = vb1.C.clone(), vb2.C.clone()
    plan = bg.Plan(M, N, K, prec=prec, force_ksplit=ksplit)
    # eager results on both batches
    plan.run(vb1.dA, vb1.dB, vb1.dC, alpha, beta)
    plan.run(vb2.dA, vb2.dB, vb2.dC, alpha, beta)
    torch.cuda.synchronize()
    E1, E2 = vb1.C.clone(), vb2.C.clone()
    vb1.C.copy_(C1)
    g = GraphedBatch(plan, vb1.dA, vb1.dB, vb1.dC, alpha, beta, warmup=1)
    torch.cuda.synchronize()
    assert torch.equal(vb1.C, E1)  # the warm-up run
    for rep in range(3):
        vb1.C.copy_(C1)            # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vb1.C.view(torch.int32), E1.view(torch.int32)), ("replay", rep)
    g.set_pointers(vb2.dA, vb2.dB, vb2.dC)
    for rep in range(2):
        vb2.C.copy_(C2)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vb2.C.view(torch.int32), E2.view(torch.int32)), ("rebound replay", rep)
    assert torch.equal(vb1.C.view(torch.int32), E1.view(torch.int32))  # untouched after rebinding


def test_graph_replay_back_to_back_accumulates(bg, gpu):
    """Replays issued back to back with no host synchronisation between them
    order like eager launches: C <- C + A*B applied R times."""
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.graph import GraphedBatch
    say("[graph] back-to-back replays")
    M, N, K = w.shapes_uniform(48, 16, 256, 5)
    vb = w.VBatch(M, N, K).fill("uniform", seed=5, c_dist="uniform")
    C0 = vb.C.clone()
    plan = bg.Plan(M, N, K)
    for _ in range(4):
        plan.run(vb.dA, vb.dB, vb.dC, 1.0, 1.0)
    torch.cuda.synchronize()
    E = vb.C.clone()
    vb.C.copy_(C0)
    g = GraphedBatch(plan, vb.dA, vb.dB, vb.dC, 1.0, 1.0, warmup=1)
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    assert torch.equal(vb.C.view(torch.int32), E.view(torch.int32))
