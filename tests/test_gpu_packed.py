"""Packed operands for the split precisions (needs an MI355X).

The pack pre-pass quantises A and B once per run into k-contiguous rows and the
GEMM kernels stream those rows; it applies the same per-element rounding (bf16
hi/lo split, x * 2^e to FP16, x * 448/amax to e4m3) as the in-loop
quantisation, and for bf16x3 and FP16 the MFMA sequence per output element is
unchanged.  So a packed run must equal the unpacked run bit for bit -- and the
unpacked run is checked against the FP64 oracle and a numpy emulation of the
quantisation in test_gpu_precision.py.  Packed FP8 runs on the 32x32x64
f8f6f4 MFMA (the in-loop path on the 32x32x16 FP8 one), which accumulates
differently: it is held to the FP8 accumulation bound instead.  Cases: every transpose pair, ragged M / N / K (K not a
multiple of the 64-element row padding, K = 0, M or N = 1), several pack jobs
per operand (M, N > 64 and K > 256), leading dimensions past the row count
whose padding holds 1e30 (a pack that read it would blow the result up),
split-K, mixed and per-shape launches, every tile shape on its own (the
packed kernels stage each shape's k-tiles through LDS in 1 KB pieces whose
split between the operands depends on the shape), a second run on new operand
values, and graph replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SHAPES = (
    [64, 100, 17, 130, 3, 1, 700, 96, 257, 40, 128],
    [64, 33, 140, 128, 7, 90, 520, 1, 300, 40, 256],
    [32, 200, 9, 257, 5, 70, 1100, 64, 0, 1, 600],
)


def poison_padding(vb):
    for pool, off, rows, cols, ld in ((vb.A, vb.offA, vb.a_rows, vb.a_cols, vb.lda),
                                      (vb.B, vb.offB, vb.b_rows, vb.b_cols, vb.ldb)):
        for b in range(vb.batch):
            if int(cols[b]) == 0:
                continue
            v = pool[int(off[b]): int(off[b]) + int(ld[b]) * int(cols[b])].view(int(cols[b]), int(ld[b]))
            v[:, int(rows[b]):] = 1e30


@pytest.mark.parametrize("prec", ["bf16x3", "fp16", "fp8"])
@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_packed_equals_unpacked(bg, gpu, prec, trans):
    import torch
    from batchgemm_amd.workloads import VBatch
    M, N, K = (np.asarray(x, np.int32) for x in SHAPES)
    fixed = [{"policy": "fixed", "fixed_tile": t, "mixed_tiles": -1}
             for t in ("64x64", "128x64", "64x128", "128x128") + (("256x128",) if prec == "bf16x3" else ())]
    for opts in ({"mixed_tiles": -1}, {"mixed_tiles": 1, "force_ksplit": 3}, *fixed):
        say(f"[packed] {prec} trans={trans} {opts}")
        vb = VBatch(M, N, K, trans[0], trans[1], pad_ab=3)
        plans = {pk: bg.Plan(M, N, K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], prec=prec,
                             pack_operands=pk, **opts) for pk in (1, -1)}
        kinds = list(plans[1].kernels()[:, 0])
        assert 6 in kinds and 7 in kinds and plans[1].packed_bytes > 0
        assert 6 not in plans[-1].kernels()[:, 0] and plans[-1].packed_bytes == 0
        for seed in (41, 42):  # the second run packs new operand values
            vb.fill("normal", seed=seed, outlier_frac=0.01, c_dist="uniform")
            poison_padding(vb)
            C0 = vb.C.clone()
            out = {}
            for pk, p in plans.items():
                vb.C.copy_(C0)
                p.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
                torch.cuda.synchronize()
                out[pk] = vb.C.clone()
            assert torch.isfinite(out[1]).all()
            if prec == "fp8":
                # Packed FP8 runs on the 32x32x64 f8f6f4 MFMA, the in-loop
                # path on the 32x32x16 FP8 one: the same quantised operands,
                # but the two accumulate their products differently inside
                # the MFMA (each within the FP8 accumulation bound against
                # exact products, checked in test_gpu_precision.py and
                # test_gpu_workloads.py).  Here: equal to 2 x 2e-4 relative.
                d = float(torch.linalg.norm((out[1] - out[-1]).double()))
                assert d <= 4e-4 * float(torch.linalg.norm(out[-1].double())), (seed, d)
            else:
                assert torch.equal(out[1].view(torch.int32), out[-1].view(torch.int32)), seed


@pytest.mark.parametrize("prec", ["bf16x3", "fp8"])
def test_packed_graph_replays(bg, gpu, prec):
    """A packed plan captured once replays the pack pre-pass with the GEMM:
    replays on rebound operands give the eager results bit for bit."""
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.graph import GraphedBatch
    say(f"[packed graph] {prec}")
    M, N, K = w.shapes_unbalanced(120, seed=43)
    vb1 = w.VBatch(M, N, K).fill("normal", seed=43, c_dist="uniform")
    vb2 = w.VBatch(M, N, K).fill("normal", seed=44, outlier_frac=0.01, c_dist="uniform")
    C1, C2 = vb1.C.clone(), vb2.C.clone()
    plan = bg.Plan(M, N, K, prec=prec, pack_operands=1)
    assert 6 in plan.kernels()[:, 0]
    plan.run(vb1.dA, vb1.dB, vb1.dC, 1.5, 0.5)
    plan.run(vb2.dA, vb2.dB, vb2.dC, 1.5, 0.5)
    torch.cuda.synchronize()
    E1, E2 = vb1.C.clone(), vb2.C.clone()
    vb1.C.copy_(C1)
    g = GraphedBatch(plan, vb1.dA, vb1.dB, vb1.dC, 1.5, 0.5, warmup=1)
    torch.cuda.synchronize()
    for rep in range(3):
        vb1.C.copy_(C1)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vb1.C.view(torch.int32), E1.view(torch.int32)), ("replay", rep)
    g.set_pointers(vb2.dA, vb2.dB, vb2.dC)
    for rep in range(2):
        vb2.C.copy_(C2)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vb2.C.view(torch.int32), E2.view(torch.int32)), ("rebound replay", rep)


@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_packed_fp16_256x128_equals_in_loop(bg, gpu, trans):
    """Packed FP16 on the 256x128 LDS-DMA tile (the AUTO choice for its big
    matrices) gives the bits of FP16 quantised in the k-loop on 128x128
    tiles: the same quantised operands, and the k grouping into
    32x32x16 f16 MFMAs is the same for every tile shape and k-tile depth."""
    import torch
    from batchgemm_amd.workloads import VBatch
    M, N, K = (np.asarray(x, np.int32) for x in SHAPES)
    say(f"[packed] fp16 256x128 trans={trans}")
    vb = VBatch(M, N, K, trans[0], trans[1], pad_ab=3)
    common = dict(prec="fp16", policy="fixed", mixed_tiles=-1)
    packed = bg.Plan(M, N, K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], pack_operands=1,
                     fixed_tile="256x128", **common)
    in_loop = bg.Plan(M, N, K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], pack_operands=-1,
                      fixed_tile="128x128", **common)
    assert (packed.kernels()[:, 1] == 6).any() and packed.packed_bytes > 0
    for seed in (43, 44):
        vb.fill("normal", seed=seed, outlier_frac=0.01, c_dist="uniform")
        poison_padding(vb)
        C0 = vb.C.clone()
        out = []
        for p in (packed, in_loop):
            vb.C.copy_(C0)
            p.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
            torch.cuda.synchronize()
            out.append(vb.C.clone())
        assert torch.isfinite(out[0]).all()
        assert torch.equal(out[0].view(torch.int32), out[1].view(torch.int32)), seed
