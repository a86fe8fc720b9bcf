"""GPU parity of the FP32 paths against the CPU oracle (needs an MI355X).

Bar: relative Frobenius error <= 1e-5 per matrix against the FP64 oracle for
the MFMA path; bit-exact against the oracle's sequential-fma tile model for the
reference-compatible VBG_PREC_FP32_SEQ path."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_check(orc, vb, alpha, beta, C0, mats=None, tol=TOL):
    """Compare vb.C (after the GPU run) with the FP64 oracle on C0 (host copies
    of the C pool before the run).  Returns the max per-matrix error."""
    mats = range(vb.batch) if mats is None else mats
    worst = 0.0
    for b in mats:
        M, N, K = int(vb.M[b]), int(vb.N[b]), int(vb.K[b])
        if M == 0 or N == 0:
            continue
        A = vb.host_flat("A", b); B = vb.host_flat("B", b)
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + int(vb.ldc[b]) * N]
        ref = orc.dgemm(vb.transA, vb.transB, A, B, c0, alpha, beta, M, N, K, int(vb.lda[b]),
                        int(vb.ldb[b]), int(vb.ldc[b]))
        got = vb.host_flat("C", b)
        err = orc.rel_fro(got, ref)
        worst = max(worst, err)
        assert err <= tol, (b, M, N, K, err)
    return worst


def run_case(bg, orc, M, N, K, ta, tb, alpha, beta, seed, **opts):
    from batchgemm_amd.workloads import VBatch
    import torch
    vb = VBatch(M, N, K, ta, tb).fill("normal", seed=seed, c_dist="uniform")
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc, ta, tb, **opts)
    plan.run(vb.dA, vb.dB, vb.dC, alpha, beta)
    torch.cuda.synchronize()
    return vb, C0, plan


EDGE_SHAPES = ([1, 2, 15, 16, 17, 33, 63, 64, 65, 127, 129, 200, 31],
               [1, 3, 16, 15, 31, 64, 65, 1, 130, 128, 127, 77, 33],
               [1, 0, 7, 16, 33, 40, 64, 129, 3, 200, 255, 64, 1])


@pytest.mark.parametrize("mixed", [1, -1])
@pytest.mark.parametrize("tile", ["16x16", "32x32", "64x64", "128x64", "64x128", "128x128",
                                  "16x128", "128x16"])
@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
def test_fp32_every_tile_every_transpose(bg, orc, gpu, tile, trans, mixed):
    """Every tile shape and transpose, in the mixed launch (items of several
    shapes in one persistent kernel) and in launches of one shape each."""
    say(f"[fp32] tile={tile} trans={trans} mixed={mixed}")
    M, N, K = EDGE_SHAPES
    vb, C0, plan = run_case(bg, orc, M, N, K, trans[0], trans[1], 1.5, 0.5, 1, policy="fixed",
                            fixed_tile=tile, mixed_tiles=mixed)
    kt = plan.kernels()
    assert (kt[:, 1] == bg._lib.TILE_MIXED).any() == (mixed > 0 and tile not in ("16x16", "32x32"))
    oracle_check(orc, vb, 1.5, 0.5, C0)


@pytest.mark.parametrize("xcd", [0, 1])
def test_fp32_auto_config1(bg, orc, gpu, xcd):
    """B=64, M,N,K ~ U{16..256}, A,B ~ U[-1,1], alpha=1, beta=0."""
    from batchgemm_amd import workloads as w
    import torch
    say(f"[fp32] config1 xcd={xcd}")
    M, N, K = w.shapes_uniform(64, 16, 256, 0)
    vb = w.VBatch(M, N, K).fill("uniform", seed=0)
    vb.C.fill_(float("nan"))  # beta = 0 must not read C
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(M, N, K, xcd_queues=xcd)
    for _ in range(3):  # repeated launches reuse the self-resetting queue counters
        plan.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
    torch.cuda.synchronize()
    oracle_check(orc, vb, 1.0, 0.0, C0)


def test_fp32_seq_bit_exact_reference_semantics(bg, orc, gpu):
    """alpha = beta = 1: C <- C + A*B as one fmaf chain per element, exactly
    the reference's correct tiles."""
    from batchgemm_amd.workloads import VBatch
    import torch
    say("[seq] bit-exact")
    M = [32, 64, 17, 100, 1]; N = [32, 96, 5, 33, 1]; K = [8, 64, 13, 257, 1]
    vb = VBatch(M, N, K).fill("normal", seed=5, c_dist="normal")
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(M, N, K, prec="fp32_seq")
    plan.run(vb.dA, vb.dB, vb.dC, 1.0, 1.0)
    torch.cuda.synchronize()
    for b in range(vb.batch):
        m, n, k = M[b], N[b], K[b]
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + m * n]
        ref = orc.ref_tile_seq(False, m, n, k, vb.host_flat("A", b), vb.host_flat("B", b), c0)
        got = vb.host_flat("C", b)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), b


def test_fp32_seq_general_alpha_beta(bg, orc, gpu):
    say("[seq] alpha/beta")
    M, N, K = EDGE_SHAPES
    for trans in ("NN", "TT"):
        vb, C0, _ = run_case(bg, orc, M, N, K, trans[0], trans[1], -0.75, 2.0, 3, prec="fp32_seq")
        oracle_check(orc, vb, -0.75, 2.0, C0)


def test_fp32_zero_k_and_empty(bg, orc, gpu):
    import torch
    from batchgemm_amd.workloads import VBatch
    say("[fp32] K=0 / empty")
    vb = VBatch([5, 70], [6, 130], [0, 0]).fill("normal", seed=2, c_dist="normal")
    C0 = vb.C.clone()
    bg.Plan(vb.M, vb.N, vb.K).run(vb.dA, vb.dB, vb.dC, 3.0, 0.5)
    torch.cuda.synchronize()
    for b in range(2):
        assert torch.allclose(vb.c(b), 0.5 * C0[int(vb.offC[b]):int(vb.offC[b]) + 5 * 6 if b == 0 else int(vb.offC[b]) + 70 * 130].view(vb.N[b], vb.M[b]).t())
    p = bg.Plan([], [], [])
    e = torch.zeros(1, dtype=torch.int64, device="cuda")
    p.run(e, e, e, 1.0, 0.0)


@pytest.mark.parametrize("tile", ["16x16", "32x32", "64x64", "128x128", "16x128", "128x16"])
def test_fp32_split_k_forced(bg, orc, gpu, tile):
    """K cut into slices whose FP32 partial panels are summed by the split-K
    reduction (alpha and beta applied there)."""
    say(f"[fp32] split-K tile={tile}")
    M, N, K = EDGE_SHAPES
    K = [k * 9 + 1 for k in K]
    vb, C0, plan = run_case(bg, orc, M, N, K, "N", "N", 1.5, 0.5, 8, policy="fixed",
                            fixed_tile=tile, force_ksplit=3)
    assert (plan.ksplit() > 1).sum() >= 8
    oracle_check(orc, vb, 1.5, 0.5, C0)


def test_kernel_timing_plan(bg, orc, gpu):
    """A timing plan brackets each kernel with events: one positive duration
    per kernel of query 7, results unchanged, refused under graph capture."""
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd._lib import VbgError
    say("[fp32] kernel timing")
    M, N, K = w.shapes_unbalanced(96, seed=9)
    vb = w.VBatch(M, N, K).fill("normal", seed=9, c_dist="uniform")
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(M, N, K, timing_slots=3, force_ksplit=2)
    plan.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
    torch.cuda.synchronize()
    oracle_check(orc, vb, 1.5, 0.5, C0)
    for _ in range(4):
        plan.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
    t = plan.kernel_times()
    kt = plan.kernels()
    assert len(t) == len(kt) and kt[-1, 0] == 5 and (t > 0).all(), (t, kt)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(VbgError):
        with torch.cuda.graph(g, stream=s):
            plan.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5, stream=s)


@pytest.mark.parametrize("prec,tile", [("fp32", "128x128"), ("fp32", "64x64"), ("bf16x3", "128x128")])
def test_result_independent_of_queue_walk(bg, gpu, prec, tile):
    """The same matrices give the same bits in a launch of at most two items
    per workgroup, where a workgroup whose queue is drained reads every queue
    head at once, and at the head of a large batch, where it claims queue by
    queue: an item's result does not depend on the workgroup that runs it."""
    import torch
    from batchgemm_amd import workloads as w
    say(f"[walk] prec={prec} tile={tile}")
    M0, N0, K0 = w.shapes_uniform(24, 16, 300, seed=7)
    M1, N1, K1 = w.shapes_uniform(1500, 64, 300, seed=8)
    M, N, K = (np.concatenate([a, b]) for a, b in ((M0, M1), (N0, N1), (K0, K1)))
    vb = w.VBatch(M, N, K).fill("normal", seed=3, c_dist="uniform")
    C0 = vb.C.clone()
    opts = dict(prec=prec, policy="fixed", fixed_tile=tile, mixed_tiles=-1, split_k_min=-1)
    small, big = bg.Plan(M0, N0, K0, **opts), bg.Plan(M, N, K, **opts)
    # rows (kind, tile, first item, items, grid); FP32 adds edge-strip launches
    ks, kb = small.kernels(), big.kernels()
    assert (ks[:, 3] <= 2 * ks[:, 4]).all() and kb[0, 3] > 2 * kb[0, 4], (ks, kb)
    small.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
    torch.cuda.synchronize()
    got_small = [vb.host_flat("C", b) for b in range(len(M0))]
    vb.C.copy_(C0)
    big.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
    torch.cuda.synchronize()
    for b in range(len(M0)):
        assert np.array_equal(got_small[b].view(np.int32), vb.host_flat("C", b).view(np.int32)), b
