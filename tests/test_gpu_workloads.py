"""GPU parity at the benchmark workloads' own sizes (needs an MI355X).

The unbalanced batch: B = 2048 column-major GEMMs, 90% with M,N,K ~ U{16..128}
and 10% with M,N,K ~ U{512..2048} (the plan the benchmark times: 24,750 work
items in a 128x128 launch and a mixed-shape launch), alpha = 1.5, beta = 0.5,
C ~ U[-1,1], A,B ~ N(0,1) with 1% of the entries scaled by 1e3 (the precision
workload's inputs).  Every path runs on that batch -- FP32, bf16x3, FP16 with
per-matrix 2^e scaling and FP8 e4m3 with per-matrix amax scaling, the last
two both on packed operands and quantised in the k-loop -- and is checked on
every large matrix (all of M,N,K in 512..2048) plus 64 small ones, each
against FP64 products of the same inputs (numpy float64 on the host, the
product that pins the oracle's orc_dgemm).  (This batch needs no split-K:
every launch has far more items than workgroup slots; split-K is covered by
the forced-split tests of every tile.)

Bounds, per matrix, with ref = alpha*A*B + beta*C0, R = ||alpha |A| |B|||_F /
||ref||_F and u = 2^-24:
  * FP32: 1e-5 (the target; FP32 accumulation gives about u*sqrt(K)*R).
  * bf16x3: the format's own error -- lo = bf16(x - hi) keeps 16 significant
    bits of x, so each product carries at most 2*2^-18 from the two lo
    roundings plus the dropped lo*lo term (<= 2^-18) -- plus FP32
    accumulation: (3*2^-18 + 4*u*sqrt(K)) * R.  A small matrix dominated by a
    few 1e3-scaled products can exceed 1e-5 on the format alone (about 1% of
    21 x 38 x 115 outlier matrices do, emulated exactly); the path's target,
    <= 1e-5, is asserted on the pooled error over all checked matrices.
  * FP16 against the oracle's emulation of its quantisation (same quantised
    operands multiplied exactly in FP64, times the kernels' post-scale): the
    FP16 MFMA forms exact products and accumulates in FP32 (crafted sums on
    the GPU round exactly as FP32 does), so the difference is FP32 summation,
    bounded by the probabilistic bound (Higham & Mary 2019) 4*u*sqrt(K)*R_q,
    R_q as R on the quantised operands (R is used: the two differ by the
    quantisation only).
  * FP8 against the emulation: the gfx950 FP8 MFMA sums each lane's 8
    products at reduced precision before its FP32 accumulation -- crafted sums
    lose a product of 2^-16 times the group's largest (and seven of them, 2^-13
    together) while the same products in another group, or across MFMAs,
    add exactly as FP32 -- so each group of 8 can be off by up to about 2^-13
    of its largest product, with signs that vary.  Per element that is at most
    about 2^-13 * sqrt(sum_k (a_ik b_kj)^2) in the root-sum-square sense, and
    the bound is 4 * 2^-13 * sqrt(sum((A*A) @ (B*B))) / ||ref||_F -- K-free, as
    the measured error is (about 1.5e-5 relative on N(0,1) inputs at K = 16 ..
    2048, against a bound of about 5e-4 there).
  * FP16 / FP8 against FP64: the formats' own rounding, bounded loosely (2e-3,
    0.2) and printed.

The batching workload: one batch of 50 (B ~ U{256..8192}, M,N,K log-uniform
in [8, 1024], U[-1,1] entries) in bf16x3, every matrix against FP64 with the
bf16x3 bound above, and <= 1e-5 pooled.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALPHA, BETA = 1.5, 0.5
U = 2.0 ** -24


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16x3_bound(K, R):
    return (3 * 2.0 ** -18 + 4 * U * np.sqrt(float(K))) * R


class Unbalanced:
    """The unbalanced batch on the GPU, the selected matrices' inputs on the
    host, and their FP64 references (computed once, shared by the paths)."""

    def __init__(self):
        import torch
        from batchgemm_amd import workloads as w
        self.M, self.N, self.K = w.shapes_unbalanced(2048, seed=0)
        self.vb = w.VBatch(self.M, self.N, self.K).fill(
            "normal", seed=3, outlier_frac=0.01, outlier_scale=1e3, c_dist="uniform")
        self.C0 = self.vb.C.clone()
        big = np.where((self.M >= 512) & (self.N >= 512) & (self.K >= 512))[0]
        rng = np.random.default_rng(5)
        small = rng.choice(np.where(self.M < 512)[0], 64, replace=False)
        self.sel = np.concatenate([big, np.sort(small)])
        self.nbig = len(big)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.A, self.B, self.ref, self.R, self.sq = {}, {}, {}, {}, {}
        self.c0 = {}
        for b in self.sel:
            A = self.vb.a(b).cpu().numpy().astype(np.float64)
            B = self.vb.b(b).cpu().numpy().astype(np.float64)
            self.A[b], self.B[b] = A.astype(np.float32), B.astype(np.float32)
            self.c0[b] = self.vb.c(b).cpu().numpy().astype(np.float64)
            ref = ALPHA * (A @ B) + BETA * self.c0[b]
            nref = np.linalg.norm(ref)
            self.ref[b] = ref
            self.R[b] = ALPHA * np.linalg.norm(np.abs(A) @ np.abs(B)) / nref
            self.sq[b] = ALPHA * np.sqrt(np.sum((A * A) @ (B * B))) / nref
        say(f"[workload] unbalanced B=2048: {self.nbig} large + {len(small)} small matrices checked; "
            f"FP64 references {time.perf_counter() - t0:.1f} s")

    def run(self, prec, **opts):
        import torch
        import batchgemm_amd as bg
        self.vb.C.copy_(self.C0)
        plan = bg.Plan(self.M, self.N, self.K, prec=prec, **opts)
        kinds = [bg.Plan.KERNEL_KINDS[int(k)] for k in plan.kernels()[:, 0]]
        say(f"[workload] {prec} {opts}: {plan.n_items} items, kernels {kinds}")
        plan.run(self.vb.dA, self.vb.dB, self.vb.dC, ALPHA, BETA)
        torch.cuda.synchronize()
        return plan

    def got(self, b):
        return self.vb.c(b).cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def unbalanced(bg, gpu):
    w = Unbalanced()
    yield w
    del w.vb


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_unbalanced_batch_fp32_and_bf16x3_vs_fp64(unbalanced, prec):
    w = unbalanced
    plan = w.run(prec)
    if prec == "fp32":  # the benchmark's plan: every tile shape in one mixed launch
        assert plan.n_launches == 1 and plan.n_items > 20000
        from batchgemm_amd import _lib
        assert int(plan.kernels()[0, 1]) == _lib.TILE_MIXED
    num = den = 0.0
    worst = (0.0, None, 0.0)
    for b in w.sel:
        d = np.linalg.norm(w.got(b) - w.ref[b]); n = np.linalg.norm(w.ref[b])
        num += d * d; den += n * n
        e = d / n
        bound = 1e-5 if prec == "fp32" else bf16x3_bound(w.K[b], w.R[b])
        if e / bound > worst[0]:
            worst = (e / bound, b, e)
        assert e <= bound, (prec, int(b), int(w.M[b]), int(w.N[b]), int(w.K[b]), e, bound)
    pooled = (num / den) ** 0.5
    r, b, e = worst
    say(f"  {prec}: pooled rel err vs FP64 {pooled:.3g}; worst matrix {b} ({w.M[b]}x{w.N[b]}x{w.K[b]}) "
        f"err {e:.3g} = {r:.3g} of its bound")
    assert pooled <= 1e-5


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
@pytest.mark.parametrize("pack", [1, -1], ids=["packed", "in_loop"])
def test_unbalanced_batch_scaled_paths_vs_emulation(orc, unbalanced, prec, pack):
    w = unbalanced
    plan = w.run(prec, pack_operands=pack)
    assert (plan.packed_bytes > 0) == (pack > 0)
    worst = (0.0, None, 0.0, 0.0)
    worst_fp64 = 0.0
    for b in w.sel:
        got = w.got(b)
        emul = ALPHA * orc.split_product(prec, w.A[b], w.B[b]) + BETA * w.c0[b]
        e = float(np.linalg.norm(got - emul) / np.linalg.norm(emul))
        if prec == "fp16":
            bound = 4 * U * np.sqrt(float(w.K[b])) * w.R[b]
        else:
            bound = 4 * 2.0 ** -13 * w.sq[b]
        if e / bound > worst[0]:
            worst = (e / bound, b, e, bound)
        assert e <= bound, (prec, int(b), int(w.M[b]), int(w.N[b]), int(w.K[b]), e, bound)
        e64 = float(np.linalg.norm(got - w.ref[b]) / np.linalg.norm(w.ref[b]))
        worst_fp64 = max(worst_fp64, e64)
        assert e64 <= {"fp16": 2e-3, "fp8": 0.2}[prec], (prec, int(b), e64)
    r, b, e, bound = worst
    say(f"  {prec} pack={pack}: vs emulation, worst matrix {b} ({w.M[b]}x{w.N[b]}x{w.K[b]}) err {e:.3g} "
        f"= {r:.3g} of its bound {bound:.3g}; max vs FP64 {worst_fp64:.3g}")


def test_batching_workload_bf16x3_vs_fp64(bg, gpu):
    """One batch of the batching workload (the first of the 50, seed 0) in
    bf16x3 on its default plan, every matrix against FP64."""
    import torch
    from batchgemm_amd import workloads as w
    M, N, K = w.autotune_batches(50, seed=0)[0]
    vb = w.VBatch(M, N, K).fill("uniform", seed=7, c_dist="uniform")
    C0 = vb.C.clone()
    plan = bg.Plan(M, N, K, prec="bf16x3")
    say(f"[workload] batching batch 0: B={len(M)}, {plan.n_items} items, {plan.n_launches} launches")
    plan.run(vb.dA, vb.dB, vb.dC, ALPHA, BETA)
    torch.cuda.synchronize()
    Ah, Bh, Ch, C0h = vb.A.cpu().numpy(), vb.B.cpu().numpy(), vb.C.cpu().numpy(), C0.cpu().numpy()
    num = den = worst = 0.0
    for b in range(len(M)):
        m, n, k = int(M[b]), int(N[b]), int(K[b])
        A = Ah[vb.offA[b]: vb.offA[b] + m * k].reshape(k, m).T.astype(np.float64)
        B = Bh[vb.offB[b]: vb.offB[b] + k * n].reshape(n, k).T.astype(np.float64)
        c0 = C0h[vb.offC[b]: vb.offC[b] + m * n].reshape(n, m).T.astype(np.float64)
        got = Ch[vb.offC[b]: vb.offC[b] + m * n].reshape(n, m).T.astype(np.float64)
        ref = ALPHA * (A @ B) + BETA * c0
        d, nr = np.linalg.norm(got - ref), np.linalg.norm(ref)
        num += d * d; den += nr * nr
        R = ALPHA * np.linalg.norm(np.abs(A) @ np.abs(B)) / nr
        worst = max(worst, d / nr)
        assert d / nr <= bf16x3_bound(k, R), (b, m, n, k, d / nr)
    pooled = (num / den) ** 0.5
    say(f"  bf16x3 batching batch: pooled rel err vs FP64 {pooled:.3g}, worst matrix {worst:.3g} "
        f"over {len(M)} matrices")
    assert pooled <= 1e-5
