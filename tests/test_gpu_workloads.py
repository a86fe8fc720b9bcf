"""GPU parity at the benchmark workloads' own sizes (needs an MI355X).

The unbalanced batch: B = 2048 column-major GEMMs, 90% with M,N,K ~ U{16..128}
and 10% with M,N,K ~ U{512..2048} (the plan the benchmark times: about 24,750
work items, split-K on the largest matrices, several launches), alpha = 1.5,
beta = 0.5, C ~ U[-1,1], A,B ~ N(0,1) with 1% of the entries scaled by 1e3
(the precision workload's inputs).  Every path runs on that batch -- FP32,
bf16x3, FP16 with per-matrix 2^e scaling and FP8 e4m3 with per-matrix amax
scaling, the last two both on packed operands and quantised in the k-loop --
and is checked on every large matrix (all of M,N,K in 512..2048) plus 64
small ones, each against FP64 products of the same inputs (numpy float64 on
the host, the product that pins the oracle's orc_dgemm).

Bounds.  FP32 and bf16x3: relative Frobenius error <= 1e-5 against FP64 (the
path's target).  FP16 and FP8 against the oracle's emulation of their
operand quantisation (the same quantised operands multiplied exactly in FP64,
times the kernels' post-scale): the only difference left is the MFMA's FP32
accumulation, bounded per matrix by the probabilistic summation bound
(Higham & Mary 2019)

    err <= c * u * sqrt(K) * rho,  u = 2^-24,  rho = ||(|A| |B|)||_F / ||A B||_F

(rho computed from the FP64 products of the unquantised operands), with c = 4
for FP16 and c = 64 for FP8: the gfx950 FP8 MFMA's own accumulation differs
from exact-product FP32 summation by about 16x the FP32 figure (measured at K
<= 257, where the identity test in test_gpu_precision.py shows the operand
quantisation itself exact element by element).  At K = 2048 with N(0,1)
operands rho is about 29, so the bound is about 3.1e-4 (FP16) and 5e-3 (FP8);
the measured errors are printed next to it.  Against FP64 the formats' own
rounding is bounded loosely (FP16 2e-3, FP8 0.2) and printed.

The batching workload: one batch of 50 (B ~ U{256..8192}, M,N,K log-uniform
in [8, 1024], U[-1,1] entries) in bf16x3, every matrix against FP64, <= 1e-5.
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
        self.A, self.B, self.c0, self.exact, self.rho = {}, {}, {}, {}, {}
        for b in self.sel:
            A = self.vb.a(b).cpu().numpy().astype(np.float64)
            B = self.vb.b(b).cpu().numpy().astype(np.float64)
            self.A[b], self.B[b] = A.astype(np.float32), B.astype(np.float32)
            self.c0[b] = self.vb.c(b).cpu().numpy().astype(np.float64)
            P = A @ B
            self.exact[b] = P
            self.rho[b] = np.linalg.norm(np.abs(A) @ np.abs(B)) / max(np.linalg.norm(P), 1e-300)
        say(f"[workload] unbalanced B=2048: {self.nbig} large + {len(small)} small matrices checked; "
            f"FP64 references {time.perf_counter() - t0:.1f} s")

    def run(self, prec, **opts):
        import torch
        import batchgemm_amd as bg
        self.vb.C.copy_(self.C0)
        plan = bg.Plan(self.M, self.N, self.K, prec=prec, **opts)
        kinds = [bg.Plan.KERNEL_KINDS[int(k)] for k in plan.kernels()[:, 0]]
        say(f"[workload] {prec} {opts}: {plan.n_items} items, kernels {kinds}, "
            f"{int((plan.ksplit() > 1).sum())} split-K matrices")
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


def check_vs_fp64(w, prec, bound):
    worst = (0.0, None)
    for b in w.sel:
        ref = ALPHA * w.exact[b] + BETA * w.c0[b]
        e = float(np.linalg.norm(w.got(b) - ref) / np.linalg.norm(ref))
        if e > worst[0]:
            worst = (e, b)
        assert e <= bound, (prec, int(b), int(w.M[b]), int(w.N[b]), int(w.K[b]), e)
    b = worst[1]
    say(f"  {prec}: max rel err vs FP64 {worst[0]:.3g} (matrix {b}: "
        f"{w.M[b]}x{w.N[b]}x{w.K[b]}), bound {bound:g}")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_unbalanced_batch_fp32_and_bf16x3_vs_fp64(unbalanced, prec):
    plan = unbalanced.run(prec)
    if prec == "fp32":  # the benchmark's plan: split-K reduction and several launches
        assert (plan.ksplit() > 1).any() and plan.n_launches >= 2
        assert plan.n_items > 20000
    check_vs_fp64(unbalanced, prec, 1e-5)


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
@pytest.mark.parametrize("pack", [1, -1], ids=["packed", "in_loop"])
def test_unbalanced_batch_scaled_paths_vs_emulation(orc, unbalanced, prec, pack):
    w = unbalanced
    plan = w.run(prec, pack_operands=pack)
    assert (plan.packed_bytes > 0) == (pack > 0)
    c = {"fp16": 4.0, "fp8": 64.0}[prec]
    worst_ratio = (0.0, None, 0.0, 0.0)
    worst_fp64 = 0.0
    for b in w.sel:
        got = w.got(b)
        emul = ALPHA * orc.split_product(prec, w.A[b], w.B[b]) + BETA * w.c0[b]
        e = float(np.linalg.norm(got - emul) / np.linalg.norm(emul))
        bound = c * U * np.sqrt(float(w.K[b])) * w.rho[b]
        if e / bound > worst_ratio[0]:
            worst_ratio = (e / bound, b, e, bound)
        assert e <= bound, (prec, int(b), int(w.M[b]), int(w.N[b]), int(w.K[b]), e, bound)
        ref = ALPHA * w.exact[b] + BETA * w.c0[b]
        e64 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        worst_fp64 = max(worst_fp64, e64)
        assert e64 <= {"fp16": 2e-3, "fp8": 0.2}[prec], (prec, int(b), e64)
    r, b, e, bound = worst_ratio
    say(f"  {prec} pack={pack}: worst err/bound {r:.3g} vs emulation (matrix {b}: "
        f"{w.M[b]}x{w.N[b]}x{w.K[b]}, err {e:.3g}, bound {bound:.3g}); max vs FP64 {worst_fp64:.3g}")


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
    worst = 0.0
    for b in range(len(M)):
        m, n, k = int(M[b]), int(N[b]), int(K[b])
        A = Ah[vb.offA[b]: vb.offA[b] + m * k].reshape(k, m).T.astype(np.float64)
        B = Bh[vb.offB[b]: vb.offB[b] + k * n].reshape(n, k).T.astype(np.float64)
        c0 = C0h[vb.offC[b]: vb.offC[b] + m * n].reshape(n, m).T.astype(np.float64)
        got = Ch[vb.offC[b]: vb.offC[b] + m * n].reshape(n, m).T.astype(np.float64)
        ref = ALPHA * (A @ B) + BETA * c0
        e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        worst = max(worst, e)
        assert e <= 1e-5, (b, m, n, k, e)
    say(f"  bf16x3 batching batch: max rel err vs FP64 {worst:.3g} over {len(M)} matrices")
