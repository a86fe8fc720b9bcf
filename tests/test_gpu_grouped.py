"""Grouped GEMM forward/backward on the GPU (needs an MI355X).

Forward and both gradients are checked against FP64 products of the same BF16
values (<= 1e-5 relative with FP32 outputs, which is the FP32-accumulation
bound; BF16 outputs add one BF16 rounding, <= 4e-3), against FP64 central
finite differences of L = sum(C * dC) on a 4-group subset (M_g <= 64,
K = N = 128), and against a looped torch.matmul."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(m_sizes, K, N, seed=0, std=0.02):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    S = int(sum(m_sizes)); G = len(m_sizes)
    A = (torch.randn(S, K, generator=g, device="cuda") * std).to(torch.bfloat16)
    B = (torch.randn(G, K, N, generator=g, device="cuda") * std).to(torch.bfloat16)
    dC = torch.randn(S, N, generator=g, device="cuda").to(torch.bfloat16)
    return A, B, dC


def ref64(A, B, dC, m_sizes):
    A = A.double().cpu().numpy(); B = B.double().cpu().numpy(); dC = dC.double().cpu().numpy()
    off = np.concatenate([[0], np.cumsum(m_sizes)])
    C = np.zeros((A.shape[0], B.shape[2])); dA = np.zeros_like(A); dB = np.zeros_like(B)
    for g in range(len(m_sizes)):
        s, e = off[g], off[g + 1]
        C[s:e] = A[s:e] @ B[g]
        dA[s:e] = dC[s:e] @ B[g].T
        dB[g] = A[s:e].T @ dC[s:e]
    return C, dA, dB


def rel(x, r):
    x = np.asarray(x, np.float64); d = np.linalg.norm(r)
    return float(np.linalg.norm(x - r) / (d if d > 0 else 1.0))


@pytest.mark.parametrize("m_sizes", [[5, 0, 64, 1, 130, 33, 200, 7], [64, 17, 3, 48],
                                     [1, 2, 16, 0, 9, 3, 16, 1]])
@pytest.mark.parametrize("KN", [(128, 128), (256, 72), (64, 320)])
def test_grouped_fwd_bwd_vs_fp64(bg, gpu, m_sizes, KN):
    import torch
    from batchgemm_amd.grouped import GroupedPlan
    K, N = KN
    say(f"[grouped] m={m_sizes} K={K} N={N}")
    A, B, dC = make(m_sizes, K, N, seed=len(m_sizes) + K)
    p = GroupedPlan(m_sizes, K, N)
    C32 = p.forward(A, B, out_dtype=torch.float32)
    C16 = p.forward(A, B, out_dtype=torch.bfloat16)
    dA, dB = p.backward(A, B, dC, grad_dtype=torch.float32)
    dA16, dB16 = p.backward(A, B, dC, grad_dtype=torch.bfloat16)
    torch.cuda.synchronize()
    C, rA, rB = ref64(A, B, dC, m_sizes)
    assert rel(C32.cpu().numpy(), C) <= 1e-5
    assert rel(dA.cpu().numpy(), rA) <= 1e-5
    assert rel(dB.cpu().numpy(), rB) <= 1e-5
    assert rel(C16.float().cpu().numpy(), C) <= 4e-3
    assert rel(dA16.float().cpu().numpy(), rA) <= 4e-3
    assert rel(dB16.float().cpu().numpy(), rB) <= 4e-3
    for g, m in enumerate(m_sizes):  # empty groups: exact zero gradient
        if m == 0:
            assert torch.count_nonzero(dB[g]) == 0
        elif m <= 16:  # direct small-group dB: exact products, FP32 sums
            assert rel(dB[g].cpu().numpy(), rB[g]) <= 1e-6, (g, m)


def test_grouped_finite_differences_fp64(bg, gpu):
    """Directional FP64 central differences of L(A, B) = sum(C * dC) on a
    4-group subset: L is bilinear, so the difference quotient is the exact
    directional derivative up to FP64 rounding."""
    import torch
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] finite differences")
    m_sizes = [64, 1, 37, 20]
    K = N = 128
    A, B, dC = make(m_sizes, K, N, seed=3)
    dA, dB = GroupedPlan(m_sizes, K, N).backward(A, B, dC, grad_dtype=torch.float32)
    torch.cuda.synchronize()
    A64, B64, dC64 = (t.double().cpu().numpy() for t in (A, B, dC))
    off = np.concatenate([[0], np.cumsum(m_sizes)])

    def L(a, b):
        return sum(float(np.sum((a[off[g]:off[g + 1]] @ b[g]) * dC64[off[g]:off[g + 1]]))
                   for g in range(len(m_sizes)))

    rng = np.random.default_rng(0)
    eps = 1e-3
    for _ in range(4):
        VA = rng.standard_normal(A64.shape); VB = rng.standard_normal(B64.shape)
        fd_a = (L(A64 + eps * VA, B64) - L(A64 - eps * VA, B64)) / (2 * eps)
        fd_b = (L(A64, B64 + eps * VB) - L(A64, B64 - eps * VB)) / (2 * eps)
        ga = dA.double().cpu().numpy(); gb = dB.double().cpu().numpy()
        got_a = float(np.sum(ga * VA)); got_b = float(np.sum(gb * VB))
        # <E, V> for a gradient error E and a unit-variance random direction V
        # has standard deviation ||E||_F: allow 4 sigma of a 1e-5 relative error
        assert abs(got_a - fd_a) <= 4e-5 * np.linalg.norm(ga), (got_a, fd_a)
        assert abs(got_b - fd_b) <= 4e-5 * np.linalg.norm(gb), (got_b, fd_b)
        assert abs(fd_a) > 1e-3 * np.linalg.norm(ga)  # the direction is not degenerate


def test_grouped_autograd_vs_looped_matmul(bg, gpu):
    import torch
    from batchgemm_amd.grouped import grouped_gemm
    say("[grouped] autograd vs looped matmul")
    m_sizes = [96, 0, 31, 128, 5]
    K, N = 192, 136
    A, B, dC = make(m_sizes, K, N, seed=7)
    a = A.clone().requires_grad_(True); b = B.clone().requires_grad_(True)
    C = grouped_gemm(a, b, m_sizes, out_dtype=torch.float32, grad_dtype=torch.float32)
    C.backward(dC.float())
    # looped torch.matmul in float32 on the same BF16 values
    a2 = A.float().requires_grad_(True); b2 = B.float().requires_grad_(True)
    outs, s = [], 0
    for g, m in enumerate(m_sizes):
        outs.append(a2[s:s + m] @ b2[g]); s += m
    C2 = torch.cat(outs)
    C2.backward(dC.float())
    # C is FP32; the gradients come back in the inputs' dtype (BF16), so they
    # carry one BF16 rounding (2^-9 relative) on top of FP32 accumulation
    assert rel(C.detach().float().cpu().numpy(), C2.detach().double().cpu().numpy()) <= 1e-5
    for x, y in [(a.grad, a2.grad), (b.grad, b2.grad)]:
        assert x.dtype == torch.bfloat16
        assert rel(x.float().cpu().numpy(), y.double().cpu().numpy()) <= 4e-3
        want = y.to(torch.bfloat16).float()
        assert rel(x.float().cpu().numpy(), want.double().cpu().numpy()) <= 4e-3


def test_grouped_config4_vs_fp64(bg, gpu):
    """G=128, K=2048, N=1408, Zipf(1.2) group sizes summing to 32768 (one
    group of ~32.6k rows, most of 1 row): C, dA, dB of every group against
    FP64 products of the same BF16 values (torch float64 on the GPU).  The
    largest group's dB reduces over ~3.3e4 rows and is split along K; FP32
    accumulation over reductions that long gives ~sqrt(n)*2^-24 ~ 1e-5, so
    the bar is 5e-5."""
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] config4")
    m_sizes = w.zipf_groups(128, 32768, 1.2, seed=0)
    assert m_sizes.sum() == 32768 and m_sizes.min() >= 1
    K, N = 2048, 1408
    A, B, dC = make(m_sizes, K, N, seed=0)
    p = GroupedPlan(m_sizes, K, N)
    C = p.forward(A, B, out_dtype=torch.float32)
    dA, dB = p.backward(A, B, dC, grad_dtype=torch.float32)
    torch.cuda.synchronize()
    off = np.concatenate([[0], np.cumsum(m_sizes)])
    A64, B64, dC64 = A.double(), B.double(), dC.double()
    worst = 0.0
    for g in range(128):
        s, e = int(off[g]), int(off[g + 1])
        for got, ref in [(C[s:e], A64[s:e] @ B64[g]), (dA[s:e], dC64[s:e] @ B64[g].T),
                         (dB[g], A64[s:e].T @ dC64[s:e])]:
            err = float(torch.linalg.norm(got.double() - ref) / torch.linalg.norm(ref))
            worst = max(worst, err)
            assert err <= 5e-5, (g, int(m_sizes[g]), err)
    say(f"  worst rel err {worst:.3g}")


def test_grouped_forced_split_k(bg, gpu):
    import torch
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] forced split-K")
    m_sizes = [300, 0, 64, 1, 129]
    K, N = 256, 136
    A, B, dC = make(m_sizes, K, N, seed=5)
    p = GroupedPlan(m_sizes, K, N, force_ksplit=3)
    C32 = p.forward(A, B, out_dtype=torch.float32)
    C16 = p.forward(A, B, out_dtype=torch.bfloat16)
    dA, dB = p.backward(A, B, dC, grad_dtype=torch.float32)
    torch.cuda.synchronize()
    C, rA, rB = ref64(A, B, dC, m_sizes)
    assert rel(C32.cpu().numpy(), C) <= 1e-5
    assert rel(C16.float().cpu().numpy(), C) <= 4e-3
    assert rel(dA.cpu().numpy(), rA) <= 1e-5
    assert rel(dB.cpu().numpy(), rB) <= 1e-5


def test_grouped_small_groups_every_bucket(bg, gpu):
    """Many groups of 0..16 rows: every row bucket (1, 2, 4, 8, 16) of the
    small-group kernels holds enough groups that the forward streams them too
    (N = 1280: 5 column blocks per group), with a few big groups on the GEMM
    plans beside them.  C, dA, dB against FP64, FP32 and BF16 outputs."""
    import torch
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] small groups in every bucket")
    rng = np.random.default_rng(9)
    m_sizes = [int(x) for x in rng.integers(1, 17, size=300)] + [0, 0, 40, 130, 0]
    rng.shuffle(m_sizes)
    K, N = 72, 1280
    A, B, dC = make(m_sizes, K, N, seed=9)
    p = GroupedPlan(m_sizes, K, N)
    C32 = p.forward(A, B, out_dtype=torch.float32)
    C16 = p.forward(A, B, out_dtype=torch.bfloat16)
    dA, dB = p.backward(A, B, dC, grad_dtype=torch.float32)
    dA16, dB16 = p.backward(A, B, dC, grad_dtype=torch.bfloat16)
    torch.cuda.synchronize()
    C, rA, rB = ref64(A, B, dC, m_sizes)
    assert rel(C32.cpu().numpy(), C) <= 1e-5
    assert rel(C16.float().cpu().numpy(), C) <= 4e-3
    assert rel(dA.cpu().numpy(), rA) <= 1e-5
    assert rel(dB.cpu().numpy(), rB) <= 1e-5
    assert rel(dA16.float().cpu().numpy(), rA) <= 4e-3
    assert rel(dB16.float().cpu().numpy(), rB) <= 4e-3
    off = np.concatenate([[0], np.cumsum(m_sizes)])
    for g in np.where(np.asarray(m_sizes) == 0)[0]:  # empty groups: zero dB_g
        assert float(dB[g].abs().max()) == 0.0
    for g in range(len(m_sizes)):  # per-group check catches a bucket that is off
        s, e = int(off[g]), int(off[g + 1])
        if e > s:
            assert rel(C32[s:e].cpu().numpy(), C[s:e]) <= 1e-5, (g, m_sizes[g])


def test_grouped_small_groups_ragged_streams(bg, gpu):
    """The streaming small-group kernels at ragged sizes: N = 1000 leaves the
    forward's last 256-column block with 232 columns and dA's lanes with 3 or 4
    column steps; K = 1096 gives the forward a second staged chunk of 72 k
    (1024 + 72) and dA a last row block of 8 rows.  The 1- and 2-row buckets
    hold enough groups (80 x 4 and 70 x 4 blocks) that the forward streams
    them too.  C, dA, dB against FP64, per group."""
    import torch
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] ragged streams K=1096 N=1000")
    m_sizes = [1] * 80 + [2] * 70 + [40]
    rng = np.random.default_rng(11)
    rng.shuffle(m_sizes)
    m_sizes = [int(x) for x in m_sizes]
    K, N = 1096, 1000
    A, B, dC = make(m_sizes, K, N, seed=11)
    p = GroupedPlan(m_sizes, K, N)
    C32 = p.forward(A, B, out_dtype=torch.float32)
    dA, dB = p.backward(A, B, dC, grad_dtype=torch.float32)
    torch.cuda.synchronize()
    C, rA, rB = ref64(A, B, dC, m_sizes)
    off = np.concatenate([[0], np.cumsum(m_sizes)])
    C32, dA, dB = C32.cpu().numpy(), dA.cpu().numpy(), dB.cpu().numpy()
    for g in range(len(m_sizes)):
        s, e = int(off[g]), int(off[g + 1])
        assert rel(C32[s:e], C[s:e]) <= 1e-5, ("C", g, m_sizes[g])
        assert rel(dA[s:e], rA[s:e]) <= 1e-5, ("dA", g, m_sizes[g])
        assert rel(dB[g], rB[g]) <= 1e-5, ("dB", g, m_sizes[g])


def test_grouped_graph_replays_match_eager(bg, gpu):
    """The differentiation workload's passes -- few-row groups on the plan's
    side stream beside the large group's GEMM, forked from and joined to the
    caller's stream -- captured into a HIP graph: every replay equals the
    eager run bit for bit, with eager work and synchronisations between
    replays."""
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.grouped import GroupedPlan
    say("[grouped] graph replays, config4 shapes")
    m_sizes = w.zipf_groups(128, 32768, 1.2, seed=0)
    K, N = 2048, 1408
    A, B, dC = make(m_sizes, K, N, seed=3)
    p = GroupedPlan(m_sizes, K, N)
    C_e = p.forward(A, B)
    dA_e, dB_e = p.backward(A, B, dC)
    torch.cuda.synchronize()
    Cg = torch.empty_like(C_e)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p.forward(A, B, out=Cg)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p.forward(A, B, out=Cg)
        dAg, dBg = p.backward(A, B, dC)
    for rep in range(3):
        for t in (Cg, dAg, dBg):
            t.zero_()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for name, got, ref in (("C", Cg, C_e), ("dA", dAg, dA_e), ("dB", dBg, dB_e)):
            assert torch.equal(got.view(torch.int16), ref.view(torch.int16)), (name, rep)
