"""GPU parity of the split-precision paths (needs an MI355X).

Each path is checked twice:
  * against a numpy emulation of its operand quantisation (bf16 hi/lo split,
    FP16 with the per-matrix power-of-two scale, OCP e4m3 with the 448/amax
    scale; the rounding helpers are pinned to torch's dtype conversions in
    test_oracle.py) evaluated in FP64 -- this pins the kernel's quantisation,
    scale and MFMA operand layout; the only difference left is accumulation
    inside the MFMA: 2e-6 for bf16x3 / FP16 (FP32 accumulation of exact
    products) and, for FP8, a bound from the FP8 MFMA's own summation (each
    lane's 8 products are summed at reduced precision, up to about 2^-13 of
    the group's largest product, before the FP32 accumulation; see
    emul_bound) -- its operand quantisation matches the emulation element by
    element (the identity test below);
  * against the FP64 oracle on the unquantised operands: <= 1e-5 for bf16x3
    (the target of the path); FP16 and FP8 errors are printed and bounded
    loosely (they are what the formats give).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def op_matrix(vb, which, b):
    flat = vb.host_flat(which, b)
    if which == "A":
        rows, cols, ld, t = int(vb.a_rows[b]), int(vb.a_cols[b]), int(vb.lda[b]), vb.transA
    else:
        rows, cols, ld, t = int(vb.b_rows[b]), int(vb.b_cols[b]), int(vb.ldb[b]), vb.transB
    X = flat.reshape(cols, ld)[:, :rows].T if cols else np.zeros((rows, 0), np.float32)
    return X.T if t == "T" else X


def emulate(orc, prec, A, B):
    """Quantised operands and the post-scale, as the kernel builds them."""
    return orc.split_product(prec, A, B)


FP64_BOUND = {"bf16x3": 1e-5, "fp16": 2e-3, "fp8": 0.2}
EMUL_BOUND = {"bf16x3": 2e-6, "fp16": 2e-6}


def emul_bound(prec, A, B, ref, alpha=1.0):
    """Bound on the relative difference from the emulation.  bf16x3 / FP16:
    exact products summed in FP32 (2e-6 covers K <= 257 here).  FP8: the FP8
    MFMA sums each lane's 8 products at reduced precision (up to about 2^-13
    of the group's largest product), so the bound is
    4 * 2^-13 * alpha * sqrt(sum((A*A) @ (B*B))) / ||ref||_F."""
    if prec != "fp8":
        return EMUL_BOUND[prec]
    A = np.asarray(A, np.float64); B = np.asarray(B, np.float64)
    nr = np.linalg.norm(ref)
    return 4 * 2.0 ** -13 * alpha * np.sqrt(np.sum((A * A) @ (B * B))) / (nr if nr > 0 else 1.0)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16", "fp8"])
@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("outliers", [0.0, 0.01])
def test_split_precision_paths(bg, orc, gpu, prec, trans, outliers):
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[{prec}] trans={trans} outliers={outliers}")
    M = [64, 100, 17, 130, 1, 255, 96]
    N = [64, 33, 140, 128, 7, 64, 1]
    K = [32, 200, 9, 257, 5, 100, 64]
    vb = VBatch(M, N, K, trans[0], trans[1]).fill("normal", seed=9, outlier_frac=outliers,
                                                   c_dist="uniform")
    C0 = vb.C.cpu().numpy()
    alpha, beta = 1.5, 0.5
    plan = bg.Plan(vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], prec=prec)
    plan.run(vb.dA, vb.dB, vb.dC, alpha, beta)
    torch.cuda.synchronize()
    worst_emul = worst_fp64 = 0.0
    for b in range(vb.batch):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + M[b] * N[b]].reshape(N[b], M[b]).T.astype(np.float64)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T.astype(np.float64)
        emul = alpha * emulate(orc, prec, A, B) + beta * c0
        exact = alpha * (A.astype(np.float64) @ B.astype(np.float64)) + beta * c0
        e1 = orc.rel_fro(got, emul); e2 = orc.rel_fro(got, exact)
        worst_emul = max(worst_emul, e1); worst_fp64 = max(worst_fp64, e2)
        assert e1 <= emul_bound(prec, A, B, emul, alpha), (b, M[b], N[b], K[b], e1)
        assert e2 <= FP64_BOUND[prec], (b, e2)
    say(f"  max rel err vs emulation {worst_emul:.3g}, vs FP64 {worst_fp64:.3g}")


def test_split_repeated_launch_and_plan_reuse(bg, orc, gpu):
    """The amax pre-pass re-zeroes its accumulators every launch: scaling the
    inputs down between launches must change the scales."""
    import torch
    from batchgemm_amd.workloads import VBatch
    say("[fp8] plan reuse")
    M, N, K = [70, 64], [50, 64], [90, 64]
    vb = VBatch(M, N, K).fill("normal", seed=4)
    plan = bg.Plan(M, N, K, prec="fp8")
    plan.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
    vb.A.mul_(1e-3)
    plan.run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
    torch.cuda.synchronize()
    for b in range(2):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T
        emul = emulate(orc, "fp8", A, B)
        assert orc.rel_fro(got, emul) <= emul_bound("fp8", A, B, emul)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16", "fp8"])
def test_quantisation_elementwise_via_identity(bg, orc, gpu, prec):
    """C = A * I: every output is one product (plus exact zeros), so C is the
    kernel's dequantised A element by element -- bf16 hi + lo, FP16 with the
    power-of-two scale, e4m3 with the 448/amax scale (B = I has amax 1, whose
    scaled value 448 / 2^14 is exact in every format)."""
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[{prec}] identity")
    M, K = [96, 40], [64, 130]
    vb = VBatch(M, K, K).fill("normal", seed=12, outlier_frac=0.01)
    for b in range(2):
        vb.b(b).copy_(torch.eye(K[b], device="cuda"))
    bg.Plan(vb.M, vb.N, vb.K, prec=prec).run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
    torch.cuda.synchronize()
    for b in range(2):
        A = op_matrix(vb, "A", b)
        got = vb.host_flat("C", b).reshape(K[b], M[b]).T
        if prec == "bf16x3":
            hi, lo = orc.bf16_split(A)
            want = hi.astype(np.float64) + lo
        elif prec == "fp16":
            e = orc.fp16_pow2_exponent(np.abs(A).max())
            want = orc.fp16_scaled(A, e).astype(np.float64) / 2.0 ** e
        else:
            s = orc.fp8_scale(np.abs(A).max())
            want = orc.fp8_scaled(A, s).astype(np.float64) / np.float64(s)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16", "fp8"])
def test_split_precision_split_k(bg, orc, gpu, prec):
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[{prec}] split-K")
    M, N, K = [64, 100, 130], [64, 33, 128], [700, 333, 1025]
    vb = VBatch(M, N, K).fill("normal", seed=21, c_dist="uniform")
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(M, N, K, prec=prec, force_ksplit=4)
    assert (plan.ksplit() == 4).all()
    plan.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
    torch.cuda.synchronize()
    for b in range(3):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + M[b] * N[b]].reshape(N[b], M[b]).T.astype(np.float64)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T
        emul = 1.5 * emulate(orc, prec, A, B) + 0.5 * c0
        assert orc.rel_fro(got, emul) <= emul_bound(prec, A, B, emul, 1.5)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16", "fp8"])
def test_mixed_launch_equals_per_shape_launches(bg, gpu, prec):
    """Every element's k order is the same whatever tile shape computes it, so
    the mixed launch with edge tiles equals one-launch-per-shape without edge
    tiles bit for bit."""
    import torch
    from batchgemm_amd import workloads as w
    print(f"[mixed] {prec}", flush=True)
    M, N, K = w.shapes_unbalanced(200, seed=31)
    vb = w.VBatch(M, N, K).fill("normal", seed=31, outlier_frac=0.01, c_dist="uniform")
    C0 = vb.C.clone()
    outs = []
    # split-K off in both: the two launch forms have different workgroup slot
    # counts (occupancy per tile shape), so the planner may split K differently,
    # which changes the summation order
    # (the 64x64 .. 128x128 shapes only: the 256x256 tile of packed FP16 keeps a
    # launch of its own in both forms and could take a different share of the
    # matrices in each)
    no256 = sum(1 << t for t in range(9))
    for opts in ({"mixed_tiles": 1}, {"mixed_tiles": -1, "edge_tiles": -1}):
        vb.C.copy_(C0)
        p = bg.Plan(M, N, K, prec=prec, split_k_min=-1, tile_mask=no256, **opts)
        p.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
        torch.cuda.synchronize()
        outs.append((vb.C.clone(), p.kernels()))
    (c0, k0), (c1, k1) = outs
    mixed_id = bg._lib.TILE_MIXED
    assert (k0[:, 1] == mixed_id).any() and not (k1[:, 1] == mixed_id).any()
    assert torch.equal(c0.view(torch.int32), c1.view(torch.int32))


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
@pytest.mark.parametrize("trans", ["NN", "TN"])
def test_scaled_paths_padded_leading_dimension(bg, orc, gpu, prec, trans):
    """A and B with lda, ldb past their row counts: the amax pre-pass walks
    them column by column and must see every element (and no padding)."""
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[{prec}] padded ld trans={trans}")
    M = [64, 100, 17, 130, 3]
    N = [64, 33, 140, 128, 7]
    K = [32, 200, 9, 257, 5]
    vb = VBatch(M, N, K, trans[0], trans[1], pad_ab=3).fill("normal", seed=13, outlier_frac=0.01,
                                                          c_dist="uniform")
    # padding holds huge values: an amax that read it would shrink every scale
    for pool, off, rows, cols, ld in ((vb.A, vb.offA, vb.a_rows, vb.a_cols, vb.lda),
                                      (vb.B, vb.offB, vb.b_rows, vb.b_cols, vb.ldb)):
        for b in range(vb.batch):
            v = pool[int(off[b]): int(off[b]) + int(ld[b]) * int(cols[b])].view(int(cols[b]), int(ld[b]))
            v[:, int(rows[b]):] = 1e30
    C0 = vb.C.cpu().numpy()
    plan = bg.Plan(vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], prec=prec)
    plan.run(vb.dA, vb.dB, vb.dC, 1.0, 0.5)
    torch.cuda.synchronize()
    for b in range(vb.batch):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + M[b] * N[b]].reshape(N[b], M[b]).T.astype(np.float64)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T.astype(np.float64)
        emul = emulate(orc, prec, A, B) + 0.5 * c0
        assert orc.rel_fro(got, emul) <= emul_bound(prec, A, B, emul), b


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
def test_scaled_paths_tiny_magnitudes(bg, orc, gpu, prec):
    """Operands near 2^-58: the product of the two operand scales (2^67 each
    for FP16, about 2^65 for FP8) overflows FP32, so a post-scale computed as
    1/(sA*sB) would be 1/inf = 0 and C would silently come out zero.  The
    result itself (about 2^-110) is an ordinary FP32 number, and the kernels
    must produce it: FP16 applies 2^-(eA+eB) exactly, FP8 (1/sA)*(1/sB)."""
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[{prec}] tiny magnitudes")
    M, N, K = [64, 40, 130], [64, 96, 33], [128, 70, 257]
    vb = VBatch(M, N, K).fill("normal", seed=17)
    vb.A.mul_(2.0 ** -58)
    vb.B.mul_(2.0 ** -58)
    bg.Plan(M, N, K, prec=prec).run(vb.dA, vb.dB, vb.dC, 1.0, 0.0)
    torch.cuda.synchronize()
    for b in range(vb.batch):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T.astype(np.float64)
        exact = A.astype(np.float64) @ B.astype(np.float64)
        assert np.abs(got).max() > 0.0, "post-scale flushed the result to zero"
        emul = emulate(orc, prec, A, B)
        assert orc.rel_fro(got, emul) <= emul_bound(prec, A, B, emul), b
        assert orc.rel_fro(got, exact) <= FP64_BOUND[prec], b


@pytest.mark.parametrize("pad", [0, 3])
@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_bf16x3_256x128_tile(bg, orc, gpu, trans, pad):
    """The bf16x3 family's 256x128 tile (16-deep k-tiles, 128 x 64 per wave)
    on ragged shapes -- rows past M in a partial 256-row tile, tiles with 1,
    2 or 3 rows inside the operand and M < 4 (the m-contiguous loader's
    element-by-element path), K not a multiple of 16, K = 0, leading
    dimensions padded with 1e30 -- against the emulation and FP64, and
    bit-equal to the 128x128 tile (the k grouping into MFMAs is the same)."""
    import torch
    from batchgemm_amd.workloads import VBatch
    say(f"[bf16x3 256x128] trans={trans} pad={pad}")
    M = [256, 300, 17, 130, 513, 64, 258, 3, 1, 259, 20]
    N = [128, 33, 140, 129, 250, 1, 40, 50, 3, 2, 130]
    K = [64, 200, 9, 257, 35, 0, 48, 17, 33, 16, 1]
    vb = VBatch(M, N, K, trans[0], trans[1], pad_ab=pad).fill("normal", seed=19, outlier_frac=0.01,
                                                                c_dist="uniform")
    if pad:
        # padding holds huge values: a load that used it would spoil the product
        for pool, off, rows, cols, ld in ((vb.A, vb.offA, vb.a_rows, vb.a_cols, vb.lda),
                                          (vb.B, vb.offB, vb.b_rows, vb.b_cols, vb.ldb)):
            for b in range(vb.batch):
                v = pool[int(off[b]): int(off[b]) + int(ld[b]) * int(cols[b])].view(int(cols[b]), int(ld[b]))
                v[:, int(rows[b]):] = 1e30
    C0 = vb.C.clone()
    out = {}
    for tile in ("256x128", "128x128"):
        vb.C.copy_(C0)
        p = bg.Plan(vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc, trans[0], trans[1], prec="bf16x3",
                    policy="fixed", fixed_tile=tile, mixed_tiles=-1, split_k_min=-1)
        assert (p.tiles() == bg.TILES[tile]).all()
        p.run(vb.dA, vb.dB, vb.dC, 1.5, 0.5)
        torch.cuda.synchronize()
        out[tile] = vb.C.clone()
    assert torch.equal(out["256x128"].view(torch.int32), out["128x128"].view(torch.int32))
    C0h = C0.cpu().numpy()
    vb.C.copy_(out["256x128"])
    for b in range(vb.batch):
        A = op_matrix(vb, "A", b); B = op_matrix(vb, "B", b)
        c0 = C0h[int(vb.offC[b]): int(vb.offC[b]) + M[b] * N[b]].reshape(N[b], M[b]).T.astype(np.float64)
        got = vb.host_flat("C", b).reshape(N[b], M[b]).T.astype(np.float64)
        emul = 1.5 * emulate(orc, "bf16x3", A, B) + 0.5 * c0
        exact = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.5 * c0
        assert orc.rel_fro(got, emul) <= EMUL_BOUND["bf16x3"], b
        # a small matrix (K = 9, 1% outliers) can put the bf16x3 algorithm
        # itself past the FP64 bound: then the kernel stays within the
        # emulation bound of the algorithm's own error
        bound = max(FP64_BOUND["bf16x3"], orc.rel_fro(emul, exact) + EMUL_BOUND["bf16x3"])
        assert orc.rel_fro(got, exact) <= bound, b


def test_256x128_tile_refused_outside_bf16x3(bg, gpu):
    from batchgemm_amd._lib import VbgError
    with pytest.raises(VbgError):
        bg.Plan([300], [200], [100], prec="fp32", policy="fixed", fixed_tile="256x128")
