"""Pins the CPU oracle against the reference's golden data, hand-derived
known answers, brute force and library routines (CPU only)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DATA = os.path.join(GOLD, "data_MN_K_512_128")


def logs():
    with open(os.path.join(GOLD, "reference_logs.json")) as f:
        return json.load(f)


# ------------------------------------------------------------ a1 + a2
def test_dataset_and_flop_accounting_match_reference_logs(orc):
    """gemm.cuh:34-48 + :112-116 reproduce every logged GFLOP total
    (ROCm/framework/log-test-verify3, ROCm/default/log-default)."""
    M, N, K = orc.read_dataset(DATA)
    assert len(M) == 4096
    d = logs()
    for src in ("rocm_framework", "rocm_hipblas_loop"):
        rows = d[src]["rows"]
        assert len(rows) == len(d["batches"])
        for B, row in zip(d["batches"], rows):
            g = orc.flops_ref(M[:B], N[:B], K[:B])
            assert "%f" % g == row["gflop"], (src, B)


def test_dataset_shape_facts(orc):
    M, N, K = orc.read_dataset(DATA)
    assert set(np.unique(M)) <= {16, 32, 64, 128, 256, 512}
    assert set(np.unique(N)) <= {16, 32, 64, 128, 256, 512}
    assert set(np.unique(K)) <= {16, 32, 64, 128}
    # 2*sum(MNK) of the first 8 and 256 dataset lines (the reference's count
    # without its 2MN term)
    assert abs(orc.flops(M[:8], N[:8], K[:8]) / 1e9 - 0.051708) < 5e-7
    assert abs(orc.flops(M[:256], N[:256], K[:256]) / 1e9 - 0.738640) < 5e-7


# ------------------------------------------------------------ a3..a7 planner
def test_ref_plan_hand_derived_cascade_with_quirk(orc):
    """Six 64x64 matrices, thres=100, walked through gemm.cuh:141-229 by hand.
    The :180 read of kernel_thread[t_strategy[i]] (the array indexed by a tile
    number) changes the outcome: with it t=[1]*6 and total 104; reading
    kernel_thread[i] instead would give t=[0,1,1,1,1,1] and total 112."""
    p = orc.ref_plan([64] * 6, [64] * 6, [32] * 6, 100)
    assert p["t_init"].tolist() == [4] * 6
    assert p["wf_init"] == 24
    assert p["t"].tolist() == [1] * 6
    assert p["wf_final"] == 104
    assert p["kernel_thread"].tolist() == [128] * 6
    assert p["grid"] == (4, 2, 6)


def test_ref_plan_hand_derived_no_cascade(orc):
    """Mixed shapes, thres just below the initial total: no demotion."""
    M = [64, 32, 16, 48, 128, 96]
    p = orc.ref_plan(M, M, [16] * 6, 59)
    assert p["t"].tolist() == [4, 2, 0, 0, 4, 2]
    assert p["kernel_thread"].tolist() == [256, 128, 128, 128, 256, 128]
    assert p["wf_init"] == p["wf_final"] == 60
    assert p["grid"] == (3, 3, 6)
    # one more wavefront of headroom starts the cascade at the first idx-4 matrix
    q = orc.ref_plan(M, M, [16] * 6, 60)
    assert q["t"][0] == 3 and q["wf_final"] > 60


def test_ref_plan_undefined_cases(orc):
    assert orc.ref_plan([20], [20], [16], 6000) is None          # no dividing tile
    assert orc.ref_plan([64], [64], [16], 6000) is None          # :180 reads kernel_thread[4], batch 1
    assert orc.ref_plan([64, 16], [64, 16], [16, 16], 0) is not None  # cascade never starts


def test_ref_plan_dataset_facts(orc):
    """Initial tile counts and wavefront totals on the shipped dataset, and the
    B=256 grid with its 2,498 useful blocks."""
    M, N, K = orc.read_dataset(DATA)
    totals = {8: 320, 16: 522, 32: 1116, 64: 2110, 128: 4356, 256: 7822}
    for B, tot in totals.items():
        p = orc.ref_plan(M[:B], N[:B], K[:B], 6000)
        assert p["wf_init"] == tot, B
        # the cascade fires exactly when the initial total is <= 6000
        assert (not np.array_equal(p["t"], p["t_init"])) == (tot <= 6000), B
    p = orc.ref_plan(M[:256], N[:256], K[:256], 6000)
    assert np.bincount(p["t_init"], minlength=5).tolist() == [47, 39, 36, 33, 101]
    assert p["grid"] == (32, 16, 256)
    tile = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64)]
    useful = sum((m // tile[t][0]) * (n // tile[t][1]) for m, n, t in zip(M[:256], N[:256], p["t"]))
    assert useful == 2498


def test_ref_plan_invariants_random(orc):
    """Brute-force invariants over random dataset prefixes and thresholds."""
    M, N, K = orc.read_dataset(DATA)
    tile = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64)]
    rng = np.random.default_rng(1)
    for _ in range(60):
        B = int(rng.integers(5, 300))
        s = int(rng.integers(0, 4096 - B))
        thres = int(rng.integers(0, 12000))
        m, n, k = M[s:s + B], N[s:s + B], K[s:s + B]
        p = orc.ref_plan(m, n, k, thres)
        if p is None:
            continue
        for i in range(B):  # initial choice = largest dividing tile (brute force)
            best = max(j for j in range(5) if m[i] % tile[j][0] == 0 and n[i] % tile[j][1] == 0
                       and tile[j][0] <= m[i] and tile[j][1] <= n[i])
            assert p["t_init"][i] == best
        assert np.all(p["t"] <= p["t_init"])
        assert np.all(p["t_init"] - p["t"] <= 4)
        if p["wf_init"] > thres:
            assert np.array_equal(p["t"], p["t_init"])
        gx = max(int(a) // tile[t][0] for a, t in zip(m, p["t"]))
        gy = max(int(b) // tile[t][1] for b, t in zip(n, p["t"]))
        assert p["grid"] == (gx, gy, B)


# ------------------------------------------------------------ FP64 oracle
def _colmajor(rng, rows, cols, ld):
    x = rng.standard_normal((cols, ld)).astype(np.float32)
    return x.ravel(), x[:, :rows].T  # flat storage, (rows, cols) view


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_dgemm_matches_numpy(orc, ta, tb):
    rng = np.random.default_rng(7)
    for M, N, K in [(1, 1, 1), (5, 3, 7), (33, 17, 65), (64, 1, 9), (0, 4, 4), (4, 4, 0)]:
        ar, ac = (K, M) if ta == "T" else (M, K)
        br, bc = (N, K) if tb == "T" else (K, N)
        lda, ldb, ldc = ar + 3, br + 1, M + 2
        A, Av = _colmajor(rng, ar, ac, lda)
        B, Bv = _colmajor(rng, br, bc, ldb)
        C, Cv = _colmajor(rng, M, N, ldc)
        opA = Av.T if ta == "T" else Av
        opB = Bv.T if tb == "T" else Bv
        for alpha, beta in [(1.0, 0.0), (1.5, 0.5), (-2.0, 1.0)]:
            out = orc.dgemm(ta, tb, A, B, C, alpha, beta, M, N, K, lda, ldb, ldc)
            ref = alpha * (opA.astype(np.float64) @ opB.astype(np.float64)) + beta * Cv
            got = out.reshape(max(N, 1), ldc).T[:M, :N] if N else np.zeros((M, 0))
            np.testing.assert_allclose(got, ref.reshape(M, N), rtol=1e-12, atol=1e-12)


def test_dgemm_beta_zero_ignores_nan(orc):
    A = np.ones(4, np.float32); B = np.ones(4, np.float32); C = np.full(4, np.nan, np.float32)
    out = orc.dgemm("N", "N", A, B, C, 1.0, 0.0, 2, 2, 2, 2, 2, 2)
    assert np.all(out == 2.0)


def test_host_sgemm_baseline_matches_fp64(orc):
    rng = np.random.default_rng(3)
    M = rng.integers(1, 90, 12); N = rng.integers(1, 90, 12); K = rng.integers(0, 300, 12)
    for ta in "NT":
        for tb in "NT":
            As, Bs, Cs, C0 = [], [], [], []
            lda, ldb = [], []
            for m, n, k in zip(M, N, K):
                ar, ac = (k, m) if ta == "T" else (m, k)
                br, bc = (n, k) if tb == "T" else (k, n)
                As.append(rng.uniform(-1, 1, max(1, ar) * ac).astype(np.float32)); lda.append(max(1, ar))
                Bs.append(rng.uniform(-1, 1, max(1, br) * bc).astype(np.float32)); ldb.append(max(1, br))
                Cs.append(rng.uniform(-1, 1, m * n).astype(np.float32)); C0.append(Cs[-1].copy())
            ref = orc.dgemm_vbatched(ta, tb, M, N, K, 1.5, As, lda, Bs, ldb, 0.5, C0, M)
            orc.sgemm_vbatched(ta, tb, M, N, K, 1.5, As, lda, Bs, ldb, 0.5, Cs, M)
            for c, r in zip(Cs, ref):
                assert orc.rel_fro(c, r[:len(c)]) < 1e-6


# ------------------------------------------------------------ tile semantics
def test_ref_tile_seq_reproduces_log_verify(orc):
    """ROCm/tiling/log-verify: after a warm-up and 1000 timed launches of
    C += A*B on A=i+1, B=i+2, C=i+3, C = (i+3) + 1001*K*(i+1)*(i+2).  The
    sequential-fma tile model reproduces the first two logged values exactly
    (K = 256 and 64 for the first two matrices of that run)."""
    rows = logs()["wang_log_verify"]["rows"]
    for row, K in zip(rows[:2], [256, 64]):
        a, b = row["A"], row["B"]
        A = np.full(32 * K, a, np.float32); B = np.full(K * 32, b, np.float32)
        C = np.full(32 * 32, a + 2, np.float32)
        for _ in range(1001):
            C = orc.ref_tile_seq(False, 32, 32, K, A, B, C)
        assert np.all(C == row["C"])


def test_ref_tile_seq_fma_chain(orc):
    """Each element equals the step-by-step chain c <- round32(a*b + c): the
    product is exact in float64 and the float64 sum rounded to float32 is the
    fused result except in double-rounding corner cases (none at these sizes)."""
    rng = np.random.default_rng(11)
    M, N, K = 7, 5, 40
    for rowmajor in (False, True):
        A = rng.standard_normal(M * K).astype(np.float32)
        B = rng.standard_normal(K * N).astype(np.float32)
        C = rng.standard_normal(M * N).astype(np.float32)
        got = orc.ref_tile_seq(rowmajor, M, N, K, A, B, C)
        for i in range(M):
            for j in range(N):
                ci = i * N + j if rowmajor else i + j * M
                c = C[ci]
                for k in range(K):
                    a = A[i * K + k] if rowmajor else A[i + k * M]
                    b = B[k * N + j] if rowmajor else B[k + j * K]
                    c = np.float32(np.float64(a) * np.float64(b) + np.float64(c))
                assert got[ci] == c


# ------------------------------------------------------------ quantisation emulation
def test_bf16_round_matches_torch():
    import torch
    from oracle import oracle as o
    x = np.random.default_rng(0).standard_normal(100000).astype(np.float32) * 10.0 ** np.random.default_rng(1).integers(-20, 20, 100000)
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(o.bf16_round(x), ref)
    hi, lo = o.bf16_split(x)
    assert np.array_equal(hi, ref)
    assert np.array_equal(lo, torch.from_numpy((x - hi).astype(np.float32)).to(torch.bfloat16).float().numpy())


def test_e4m3_round_matches_torch():
    import torch
    from oracle import oracle as o
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-4, 2.6, 200000)).astype(np.float32)
    x = np.clip(x, -448, 448).astype(np.float32)
    x[:5] = [0.0, 2.0 ** -9, 2.0 ** -10, 3 * 2.0 ** -11, 448.0]
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).float().numpy()
    assert np.array_equal(o.e4m3_round(x), ref)


def test_fp16_scaled_matches_numpy():
    from oracle import oracle as o
    x = np.random.default_rng(4).standard_normal(5000).astype(np.float32) * 3e3
    e = o.fp16_pow2_exponent(np.abs(x).max())
    assert 2 ** 14 <= np.abs(x).max() * 2.0 ** e < 2 ** 15
    q = o.fp16_scaled(x, e)
    assert np.array_equal(q, (x * np.float32(2.0 ** e)).astype(np.float16).astype(np.float32))


def test_post_scale_undoes_both_scales():
    """FP16: 2^-(eA+eB) is exact wherever float32 holds it, including where
    the product of the two scales 2^eA * 2^eB overflows float32 (eA + eB >=
    128: tiny operands).  FP8: (1/sA)(1/sB) is 1/(sA sB) to float32 rounding,
    also where sA * sB overflows."""
    from oracle import oracle as o
    for ea, eb in [(0, 0), (14, -3), (70, 70), (126, 20), (-60, -60)]:
        p = o.post_scale("fp16", ea, eb)
        assert p > 0.0 and p * 2.0 ** ea * 2.0 ** eb == 1.0
    assert o.post_scale("fp16", 126, 126) == 0.0  # 2^-252: below float32's range
    for sa, sb in [(448.0, 1.0), (3.7, 0.01), (2.0 ** 65, 2.0 ** 64.5), (1e-20, 7.0)]:
        sa, sb = float(np.float32(sa)), float(np.float32(sb))
        p = o.post_scale("fp8", sa, sb)
        # two float32 roundings: within 2 units in the last place of p (a
        # subnormal p, as for the 2^65 scales, has a coarser last place)
        assert p > 0.0 and abs(p - 1.0 / (sa * sb)) <= 2 * float(np.spacing(np.float32(p)))
    # scales stay finite for tiny operands
    assert o.fp16_pow2_exponent(2.0 ** -140) == 126
    assert np.isfinite(o.fp8_scale(1e-40)) and o.fp8_scale(1e-40) == np.float32(2.0 ** 126)


def test_split_product_exact_on_representable_inputs():
    """Where every quantisation step is exact the emulated product is the
    exact product: bf16x3 on values with 16 significant bits (hi + lo holds
    them, lo*lo = 0 is the only dropped term and vanishes when lo has no
    bits below the product's); FP16 on small integers (amax 7 -> scale 2^12,
    7 * 2^12 < 2^15); FP8 on integers in [-7, 7] with amax 7 (scale 448/7 = 64,
    and 64 k, |k| <= 7, is representable in e4m3)."""
    from oracle import oracle as o
    rng = np.random.default_rng(8)
    A = rng.integers(-7, 8, (13, 40)).astype(np.float32); A[0, 0] = 7
    B = rng.integers(-7, 8, (40, 9)).astype(np.float32); B[0, 0] = -7
    exact = A.astype(np.float64) @ B.astype(np.float64)
    for prec in ("fp16", "fp8", "bf16x3"):
        assert np.array_equal(o.split_product(prec, A, B), exact), prec
    # bf16x3 keeps 16 significant bits of each operand: hi + lo = x exactly
    x = (rng.integers(1 << 15, 1 << 16, (6, 30)) * 2.0 ** -20).astype(np.float32)
    y = (rng.integers(1 << 15, 1 << 16, (30, 5)) * 2.0 ** -3).astype(np.float32)
    hi, lo = o.bf16_split(x)
    assert np.array_equal(hi.astype(np.float64) + lo, x.astype(np.float64))
    full = x.astype(np.float64) @ y.astype(np.float64)
    xh, xl = o.bf16_split(x); yh, yl = o.bf16_split(y)
    dropped = xl.astype(np.float64) @ yl.astype(np.float64)
    assert np.allclose(o.split_product("bf16x3", x, y) + dropped, full, rtol=1e-15, atol=0)


def test_wang_order_matches_reference_sorted_dataset(orc):
    """sort.cpp's output for the first 256 dataset lines is shipped with the
    reference (ROCm/data/data_MN_K_512_128_sort)."""
    M, N, K = orc.read_dataset(os.path.join(GOLD, "data_MN_K_512_128"))
    gold = np.loadtxt(os.path.join(GOLD, "data_MN_K_512_128_sort"), dtype=np.int64)
    o = orc.wang_order(M[:256], N[:256], K[:256])
    assert np.array_equal(np.stack([M[:256][o], N[:256][o], K[:256][o]], 1), gold)
    # stable: equal keys keep their input order
    assert all(o[i] < o[i + 1] for i in range(255)
               if (K[:256][o[i]], M[:256][o[i]] % 64, N[:256][o[i]] % 64) ==
               (K[:256][o[i + 1]], M[:256][o[i + 1]] % 64, N[:256][o[i + 1]] % 64))


def test_wang_plan_invariants(orc):
    M, N, K = orc.read_dataset(os.path.join(GOLD, "data_MN_K_512_128"))
    for B in (8, 64, 256):
        for thres in (2000, 65536, 1 << 20):
            st = orc.wang_plan(M[:B], N[:B], K[:B], thres)
            ts = np.array([16, 32, 64])[st]
            assert ((M[:B] % ts == 0) & (N[:B] % ts == 0)).all()   # chosen tiles divide
            assert st.max() <= 2 and st.min() >= 0
    # hand-derived: six 128x128 matrices, thres 40.  No tile gives TLP < 40,
    # so the global tile is 64 (index 2); the initial workgroup count is
    # 6 * 4 = 24 < 40, so the demotion walks the sorted order (K descending:
    # 70, 60, 50, 30, 20, 10) from its end: K=10 -> 32x32 (24 + 12 = 36),
    # K=20 -> 32x32 (48 >= 40): stop.
    st = orc.wang_plan([128] * 6, [128] * 6, [10, 50, 30, 70, 20, 60], 40)
    assert st.tolist() == [1, 2, 2, 2, 1, 2]


def test_li_plan_hand_derived(orc):
    """M = (128, 64, 16, 32), N = (128, 128, 16, 256), thres 1000: TLP is
    33024 (16x16), 8192 (32x32), 1536 (64x64), 512 (128x64) -> global tile
    128x64; 128x128 fits it; 64x128 takes 64x64 (tm == M); 16x16 takes 16x16;
    32x256 takes 32x32."""
    assert orc.li_plan([128, 64, 16, 32], [128, 128, 16, 256], 1000).tolist() == [3, 2, 0, 1]
    # no tile below the threshold: the last (128x128) is the global tile
    assert orc.li_plan([256], [256], 1).tolist() == [5]
