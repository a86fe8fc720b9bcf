"""C-ABI library on the CPU: exports, host-only planning, argument checks,
and the library's reference planner against the oracle's (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DATA = os.path.join(HERE, "golden", "data_MN_K_512_128")


def header_symbols():
    src = open(os.path.join(ROOT, "include", "vbgemm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vbg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(bg):
    syms = header_symbols()
    assert len(syms) >= 14
    L = ctypes.CDLL(bg.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(syms) == sorted(bg._lib.EXPORTS)
    assert bg.lib().vbg_abi_version() == 2


def test_status_strings(bg):
    L = bg.lib()
    for code, name in [(0, b"VBG_OK"), (-1, b"VBG_ERR_INVALID_ARG"), (-6, b"VBG_ERR_UNDEFINED")]:
        assert L.vbg_status_string(code) == name


def _cover(plan, M, N, sample=40):
    """Every output element of every matrix covered exactly once per K slice;
    no item outside its matrix (exact on a sample, area totals on all)."""
    import batchgemm_amd
    shapes = batchgemm_amd.TILE_SHAPES
    items = plan.tiled_items()
    ks = plan.ksplit()
    area = np.zeros(len(M), np.int64)
    rng = np.random.default_rng(0)
    pick = set(rng.choice(len(M), min(sample, len(M)), replace=False).tolist())
    grids = {b: np.zeros((int(ks[b]), int(M[b]), int(N[b])), np.int16) for b in pick}
    for mat, t, tm, tn, sl in items:
        bm, bn = shapes[t]
        r0, c0 = tm * bm, tn * bn
        assert 0 <= r0 < M[mat] and 0 <= c0 < N[mat] and 0 <= sl < ks[mat]
        r1, c1 = min(r0 + bm, M[mat]), min(c0 + bn, N[mat])
        area[mat] += (r1 - r0) * (c1 - c0)
        if mat in grids:
            grids[mat][sl, r0:r1, c0:c1] += 1
    for b in range(len(M)):
        assert area[b] == int(M[b]) * int(N[b]) * int(ks[b]) if M[b] and N[b] else area[b] == 0
    for b, g in grids.items():
        assert g.size == 0 or (g.min() == 1 and g.max() == 1), b


@pytest.mark.parametrize("policy", ["auto", "reference", "wang", "li", "fixed"])
def test_host_plan_covers_every_tile_once(bg, policy):
    from batchgemm_amd import workloads as w
    for seed, (M, N, K) in enumerate([w.shapes_uniform(64, 16, 256, 0),
                                      w.shapes_unbalanced(512, seed=1),
                                      w.shapes_loguniform(300, seed=2),
                                      w.read_dataset(DATA, 256)]):
        for edges in (1, -1):
            p = bg.Plan(M, N, K, host_only=True, policy=policy, tlp_thres=6000, edge_tiles=edges)
            _cover(p, M, N)
            assert p.n_items > 0 and 1 <= p.n_launches <= 6


def test_host_plan_zero_sized_matrices(bg):
    p = bg.Plan([0, 5, 7], [3, 0, 9], [4, 4, 0], host_only=True)
    t = p.tiles()
    assert t[0] == -1 and t[1] == -1 and t[2] >= 0     # K = 0 still writes beta*C
    p = bg.Plan([], [], [], host_only=True)
    assert p.n_items == 0


def test_host_plan_xcd_queues_balance(bg):
    from batchgemm_amd import workloads as w
    M, N, K = w.shapes_unbalanced(2048, seed=0)
    p = bg.Plan(M, N, K, host_only=True)
    assert p.grid >= 256
    p1 = bg.Plan(M, N, K, host_only=True, xcd_queues=0)
    assert p1.n_items == p.n_items


def test_invalid_arguments_return_errors(bg):
    from batchgemm_amd._lib import VbgError
    with pytest.raises(VbgError) as e:
        bg.Plan([4], [4], [-1], host_only=True)
    assert e.value.code == -1
    with pytest.raises(VbgError):
        bg.Plan([4], [4], [4], lda=[3], host_only=True)       # lda < M
    with pytest.raises(VbgError):
        bg.Plan([4], [4], [4], ldc=[2], host_only=True)
    with pytest.raises(VbgError):
        bg.Plan([4], [4], [4], transA="X", host_only=True)
    with pytest.raises(VbgError):
        bg.Plan([4], [4], [4], prec=9, host_only=True)
    with pytest.raises(VbgError):
        bg.Plan([4], [4], [4], host_only=True, policy="fixed", fixed_tile=99)
    with pytest.raises(VbgError):  # the skinny tiles are built for FP32 only
        bg.Plan([4], [4], [4], host_only=True, policy="fixed", fixed_tile="16x128", prec="bf16x3")
    p = bg.Plan([4], [4], [4], host_only=True)
    L = bg.lib()
    rc = L.vbg_sgemm_vbatched(p._p, 1.0, None, None, 0.0, None, None)
    assert rc == -5  # host-only plan cannot run


def test_library_ref_plan_matches_oracle(bg, orc):
    M, N, K = orc.read_dataset(DATA)
    rng = np.random.default_rng(5)
    cases = [(M[:B], N[:B], K[:B], 6000) for B in (8, 16, 32, 64, 128, 256)]
    for _ in range(100):
        B = int(rng.integers(1, 200)); s = int(rng.integers(0, 4096 - B))
        cases.append((M[s:s + B], N[s:s + B], K[s:s + B], int(rng.integers(0, 9000))))
    cases.append(([64] * 6, [64] * 6, [32] * 6, 100))
    cases.append(([20], [20], [16], 6000))
    for m, n, k, th in cases:
        a = bg.ref_plan(m, n, k, th)
        b = orc.ref_plan(m, n, k, th)
        assert (a is None) == (b is None)
        if a is None:
            continue
        assert np.array_equal(a["t"], b["t"])
        assert np.array_equal(a["kernel_thread"], b["kernel_thread"])
        assert a["grid"] == b["grid"]
        assert (a["wf_init"], a["wf_final"]) == (b["wf_init"], b["wf_final"])


def test_reference_policy_follows_reference_planner(bg, orc):
    M, N, K = orc.read_dataset(DATA)
    for B in (8, 64, 256):
        p = bg.Plan(M[:B], N[:B], K[:B], host_only=True, policy="reference", wavefront_thres=6000)
        ref = orc.ref_plan(M[:B], N[:B], K[:B], 6000)
        to_tile = [0, 1, 1, 2, 2]
        assert np.array_equal(p.tiles(), np.array([to_tile[t] for t in ref["t"]]))


def test_host_plan_split_k_and_edges_cover(bg):
    from batchgemm_amd import workloads as w
    M, N, K = w.shapes_unbalanced(300, seed=3)
    for prec in ("fp32", "bf16x3"):
        p = bg.Plan(M, N, K, host_only=True, prec=prec, force_ksplit=3)
        assert (p.ksplit() > 1).any()
        _cover(p, M, N)


def test_edge_tiles_cut_padding(bg):
    """A 130 x 190 matrix on 128x128 tiles: its last row strip (2 rows) and
    column strip (62 columns) go to 64-wide tiles, cutting the padded area."""
    shapes = bg.TILE_SHAPES
    def padded(p):
        return sum(shapes[t][0] * shapes[t][1] for _, t, _, _, _ in p.tiled_items())
    on = bg.Plan([130], [190], [64], host_only=True, policy="fixed", fixed_tile="128x128")
    off = bg.Plan([130], [190], [64], host_only=True, policy="fixed", fixed_tile="128x128", edge_tiles=-1)
    assert padded(on) < padded(off) and padded(off) == 4 * 128 * 128
    # 64x64 corner, 128x64 right strip, 64x128 bottom strip, 128x128 interior
    assert sorted({int(t) for _, t, _, _, _ in on.tiled_items()}) == [2, 3, 4, 5]


def test_kernel_table_and_item_flops(bg):
    """Query 7 lists the kernels of a run in order; per-item algorithmic flops
    sum to 2*sum(MNK) (padding excluded, split-K slices partition K)."""
    from batchgemm_amd import workloads as w
    from batchgemm_amd._lib import VbgError
    M, N, K = w.shapes_unbalanced(300, seed=4)
    total = 2.0 * np.sum(M.astype(np.float64) * N * K)
    for prec, ks in (("fp32", 0), ("fp32", 3), ("fp16", 0), ("bf16x3", 5), ("fp32_seq", 0)):
        p = bg.Plan(M, N, K, host_only=True, prec=prec, force_ksplit=ks)
        kt = p.kernels()
        kinds = list(kt[:, 0])
        assert (kinds[0] == 0) == (prec == "fp16")
        assert (kinds[-1] == 5) == (ks > 1)
        packed = p.packed_bytes > 0
        assert (6 in kinds) == packed
        if packed:  # the pack pre-pass follows the scale pre-pass, before every GEMM launch
            assert kinds.index(6) == (1 if prec == "fp16" else 0)
            assert all(k == 7 for k in kinds if k not in (0, 5, 6))
        body = kt[[k not in (0, 5, 6) for k in kinds]]
        assert len(body) == p.n_launches
        assert body[0, 2] == 0 and body[:, 3].sum() == p.n_items
        assert np.array_equal(body[1:, 2], np.cumsum(body[:, 3])[:-1])
        fl = p.item_flops()
        assert abs(fl.sum() - total) <= 1e-9 * total, prec
        # k slices of one output tile partition [0, K)
        ti, kr = p.tiled_items(), p.items()
        spans = {}
        for (mat, t, tm, tn, sl), (_, _, _, kb, ke) in zip(ti, kr):
            spans.setdefault((mat, t, tm, tn), []).append((kb, ke))
        for (mat, *_), sp in spans.items():
            sp.sort()
            assert sp[0][0] == 0 and sp[-1][1] == K[mat]
            assert all(a[1] == b[0] for a, b in zip(sp, sp[1:]))
        with pytest.raises(VbgError) as e:
            p.kernel_times()
        assert e.value.code == -5


def test_packed_operand_workspace(bg):
    """Packed operands: per matrix, A (M rows) then B (N rows) of K rounded up
    to 64 elements at 4 / 2 / 1 bytes (bf16x3 hi+lo / FP16 / FP8), each
    256-byte aligned; by default FP16 / FP8 pack a high-intensity batch (the
    unbalanced one) and not a low-intensity one, bf16x3 never;
    pack_operands forces either way."""
    from batchgemm_amd import workloads as w
    M, N, K = w.shapes_unbalanced(300, seed=4)
    M[3] = 0  # an empty matrix and a K = 0 one hold no packed operands
    K[5] = 0
    for prec, eb in (("bf16x3", 4), ("fp16", 2), ("fp8", 1)):
        kp = -(-K.astype(np.int64) // 64) * 64
        live = (M > 0) & (N > 0) & (K > 0)
        al = lambda x: -(-x // 256) * 256
        want = int(np.sum(np.where(live, al(M * kp * eb) + al(N * kp * eb), 0)))
        p = bg.Plan(M, N, K, host_only=True, prec=prec, pack_operands=1)
        assert p.packed_bytes == want, prec
        dflt = bg.Plan(M, N, K, host_only=True, prec=prec).packed_bytes
        assert dflt == (0 if prec == "bf16x3" else want), prec
        assert bg.Plan(M, N, K, host_only=True, prec=prec, pack_operands=-1).packed_bytes == 0
    Ms, Ns, Ks = [np.full(64, 40, np.int32)] * 3  # AI = 2*40^3 / (4*2*40^2) = 10 flop/B
    assert bg.Plan(Ms, Ns, Ks, host_only=True, prec="fp8").packed_bytes == 0
    assert bg.Plan(Ms, Ns, Ks, host_only=True, prec="fp8", pack_operands=1).packed_bytes > 0
    assert bg.Plan(Ms, Ns, Ks, host_only=True, prec="fp32", pack_operands=1).packed_bytes == 0
    # the default does not pack past 8 GiB of workspace (80 x 8192^3 FP8: 10 GiB), a request does
    Mb = np.full(80, 8192, np.int32)
    assert bg.Plan(Mb, Mb, Mb, host_only=True, prec="fp8").packed_bytes == 0
    assert bg.Plan(Mb[:8], Mb[:8], Mb[:8], host_only=True, prec="fp8").packed_bytes == 8 * 2 * 8192 * 8192


def test_driver_gflop_count_matches_reference_logs(bg):
    """The driver's GFLOP count is the reference's sum(2MNK + 2MN) (the
    numbers ROCm/framework/log-test-verify3 prints for B = 8..256)."""
    import json
    from batchgemm_amd import workloads as w
    from batchgemm_amd.driver import reference_gflop
    logs = json.load(open(os.path.join(HERE, "golden", "reference_logs.json")))
    for B, row in zip(logs["batches"], logs["rocm_framework"]["rows"]):
        M, N, K = w.read_dataset(DATA, B)
        assert f"{reference_gflop(M, N, K):f}" == row["gflop"], B


def test_autotune_keys_and_cache(bg, tmp_path):
    """Batches of one distribution share coarse keys; the JSON cache
    round-trips; an unknown key falls back to the default heuristic."""
    from batchgemm_amd import autotune as at
    batches = at.batching_workload(12, seed=3)
    for M, N, K in batches:
        assert 256 <= len(M) <= 8192 and M.min() >= 8 and max(M.max(), N.max(), K.max()) <= 1024
    keys = {at.distribution_key(M, N, K, "fp32") for M, N, K in batches}
    assert len(keys) <= 7            # one per log2(B) bucket at most
    assert at.distribution_key(*batches[0], "fp32") != at.distribution_key(*batches[0], "bf16x3")
    path = str(tmp_path / "tune.json")
    t = at.Autotuner(cache_path=path, tune_missing=False)
    k = at.distribution_key(*batches[0], "fp32")
    assert t.choice(k) == ("auto", {})
    t.table[k] = {"choice": "auto:ge64", "ms": {}, "gain": 1.1}
    t.save()
    t2 = at.Autotuner(cache_path=path, tune_missing=False)
    assert t2.choice(k) == ("auto:ge64", {"tile_mask": 0b111100})
    p = t2.plan(*batches[0], prec="fp32", host_only=True)
    assert set(int(x) for x in p.tiles()) <= {2, 3, 4, 5}
    names = [n for n, _ in at.CANDIDATES]
    assert len(set(names)) == len(names) and names[0] == at.DEFAULT


def test_driver_sequential_model():
    """The host replay of the reference tiles' fma sequence equals the
    closed form while sums stay below 2^24 and departs from it above."""
    from batchgemm_amd.driver import sequential_fp32
    K = np.array([16, 64, 128, 8] + [128] * 252)
    v = sequential_fp32(K, 11)
    exact = np.array([(i + 3) + 11 * int(K[i]) * (i + 1) * (i + 2) for i in range(len(K))], np.float64)
    small = exact < 2 ** 24
    assert small[:4].all() and np.array_equal(v[small].astype(np.float64), exact[small])
    assert (v[~small].astype(np.float64) != exact[~small]).any()
    assert np.allclose(v, exact, rtol=2e-4)


def test_opts_struct_matches_header(bg):
    """The ctypes vbg_opts has the header's fields in the header's order."""
    src = open(os.path.join(ROOT, "include", "vbgemm.h")).read()
    body = re.search(r"typedef struct vbg_opts \{(.*?)\} vbg_opts;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b(?:int|unsigned)\s+([a-z_0-9]+)\s*;", body)
    assert fields == [f for f, _ in bg._lib.vbg_opts._fields_]
    assert ctypes.sizeof(bg._lib.vbg_opts) == 4 * len(fields)


def test_wang_policy_matches_oracle_and_ignores_input_order(bg, orc):
    """The library's Wang policy (ROCm/tiling/gemm_wang.cuh:104-182 on the
    sort.cpp order) against the oracle's restatement, and the same choice
    for any order of the batch."""
    M, N, K = orc.read_dataset(DATA)
    cases = [(M[:B], N[:B], K[:B], th) for B in (8, 64, 256) for th in (40, 4000, 65536, 1 << 20)]
    cases.append(([128] * 6, [128] * 6, [10, 50, 30, 70, 20, 60], 40))
    rng = np.random.default_rng(3)
    # distinct sort keys (K) so the sorted order, and hence the choice, cannot
    # depend on the input order (ties keep input order, as in the reference)
    mm = 32 * rng.integers(1, 9, 120); nn = 32 * rng.integers(1, 9, 120)
    kk = rng.permutation(np.arange(8, 8 + 120))
    cases += [(mm, nn, kk, th) for th in (4000, 65536, 400000)]
    for i, (m, n, k, th) in enumerate(cases):
        m, n, k = (np.asarray(x, np.int32) for x in (m, n, k))
        p = bg.Plan(m, n, k, host_only=True, policy="wang", tlp_thres=th)
        want = orc.wang_plan(m, n, k, th)
        assert np.array_equal(p.tiles(), want), (len(m), th)
        if len(set(k.tolist())) == len(k):
            perm = rng.permutation(len(m))
            q = bg.Plan(m[perm], n[perm], k[perm], host_only=True, policy="wang", tlp_thres=th)
            assert np.array_equal(q.tiles(), p.tiles()[perm]), i


def test_li_policy_matches_oracle(bg, orc):
    """The library's Li policy (CUDA/batching/gemm.cu:99-127) against the
    oracle's restatement on dataset prefixes and a hand-derived case."""
    M, N, K = orc.read_dataset(DATA)
    cases = [(M[:B], N[:B], K[:B], th) for B in (8, 64, 256) for th in (1000, 65536, 1 << 20, 1 << 24)]
    cases.append(([128, 64, 16, 32], [128, 128, 16, 256], [8, 8, 8, 8], 1000))
    for m, n, k, th in cases:
        p = bg.Plan(m, n, k, host_only=True, policy="li", tlp_thres=th)
        assert np.array_equal(p.tiles(), orc.li_plan(m, n, th)), (len(m), th)


def test_package_dataset_is_the_reference_dataset(bg):
    """The driver's default input (package data) is byte for byte the dataset
    fixture copied from the reference (ROCm/data/data_MN_K_512_128)."""
    from batchgemm_amd import workloads
    path = workloads.dataset_path()
    assert os.path.dirname(path).endswith(os.path.join("batchgemm_amd", "data"))
    assert open(path, "rb").read() == open(DATA, "rb").read()


def test_256x128_tile_is_bf16x3_bf16_and_packed_fp16_only(bg):
    """AUTO plans use the 256x128 tile for bf16x3, BF16 and packed FP16
    operands only (unless the tile mask leaves it out), and packed FP16 never
    the 256x256 one; FIXED refuses 256x128 for the other precisions and for
    FP16 quantised in the k-loop."""
    from batchgemm_amd import workloads as w
    M, N, K = w.shapes_unbalanced(256, seed=4)
    all_but = sum(1 << t for t in range(6))
    for prec in ("fp32", "bf16x3", "fp16", "fp8"):
        p = bg.Plan(M, N, K, prec=prec, host_only=True)
        uses = (p.tiles() == 6).any()
        assert uses == (prec in ("bf16x3", "fp16")), prec
        assert not (p.tiles() == 9).any(), prec
        assert not (bg.Plan(M, N, K, prec=prec, host_only=True, tile_mask=all_but).tiles() == 6).any()
    assert bg.Plan(M, N, K, prec="fp16", host_only=True).packed_bytes > 0
    assert not (bg.Plan(M, N, K, prec="fp16", host_only=True, pack_operands=-1).tiles() == 6).any()
    for prec, o in (("bf16x3", {}), ("fp16", {})):
        p = bg.Plan(M, N, K, prec=prec, host_only=True, policy="fixed", fixed_tile="256x128", **o)
        assert (p.tiles()[(M > 0) & (N > 0)] == 6).all()
        _cover(p, M, N)
    for prec, o in (("fp32", {}), ("fp8", {}), ("fp16", {"pack_operands": -1})):
        with pytest.raises(bg.VbgError):
            bg.Plan(M, N, K, prec=prec, host_only=True, policy="fixed", fixed_tile="256x128", **o)
