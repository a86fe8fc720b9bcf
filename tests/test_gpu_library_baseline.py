"""The per-matrix library baseline as a sanity oracle on the GPU (needs an
MI355X): one library GEMM per matrix (torch.addmm, rocBLAS / hipBLASLt), the
counterpart of the reference's hipBLAS loop ROCm/default/gemm.cuh:64-85, is
checked against FP64 and against the batched HIP path on the same inputs,
and the reference-protocol driver runs it with its known-answer check.

The library's own arithmetic is not pinned by any reference test (parity
unpinned); these tests show that the comparator computes the same GEMM as
the batched path, so the timings compare like with like."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fp64(vb, C0, b, alpha, beta):
    A = vb.a(b).double().cpu().numpy(); B = vb.b(b).double().cpu().numpy()
    m, n = int(vb.M[b]), int(vb.N[b])
    c0 = C0[int(vb.offC[b]): int(vb.offC[b]) + m * n].view(n, m).t().double().cpu().numpy()
    return alpha * (A @ B) + beta * c0


@pytest.mark.parametrize("case", ["uniform_b64", "dataset_b256"])
def test_library_loop_matches_fp64_and_batched_path(bg, gpu, case):
    import torch
    from batchgemm_amd import workloads as w
    from batchgemm_amd.baseline import library_loop
    if case == "uniform_b64":  # plain FP32 workload: B = 64, M,N,K ~ U{16..256}, alpha 1, beta 0
        M, N, K = w.shapes_uniform(64, 16, 256, seed=0)
        alpha, beta, dist = 1.0, 0.0, "uniform"
    else:  # the reference's dataset, first 256 triples, with the loop's alpha = beta = 1
        M, N, K = w.read_dataset(w.dataset_path(), 256)
        alpha, beta, dist = 1.0, 1.0, "normal"
    say(f"[library] {case}")
    vb = w.VBatch(M, N, K).fill(dist, seed=11, c_dist="uniform")
    C0 = vb.C.clone()
    library_loop(vb, alpha, beta)()
    torch.cuda.synchronize()
    C_lib = vb.C.clone()
    vb.C.copy_(C0)
    bg.Plan(M, N, K).run(vb.dA, vb.dB, vb.dC, alpha, beta)
    torch.cuda.synchronize()
    C_hip = vb.C.clone()
    worst_lib = worst_pair = 0.0
    for b in range(len(M)):
        ref = _fp64(vb, C0, b, alpha, beta)
        m, n = int(M[b]), int(N[b])
        sl = slice(int(vb.offC[b]), int(vb.offC[b]) + m * n)
        lib = C_lib[sl].view(n, m).t().double().cpu().numpy()
        hip = C_hip[sl].view(n, m).t().double().cpu().numpy()
        e_lib = np.linalg.norm(lib - ref) / np.linalg.norm(ref)
        e_pair = np.linalg.norm(lib - hip) / np.linalg.norm(ref)
        worst_lib, worst_pair = max(worst_lib, e_lib), max(worst_pair, e_pair)
        assert e_lib <= 1e-5, (b, m, n, int(K[b]), e_lib)
        assert e_pair <= 2e-5, (b, m, n, int(K[b]), e_pair)
    say(f"  library vs FP64 {worst_lib:.3g}; library vs batched path {worst_pair:.3g}")


def test_driver_per_matrix_library_known_answer(bg, gpu, capsys):
    """`driver B 6000 6000 --per-matrix-library --verify`: the reference's
    constant-data protocol through the library loop, every element against
    the closed form C_i = (i+3) + 11 K_i (i+1)(i+2)."""
    from batchgemm_amd import driver
    print("[library] driver --per-matrix-library", file=sys.stderr, flush=True)
    rc = driver.main(["64", "6000", "6000", "--per-matrix-library", "--verify"])
    out = capsys.readouterr().out.splitlines()
    assert rc == 0
    assert out[0].startswith("time======") and out[1].startswith("gflops==")
