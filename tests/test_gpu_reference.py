"""GPU parity against the reference's own framework kernels (needs an MI355X).

The reference's dispatcher gemm<256> (ROCm/framework/kernel.h:9-42) is built
from its sources by oracle/build_ref.sh into oracle/_ref/libref_framework.so
and launched on the same device buffers as the product path.  Only its two
tiles that compute a correct product are compared:

- 32x32, column-major (ROCm/framework/kernel_128_new.h:131-201), launched with
  the reference's own 4096 bytes of LDS (ROCm/framework/gemm.cuh:238);
- 32x64, row-major (ROCm/framework/kernel_256_new.h:2-72), launched with
  8192 bytes, since it indexes 6144 bytes of LDS (kernel_256_new.h:5,58).

Both accumulate every element as one fma chain over k = 0..K-1 seeded from C,
so the reference-compatible VBG_PREC_FP32_SEQ path with alpha = beta = 1 must
equal them bit for bit, and the MFMA path must agree within 1e-5 relative
Frobenius error.  A row-major product C = A*B is run on the product path as
the column-major C^T = B^T * A^T over the same memory.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_framework.so")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REF_SO),
                                 reason="reference kernels not built (oracle/build_ref.sh "
                                        "needs the reference source tree)")]

TILE = {2: (32, 32), 3: (32, 64)}
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_ref = None


def ref_lib():
    global _ref
    if _ref is None:
        import torch  # noqa: F401  (one HIP runtime: torch's, loaded first)
        L = ctypes.CDLL(REF_SO)
        L.ref_framework_gemm.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 8
        L.ref_framework_gemm.restype = ctypes.c_int
        _ref = L
    return _ref


def make_problem(Mr, Nr, K, T, seed):
    """Buffers for the reference's problems as column-major product-path
    problems: tile 2 keeps (M, N); tile 3 (row-major) becomes (N, M) with the
    operands swapped."""
    from batchgemm_amd.workloads import VBatch
    T = np.asarray(T)
    Mo = np.where(T == 3, Nr, Mr)
    No = np.where(T == 3, Mr, Nr)
    return VBatch(Mo, No, K).fill("normal", seed=seed, c_dist="normal")


def run_reference(vb, Mr, Nr, K, T, lds_bytes):
    """Launch gemm<256> on vb's buffers after checking, on the host, every
    precondition under which both tiles stay in bounds."""
    import torch
    Mr, Nr, K, T = (np.asarray(x, np.int32) for x in (Mr, Nr, K, T))
    assert len(Mr) == vb.batch > 0
    tm = np.array([TILE[int(t)][0] for t in T]); tn = np.array([TILE[int(t)][1] for t in T])
    assert (Mr % tm == 0).all() and (Nr % tn == 0).all() and (Mr > 0).all() and (Nr > 0).all()
    assert (K >= 8).all() and (K % 8 == 0).all()
    assert lds_bytes >= (6144 if (T == 3).any() else 4096)
    for pool in (vb.A, vb.B, vb.C):
        assert pool.data_ptr() % 16 == 0
    assert (np.asarray(vb.offA) % 4 == 0).all() and (np.asarray(vb.offC) % 4 == 0).all()
    gx = int((Mr // tm).max()); gy = int((Nr // tn).max())
    dev = vb.C.device
    dM, dN, dK, dT = (torch.tensor(x, dtype=torch.int32, device=dev) for x in (Mr, Nr, K, T))
    row = torch.tensor(T == 3, device=dev)
    dA = torch.where(row, vb.dB, vb.dA)      # row-major A is the column-major B'
    dB = torch.where(row, vb.dA, vb.dB)
    stream = torch.cuda.current_stream().cuda_stream
    rc = ref_lib().ref_framework_gemm(vb.batch, gx, gy, lds_bytes, dM.data_ptr(), dN.data_ptr(),
                                      dK.data_ptr(), dA.data_ptr(), dB.data_ptr(),
                                      vb.dC.data_ptr(), dT.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rc == 0, rc


def compare(bg, orc, vb, Mr, Nr, K, T, lds_bytes):
    """Reference kernels vs FP32_SEQ (bit-exact) and vs the MFMA path (1e-5);
    the reference's output is itself checked against the FP64 oracle."""
    import torch
    C0 = vb.C.clone()
    run_reference(vb, Mr, Nr, K, T, lds_bytes)
    Cref = vb.C.clone()
    vb.C.copy_(C0)
    bg.Plan(vb.M, vb.N, vb.K, prec="fp32_seq").run(vb.dA, vb.dB, vb.dC, 1.0, 1.0)
    torch.cuda.synchronize()
    seq = vb.C.clone()
    vb.C.copy_(C0)
    bg.Plan(vb.M, vb.N, vb.K).run(vb.dA, vb.dB, vb.dC, 1.0, 1.0)
    torch.cuda.synchronize()
    mfma = vb.C.clone()
    c0h = C0.cpu().numpy(); refh = Cref.cpu().numpy(); mfh = mfma.cpu().numpy()
    worst = 0.0
    for b in range(vb.batch):
        m, n, k = int(vb.M[b]), int(vb.N[b]), int(vb.K[b])
        s = slice(int(vb.offC[b]), int(vb.offC[b]) + m * n)
        exact = orc.dgemm("N", "N", vb.host_flat("A", b), vb.host_flat("B", b), c0h[s], 1.0, 1.0,
                          m, n, k, m, k, m)
        assert orc.rel_fro(refh[s], exact) <= TOL, ("reference vs FP64", b)
        e = orc.rel_fro(mfh[s], refh[s].astype(np.float64))
        worst = max(worst, e)
        assert e <= TOL, ("MFMA vs reference", b, m, n, k, e)
    assert torch.equal(seq.view(torch.int32), Cref.view(torch.int32)), "FP32_SEQ differs from reference"
    return worst


def test_reference_32x32_tile(bg, orc, gpu):
    say("[reference] 32x32 column-major tile, 4096 B LDS")
    rng = np.random.default_rng(11)
    B = 40
    Mr = 32 * rng.integers(1, 9, B); Nr = 32 * rng.integers(1, 9, B); K = 8 * rng.integers(1, 65, B)
    T = np.full(B, 2)
    vb = make_problem(Mr, Nr, K, T, seed=11)
    compare(bg, orc, vb, Mr, Nr, K, T, 4096)


def test_reference_32x64_tile(bg, orc, gpu):
    say("[reference] 32x64 row-major tile, 8192 B LDS")
    rng = np.random.default_rng(12)
    B = 40
    Mr = 32 * rng.integers(1, 9, B); Nr = 64 * rng.integers(1, 5, B); K = 8 * rng.integers(1, 65, B)
    T = np.full(B, 3)
    vb = make_problem(Mr, Nr, K, T, seed=12)
    compare(bg, orc, vb, Mr, Nr, K, T, 8192)


@pytest.mark.parametrize("B", [64, 128, 256])
def test_reference_dispatcher_on_dataset(bg, orc, gpu, B):
    """The reference planner (wavefront_thres 6000, ROCm/framework/run.sh:4)
    on the first B dataset triples; the matrices it gives a correct tile are
    launched through the reference's dispatcher with its grid rule."""
    say(f"[reference] dispatcher, dataset B={B}")
    M, N, K = orc.read_dataset(os.path.join(ROOT, "tests", "golden", "data_MN_K_512_128"))
    M, N, K = M[:B], N[:B], K[:B]
    t = bg.ref_plan(M, N, K, 6000)["t"]
    keep = np.isin(t, (2, 3))
    assert keep.sum() >= 18
    Mr, Nr, Kr, T = M[keep], N[keep], K[keep], t[keep]
    vb = make_problem(Mr, Nr, Kr, T, seed=B)
    compare(bg, orc, vb, Mr, Nr, Kr, T, 8192 if (T == 3).any() else 4096)
