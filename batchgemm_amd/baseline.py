"""The per-matrix library baseline: one vendor GEMM launch per matrix.

Counterpart of the reference's comparison program ROCm/default/gemm.cuh:64-85,
a loop of hipblasSgemm(OP_N, OP_N, M, N, K, alpha, A, M, B, K, beta, C, M) over
the batch.  Here each call is torch.addmm on the same column-major buffers the
batched plan uses (rocBLAS / hipBLASLt FP32 GEMM underneath): the "per-matrix
launch" comparator of the unbalanced workload and a sanity check of the
batched path's results on the GPU.
"""
from __future__ import annotations


def library_loop(vb, alpha: float, beta: float, transA: str = "N", transB: str = "N"):
    """A callable that runs C_b = alpha * op(A_b) op(B_b) + beta * C_b for every
    matrix of the VBatch `vb` as one library call each, in batch order."""
    import torch
    ta, tb = transA.upper() == "T", transB.upper() == "T"
    mats = []
    for b in range(vb.batch):
        a, bm, c = vb.a(b), vb.b(b), vb.c(b)
        mats.append((a.t() if ta else a, bm.t() if tb else bm, c))

    def run():
        for a, bm, c in mats:
            if c.numel() == 0:
                continue
            if a.shape[1] == 0:  # K = 0: C = beta * C
                c.mul_(beta)
                continue
            torch.addmm(c, a, bm, beta=beta, alpha=alpha, out=c)
    return run
