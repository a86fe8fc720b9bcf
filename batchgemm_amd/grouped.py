"""Grouped GEMM forward/backward over the C ABI, and its torch.autograd.Function.

Groups are stacked along rows: A is [S, K] with S = sum(m_sizes), group g owns
rows [off_g, off_g + m_g); B is [G, K, N]; C = A_g @ B_g is [S, N].  Backward:
dA_g = dC_g @ B_g^T, dB_g = A_g^T @ dC_g (both computed by the library's
grouped kernels on the same variable-size work map, FP32 accumulate).
Operands are BF16, contiguous, 16-byte aligned; K and N multiples of 8.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from ._lib import check, default_opts, lib
from .api import _stream_ptr

_DT = {torch.float32: 0, torch.bfloat16: 1}


class GroupedPlan:
    def __init__(self, m_sizes: Sequence[int], K: int, N: int, **opts):
        self.m_sizes = np.asarray(m_sizes, np.int32)
        self.G, self.K, self.N = len(self.m_sizes), int(K), int(N)
        self.S = int(self.m_sizes.sum())
        o = default_opts(**opts)
        p = ctypes.c_void_p()
        m = np.ascontiguousarray(self.m_sizes)
        check(lib().vbg_grouped_plan_create(ctypes.byref(p), self.G,
                                            m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            self.K, self.N, ctypes.byref(o)),
              "vbg_grouped_plan_create")
        self._p = p

    def _check(self, A, B):
        if A.dtype != torch.bfloat16 or B.dtype != torch.bfloat16:
            raise TypeError("grouped GEMM operands must be bfloat16")
        if tuple(A.shape) != (self.S, self.K) or tuple(B.shape) != (self.G, self.K, self.N):
            raise ValueError(f"expected A {(self.S, self.K)} and B {(self.G, self.K, self.N)}, "
                             f"got {tuple(A.shape)} and {tuple(B.shape)}")
        if not (A.is_contiguous() and B.is_contiguous()):
            raise ValueError("operands must be contiguous")

    def forward(self, A: torch.Tensor, B: torch.Tensor, out_dtype=torch.bfloat16,
                out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        self._check(A, B)
        if out is None:
            out = torch.empty((self.S, self.N), dtype=out_dtype, device=A.device)
        check(lib().vbg_grouped_gemm_fwd(self._p, A.data_ptr(), B.data_ptr(), out.data_ptr(),
                                         _DT[out.dtype], _stream_ptr(stream)), "vbg_grouped_gemm_fwd")
        return out

    def backward(self, A: torch.Tensor, B: torch.Tensor, dC: torch.Tensor, grad_dtype=torch.bfloat16,
                 need_dA=True, need_dB=True, stream=None):
        self._check(A, B)
        if dC.dtype != torch.bfloat16:
            dC = dC.to(torch.bfloat16)  # dC is an MFMA operand: BF16 only
        dC = dC.contiguous()
        if tuple(dC.shape) != (self.S, self.N):
            raise ValueError("dC shape mismatch")
        dA = torch.empty((self.S, self.K), dtype=grad_dtype, device=A.device) if need_dA else None
        dB = torch.empty((self.G, self.K, self.N), dtype=grad_dtype, device=A.device) if need_dB else None
        check(lib().vbg_grouped_gemm_bwd(self._p, A.data_ptr(), B.data_ptr(), dC.data_ptr(), 1,
                                         dA.data_ptr() if dA is not None else None,
                                         dB.data_ptr() if dB is not None else None,
                                         _DT[grad_dtype], _stream_ptr(stream)), "vbg_grouped_gemm_bwd")
        return dA, dB

    def close(self):
        if getattr(self, "_p", None):
            lib().vbg_grouped_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_plans: dict = {}


def grouped_plan(m_sizes, K: int, N: int) -> GroupedPlan:
    """The cached plan for these group sizes on the current device and stream.
    A plan's device pointer arrays and work-queue counters are per-run state,
    so calls that may overlap in time (different streams) get different
    plans; calls on one stream run in order and share one."""
    key = (tuple(int(m) for m in m_sizes), int(K), int(N), torch.cuda.current_device(),
           int(torch.cuda.current_stream().cuda_stream))
    p = _plans.get(key)
    if p is None:
        p = _plans[key] = GroupedPlan(m_sizes, K, N)
    return p


class GroupedGemmFunction(torch.autograd.Function):
    """C = grouped(A, B); backward through the library's grouped kernels."""

    @staticmethod
    def forward(ctx, A, B, plan: GroupedPlan, out_dtype, grad_dtype):
        ctx.save_for_backward(A, B)
        ctx.plan = plan
        ctx.grad_dtype = grad_dtype
        return plan.forward(A.contiguous(), B.contiguous(), out_dtype=out_dtype)

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        dA, dB = ctx.plan.backward(A.contiguous(), B.contiguous(), dC, grad_dtype=ctx.grad_dtype,
                                   need_dA=ctx.needs_input_grad[0], need_dB=ctx.needs_input_grad[1])
        if dA is not None and dA.dtype != A.dtype:
            dA = dA.to(A.dtype)
        if dB is not None and dB.dtype != B.dtype:
            dB = dB.to(B.dtype)
        return dA, dB, None, None, None


def grouped_gemm(A: torch.Tensor, B: torch.Tensor, m_sizes, out_dtype=torch.bfloat16,
                 grad_dtype=torch.bfloat16) -> torch.Tensor:
    """Differentiable grouped GEMM: rows of A grouped by `m_sizes`, times B[g]."""
    plan = grouped_plan(m_sizes, A.shape[1], B.shape[2])
    return GroupedGemmFunction.apply(A, B, plan, out_dtype, grad_dtype)


def bench_config4(steps: int = 20, warmup: int = 3, seed: int = 0, library_reference: bool = False) -> dict:
    """Grouped forward and backward on the differentiation workload: G = 128
    groups, K = 2048, N = 1408, Zipf(1.2) row counts summing to 32768, BF16
    operands ~ N(0, 0.02), BF16 dC ~ N(0, 1), BF16 outputs; mean ms per call
    over `steps` calls after `warmup`, from HIP events on the current stream.
    With `library_reference`, also the largest group's products through
    torch.mm (hipBLASLt) on the same device."""
    from .workloads import zipf_groups
    m_sizes = zipf_groups(128, 32768, 1.2, seed=seed)
    K, N, S = 2048, 1408, 32768
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = (torch.randn(S, K, generator=g, device="cuda") * 0.02).to(torch.bfloat16)
    B = (torch.randn(128, K, N, generator=g, device="cuda") * 0.02).to(torch.bfloat16)
    dC = torch.randn(S, N, generator=g, device="cuda").to(torch.bfloat16)
    plan = GroupedPlan(m_sizes, K, N)
    C = torch.empty(S, N, dtype=torch.bfloat16, device="cuda")
    dA, dB = torch.empty_like(A), torch.empty_like(B)
    L = lib()

    def fwd():
        check(L.vbg_grouped_gemm_fwd(plan._p, A.data_ptr(), B.data_ptr(), C.data_ptr(), 1,
                                     _stream_ptr(None)), "vbg_grouped_gemm_fwd")

    def bwd():
        check(L.vbg_grouped_gemm_bwd(plan._p, A.data_ptr(), B.data_ptr(), dC.data_ptr(), 1, dA.data_ptr(),
                                     dB.data_ptr(), 1, _stream_ptr(None)), "vbg_grouped_gemm_bwd")

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    f_ms, b_ms = timed(fwd), timed(bwd)
    gfl = 2.0 * S * K * N / 1e9
    out = {"workload": "G=128, K=2048, N=1408, Zipf(1.2) rows summing to 32768, BF16 in/out, FP32 accumulate",
           "fwd_ms": round(f_ms, 4), "bwd_ms": round(b_ms, 4), "fwd_bwd_ms": round(f_ms + b_ms, 4),
           "fwd_tflops": round(gfl / f_ms, 1), "bwd_tflops": round(2 * gfl / b_ms, 1)}
    if library_reference:
        # The largest group's three products as single 2-D torch.mm calls
        # (hipBLASLt BF16 GEMMs, FP32 accumulate, BF16 out) on the same
        # tensors: a same-device reference point for the dense part of each
        # pass (the small groups are not included).
        g0 = int(np.argmax(m_sizes))
        r0 = int(np.sum(m_sizes[:g0]))
        Ag, dCg, Bg = A[r0:r0 + m_sizes[g0]], dC[r0:r0 + m_sizes[g0]], B[g0]
        lib_f = timed(lambda: torch.mm(Ag, Bg))
        lib_da = timed(lambda: torch.mm(dCg, Bg.t()))
        lib_db = timed(lambda: torch.mm(Ag.t(), dCg))
        gfl0 = 2.0 * int(m_sizes[g0]) * K * N / 1e9
        out["library_largest_group"] = {
            "what": f"torch.mm (hipBLASLt) on the largest group, {int(m_sizes[g0])} x {K} x {N} BF16",
            "fwd_ms": round(lib_f, 4), "da_ms": round(lib_da, 4), "db_ms": round(lib_db, 4),
            "fwd_tflops": round(gfl0 / lib_f, 1), "da_tflops": round(gfl0 / lib_da, 1),
            "db_tflops": round(gfl0 / lib_db, 1)}
    return out


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="grouped BF16 forward/backward timing (differentiation workload)")
    ap.add_argument("--bench", action="store_true", help="time the G=128, K=2048, N=1408 workload")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--library-reference", action="store_true",
                    help="also time the largest group's products with torch.mm (hipBLASLt)")
    args = ap.parse_args()
    if not args.bench:
        ap.error("nothing to do: pass --bench")
    print(json.dumps(bench_config4(args.steps, args.warmup, library_reference=args.library_reference)),
          flush=True)
