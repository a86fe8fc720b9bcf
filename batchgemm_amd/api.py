"""Host-side API over the C ABI: plans, pointer arrays and torch-tensor entry points.

Mirrors the reference's host flow (ROCm/framework/gemm.cuh): per-matrix
dimension arrays and device pointer arrays in, one batched launch out -- with
hipBLAS column-major semantics (ROCm/default/gemm.cuh:82) and explicit
lda/ldb/ldc, alpha, beta.

Matrix convention for the tensor entry points: a column-major matrix with
``rows`` x ``cols`` and leading dimension ``ld`` is a 2-D torch tensor of shape
(rows, cols) with strides (1, ld) -- e.g. ``colmajor_empty(rows, cols)`` or the
``.t()`` of a row-major tensor.
"""
from __future__ import annotations

import contextlib
import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import PREC, check, default_opts, lib


def _lib_tile_shapes():
    return _lib.TILE_SHAPES


def _iarr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _stream_ptr(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream) if hasattr(stream, "cuda_stream") else int(stream)


def colmajor_empty(rows: int, cols: int, ld: int | None = None, device="cuda",
                   dtype=torch.float32) -> torch.Tensor:
    ld = max(1, rows) if ld is None else ld
    storage = torch.empty((max(cols, 1), ld), device=device, dtype=dtype)
    return storage.t()[:rows, :cols]


def ld_of(x: torch.Tensor) -> int:
    """Leading dimension of a column-major (strides (1, ld)) 2-D tensor."""
    if x.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    rows, cols = x.shape
    if cols <= 1:
        return max(1, rows, x.stride(1) if cols == 1 else 1)
    if x.stride(0) != 1 and rows > 1:
        raise ValueError(f"matrix is not column-major (strides {x.stride()})")
    return int(x.stride(1))


def pointer_array(ptrs: Sequence[int] | Sequence[torch.Tensor], device="cuda") -> torch.Tensor:
    """Device int64 array of device pointers (the reference's dev_A/dev_B/dev_C)."""
    vals = [p.data_ptr() if isinstance(p, torch.Tensor) else int(p) for p in ptrs]
    return torch.tensor(vals, dtype=torch.int64).to(device)


class Plan:
    """A planned batch: shapes, layouts and precision fixed; pointers per call.

    Creating a plan uploads the work map; ``run`` allocates nothing and may be
    captured in a HIP graph.
    """

    def __init__(self, M, N, K, lda=None, ldb=None, ldc=None, transA="N", transB="N",
                 prec="fp32", host_only=False, **opts):
        M = np.asarray(M, np.int32); N = np.asarray(N, np.int32); K = np.asarray(K, np.int32)
        self.batch = int(len(M))
        ta, tb = transA.upper() == "T", transB.upper() == "T"
        if lda is None:
            lda = np.maximum(1, K if ta else M)
        if ldb is None:
            ldb = np.maximum(1, N if tb else K)
        if ldc is None:
            ldc = np.maximum(1, M)
        self.M, self.N, self.K = M, N, K
        self.lda = np.asarray(lda, np.int32); self.ldb = np.asarray(ldb, np.int32)
        self.ldc = np.asarray(ldc, np.int32)
        self.transA, self.transB = transA.upper(), transB.upper()
        self.prec = prec
        o = default_opts(host_only=int(host_only), **opts)
        self._keep = [_iarr(v) for v in (M, N, K, self.lda, self.ldb, self.ldc)]
        p = ctypes.c_void_p()
        check(lib().vbg_plan_create(ctypes.byref(p), self.batch, *[k[1] for k in self._keep],
                                    self.transA.encode(), self.transB.encode(),
                                    PREC[prec] if isinstance(prec, str) else int(prec),
                                    ctypes.byref(o)), "vbg_plan_create")
        self._p = p

    # -- inspection
    def query(self, what: int) -> int:
        return check(lib().vbg_plan_query(self._p, what, None, 0), "vbg_plan_query")

    @property
    def n_items(self) -> int:
        return self.query(0)

    @property
    def n_launches(self) -> int:
        return self.query(1)

    @property
    def grid(self) -> int:
        return self.query(3)

    def tiles(self) -> np.ndarray:
        out = np.zeros(max(self.batch, 1), np.int32)
        check(lib().vbg_plan_query(self._p, 2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(out)), "vbg_plan_query")
        return out[:self.batch]

    def ksplit(self) -> np.ndarray:
        out = np.zeros(max(self.batch, 1), np.int32)
        check(lib().vbg_plan_query(self._p, 5, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(out)), "vbg_plan_query")
        return out[:self.batch]

    def tiled_items(self) -> np.ndarray:
        """Work items as rows (mat, tile id, tile_m, tile_n, k slice)."""
        n = self.n_items
        out = np.zeros(max(5 * n, 1), np.int32)
        check(lib().vbg_plan_query(self._p, 6, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(out)), "vbg_plan_query")
        return out[:5 * n].reshape(n, 5)

    def items(self) -> np.ndarray:
        n = self.n_items
        out = np.zeros(max(5 * n, 1), np.int32)
        check(lib().vbg_plan_query(self._p, 4, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(out)), "vbg_plan_query")
        return out[:5 * n].reshape(n, 5)

    KERNEL_KINDS = ("scale pre-pass", "fp32 mfma", "fp32 sequential", "split precision", "bf16",
                    "split-K reduction", "operand pack", "split precision (packed operands)")

    @property
    def packed_bytes(self) -> int:
        """Bytes of the plan's packed-operand workspace (0: operands are
        quantised on the way into LDS instead)."""
        return self.query(8)

    def kernels(self) -> np.ndarray:
        """Kernels a run issues, in order, as rows (kind, tile id, first item,
        item count, grid); kind indexes ``KERNEL_KINDS``, items index
        ``items()`` / ``tiled_items()``."""
        out = np.zeros(5 * 16, np.int32)
        n = check(lib().vbg_plan_query(self._p, 7, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       len(out)), "vbg_plan_query")
        return out[:n].reshape(-1, 5)

    def kernel_times(self) -> np.ndarray:
        """Mean ms per kernel (rows of ``kernels()``) over the recorded runs of
        a plan made with ``timing_slots > 0``; waits for those runs."""
        out = np.zeros(16, np.float32)
        n = check(lib().vbg_plan_launch_times(self._p, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float)), len(out)), "vbg_plan_launch_times")
        return out[:n].astype(np.float64)

    def item_flops(self) -> np.ndarray:
        """Algorithmic flops of every work item (2 * rows * cols * k of its
        part of the matrix, padding excluded), in ``items()`` order."""
        shapes = np.asarray(_lib_tile_shapes(), np.int64)
        ti = self.tiled_items().astype(np.int64)
        kr = self.items().astype(np.int64)
        if len(ti) == 0:
            return np.zeros(0)
        bm, bn = shapes[ti[:, 1], 0], shapes[ti[:, 1], 1]
        mat = ti[:, 0]
        rows = np.minimum(bm, self.M[mat].astype(np.int64) - ti[:, 2] * bm)
        cols = np.minimum(bn, self.N[mat].astype(np.int64) - ti[:, 3] * bn)
        return 2.0 * rows * cols * (kr[:, 4] - kr[:, 3])

    # -- execution
    def run(self, dA: torch.Tensor, dB: torch.Tensor, dC: torch.Tensor, alpha=1.0, beta=0.0,
            stream=None) -> None:
        """dA/dB/dC: device int64 tensors of `batch` device pointers."""
        for t in (dA, dB, dC):
            if t.dtype != torch.int64 or not t.is_cuda or t.numel() < self.batch:
                raise ValueError("pointer arrays must be device int64 tensors of length batch")
        check(lib().vbg_sgemm_vbatched(self._p, float(alpha), dA.data_ptr(), dB.data_ptr(),
                                       float(beta), dC.data_ptr(), _stream_ptr(stream)),
              "vbg_sgemm_vbatched")

    def close(self):
        if getattr(self, "_p", None):
            lib().vbg_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_plan(M, N, K, wavefront_thres: int = 6000):
    """The reference planner (ROCm/framework/gemm.cuh:121-229) as implemented
    in the library.  Returns None where the reference is undefined."""
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    b = len(m)
    t = np.zeros(b, np.int32); kt = np.zeros(b, np.int32); grid = np.zeros(3, np.int32)
    wf = np.zeros(2, np.int64)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib().vbg_ref_plan(b, pm, pn, pk, int(wavefront_thres), t.ctypes.data_as(ip),
                            kt.ctypes.data_as(ip), grid.ctypes.data_as(ip),
                            wf.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc == _lib.VBG_ERR_UNDEFINED:
        return None
    check(rc, "vbg_ref_plan")
    return dict(t=t, kernel_thread=kt, grid=tuple(int(g) for g in grid), wf_init=int(wf[0]),
                wf_final=int(wf[1]))


def sgemm_vbatched(As: Sequence[torch.Tensor], Bs: Sequence[torch.Tensor],
                   Cs: Sequence[torch.Tensor], alpha: float = 1.0, beta: float = 0.0,
                   transA: str = "N", transB: str = "N", prec: str = "fp32", stream=None,
                   **opts) -> Plan:
    """C_b = alpha * op(A_b) @ op(B_b) + beta * C_b for column-major tensors
    (strides (1, ld)).  Plans, runs asynchronously, returns the plan (reuse it
    with ``plan.run`` for repeated batches of the same shapes)."""
    if not (len(As) == len(Bs) == len(Cs)):
        raise ValueError("As, Bs, Cs must have the same length")
    ta, tb = transA.upper() == "T", transB.upper() == "T"
    M, N, K, lda, ldb, ldc = [], [], [], [], [], []
    for a, b, c in zip(As, Bs, Cs):
        for x in (a, b, c):
            if x.dtype != torch.float32 or not x.is_cuda:
                raise ValueError("operands must be float32 device tensors")
        m, n = c.shape
        k = a.shape[0] if ta else a.shape[1]
        if (a.shape[1] if ta else a.shape[0]) != m:
            raise ValueError("A rows do not match C")
        if (b.shape[1] if tb else b.shape[0]) != k or (b.shape[0] if tb else b.shape[1]) != n:
            raise ValueError("B shape does not match")
        M.append(m); N.append(n); K.append(k)
        lda.append(ld_of(a)); ldb.append(ld_of(b)); ldc.append(ld_of(c))
    plan = Plan(M, N, K, lda, ldb, ldc, transA, transB, prec, **opts)
    dev = Cs[0].device if len(Cs) else "cuda"
    # The kernels read the pointer arrays asynchronously on `stream`: allocate
    # them on that stream (the caching allocator then reuses their blocks only
    # in that stream's order) and keep them alive on the plan.
    ctx = torch.cuda.stream(stream) if isinstance(stream, torch.cuda.Stream) else contextlib.nullcontext()
    with ctx:
        ptrs = (pointer_array(As, dev), pointer_array(Bs, dev), pointer_array(Cs, dev))
    plan._pointer_arrays = ptrs
    plan.run(*ptrs, alpha, beta, stream)
    return plan
