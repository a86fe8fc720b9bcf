"""Repeated batches as HIP graphs.

A plan's run is a fixed sequence of kernel launches whose arguments are the
plan's device work map, three device pointer arrays and alpha/beta; it
allocates nothing and issues no memset (queue counters reset themselves, the
FP16/FP8 amax buffer is cleared by a kernel), so it can be captured once and
replayed.  ``GraphedBatch`` captures ``plan.run`` on a private stream and
replays it on the caller's stream; operands of the same shapes are bound by
rewriting the captured pointer arrays in place (``set_pointers``), not by
capturing again.  Plans with kernel timing enabled refuse capture.

The reference runs every batch as a fresh launch with a device
synchronisation after it (ROCm/framework/gemm.cuh:243-250); replaying a
captured batch removes the per-launch host overhead for repeated batches.
"""
from __future__ import annotations

import torch

from .api import Plan


class GraphedBatch:
    """``plan.run(dA, dB, dC, alpha, beta)`` captured once, replayed many times.

    Construction runs the batch ``warmup`` times eagerly (so that every kernel
    of the plan is loaded before capture), then captures one run.  The pointer
    arrays are copied: the graph reads this object's own arrays, which
    ``set_pointers`` overwrites on the current stream.
    """

    def __init__(self, plan: Plan, dA: torch.Tensor, dB: torch.Tensor, dC: torch.Tensor,
                 alpha: float = 1.0, beta: float = 0.0, warmup: int = 1):
        for t in (dA, dB, dC):
            if t.dtype != torch.int64 or not t.is_cuda or t.numel() < plan.batch:
                raise ValueError("pointer arrays must be device int64 tensors of length batch")
        self.plan = plan
        self.alpha, self.beta = float(alpha), float(beta)
        self.dA, self.dB, self.dC = dA.clone(), dB.clone(), dC.clone()
        stream = torch.cuda.Stream(device=dC.device)
        stream.wait_stream(torch.cuda.current_stream(dC.device))
        with torch.cuda.stream(stream):
            for _ in range(warmup):
                plan.run(self.dA, self.dB, self.dC, self.alpha, self.beta, stream=stream)
        stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            plan.run(self.dA, self.dB, self.dC, self.alpha, self.beta, stream=stream)
        torch.cuda.current_stream(dC.device).wait_stream(stream)

    def set_pointers(self, dA: torch.Tensor | None = None, dB: torch.Tensor | None = None,
                     dC: torch.Tensor | None = None) -> None:
        """Bind other operands of the same shapes and layouts (device pointer
        arrays); takes effect for replays issued after it on the current stream."""
        for mine, new in ((self.dA, dA), (self.dB, dB), (self.dC, dC)):
            if new is not None:
                mine.copy_(new[: mine.numel()])

    def replay(self) -> None:
        """One run of the captured batch, ordered on the current stream."""
        self.graph.replay()
