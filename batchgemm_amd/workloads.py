"""Batches of column-major matrices in pooled device memory, plus the seeded
shape generators of the benchmark workloads and the reference dataset reader.

One pool per operand holds every matrix of a batch (each matrix 256-byte
aligned), so a batch of thousands of matrices costs three allocations, and the
device pointer arrays are base + offset.  This replaces the reference's one
hipMalloc per matrix (ROCm/framework/gemm.cuh:58-62).
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch

ALIGN = 64  # floats (256 bytes)


# ----------------------------------------------------------------- shapes
def read_dataset(path: str, batch: int | None = None):
    """Whitespace text of "M N K" triples, first `batch` lines
    (ROCm/framework/gemm.cuh:34-48)."""
    vals = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if batch is not None:
        if batch > len(vals):
            raise ValueError(f"{path} has {len(vals)} triples, {batch} requested")
        vals = vals[:batch]
    return [vals[:, i].astype(np.int32) for i in range(3)]


def shapes_uniform(batch: int, lo: int, hi: int, seed: int):
    """M, N, K iid uniform integers in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return [rng.integers(lo, hi + 1, size=batch, dtype=np.int64).astype(np.int32) for _ in range(3)]


def shapes_unbalanced(batch: int = 2048, frac_big: float = 0.1, small=(16, 128), big=(512, 2048),
                      seed: int = 0):
    """round(frac_big*batch) matrices with M,N,K ~ U{big}, the rest ~ U{small},
    at random positions in the batch."""
    rng = np.random.default_rng(seed)
    nbig = int(round(frac_big * batch))
    isbig = np.zeros(batch, bool)
    isbig[rng.permutation(batch)[:nbig]] = True
    out = []
    for _ in range(3):
        s = rng.integers(small[0], small[1] + 1, size=batch)
        b = rng.integers(big[0], big[1] + 1, size=batch)
        out.append(np.where(isbig, b, s).astype(np.int32))
    return out


def shapes_loguniform(batch: int, lo: int = 8, hi: int = 1024, seed: int = 0):
    """M, N, K iid log-uniform on [lo, hi], rounded to integers."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        x = np.exp(rng.uniform(math.log(lo), math.log(hi), size=batch))
        out.append(np.clip(np.rint(x), lo, hi).astype(np.int32))
    return out


def autotune_batches(n: int = 50, seed: int = 0, bmin: int = 256, bmax: int = 8192):
    """`n` successive batches: size ~ U{bmin..bmax}, M,N,K log-uniform [8, 1024]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(bmin, bmax + 1, size=n)
    return [shapes_loguniform(int(b), seed=seed * 1000 + i + 1) for i, b in enumerate(sizes)]


def zipf_groups(G: int = 128, total: int = 32768, s: float = 1.2, seed: int = 0):
    """Per-group row counts: Zipf(s) draws normalised to sum `total`, each >= 1,
    the rounding remainder given to the largest group."""
    rng = np.random.default_rng(seed)
    z = rng.zipf(s, size=G).astype(np.float64)
    m = np.maximum(1, np.floor(z / z.sum() * total)).astype(np.int64)
    m[np.argmax(m)] += total - m.sum()
    if m.min() < 1:  # the largest group absorbed a negative remainder below 1
        raise ValueError("cannot normalise Zipf draws")
    return m.astype(np.int32)


def flops(M, N, K) -> int:
    return int(2 * np.sum(np.asarray(M, np.int64) * np.asarray(N, np.int64) * np.asarray(K, np.int64)))


def algo_bytes(M, N, K, beta_nonzero: bool, elem: int = 4) -> int:
    """Compulsory HBM traffic: A, B read once, C written once (and read if beta != 0)."""
    M = np.asarray(M, np.int64); N = np.asarray(N, np.int64); K = np.asarray(K, np.int64)
    return int(elem * np.sum(M * K + K * N + M * N * (2 if beta_nonzero else 1)))


# ----------------------------------------------------------------- batches
def _offsets(sizes):
    sizes = np.asarray(sizes, np.int64)
    padded = (np.maximum(sizes, 1) + ALIGN - 1) // ALIGN * ALIGN
    off = np.zeros(len(sizes), np.int64)
    if len(sizes) > 1:
        off[1:] = np.cumsum(padded)[:-1]
    return off, int(padded.sum()) if len(sizes) else 0


class VBatch:
    """A batch of column-major FP32 matrices A_b (stored rows x cols with
    ld = rows), B_b, C_b in three device pools."""

    def __init__(self, M, N, K, transA="N", transB="N", device="cuda", pad_ab: int = 0):
        self.M = np.asarray(M, np.int32); self.N = np.asarray(N, np.int32)
        self.K = np.asarray(K, np.int32)
        self.batch = len(self.M)
        self.transA, self.transB = transA.upper(), transB.upper()
        ta, tb = self.transA == "T", self.transB == "T"
        self.a_rows = np.where(ta, self.K, self.M); self.a_cols = np.where(ta, self.M, self.K)
        self.b_rows = np.where(tb, self.N, self.K); self.b_cols = np.where(tb, self.K, self.N)
        # pad_ab > 0: A and B stored with leading dimension rows + pad_ab
        self.lda = (np.maximum(1, self.a_rows) + pad_ab).astype(np.int32)
        self.ldb = (np.maximum(1, self.b_rows) + pad_ab).astype(np.int32)
        self.ldc = np.maximum(1, self.M).astype(np.int32)
        self.offA, nA = _offsets(self.lda.astype(np.int64) * self.a_cols)
        self.offB, nB = _offsets(self.ldb.astype(np.int64) * self.b_cols)
        self.offC, nC = _offsets(self.ldc.astype(np.int64) * self.N)
        self.device = torch.device(device)
        self.A = torch.empty(max(nA, 1), dtype=torch.float32, device=device)
        self.B = torch.empty(max(nB, 1), dtype=torch.float32, device=device)
        self.C = torch.empty(max(nC, 1), dtype=torch.float32, device=device)
        self.dA = self._ptrs(self.A, self.offA)
        self.dB = self._ptrs(self.B, self.offB)
        self.dC = self._ptrs(self.C, self.offC)

    def _ptrs(self, pool, off):
        p = pool.data_ptr() + 4 * off
        return torch.from_numpy(p.astype(np.int64)).to(self.device)

    @property
    def nbytes(self) -> int:
        return 4 * (self.A.numel() + self.B.numel() + self.C.numel())

    def fill(self, dist: str = "uniform", seed: int = 0, outlier_frac: float = 0.0,
             outlier_scale: float = 1e3, c_dist: str | None = None):
        """A, B ~ U[-1,1] ("uniform") or N(0,1) ("normal"); optional Bernoulli
        outliers scaled by `outlier_scale`; C ~ `c_dist` (default zeros)."""
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        for pool in (self.A, self.B):
            if dist == "uniform":
                pool.uniform_(-1.0, 1.0, generator=g)
            elif dist == "normal":
                pool.normal_(0.0, 1.0, generator=g)
            else:
                raise ValueError(dist)
            if outlier_frac > 0:
                mask = torch.rand(pool.shape, generator=g, device=self.device) < outlier_frac
                pool[mask] *= outlier_scale
        if c_dist is None:
            self.C.zero_()
        elif c_dist == "uniform":
            self.C.uniform_(-1.0, 1.0, generator=g)
        elif c_dist == "normal":
            self.C.normal_(0.0, 1.0, generator=g)
        else:
            raise ValueError(c_dist)
        return self

    def _view(self, pool, off, rows, cols, ld):
        flat = pool[int(off): int(off) + int(ld) * int(cols)]
        return flat.view(int(cols), int(ld)).t()[: int(rows)] if cols > 0 else flat[:0].view(0, 0)

    def a(self, b):
        return self._view(self.A, self.offA[b], self.a_rows[b], self.a_cols[b], self.lda[b])

    def b(self, b):
        return self._view(self.B, self.offB[b], self.b_rows[b], self.b_cols[b], self.ldb[b])

    def c(self, b):
        return self._view(self.C, self.offC[b], self.M[b], self.N[b], self.ldc[b])

    def host_flat(self, which: str, b: int) -> np.ndarray:
        """Column-major storage of one matrix as a flat float32 numpy array."""
        pool, off, n = {
            "A": (self.A, self.offA, int(self.lda[b]) * int(self.a_cols[b])),
            "B": (self.B, self.offB, int(self.ldb[b]) * int(self.b_cols[b])),
            "C": (self.C, self.offC, int(self.ldc[b]) * int(self.N[b])),
        }[which]
        return pool[int(off[b]): int(off[b]) + n].cpu().numpy()


def dataset_path() -> str:
    """The reference's dataset (4096 "M N K" lines, ROCm/data/data_MN_K_512_128),
    shipped as package data: the benchmark driver's default input, as the
    reference's driver reads ../data/data_MN_K_512_128 (ROCm/framework/gemm.cuh:35)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "data_MN_K_512_128")
