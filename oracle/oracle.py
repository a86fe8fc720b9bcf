"""CPU oracle for the variable-size batched SGEMM path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``batchgemm_amd``) imports this module; only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use it, and
only as the checker.

Two layers:
  * ``liboracle.so`` (``oracle.c``) -- the C restatement of the reference path:
    dataset ingest and FLOP accounting (ROCm/framework/gemm.cuh:34-48,112-116),
    the planner (gemm.cuh:121-229), the sequential-fma tile semantics
    (ROCm/framework/kernel_128_new.h:131-201, kernel_256_new.h:2-72), an FP64
    column-major GEMM with hipBLAS conventions (ROCm/default/gemm.cuh:82) and the
    host per-matrix SGEMM loop used as the CPU baseline.
  * numpy emulation of the operand quantisation used by the split-precision
    paths: bf16 round-to-nearest-even and the hi/lo split, FP16 with a per-matrix
    power-of-two scale, FP8 e4m3 (OCP "fn") with a per-matrix amax scale.  These
    have no counterpart in the reference (it is FP32 only); they are pinned in
    tests against torch's own dtype conversions (a library routine), not against
    a retyped formula.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORC_LIBRARY selects another build of oracle.c (the sanitized host test run
# loads liboracle_asan.so this way)
_LIB_PATH = os.environ.get("ORC_LIBRARY") or os.path.join(_HERE, "liboracle.so")
_lib = None

ORC_OK = 0
ORC_UNDEFINED = -3


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_read_dataset.argtypes = [ctypes.c_char_p, ctypes.c_int, ip, ip, ip]
        L.orc_flops_ref.argtypes = [ctypes.c_int, ip, ip, ip]
        L.orc_flops_ref.restype = ctypes.c_double
        L.orc_flops.argtypes = [ctypes.c_int, ip, ip, ip]
        L.orc_flops.restype = ctypes.c_int64
        L.orc_ref_plan.argtypes = [ctypes.c_int, ip, ip, ip, ctypes.c_int, ip, ip, ip, ip,
                                   ctypes.POINTER(ctypes.c_int64)]
        fpp = ctypes.POINTER(ctypes.c_void_p)
        L.orc_dgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.orc_dgemm_vbatched.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ip, ip, ip,
                                         ctypes.c_double, fpp, ip, fpp, ip, ctypes.c_double,
                                         fpp, ip, fpp]
        L.orc_sgemm_vbatched.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ip, ip, ip,
                                         ctypes.c_float, fpp, ip, fpp, ip, ctypes.c_float, fpp,
                                         ip, ctypes.c_int]
        L.orc_ref_tile_seq.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _iarr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _parr(arrs):
    p = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return p


# ---------------------------------------------------------------- dataset / flops
def read_dataset(path: str, max_lines: int = 1 << 20):
    M = np.zeros(max_lines, np.int32)
    N = np.zeros(max_lines, np.int32)
    K = np.zeros(max_lines, np.int32)
    n = lib().orc_read_dataset(path.encode(), max_lines,
                               M.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               N.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               K.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if n < 0:
        raise OSError(f"cannot read {path}")
    return M[:n].copy(), N[:n].copy(), K[:n].copy()


def flops_ref(M, N, K) -> float:
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    return lib().orc_flops_ref(len(m), pm, pn, pk)


def flops(M, N, K) -> int:
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    return int(lib().orc_flops(len(m), pm, pn, pk))


# ---------------------------------------------------------------- planner
def ref_plan(M, N, K, thres: int):
    """Reference planner (gemm.cuh:121-229).  Returns None where the reference
    is undefined, else dict(t=..., t_init=..., kernel_thread=..., grid=(x,y,z),
    wf_init=..., wf_final=...)."""
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    b = len(m)
    t = np.zeros(b, np.int32); kt = np.zeros(b, np.int32); ti = np.zeros(b, np.int32)
    grid = np.zeros(3, np.int32); wf = np.zeros(2, np.int64)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib().orc_ref_plan(b, pm, pn, pk, int(thres), t.ctypes.data_as(ip), kt.ctypes.data_as(ip),
                            ti.ctypes.data_as(ip), grid.ctypes.data_as(ip),
                            wf.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc == ORC_UNDEFINED:
        return None
    if rc != ORC_OK:
        raise ValueError(f"orc_ref_plan rc={rc}")
    return dict(t=t, t_init=ti, kernel_thread=kt, grid=tuple(int(g) for g in grid),
                wf_init=int(wf[0]), wf_final=int(wf[1]))


# ---------------------------------------------------------------- GEMM oracles
def dgemm(ta, tb, A, B, C, alpha, beta, M, N, K, lda, ldb, ldc):
    """FP64 column-major GEMM.  A, B, C are flat float32 arrays holding the
    column-major storage; returns a flat float64 array with ld = ldc."""
    A = np.ascontiguousarray(A, np.float32); B = np.ascontiguousarray(B, np.float32)
    C = np.ascontiguousarray(C, np.float32)
    out = np.zeros(max(ldc * N, 1), np.float64)
    rc = lib().orc_dgemm(ta.encode(), tb.encode(), M, N, K, alpha, A.ctypes.data, lda,
                         B.ctypes.data, ldb, beta, C.ctypes.data, ldc, out.ctypes.data, ldc)
    assert rc == 0
    return out


def dgemm_vbatched(ta, tb, M, N, K, alpha, As, lda, Bs, ldb, beta, Cs, ldc):
    b = len(M)
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    la, pla = _iarr(lda); lb, plb = _iarr(ldb); lc, plc = _iarr(ldc)
    As = [np.ascontiguousarray(a, np.float32) for a in As]
    Bs = [np.ascontiguousarray(x, np.float32) for x in Bs]
    Cs = [np.ascontiguousarray(x, np.float32) for x in Cs]
    outs = [np.zeros(max(int(lc[i]) * int(n[i]), 1), np.float64) for i in range(b)]
    rc = lib().orc_dgemm_vbatched(ta.encode(), tb.encode(), b, pm, pn, pk, alpha, _parr(As), pla,
                                  _parr(Bs), plb, beta, _parr(Cs), plc, _parr(outs))
    assert rc == 0
    return outs


def sgemm_vbatched(ta, tb, M, N, K, alpha, As, lda, Bs, ldb, beta, Cs, ldc, nthreads=0):
    """Host per-matrix FP32 SGEMM loop (the CPU baseline).  Updates Cs in place."""
    b = len(M)
    m, pm = _iarr(M); n, pn = _iarr(N); k, pk = _iarr(K)
    la, pla = _iarr(lda); lb, plb = _iarr(ldb); lc, plc = _iarr(ldc)
    for x in list(As) + list(Bs) + list(Cs):
        assert x.dtype == np.float32 and x.flags.c_contiguous
    rc = lib().orc_sgemm_vbatched(ta.encode(), tb.encode(), b, pm, pn, pk, alpha, _parr(As), pla,
                                  _parr(Bs), plb, beta, _parr(Cs), plc, int(nthreads))
    assert rc == 0


def ref_tile_seq(rowmajor: bool, M, N, K, A, B, C):
    """Reference tile semantics: C <- C + A*B as one sequential fma chain per
    element over ascending k, seeded from C.  Returns the new C (float32)."""
    A = np.ascontiguousarray(A, np.float32).ravel()
    B = np.ascontiguousarray(B, np.float32).ravel()
    C = np.array(C, np.float32).ravel()
    rc = lib().orc_ref_tile_seq(int(rowmajor), M, N, K, A.ctypes.data, B.ctypes.data, C.ctypes.data)
    assert rc == 0
    return C


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------- quantisation emulation
def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def bf16_split(x):
    """hi = bf16(x), lo = bf16(x - hi); x - hi is exact in float32."""
    x = np.asarray(x, np.float32)
    hi = bf16_round(x)
    lo = bf16_round((x - hi).astype(np.float32))
    return hi, lo


FP16_EXP_MAX = 126  # scale 2**e stays a finite float32 for a tiny amax


def fp16_pow2_exponent(amax: float) -> int:
    """Per-matrix scale exponent e with amax * 2**e in [2**14, 2**15):
    amax = f * 2**x with f in [0.5, 1)  ->  e = 15 - x, capped at 126 (amax
    below 2**-111 keeps a finite scale).  amax == 0 -> 0."""
    amax = float(np.float32(amax))
    if amax == 0.0 or not np.isfinite(amax):
        return 0
    _, x = np.frexp(np.float32(amax))
    return int(min(15 - x, FP16_EXP_MAX))


def fp16_scaled(x, e: int):
    """Quantise x * 2**e to IEEE binary16 (round to nearest even)."""
    xs = np.ldexp(np.asarray(x, np.float32), e).astype(np.float32)
    return xs.astype(np.float16).astype(np.float32)


E4M3_MAX = 448.0


FP8_SCALE_MAX = np.float32(2.0 ** 126)


def fp8_scale(amax: float) -> np.float32:
    """Per-matrix FP8 scale s = 448 / amax in float32, capped at 2**126
    (amax == 0 -> 1)."""
    amax = np.float32(amax)
    if amax == 0 or not np.isfinite(amax):
        return np.float32(1.0)
    with np.errstate(over="ignore"):
        s = np.float32(np.float32(E4M3_MAX) / amax)
    return np.float32(min(s, FP8_SCALE_MAX))


def post_scale(prec: str, a, b) -> float:
    """The factor that undoes both operand scales, as the scaled paths apply
    it to the FP32 result.  FP16: a, b are the exponents eA, eB and the factor
    is 2**-(eA + eB), computed exactly in FP64 and rounded once to float32
    (a power of two: exact down to 2**-149).  FP8: a, b are the scales sA, sB
    and the factor is (1/sA) * (1/sB) with each step rounded to float32."""
    if prec == "fp16":
        return float(np.float32(np.ldexp(1.0, -(int(a) + int(b)))))
    if prec == "fp8":
        ia = np.float32(np.float32(1.0) / np.float32(a))
        ib = np.float32(np.float32(1.0) / np.float32(b))
        return float(np.float32(ia * ib))
    raise ValueError(prec)


def e4m3_round(x):
    """float32 -> OCP e4m3 ("fn": bias 7, max 448, subnormals down to 2**-9),
    round to nearest even, saturating at +-448.  Returned as float32."""
    x = np.asarray(x, np.float32)
    a = np.abs(x).astype(np.float64)
    nz = a > 0
    # a = f * 2**x, f in [0.5, 1): the binade exponent is x - 1, clamped to the
    # smallest normal exponent (-6) so that subnormals share its quantum
    _, xexp = np.frexp(np.where(nz, a, 1.0))
    e = np.maximum(xexp.astype(np.int64) - 1, -6)
    q = np.ldexp(1.0, e - 3)  # quantum: 3 mantissa bits
    out = np.where(nz, np.round(a / q) * q, 0.0)  # np.round: ties to even
    out = np.minimum(out, E4M3_MAX)
    return (np.sign(x) * out).astype(np.float32)


def fp8_scaled(x, s):
    xs = (np.asarray(x, np.float32) * np.float32(s)).astype(np.float32)
    return e4m3_round(xs)


def split_product(prec: str, A, B):
    """op(A) @ op(B) as a split-precision path computes it, up to its FP32
    accumulation: the operands quantised as the kernels quantise them (bf16x3:
    hi/lo split, C = hi*hi + hi*lo + lo*hi; fp16 / fp8: per-matrix scale from
    each operand's amax, then rounding), multiplied exactly in FP64 and scaled
    back by the kernels' post-scale.  A, B: float32 arrays (M x K, K x N)."""
    A = np.asarray(A, np.float32); B = np.asarray(B, np.float32)
    d = lambda x: x.astype(np.float64)
    if prec == "bf16x3":
        ah, al = bf16_split(A); bh, bl = bf16_split(B)
        return d(ah) @ d(bh) + d(ah) @ d(bl) + d(al) @ d(bh)
    amax_a = float(np.abs(A).max()) if A.size else 0.0
    amax_b = float(np.abs(B).max()) if B.size else 0.0
    if prec == "fp16":
        ea, eb = fp16_pow2_exponent(amax_a), fp16_pow2_exponent(amax_b)
        return (d(fp16_scaled(A, ea)) @ d(fp16_scaled(B, eb))) * post_scale("fp16", ea, eb)
    if prec == "fp8":
        sa, sb = fp8_scale(amax_a), fp8_scale(amax_b)
        return (d(fp8_scaled(A, sa)) @ d(fp8_scaled(B, sb))) * post_scale("fp8", sa, sb)
    raise ValueError(prec)


def rel_fro(x, ref) -> float:
    x = np.asarray(x, np.float64).ravel(); ref = np.asarray(ref, np.float64).ravel()
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (d if d > 0 else 1.0))


def wang_order(M, N, K) -> np.ndarray:
    """Order of ROCm/data/sort.cpp:17-60 (direction 1): stable by K
    descending, then stable by tile class -- M and N both multiples of 64
    (class 0), of 32 (class 1), the rest (class 2).  Returns the
    permutation (indices into the batch)."""
    M = np.asarray(M, np.int64); N = np.asarray(N, np.int64); K = np.asarray(K, np.int64)
    cls = np.where((M % 64 == 0) & (N % 64 == 0), 0, np.where((M % 32 == 0) & (N % 32 == 0), 1, 2))
    return np.lexsort((np.arange(len(M)), -K, cls))  # last key primary; index keeps it stable


def wang_plan(M, N, K, thres: int) -> np.ndarray:
    """Tile index per matrix (0/1/2 = 16x16/32x32/64x64) of the Wang
    planner, ROCm/tiling/gemm_wang.cuh:104-182 run on the sort.cpp order
    (gemm_wang.cuh:33 reads the sorted file), with one threshold for both of
    its TLP tests; workgroup counts round up for sides the tile does not
    divide (the reference assumes it divides)."""
    M = [int(x) for x in M]; N = [int(x) for x in N]
    B = len(M)
    ts = (16, 32, 64)
    t = 0
    while t < 3:
        tlp = sum((M[j] // ts[t]) * (N[j] // ts[t]) * 128 for j in range(B))
        if tlp < thres:
            break
        t += 1
    t = min(t, 2)
    fits = lambda j, k: ts[k] <= M[j] and ts[k] <= N[j] and M[j] % ts[k] == 0 and N[j] % ts[k] == 0
    st = [0] * B
    for j in range(B):
        if fits(j, t):
            st[j] = t
        else:
            for k in range(t):
                if fits(j, k):
                    st[j] = k
    cdiv = lambda a, b: -(-a // b)
    wg = sum(cdiv(M[j], ts[st[j]]) * cdiv(N[j], ts[st[j]]) for j in range(B))
    order = wang_order(M, N, K)
    for jj in range(B - 1, -1, -1):
        if wg >= thres:
            break
        j = int(order[jj])
        if st[j] == 0:
            continue
        wg += cdiv(M[j], ts[st[j]]) * cdiv(N[j], ts[st[j]]) * 3
        st[j] -= 1
    return np.array(st, np.int32)


LI_TILES = ((16, 16), (32, 32), (64, 64), (128, 64), (64, 128), (128, 128))


def li_plan(M, N, thres: int) -> np.ndarray:
    """Tile index per matrix over LI_TILES (the library's tile ids) of the Li
    tiling strategy, CUDA/batching/gemm.cu:99-127: the first tile whose
    TLP = sum (M/tm)(N/tn)*256 is below thres (the last if none); per
    matrix that tile when tm <= M and tn <= N, else the last smaller tile
    with tm == M and tn <= N, else 16x16."""
    M = [int(x) for x in M]; N = [int(x) for x in N]
    t = 0
    while t < 6:
        if sum((M[j] // LI_TILES[t][0]) * (N[j] // LI_TILES[t][1]) * 256 for j in range(len(M))) < thres:
            break
        t += 1
    t = min(t, 5)
    out = []
    for m, n in zip(M, N):
        s = 0
        if LI_TILES[t][0] <= m and LI_TILES[t][1] <= n:
            s = t
        else:
            for k in range(t):
                if LI_TILES[k][0] == m and LI_TILES[k][1] <= n:
                    s = k
        out.append(s)
    return np.array(out, np.int32)
