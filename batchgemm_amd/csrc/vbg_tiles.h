// vbg_tiles.h -- device building blocks shared by the gfx950 GEMM kernels:
// vector types, the global->register->LDS operand loaders with edge masking,
// and the persistent work-queue driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vbg_internal.h"

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
// 4-byte-aligned float4 for HBM accesses at arbitrary element offsets (ld and
// tile origins need not be multiples of 4); gfx950 runs these as dwordx4.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned v4u32_t __attribute__((ext_vector_type(4)));

#define VBG_THREADS 256

// Epilogue: the next chunk of C is loaded before the current one is stored
// (unbalanced B=2048 batch, three runs each: bf16x3 3.76 vs 3.79 ms, FP32
// 7.51-7.52 vs 7.52-7.53 ms).
#ifndef VBG_EPI_PIPE
#define VBG_EPI_PIPE 1
#endif


// ---------------------------------------------------------------------------
// Wave-uniform views.  Work items, descriptors and operand pointers are the
// same for every lane of a workgroup, but the compiler cannot prove it for
// values loaded from global memory (the kernel also stores to global memory);
// readfirstlane moves them to scalar registers, so buffer resources, k offsets
// and loop bounds stay scalar and no per-lane address arithmetic is generated.
__device__ __forceinline__ int vbg_u(int x) { return __builtin_amdgcn_readfirstlane(x); }

template <class T>
__device__ __forceinline__ T* vbg_uptr(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t vbg_u64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Element i (wave-uniform) of plan data that no kernel writes while it runs:
// work items, descriptors, operand pointer arrays, offsets and scales.  Read
// through the constant address space, which the compiler turns into scalar
// loads straight into SGPRs, counted by lgkmcnt -- so these loads never wait
// behind (or hold up) the vector memory operations in flight, such as the
// next item's claim or the previous item's C stores.
template <class T>
__device__ __forceinline__ T vbg_cload(const T* p, int64_t i) {
  static_assert(sizeof(T) % 4 == 0, "plan data is made of 4-byte words");
  typedef const __attribute__((address_space(4))) uint32_t CW;
  const CW* w = (const CW*)(uintptr_t)(p + i);
  struct Words {
    uint32_t x[sizeof(T) / 4];
  } r;
#pragma unroll
  for (int j = 0; j < (int)(sizeof(T) / 4); ++j) r.x[j] = w[j];
  return __builtin_bit_cast(T, r);
}

__device__ __forceinline__ vbg_mat_desc vbg_udesc(const vbg_mat_desc& d) {
  return {vbg_u(d.M), vbg_u(d.N), vbg_u(d.K), vbg_u(d.lda), vbg_u(d.ldb), vbg_u(d.ldc),
          vbg_u(d.tiles_n), vbg_u(d.ksplit)};
}

__device__ __forceinline__ vbg_item vbg_uitem(const vbg_item& it) {
  int packed;  // kslice | tile << 16, one scalar register
  __builtin_memcpy(&packed, &it.kslice, sizeof(packed));
  packed = vbg_u(packed);
  return {vbg_u(it.mat), vbg_u(it.tm), vbg_u(it.tn), (short)(packed & 0xffff), (short)(packed >> 16)};
}

// ---------------------------------------------------------------------------
// Buffer resource over one operand: raw buffer loads take a per-lane 32-bit
// byte offset (VGPR) plus a wave-uniform byte offset (SGPR), and return 0 for
// any access past `bytes` instead of faulting.  Third dword 0x00020000 is the
// gfx9-family data-format word for untyped dword access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vbg_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// ---------------------------------------------------------------------------
// LDS-DMA: buffer loads whose destination is LDS (buffer_load_dwordx4 ... lds).
// One LDS-DMA instruction: 16 bytes per lane from the buffer `rs` at byte
// offset voff + soff into LDS at lds_base + 16 * lane.  Issued as inline asm
// (it has no VGPR result: its destination is LDS) so that the compiler's
// load-counter model does not see it and inserts no waits of its own for it
// (merging its per-buffer DMA tracking at the loop back-edge, it drained every
// DMA in flight at the top of each unrolled group of k-tiles).  The waits are
// the explicit vmcnt(N) before each k-tile's barrier; no compiler-visible
// vector-memory load is issued between the first DMA and the final vmcnt(0),
// so those counts are exact.  One wait state separates the M0 write from the
// DMA that reads it.
__device__ __forceinline__ void vbg_dma16(v4u32_t rs, uint32_t lds_base, int voff, int soff) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, %3 offen lds"
               ::"s"(lds_base), "v"(voff), "s"(rs), "s"(soff)
               : "memory", "m0");
}

// Buffer descriptor words (base, stride 0, `bytes` records, raw dword
// format) in scalar registers, for vbg_dma16.
__device__ __forceinline__ v4u32_t vbg_rsrc_words(const void* p, uint32_t bytes) {
  const uint64_t a = (uint64_t)p;
  return v4u32_t{(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a),
                 (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32) & 0xffffu),
                 (uint32_t)__builtin_amdgcn_readfirstlane(bytes), 0x00020000u};
}

// s_waitcnt vmcnt(N) lgkmcnt(0), then s_barrier, with no memory access moved
// across either; earlier DMAs beyond the newest N stay in flight
// (__syncthreads() would wait for all of them).  The wait is the builtin, not
// inline asm, so the compiler's own wait tracking learns that every LDS read
// issued before it has completed (after an opaque asm it would instead wait
// for all LDS reads at the next use of any of them).
template <int N>
__device__ __forceinline__ void vbg_wait_vm_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  // gfx9 s_waitcnt immediate: vmcnt[3:0] at 3:0 and vmcnt[5:4] at 15:14,
  // expcnt at 6:4 (7: no wait), lgkmcnt at 11:8
  constexpr int imm = (N & 15) | ((N >> 4) << 14) | (7 << 4);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(imm);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---------------------------------------------------------------------------
// Operand tile loader.  Stages a ROWS x BK slice of one operand -- rows are the
// operand's M (for A) or N (for B) index "mn", columns the reduction index k --
// through registers.  r[v] holds 4 consecutive k of one row (coords(v) names
// the row and the k-quad), which the tile writes to LDS laid out [mn][k] (k
// contiguous), so that an MFMA lane reads consecutive k with one ds_read.
//
// KC = true : element (mn, k) at p[mn*ld + k]  (k contiguous in HBM: A^T, B);
//             one 16-byte buffer load per r[v].
// KC = false: element (mn, k) at p[mn + k*ld]  (mn contiguous in HBM: A, B^T);
//             4 dword loads per r[v], one per k, each wave-coalesced along mn.
//
// Addressing: init() fixes each lane's byte offset for its row and k-quad once
// per work item (rows past MN clamped to MN-1: they only feed outputs that are
// never stored); a whole k-tile's position is a wave-uniform scalar offset, so
// the main loop issues no vector address arithmetic.  The scalar offset is not
// range-checked by the buffer unit, so it is only used where every address is
// inside the operand.  The last (partial) k-tile addresses each element with
// its k clamped to K-1, entirely in the range-checked vector offset, and
// zeroes k >= K by a select: no load leaves the operand, whatever lies after
// it.  The buffer extent is the operand's own span (last element + 1).
// Operands must span less than 2^31 bytes (the plan checks).
template <int ROWS, int BK, int P, bool KC>
struct TileLoader {
  static constexpr int NV = ROWS * BK / (4 * VBG_THREADS);  // float4 groups per thread
  static_assert(NV >= 1 && ROWS * BK % (4 * VBG_THREADS) == 0, "tile too small for the block");
  typedef v4f Regs[NV];
  __amdgpu_buffer_rsrc_t rsrc;
  int voff[NV];
  int ld4;

  // MC with ROWS >= 32: each half-wave takes 32 consecutive rows (128-byte
  // coalesced loads) and the two halves of a wave take neighbouring k-quads,
  // so one LDS store instruction covers every bank.
  __device__ __forceinline__ static void coords(int v, int& row, int& kq) {
    const int idx = threadIdx.x + VBG_THREADS * v;
    if (KC) {
      row = idx / (BK / 4);
      kq = idx % (BK / 4);
    } else if (ROWS >= 32) {
      constexpr int RB = ROWS / 32;
      const int l = idx & 63, w = idx >> 6;
      row = (l & 31) + 32 * (w % RB);
      kq = 2 * (w / RB) + (l >> 5);
    } else {
      row = idx % ROWS;
      kq = idx / ROWS;
    }
  }

  // p: operand base; MN rows; K the operand's full reduction length.
  __device__ __forceinline__ void init(const float* p, int ld, int mn0, int MN, int K) {
    // k contiguous: rows x K at ld; mn contiguous: K columns of MN at ld
    const uint32_t extent =
        (MN <= 0 || K <= 0) ? 0u
        : KC ? ((uint32_t)(MN - 1) * (uint32_t)ld + (uint32_t)K) * 4u
             : ((uint32_t)(K - 1) * (uint32_t)ld + (uint32_t)MN) * 4u;
    rsrc = vbg_rsrc(p, extent);
    ld4 = ld * 4;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int row, kq;
      coords(v, row, kq);
      const int mnc = min(mn0 + row, MN - 1);
      voff[v] = KC ? (mnc * ld + 4 * kq) * 4 : (mnc + 4 * kq * ld) * 4;
    }
  }

  // A whole k-tile starting at k0 (k0 + BK <= K).
  __device__ __forceinline__ void load(Regs& r, int k0) const {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (KC) {
        r[v] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff[v], k0 * 4, 0));
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          r[v][e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff[v], (k0 + e) * ld4, 0));
      }
    }
  }

  // The last k-tile [k0, K) (K: the end of the item's k range, at most the
  // operand's K): k >= K reads as zero, and its address is that of k = K-1.
  __device__ __forceinline__ void load_tail(Regs& r, int k0, int K) const {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int row, kq;
      coords(v, row, kq);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + 4 * kq + e;
        const int dk = min(k, K - 1) - 4 * kq;  // k offset from the lane's k-quad origin
        const unsigned bits = __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff[v] + dk * (KC ? 4 : ld4), 0, 0);
        const float x = __builtin_bit_cast(float, bits);
        r[v][e] = k < K ? x : 0.0f;
      }
    }
  }

  __device__ __forceinline__ void fetch(Regs& r, int k0, int K) const {
    if (k0 + BK <= K) load(r, k0);
    else load_tail(r, k0, K);
  }

  __device__ __forceinline__ static void store(const Regs& r, float* lds) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int row, kq;
      coords(v, row, kq);
      *(v4f*)(lds + row * P + 4 * kq) = r[v];
    }
  }
};

// ---------------------------------------------------------------------------
// Persistent driver: each workgroup pops work items until every queue is
// empty.  Queue q holds items [q_begin[q], q_begin[q+1]) in cost order; a
// workgroup starts on queue blockIdx.x % nq (its XCD when nq == 8, since
// workgroup ids are dealt round-robin to the XCDs) and moves on to the next
// queues once its own is drained.  The last workgroup to finish resets the
// counters, so consecutive launches on one plan need no memset.
// WALK picks how a workgroup moves past drained queues: 0 claims queue by
// queue, 1 looks at every queue at once (pop_look below), 2 takes 1 when the
// launch's look_all is set and 0 otherwise.
template <int WALK, class Body>
__device__ __forceinline__ void vbg_persistent_loop(const vbg_args& a, Body&& body) {
  __shared__ int s_item;
  const int nq = a.nq;
  int q = blockIdx.x % nq;
  int tried = 0;
  // thread 0: the next item from queue q onwards (-1 once every queue is empty)
  auto pop = [&]() -> int {
    while (tried < nq) {
      const int lo = a.q_begin[q], hi = a.q_begin[q + 1];
      if (lo < hi) {
        const unsigned h = atomicAdd(&a.counters[q], 1u);
        if ((int)h < hi - lo) return lo + (int)h;
      }
      q = (q + 1 == nq) ? 0 : q + 1;
      ++tried;
    }
    return -1;
  };
  // wave 0, launches of few items per workgroup: claim from the own queue
  // while it has items, then read every queue's head at once (a lane per
  // queue) and claim from the first queue at or after q that still has
  // items, so the last pop of a workgroup costs one read instead of a
  // claim per queue.  Heads only grow during a launch, so a queue read as
  // drained is drained.  Every lane follows the same path and values.
  auto pop_look = [&]() -> int {
    const int lane = threadIdx.x;
    for (;;) {
      if (tried == 0) {
        int v = -1;
        if (lane == 0) {
          const int lo = a.q_begin[q], hi = a.q_begin[q + 1];
          if (lo < hi) {
            const unsigned h = atomicAdd(&a.counters[q], 1u);
            if ((int)h < hi - lo) v = lo + (int)h;
          }
        }
        v = __shfl(v, 0);
        if (v >= 0) return v;
        tried = 1;
      }
      int left = 0;
      if (lane < nq) {
        const int cnt = a.q_begin[lane + 1] - a.q_begin[lane];
        const unsigned h =
            __hip_atomic_load(&a.counters[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        left = (int)h < cnt;
      }
      unsigned long long m = __ballot(left);
      if (m == 0) return -1;
      // rotate so that bit 0 is queue q, then take the lowest set bit (only
      // bits 0 .. nq-1 are set, nq <= 64)
      if (q > 0) m = ((m >> q) | (m << (nq - q))) & (nq == 64 ? ~0ull : (1ull << nq) - 1);
      const int qq = q + __builtin_ctzll(m);
      q = qq >= nq ? qq - nq : qq;
      int v = -1;
      if (lane == 0) {
        const int lo = a.q_begin[q], hi = a.q_begin[q + 1];
        const unsigned h = atomicAdd(&a.counters[q], 1u);
        if ((int)h < hi - lo) v = lo + (int)h;
      }
      v = __shfl(v, 0);
      if (v >= 0) return v;
    }
  };
  // the item index is wave-uniform; saying so (readfirstlane) lets the
  // compiler keep the item, its descriptor, pointers and k offsets in scalar
  // registers.  (Claiming the next item while the current one runs measured
  // no better: FP32 -1.5% on the batching workload, the split precisions
  // within +-1%.)
  for (;;) {
    if (WALK == 1 || (WALK == 2 && a.look_all)) {
      if (threadIdx.x < 64) {
        const int v = pop_look();
        if (threadIdx.x == 0) s_item = v;
      }
    } else if (threadIdx.x == 0) {
      s_item = pop();
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it < 0) break;
    body(vbg_uitem(vbg_cload(a.items, it)));
  }
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned d = atomicAdd(&a.counters[nq], 1u);
    if (d == gridDim.x - 1) {
      __threadfence();
      for (int i = 0; i < nq; ++i) atomicExch(&a.counters[i], 0u);
      atomicExch(&a.counters[nq], 0u);
    }
  }
}

// k range of an item's slice: slices cover [0, K) in whole multiples of
// VBG_KSLICE_ALIGN (a multiple of every kernel's BK, so the slice boundaries
// do not depend on which tile shape computes a part of the matrix), the last
// one taking the remainder.
struct KRange {
  int kb, ke;
};
template <int BK>
__device__ __forceinline__ KRange vbg_k_range(int K, int ksplit, int slice) {
  static_assert(VBG_KSLICE_ALIGN % BK == 0, "k slices must be whole k-tiles");
  if (ksplit <= 1) return {0, K};
  const int chunks = (K + VBG_KSLICE_ALIGN - 1) / VBG_KSLICE_ALIGN;
  const int per = (chunks + ksplit - 1) / ksplit;
  const int kb = min(K, slice * per * VBG_KSLICE_ALIGN);
  return {kb, min(K, kb + per * VBG_KSLICE_ALIGN)};
}

// Where an item's tile goes: the matrix's C with alpha*scale and beta, or --
// for one K slice of a split matrix -- its FP32 partial panel in the workspace
// (column-major, ld = M; slice s at offset s*M*N), scaled by `scale` only; the
// split-K reduction applies alpha and beta.
struct OutTile {
  float* p;
  int ld;
  float alpha, beta;
  bool partial;
};
__device__ __forceinline__ OutTile vbg_out(const vbg_args& a, const vbg_mat_desc& d,
                                           const vbg_item& it, float scale) {
  if (d.ksplit > 1)
    return {a.ws + vbg_u64(vbg_cload(a.ws_off, it.mat)) + (size_t)it.kslice * d.M * d.N, d.M, scale, 0.0f, true};
  return {vbg_uptr(vbg_cload(a.C, it.mat)), d.ldc, a.alpha * scale, a.beta, false};
}

// Epilogue for a 4-row column segment: C[row..row+3, col] = alpha*v + beta*C.
// Column-major C: the 4 rows are contiguous.
__device__ __forceinline__ void vbg_store4(float* C, int ldc, int row, int col, int M, int N, v4f v,
                                           float alpha, float beta) {
  if (col >= N || row >= M) return;
  float* c = C + (size_t)col * ldc + row;
  if (row + 3 < M) {
    v4f o = v * alpha;
    if (beta != 0.0f) {
      const v4f old = *(const f4u*)c;
      o += beta * old;
    }
    *(f4u*)c = o;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (row + e < M) {
        float o = alpha * v[e];
        if (beta != 0.0f) o += beta * c[e];
        c[e] = o;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Tile epilogue over NSEG 4-row column segments of the output (segment s:
// rows row(s)..row(s)+3 of column col(s), values v(s)): C = alpha*v + beta*C.
// Buffer accesses over the output's extent (ld * N floats); segments are
// handled in chunks of 8 whose C loads (beta != 0) are all issued before any
// store, so a tile costs two memory round trips instead of one per segment.
// A segment crossing row M is written element by element.
// C segments are handled in chunks of up to 8 (32 registers of old values).
template <int NSEG>
struct EpiChunk {
  static constexpr int CH = NSEG < 8 ? NSEG : 8;
  static_assert(NSEG % CH == 0, "segments must split into whole chunks");
};

// Loads the old C values of segments [c0, c0 + CH) (zero where beta == 0 or
// outside the matrix): 128-bit buffer loads for whole segments, per-element
// loads at the M edge.
template <int NSEG, class Seg>
__device__ __forceinline__ void vbg_c_load(const OutTile& o, int M, int N, Seg&& seg, int c0,
                                           v4f (&old)[EpiChunk<NSEG>::CH]) {
  constexpr int CH = EpiChunk<NSEG>::CH;
  const __amdgpu_buffer_rsrc_t rs = vbg_rsrc(o.p, (uint32_t)o.ld * (uint32_t)N * 4u);
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    int row, col;
    v4f v;
    seg(c0 + j, row, col, v);
    old[j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    if (o.beta != 0.0f && col < N && row < M) {
      if (row + 3 < M) {
        old[j] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, (col * o.ld + row) * 4, 0, 0));
      } else {
        const float* q = o.p + (size_t)col * o.ld + row;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (row + e < M) old[j][e] = q[e];
      }
    }
  }
}

// C <- alpha * acc + beta * C over the tile's segments.  All C loads of a
// chunk are issued before any of its stores, and the next chunk's loads are
// issued before this chunk's products and stores, so their round trip
// overlaps them.  With PRE, `pre` holds the first chunk's old values, loaded
// earlier (vbg_c_load) to overlap that round trip with the last k-tile.
template <int NSEG, bool PRE, class Seg>
__device__ __forceinline__ void vbg_epilogue_impl(const OutTile& o, int M, int N, Seg&& seg,
                                                  const v4f (&pre)[EpiChunk<NSEG>::CH]) {
  constexpr int CH = EpiChunk<NSEG>::CH, NC = NSEG / CH;
  const __amdgpu_buffer_rsrc_t rs = vbg_rsrc(o.p, (uint32_t)o.ld * (uint32_t)N * 4u);
  v4f old[2][CH];  // chunk c in old[c % 2] (indices fixed once the loop is unrolled)
  if (PRE) {
#pragma unroll
    for (int j = 0; j < CH; ++j) old[0][j] = pre[j];
  } else {
    vbg_c_load<NSEG>(o, M, N, seg, 0, old[0]);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (VBG_EPI_PIPE && c + 1 < NC) vbg_c_load<NSEG>(o, M, N, seg, (c + 1) * CH, old[(c + 1) % 2]);
    if (!VBG_EPI_PIPE && c > 0) vbg_c_load<NSEG>(o, M, N, seg, c * CH, old[c % 2]);
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      int row, col;
      v4f v;
      seg(c * CH + j, row, col, v);
      if (!(col < N && row < M)) continue;
      v4f r = v * o.alpha;
      if (o.beta != 0.0f) r += o.beta * old[c % 2][j];
      if (row + 3 < M) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32_t, r), rs, (col * o.ld + row) * 4, 0, 0);
      } else {
        // Edge segments go through plain global accesses: per-element dword
        // buffer stores of a vector's lanes were compiled to store lane 0 to
        // every offset (hipcc for gfx950), and edges are rare enough that the
        // flat path costs nothing measurable.
        float* q = o.p + (size_t)col * o.ld + row;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (row + e < M) q[e] = r[e];
      }
    }
  }
}

template <int NSEG, class Seg>
__device__ __forceinline__ void vbg_epilogue(const OutTile& o, int M, int N, Seg&& seg,
                                             const v4f (&pre)[EpiChunk<NSEG>::CH]) {
  vbg_epilogue_impl<NSEG, true>(o, M, N, seg, pre);
}

template <int NSEG, class Seg>
__device__ __forceinline__ void vbg_epilogue(const OutTile& o, int M, int N, Seg&& seg) {
  v4f unused[EpiChunk<NSEG>::CH];
  vbg_epilogue_impl<NSEG, false>(o, M, N, seg, unused);
}
