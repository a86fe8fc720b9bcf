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

#define VBG_THREADS 256

// ---------------------------------------------------------------------------
// Operand tile loader.  Stages a ROWS x BK slice of one operand -- rows are the
// operand's M (for A) or N (for B) index "mn", columns the reduction index k --
// through registers into LDS laid out [mn][k] with pitch P (k contiguous), so
// that an MFMA lane reads 4 consecutive k with one ds_read_b128.
//
// KC = true : element (mn, k) at p[mn*ld + k]  (k contiguous in HBM: A^T, B)
// KC = false: element (mn, k) at p[mn + k*ld]  (mn contiguous in HBM: A, B^T)
//
// Rows past MN and k past K read as zero.  Every address stays inside the
// operand: out-of-range rows are clamped to MN-1 and k to K-1 before the load
// and the loaded value is then replaced by zero.
template <int ROWS, int BK, int P, bool KC>
struct TileLoader {
  static constexpr int NV = ROWS * BK / (4 * VBG_THREADS);  // float4 groups per thread
  static_assert(NV >= 1 && ROWS * BK % (4 * VBG_THREADS) == 0, "tile too small for the block");
  v4f r[NV];

  __device__ __forceinline__ static void coords(int v, int& row, int& kq) {
    const int idx = threadIdx.x + VBG_THREADS * v;
    if (KC) {
      row = idx / (BK / 4);
      kq = idx % (BK / 4);
    } else {
      row = idx % ROWS;
      kq = idx / ROWS;
    }
  }

  // ktail: this k-tile may extend past K (only the last one can).
  __device__ __forceinline__ void load(const float* __restrict__ p, int ld, int mn0, int k0, int MN,
                                       int K, bool ktail) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int row, kq;
      coords(v, row, kq);
      const int mn = mn0 + row;
      const bool ok = mn < MN;
      const int mnc = ok ? mn : MN - 1;
      const int k = k0 + 4 * kq;
      if (KC) {
        const float* q = p + (size_t)mnc * ld;
        if (!ktail) {
          r[v] = *(const f4u*)(q + k);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int kk = k + e;
            const float x = q[kk < K ? kk : K - 1];
            r[v][e] = kk < K ? x : 0.0f;
          }
        }
      } else {
        const float* q = p + mnc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          int kk = k + e;
          bool okk = true;
          if (ktail) {
            okk = kk < K;
            kk = okk ? kk : K - 1;
          }
          const float x = q[(size_t)kk * ld];
          r[v][e] = okk ? x : 0.0f;
        }
      }
      if (!ok) r[v] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }

  __device__ __forceinline__ void store(float* lds) const {
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
template <class Body>
__device__ __forceinline__ void vbg_persistent_loop(const vbg_args& a, Body&& body) {
  __shared__ int s_item;
  const int nq = a.nq;
  int q = blockIdx.x % nq;
  int tried = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      while (tried < nq) {
        const int lo = a.q_begin[q], hi = a.q_begin[q + 1];
        if (lo < hi) {
          const unsigned h = atomicAdd(&a.counters[q], 1u);
          if ((int)h < hi - lo) {
            it = lo + (int)h;
            break;
          }
        }
        q = (q + 1 == nq) ? 0 : q + 1;
        ++tried;
      }
      s_item = it;
    }
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it < 0) break;
    body(a.items[it]);
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

// k range of an item's slice: slices cover [0, K) in chunks that are whole
// multiples of BK, the last one taking the remainder.
struct KRange {
  int kb, ke;
};
template <int BK>
__device__ __forceinline__ KRange vbg_k_range(int K, int ksplit, int slice) {
  if (ksplit <= 1) return {0, K};
  const int ktiles = (K + BK - 1) / BK;
  const int per = (ktiles + ksplit - 1) / ksplit;
  const int kb = min(K, slice * per * BK);
  return {kb, min(K, kb + per * BK)};
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
