// vbg_grouped.hip -- grouped GEMM forward / backward over the variable-size
// work map (C ABI: vbg_grouped_*).  No counterpart in the reference, whose
// framework is a forward FP32 batched GEMM only.
//
// Row-major BF16 storage, groups stacked along rows (group g owns rows
// [off_g, off_g + M_g) of A, C, dC, dA; B and dB are [G, K, N]).  Each pass is
// a column-major variable-size batch on the BF16 tile family:
//   forward  C_g^T  (N x M_g) = B_g^T (N x K)  * A_g^T (K x M_g)   NN
//   dA       dA_g^T (K x M_g) = B_g   (K x N)  * dC_g^T (N x M_g)  TN
//   dB       dB_g^T (N x K)   = dC_g^T (N x M_g) * A_g  (M_g x K)  NT
// (a row-major R x C array is the column-major C x R array with ld = C).  In
// dB the ragged M_g is the reduction length; M_g = 0 writes a zero dB_g.
// Per-group pointers are derived on the device from the call's base pointers
// by a small kernel, so a call allocates nothing and issues no memset.
// Groups of at most kSmallRows rows take all three passes from direct
// streaming kernels instead of GEMM items, which would pad M_g up to a 64-wide
// tile side (forward, dA) or a 64-deep k-tile (dB) in every output tile.  For
// them each pass costs one sweep over a K x N matrix -- reading B_g (forward,
// dA) or writing dB_g -- so the kernels are laid out for contiguous HBM rows.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vbgemm.h"

namespace {

enum { PASS_FWD = 0, PASS_DA = 1, PASS_DB = 2 };

constexpr int kSmallRows = 16;

// Entry j of the pass's pointer arrays serves group gidx[j] (or j itself).
__global__ void __launch_bounds__(256)
    vbg_grouped_ptrs(int pass, int n, const int* gidx, int K, int N, const int64_t* rowoff,
                     const char* a, const char* b, const char* c, char* x, int c_es, int x_es,
                     const void** pa, const void** pb, void** pc) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int64_t g = gidx ? gidx[j] : j;
  const int64_t off = rowoff[g];
  const int64_t kn = (int64_t)K * N;
  if (pass == PASS_FWD) {  // a = A, b = B, x = C
    pa[j] = b + 2 * g * kn;
    pb[j] = a + 2 * off * K;
    pc[j] = x + (int64_t)x_es * off * N;
  } else if (pass == PASS_DA) {  // b = B, c = dC, x = dA
    pa[j] = b + 2 * g * kn;
    pb[j] = c + (int64_t)c_es * off * N;
    pc[j] = x + (int64_t)x_es * off * K;
  } else {  // c = dC, a = A, x = dB
    pa[j] = c + (int64_t)c_es * off * N;
    pb[j] = a + 2 * off * K;
    pc[j] = x + (int64_t)x_es * g * kn;
  }
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __builtin_bit_cast(float, (unsigned)h << 16);
}

typedef unsigned short u16;

// Row-count buckets of the small groups: each small-group kernel is built for
// NR rows (1, 2, 4, 8, 16) and run over the groups whose M_g fits, rows past
// M_g reading as zero, so its inner loops are straight-line code.
constexpr int kBuckets = 5;
__host__ __device__ constexpr int bucket_rows(int b) { return 1 << b; }

__device__ __forceinline__ void bf16x8_to_f32(const uint4 raw, float (&f)[8]) {
  const unsigned w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f[2 * e] = __builtin_bit_cast(float, w[e] << 16);
    f[2 * e + 1] = __builtin_bit_cast(float, w[e] & 0xffff0000u);
  }
}

__device__ __forceinline__ uint4 f32x8_to_bf16(const float (&f)[8]) {
  unsigned w[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const __bf16 lo = (__bf16)f[2 * e], hi = (__bf16)f[2 * e + 1];
    w[e] = (unsigned)__builtin_bit_cast(unsigned short, lo) |
           ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
  }
  return uint4{w[0], w[1], w[2], w[3]};
}

// Forward of the small groups: C[off_g + r][c] = sum_k A[off_g + r][k] *
// B_g[k][c], r < M_g <= NR.  The cost is streaming B_g (K x N) once.  Block
// (blockIdx.x, group blockIdx.y): 256 consecutive columns over all of K; lane
// t takes columns 8 (t % 32) .. + 7 at k = t / 32 (mod 8), so each half-wave
// reads 512 contiguous bytes of one B_g row.  A's rows are staged as FP32 in
// LDS per chunk of KC k (one batch of 16-byte loads per thread); the 8
// k-phases are summed through LDS at the end.
//   The B_g stream is pipelined: batch b+1 (UNR rows per lane) is loaded
// before batch b is consumed, with the loads unconditional (clamped to the
// chunk's last row) so every path has the same loads in flight.  On the
// 1-row bucket of G = 128, K = 2048, N = 1408 this measured the same as one
// batch at a time (129 vs 130 us, about 5.5 TB/s): that stream runs at the
// HBM rate either way.
template <int NR, bool OUT_BF16>
__global__ void __launch_bounds__(256)
    vbg_grouped_fwd_small(const int* groups, int n_groups, const int* msize, const int64_t* rowoff,
                          int K, int N, const u16* A, const u16* B, void* C) {
  constexpr int KC = NR >= 16 ? 512 : 1024;  // k per staged chunk of A (at most 32 KB)
  constexpr int UPR = KC / 8;                // 16-byte units per staged row
  constexpr int NST = (NR * UPR + 255) / 256;
  // rows of B_g per lane per batch: fewer for wide buckets (registers)
  constexpr int UNR = NR >= 16 ? 2 : NR >= 8 ? 4 : 8;
  __shared__ float As[KC][NR];
  __shared__ float red[8][256];
  const int t = threadIdx.x, cg = t & 31, kp = t >> 5;
  const int c = blockIdx.x * 256 + 8 * cg;
  for (int j = blockIdx.y; j < n_groups; j += gridDim.y) {
    const int g = groups[j];
    const int mg = msize[g];
    const int64_t off = rowoff[g];
    const u16* Bg = B + (int64_t)g * K * N;
    float acc[NR][8];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[r][e] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += KC) {
      __syncthreads();  // the previous chunk's reads of As are done
      {
        uint4 raw[NST];
        unsigned keep[NST];
#pragma unroll
        for (int i = 0; i < NST; ++i) {
          // unconditional load at a clamped address (a group here has M_g >= 1),
          // masked in the second loop, so the batch's loads are all in flight
          // at once (a conditional overwrite of the loaded registers would wait
          // for each load before issuing the next)
          const int u = t + 256 * i, uc = min(u, NR * UPR - 1), r = uc / UPR, k = k0 + 8 * (uc % UPR);
          raw[i] = *(const uint4*)(A + (off + min(r, mg - 1)) * K + min(k, K - 8));
          keep[i] = (u < NR * UPR && r < mg && k < K) ? 0xffffffffu : 0u;
        }
#pragma unroll
        for (int i = 0; i < NST; ++i) {
          const int u = t + 256 * i, r = u / UPR, kk = 8 * (u % UPR);
          if (u < NR * UPR) {
            float v[8];
            const uint4 x = raw[i];
            bf16x8_to_f32(uint4{x.x & keep[i], x.y & keep[i], x.z & keep[i], x.w & keep[i]}, v);
#pragma unroll
            for (int e = 0; e < 8; ++e) As[kk + e][r] = v[e];
          }
        }
      }
      __syncthreads();
      if (c < N) {
        // this lane's rows of the chunk: kk = kp + 8 q, q < nq (K % 8 == 0)
        const int nq = min(KC, K - k0) / 8;
        const int nb = (nq + UNR - 1) / UNR;
        const u16* bp = Bg + (int64_t)(k0 + kp) * N + c;
        const int64_t step = 8 * (int64_t)N;
        uint4 b0[UNR], b1[UNR];
        auto load = [&](uint4 (&dst)[UNR], int b) {
#pragma unroll
          for (int i = 0; i < UNR; ++i) dst[i] = *(const uint4*)(bp + min(b * UNR + i, nq - 1) * step);
        };
        auto use = [&](const uint4 (&src)[UNR], int b) {
#pragma unroll
          for (int i = 0; i < UNR; ++i) {
            const int q = b * UNR + i;
            if (q < nq) {  // uniform over the block
              float bv[8];
              bf16x8_to_f32(src[i], bv);
#pragma unroll
              for (int r = 0; r < NR; ++r) {
                const float a = As[kp + 8 * q][r];
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[r][e] = fmaf(a, bv[e], acc[r][e]);
              }
            }
          }
        };
        // pairs of batches, the odd last one peeled: no path reaches the loop
        // header with loads in flight into registers the next pair reloads
        load(b0, 0);
        int b = 0;
        for (; b + 1 < nb; b += 2) {
          load(b1, b + 1);
          use(b0, b);
          load(b0, min(b + 2, nb - 1));
          use(b1, b + 1);
        }
        if (b < nb) use(b0, b);
      }
    }
    // sum the 8 k-phases: red[kp][column] per row, then thread t owns column t
    // (a counted loop over all NR rows, so acc is indexed statically; rows
    // past M_g are not stored)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < 8; ++e) red[kp][8 * cg + e] = acc[r][e];
      __syncthreads();
      const int col = blockIdx.x * 256 + t;
      if (col < N && r < mg) {
        float v = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) v += red[q][t];
        const int64_t o = (off + r) * N + col;
        if (OUT_BF16) {
          const __bf16 h = (__bf16)v;
          ((u16*)C)[o] = __builtin_bit_cast(unsigned short, h);
        } else {
          ((float*)C)[o] = v;
        }
      }
    }
  }
}

// dA of the small groups: dA[off_g + r][k] = sum_c dC[off_g + r][c] *
// B_g[k][c], r < M_g <= NR.  The cost is streaming B_g once.  Block
// (blockIdx.x, group blockIdx.y): 8 * RPL rows of B_g; a half-wave takes one
// row at a time, 8 consecutive columns per lane (512 contiguous bytes per row
// per step), each lane keeping the partial sums of its RPL rows.  dC's rows
// are staged as FP32 in dynamic LDS, cw columns at a time (all of N when it
// fits in 48 KB, so the block's loads run without barriers between them); the
// 32 lanes of a row are summed by xor-shuffles.
template <int NR>
struct DaSmall {
  static constexpr int RPL = 4;                // rows per lane (8 RPL rows per block)
  static constexpr int ROWS = 8 * RPL;         // rows of B_g per block
  static constexpr int LDS_BYTES = 48 * 1024;
  static int cw(int N) {  // staged columns: a multiple of 256, or all of N
    const int cap = LDS_BYTES / (4 * NR) / 256 * 256;
    return N <= cap ? N : cap;
  }
};

template <int NR, bool OUT_BF16>
__global__ void __launch_bounds__(256)
    vbg_grouped_da_small(const int* groups, int n_groups, const int* msize, const int64_t* rowoff,
                         int K, int N, int cw, const u16* B, const u16* dC, void* dA) {
  constexpr int RPL = DaSmall<NR>::RPL, ROWS = DaSmall<NR>::ROWS;
  extern __shared__ float Ds[];  // [NR][cw]
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  const int cq = 8 * (l & 31);
  int krow[RPL];
#pragma unroll
  for (int i = 0; i < RPL; ++i) krow[i] = min(blockIdx.x * ROWS + 2 * RPL * w + 2 * i + (l >> 5), K - 1);
  for (int j = blockIdx.y; j < n_groups; j += gridDim.y) {
    const int g = groups[j];
    const int mg = msize[g];
    const int64_t off = rowoff[g];
    const u16* Bg = B + (int64_t)g * K * N;
    float acc[RPL][NR];
#pragma unroll
    for (int i = 0; i < RPL; ++i)
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[i][r] = 0.0f;
    for (int c0 = 0; c0 < N; c0 += cw) {
      __syncthreads();  // the previous chunk's reads of Ds are done
      // dC rows [r][c0, c0 + cw) as FP32, 16-byte loads (N and cw are
      // multiples of 8), four per thread per batch
      const int upr = cw / 8, nu = NR * upr;
      for (int u0 = 0; u0 < nu; u0 += 4 * 256) {
        uint4 raw[4];
        unsigned keep[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          // unconditional load at a clamped address (a group here has M_g >= 1),
          // masked in the second loop, so the batch's loads are all in flight
          const int u = u0 + t + 256 * i, uc = min(u, nu - 1), r = uc / upr, c = c0 + 8 * (uc % upr);
          raw[i] = *(const uint4*)(dC + (off + min(r, mg - 1)) * N + min(c, N - 8));
          keep[i] = (u < nu && r < mg && c < N) ? 0xffffffffu : 0u;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = u0 + t + 256 * i;
          if (u < nu) {
            float v[8];
            const uint4 x = raw[i];
            bf16x8_to_f32(uint4{x.x & keep[i], x.y & keep[i], x.z & keep[i], x.w & keep[i]}, v);
            float* d = Ds + (u / upr) * cw + 8 * (u % upr);
            *(float4*)d = float4{v[0], v[1], v[2], v[3]};
            *(float4*)(d + 4) = float4{v[4], v[5], v[6], v[7]};
          }
        }
      }
      __syncthreads();
      // this lane's column steps cc = cq + 256 s, s < ns; the next step's B
      // loads are issued before this step's products (unconditional, clamped
      // to the lane's last step), so two steps of loads are in flight
      const int cn = min(cw, N - c0);
      const int ns = cq < cn ? (cn - cq + 255) / 256 : 0;
      if (ns > 0) {
        const u16* bp[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) bp[i] = Bg + (int64_t)krow[i] * N + c0 + cq;
        uint4 b0[RPL], b1[RPL];
        auto load = [&](uint4 (&dst)[RPL], int s) {
#pragma unroll
          for (int i = 0; i < RPL; ++i) dst[i] = *(const uint4*)(bp[i] + 256 * min(s, ns - 1));
        };
        auto use = [&](const uint4 (&src)[RPL], int s) {
          const int cc = cq + 256 * s;
          float d[NR][8];
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const float4 d0 = *(const float4*)&Ds[r * cw + cc], d1 = *(const float4*)&Ds[r * cw + cc + 4];
            d[r][0] = d0.x; d[r][1] = d0.y; d[r][2] = d0.z; d[r][3] = d0.w;
            d[r][4] = d1.x; d[r][5] = d1.y; d[r][6] = d1.z; d[r][7] = d1.w;
          }
#pragma unroll
          for (int i = 0; i < RPL; ++i) {
            float b[8];
            bf16x8_to_f32(src[i], b);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              float sum = acc[i][r];
#pragma unroll
              for (int e = 0; e < 8; ++e) sum = fmaf(b[e], d[r][e], sum);
              acc[i][r] = sum;
            }
          }
        };
        load(b0, 0);
        int s = 0;
        for (; s + 1 < ns; s += 2) {
          load(b1, s + 1);
          use(b0, s);
          load(b0, min(s + 2, ns - 1));
          use(b1, s + 1);
        }
        if (s < ns) use(b0, s);
      }
    }
    const int kfirst = blockIdx.x * ROWS + 2 * RPL * w + (l >> 5);
#pragma unroll
    for (int i = 0; i < RPL; ++i)
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        float v = acc[i][r];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        v += __shfl_xor(v, 16);
        const int kr = kfirst + 2 * i;  // krow[i] before the clamp
        if ((l & 31) == 0 && r < mg && kr < K) {
          const int64_t o = (off + r) * K + kr;
          if (OUT_BF16) {
            const __bf16 h = (__bf16)v;
            ((u16*)dA)[o] = __builtin_bit_cast(unsigned short, h);
          } else {
            ((float*)dA)[o] = v;
          }
        }
      }
  }
}

// dB of the small groups: dB_g[k][n] = sum_{r < M_g} A[off_g + r][k] *
// dC[off_g + r][n], M_g <= NR (M_g = 0 writes zeros).  The cost is writing
// dB_g (K x N).  Thread: 4 consecutive k x 8 consecutive n (each dC load
// serves 4 outputs rows); consecutive threads take consecutive n, so every
// row's stores are contiguous.  BF16 products are exact in FP32, so only the
// row-order additions round.
template <int NR, bool OUT_BF16>
__global__ void __launch_bounds__(256)
    vbg_grouped_db_small(const int* groups, int n_groups, const int* msize, const int64_t* rowoff,
                         int K, int N, const u16* A, const u16* dC, void* dB) {
  const int64_t n8 = N / 8, k4 = (K + 3) / 4;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= k4 * n8) return;
  const int kq = (int)(idx / n8), n0 = (int)(idx - kq * n8) * 8;
  const int kb = 4 * kq;
  for (int j = blockIdx.y; j < n_groups; j += gridDim.y) {
    const int g = groups[j];
    const int mg = msize[g];
    const int64_t off = rowoff[g];
    float acc[4][8];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[q][e] = 0.0f;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= mg) break;
      float d[8];
      bf16x8_to_f32(*(const uint4*)(dC + (off + r) * N + n0), d);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float a = kb + q < K ? bf16_to_f32(A[(off + r) * K + kb + q]) : 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[q][e] = fmaf(a, d[e], acc[q][e]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (kb + q >= K) break;
      const int64_t o = (int64_t)g * K * N + (int64_t)(kb + q) * N + n0;
      if (OUT_BF16) {
        *(uint4*)((u16*)dB + o) = f32x8_to_bf16(acc[q]);
      } else {
        float4* p = (float4*)((float*)dB + o);
        p[0] = float4{acc[q][0], acc[q][1], acc[q][2], acc[q][3]};
        p[1] = float4{acc[q][4], acc[q][5], acc[q][6], acc[q][7]};
      }
    }
  }
}

thread_local std::string g_gerr;

}  // namespace

struct vbg_grouped_plan {
  int G = 0, K = 0, N = 0;
  int device = -1;
  vbg_plan* plans[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  void* dbuf = nullptr;
  int64_t* rowoff = nullptr;
  void** ptrs[3][3] = {};  // [pass][A', B', C']
  int* msize = nullptr;    // [G] rows per group
  int* big = nullptr;      // groups of more than kSmallRows rows: the dA / dB GEMM plans
  int* fwd_gemm = nullptr; // the forward GEMM plan's groups: big ones and thin buckets' ones
  int* small = nullptr;    // groups of 1..kSmallRows rows, sorted by row bucket
  int* zero = nullptr;     // groups of 0 rows (dB_g = 0)
  int n_big = 0, n_fwd_gemm = 0, n_zero = 0;
  int bucket_off[kBuckets + 1] = {};  // small[bucket_off[b] .. bucket_off[b+1]) have M_g <= 2^b
  bool fwd_small[kBuckets] = {};      // the forward runs bucket b on vbg_grouped_fwd_small
  bool beside[3] = {};                // the pass's streaming kernels run beside its GEMM plan
  hipStream_t side = nullptr;         // their stream (forked from and joined to the caller's)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

static int gfail(int code, const std::string& m) {
  g_gerr = m;
  return code;
}

// One small-group kernel launch per row bucket that holds groups.
template <bool OB>
static void launch_small(int pass, int b, const int* groups, int n, const vbg_grouped_plan* p,
                         const void* a, const void* bb, const void* c, void* x, hipStream_t s) {
  const int K = p->K, N = p->N;
  const unsigned gy = (unsigned)std::min(n, 65535);
  const u16 *A = (const u16*)a, *B = (const u16*)bb, *dC = (const u16*)c;
#define VBG_SMALL(NR)                                                                                   \
  do {                                                                                                  \
    if (pass == PASS_FWD)                                                                               \
      hipLaunchKernelGGL((vbg_grouped_fwd_small<NR, OB>), dim3((N + 255) / 256, gy), dim3(256), 0, s,  \
                         groups, n, p->msize, p->rowoff, K, N, A, B, x);                                \
    else if (pass == PASS_DA)                                                                           \
      hipLaunchKernelGGL((vbg_grouped_da_small<NR, OB>),                                                \
                         dim3((K + DaSmall<NR>::ROWS - 1) / DaSmall<NR>::ROWS, gy), dim3(256),          \
                         (size_t)4 * NR * DaSmall<NR>::cw(N), s, groups, n, p->msize, p->rowoff, K, N,  \
                         DaSmall<NR>::cw(N), B, dC, x);                                                 \
    else                                                                                                \
      hipLaunchKernelGGL((vbg_grouped_db_small<NR, OB>),                                                \
                         dim3((unsigned)(((int64_t)(K + 3) / 4 * (N / 8) + 255) / 256), gy), dim3(256), \
                         0, s, groups, n, p->msize, p->rowoff, K, N, A, dC, x);                         \
  } while (0)
  switch (b) {
    case 0: VBG_SMALL(1); break;
    case 1: VBG_SMALL(2); break;
    case 2: VBG_SMALL(4); break;
    case 3: VBG_SMALL(8); break;
    default: VBG_SMALL(16); break;
  }
#undef VBG_SMALL
}

extern "C" {

int vbg_grouped_plan_destroy(vbg_grouped_plan* p) {
  if (!p) return VBG_OK;
  for (auto& pp : p->plans)
    for (auto* q : pp) vbg_plan_destroy(q);
  if (p->dbuf) (void)hipFree(p->dbuf);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  delete p;
  return VBG_OK;
}

int vbg_grouped_plan_create(vbg_grouped_plan** plan, int G, const int* M_host, int K, int N,
                            const vbg_opts* opts) {
  if (!plan) return VBG_ERR_INVALID_ARG;
  *plan = nullptr;
  if (G <= 0 || !M_host || K <= 0 || N <= 0) return gfail(VBG_ERR_INVALID_ARG, "bad grouped shape");
  if (K % 8 || N % 8) return gfail(VBG_ERR_NOT_SUPPORTED, "grouped GEMM needs K and N multiples of 8");
  std::vector<int64_t> off(G);
  int64_t rows = 0;
  for (int g = 0; g < G; ++g) {
    if (M_host[g] < 0) return gfail(VBG_ERR_INVALID_ARG, "negative group size");
    off[g] = rows;
    rows += M_host[g];
  }
  if (rows * (int64_t)std::max(K, N) >= ((int64_t)1 << 31))
    return gfail(VBG_ERR_NOT_SUPPORTED, "stacked operand exceeds 2^31 elements");
  auto* p = new vbg_grouped_plan();
  p->G = G;
  p->K = K;
  p->N = N;
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    delete p;
    return gfail(VBG_ERR_HIP, hipGetErrorString(e));
  }
  std::vector<int> big, zero, bigN, bigK, bigM;
  std::vector<std::vector<int>> bucket(kBuckets);
  for (int g = 0; g < G; ++g) {
    const int mg = M_host[g];
    if (mg > kSmallRows) {
      big.push_back(g);
      bigN.push_back(N);
      bigK.push_back(K);
      bigM.push_back(mg);
    } else if (mg == 0) {
      zero.push_back(g);
    } else {
      int b = 0;
      while (bucket_rows(b) < mg) ++b;
      bucket[b].push_back(g);
    }
  }
  std::vector<int> small;
  p->bucket_off[0] = 0;
  for (int b = 0; b < kBuckets; ++b) {
    small.insert(small.end(), bucket[b].begin(), bucket[b].end());
    p->bucket_off[b + 1] = (int)small.size();
  }
  // The forward's streaming kernel has one block per (group, 256 columns) and
  // nothing else in flight: a bucket with fewer blocks than kFwdMinBlocks (one
  // per CU) would be one long latency-bound launch, so its groups join the
  // forward GEMM plan instead (measured on config 4: a lone 5-row group took
  // 173 us on 6 blocks).
  const int kFwdMinBlocks = 256;
  std::vector<int> fg = big, fgN = bigN, fgK = bigK, fgM = bigM;
  for (int b = 0; b < kBuckets; ++b) {
    p->fwd_small[b] = (int64_t)bucket[b].size() * ((N + 255) / 256) >= kFwdMinBlocks;
    if (p->fwd_small[b]) continue;
    for (int g : bucket[b]) {
      fg.push_back(g);
      fgN.push_back(N);
      fgK.push_back(K);
      fgM.push_back(M_host[g]);
    }
  }
  // The GEMM plans cover the big groups (entry j = group big[j]):
  // forward: (m=N, n=M_g, k=K) NN, lda=N, ldb=K, ldc=N
  // dA:      (m=K, n=M_g, k=N) TN, lda=N, ldb=N, ldc=K
  // dB:      (m=N, n=K, k=M_g) NT, lda=N, ldb=K, ldc=N
  // Passes whose small groups run on the streaming kernels: those kernels
  // go on a side stream, beside the pass's GEMM plan, whose persistent grid
  // is then 3/4 of its default so that the streaming blocks find CUs free
  // (a GEMM workgroup of the 256x256 tile holds a whole CU).  Measured on the
  // differentiation workload (G=128, K=2048, N=1408, Zipf(1.2) rows summing
  // to 32768, 30 calls, two runs): forward 0.385 -> 0.352-0.358 ms, backward
  // 0.862 -> 0.830-0.833 ms; side stream alone 0.400 / 0.887, 3/4 grid alone
  // 0.417 / 0.939 ms, and 184 or 200 of the 256 workgroups lose to 192.
  bool small_any = false;
  for (int b = 0; b < kBuckets; ++b) small_any |= p->fwd_small[b] && !bucket[b].empty();
  p->beside[PASS_FWD] = small_any && !fg.empty();
  p->beside[PASS_DA] = !small.empty() && !big.empty();
  p->beside[PASS_DB] = (!small.empty() || !zero.empty()) && !big.empty();
  vbg_opts gopt;
  if (opts) gopt = *opts; else vbg_opts_default(&gopt);
  // one GEMM plan; beside a streaming pass (and with the grid left to the
  // planner) it is made twice, the second time with 3/4 of the first's grid
  auto gemm_plan = [&](vbg_plan** pl, bool beside, auto&& create) -> int {
    vbg_opts lo = gopt;
    int r = create(pl, &lo);
    if (r != VBG_OK || !beside || gopt.grid_wgs != 0) return r;
    const int g = vbg_plan_query(*pl, 3, nullptr, 0);
    if (g < 4) return r;
    vbg_plan_destroy(*pl);
    *pl = nullptr;
    lo.grid_wgs = g * 3 / 4;
    return create(pl, &lo);
  };
  int rc = VBG_OK;
  for (int o = 0; o < 2 && rc == VBG_OK && !fg.empty(); ++o) {
    const int prec = o ? VBG_PREC_BF16_OUT_BF16 : VBG_PREC_BF16;
    rc = gemm_plan(&p->plans[PASS_FWD][o], p->beside[PASS_FWD], [&](vbg_plan** pl, const vbg_opts* lo) {
      return vbg_plan_create(pl, (int)fg.size(), fgN.data(), fgM.data(), fgK.data(), fgN.data(), fgK.data(),
                             fgN.data(), 'N', 'N', prec, lo);
    });
  }
  if (rc == VBG_OK && !big.empty()) {
    const int nb = (int)big.size();
    for (int o = 0; o < 2 && rc == VBG_OK; ++o) {
      const int prec = o ? VBG_PREC_BF16_OUT_BF16 : VBG_PREC_BF16;
      if (rc == VBG_OK)
        rc = gemm_plan(&p->plans[PASS_DA][o], p->beside[PASS_DA], [&](vbg_plan** pl, const vbg_opts* lo) {
          return vbg_plan_create(pl, nb, bigK.data(), bigM.data(), bigN.data(), bigN.data(), bigN.data(),
                                 bigK.data(), 'T', 'N', prec, lo);
        });
      if (rc == VBG_OK)
        rc = gemm_plan(&p->plans[PASS_DB][o], p->beside[PASS_DB], [&](vbg_plan** pl, const vbg_opts* lo) {
          return vbg_plan_create(pl, nb, bigN.data(), bigK.data(), bigM.data(), bigN.data(), bigK.data(),
                                 bigN.data(), 'N', 'T', prec, lo);
        });
    }
  }
  if (rc == VBG_OK && (p->beside[PASS_FWD] || p->beside[PASS_DA] || p->beside[PASS_DB])) {
    hipError_t se = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
    if (se == hipSuccess) se = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (se == hipSuccess) se = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
    if (se != hipSuccess) {
      vbg_grouped_plan_destroy(p);
      return gfail(VBG_ERR_HIP, hipGetErrorString(se));
    }
  }
  if (rc != VBG_OK) {
    g_gerr = vbg_last_error();
    vbg_grouped_plan_destroy(p);
    return rc;
  }
  // device layout: rowoff[G] | 9 pointer arrays [G] | msize[G] | big[G] | small[G] | zero[G] | fwd_gemm[G]
  const size_t ints_off = sizeof(int64_t) * G + 9 * sizeof(void*) * (size_t)G;
  const size_t bytes = ints_off + 5 * sizeof(int) * (size_t)G;
  std::vector<char> host(bytes, 0);
  std::memcpy(host.data(), off.data(), sizeof(int64_t) * G);
  int* hints = (int*)(host.data() + ints_off);
  std::memcpy(hints, M_host, sizeof(int) * G);
  if (!big.empty()) std::memcpy(hints + G, big.data(), sizeof(int) * big.size());
  if (!small.empty()) std::memcpy(hints + 2 * G, small.data(), sizeof(int) * small.size());
  if (!zero.empty()) std::memcpy(hints + 3 * G, zero.data(), sizeof(int) * zero.size());
  if (!fg.empty()) std::memcpy(hints + 4 * G, fg.data(), sizeof(int) * fg.size());
  e = hipMalloc(&p->dbuf, bytes);
  if (e == hipSuccess) e = hipMemcpy(p->dbuf, host.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    vbg_grouped_plan_destroy(p);
    return gfail(VBG_ERR_HIP, hipGetErrorString(e));
  }
  p->rowoff = (int64_t*)p->dbuf;
  void** base = (void**)((char*)p->dbuf + sizeof(int64_t) * G);
  for (int s = 0; s < 3; ++s)
    for (int t = 0; t < 3; ++t) p->ptrs[s][t] = base + (size_t)(3 * s + t) * G;
  int* dints = (int*)((char*)p->dbuf + ints_off);
  p->msize = dints;
  p->big = dints + G;
  p->small = dints + 2 * G;
  p->zero = dints + 3 * G;
  p->fwd_gemm = dints + 4 * G;
  p->n_big = (int)big.size();
  p->n_fwd_gemm = (int)fg.size();
  p->n_zero = (int)zero.size();
  *plan = p;
  return VBG_OK;
}

static bool aligned16(const void* q) { return ((uintptr_t)q & 15) == 0; }

static int run_pass(vbg_grouped_plan* p, int pass, const void* a, const void* b, const void* c,
                    void* x, int c_es, int x_es, int out_bf16, hipStream_t s) {
  // the GEMM plans cover the big groups, and for the forward the groups of
  // thin buckets too (entry j = group gidx[j])
  const int n = pass == PASS_FWD ? p->n_fwd_gemm : p->n_big;
  const int* gidx = pass == PASS_FWD ? p->fwd_gemm : p->big;
  const hipStream_t s0 = s;
  const bool side = p->side && p->beside[pass] && n > 0;
  if (side) {
    hipError_t e = hipEventRecord(p->ev_fork, s0);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->side, p->ev_fork, 0);
    if (e != hipSuccess) return gfail(VBG_ERR_HIP, hipGetErrorString(e));
  }
  if (n > 0) {
    hipLaunchKernelGGL(vbg_grouped_ptrs, dim3((n + 255) / 256), dim3(256), 0, s, pass, n,
                       gidx, p->K, p->N, p->rowoff, (const char*)a, (const char*)b, (const char*)c,
                       (char*)x, c_es, x_es, (const void**)p->ptrs[pass][0], (const void**)p->ptrs[pass][1],
                       p->ptrs[pass][2]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gfail(VBG_ERR_HIP, hipGetErrorString(e));
    const int rc = vbg_sgemm_vbatched(p->plans[pass][out_bf16], 1.0f,
                                      (const float* const*)p->ptrs[pass][0],
                                      (const float* const*)p->ptrs[pass][1], 0.0f,
                                      (float* const*)p->ptrs[pass][2], s);
    if (rc != VBG_OK) return rc;
  }
  if (side) s = p->side;
  // the small groups, one launch per row bucket; dB also zeroes the empty groups
  for (int bk = 0; bk < kBuckets; ++bk) {
    const int nb = p->bucket_off[bk + 1] - p->bucket_off[bk];
    if (nb == 0 || (pass == PASS_FWD && !p->fwd_small[bk])) continue;
    const int* groups = p->small + p->bucket_off[bk];
    if (out_bf16) launch_small<true>(pass, bk, groups, nb, p, a, b, c, x, s);
    else launch_small<false>(pass, bk, groups, nb, p, a, b, c, x, s);
  }
  if (pass == PASS_DB && p->n_zero > 0) {
    if (out_bf16) launch_small<true>(pass, 0, p->zero, p->n_zero, p, a, b, c, x, s);
    else launch_small<false>(pass, 0, p->zero, p->n_zero, p, a, b, c, x, s);
  }
  hipError_t e = hipGetLastError();
  if (side && e == hipSuccess) {
    e = hipEventRecord(p->ev_join, p->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, p->ev_join, 0);
  }
  if (e != hipSuccess) return gfail(VBG_ERR_HIP, hipGetErrorString(e));
  return VBG_OK;
}

int vbg_grouped_gemm_fwd(vbg_grouped_plan* p, const void* A, const void* B, void* C, int c_dtype,
                         void* stream) {
  if (!p) return gfail(VBG_ERR_INVALID_ARG, "null plan");
  if (c_dtype != 0 && c_dtype != 1) return gfail(VBG_ERR_INVALID_ARG, "c_dtype must be 0 or 1");
  if (!A || !B || !C) return gfail(VBG_ERR_INVALID_ARG, "null operand");
  if (!aligned16(A) || !aligned16(B) || !aligned16(C))
    return gfail(VBG_ERR_NOT_SUPPORTED, "grouped operands must be 16-byte aligned");
  return run_pass(p, PASS_FWD, A, B, nullptr, C, 0, c_dtype ? 2 : 4, c_dtype, (hipStream_t)stream);
}

int vbg_grouped_gemm_bwd(vbg_grouped_plan* p, const void* A, const void* B, const void* dC,
                         int dc_dtype, void* dA, void* dB, int grad_dtype, void* stream) {
  if (!p) return gfail(VBG_ERR_INVALID_ARG, "null plan");
  if (dc_dtype != 1)
    return gfail(VBG_ERR_NOT_SUPPORTED, "dC must be BF16 (dc_dtype = 1): it is an MFMA operand");
  if (grad_dtype != 0 && grad_dtype != 1) return gfail(VBG_ERR_INVALID_ARG, "grad_dtype must be 0 or 1");
  if (!A || !B || !dC || (!dA && !dB)) return gfail(VBG_ERR_INVALID_ARG, "null operand");
  for (const void* q : {A, B, dC, (const void*)dA, (const void*)dB})
    if (q && !aligned16(q)) return gfail(VBG_ERR_NOT_SUPPORTED, "grouped operands must be 16-byte aligned");
  const int es = grad_dtype ? 2 : 4;
  int rc = VBG_OK;
  if (dA) rc = run_pass(p, PASS_DA, A, B, dC, dA, 2, es, grad_dtype, (hipStream_t)stream);
  if (rc == VBG_OK && dB) rc = run_pass(p, PASS_DB, A, B, dC, dB, 2, es, grad_dtype, (hipStream_t)stream);
  return rc;
}

}  // extern "C"
