// vbg_bf16.hip -- BF16-operand tiles (v_mfma_f32_32x32x16_bf16, FP32
// accumulate) for the grouped forward / backward passes.  Operands are BF16 in
// HBM and are staged through LDS unconverted; C is written as FP32 or BF16.
//
// Loads are 16-byte vectors (8 BF16), so every leading dimension must be a
// multiple of 8 and every operand pointer 16-byte aligned (the plan checks the
// leading dimensions; the grouped entry points check the pointers).
#include <utility>

#include "vbg_dma.h"
#include "vbg_launch.h"
#include "vbg_tiles.h"
#include "../../include/vbgemm.h"

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned u32;
typedef u32 v4u __attribute__((ext_vector_type(4)));
typedef unsigned short u16;


// LDS row layout shared by loader and tile: rows of BK BF16 (BK*2 bytes) in
// 16-byte units; unit u of row r sits at unit u ^ ((r / 2) % 8): rows are 128
// bytes, so rows 2j and 2j+1 share a 256-byte bank line, and any 16
// consecutive rows read at one logical unit land in 16 distinct 16-byte bank
// slots.  No row padding.
template <int PITCH>
__device__ __forceinline__ int bswz(int row, int byte) {
  return row * PITCH + ((((byte >> 4) ^ (row >> 1)) & 7) << 4) + (byte & 15);
}

// BF16 operand loader.  LDS plane layout [mn][k], k contiguous, PITCH bytes.
// KC (k contiguous in HBM): a lane loads 8 consecutive k of one row.
// MC (mn contiguous in HBM): a lane owns a block of 4 consecutive mn x KB
// consecutive k (KB = 8 for 128-row tiles, 4 for 64-row ones): KB 8-byte
// loads (4 mn each, one per k), a 16-bit transpose in registers (one byte
// permute per output word), then 4 LDS writes of KB k each (16 or 8 bytes).
// Lanes 0..31 of a wave cover 128 consecutive mn at one k: 256-byte loads.
//   Every k-tile issues the same vector loads (no per-lane or per-tile
// branches around them), so the k-loop's loads in flight are the same on
// every path and the compiler can wait for exactly the oldest tile: a lane's
// byte offset is fixed once per work item and the k-tile position is a
// wave-uniform scalar offset.  Rows past MN are clamped (to MN - 1 for KC,
// to a 4-aligned row for MC); they only feed outputs that are never stored.
// The scalar offset is not range-checked by the buffer unit, so a lane whose
// k lies past the end of the item's k range has its vector offset pulled back
// to the last k inside it (KC: to the last 8-aligned vector that holds a k
// below the end; MC: to k = end - 1); those values are zeroed by selects when
// the tile that straddles the end is stored to LDS.  The pull-back is a
// no-op for every whole tile, so all tiles still issue the same loads.  A KC
// vector that straddles K may run past the row's K: rows start 16-byte
// aligned (ld % 8 == 0, 16-byte-aligned base), so that vector never leaves
// the 16-byte block of its first element.
template <int ROWS, int BK, int PITCH, bool KC>
struct Bf16Loader {
  static constexpr int KB = ROWS * BK / (4 * VBG_THREADS);  // MC: k per lane
  static_assert(KC || KB == 4 || KB == 8, "MC block of 4 mn x 4 or 8 k per lane");
  static constexpr int NV = KC ? ROWS * BK / 8 / VBG_THREADS : 1;
  static_assert(!KC || (NV >= 1 && ROWS * BK % (8 * VBG_THREADS) == 0), "tile too small for the block");
  typedef unsigned v2u_t __attribute__((ext_vector_type(2)));
  struct Regs {
    v4u kc[KC ? NV : 1];
    v2u_t mc[KC ? 1 : KB];
    int k0, kend;  // the k-tile these registers hold; k >= kend reads as zero
  };
  __amdgpu_buffer_rsrc_t rsrc;
  int voff[NV];
  int ld2;

  // MC lane map: each 8-lane group takes 4 consecutive mn-blocks (t = lane %
  // 4) at 2 consecutive k-groups (lane / 4 % 2).  Its 16-byte LDS writes of
  // one mn offset then fall in 8 distinct swizzled units (unit = kg ^ (row /
  // 2) over rows 4*blk + q), so they use 8 distinct bank groups; the 32 lanes
  // of one k-group still cover 4 * ROWS / 16 * 4 consecutive mn (256 or 128
  // contiguous bytes per load instruction).
  __device__ __forceinline__ static void coords(int v, int& row, int& k) {
    const int idx = threadIdx.x + VBG_THREADS * v;
    if (KC) {
      row = idx / (BK / 8);
      k = 8 * (idx % (BK / 8));
    } else {
      constexpr int BPW = ROWS / 16;  // 4-block groups along mn
      const int l = idx & 63, w = idx >> 6, g = l >> 3, s = l & 7;
      row = 4 * (4 * (g % BPW) + (s & 3));
      k = KB * (2 * (w * (8 / BPW) + g / BPW) + (s >> 2));
    }
  }

  __device__ __forceinline__ void init(const u16* p, int ld, int mn0, int MN, int K) {
    // the operand's span: KC rows x K (vectors of 8 k), MC K columns of MN
    // (vectors of 4 mn) at ld
    const uint32_t extent =
        (MN <= 0 || K <= 0) ? 0u
        : KC ? ((uint32_t)(MN - 1) * (uint32_t)ld + (uint32_t)((K + 7) & ~7)) * 2u
             : ((uint32_t)(K - 1) * (uint32_t)ld + (uint32_t)((MN + 3) & ~3)) * 2u;
    rsrc = vbg_rsrc(p, extent);
    ld2 = ld * 2;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int row, kk;
      coords(v, row, kk);
      if (KC) {
        voff[v] = (min(mn0 + row, MN - 1) * ld + kk) * 2;
      } else {
        voff[v] = (min(mn0 + row, (MN - 1) & ~3) + kk * ld) * 2;
      }
    }
  }

  __device__ __forceinline__ v4u ld128(int off, int soff) const {
    return __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, soff, 0));
  }
  __device__ __forceinline__ v2u_t ld64(int off, int soff) const {
    return __builtin_bit_cast(v2u_t, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, soff, 0));
  }

  __device__ __forceinline__ void fetch(Regs& r, int k0, int kend) const {
    r.k0 = k0;
    r.kend = kend;
    if constexpr (KC) {
      const int klast = (kend - 1) & ~7;  // last vector start with a k below kend
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        int row, kk;
        coords(v, row, kk);
        const int back = max(0, k0 + kk - klast);  // 0 unless the vector lies past kend
        r.kc[v] = ld128(voff[v] - 2 * back, k0 * 2);
      }
    } else {
      int row, kk;
      coords(0, row, kk);
      // k-row of load j relative to k0: kk + j, clamped to the last one below kend
      const int kmax = kend - 1 - k0;
#pragma unroll
      for (int j = 0; j < KB; ++j) r.mc[j] = ld64(voff[0] + (min(kk + j, kmax) - kk) * ld2, k0 * ld2);
    }
  }

  __device__ __forceinline__ void store(Regs& r, char* plane) const {
    const int K = r.kend;
    const bool tail = r.k0 + BK > K;  // wave-uniform: the tile straddles the k end
    if constexpr (KC) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        int row, kk;
        coords(v, row, kk);
        v4u x = r.kc[v];
        if (tail) {
          const int lim = K - (r.k0 + kk);  // elements of this vector below K
#pragma unroll
          for (int e = 0; e < 4; ++e)
            x[e] &= (2 * e < lim ? 0xffffu : 0u) | (2 * e + 1 < lim ? 0xffff0000u : 0u);
        }
        *(v4u*)(plane + bswz<PITCH>(row, 2 * kk)) = x;
      }
    } else {
      int row, kk;
      coords(0, row, kk);
      if (tail) {
#pragma unroll
        for (int j = 0; j < KB; ++j)
          if (r.k0 + kk + j >= K) r.mc[j] = v2u_t{0u, 0u};
      }
      // element q of the 4-mn vector at k + j is 16-bit half q % 2 of word q / 2
      auto pair = [&](int j, int q) -> u32 {
        const u32 a = r.mc[j][q >> 1], b = r.mc[j + 1][q >> 1];
        return (q & 1) ? __builtin_amdgcn_perm(b, a, 0x07060302u) : __builtin_amdgcn_perm(b, a, 0x05040100u);
      };
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        char* dst = plane + bswz<PITCH>(row + q, 2 * kk);
        if constexpr (KB == 8) {
          *(v4u*)dst = v4u{pair(0, q), pair(2, q), pair(4, q), pair(6, q)};
        } else {
          *(v2u_t*)dst = v2u_t{pair(0, q), pair(2, q)};
        }
      }
    }
  }
};

__device__ __forceinline__ float bf2f(u16 x) { return __uint_as_float((u32)x << 16); }

// 4-row column segment of a column-major BF16 C: round alpha*v + beta*C to BF16.
typedef u16 u16x4 __attribute__((ext_vector_type(4), aligned(2)));

__device__ __forceinline__ u16 f2bf(float x) {
  const __bf16 h = (__bf16)x;
  return *(const u16*)&h;
}

__device__ __forceinline__ void store4_bf16(u16* C, int ldc, int row, int col, int M, int N, v4f v,
                                            float alpha, float beta) {
  if (col >= N || row >= M) return;
  u16* c = C + (size_t)col * ldc + row;
  if (row + 3 < M) {  // 4 consecutive rows of a column: one 8-byte access
    v4f o = v * alpha;
    if (beta != 0.0f) {
      const u16x4 old = *(const u16x4*)c;
      o += beta * v4f{bf2f(old[0]), bf2f(old[1]), bf2f(old[2]), bf2f(old[3])};
    }
    *(u16x4*)c = u16x4{f2bf(o[0]), f2bf(o[1]), f2bf(o[2]), f2bf(o[3])};
    return;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (row + e < M) {
      float o = alpha * v[e];
      if (beta != 0.0f) o += beta * bf2f(c[e]);
      const __bf16 h = (__bf16)o;
      c[e] = *(const u16*)&h;
    }
  }
}

// Register stages of the k-loop: 2 = tile kt + 2 loads while tile kt + 1 is
// stored between the MFMA steps of tile kt; 1 = tile kt + 1 loads during tile
// kt and is stored after it.  Two stages only pay once every path has the same
// loads in flight (branch-free fetch, unconditional clamped loads): before
// that the compiler drained all loads (vmcnt(0)) ahead of each store and the
// second stage did nothing.  Measured on the grouped config (G = 128, K =
// 2048, N = 1408), two runs each: 2 stages fwd 0.469 / 0.473 ms, bwd 1.014 /
// 1.012; 1 stage fwd 0.478 / 0.478, bwd 1.027 / 1.026.
#ifndef VBG_BF16_STAGES
#define VBG_BF16_STAGES 2
#endif

template <int BM, int BN, bool TA, bool TB, bool OUT_BF16>
struct Bf16Tile {
  static constexpr int BK = 64;
  static constexpr int PITCH = BK * 2;  // 128 bytes: 8 swizzled 16-byte units
  static_assert(PITCH == 128, "bswz assumes 8 units per row");
  static constexpr int PLANE_A = BM * PITCH, PLANE_B = BN * PITCH;
  static constexpr int STAGE = PLANE_A + PLANE_B;
  static constexpr int LDS_FLOATS = 2 * STAGE / 4;
  static constexpr int WTM = BM / 2, WTN = BN / 2;
  static constexpr int TM = WTM / 32, TN = WTN / 32;
  static_assert(TM >= 1 && TN >= 1, "BF16 tile needs >= 64x64");
  using LA = Bf16Loader<BM, BK, PITCH, TA>;
  using LB = Bf16Loader<BN, BK, PITCH, !TB>;

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds_f) {
    char* lds = (char*)lds_f;
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const u16* __restrict__ A = (const u16*)vbg_uptr(vbg_cload(a.A, it.mat));
    const u16* __restrict__ B = (const u16*)vbg_uptr(vbg_cload(a.B, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    const int l32 = lane & 31, h = lane >> 5;

    v16f acc[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int s = 0; s < TN; ++s)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[r][s][e] = 0.0f;

    LA la;
    LB lb;
    la.init(A, d.lda, m0, d.M, d.K);
    lb.init(B, d.ldb, n0, d.N, d.K);
    auto compute = [&](const char* As, auto&& mid) {
      const char* Bs = As + PLANE_A;
#pragma unroll
      for (int s = 0; s < BK / 16; ++s) {
        const int koff = (h * (BK / 2) + 8 * s) * 2;
        v8bf av[TM], bv[TN];
#pragma unroll
        for (int r = 0; r < TM; ++r) av[r] = *(const v8bf*)(As + bswz<PITCH>(wr * WTM + r * 32 + l32, koff));
#pragma unroll
        for (int c = 0; c < TN; ++c) bv[c] = *(const v8bf*)(Bs + bswz<PITCH>(wc * WTN + c * 32 + l32, koff));
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
          for (int c = 0; c < TN; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[r], bv[c], acc[r][c], 0, 0, 0);
        mid(s);
      }
    };
    if constexpr (VBG_BF16_STAGES >= 2) {
      // Two register stages: while the MFMAs read LDS buffer kt & 1, tile
      // kt + 1 (loaded one step earlier) goes into the other buffer between
      // the MFMA steps and the loads of tile kt + 2 are in flight.
      typename LA::Regs ra0, ra1;
      typename LB::Regs rb0, rb1;
      // The loads and stores are unconditional (the k-tile index is clamped
      // to the last tile; a surplus tile lands in the LDS buffer no wave reads
      // again), so every path has the same loads in flight and the compiler
      // can wait for tile kt + 1 alone (vmcnt(n)) instead of draining tile
      // kt + 2's loads too (vmcnt(0)), which would undo the second stage.
      auto step = [&](int kt, typename LA::Regs& nxa, typename LB::Regs& nxb, typename LA::Regs& lda_,
                      typename LB::Regs& ldb_) {
        const int cur = kt & 1;
        const int kf = min(kt + 2, nk - 1);
        la.fetch(lda_, kb + kf * BK, ke);
        lb.fetch(ldb_, kb + kf * BK, ke);
        char* other = lds + (cur ^ 1) * STAGE;
        compute(lds + cur * STAGE, [&](int s) {
          if (s == 0) la.store(nxa, other);
          if (s == BK / 32) lb.store(nxb, other + PLANE_A);
        });
        __syncthreads();
      };
      if (nk > 0) {
        la.fetch(ra0, kb, ke);
        lb.fetch(rb0, kb, ke);
        la.store(ra0, lds);
        lb.store(rb0, lds + PLANE_A);
        const int k1 = min(1, nk - 1);
        la.fetch(ra1, kb + k1 * BK, ke);
        lb.fetch(rb1, kb + k1 * BK, ke);
      }
      __syncthreads();
      // The odd last step is peeled out of the two-step loop: with it inside
      // (as a conditional second step), the exit path rejoined the loop header
      // with the first step's loads still in flight into registers the next
      // step reuses, and the compiler waited for every load (vmcnt(0)) at the
      // top of each step.  Grouped backward (G = 128, K = 2048, N = 1408):
      // 1.020 -> 1.012 ms.
      int kt = 0;
      for (; kt + 1 < nk; kt += 2) {
        step(kt, ra1, rb1, ra0, rb0);
        step(kt + 1, ra0, rb0, ra1, rb1);
      }
      if (kt < nk) step(kt, ra1, rb1, ra0, rb0);
    } else {
      typename LA::Regs ra;
      typename LB::Regs rb;
      if (nk > 0) {
        la.fetch(ra, kb, ke);
        lb.fetch(rb, kb, ke);
        la.store(ra, lds);
        lb.store(rb, lds + PLANE_A);
      }
      __syncthreads();
      for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
          const int k0 = kb + (kt + 1) * BK;
          la.fetch(ra, k0, ke);
          lb.fetch(rb, k0, ke);
        }
        compute(lds + cur * STAGE, [](int) {});
        if (more) {
          char* nxt = lds + (cur ^ 1) * STAGE;
          la.store(ra, nxt);
          lb.store(rb, nxt + PLANE_A);
        }
        __syncthreads();
      }
    }
    const OutTile o = vbg_out(a, d, it, 1.0f);
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int row = m0 + wr * WTM + r * 32 + 8 * g + 4 * h;
          const int col = n0 + wc * WTN + c * 32 + l32;
          v4f v = {acc[r][c][4 * g], acc[r][c][4 * g + 1], acc[r][c][4 * g + 2], acc[r][c][4 * g + 3]};
          if (OUT_BF16 && !o.partial)
            store4_bf16((u16*)o.p, o.ld, row, col, d.M, d.N, v, o.alpha, o.beta);
          else
            vbg_store4(o.p, o.ld, row, col, d.M, d.N, v, o.alpha, o.beta);
        }
  }
};

// Two workgroups per CU (the 64 KB LDS of the 128x128 tile allows no more):
// the register allocation is held to two waves per SIMD.
template <class Tile>
__global__ void __launch_bounds__(VBG_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
vbg_bf16_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(16))) float lds[Tile::LDS_FLOATS];
  vbg_persistent_loop<2>(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

}  // namespace

template <class Tile>
static int run_or_occ(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, vbg_bf16_kernel<Tile>, VBG_THREADS, 0) != hipSuccess)
      nb = 1;
    *occ = nb > 0 ? nb : 1;
    return 0;
  }
  hipLaunchKernelGGL(vbg_bf16_kernel<Tile>, dim3(grid), dim3(VBG_THREADS), 0, s, *a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// The LDS-DMA tiles (vbg_dma.h) by default; the register-staged Bf16Tile
// with VBG_BF16_DMA=0 (shapes up to 128x128).
#ifndef VBG_BF16_DMA
#define VBG_BF16_DMA 1
#endif
template <int BM, int BN, bool TA, bool TB, bool OB>
static int run_tile(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if constexpr (VBG_BF16_DMA || BM * BN > 128 * 128)
    return vbg_run_dmatile<DmaTile<BM, BN, TA, TB, OB ? DM_BF16_OUT : DM_BF16>>(a, grid, s, occ);
  else
    return run_or_occ<Bf16Tile<BM, BN, TA, TB, OB>>(a, grid, s, occ);
}

template <int BM, int BN, bool OB>
static int dispatch_tt(bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (!ta && !tb) return run_tile<BM, BN, false, false, OB>(a, grid, s, occ);
  if (!ta && tb) return run_tile<BM, BN, false, true, OB>(a, grid, s, occ);
  if (ta && !tb) return run_tile<BM, BN, true, false, OB>(a, grid, s, occ);
  return run_tile<BM, BN, true, true, OB>(a, grid, s, occ);
}

template <bool OB>
static int dispatch_tile(int tile, bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s,
                         int* occ) {
  switch (tile) {
    case T64: return dispatch_tt<64, 64, OB>(ta, tb, a, grid, s, occ);
    case T128x64: return dispatch_tt<128, 64, OB>(ta, tb, a, grid, s, occ);
    case T64x128: return dispatch_tt<64, 128, OB>(ta, tb, a, grid, s, occ);
    case T128: return dispatch_tt<128, 128, OB>(ta, tb, a, grid, s, occ);
    case T256x128: return dispatch_tt<256, 128, OB>(ta, tb, a, grid, s, occ);
    case T256: return dispatch_tt<256, 256, OB>(ta, tb, a, grid, s, occ);
    default: return -1;
  }
}

int vbg_launch_bf16(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream) {
  const bool ob = a->prec == VBG_PREC_BF16_OUT_BF16;
  return ob ? dispatch_tile<true>(tile, ta, tb, a, grid, (hipStream_t)stream, nullptr)
            : dispatch_tile<false>(tile, ta, tb, a, grid, (hipStream_t)stream, nullptr);
}

int vbg_occupancy_bf16(int prec, int tile, int ta, int tb) {
  int occ = 1;
  if (prec == VBG_PREC_BF16_OUT_BF16)
    dispatch_tile<true>(tile, ta, tb, nullptr, 0, nullptr, &occ);
  else
    dispatch_tile<false>(tile, ta, tb, nullptr, 0, nullptr, &occ);
  return occ;
}
