// vbg_split.hip -- reduced-precision MFMA paths with FP32 operands in HBM and
// FP32 accumulate, quantised on the HBM -> LDS path (no extra pass over the
// operands except the per-matrix amax of the scaled modes):
//   bf16x3 : x = hi + lo, hi = bf16(x), lo = bf16(x - hi) (x - hi is exact);
//            C += hi_a*hi_b + hi_a*lo_b + lo_a*hi_b on v_mfma_f32_32x32x16_bf16
//   fp16   : x * 2^e (per-matrix e: amax*2^e in [2^14, 2^15)) rounded to FP16,
//            v_mfma_f32_32x32x16_f16, result scaled by 2^-(eA+eB)
//   fp8    : x * (448/amax) rounded to OCP e4m3, v_mfma_f32_32x32x16_fp8_fp8,
//            result scaled by (1/sA)*(1/sB)
// No counterpart in the reference (its kernels are FP32 only); the tiling,
// work map and persistent driver are shared with the FP32 family.
#include <type_traits>

#include "vbg_dma.h"
#include "vbg_launch.h"
#include "vbg_tiles.h"
#include "../../include/vbgemm.h"


namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

// bf16x3 split of 4 floats: hi = bf16_rne(x), lo = bf16_rne(x - hi).  hi is
// converted once, two values per v_cvt_pk_bf16_f32, and recovered as floats
// from the packed word itself (low half: shift left 16; high half: mask), so
// each pair costs 2 conversions + 2 integer ops + 2 subtractions.
__device__ __forceinline__ void bf16x3_split4(v4f x, v4bf& hi, v4bf& lo) {
  const v2bf h01 = __builtin_convertvector(v2f{x[0], x[1]}, v2bf);
  const v2bf h23 = __builtin_convertvector(v2f{x[2], x[3]}, v2bf);
  const unsigned u01 = __builtin_bit_cast(unsigned, h01), u23 = __builtin_bit_cast(unsigned, h23);
  const v2f r01 = {x[0] - __builtin_bit_cast(float, u01 << 16),
                   x[1] - __builtin_bit_cast(float, u01 & 0xffff0000u)};
  const v2f r23 = {x[2] - __builtin_bit_cast(float, u23 << 16),
                   x[3] - __builtin_bit_cast(float, u23 & 0xffff0000u)};
  hi = __builtin_shufflevector(h01, h23, 0, 1, 2, 3);
  lo = __builtin_shufflevector(__builtin_convertvector(r01, v2bf), __builtin_convertvector(r23, v2bf),
                               0, 1, 2, 3);
}
typedef unsigned v2u32_t __attribute__((ext_vector_type(2)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));

enum { MODE_BF16X3 = 0, MODE_FP16 = 1, MODE_FP8 = 2 };

template <int MODE>
struct Mode;
// STAGES: register stages of the pipeline.  2: tile kt+1 is quantised between
// the MFMA steps of tile kt while tile kt+2 loads; 1: tile kt+1 loads during
// tile kt and is quantised after it.  Measured on the unbalanced B=2048 batch:
// bf16x3 is faster with 1 (its 64 KB of LDS holds occupancy at 2 either way,
// and the second stage costs registers), FP16/FP8 with 2.  WAVES: the
// register-allocation target handed to the compiler (bf16x3's 3 keeps its
// accumulators in AGPRs, the allocation that was measured fastest).
// LDS_STAGES: LDS buffers (2: the next tile is stored while the current one
// is read; 1: one buffer and a second barrier per k-tile, half the LDS).
// bf16x3 measured on the unbalanced B=2048 batch: waves 3 / LDS 2 (occupancy
// 2, no spill) 211.6 TFLOP/s; waves 3 / LDS 1 (occupancy 3, 40 B/lane spill)
// 213.6 -- within noise, so occupancy is not what limits this kernel; BK 64
// (one LDS buffer) 206.7; two register stages (waves 2, accumulators in VGPRs)
// 207.5 against 217.0 for one stage.
#ifndef VBG_BF16X3_WAVES
#define VBG_BF16X3_WAVES 3
#endif
#ifndef VBG_BF16X3_LDS_STAGES
#define VBG_BF16X3_LDS_STAGES 2
#endif
#ifndef VBG_BF16X3_BK
#define VBG_BF16X3_BK 32
#endif
#ifndef VBG_FP16_BK
#define VBG_FP16_BK 32
#endif
#ifndef VBG_FP16_STAGES
#define VBG_FP16_STAGES 2
#endif
#ifndef VBG_BF16X3_STAGES
#define VBG_BF16X3_STAGES 1
#endif
// One-register-stage loop (bf16x3): PRIO raises the wave's issue priority
// while it issues a k-tile's MFMAs; FENCE keeps the quantisation of the next
// tile after those MFMAs (the scheduler otherwise interleaves it into them,
// which puts waits on the next tile's loads in the middle of the MFMA
// stream).  Measured on the unbalanced B=2048 batch, bf16x3, two runs each:
// both on 239.3 / 240.4 TFLOP/s, both off 236.8 / 236.7.
#ifndef VBG_SPLIT_PRIO
#define VBG_SPLIT_PRIO 1
#endif
#ifndef VBG_SPLIT_FENCE
#define VBG_SPLIT_FENCE 1
#endif
// The two-stage loop loads tile kt+2 only while one exists, so the compiler
// drains every load (vmcnt(0)) before storing tile kt+1.  Making that load
// unconditional (clamped to the last tile), which lets it wait for tile kt+1
// alone, measured slower here (in-loop FP16 / FP8 GEMM kernels 3.38 -> 3.51 /
// 3.48 ms; packed and bf16x3 two-stage variants lost too), unlike the BF16
// tiles, so the split family keeps the conditional load.
// Register stages on packed operands (no conversion between load and LDS).
// Measured on the unbalanced B=2048 batch, GEMM kernel ms with 2 / 1 stages:
// bf16x3 3.97 / 3.58, FP16 2.15 / 2.12, FP8 1.88 / 1.83.
#ifndef VBG_PK_STAGES
#define VBG_PK_STAGES 1
#endif
template <>
struct Mode<MODE_BF16X3> {
  static constexpr int EB = 2, PLANES = 2, STAGES = VBG_BF16X3_STAGES, WAVES = VBG_BF16X3_WAVES,
                       LDS_STAGES = VBG_BF16X3_LDS_STAGES, BK = VBG_BF16X3_BK;
};
template <>
struct Mode<MODE_FP16> {
  static constexpr int EB = 2, PLANES = 1, STAGES = VBG_FP16_STAGES, WAVES = 2, LDS_STAGES = 2,
                       BK = VBG_FP16_BK;
};
template <>
struct Mode<MODE_FP8> {
  static constexpr int EB = 1, PLANES = 1, STAGES = VBG_FP16_STAGES, WAVES = 2, LDS_STAGES = 2,
                       BK = VBG_FP16_BK;
};

// Loader of a packed operand: ROWS rows of Kp = K rounded up to
// VBG_PACK_KALIGN elements (k contiguous, EB bytes each, zero past K), PLANES
// planes of MN*Kp*EB bytes one after the other.  A k-tile row is BK*EB bytes
// = UPR 16-byte units; consecutive threads take consecutive units of a row.
// Rows past MN read row MN-1 (their results are never stored).
template <int ROWS, int BK, int EB, int PLANES>
struct PkLoader {
  static constexpr bool PACKED = true;
  static constexpr int UPR = BK * EB / 16;
  static constexpr int PER_PLANE = ROWS * UPR;
  static constexpr int UNITS = PER_PLANE * PLANES;
  static constexpr int NV = (UNITS + VBG_THREADS - 1) / VBG_THREADS;
  static constexpr bool FULL = UNITS % VBG_THREADS == 0;
  static_assert(UPR >= 1 && (FULL || NV == 1), "partial use of the block only for one unit per thread");
  typedef v4u32_t Regs[NV];
  __amdgpu_buffer_rsrc_t rsrc;
  int voff[NV];

  __device__ __forceinline__ static bool active(int v) {
    return FULL || (int)threadIdx.x + VBG_THREADS * v < UNITS;
  }
  __device__ __forceinline__ static void coords(int v, int& plane, int& row, int& u) {
    const int idx = threadIdx.x + VBG_THREADS * v;
    plane = idx / PER_PLANE;
    const int r = idx % PER_PLANE;
    row = r / UPR;
    u = r % UPR;
  }
  __device__ __forceinline__ void init(const char* base, int Kp, int mn0, int MN) {
    const uint32_t plane_bytes = (uint32_t)MN * (uint32_t)Kp * EB;
    rsrc = vbg_rsrc(base, plane_bytes * PLANES);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      int plane, row, u;
      coords(v, plane, row, u);
      voff[v] = (int)(plane * plane_bytes) + (min(mn0 + row, MN - 1) * Kp * EB) + 16 * u;
    }
  }
  // k-tile [k0, k0 + BK) (inside the packed rows: Kp is a multiple of BK).
  __device__ __forceinline__ void fetch(Regs& r, int k0, int) const {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (active(v)) r[v] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff[v], k0 * EB, 0);
  }
};

// LDS layout per operand plane: [mn][k] with k contiguous, rows of BK*EB bytes
// split into NU = BK/8 units of 8 k (16 B for BF16/FP16, 8 B for FP8: one MFMA
// operand read); unit u of row r sits at unit u ^ ((r / RPB) % NU), RPB =
// 256 / row bytes (at least 1), so the rows an MFMA operand read touches hit
// distinct LDS banks without row padding.  MFMA step s of a k-tile sums k in
// [16s, 16s + 16), lane half h supplying k = 16s + 8h + [0, 8): the grouping of
// k into MFMAs is the same for every k-tile depth, so tiles of different
// depths (the LDS-DMA tiles on packed operands) give the same bits.  With CPRE, the one-register-stage loop loads
// the epilogue's first chunk of C during the last k-tile, into registers the
// next tile's prefetch would otherwise hold.
//   With PK the operands come already quantised from the packed workspace
// (the operand pack pre-pass); TA / TB are then unused.
//   The 256x128 tile (bf16x3): each wave owns 128 x 64 (eight 32x32 blocks,
// 128 accumulators, kept in VGPRs at two waves per SIMD), over 16-deep
// k-tiles so that its staging registers and double-buffered LDS (48 KB) fit:
// 25% fewer operand bytes per flop than 128x128.  Measured: 1024^3 bf16x3
// batches 311 vs 283 TFLOP/s; the unbalanced B=2048 batch on one fixed shape
// 3.71 vs 3.92 ms.
template <int BM, int BN, int MODE, bool TA, bool TB, bool CPRE = true, bool PK = false>
struct SplitTile {
  static constexpr bool TALL = BM * BN > 128 * 128;
  static_assert(!TALL || MODE == MODE_BF16X3, "256x128 tiles are built for bf16x3");
  static constexpr int BK = TALL ? 16 : Mode<MODE>::BK;
  static_assert(BK == 16 || BK == 32 || BK == 64, "k-tile depth 16, 32 or 64");
  static constexpr int MODE_ID = MODE;
  static constexpr int WAVES = TALL ? 2 : Mode<MODE>::WAVES;
  static constexpr int EB = Mode<MODE>::EB, PLANES = Mode<MODE>::PLANES;
  static constexpr int PITCH = BK * EB;  // bytes per LDS row (no padding: swizzled)
  static constexpr int UNIT = 8 * EB;    // bytes per MFMA operand read
  static constexpr int NU = PITCH / UNIT;
  static constexpr int RPB = PITCH >= 256 ? 1 : 256 / PITCH;
  __device__ __forceinline__ static int swz(int row, int byte) {
    return row * PITCH + ((((byte / UNIT) ^ (row / RPB)) & (NU - 1)) * UNIT) + (byte % UNIT);
  }
  static constexpr int PLANE_A = BM * PITCH, PLANE_B = BN * PITCH;
  static constexpr int STAGE = PLANES * (PLANE_A + PLANE_B);  // bytes
  static constexpr int LDS_STAGES = Mode<MODE>::LDS_STAGES;
  static constexpr int LDS_FLOATS = LDS_STAGES * STAGE / 4;
  static constexpr int WTM = BM / 2, WTN = BN / 2;
  static constexpr int TM = WTM / 32, TN = WTN / 32;
  static_assert(TM >= 1 && TN >= 1, "split tile needs >= 64x64");
  using LA = std::conditional_t<PK, PkLoader<BM, BK, EB, PLANES>, TileLoader<BM, BK, 4, TA>>;
  using LB = std::conditional_t<PK, PkLoader<BN, BK, EB, PLANES>, TileLoader<BN, BK, 4, !TB>>;
  static constexpr int STAGES = PK ? VBG_PK_STAGES : Mode<MODE>::STAGES;
  static_assert(LDS_STAGES == 2 || STAGES == 1, "two register stages need two LDS buffers");

  // Quantise 4 consecutive k of one row and write them to the plane(s).
  template <class L>
  __device__ __forceinline__ static void store(const typename L::Regs& rr, char* plane0,
                                               int plane_bytes, float s) {
    if constexpr (PK) {
      (void)s;
#pragma unroll
      for (int v = 0; v < L::NV; ++v) {
        if (!L::active(v)) continue;
        int plane, row, u;
        L::coords(v, plane, row, u);
        char* dst = plane0 + plane * plane_bytes;
        if constexpr (UNIT == 16) {
          *(v4u32_t*)(dst + swz(row, 16 * u)) = rr[v];
        } else {  // 8-byte units (FP8): the two halves land in different units
          *(v2u32_t*)(dst + swz(row, 16 * u)) = v2u32_t{rr[v][0], rr[v][1]};
          *(v2u32_t*)(dst + swz(row, 16 * u + 8)) = v2u32_t{rr[v][2], rr[v][3]};
        }
      }
    } else {
#pragma unroll
    for (int v = 0; v < L::NV; ++v) {
      int row, kq;
      L::coords(v, row, kq);
      const v4f x = rr[v];
      char* dst = plane0 + swz(row, 4 * kq * EB);
      if (MODE == MODE_BF16X3) {
        v4bf hi, lo;
        bf16x3_split4(x, hi, lo);  // x - hi is exact
        *(v4bf*)dst = hi;
        *(v4bf*)(dst + plane_bytes) = lo;
      } else if (MODE == MODE_FP16) {
        const v4h q = {(_Float16)(x[0] * s), (_Float16)(x[1] * s), (_Float16)(x[2] * s),
                       (_Float16)(x[3] * s)};
        *(v4h*)dst = q;
      } else {
        int p = __builtin_amdgcn_cvt_pk_fp8_f32(x[0] * s, x[1] * s, 0, false);
        p = __builtin_amdgcn_cvt_pk_fp8_f32(x[2] * s, x[3] * s, p, true);
        *(int*)dst = p;
      }
    }
    }
  }

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds_f) {
    char* lds = (char*)lds_f;
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const float* __restrict__ A = PK ? nullptr : vbg_uptr(vbg_cload(a.A, it.mat));
    const float* __restrict__ B = PK ? nullptr : vbg_uptr(vbg_cload(a.B, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    float sA = 1.0f, sB = 1.0f, post = 1.0f;
    if (MODE != MODE_BF16X3) {
      sA = __builtin_bit_cast(float, vbg_u(__builtin_bit_cast(int, vbg_cload(a.scales, 4 * it.mat))));
      sB = __builtin_bit_cast(float, vbg_u(__builtin_bit_cast(int, vbg_cload(a.scales, 4 * it.mat + 1))));
      post = __builtin_bit_cast(float, vbg_u(__builtin_bit_cast(int, vbg_cload(a.scales, 4 * it.mat + 2))));
    }

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

    // Two register stages: while the MFMAs consume LDS buffer kt&1, tile kt+1
    // (loaded from HBM one iteration earlier) is quantised into the other LDS
    // buffer between the MFMA steps, and the loads of tile kt+2 are in flight.
    LA la;
    LB lb;
    if constexpr (PK) {
      const int Kp = (d.K + VBG_PACK_KALIGN - 1) & ~(VBG_PACK_KALIGN - 1);
      la.init(a.pk + vbg_cload(a.pk_off, 2 * it.mat), Kp, m0, d.M);
      lb.init(a.pk + vbg_cload(a.pk_off, 2 * it.mat + 1), Kp, n0, d.N);
    } else {
      la.init(A, d.lda, m0, d.M, d.K);
      lb.init(B, d.ldb, n0, d.N, d.K);
    }
    typename LA::Regs la0, la1;
    typename LB::Regs lb0, lb1;
    auto load = [&](typename LA::Regs& ra, typename LB::Regs& rb, int kt) {
      const int k0 = kb + kt * BK;
      la.fetch(ra, k0, ke);
      lb.fetch(rb, k0, ke);
    };
    auto compute = [&](const char* As, auto&& mid) {
      const char* Bs = As + PLANES * PLANE_A;
#pragma unroll
      for (int s = 0; s < BK / 16; ++s) {
        const int koff = (16 * s + 8 * h) * EB;  // MFMA s: k in [16s, 16s + 16), lane half h: 8 of them
        if (MODE == MODE_FP8) {
          long av[TM], bv[TN];
#pragma unroll
          for (int r = 0; r < TM; ++r) av[r] = *(const long*)(As + swz(wr * WTM + r * 32 + l32, koff));
#pragma unroll
          for (int c = 0; c < TN; ++c) bv[c] = *(const long*)(Bs + swz(wc * WTN + c * 32 + l32, koff));
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(av[r], bv[c], acc[r][c], 0, 0, 0);
        } else if (MODE == MODE_FP16) {
          v8h av[TM], bv[TN];
#pragma unroll
          for (int r = 0; r < TM; ++r) av[r] = *(const v8h*)(As + swz(wr * WTM + r * 32 + l32, koff));
#pragma unroll
          for (int c = 0; c < TN; ++c) bv[c] = *(const v8h*)(Bs + swz(wc * WTN + c * 32 + l32, koff));
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av[r], bv[c], acc[r][c], 0, 0, 0);
        } else {
          v8bf ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
          for (int r = 0; r < TM; ++r) {
            const char* p = As + swz(wr * WTM + r * 32 + l32, koff);
            ah[r] = *(const v8bf*)p;
            al[r] = *(const v8bf*)(p + PLANE_A);
          }
#pragma unroll
          for (int c = 0; c < TN; ++c) {
            const char* p = Bs + swz(wc * WTN + c * 32 + l32, koff);
            bh[c] = *(const v8bf*)p;
            bl[c] = *(const v8bf*)(p + PLANE_B);
          }
          // the three products outermost: consecutive MFMAs on one accumulator
          // are TM*TN issues apart instead of back to back
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bl[c], acc[r][c], 0, 0, 0);
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[r], bh[c], acc[r][c], 0, 0, 0);
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bh[c], acc[r][c], 0, 0, 0);
        }
        mid(s);
      }
    };
    // one k-tile: `nxt` holds tile kt+1 (to quantise now), `ldr` receives kt+2
    auto step = [&](int kt, typename LA::Regs& nxa, typename LB::Regs& nxb, typename LA::Regs& lda_,
                    typename LB::Regs& ldb_) {
      const int cur = kt & 1;
      if (kt + 2 < nk) load(lda_, ldb_, kt + 2);
      char* other = lds + (cur ^ 1) * STAGE;
      const bool more = kt + 1 < nk;
      compute(lds + cur * STAGE, [&](int s) {
        if (!more) return;
        if (s == 0) store<LA>(nxa, other, PLANE_A, sA);
        if (s == BK / 16 - 1) store<LB>(nxb, other + PLANES * PLANE_A, PLANE_B, sB);
      });
      __syncthreads();
    };
    constexpr int NSEG = TM * TN * 4;
    const OutTile o = vbg_out(a, d, it, post);
    auto seg = [&](int sg, int& row, int& col, v4f& v) {
      const int r = sg / (TN * 4), c = (sg / 4) % TN, g = sg % 4;
      row = m0 + wr * WTM + r * 32 + 8 * g + 4 * h;
      col = n0 + wc * WTN + c * 32 + l32;
      v = v4f{acc[r][c][4 * g], acc[r][c][4 * g + 1], acc[r][c][4 * g + 2], acc[r][c][4 * g + 3]};
    };
    if constexpr (STAGES >= 2) {
    if (nk > 0) {
      load(la0, lb0, 0);
      store<LA>(la0, lds, PLANE_A, sA);
      store<LB>(lb0, lds + PLANES * PLANE_A, PLANE_B, sB);
    }
    if (nk > 1) load(la1, lb1, 1);
    __syncthreads();
    // the odd last step outside the two-step loop, so no path rejoins the
    // loop header with a step's loads still in flight (which made the
    // compiler drain every load at the top of each step); in-loop FP16 / FP8
    // GEMM kernels on the unbalanced B=2048 batch 3.35 / 3.36 -> 3.34 / 3.33 ms
    int kt = 0;
    for (; kt + 1 < nk; kt += 2) {
      step(kt, la1, lb1, la0, lb0);
      step(kt + 1, la0, lb0, la1, lb1);
    }
    if (kt < nk) step(kt, la1, lb1, la0, lb0);
    vbg_epilogue<NSEG>(o, d.M, d.N, seg);
    } else {
    // one register stage: tile kt+1 is loaded while tile kt computes and is
    // quantised into the other LDS buffer after the MFMAs
    (void)la1;
    (void)lb1;
    (void)step;
    if (nk > 0) {
      load(la0, lb0, 0);
      store<LA>(la0, lds, PLANE_A, sA);
      store<LB>(lb0, lds + PLANES * PLANE_A, PLANE_B, sB);
    }
    __syncthreads();
    v4f cpre[EpiChunk<NSEG>::CH];
    auto ktile = [&](int kt, auto last, auto pre) {
      const int cur = LDS_STAGES == 2 ? (kt & 1) : 0;
      if constexpr (!decltype(last)::value) load(la0, lb0, kt + 1);
      if constexpr (decltype(pre)::value) vbg_c_load<NSEG>(o, d.M, d.N, seg, 0, cpre);
#if VBG_SPLIT_PRIO
      __builtin_amdgcn_s_setprio(1);
#endif
      compute(lds + cur * STAGE, [](int) {});
#if VBG_SPLIT_PRIO
      __builtin_amdgcn_s_setprio(0);
#endif
#if VBG_SPLIT_FENCE
      __builtin_amdgcn_sched_barrier(0);
#endif
      if (LDS_STAGES == 1) __syncthreads();  // every wave is done reading the buffer
      if constexpr (!decltype(last)::value) {
        char* other = lds + (LDS_STAGES == 2 ? (cur ^ 1) : 0) * STAGE;
        store<LA>(la0, other, PLANE_A, sA);
        store<LB>(lb0, other + PLANES * PLANE_A, PLANE_B, sB);
      }
      __syncthreads();
    };
    for (int kt = 0; kt + 1 < nk; ++kt) ktile(kt, std::false_type{}, std::false_type{});
    if constexpr (CPRE) {
      if (nk > 0) {
        ktile(nk - 1, std::true_type{}, std::true_type{});
        vbg_epilogue<NSEG>(o, d.M, d.N, seg, cpre);
      } else {
        vbg_epilogue<NSEG>(o, d.M, d.N, seg);
      }
    } else {
      if (nk > 0) ktile(nk - 1, std::true_type{}, std::false_type{});
      vbg_epilogue<NSEG>(o, d.M, d.N, seg);
    }
    }
  }
};

// Mixed launch: items of the 64x64 .. 128x128 shapes in one persistent grid,
// each naming its own (wave-uniform) shape; LDS sized for the largest.
template <int MODE, bool TA, bool TB, bool PK = false>
struct SplitMixed {
  using Sq128 = SplitTile<128, 128, MODE, TA, TB, false, PK>;
  using Tall = SplitTile<128, 64, MODE, TA, TB, false, PK>;
  using Wide = SplitTile<64, 128, MODE, TA, TB, false, PK>;
  using Sq64 = SplitTile<64, 64, MODE, TA, TB, false, PK>;
  static constexpr int MODE_ID = MODE;
  static constexpr int WAVES = 2;  // two workgroups per CU, as the 128x128 body alone
  static constexpr int LDS_FLOATS = Sq128::LDS_FLOATS;
  static_assert(Sq128::LDS_FLOATS >= Tall::LDS_FLOATS && Sq128::LDS_FLOATS >= Sq64::LDS_FLOATS,
                "the 128x128 body must need the most LDS");
  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    switch (it.tile) {
      case T128: Sq128::run(a, it, lds); break;
      case T128x64: Tall::run(a, it, lds); break;
      case T64x128: Wide::run(a, it, lds); break;
      default: Sq64::run(a, it, lds); break;
    }
  }
};

template <class Tile>
__global__ void __launch_bounds__(VBG_THREADS) __attribute__((amdgpu_waves_per_eu(Tile::WAVES)))
vbg_split_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(16))) float lds[Tile::LDS_FLOATS];
  vbg_persistent_loop<2>(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

// ---------------------------------------------------------------- operand pack
// Operand pack pre-pass: job j quantises rows [row0, row0 + 64) x k [k0, k0 +
// VBG_PACK_K) of one operand into the packed layout PkLoader reads -- rows of
// Kp = K rounded up to VBG_PACK_KALIGN, k contiguous, zeros past K -- with the
// same arithmetic as the in-loop quantisation (bf16x3 split, x * scale to FP16
// or e4m3), so a packed run gives the same bits as an unpacked one.  Operands
// stored with k contiguous are converted in place order; the others are
// transposed through LDS, 64 x 64 at a time.  Thread t owns row t / 4 and 16
// consecutive k of each 64-wide step.
template <int MODE>
__global__ void __launch_bounds__(VBG_THREADS)
    vbg_pack_kernel(const vbg_pack_job* jobs, const vbg_mat_desc* desc, const float* const* A,
                    const float* const* B, int ta, int tb, const float* scales, char* pk,
                    const int64_t* pk_off) {
  constexpr int EB = Mode<MODE>::EB;
  __shared__ float tr[64][65];
  const vbg_pack_job j = jobs[blockIdx.x];
  const vbg_mat_desc d = desc[j.mat];
  const bool op_a = j.op == 0;
  const int rows = op_a ? d.M : d.N;
  const bool kc = op_a ? ta != 0 : tb == 0;  // stored with k contiguous
  const int ld = op_a ? d.lda : d.ldb;
  const float* p = op_a ? A[j.mat] : B[j.mat];
  const float s = MODE == MODE_BF16X3 ? 1.0f : scales[4 * j.mat + j.op];
  const int Kp = (d.K + VBG_PACK_KALIGN - 1) & ~(VBG_PACK_KALIGN - 1);
  char* out = pk + pk_off[2 * j.mat + j.op];
  const int64_t plane = (int64_t)rows * Kp * EB;
  // the stored operand ends after its last element: rows x K at ld (k
  // contiguous) or K columns of `rows` at ld; past that, buffer loads read 0
  const uint32_t extent = kc ? ((uint32_t)(rows - 1) * (uint32_t)ld + (uint32_t)d.K) * 4u
                             : ((uint32_t)(d.K - 1) * (uint32_t)ld + (uint32_t)rows) * 4u;
  const __amdgpu_buffer_rsrc_t rs = vbg_rsrc(p, extent);
  const int t = threadIdx.x;
  const int orow = t >> 2, oq = (t & 3) * 16;
  const int r = j.row0 + orow;
  const int kend = min(Kp, j.k0 + VBG_PACK_K);
  for (int kb = j.k0; kb < kend; kb += 64) {
    float x[16];
    // Addresses use k clamped to the last column (k >= K is zeroed by the
    // selects), so no offset leaves the operand's 32-bit range.
    if (kc) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = kb + oq + 4 * q;
        const v4f v = __builtin_bit_cast(
            v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, (min(r, rows - 1) * ld + min(k, d.K - 1)) * 4, 0, 0));
#pragma unroll
        for (int e = 0; e < 4; ++e) x[4 * q + e] = k + e < d.K ? v[e] : 0.0f;
      }
    } else {
      __syncthreads();  // the previous step's reads of tr are done
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = (t >> 4) + 16 * i, rr = 4 * (t & 15);
        const int k = kb + kk;
        const v4f v = __builtin_bit_cast(
            v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, (min(k, d.K - 1) * ld + j.row0 + rr) * 4, 0, 0));
#pragma unroll
        for (int e = 0; e < 4; ++e) tr[kk][rr + e] = k < d.K ? v[e] : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] = tr[oq + e][orow];
    }
    if (r >= rows) continue;
    char* dst = out + (int64_t)r * Kp * EB + (int64_t)(kb + oq) * EB;
    if (MODE == MODE_BF16X3) {
      v4bf hi[4], lo[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        bf16x3_split4(v4f{x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]}, hi[q], lo[q]);
      *(v8bf*)dst = __builtin_shufflevector(hi[0], hi[1], 0, 1, 2, 3, 4, 5, 6, 7);
      *(v8bf*)(dst + 16) = __builtin_shufflevector(hi[2], hi[3], 0, 1, 2, 3, 4, 5, 6, 7);
      *(v8bf*)(dst + plane) = __builtin_shufflevector(lo[0], lo[1], 0, 1, 2, 3, 4, 5, 6, 7);
      *(v8bf*)(dst + plane + 16) = __builtin_shufflevector(lo[2], lo[3], 0, 1, 2, 3, 4, 5, 6, 7);
    } else if (MODE == MODE_FP16) {
      v8h h0, h1;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        h0[e] = (_Float16)(x[e] * s);
        h1[e] = (_Float16)(x[8 + e] * s);
      }
      *(v8h*)dst = h0;
      *(v8h*)(dst + 16) = h1;
    } else {
      v4u32_t w;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int pq = __builtin_amdgcn_cvt_pk_fp8_f32(x[4 * q] * s, x[4 * q + 1] * s, 0, false);
        pq = __builtin_amdgcn_cvt_pk_fp8_f32(x[4 * q + 2] * s, x[4 * q + 3] * s, pq, true);
        w[q] = (unsigned)pq;
      }
      *(v4u32_t*)dst = w;
    }
  }
}

// ---------------------------------------------------------------- amax / scales
// Chunk of one stored operand: elements [begin, end) of the packed matrix
// (ld == rows), or columns [begin, end) otherwise.
__device__ __forceinline__ float block_max(float m) {
  __shared__ float red[VBG_THREADS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < VBG_THREADS / 64; ++w) m = fmaxf(m, red[w]);
  return m;
}

__global__ void __launch_bounds__(VBG_THREADS) vbg_amax_zero(unsigned* amax, int n) {
  const int i = blockIdx.x * VBG_THREADS + threadIdx.x;
  if (i < n) amax[i] = 0u;
}

__global__ void __launch_bounds__(VBG_THREADS)
    vbg_amax_kernel(const vbg_chunk* ch, const vbg_mat_desc* desc, const float* const* A,
                    const float* const* B, int ta, int tb, unsigned* amax) {
  const vbg_chunk c = ch[blockIdx.x];
  const vbg_mat_desc d = desc[c.mat];
  const float* p = c.op == 0 ? A[c.mat] : B[c.mat];
  const int rows = c.op == 0 ? (ta ? d.K : d.M) : (tb ? d.N : d.K);
  const int ld = c.op == 0 ? d.lda : d.ldb;
  // Max of |x| over a contiguous run of n floats: 16-byte buffer loads, four
  // in flight per thread, over the whole float4s [0, n & ~3) (any 4-byte
  // alignment: buffer loads need no 16-byte alignment); loads that fall past
  // the run read 0 or in-run values, neither of which changes the max.  The
  // last n % 4 floats are read one by one.
  auto run_max = [](const float* q, int n, float m) {
    const __amdgpu_buffer_rsrc_t rs = vbg_rsrc(q, (uint32_t)n * 4u);
    const int n4 = n & ~3;
    for (int j = n4 + (int)threadIdx.x; j < n; j += VBG_THREADS) m = fmaxf(m, fabsf(q[j]));
    for (int i = 4 * threadIdx.x; i < n4; i += 16 * VBG_THREADS) {
      v4f x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        x[u] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, (i + u * 4 * VBG_THREADS) * 4, 0, 0));
#pragma unroll
      for (int u = 0; u < 4; ++u)
        m = fmaxf(fmaxf(m, fmaxf(fabsf(x[u][0]), fabsf(x[u][1]))), fmaxf(fabsf(x[u][2]), fabsf(x[u][3])));
    }
    return m;
  };
  float m = 0.0f;
  if (ld == rows) {  // the chunk is elements [begin, end) of the packed matrix
    m = run_max(p + c.begin, (int)(c.end - c.begin), m);
  } else {           // the chunk is columns [begin, end)
    for (int64_t col = c.begin; col < c.end; ++col) m = run_max(p + col * ld, rows, m);
  }
  m = block_max(m);
  if (threadIdx.x == 0) atomicMax(&amax[2 * c.mat + c.op], __float_as_uint(m));
}

// scales[4b..4b+2] = {sA, sB, post}, post the factor that undoes both scales.
// FP16: s = 2^e, e = 15 - x for amax = f * 2^x (f in [0.5, 1)), capped at
// 2^126 so a tiny amax keeps a finite scale; post = 2^-(eA + eB) exactly
// (the product sA*sB itself would overflow once eA + eB >= 128, i.e. for
// amaxA * amaxB below about 2^-98).  FP8: s = 448/amax, capped at 2^126;
// post = (1/sA) * (1/sB), each factor in range.
__global__ void __launch_bounds__(VBG_THREADS)
    vbg_scales_kernel(int prec, int batch, const unsigned* amax, float* scales) {
  const int b = blockIdx.x * VBG_THREADS + threadIdx.x;
  if (b >= batch) return;
  float s[2], inv[2];
  int e[2];
#pragma unroll
  for (int op = 0; op < 2; ++op) {
    const float m = __uint_as_float(amax[2 * b + op]);
    e[op] = 0;
    s[op] = 1.0f;
    inv[op] = 1.0f;
    if (!(m > 0.0f) || isinf(m)) {
      // no finite nonzero entry: unscaled
    } else if (prec == VBG_PREC_FP16_SCALED) {
      int x;
      frexpf(m, &x);  // m = f * 2^x, f in [0.5, 1)
      e[op] = min(15 - x, 126);
      s[op] = ldexpf(1.0f, e[op]);
    } else {
      s[op] = fminf(448.0f / m, 0x1p126f);
      inv[op] = 1.0f / s[op];
    }
  }
  scales[4 * b] = s[0];
  scales[4 * b + 1] = s[1];
  scales[4 * b + 2] = prec == VBG_PREC_FP16_SCALED ? ldexpf(1.0f, -(e[0] + e[1])) : inv[0] * inv[1];
  scales[4 * b + 3] = 0.0f;
}

}  // namespace

template <class Tile>
static int run_or_occ(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, vbg_split_kernel<Tile>, VBG_THREADS, 0) !=
        hipSuccess)
      nb = 1;
    *occ = nb > 0 ? nb : 1;
    return 0;
  }
  hipLaunchKernelGGL(vbg_split_kernel<Tile>, dim3(grid), dim3(VBG_THREADS), 0, s, *a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int BM, int BN, int MODE>
static int dispatch_tt(bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (!ta && !tb) return run_or_occ<SplitTile<BM, BN, MODE, false, false>>(a, grid, s, occ);
  if (!ta && tb) return run_or_occ<SplitTile<BM, BN, MODE, false, true>>(a, grid, s, occ);
  if (ta && !tb) return run_or_occ<SplitTile<BM, BN, MODE, true, false>>(a, grid, s, occ);
  return run_or_occ<SplitTile<BM, BN, MODE, true, true>>(a, grid, s, occ);
}

template <int MODE>
static int dispatch_mixed(bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (!ta && !tb) return run_or_occ<SplitMixed<MODE, false, false>>(a, grid, s, occ);
  if (!ta && tb) return run_or_occ<SplitMixed<MODE, false, true>>(a, grid, s, occ);
  if (ta && !tb) return run_or_occ<SplitMixed<MODE, true, false>>(a, grid, s, occ);
  return run_or_occ<SplitMixed<MODE, true, true>>(a, grid, s, occ);
}

template <int MODE>
static int dispatch_tile(int tile, bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s,
                         int* occ) {
  switch (tile) {
    case T64: return dispatch_tt<64, 64, MODE>(ta, tb, a, grid, s, occ);
    case T128x64: return dispatch_tt<128, 64, MODE>(ta, tb, a, grid, s, occ);
    case T64x128: return dispatch_tt<64, 128, MODE>(ta, tb, a, grid, s, occ);
    case T128: return dispatch_tt<128, 128, MODE>(ta, tb, a, grid, s, occ);
    case T256x128:
      if constexpr (MODE == MODE_BF16X3) return dispatch_tt<256, 128, MODE>(ta, tb, a, grid, s, occ);
      return -1;
    case T_MIXED: return dispatch_mixed<MODE>(ta, tb, a, grid, s, occ);
    default: return -1;
  }
}

// Packed FP16 / FP8 run on the shared LDS-DMA tiles (vbg_dma.h): FP16 on the
// 32x32x16 MFMA with the 256x256 shape for the big matrices, FP8 on the
// 32x32x64 f8f6f4 MFMA.  Packed bf16x3 (two planes per operand) keeps the
// register-staged loop.  Unbalanced B=2048 batch, GEMM kernel ms: the earlier
// 128x128 DMA ring (64-byte rows, 4 stages, one wait per tile) FP16 2.21 /
// FP8 1.60; these tiles FP16 1.94 + 0.08 / FP8 1.59; register staging FP16
// 2.19, FP8 1.85.
#ifndef VBG_PK_DMA
#define VBG_PK_DMA 1
#endif
template <int MODE>
static int dispatch_tile_pk(int tile, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if constexpr (VBG_PK_DMA && MODE != MODE_BF16X3) {
    constexpr int DM = MODE == MODE_FP16 ? DM_PK_FP16 : DM_PK_FP8;
    switch (tile) {
      case T64: return vbg_run_dmatile<DmaTile<64, 64, true, false, DM>>(a, grid, s, occ);
      case T128x64: return vbg_run_dmatile<DmaTile<128, 64, true, false, DM>>(a, grid, s, occ);
      case T64x128: return vbg_run_dmatile<DmaTile<64, 128, true, false, DM>>(a, grid, s, occ);
      case T128: return vbg_run_dmatile<DmaTile<128, 128, true, false, DM>>(a, grid, s, occ);
      case T256x128:
        // FP16 only: FP8's one-step k-tile needs an even ring, which at
        // 256x128 leaves one workgroup per CU (measured slower)
        if constexpr (DM == DM_PK_FP16) return vbg_run_dmatile<DmaTile<256, 128, true, false, DM>>(a, grid, s, occ);
        return -1;
      case T256:
        if constexpr (DM == DM_PK_FP16) return vbg_run_dmatile<DmaTile<256, 256, true, false, DM>>(a, grid, s, occ);
        return -1;
      case T_MIXED: return vbg_run_dmatile<DmaTileMixed<DM>>(a, grid, s, occ);
      default: return -1;
    }
  }
  switch (tile) {
    case T64: return run_or_occ<SplitTile<64, 64, MODE, false, false, true, true>>(a, grid, s, occ);
    case T128x64: return run_or_occ<SplitTile<128, 64, MODE, false, false, true, true>>(a, grid, s, occ);
    case T64x128: return run_or_occ<SplitTile<64, 128, MODE, false, false, true, true>>(a, grid, s, occ);
    case T128: return run_or_occ<SplitTile<128, 128, MODE, false, false, true, true>>(a, grid, s, occ);
    case T256x128:
      if constexpr (MODE == MODE_BF16X3)
        return run_or_occ<SplitTile<256, 128, MODE, false, false, true, true>>(a, grid, s, occ);
      return -1;
    case T_MIXED: return run_or_occ<SplitMixed<MODE, false, false, true>>(a, grid, s, occ);
    default: return -1;
  }
}

static int split_pk_dispatch(int prec, int tile, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  switch (prec) {
    case VBG_PREC_BF16X3: return dispatch_tile_pk<MODE_BF16X3>(tile, a, grid, s, occ);
    case VBG_PREC_FP16_SCALED: return dispatch_tile_pk<MODE_FP16>(tile, a, grid, s, occ);
    case VBG_PREC_FP8_E4M3: return dispatch_tile_pk<MODE_FP8>(tile, a, grid, s, occ);
    default: return -1;
  }
}

static int split_dispatch(int prec, int tile, bool ta, bool tb, const vbg_args* a, int grid,
                          hipStream_t s, int* occ) {
  switch (prec) {
    case VBG_PREC_BF16X3: return dispatch_tile<MODE_BF16X3>(tile, ta, tb, a, grid, s, occ);
    case VBG_PREC_FP16_SCALED: return dispatch_tile<MODE_FP16>(tile, ta, tb, a, grid, s, occ);
    case VBG_PREC_FP8_E4M3: return dispatch_tile<MODE_FP8>(tile, ta, tb, a, grid, s, occ);
    default: return -1;
  }
}

int vbg_launch_split(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream) {
  return split_dispatch(a->prec, tile, ta != 0, tb != 0, a, grid, (hipStream_t)stream, nullptr);
}

int vbg_occupancy_split(int prec, int tile, int ta, int tb) {
  int occ = 1;
  split_dispatch(prec, tile, ta != 0, tb != 0, nullptr, 0, nullptr, &occ);
  return occ;
}

int vbg_launch_split_pk(int tile, const vbg_args* a, int grid, void* stream) {
  return split_pk_dispatch(a->prec, tile, a, grid, (hipStream_t)stream, nullptr);
}

int vbg_occupancy_split_pk(int prec, int tile) {
  int occ = 1;
  split_pk_dispatch(prec, tile, nullptr, 0, nullptr, &occ);
  return occ;
}

int vbg_launch_pack(int prec, int ta, int tb, const vbg_pack_job* jobs, int n_jobs,
                    const vbg_mat_desc* desc, const float* const* A, const float* const* B,
                    const float* scales, char* pk, const int64_t* pk_off, void* stream) {
  if (n_jobs <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  switch (prec) {
    case VBG_PREC_BF16X3:
      hipLaunchKernelGGL(vbg_pack_kernel<MODE_BF16X3>, dim3(n_jobs), dim3(VBG_THREADS), 0, s, jobs, desc, A, B,
                         ta, tb, scales, pk, pk_off);
      break;
    case VBG_PREC_FP16_SCALED:
      hipLaunchKernelGGL(vbg_pack_kernel<MODE_FP16>, dim3(n_jobs), dim3(VBG_THREADS), 0, s, jobs, desc, A, B,
                         ta, tb, scales, pk, pk_off);
      break;
    case VBG_PREC_FP8_E4M3:
      hipLaunchKernelGGL(vbg_pack_kernel<MODE_FP8>, dim3(n_jobs), dim3(VBG_THREADS), 0, s, jobs, desc, A, B,
                         ta, tb, scales, pk, pk_off);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int vbg_launch_scales(int prec, int batch, int ta, int tb, const vbg_mat_desc* desc,
                      const vbg_chunk* chunks, int n_chunks, const float* const* A,
                      const float* const* B, unsigned* amax, float* scales, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nb = (2 * batch + VBG_THREADS - 1) / VBG_THREADS;
  hipLaunchKernelGGL(vbg_amax_zero, dim3(nb > 0 ? nb : 1), dim3(VBG_THREADS), 0, s, amax, 2 * batch);
  if (n_chunks > 0)
    hipLaunchKernelGGL(vbg_amax_kernel, dim3(n_chunks), dim3(VBG_THREADS), 0, s, chunks, desc, A, B,
                       ta, tb, amax);
  const int sb = (batch + VBG_THREADS - 1) / VBG_THREADS;
  hipLaunchKernelGGL(vbg_scales_kernel, dim3(sb > 0 ? sb : 1), dim3(VBG_THREADS), 0, s, prec, batch,
                     amax, scales);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
