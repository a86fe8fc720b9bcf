// vbg_dma.h -- LDS-DMA staged MFMA tiles, shared by the BF16-operand family
// (the grouped passes) and the split precisions on packed operands (FP16,
// FP8).
//
// The operands go from HBM straight into LDS (buffer_load_dwordx4 ... lds:
// lane l's 16 bytes land at the instruction's LDS base + 16 l), through a ring
// of S stage buffers, so S-1 k-tiles are in flight without holding any of them
// in VGPRs.  Operands keep their HBM orientation in LDS:
//   KC (k contiguous): rows of one mn over the k-tile (RB = 64 bytes: four
//   16-byte units, unit u of row r at u ^ ((r / 4) % 4)); an MFMA operand is
//   one or two ds_read_b128.
//   MC (mn contiguous, BF16 only): rows of one k over the tile's mn (2*ROWS
//   bytes, in 64-byte blocks, block b of k-row r at b ^ f(r)); an MFMA operand
//   of 8 k is two ds_read_b64_tr_b16, the transposing LDS read of gfx950:
//   within each 16-lane group, lane 4i + q addresses 4 consecutive mn at k-row
//   i, and lane j receives column j of that 4 x 16 block (mn j, 4 consecutive
//   k).
//   The swizzles put the rows that one read instruction touches in distinct
// LDS banks.  A DMA instruction fills 1 KB of consecutive LDS positions; each
// lane fetches the global 16 bytes that belong at its position.  Rows past MN
// read the last row (KC) or last 8-element unit (MC): they only feed outputs
// that are never stored.  The k-tile that crosses the end of the item's k
// range puts its whole offset in the range-checked vector offset and sends
// each 16-byte unit that starts at k >= end past the buffer's extent (no
// memory access); once it has landed, each lane zeroes its units at k >= end
// in LDS, element by element in a BF16 KC unit that straddles the end
// (K % 8 != 0; packed operands are zero past K already).
//   Waits: each k-tile's barrier is preceded by vmcnt(N), N = the DMAs issued
// after that tile's (see vbg_dma16).
#pragma once
#include <utility>

#include "vbg_tiles.h"

typedef __bf16 vbg_v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 vbg_v8h __attribute__((ext_vector_type(8)));
typedef int vbg_v8i __attribute__((ext_vector_type(8)));

// f(integral_constant<int, j>) for j = 0 .. N-1, in order.
template <class F, int... J>
__device__ __forceinline__ void vbg_static_for_impl(F&& f, std::integer_sequence<int, J...>) {
  (f(std::integral_constant<int, J>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void vbg_static_for(F&& f) {
  vbg_static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// s_waitcnt vmcnt(n * NI) + barrier for a run-time n in [0, MAXN].
template <int NI, int MAXN>
__device__ __forceinline__ void vbg_wait_tiles(int n) {
  if constexpr (MAXN > 0) {
    if (n >= MAXN) {
      vbg_wait_vm_barrier<MAXN * NI>();
      return;
    }
    vbg_wait_tiles<NI, MAXN - 1>(n);
  } else {
    vbg_wait_vm_barrier<0>();
  }
}

// 4-row column segment of a column-major BF16 C: C = bf16(alpha * v + beta * C).
__device__ __forceinline__ float vbg_bf2f(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ uint16_t vbg_f2bf(float x) {
  const __bf16 b = (__bf16)x;
  return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ void vbg_store4_bf16(uint16_t* C, int ldc, int row, int col, int M, int N, v4f v,
                                                float alpha, float beta) {
  typedef uint16_t u16x4 __attribute__((ext_vector_type(4), aligned(2)));
  if (col >= N || row >= M) return;
  uint16_t* c = C + (size_t)col * ldc + row;
  if (row + 3 < M) {  // 4 consecutive rows of a column: one 8-byte access
    v4f r = v * alpha;
    if (beta != 0.0f) {
      const u16x4 old = *(const u16x4*)c;
      r += beta * v4f{vbg_bf2f(old[0]), vbg_bf2f(old[1]), vbg_bf2f(old[2]), vbg_bf2f(old[3])};
    }
    *(u16x4*)c = u16x4{vbg_f2bf(r[0]), vbg_f2bf(r[1]), vbg_f2bf(r[2]), vbg_f2bf(r[3])};
    return;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (row + e < M) {
      float r = alpha * v[e];
      if (beta != 0.0f) r += beta * vbg_bf2f(c[e]);
      c[e] = vbg_f2bf(r);
    }
  }
}

// Operand modes of the DMA tiles.
//   BF16 / BF16_OUT: BF16 operands at the plan's A / B pointers, stored k- or
//     mn-contiguous (TA / TB), C in FP32 / BF16; v_mfma_f32_32x32x16_bf16.
//   PK_FP16 / PK_FP8: operands quantised into the plan's packed workspace
//     (both k-contiguous rows of Kp = K rounded up to 64 elements, zero past
//     K), C in FP32 times the per-matrix post-scale; FP16 on
//     v_mfma_f32_32x32x16_f16, FP8 e4m3 on v_mfma_scale_f32_32x32x64_f8f6f4
//     with unit scales (twice the k per MFMA of the 32x32x16 FP8 form).
enum { DM_BF16 = 0, DM_BF16_OUT = 1, DM_PK_FP16 = 2, DM_PK_FP8 = 3 };
template <int DM>
struct DmaMode {
  static constexpr bool PK = DM == DM_PK_FP16 || DM == DM_PK_FP8;
  static constexpr bool OUT_BF16 = DM == DM_BF16_OUT;
  static constexpr int EB = DM == DM_PK_FP8 ? 1 : 2;  // bytes per element
  static constexpr int LB = DM == DM_PK_FP8 ? 32 : 16;  // bytes per lane per MFMA operand
  static constexpr int STEP_K = LB * 2 / EB;            // k per MFMA (two lane halves)
};

// One operand's plane of a stage: ROWS rows of RB bytes (KC) or RB / EB
// k-rows of ROWS elements (MC).
template <int ROWS, int RB, int EB, bool KC>
struct DmaPlane {
  static_assert(RB == 64, "64-byte k-tile rows");
  static constexpr bool IS_KC = KC;
  static constexpr int BK = RB / EB;  // k per tile
  static constexpr int BYTES = ROWS * RB;
  static constexpr int PITCH = KC ? RB : ROWS * EB;  // bytes per LDS row
  static constexpr int NU = PITCH / 16;              // KC: 16-byte units per row
  static constexpr int RPB = 256 / PITCH;            // KC: rows per 256-byte bank line
  static constexpr int NB = PITCH / 64;              // MC: 64-byte blocks per k-row
  static_assert(KC || (EB == 2 && (NB == 2 || NB % 4 == 0)), "MC: BF16 rows of 2 or 4k blocks");
  __device__ __forceinline__ static int kc_unit(int row, int u) { return (u ^ (row / RPB)) & (NU - 1); }
  __device__ __forceinline__ static int mc_f(int r) { return NB >= 4 ? (r & 3) : ((r >> 1) & 1); }

  // The 16 bytes lane `lane` of wave `wave` fetches for this plane's chunk
  // 4 i + wave (a chunk is 1 KB of consecutive LDS positions): its mn row
  // (KC) or k-row (MC) advances by a fixed step with i while its unit or
  // block stays the same (the swizzles repeat every 4 KB), so a lane keeps
  // one origin per plane: KC row r0 + i * RSTEP at byte u16 within the row;
  // MC byte offset base0 + i * KSTEP * ld * EB, k-row kk0 + i * KSTEP.
  static constexpr int RSTEP = KC ? 4096 / PITCH : 0;
  static constexpr int KSTEP = KC ? 0 : 4096 / PITCH;
  struct Src {
    int r0, u16;  // KC: tile row of chunk 0, byte of the unit within the row
    int base0;    // MC: byte offset of chunk 0 (k-tile at k = 0)
    int kk0;      // first k of chunk 0's unit within the k-tile
  };
  __device__ __forceinline__ static Src src(int wave, int lane, int mn0, int MN, int ld) {
    const int q = wave * 1024 + 16 * lane;
    Src o{};
    if (KC) {
      const int row = q / PITCH, unit = kc_unit(row, (q % PITCH) >> 4);
      o.r0 = mn0 + row;
      o.u16 = 16 * unit;
      o.kk0 = 16 * unit / EB;
    } else {
      const int krow = q / PITCH, pb = (q % PITCH) >> 6, within = q & 63;
      const int lb = pb ^ mc_f(krow);
      const int mn = min(mn0 + 32 * lb + (within >> 1), (MN - 1) & ~7);
      o.base0 = (mn + krow * ld) * EB;
      o.kk0 = krow;
    }
    return o;
  }
  // Byte offset (k-tile at k = 0) of chunk i: KC rows past MN read row MN - 1
  // (they only feed outputs that are never stored), so no offset leaves the
  // operand whatever k offset is added within a whole tile.
  __device__ __forceinline__ static int off(const Src& o, int i, int MN, int ld) {
    if (KC) return min(o.r0 + i * RSTEP, MN - 1) * (EB * ld) + o.u16;
    return o.base0 + i * KSTEP * (EB * ld);
  }
  __device__ __forceinline__ static int kk(const Src& o, int i) { return o.kk0 + i * KSTEP; }
  // bytes of one k step in HBM
  __device__ __forceinline__ static int kstride(int ld) { return KC ? EB : EB * ld; }
  // Operand extent in bytes: KC MN rows of K at ld, MC K rows of MN at ld,
  // the last row rounded up to whole 16-byte units (the range check works per
  // dword and would zero valid elements that share a dword with the end).
  // With ld a multiple of the unit and a 16-byte-aligned base, that unit lies
  // in the 16-byte block of the operand's last element.
  __device__ __forceinline__ static uint32_t extent(int MN, int K, int ld) {
    if (MN <= 0 || K <= 0) return 0u;
    const int ue = 16 / EB;
    return KC ? ((uint32_t)(MN - 1) * (uint32_t)ld + (uint32_t)((K + ue - 1) / ue * ue)) * EB
              : ((uint32_t)(K - 1) * (uint32_t)ld + (uint32_t)((MN + ue - 1) / ue * ue)) * EB;
  }
  // 16 bytes of the MFMA operand of fragment rows fb .. fb+31 (fb % 32 == 0):
  // KC unit u of row fb + l32.
  __device__ __forceinline__ static v4u32_t read_kc(const char* plane, int fb, int u, int lane) {
    const int row = fb + (lane & 31);
    return *(const v4u32_t*)(plane + row * PITCH + (kc_unit(row, u) << 4));
  }
  // MC (BF16): the 8 k of MFMA step s for lane (l32, h): mn fb + l32, k 16 s +
  // 8 h + [0, 8).
  __device__ __forceinline__ static vbg_v8bf read_mc(const char* plane, int fb, int s, int lane) {
    typedef __bf16 v4bf_t __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) v4bf_t* LP;
    const int G = lane >> 4, j = lane & 15;
    const int kr = 16 * s + 8 * (G >> 1) + (j >> 2);
    const int lb = fb >> 5, off = 32 * (G & 1) + 8 * (j & 3);
    const char* p0 = plane + kr * PITCH + ((lb ^ mc_f(kr)) << 6) + off;
    const char* p1 = plane + (kr + 4) * PITCH + ((lb ^ mc_f(kr + 4)) << 6) + off;
    const v4bf_t x0 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((LP)(p0));
    const v4bf_t x1 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((LP)(p1));
    return __builtin_shufflevector(x0, x1, 0, 1, 2, 3, 4, 5, 6, 7);
  }
};

// Shapes.  The bytes a CU can move from HBM/L2 into LDS per cycle through the
// buffer-load address path, not the MFMA rate, bound the smaller tiles (a
// 128x128 BF16 tile needs about as many bytes per MFMA cycle as that path
// delivers), so the large shapes carry the big matrices.  k-tiles are 64-byte
// rows (32 k of BF16 / FP16, 64 of FP8):
//   256x256: 4 waves of 128 x 128 (one per SIMD, 256 accumulators each), a
//            ring of 4 x 32 KB, one workgroup per CU;
//   256x128: 4 waves of 128 x 64, a ring of 3 x 24 KB, two workgroups per CU;
//   64x64 .. 128x128: 4 waves of up to 64 x 64, a ring of 4 x <= 16 KB, two
//            workgroups per CU.
template <int BM, int BN, bool TA, bool TB, int DM>
struct DmaTile {
  using MD = DmaMode<DM>;
  static constexpr int RB = 64;
  static constexpr int EB = MD::EB;
  static constexpr int BK = RB / EB;
  static constexpr bool HUGE = BM * BN >= 256 * 256;
  static constexpr bool BIG = BM * BN > 128 * 128;
  static constexpr int KS = BK / MD::STEP_K;  // MFMAs along k per k-tile
  static constexpr int S = BIG && !HUGE ? 3 : 4;
  static constexpr int WAVES = HUGE ? 1 : 2;  // waves per SIMD
  static_assert(S * KS % 2 == 0, "the register sets alternate by MFMA step across the ring");
  static_assert(!MD::PK || (TA && !TB), "packed operands are k-contiguous rows");
  using PA = DmaPlane<BM, RB, EB, TA>;   // A is k-contiguous when transposed
  using PB = DmaPlane<BN, RB, EB, !TB>;  // B is k-contiguous when not transposed
  static constexpr int PLANE_A = PA::BYTES, PLANE_B = PB::BYTES;
  static constexpr int STAGE = PLANE_A + PLANE_B;
  static constexpr int LDS_BYTES = S * STAGE;
  static constexpr int CA = PLANE_A / 1024;    // 1 KB chunks of A per stage
  static constexpr int NI = STAGE / 1024 / 4;  // DMA instructions per wave per stage
  static_assert(STAGE % 4096 == 0 && CA % 4 == 0, "whole 1 KB chunks, four per wave row");
  static_assert(NI * (S - 1) < 64, "vmcnt range");
  static constexpr int WTM = BM / 2, WTN = BN / 2;
  static constexpr int TM = WTM / 32, TN = WTN / 32;
  static_assert(TM >= 1 && TN >= 1, "tile needs >= 64x64");
  typedef std::conditional_t<DM == DM_PK_FP8, vbg_v8i, std::conditional_t<DM == DM_PK_FP16, vbg_v8h, vbg_v8bf>> Op;
  struct Frags {
    Op a[TM], b[TN];
  };

  // The MFMA operand for MFMA step s of a k-tile (STEP_K k each): lane
  // (l32, h) reads units (2 s + h) * UPO + j, j < UPO, of its row (KC), or
  // the transposed k-rows (MC).  A and B lanes read the same k positions, so
  // the products pair the right k whatever order they sit in.
  template <class P>
  __device__ __forceinline__ static Op read_op(const char* plane, int fb, int s, int lane) {
    if constexpr (!P::IS_KC) {
      return __builtin_bit_cast(Op, P::read_mc(plane, fb, s, lane));
    } else {
      constexpr int UPO = MD::LB / 16;
      const int u0 = (2 * s + (lane >> 5)) * UPO;
      if constexpr (UPO == 1) {
        return __builtin_bit_cast(Op, P::read_kc(plane, fb, u0, lane));
      } else {
        const v4u32_t x0 = P::read_kc(plane, fb, u0, lane), x1 = P::read_kc(plane, fb, u0 + 1, lane);
        return __builtin_bit_cast(Op, __builtin_shufflevector(x0, x1, 0, 1, 2, 3, 4, 5, 6, 7));
      }
    }
  }
  __device__ __forceinline__ static v16f mfma(const Op& x, const Op& y, const v16f& c) {
    if constexpr (DM == DM_PK_FP8)
      return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x, y, c, 0, 0, 0, 127, 0, 127);
    else if constexpr (DM == DM_PK_FP16)
      return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
  }

  __device__ __forceinline__ static void run(const vbg_args& a, const vbg_item& it, char* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    const int lane = threadIdx.x & 63, wave = vbg_u(threadIdx.x >> 6);
    const int wr = wave & 1, wc = wave >> 1;
    const int l32 = lane & 31, h = lane >> 5;

    // Operands: the plan's pointers, or the packed workspace (rows of Kp).
    const void* A;
    const void* B;
    int lda, ldb;
    if constexpr (MD::PK) {
      const int Kp = (d.K + VBG_PACK_KALIGN - 1) & ~(VBG_PACK_KALIGN - 1);
      A = vbg_uptr(a.pk + vbg_u64(vbg_cload(a.pk_off, 2 * it.mat)));
      B = vbg_uptr(a.pk + vbg_u64(vbg_cload(a.pk_off, 2 * it.mat + 1)));
      lda = ldb = Kp;
    } else {
      A = vbg_uptr(vbg_cload(a.A, it.mat));
      B = vbg_uptr(vbg_cload(a.B, it.mat));
      lda = d.lda;
      ldb = d.ldb;
    }
    // DMA sources: chunk 4 i + wave of each stage is this wave's i-th (the
    // first CA chunks are A's), so whether an instruction fills A or B is
    // known at compile time.
    const v4u32_t rsA = vbg_rsrc_words(A, PA::extent(d.M, d.K, lda));
    const v4u32_t rsB = vbg_rsrc_words(B, PB::extent(d.N, d.K, ldb));
    const int ksA = PA::kstride(lda), ksB = PB::kstride(ldb);
    const typename PA::Src sa = PA::src(wave, lane, m0, d.M, lda);
    const typename PB::Src sb = PB::src(wave, lane, n0, d.N, ldb);
    constexpr int IA = CA / 4;  // this wave's chunks of A per stage
    auto off = [&](int i) { return i < IA ? PA::off(sa, i, d.M, lda) : PB::off(sb, i - IA, d.N, ldb); };
    auto kk = [&](int i) { return i < IA ? PA::kk(sa, i) : PB::kk(sb, i - IA); };
    auto stage = [&](int j) { return lds + j * STAGE; };
    auto dma = [&](char* stg, int t) {
      const int k0 = kb + t * BK;
      const uint32_t lbase = (uint32_t)(uintptr_t)stg + (uint32_t)(wave * 1024);
      if (k0 + BK <= ke) {  // wave-uniform
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          if (i < IA) vbg_dma16(rsA, vbg_u(lbase + 4096 * i), off(i), vbg_u(k0 * ksA));
          else vbg_dma16(rsB, vbg_u(lbase + 4096 * i), off(i), vbg_u(k0 * ksB));
        }
      } else {
        // the whole offset in the range-checked vector offset; units at
        // k >= end go past the extent (no access)
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int ks = i < IA ? ksA : ksB;
          const int v = kk(i) < ke - k0 ? off(i) + k0 * ks : (int)0xfffffff0u;
          vbg_dma16(i < IA ? rsA : rsB, vbg_u(lbase + 4096 * i), v, 0);
        }
      }
    };
    // One whole-tile DMA instruction (chunk i of tile t into stage stg).
    auto dma_chunk = [&](char* stg, int t, auto I) {
      constexpr int i = decltype(I)::value;
      const int k0 = kb + t * BK;
      const uint32_t lbase = (uint32_t)(uintptr_t)stg + (uint32_t)(wave * 1024);
      if constexpr (i < IA) vbg_dma16(rsA, vbg_u(lbase + 4096 * i), off(i), vbg_u(k0 * ksA));
      else vbg_dma16(rsB, vbg_u(lbase + 4096 * i), off(i), vbg_u(k0 * ksB));
    };
    // Wait for tile t (n tiles issued after it), and when it crosses the end
    // of the k range zero what this lane's DMAs placed at k >= end.
    auto arrive = [&](char* stg, int t, int n) {
      vbg_wait_tiles<NI, S - 1>(n);
      const int k0 = kb + t * BK;
      if (k0 + BK <= ke) return;  // wave-uniform
      const int rel = ke - k0;
      char* mine = stg + wave * 1024 + 16 * lane;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const bool kc = i < IA ? TA : !TB;
        v4u32_t* p = (v4u32_t*)(mine + 4096 * i);
        if (kk(i) >= rel) {
          *p = v4u32_t{0u, 0u, 0u, 0u};
        } else if (!MD::PK && kc && kk(i) + 8 > rel) {
          const int lim = rel - kk(i);  // BF16 elements of this unit below the end
          v4u32_t x = *p;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            x[e] &= (2 * e < lim ? 0xffffu : 0u) | (2 * e + 1 < lim ? 0xffff0000u : 0u);
          *p = x;
        }
      }
      vbg_wait_vm_barrier<0>();
    };

    v16f acc[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[r][c][e] = 0.0f;
    auto read = [&](Frags& f, const char* stg, int s) {
#pragma unroll
      for (int r = 0; r < TM; ++r) f.a[r] = read_op<PA>(stg, wr * WTM + r * 32, s, lane);
#pragma unroll
      for (int c = 0; c < TN; ++c) f.b[c] = read_op<PB>(stg + PLANE_A, wc * WTN + c * 32, s, lane);
    };
    // operand w of a step: A fragment w (w < TM) or B fragment w - TM
    auto read_one = [&](Frags& f, const char* stg, int s, auto W) {
      constexpr int w = decltype(W)::value;
      if constexpr (w < TM) f.a[w] = read_op<PA>(stg, wr * WTM + w * 32, s, lane);
      else f.b[w - TM] = read_op<PB>(stg + PLANE_A, wc * WTN + (w - TM) * 32, s, lane);
    };
    // MFMA j of a step (row j / TN, column j % TN), then the extras of its
    // slot: the E extra instructions (LDS reads, DMAs) of a step are spread
    // over the slots after its MFMAs 1 .. TM*TN-1, in order, each pair fenced,
    // so every MFMA's pipe time covers the issue of the work beside it.
    auto interleave = [&](const Frags& f, auto E, auto&& extra) {
      constexpr int NE = decltype(E)::value;
      constexpr int SL = TM * TN - 1;
      if constexpr (SL == 0) {
        vbg_static_for<NE>([&](auto X) { extra(X); });
      } else {
        vbg_static_for<SL>([&](auto J) {
          constexpr int j = decltype(J)::value + 1;
          acc[j / TN][j % TN] = mfma(f.a[j / TN], f.b[j % TN], acc[j / TN][j % TN]);
          __builtin_amdgcn_sched_barrier(0);
          constexpr int lo = (decltype(J)::value * NE) / SL, hi = ((decltype(J)::value + 1) * NE) / SL;
          vbg_static_for<hi - lo>([&](auto X) { extra(std::integral_constant<int, lo + decltype(X)::value>{}); });
          __builtin_amdgcn_sched_barrier(0);
        });
      }
    };
    // MFMAs of one step; the first `from` were issued earlier
    auto mma = [&](const Frags& f, int from) {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          if (r * TN + c >= from) acc[r][c] = mfma(f.a[r], f.b[c], acc[r][c]);
    };

    // Pipeline.  Tile t sits in stage t % S.  The MFMA operands of each step
    // are read from LDS one step ahead, into the other register set, so the
    // LDS latency overlaps the previous step's MFMAs.  Moving to tile t + 1:
    // wait for its DMAs (and, via lgkmcnt(0), for this wave's reads of tile
    // t), meet at the barrier -- after which no wave reads stage t % S again --
    // refill that stage with tile t + S, then read tile t + 1's first step.
    // S - 1 tiles are in flight ahead of the MFMAs.  FAST: the caller knows
    // tile t + 1 is a whole tile with at least S - 2 tiles after it (the
    // branch-free steady state).
    auto next = [&](auto I, auto FAST, int t, Frags& f) {
      constexpr int i = decltype(I)::value;
      if constexpr (decltype(FAST)::value) vbg_wait_vm_barrier<(S - 2) * NI>();
      else arrive(stage((i + 1) % S), t + 1, min(S - 2, nk - 2 - t));
      if (t + S < nk) dma(stage(i), t + S);
      read(f, stage((i + 1) % S), 0);
    };
    // One k-tile: KS MFMA steps; the two register sets alternate by step
    // (global parity, so a group of the ring starts in f0).  Each step issues
    // its first MFMA, then the reads of the next step (the next tile's first
    // step, after the tile transition, in the last), then its other MFMAs.
    // Scheduling fences keep each phase whole.  (The first MFMA goes before
    // the reads because the compiler waits for every outstanding LDS read
    // before the first use of a transposing read's result, so reads issued
    // just before that use would stall it.)
    Frags f0, f1;
    constexpr int NR = TM + TN;  // operand reads of a step
    auto body = [&](auto I, auto FAST, int t) {
      constexpr int i = decltype(I)::value;
      constexpr bool fast = decltype(FAST)::value;
      vbg_static_for<KS>([&](auto SS) {
        constexpr int s = decltype(SS)::value;
        constexpr bool even = (i * KS + s) % 2 == 0;
        Frags& cur = even ? f0 : f1;
        Frags& nxt = even ? f1 : f0;
        acc[0][0] = mfma(cur.a[0], cur.b[0], acc[0][0]);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (s + 1 < KS) {
          // the next step's reads spread over this step's MFMAs
          interleave(cur, std::integral_constant<int, NR>{},
                     [&](auto W) { read_one(nxt, stage(i), s + 1, W); });
        } else if constexpr (fast) {
          // steady state: tile t + 1 is whole and tile t + S a whole tile to
          // fetch; after the barrier, the next tile's first reads, then the
          // refill of this stage, spread over the MFMAs
          vbg_wait_vm_barrier<(S - 2) * NI>();
          __builtin_amdgcn_sched_barrier(0);
          interleave(cur, std::integral_constant<int, NR + NI>{}, [&](auto X) {
            constexpr int x = decltype(X)::value;
            if constexpr (x < NR) read_one(nxt, stage((i + 1) % S), 0, std::integral_constant<int, x>{});
            else dma_chunk(stage(i), t + S, std::integral_constant<int, x - NR>{});
          });
        } else {
          if (t + 1 < nk) next(I, FAST, t, nxt);
          __builtin_amdgcn_sched_barrier(0);
          mma(cur, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
      });
    };
#pragma unroll
    for (int j = 0; j < S; ++j)
      if (j < nk) dma(stage(j), j);
    if (nk > 0) {
      arrive(stage(0), 0, min(S - 1, nk - 1));
      read(f0, stage(0), 0);
    }
    // Steady state (unrolled over the ring): every tile of the group has a
    // successor that is a whole tile with S - 2 tiles issued after it, and
    // the tile it refetches (t + S) is whole, so the group runs without
    // branches around its waits and DMAs.
    const int nwhole = (ke - kb) / BK;
    int t = 0;
    for (; t + 2 * S - 1 < nwhole; t += S)
      vbg_static_for<S>([&](auto J) { body(J, std::true_type{}, t + decltype(J)::value); });
    for (; t + S <= nk; t += S)
      vbg_static_for<S>([&](auto J) { body(J, std::false_type{}, t + decltype(J)::value); });
    vbg_static_for<S - 1>([&](auto J) {
      if (t + decltype(J)::value < nk) body(J, std::false_type{}, t + decltype(J)::value);
    });
    // no stage is refilled by the next item before every wave has read this
    // item's last tile
    vbg_wait_vm_barrier<0>();

    // Epilogue: C = alpha * post * acc + beta * C (post: the packed modes'
    // per-matrix scale), or the split-K partial panel.
    float post = 1.0f;
    if constexpr (MD::PK)
      post = __builtin_bit_cast(float, vbg_u(__builtin_bit_cast(int, vbg_cload(a.scales, 4 * it.mat + 2))));
    const OutTile o = vbg_out(a, d, it, post);
    constexpr int NSEG = TM * TN * 4;
    auto seg = [&](int sg, int& row, int& col, v4f& v) {
      const int r = sg / (TN * 4), c = (sg / 4) % TN, g = sg % 4;
      row = m0 + wr * WTM + r * 32 + 8 * g + 4 * h;
      col = n0 + wc * WTN + c * 32 + l32;
      v = v4f{acc[r][c][4 * g], acc[r][c][4 * g + 1], acc[r][c][4 * g + 2], acc[r][c][4 * g + 3]};
    };
    if (MD::OUT_BF16 && !o.partial) {
#pragma unroll
      for (int sg = 0; sg < NSEG; ++sg) {
        int row, col;
        v4f v;
        seg(sg, row, col, v);
        vbg_store4_bf16((uint16_t*)o.p, o.ld, row, col, d.M, d.N, v, o.alpha, o.beta);
      }
    } else {
      vbg_epilogue<NSEG>(o, d.M, d.N, seg);
    }
  }
};

// Mixed launch of the packed-operand tiles: the 64x64 .. 128x128 shapes share
// the ring, sized for the largest.
template <int DM>
struct DmaTileMixed {
  using Sq128 = DmaTile<128, 128, true, false, DM>;
  using Tall = DmaTile<128, 64, true, false, DM>;
  using Wide = DmaTile<64, 128, true, false, DM>;
  using Sq64 = DmaTile<64, 64, true, false, DM>;
  static constexpr int WAVES = 2;
  static constexpr int LDS_BYTES = Sq128::LDS_BYTES;
  __device__ __forceinline__ static void run(const vbg_args& a, const vbg_item& it, char* lds) {
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
vbg_dmatile_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(1024))) char lds[Tile::LDS_BYTES];
  vbg_persistent_loop<2>(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

// Launch (or, with occ, report the workgroups per CU of) a DMA-tile kernel.
template <class Tile>
int vbg_run_dmatile(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, vbg_dmatile_kernel<Tile>, VBG_THREADS, 0) != hipSuccess)
      nb = 1;
    *occ = nb > 0 ? nb : 1;
    return 0;
  }
  hipLaunchKernelGGL(vbg_dmatile_kernel<Tile>, dim3(grid), dim3(VBG_THREADS), 0, s, *a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
