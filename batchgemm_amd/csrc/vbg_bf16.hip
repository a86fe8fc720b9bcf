// vbg_bf16.hip -- BF16-operand tiles (v_mfma_f32_32x32x16_bf16, FP32
// accumulate) for the grouped forward / backward passes.  Operands are BF16 in
// HBM and are staged through LDS unconverted; C is written as FP32 or BF16.
//
// Loads are 16-byte vectors (8 BF16), so every leading dimension must be a
// multiple of 8 and every operand pointer 16-byte aligned (the plan checks the
// leading dimensions; the grouped entry points check the pointers).
#include <utility>

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

// ------------------------------------------------------ LDS-DMA staged tiles
// The operands go from HBM straight into LDS (buffer_load_dwordx4 ... lds:
// lane l's 16 bytes land at the instruction's LDS base + 16 l), through a ring
// of S stage buffers of BK = 32 k each, so S-1 k-tiles are in flight without
// holding any of them in VGPRs (the register-staged Bf16Tile holds at most two
// and spends VALU and LDS-write instructions on each).  Operands keep their
// HBM orientation in LDS:
//   KC (k contiguous): rows of 32 k (64 bytes, four 16-byte units, unit u of
//   row r at u ^ ((r / 4) % 4)); an MFMA operand is one ds_read_b128.
//   MC (mn contiguous): rows of one k over the tile's mn (2*ROWS bytes, in
//   64-byte blocks, block b of k-row r at b ^ f(r)); an MFMA operand of 8 k is
//   two ds_read_b64_tr_b16, the transposing LDS read of gfx950: within each
//   16-lane group, lane 4i + q addresses 4 consecutive mn at k-row i, and
//   lane j receives column j of that 4 x 16 block (mn j, 4 consecutive k).
//   The swizzles put the rows that one read instruction touches in distinct
//   LDS banks.  A DMA instruction fills 1 KB of consecutive LDS positions;
//   each lane fetches the global 16 bytes that belong at its position.
// Rows past MN read the last row (KC) or last 8-element unit (MC): they only
// feed outputs that are never stored.  The k-tile that crosses the end of the
// item's k range puts its whole offset in the range-checked vector offset and
// sends each 16-byte unit that starts at k >= end past the buffer's extent
// (no memory access); once it has landed, each lane zeroes its units at
// k >= end in LDS, element by element in a KC unit that straddles the end
// (K % 8 != 0).
//   Waits: each k-tile's barrier is preceded by vmcnt(N), N = the DMAs issued
// after that tile's (see vbg_dma16).
#ifndef VBG_BF16_DMA
#define VBG_BF16_DMA 1
#endif
// k depth of the 256x256 tile's stages: 64 (a ring of 2 x 64 KB) or 32 (4 x 32 KB)
#ifndef VBG_BF16_HUGE_BK
#define VBG_BF16_HUGE_BK 32
#endif

template <int ROWS, int BK, bool KC>
struct DmaPlane {
  static_assert(BK == 32 || BK == 64, "k-tile depth 32 or 64");
  static constexpr int BYTES = ROWS * BK * 2;
  static constexpr int PITCH = KC ? BK * 2 : ROWS * 2;  // bytes per LDS row
  static constexpr int NU = PITCH / 16;                 // KC: 16-byte units per row
  static constexpr int RPB = 256 / PITCH;               // KC: rows per 256-byte bank line
  static constexpr int NB = PITCH / 64;                 // MC: 64-byte blocks per k-row
  static_assert(KC || NB == 2 || NB % 4 == 0, "MC rows of 2 or a multiple of 4 blocks");
  __device__ __forceinline__ static int kc_unit(int row, int u) { return (u ^ (row / RPB)) & (NU - 1); }
  __device__ __forceinline__ static int mc_f(int r) { return NB >= 4 ? (r & 3) : ((r >> 1) & 1); }

  // The 16 bytes lane `lane` of wave `wave` fetches for this plane's chunk
  // 4 i + wave (a chunk is 1 KB of consecutive LDS positions): its mn row
  // (KC) or k-row (MC) advances by a fixed step with i while its unit or
  // block stays the same (the swizzles repeat every 4 KB), so a lane keeps
  // one origin per plane: KC row r0 + i * RSTEP at byte u16 within the row;
  // MC byte offset base0 + i * KSTEP * ld * 2, k-row kk0 + i * KSTEP.
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
      o.kk0 = 8 * unit;
    } else {
      const int krow = q / PITCH, pb = (q % PITCH) >> 6, within = q & 63;
      const int lb = pb ^ mc_f(krow);
      const int mn = min(mn0 + 32 * lb + (within >> 1), (MN - 1) & ~7);
      o.base0 = (mn + krow * ld) * 2;
      o.kk0 = krow;
    }
    return o;
  }
  // Byte offset (k-tile at k = 0) of chunk i: KC rows past MN read row MN - 1
  // (they only feed outputs that are never stored), so no offset leaves the
  // operand whatever k offset is added within a whole tile.
  __device__ __forceinline__ static int off(const Src& o, int i, int MN, int ld) {
    if (KC) return min(o.r0 + i * RSTEP, MN - 1) * (2 * ld) + o.u16;
    return o.base0 + i * KSTEP * (2 * ld);
  }
  __device__ __forceinline__ static int kk(const Src& o, int i) { return o.kk0 + i * KSTEP; }
  // bytes of one k step in HBM
  __device__ __forceinline__ static int kstride(int ld) { return KC ? 2 : 2 * ld; }
  // Operand extent in bytes: KC MN rows of K at ld, MC K rows of MN at ld,
  // the last row rounded up to whole 16-byte units (the range check works per
  // dword and would zero valid elements that share a dword with the end).
  // With ld % 8 == 0 and a 16-byte-aligned base, that unit lies in the
  // 16-byte block of the operand's last element.
  __device__ __forceinline__ static uint32_t extent(int MN, int K, int ld) {
    if (MN <= 0 || K <= 0) return 0u;
    return KC ? ((uint32_t)(MN - 1) * (uint32_t)ld + (uint32_t)((K + 7) & ~7)) * 2u
              : ((uint32_t)(K - 1) * (uint32_t)ld + (uint32_t)((MN + 7) & ~7)) * 2u;
  }
  // The MFMA operand of fragment rows fb .. fb+31 (fb % 32 == 0) at k16 step
  // s: lane (l32, h) gets mn fb + l32, k 16 s + 8 h + [0, 8).
  __device__ __forceinline__ static v8bf read(const char* plane, int fb, int s, int lane) {
    if (KC) {
      const int row = fb + (lane & 31), unit = 2 * s + (lane >> 5);
      return *(const v8bf*)(plane + row * PITCH + (kc_unit(row, unit) << 4));
    } else {
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
  }
};

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

// Shapes.  The bytes a CU can move from HBM/L2 into LDS per cycle through the
// buffer-load address path, not the MFMA rate, bound the smaller tiles (a
// 128x128 tile needs about as many bytes per MFMA cycle as that path
// delivers), so the large shapes carry the big matrices:
//   256x256: 4 waves of 128 x 128 (one per SIMD, 256 accumulators each), 64-deep
//            k-tiles (operands stored k-contiguous fetch whole 128-byte lines),
//            a ring of 2 x 64 KB, one workgroup per CU;
//   256x128: 4 waves of 128 x 64, 32-deep k-tiles, a ring of 3 x 24 KB, two
//            workgroups per CU;
//   64x64 .. 128x128: 4 waves of up to 64 x 64, 32-deep k-tiles, a ring of
//            4 x <= 16 KB, two workgroups per CU.
template <int BM, int BN, bool TA, bool TB, bool OUT_BF16>
struct Bf16DmaTile {
  static constexpr bool HUGE = BM * BN >= 256 * 256;
  static constexpr bool BIG = BM * BN > 128 * 128;
  static constexpr int BK = HUGE ? VBG_BF16_HUGE_BK : 32;
  static constexpr int KS = BK / 16;  // MFMA k-steps per k-tile (even)
  static constexpr int S = HUGE ? 131072 / ((BM + BN) * BK * 2) : BIG ? 3 : 4;
  static constexpr int WAVES = HUGE ? 1 : 2;  // waves per SIMD
  static_assert(S >= 2 && S <= 4, "stage ring of 2 .. 4 buffers");
  using PA = DmaPlane<BM, BK, TA>;   // A is k-contiguous when transposed
  using PB = DmaPlane<BN, BK, !TB>;  // B is k-contiguous when not transposed
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
  struct Frags {
    v8bf a[TM], b[TN];
  };

  __device__ __forceinline__ static void run(const vbg_args& a, const vbg_item& it, char* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const void* A = vbg_uptr(vbg_cload(a.A, it.mat));
    const void* B = vbg_uptr(vbg_cload(a.B, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    const int lane = threadIdx.x & 63, wave = vbg_u(threadIdx.x >> 6);
    const int wr = wave & 1, wc = wave >> 1;
    const int l32 = lane & 31, h = lane >> 5;

    // DMA sources: chunk 4 i + wave of each stage is this wave's i-th (the
    // first CA chunks are A's), so whether an instruction fills A or B is
    // known at compile time.
    const v4u32_t rsA = vbg_rsrc_words(A, PA::extent(d.M, d.K, d.lda));
    const v4u32_t rsB = vbg_rsrc_words(B, PB::extent(d.N, d.K, d.ldb));
    const int ksA = PA::kstride(d.lda), ksB = PB::kstride(d.ldb);
    const typename PA::Src sa = PA::src(wave, lane, m0, d.M, d.lda);
    const typename PB::Src sb = PB::src(wave, lane, n0, d.N, d.ldb);
    constexpr int IA = CA / 4;  // this wave's chunks of A per stage
    auto off = [&](int i) { return i < IA ? PA::off(sa, i, d.M, d.lda) : PB::off(sb, i - IA, d.N, d.ldb); };
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
    // Wait for tile t (n tiles issued after it), and when it crosses the end
    // of the k range zero what this lane's DMAs placed at k >= end, element by
    // element in a KC unit that straddles the end (K % 8 != 0).
    auto arrive = [&](char* stg, int t, int n) {
      vbg_wait_tiles<NI, S - 1>(n);
      const int k0 = kb + t * BK;
      if (k0 + BK <= ke) return;  // wave-uniform
      const int rel = ke - k0;
      char* mine = stg + wave * 1024 + 16 * lane;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const bool kc = i < IA ? TA : !TB;
        v4u* p = (v4u*)(mine + 4096 * i);
        if (kk(i) >= rel) {
          *p = v4u{0u, 0u, 0u, 0u};
        } else if (kc && kk(i) + 8 > rel) {
          const int lim = rel - kk(i);  // elements of this unit below the end
          v4u x = *p;
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
      for (int r = 0; r < TM; ++r) f.a[r] = PA::read(stg, wr * WTM + r * 32, s, lane);
#pragma unroll
      for (int c = 0; c < TN; ++c) f.b[c] = PB::read(stg + PLANE_A, wc * WTN + c * 32, s, lane);
    };
    // MFMAs of one k16 step; FROM skips the first `from` (issued earlier)
    auto mma = [&](const Frags& f, int from = 0) {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          if (r * TN + c >= from)
            acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[r], f.b[c], acc[r][c], 0, 0, 0);
    };
    auto mma_first = [&](const Frags& f) {
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[0], f.b[0], acc[0][0], 0, 0, 0);
    };

    // Pipeline.  Tile t sits in stage t % S.  The MFMA operands of each k16
    // step are read from LDS one step ahead, into the other register set, so
    // the LDS latency overlaps the previous step's MFMAs.  Moving to tile
    // t + 1: wait for its DMAs (and, via lgkmcnt(0), for this wave's reads of
    // tile t), meet at the barrier -- after which no wave reads stage t % S
    // again -- refill that stage with tile t + S, then read tile t + 1's
    // first step.  S - 1 tiles are in flight ahead of the MFMAs.
    // FAST: the caller knows tile t + 1 is a whole tile with at least S - 2
    // tiles after it (the branch-free steady state).
    auto next = [&](auto I, auto FAST, int t, Frags& f) {
      constexpr int i = decltype(I)::value;
      if constexpr (decltype(FAST)::value) vbg_wait_vm_barrier<(S - 2) * NI>();
      else arrive(stage((i + 1) % S), t + 1, min(S - 2, nk - 2 - t));
      if (t + S < nk) dma(stage(i), t + S);
      read(f, stage((i + 1) % S), 0);
    };
    // One k-tile: KS k16 steps, alternating between the two register sets
    // (KS is even, so every tile starts in f0).  Each step issues its first
    // MFMA, then the reads of the next step (the next tile's first step,
    // after the tile transition, in the last), then its other MFMAs.
    // Scheduling fences keep each phase whole.  (The first MFMA goes before
    // the reads because the compiler waits for every outstanding LDS read
    // before the first use of a transposing read's result, so reads issued
    // just before that use would stall it.)
    Frags f0, f1;
    auto body = [&](auto I, auto FAST, int t) {
      constexpr bool fast = decltype(FAST)::value;
      vbg_static_for<KS>([&](auto SS) {
        constexpr int s = decltype(SS)::value;
        Frags& cur = s % 2 == 0 ? f0 : f1;
        Frags& nxt = s % 2 == 0 ? f1 : f0;
        mma_first(cur);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (s + 1 < KS) read(nxt, stage(decltype(I)::value), s + 1);
        else if (fast || t + 1 < nk) next(I, FAST, t, nxt);
        __builtin_amdgcn_sched_barrier(0);
        mma(cur, 1);
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
    // successor that is a whole tile with S - 2 tiles issued after it, so the
    // group runs without branches around its waits.
    const int nwhole = (ke - kb) / BK;
    int t = 0;
    for (; t + S + S - 2 < nwhole; t += S)
      vbg_static_for<S>([&](auto J) { body(J, std::true_type{}, t + decltype(J)::value); });
    for (; t + S <= nk; t += S)
      vbg_static_for<S>([&](auto J) { body(J, std::false_type{}, t + decltype(J)::value); });
    vbg_static_for<S - 1>([&](auto J) {
      if (t + decltype(J)::value < nk) body(J, std::false_type{}, t + decltype(J)::value);
    });
    // no stage is refilled by the next item before every wave has read this
    // item's last tile
    vbg_wait_vm_barrier<0>();

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

template <class Tile>
__global__ void __launch_bounds__(VBG_THREADS) __attribute__((amdgpu_waves_per_eu(Tile::WAVES)))
vbg_bf16_dma_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(1024))) char lds[Tile::LDS_BYTES];
  vbg_persistent_loop(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

// Two workgroups per CU (the 64 KB LDS of the 128x128 tile allows no more):
// the register allocation is held to two waves per SIMD.
template <class Tile>
__global__ void __launch_bounds__(VBG_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
vbg_bf16_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(16))) float lds[Tile::LDS_FLOATS];
  vbg_persistent_loop(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

}  // namespace

template <class Tile, bool DMA>
static int run_or_occ(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  auto kern = [] {
    if constexpr (DMA) return vbg_bf16_dma_kernel<Tile>;
    else return vbg_bf16_kernel<Tile>;
  }();
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, VBG_THREADS, 0) != hipSuccess) nb = 1;
    *occ = nb > 0 ? nb : 1;
    return 0;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(VBG_THREADS), 0, s, *a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int BM, int BN, bool TA, bool TB, bool OB>
static int run_tile(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if constexpr (VBG_BF16_DMA || BM * BN > 128 * 128)
    return run_or_occ<Bf16DmaTile<BM, BN, TA, TB, OB>, true>(a, grid, s, occ);
  else
    return run_or_occ<Bf16Tile<BM, BN, TA, TB, OB>, false>(a, grid, s, occ);
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
