// vbg_fp32.hip -- FP32 tile kernels for gfx950.
//
//  * Fp32MfmaTile<BM, BN>: 256 threads = 2x2 waves, each wave owns a
//    (BM/2) x (BN/2) sub-tile built from 32x32 blocks of
//    v_mfma_f32_32x32x2_f32.  Operands stream HBM -> registers -> LDS
//    (double-buffered, one barrier per k-tile); the next k-tile's HBM loads are
//    in flight while the MFMAs consume the current LDS buffer.
//  * SeqTile: the VALU reference-compatible tile, one fmaf chain per output
//    element over ascending k.
//
// Replaces the reference's tile functions gemm_128_16x16 ... gemm_256_64x64
// (ROCm/framework/kernel_128_new.h, kernel_256_new.h) and its dispatcher
// (ROCm/framework/kernel.h:9-42): a work item names the matrix and tile, so no
// workgroup is launched for an empty bounding-box slot.
#include <type_traits>

#include "vbg_tiles.h"

namespace {

// MFMA k-slot layout used by every FP32 MFMA tile: v_mfma_f32_32x32x2_f32
// sums over two k-slots, lanes 0-31 supplying slot 0 and lanes 32-63 slot 1.
// Any assignment of actual k to (slot, step) works as long as A and B agree;
// lane half h takes k in [h*BK/2, (h+1)*BK/2), so each lane reads its k values
// as contiguous float4s from the [mn][k] LDS layout.
// CPRE: the last k-tile loads the epilogue's first chunk of C (off inside the
// mixed kernel, where four bodies share one register budget).
template <int BM, int BN, int BK, bool TA, bool TB, bool CPRE = true>
struct Fp32MfmaTile {
  static constexpr int P = BK + 4;  // LDS pitch (floats); +4 staggers rows across banks
  static constexpr int STAGE = (BM + BN) * P;
  static constexpr int LDS_FLOATS = 2 * STAGE;
  static constexpr int WTM = BM / 2, WTN = BN / 2;
  static constexpr int TM = WTM / 32, TN = WTN / 32;
  static_assert(TM >= 1 && TN >= 1, "FP32 MFMA tile needs >= 64x64");

  // A tile: op(A) rows = M index.  op N: A[i + k*lda] (mn contiguous);
  // op T: A[k + i*lda] (k contiguous).  B tile: op(B) columns = N index.
  // op N: B[k + j*ldb] (k contiguous); op T: B[j + k*ldb] (mn contiguous).
  using LA = TileLoader<BM, BK, P, TA>;
  using LB = TileLoader<BN, BK, P, !TB>;

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const float* __restrict__ A = vbg_uptr(vbg_cload(a.A, it.mat));
    const float* __restrict__ B = vbg_uptr(vbg_cload(a.B, it.mat));
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
    typename LA::Regs ra;
    typename LB::Regs rb;
    if (nk > 0) {
      la.fetch(ra, kb, ke);
      lb.fetch(rb, kb, ke);
      LA::store(ra, lds);
      LB::store(rb, lds + BM * P);
    }
    __syncthreads();

    // 32x32x2 f32 accumulator layout: register 4g+e of lane l holds
    // row 8g + 4*(l/32) + e, column l%32 of the 32x32 block.
    constexpr int NSEG = TM * TN * 4;
    const OutTile o = vbg_out(a, d, it, 1.0f);
    auto seg = [&](int sg, int& row, int& col, v4f& v) {
      const int r = sg / (TN * 4), s = (sg / 4) % TN, g = sg % 4;
      row = m0 + wr * WTM + r * 32 + 8 * g + 4 * h;
      col = n0 + wc * WTN + s * 32 + l32;
      v = v4f{acc[r][s][4 * g], acc[r][s][4 * g + 1], acc[r][s][4 * g + 2], acc[r][s][4 * g + 3]};
    };
    // one k-tile; the last one (peeled) loads the epilogue's first chunk of
    // C in place of a next tile, so that round trip overlaps its MFMAs
    v4f cpre[EpiChunk<NSEG>::CH];
    auto ktile = [&](int kt, auto last, auto pre) {
      constexpr bool LAST = decltype(last)::value;
      const int cur = kt & 1;
      if constexpr (!LAST) {
        const int k0 = kb + (kt + 1) * BK;
        la.fetch(ra, k0, ke);
        lb.fetch(rb, k0, ke);
      } else if constexpr (decltype(pre)::value) {
        vbg_c_load<NSEG>(o, d.M, d.N, seg, 0, cpre);
      }
      const float* As = lds + cur * STAGE;
      const float* Bs = As + BM * P;
#pragma unroll
      for (int c = 0; c < BK / 8; ++c) {
        v4f av[TM], bv[TN];
#pragma unroll
        for (int r = 0; r < TM; ++r)
          av[r] = *(const v4f*)(As + (wr * WTM + r * 32 + l32) * P + h * (BK / 2) + 4 * c);
#pragma unroll
        for (int s = 0; s < TN; ++s)
          bv[s] = *(const v4f*)(Bs + (wc * WTN + s * 32 + l32) * P + h * (BK / 2) + 4 * c);
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int s = 0; s < TN; ++s)
              acc[r][s] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[r][e], bv[s][e], acc[r][s], 0, 0, 0);
      }
      if constexpr (!LAST) {
        float* nxt = lds + (cur ^ 1) * STAGE;
        LA::store(ra, nxt);
        LB::store(rb, nxt + BM * P);
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
};

// Small FP32 MFMA tiles on v_mfma_f32_16x16x4_f32 (16x16 block, 4 k-slots:
// lane l supplies A[l%16][slot l/16] and B[slot l/16][l%16]; register r of
// lane l holds D[4*(l/16) + r][l%16]).
//   BMN = 32: 2x2 waves, one 16x16 block each; lane group g = l/16 takes
//             k in [g*BK/4, (g+1)*BK/4) of every k-tile.
//   BMN = 16: the four waves split each k-tile four ways (wave w takes
//             [w*BK/4, (w+1)*BK/4), lane group g a quarter of that) and the
//             partial 16x16 blocks are summed through LDS at the end -- the
//             wavefront-level reduction for tiles too small to occupy 4 waves.
template <int BMN, bool TA, bool TB>
struct Fp32SmallTile {
  static constexpr bool KSPLIT = BMN == 16;
  static constexpr int BK = KSPLIT ? 64 : 32;
  static constexpr int P = BK + 4;
  static constexpr int STAGE = 2 * BMN * P;
  static constexpr int LDS_FLOATS = 2 * STAGE;
  static constexpr int KL = KSPLIT ? BK / 16 : BK / 4;  // k per lane per k-tile
  static_assert(KL % 4 == 0, "lane k range must be whole float4s");
  static_assert(!KSPLIT || LDS_FLOATS >= 4 * 256, "reduction buffer must fit in LDS");
  using LA = TileLoader<BMN, BK, P, TA>;
  using LB = TileLoader<BMN, BK, P, !TB>;

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const float* __restrict__ A = vbg_uptr(vbg_cload(a.A, it.mat));
    const float* __restrict__ B = vbg_uptr(vbg_cload(a.B, it.mat));
    const int m0 = it.tm * BMN, n0 = it.tn * BMN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int arow = KSPLIT ? l16 : (wave & 1) * 16 + l16;
    const int bcol = KSPLIT ? l16 : (wave >> 1) * 16 + l16;
    const int koff = KSPLIT ? wave * (BK / 4) + g * KL : g * KL;

    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    LA la;
    LB lb;
    la.init(A, d.lda, m0, d.M, d.K);
    lb.init(B, d.ldb, n0, d.N, d.K);
    typename LA::Regs ra;
    typename LB::Regs rb;
    if (nk > 0) {
      la.fetch(ra, kb, ke);
      lb.fetch(rb, kb, ke);
      LA::store(ra, lds);
      LB::store(rb, lds + BMN * P);
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
      const float* As = lds + cur * STAGE;
      const float* Bs = As + BMN * P;
#pragma unroll
      for (int c = 0; c < KL / 4; ++c) {
        const v4f av = *(const v4f*)(As + arow * P + koff + 4 * c);
        const v4f bv = *(const v4f*)(Bs + bcol * P + koff + 4 * c);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], bv[e], acc, 0, 0, 0);
      }
      if (more) {
        float* nxt = lds + (cur ^ 1) * STAGE;
        LA::store(ra, nxt);
        LB::store(rb, nxt + BMN * P);
      }
      __syncthreads();
    }
    const int row0 = 4 * g;  // rows row0..row0+3 of the 16x16 block, column l16
    const OutTile o = vbg_out(a, d, it, 1.0f);
    if (KSPLIT) {
      *(v4f*)(lds + wave * 256 + lane * 4) = acc;
      __syncthreads();
      if (wave == 0) {
        v4f s = acc;
#pragma unroll
        for (int w = 1; w < 4; ++w) s += *(const v4f*)(lds + w * 256 + lane * 4);
        vbg_store4(o.p, o.ld, m0 + row0, n0 + l16, d.M, d.N, s, o.alpha, o.beta);
      }
      __syncthreads();
    } else {
      vbg_store4(o.p, o.ld, m0 + (wave & 1) * 16 + row0, n0 + (wave >> 1) * 16 + l16, d.M, d.N, acc,
                 o.alpha, o.beta);
    }
  }
};

// Skinny FP32 tiles for matrices with one side of at most 16: 16 x 128 (M <=
// 16) and 128 x 16 (N <= 16), on v_mfma_f32_16x16x4_f32.  Such a matrix on a
// square tile computes mostly padding (an 8 x 1000 x 1000 product on 128x64
// tiles issues 16x its useful MFMA work) and re-reads the short operand once
// per tile of the long side; here one item covers 128 of the long side, the
// short operand's 16 x 64 k-slab is loaded once per k-tile for all of it, and
// each wave owns two 16x16 blocks along the long side.  Within a block the
// 64-deep k-tile is spread over the lanes -- lane group g = l/16 takes k in
// [16g, 16g + 16), as four float4 LDS reads -- and the MFMA's 4-slot sum adds
// the lane groups' partial products, so the k reduction happens across lanes
// inside the matrix core with no LDS round trip.
template <bool WIDE, bool TA, bool TB>
struct Fp32SkinnyTile {
  static constexpr int BM = WIDE ? 16 : 128, BN = WIDE ? 128 : 16;
  static constexpr int BK = 64;
  static constexpr int P = BK + 4;
  static constexpr int STAGE = (BM + BN) * P;
  static constexpr int LDS_FLOATS = 2 * STAGE;
  static constexpr int KL = BK / 4;  // k per lane group per k-tile
  using LA = TileLoader<BM, BK, P, TA>;
  using LB = TileLoader<BN, BK, P, !TB>;

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const float* __restrict__ A = vbg_uptr(vbg_cload(a.A, it.mat));
    const float* __restrict__ B = vbg_uptr(vbg_cload(a.B, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    const KRange kr = vbg_k_range<BK>(d.K, d.ksplit, it.kslice);
    const int kb = kr.kb, ke = kr.ke;
    const int nk = (ke - kb + BK - 1) / BK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    // block j (0, 1) of this wave: rows / columns 32*wave + 16*j of the long side
    v4f acc[2] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
    LA la;
    LB lb;
    la.init(A, d.lda, m0, d.M, d.K);
    lb.init(B, d.ldb, n0, d.N, d.K);
    typename LA::Regs ra;
    typename LB::Regs rb;
    if (nk > 0) {
      la.fetch(ra, kb, ke);
      lb.fetch(rb, kb, ke);
      LA::store(ra, lds);
      LB::store(rb, lds + BM * P);
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
      const float* As = lds + cur * STAGE;
      const float* Bs = As + BM * P;
#pragma unroll
      for (int c = 0; c < KL / 4; ++c) {
        const int ko = g * KL + 4 * c;
        // the short side's fragment is shared by both blocks
        const v4f sv = WIDE ? *(const v4f*)(As + l16 * P + ko) : *(const v4f*)(Bs + l16 * P + ko);
        v4f lv[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int r = 32 * wave + 16 * j + l16;
          lv[j] = WIDE ? *(const v4f*)(Bs + r * P + ko) : *(const v4f*)(As + r * P + ko);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[j] = WIDE ? __builtin_amdgcn_mfma_f32_16x16x4f32(sv[e], lv[j][e], acc[j], 0, 0, 0)
                          : __builtin_amdgcn_mfma_f32_16x16x4f32(lv[j][e], sv[e], acc[j], 0, 0, 0);
      }
      if (more) {
        float* nxt = lds + (cur ^ 1) * STAGE;
        LA::store(ra, nxt);
        LB::store(rb, nxt + BM * P);
      }
      __syncthreads();
    }
    // 16x16x4 layout: register r of lane l holds D[4*(l/16) + r][l%16]
    const OutTile o = vbg_out(a, d, it, 1.0f);
    auto seg = [&](int sg, int& row, int& col, v4f& v) {
      row = m0 + (WIDE ? 0 : 32 * wave + 16 * sg) + 4 * g;
      col = n0 + (WIDE ? 32 * wave + 16 * sg : 0) + l16;
      v = acc[sg];
    };
    vbg_epilogue<2>(o, d.M, d.N, seg);
  }
};

// Reference-compatible VALU tile (32x32 outputs, 4 per thread).  Each output
// is one fmaf chain over k = 0, 1, ..., K-1 in order.  With alpha = beta = 1
// the chain is seeded from C, which is exactly the reference's correct tiles
// (ROCm/framework/kernel_128_new.h:160-187, kernel_256_new.h:28-67, fmas :43-51);
// otherwise it starts from 0 and C = alpha*acc + beta*C.
template <bool TA, bool TB>
struct SeqTile {
  static constexpr int BM = 32, BN = 32, BK = 32;
  static constexpr int PA = 36, PB = 33;  // As[k][i] (float4 reads), Bs[k][j]
  static constexpr int LDS_FLOATS = BK * PA + BK * PB;

  __device__ static float ga(const float* A, int lda, int i, int k) {
    return TA ? A[(size_t)i * lda + k] : A[i + (size_t)k * lda];
  }
  __device__ static float gb(const float* B, int ldb, int k, int j) {
    return TB ? B[j + (size_t)k * ldb] : B[(size_t)j * ldb + k];
  }

  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    const vbg_mat_desc d = vbg_udesc(vbg_cload(a.desc, it.mat));
    const float* A = vbg_uptr(vbg_cload(a.A, it.mat));
    const float* B = vbg_uptr(vbg_cload(a.B, it.mat));
    float* C = vbg_uptr(vbg_cload(a.C, it.mat));
    const int m0 = it.tm * BM, n0 = it.tn * BN;
    float* As = lds;
    float* Bs = lds + BK * PA;
    const int t = threadIdx.x;
    const int jj = t >> 3, ii = 4 * (t & 7);
    const int col = n0 + jj, row = m0 + ii;
    const bool seeded = (a.alpha == 1.0f && a.beta == 1.0f);
    float c[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      c[e] = 0.0f;
      if (seeded && col < d.N && row + e < d.M) c[e] = C[(size_t)col * d.ldc + row + e];
    }
    for (int k0 = 0; k0 < d.K; k0 += BK) {
      const int kn = min(BK, d.K - k0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int idx = t + VBG_THREADS * v;
        {  // A: i fastest for coalescing when op(A) = A
          const int i = idx & 31, k = idx >> 5;
          const int gi = m0 + i, gk = k0 + k;
          As[k * PA + i] = (gi < d.M && gk < d.K) ? ga(A, d.lda, gi, gk) : 0.0f;
        }
        {  // B: k fastest for coalescing when op(B) = B
          const int k = idx & 31, j = idx >> 5;
          const int gk = k0 + k, gj = n0 + j;
          Bs[k * PB + j] = (gj < d.N && gk < d.K) ? gb(B, d.ldb, gk, gj) : 0.0f;
        }
      }
      __syncthreads();
      for (int k = 0; k < kn; ++k) {
        const v4f av = *(const v4f*)(As + k * PA + ii);
        const float bv = Bs[k * PB + jj];
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = fmaf(av[e], bv, c[e]);
      }
      __syncthreads();
    }
    if (col < d.N) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (row + e < d.M) {
          float* p = C + (size_t)col * d.ldc + row + e;
          if (seeded) {
            *p = c[e];
          } else {
            float o = a.alpha * c[e];
            if (a.beta != 0.0f) o += a.beta * *p;
            *p = o;
          }
        }
      }
    }
  }
};

// Register-allocation target: tiles may name a minimum number of waves per
// SIMD (amdgpu_waves_per_eu); 1 leaves the compiler free.
template <class T, class = void>
struct min_waves {
  static constexpr int value = 1;
};
template <class T>
struct min_waves<T, decltype((void)T::MIN_WAVES)> {
  static constexpr int value = T::MIN_WAVES;
};

// LOOK: the queue walk of launches with few items per workgroup (look_all).
// It is a separate instantiation because a walk chosen at run time changes
// the code the compiler makes for the FP32 bodies (4% slower on large
// launches that never take it).
template <class Tile, bool LOOK>
__global__ void __launch_bounds__(VBG_THREADS) __attribute__((amdgpu_waves_per_eu(min_waves<Tile>::value)))
vbg_gemm_kernel(vbg_args a) {
  __shared__ __attribute__((aligned(16))) float lds[Tile::LDS_FLOATS];
  vbg_persistent_loop<LOOK ? 1 : 0>(a, [&](const vbg_item& it) { Tile::run(a, it, lds); });
}

// Mixed launch: one persistent grid over items of the 64x64 .. 128x128
// shapes, each item naming its own (wave-uniform) shape.  LDS is sized for
// the largest; registers are those of the largest body.
template <bool TA, bool TB>
struct Fp32Mixed {
  using Sq128 = Fp32MfmaTile<128, 128, 32, TA, TB, false>;
  using Tall = Fp32MfmaTile<128, 64, 32, TA, TB, false>;
  using Wide = Fp32MfmaTile<64, 128, 32, TA, TB, false>;
  using Sq64 = Fp32MfmaTile<64, 64, 32, TA, TB, false>;
  using SkinnyM = Fp32SkinnyTile<true, TA, TB>;   // 16 x 128
  using SkinnyN = Fp32SkinnyTile<false, TA, TB>;  // 128 x 16
  static constexpr int LDS_FLOATS = SkinnyM::LDS_FLOATS > Sq128::LDS_FLOATS ? SkinnyM::LDS_FLOATS
                                                                              : Sq128::LDS_FLOATS;
  static constexpr int MIN_WAVES = 2;  // two workgroups per CU, as the 128x128 body alone
  static_assert(LDS_FLOATS >= Tall::LDS_FLOATS && LDS_FLOATS >= Sq64::LDS_FLOATS, "LDS sized for the largest body");
  __device__ static void run(const vbg_args& a, const vbg_item& it, float* lds) {
    switch (it.tile) {
      case T128: Sq128::run(a, it, lds); break;
      case T128x64: Tall::run(a, it, lds); break;
      case T64x128: Wide::run(a, it, lds); break;
      case T16x128: SkinnyM::run(a, it, lds); break;
      case T128x16: SkinnyN::run(a, it, lds); break;
      default: Sq64::run(a, it, lds); break;
    }
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Host-side launch table.
#include "vbg_launch.h"

template <class Tile>
static hipError_t launch_one(const vbg_args& a, int grid, hipStream_t s) {
  if (a.look_all)
    hipLaunchKernelGGL((vbg_gemm_kernel<Tile, true>), dim3(grid), dim3(VBG_THREADS), 0, s, a);
  else
    hipLaunchKernelGGL((vbg_gemm_kernel<Tile, false>), dim3(grid), dim3(VBG_THREADS), 0, s, a);
  return hipGetLastError();
}

template <class Tile>
static int occupancy_one() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, vbg_gemm_kernel<Tile, false>, VBG_THREADS, 0) !=
      hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

template <class Tile>
static int run_or_occ(const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (occ) {
    *occ = occupancy_one<Tile>();
    return 0;
  }
  return launch_one<Tile>(*a, grid, s) == hipSuccess ? 0 : -2;
}

template <template <bool, bool> class TileTT>
static int dispatch_tt(bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  if (!ta && !tb) return run_or_occ<TileTT<false, false>>(a, grid, s, occ);
  if (!ta && tb) return run_or_occ<TileTT<false, true>>(a, grid, s, occ);
  if (ta && !tb) return run_or_occ<TileTT<true, false>>(a, grid, s, occ);
  return run_or_occ<TileTT<true, true>>(a, grid, s, occ);
}

template <bool TA, bool TB> using F32T16 = Fp32SmallTile<16, TA, TB>;
template <bool TA, bool TB> using F32T32 = Fp32SmallTile<32, TA, TB>;
template <bool TA, bool TB> using F32T64 = Fp32MfmaTile<64, 64, 32, TA, TB>;
template <bool TA, bool TB> using F32T128x64 = Fp32MfmaTile<128, 64, 32, TA, TB>;
template <bool TA, bool TB> using F32T64x128 = Fp32MfmaTile<64, 128, 32, TA, TB>;
template <bool TA, bool TB> using F32T128 = Fp32MfmaTile<128, 128, 32, TA, TB>;
template <bool TA, bool TB> using F32Mixed = Fp32Mixed<TA, TB>;
template <bool TA, bool TB> using F32S16x128 = Fp32SkinnyTile<true, TA, TB>;
template <bool TA, bool TB> using F32S128x16 = Fp32SkinnyTile<false, TA, TB>;

static int fp32_dispatch(int tile, bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s,
                         int* occ) {
  switch (tile) {
    case T16: return dispatch_tt<F32T16>(ta, tb, a, grid, s, occ);
    case T32: return dispatch_tt<F32T32>(ta, tb, a, grid, s, occ);
    case T64: return dispatch_tt<F32T64>(ta, tb, a, grid, s, occ);
    case T128x64: return dispatch_tt<F32T128x64>(ta, tb, a, grid, s, occ);
    case T64x128: return dispatch_tt<F32T64x128>(ta, tb, a, grid, s, occ);
    case T128: return dispatch_tt<F32T128>(ta, tb, a, grid, s, occ);
    case T16x128: return dispatch_tt<F32S16x128>(ta, tb, a, grid, s, occ);
    case T128x16: return dispatch_tt<F32S128x16>(ta, tb, a, grid, s, occ);
    case T_MIXED: return dispatch_tt<F32Mixed>(ta, tb, a, grid, s, occ);
    default: return -1;
  }
}

static int seq_dispatch(bool ta, bool tb, const vbg_args* a, int grid, hipStream_t s, int* occ) {
  return dispatch_tt<SeqTile>(ta, tb, a, grid, s, occ);
}

int vbg_launch_fp32(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream) {
  return fp32_dispatch(tile, ta != 0, tb != 0, a, grid, (hipStream_t)stream, nullptr);
}
int vbg_occupancy_fp32(int tile, int ta, int tb) {
  int occ = 1;
  fp32_dispatch(tile, ta != 0, tb != 0, nullptr, 0, nullptr, &occ);
  return occ;
}
int vbg_launch_seq(int ta, int tb, const vbg_args* a, int grid, void* stream) {
  return seq_dispatch(ta != 0, tb != 0, a, grid, (hipStream_t)stream, nullptr);
}
int vbg_occupancy_seq(int ta, int tb) {
  int occ = 1;
  seq_dispatch(ta != 0, tb != 0, nullptr, 0, nullptr, &occ);
  return occ;
}
