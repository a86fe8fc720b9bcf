// vbg_planner.cpp -- host planner (no HIP calls).
//
// Replaces the reference planner (ROCm/framework/gemm.cuh:121-229) whose
// output is a tile index per matrix plus a bounding-box grid in which ~98% of
// the workgroups exit at once (grid (32,16,256) with 2,498 useful blocks on
// the shipped dataset at B=256).  Here the planner emits an explicit work map:
//   1. a tile shape per matrix from a cost model (or one of the reference /
//      Wang / Li policies, restated for comparison and autotuning);
//   2. one work item per (matrix, tile_m, tile_n) -- never an empty slot;
//   3. matrices dealt to 8 per-XCD queues by longest-processing-time-first,
//      each queue ordered by decreasing per-item cost, so that a persistent
//      workgroup pops long items first and the tiles of one matrix share the
//      L2 of one XCD.
#include "vbg_planner.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <queue>

#include "../../include/vbgemm.h"

#ifndef VBG_BAND_SPLIT
#define VBG_BAND_SPLIT 0
#endif
#ifndef VBG_PARTIAL_MIX
#define VBG_PARTIAL_MIX 1
#endif

namespace vbg {

namespace {

// k-tile depth of each tile id (matches the kernels).
const int kTileBK[NTILES] = {64, 32, 32, 32, 32, 32, 16, 64, 64, 64};

// Relative MFMA efficiency of each tile shape on gfx950 (per-flop cost factor
// = 1/eff).  Larger tiles reuse each LDS byte and each HBM byte more times.
// 256x128 (bf16x3 only): 1.10, from 1024^3 bf16x3 batches at 311 vs 283
// TFLOP/s against 128x128 (two k-tiles of 16 against one of 32 per barrier,
// 25% fewer operand bytes per flop).  The skinny 16x128 / 128x16 tiles
// (FP32 only): 0.35, between the 16x16 and 32x32 tiles -- they stream the
// long operand at an eighth of a 128x128 tile's MFMA work per byte.  256x256
// (BF16 only): 1.40 -- half the operand bytes per flop of 128x128, the bytes
// the LDS-DMA address path can deliver per MFMA cycle being what bounds the
// BF16 tiles.
const double kTileEff[NTILES] = {0.20, 0.40, 0.70, 0.85, 0.85, 1.00, 1.10, 0.35, 0.35, 1.40};
// Fixed cost of one work item in units of a 128x128x32 k-step (queue pop,
// pipeline fill, epilogue).
const double kItemFixedDefault = 3.0;

// The cost model in force for the plan being built (the constants above, or
// the values a plan's options set: the autotuner searches them).  Set at the
// start of each planning call; thread-local, so plans built on different
// threads do not interfere.
struct CostModel {
  double eff[NTILES];
  double fixed;
  bool packed;  // split precisions: operands packed (the 256x256 tile needs them)
};
thread_local CostModel g_cm;
void set_cost_model(const vbg_opts& o) {
  const double small = o.small_tile_eff_pct > 0 ? o.small_tile_eff_pct / 100.0 : 1.0;
  for (int t = 0; t < NTILES; ++t)
    g_cm.eff[t] = kTileEff[t] * (kTileBM[t] * kTileBN[t] < 128 * 128 ? small : 1.0);
  g_cm.fixed = o.item_fixed_x10 > 0 ? o.item_fixed_x10 / 10.0 : kItemFixedDefault;
  g_cm.packed = false;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

struct Shape {
  int M, N, K;
};

double item_cost(int tile, const Shape& s) {
  const int bm = kTileBM[tile], bn = kTileBN[tile], bk = kTileBK[tile];
  const double ksteps = (double)cdiv(std::max(s.K, 1), bk);
  const double unit = 128.0 * 128.0 * 32.0;
  return g_cm.fixed + ksteps * (double)bm * bn * bk / unit / g_cm.eff[tile];
}

int num_tiles(int tile, const Shape& s) {
  return cdiv(s.M, kTileBM[tile]) * cdiv(s.N, kTileBN[tile]);
}

double matrix_cost(int tile, const Shape& s) { return num_tiles(tile, s) * item_cost(tile, s); }

// Modelled duration of a persistent launch: its items (cost, count) on
// `slots` workgroup slots.  The workgroups pop cost-ordered queues, which is
// greedy list scheduling longest-first; while a launch is only a few rounds
// long that makespan is set by whole rounds (768 equal items on 256 slots take
// 3 items' time, 769 take 4), so it is simulated.  Longer launches are
// modelled as max(total / slots, longest item).
struct CostCount {
  double cost;
  int64_t count;
};
double launch_makespan(std::vector<CostCount> items, int slots) {
  const int kMaxRounds = 16;
  double total = 0.0, longest = 0.0;
  int64_t n = 0;
  for (const CostCount& x : items) {
    if (x.count <= 0) continue;
    total += x.cost * (double)x.count;
    longest = std::max(longest, x.cost);
    n += x.count;
  }
  if (n == 0 || slots <= 0) return 0.0;
  if (n > (int64_t)kMaxRounds * slots) return std::max(total / slots, longest);
  std::sort(items.begin(), items.end(), [](const CostCount& a, const CostCount& b) { return a.cost > b.cost; });
  std::priority_queue<double, std::vector<double>, std::greater<double>> load;
  for (int i = 0; i < slots; ++i) load.push(0.0);
  double span = 0.0;
  for (const CostCount& x : items)
    for (int64_t j = 0; j < x.count; ++j) {
      const double l = load.top() + x.cost;
      load.pop();
      load.push(l);
      span = std::max(span, l);
    }
  return span;
}

bool is_split(int prec) {
  return prec == VBG_PREC_BF16X3 || prec == VBG_PREC_FP16_SCALED || prec == VBG_PREC_FP8_E4M3 ||
         prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16;
}

// The split-precision and BF16 families are built on 32x32 MFMA blocks only
// (>= 64x64); the 256x128 tile for bf16x3 only.  AUTO picks it per matrix by
// the cost model like any other shape (unbalanced B=2048 bf16x3 batch 3.567
// vs 3.644-3.684 ms on 128x128 alone, the matrices it leaves to smaller
// shapes running as a second, short launch; batching workload 97.7 vs 95.9
// TFLOP/s).  A combined kernel holding the 256x128 body with the others
// spills registers, so the shape keeps a launch of its own.
bool tile_allowed(int prec, int tile, unsigned mask) {
  if (mask && !(mask & (1u << tile))) return false;
  if (is_split(prec) && tile < T64) return false;
  const bool bf16 = prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16;
  // (packed FP8's 256x256 body, two sets of 32-byte f8f6f4 operands beside 256
  // accumulators, does not fit the register file)
  const bool pk_dma = g_cm.packed && prec == VBG_PREC_FP16_SCALED;
  if (tile == T256x128 && prec != VBG_PREC_BF16X3 && !bf16 && !pk_dma) return false;
  if (tile == T256 && !bf16 && !pk_dma) return false;
  if ((tile == T16x128 || tile == T128x16) && prec != VBG_PREC_FP32) return false;
  return true;
}

// Shapes a mixed launch holds (one kernel dispatching on each item's shape):
// 64x64 .. 128x128, and the skinny ones (FP32's mixed kernel holds them).
bool mixable(int tile) { return (tile >= T64 && tile <= T128) || tile == T16x128 || tile == T128x16; }

// AUTO: per-matrix minimum modelled cost, then, while the batch exposes fewer
// work items than resident workgroup slots, split the matrix whose single
// item is longest into the next smaller tile (more, shorter items).
void choose_auto(const std::vector<Shape>& sh, int prec, const vbg_opts& o, int slots,
                 std::vector<int>& tile) {
  const int B = (int)sh.size();
  const int big_min = o.big_tile_min_mn > 0 ? o.big_tile_min_mn : 0;
  for (int b = 0; b < B; ++b) {
    int best = -1;
    double bc = 0;
    for (int t = 0; t < NTILES; ++t) {
      if (!tile_allowed(prec, t, o.tile_mask)) continue;
      if (t == T128 && big_min && std::min(sh[b].M, sh[b].N) < big_min) continue;
      // packed FP16 runs 256x128 rather than 256x256 (two workgroups per CU
      // instead of one): unbalanced B=2048 batch 3.32 vs 3.43 ms, B=512 0.890
      // vs 0.898 ms; 64 x 2048^3 2.584 vs 2.557-2.562 ms (256x256 stays a
      // fixed-policy choice)
      if (t == T256 && prec == VBG_PREC_FP16_SCALED) continue;
      const double c = matrix_cost(t, sh[b]);
      if (best < 0 || c < bc - 1e-9) {
        best = t;
        bc = c;
      }
    }
    tile[b] = best < 0 ? T64 : best;
  }
  // Undersubscription repair.  It stops at 64x64: on small batches the
  // 32x32 / 16x16 kernels' per-item latency outweighs the extra parallelism
  // (measured on the reference's dataset at B = 8 / 16: 64x64 items run the
  // batch in 29 / 34 us, 32x32 items in 51 / 70 us).  Matrices with K >=
  // kRepairMaxK keep their tile when split-K is available: split-K then cuts
  // their long items instead, at the larger tile's LDS and MFMA efficiency
  // (measured on the grouped backward's dB pass, one 1408 x 2048 x 32636
  // matrix: 64x64 tiles 634 us; see the planner notes on split-K).
  const int kRepairMaxK = 2048;
  const bool split_ok = prec != VBG_PREC_FP32_SEQ && o.split_k_min >= 0;
  int64_t items = 0;
  for (int b = 0; b < B; ++b) items += num_tiles(tile[b], sh[b]);
  const int smaller[NTILES] = {-1, -1, -1, T64, T64, T128x64, T128, T16, T16, T256x128};
  for (int iter = 0; iter < 4 * B && items < slots; ++iter) {
    int pick = -1;
    double longest = 0;
    for (int b = 0; b < B; ++b) {
      if (split_ok && sh[b].K >= kRepairMaxK) continue;
      int s = smaller[tile[b]];
      while (s >= 0 && !tile_allowed(prec, s, o.tile_mask)) s = smaller[s];
      if (s < 0) continue;
      const double c = item_cost(tile[b], sh[b]);
      if (c > longest) {
        longest = c;
        pick = b;
      }
    }
    if (pick < 0) break;
    int s = smaller[tile[pick]];
    while (s >= 0 && !tile_allowed(prec, s, o.tile_mask)) s = smaller[s];
    items += num_tiles(s, sh[pick]) - num_tiles(tile[pick], sh[pick]);
    tile[pick] = s;
  }
}

// Reference tile index (gemm.cuh:121-134, M-tile x N-tile) -> this library's
// smallest tile covering it.
const int kRefToTile[5] = {T16, T32, T32, T64, T64};

// Wang (ROCm/tiling/gemm_wang.cuh:104-182): one global square tile from
// {16, 32, 64} -- the first whose TLP = sum (M/t)(N/t)*128 falls below
// tlp_thres -- then per matrix the largest dividing square tile not above it,
// then demotion from the batch end while the workgroup count is below
// tlp_thres (the reference's second threshold; one threshold is used here).
std::vector<int> wang_order(const std::vector<Shape>& sh) {
  std::vector<int> ord(sh.size());
  std::iota(ord.begin(), ord.end(), 0);
  auto cls = [&](int j) {
    const int m = sh[j].M, n = sh[j].N;
    if (m % 64 == 0 && n % 64 == 0) return 0;
    if (m % 32 == 0 && n % 32 == 0) return 1;
    return 2;
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sh[x].K > sh[y].K; });
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cls(x) < cls(y); });
  return ord;
}

void choose_wang(const std::vector<Shape>& sh, int thres, std::vector<int>& tile) {
  const int B = (int)sh.size();
  const int ts[3] = {16, 32, 64};
  int t;
  for (t = 0; t < 3; ++t) {
    int64_t tlp = 0;
    for (int j = 0; j < B; ++j) tlp += (int64_t)(sh[j].M / ts[t]) * (sh[j].N / ts[t]) * 128;
    if (tlp < thres) break;
  }
  if (t >= 3) t = 2;
  std::vector<int> st(B, 0);
  for (int j = 0; j < B; ++j) {
    auto fits = [&](int k) {
      return ts[k] <= sh[j].M && ts[k] <= sh[j].N && sh[j].M % ts[k] == 0 && sh[j].N % ts[k] == 0;
    };
    if (fits(t)) {
      st[j] = t;
    } else {
      for (int k = 0; k < t; ++k)
        if (fits(k)) st[j] = k;
    }
  }
  int64_t wg = 0;
  for (int j = 0; j < B; ++j) wg += (int64_t)cdiv(sh[j].M, ts[st[j]]) * cdiv(sh[j].N, ts[st[j]]);
  // The reference plans the batch in the order of ROCm/data/sort.cpp:17-60
  // (gemm_wang.cuh:33 reads the sorted file): stable by K descending, then
  // stable by tile class -- M and N both multiples of 64, then of 32, then
  // the rest; the demotion pass walks that order from its end.
  const std::vector<int> ord = wang_order(sh);
  for (int jj = B - 1; jj >= 0 && wg < thres; --jj) {
    const int j = ord[jj];
    if (st[j] == 0) continue;
    wg += (int64_t)cdiv(sh[j].M, ts[st[j]]) * cdiv(sh[j].N, ts[st[j]]) * 3;
    st[j] -= 1;
  }
  const int map[3] = {T16, T32, T64};
  for (int j = 0; j < B; ++j) tile[j] = map[st[j]];
}

// Li (CUDA/batching/gemm.cu:99-127): global tile over the six-tile set, the
// first whose TLP = sum (M/tm)(N/tn)*256 falls below tlp_thres; per matrix that
// tile if it fits (tm <= M, tn <= N), else the last smaller tile with
// tm == M and tn <= N, else 16x16.
void choose_li(const std::vector<Shape>& sh, int thres, std::vector<int>& tile) {
  const int B = (int)sh.size();
  const int kLiTiles = T128 + 1;  // Li's six tiles, 16x16 .. 128x128
  int t;
  for (t = 0; t < kLiTiles; ++t) {
    int64_t tlp = 0;
    for (int j = 0; j < B; ++j)
      tlp += (int64_t)(sh[j].M / kTileBM[t]) * (sh[j].N / kTileBN[t]) * 256;
    if (tlp < thres) break;
  }
  if (t == kLiTiles) t = kLiTiles - 1;
  for (int j = 0; j < B; ++j) {
    int s = T16;
    if (kTileBM[t] <= sh[j].M && kTileBN[t] <= sh[j].N) {
      s = t;
    } else {
      for (int k = 0; k < t; ++k)
        if (kTileBM[k] == sh[j].M && kTileBN[k] <= sh[j].N) s = k;
    }
    tile[j] = s;
  }
}

}  // namespace

int ref_plan(int batch, const int* M, const int* N, const int* K, int thres, int* t_strategy,
             int* kernel_thread, int* grid, int64_t* wavefronts) {
  (void)K;
  static const int tsz[5][2] = {{16, 16}, {16, 32}, {32, 32}, {32, 64}, {64, 64}};
  auto threads = [](int t) { return t <= 2 ? 128 : 256; };
  if (batch <= 0 || !M || !N || !t_strategy || !kernel_thread || !grid) return VBG_ERR_INVALID_ARG;
  for (int i = 0; i < batch; ++i) {
    int c = -1;
    for (int j = 4; j >= 0 && c < 0; --j)
      if (tsz[j][0] <= M[i] && tsz[j][1] <= N[i] && M[i] % tsz[j][0] == 0 && N[i] % tsz[j][1] == 0)
        c = j;
    if (c < 0) return VBG_ERR_UNDEFINED;
    t_strategy[i] = c;
  }
  int total = 0;
  for (int i = 0; i < batch; ++i) {
    const int t = t_strategy[i];
    kernel_thread[i] = threads(t);
    total += M[i] * N[i] / (tsz[t][0] * tsz[t][1]) * (kernel_thread[i] / 64);
  }
  if (wavefronts) wavefronts[0] = total;
  for (int j = 4; j > 0 && total <= thres; --j) {
    for (int i = 0; i < batch && total <= thres; ++i) {
      if (t_strategy[i] != j) continue;
      const int t = t_strategy[i];
      // gemm.cuh:180 reads the thread count at index t (a tile number).
      if (t >= batch) return VBG_ERR_UNDEFINED;
      const int old_v = M[i] * N[i] / (tsz[t][0] * tsz[t][1]) * (kernel_thread[t] / 64);
      const int nt = t > 0 ? t - 1 : 0;
      t_strategy[i] = nt;
      kernel_thread[i] = threads(nt);
      const int new_v = M[i] * N[i] / (tsz[nt][0] * tsz[nt][1]) * (kernel_thread[i] / 64);
      total += new_v - old_v;
    }
  }
  if (wavefronts) wavefronts[1] = total;
  int gx = 0, gy = 0;
  for (int i = 0; i < batch; ++i) {
    gx = std::max(gx, M[i] / tsz[t_strategy[i]][0]);
    gy = std::max(gy, N[i] / tsz[t_strategy[i]][1]);
  }
  grid[0] = gx;
  grid[1] = gy;
  grid[2] = batch;
  return VBG_OK;
}

int build_plan(HostPlan& hp, int batch, const int* M, const int* N, const int* K, const int* lda,
               const int* ldb, const int* ldc, bool ta, bool tb, int prec, const vbg_opts& o,
               int num_cus, int (*occupancy)(int, int, int, bool, bool), std::string& err,
               int64_t ws_cap_floats) {
  hp = HostPlan();
  set_cost_model(o);
  hp.batch = batch;
  hp.ta = ta;
  hp.tb = tb;
  hp.prec = prec;
  if (batch < 0) {
    err = "batch < 0";
    return VBG_ERR_INVALID_ARG;
  }
  if (batch > 0 && (!M || !N || !K || !lda || !ldb || !ldc)) {
    err = "null dimension array";
    return VBG_ERR_INVALID_ARG;
  }
  if (prec < VBG_PREC_FP32 || prec > VBG_PREC_BF16_OUT_BF16) {
    err = "unknown precision";
    return VBG_ERR_INVALID_ARG;
  }
  hp.desc.resize(batch);
  std::vector<Shape> sh(batch);
  for (int b = 0; b < batch; ++b) {
    if (M[b] < 0 || N[b] < 0 || K[b] < 0) {
      err = "negative dimension at matrix " + std::to_string(b);
      return VBG_ERR_INVALID_ARG;
    }
    const int rows_a = ta ? K[b] : M[b];
    const int rows_b = tb ? N[b] : K[b];
    if (lda[b] < std::max(1, rows_a) || ldb[b] < std::max(1, rows_b) || ldc[b] < std::max(1, M[b])) {
      err = "leading dimension too small at matrix " + std::to_string(b);
      return VBG_ERR_INVALID_ARG;
    }
    if ((prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16) && (lda[b] % 8 || ldb[b] % 8)) {
      err = "BF16 operands need lda and ldb multiples of 8 (matrix " + std::to_string(b) + ")";
      return VBG_ERR_NOT_SUPPORTED;
    }
    // The kernels address every operand with 32-bit byte offsets from its
    // base (buffer resources), so each stored matrix must span < 2 GiB.
    {
      const int64_t eb = (prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16) ? 2 : 4;
      const int64_t lim = (int64_t)1 << 31;
      const int64_t span_a = (int64_t)lda[b] * (ta ? M[b] : K[b]) * eb;
      const int64_t span_b = (int64_t)ldb[b] * (tb ? K[b] : N[b]) * eb;
      const int64_t span_c = (int64_t)ldc[b] * N[b] * 4;
      if (span_a >= lim || span_b >= lim || span_c >= lim) {
        err = "matrix " + std::to_string(b) + " spans 2 GiB or more (32-bit offsets)";
        return VBG_ERR_NOT_SUPPORTED;
      }
    }
    sh[b] = {M[b], N[b], K[b]};
    hp.desc[b] = {M[b], N[b], K[b], lda[b], ldb[b], ldc[b], 0, 1};
  }

  // without a device (host-only plans): the kernels' workgroups per CU --
  // two, one for the 256x256 tile (128 KB of LDS)
  auto occ = [&](int fam, int tile) {
    int v = occupancy ? occupancy(fam, prec, tile, ta, tb) : (tile == T256 ? 1 : 2);
    return v > 0 ? v : 1;
  };
  const int cus = num_cus > 0 ? num_cus : 256;

  // Packed operands for bf16x3 / FP16 / FP8: one pre-pass quantises every
  // operand into k-contiguous rows (4 / 2 / 1 bytes per element), which the
  // GEMM kernels then stream with no conversion on the HBM -> LDS path.  The
  // pass costs a read and a write of the operands (about 4.3 TB/s), so by
  // default an FP16 / FP8 batch packs when its FP32 arithmetic intensity,
  // 2*sum(MNK) / 4*sum(MK + KN) flop/B, reaches kPackMinAI.  Measured, ms
  // (unpacked / packed incl. the pass): unbalanced B=2048 (AI ~156) FP16
  // 3.85 / 3.57, FP8 3.86 / 3.18; log-uniform 8..1024, B=4096 (AI ~14) FP16
  // 1.25 / 1.55, FP8 1.26 / 1.43.  bf16x3 stays unpacked by default: its
  // GEMM kernel runs no faster on packed operands (3.80 / 3.95 ms on the
  // unbalanced batch; the conversion is not what bounds it), so the pass
  // (1.26 ms) is pure cost.
  const bool packable = prec == VBG_PREC_BF16X3 || prec == VBG_PREC_FP16_SCALED || prec == VBG_PREC_FP8_E4M3;
  bool pack = packable && o.pack_operands > 0;
  // The default also leaves a batch unpacked when its packed operands would
  // exceed kPackMaxBytes of plan-owned memory (asking for packing explicitly
  // has no limit beyond the 2 GiB per operand the kernels address).
  if (packable && o.pack_operands == 0 && prec != VBG_PREC_BF16X3) {
    const double kPackMinAI = 64.0, kPackMaxBytes = 8.0 * (1u << 30);
    const double ebp = prec == VBG_PREC_FP16_SCALED ? 2.0 : 1.0;
    double flops = 0.0, bytes = 0.0, packed = 0.0;
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0 || K[b] == 0) continue;
      flops += 2.0 * M[b] * N[b] * K[b];
      bytes += 4.0 * ((double)M[b] * K[b] + (double)K[b] * N[b]);
      packed += ebp * ((double)M[b] + N[b]) * (double)(cdiv(K[b], VBG_PACK_KALIGN) * VBG_PACK_KALIGN);
    }
    pack = bytes > 0.0 && flops / bytes >= kPackMinAI && packed <= kPackMaxBytes;
  }
  g_cm.packed = pack;

  // 1. tile per matrix
  std::vector<int> tile(batch, T64);
  if (prec == VBG_PREC_FP32_SEQ) {
    std::fill(tile.begin(), tile.end(), T32);
  } else {
    switch (o.policy) {
      case VBG_POLICY_AUTO:
        choose_auto(sh, prec, o, cus * occ(FAM_FP32, T128), tile);
        break;
      case VBG_POLICY_FIXED:
        if (o.fixed_tile < 0 || o.fixed_tile >= NTILES) {
          err = "fixed_tile out of range";
          return VBG_ERR_INVALID_ARG;
        }
        if ((o.fixed_tile == T256x128 || o.fixed_tile == T256) && !tile_allowed(prec, o.fixed_tile, 0)) {
          err = "the 256x128 tile is built for bf16x3, BF16 and packed FP16 operands, 256x256 for BF16 and "
                "packed FP16 operands only";
          return VBG_ERR_NOT_SUPPORTED;
        }
        if ((o.fixed_tile == T16x128 || o.fixed_tile == T128x16) && prec != VBG_PREC_FP32) {
          err = "the 16x128 and 128x16 tiles are built for FP32 only";
          return VBG_ERR_NOT_SUPPORTED;
        }
        std::fill(tile.begin(), tile.end(), o.fixed_tile);
        break;
      case VBG_POLICY_REFERENCE: {
        std::vector<int> t(batch), kt(batch);
        int grid[3];
        int rc = batch > 0 ? ref_plan(batch, M, N, K, o.wavefront_thres, t.data(), kt.data(), grid,
                                      nullptr)
                           : VBG_OK;
        if (rc == VBG_OK) {
          for (int b = 0; b < batch; ++b) tile[b] = kRefToTile[t[b]];
        } else {
          // the reference is undefined on this batch: AUTO instead
          choose_auto(sh, prec, o, cus * occ(FAM_FP32, T128), tile);
        }
        break;
      }
      case VBG_POLICY_WANG:
        choose_wang(sh, o.tlp_thres, tile);
        break;
      case VBG_POLICY_LI:
        choose_li(sh, o.tlp_thres, tile);
        break;
      default:
        err = "unknown policy";
        return VBG_ERR_INVALID_ARG;
    }
  }
  if (is_split(prec))
    for (int b = 0; b < batch; ++b) tile[b] = std::max(tile[b], (int)T64);
  hp.tile_of.assign(batch, -1);
  const int fam = prec == VBG_PREC_FP32_SEQ ? FAM_SEQ
                  : prec == VBG_PREC_FP32   ? FAM_FP32
                  : (prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16) ? FAM_BF16
                  : pack                    ? FAM_SPLIT_PK
                                            : FAM_SPLIT;

  // 1a. launch consolidation (AUTO).  Each tile shape in use is one more
  // persistent launch, and a launch costs a fixed start-up and drain time that
  // the per-matrix cost model does not see.  A launch's modelled duration is
  // max(total cost / slots, longest item) in the cost model's units (one
  // 128x128x32 k-step of one workgroup slot).  The matrices of a shape join
  // the dominant launch (retiled with its shape) when that lengthens it by
  // less than their own launch's duration plus kLaunchFixed, provided their
  // launch is short: a small fraction of the dominant one, or no longer than
  // a launch's fixed cost (tiny batches, where every launch is latency).
  // Measured: bf16x3 on the unbalanced B=2048 batch runs in one launch
  // instead of four (220 -> 223 TFLOP/s); with "small" at 10% of the
  // dominant launch's work rather than 5%, its 256x128 launch also takes the
  // 128x128 launch holding 5.5% of the work (3.62-3.66 -> 3.56-3.60 ms).
  if (o.policy == VBG_POLICY_AUTO && o.consolidate >= 0 && fam != FAM_SEQ && batch > 0) {
    const double kLaunchFixed = 8.0, kAbsorbFrac = 0.10;
    double tot[NTILES] = {}, longest[NTILES] = {};
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0) continue;
      tot[tile[b]] += matrix_cost(tile[b], sh[b]);
      longest[tile[b]] = std::max(longest[tile[b]], item_cost(tile[b], sh[b]));
    }
    auto slots = [&](int t) { return (double)cus * occ(fam, t); };
    // items of each shape as (cost, count) per matrix; a launch's duration
    // is the makespan model above
    std::vector<std::vector<CostCount>> its(NTILES);
    for (int b = 0; b < batch; ++b)
      if (M[b] > 0 && N[b] > 0) its[tile[b]].push_back({item_cost(tile[b], sh[b]), num_tiles(tile[b], sh[b])});
    double durs[NTILES] = {};
    for (int t = 0; t < NTILES; ++t) durs[t] = launch_makespan(its[t], (int)slots(t));
    auto dur = [&](int t) { return durs[t]; };
    int D = 0;
    for (int t = 1; t < NTILES; ++t)
      if (dur(t) > dur(D)) D = t;
    std::vector<int> order;
    for (int t = 0; t < NTILES; ++t)
      if (t != D && tot[t] > 0) order.push_back(t);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return dur(x) < dur(y); });
    if (tile_allowed(prec, D, o.tile_mask)) {
      for (int t : order) {
        const bool short_launch = tot[t] / slots(t) <= kAbsorbFrac * (tot[D] / slots(D)) ||
                                  dur(t) <= kLaunchFixed;
        if (!short_launch) continue;
        double add = 0.0, add_longest = 0.0;
        std::vector<CostCount> joined = its[D];
        for (int b = 0; b < batch; ++b)
          if (tile[b] == t && M[b] > 0 && N[b] > 0) {
            add += matrix_cost(D, sh[b]);
            add_longest = std::max(add_longest, item_cost(D, sh[b]));
            joined.push_back({item_cost(D, sh[b]), num_tiles(D, sh[b])});
          }
        const double before = dur(D);
        const double after = launch_makespan(joined, (int)slots(D));
        if (after - before < dur(t) + kLaunchFixed) {
          for (int b = 0; b < batch; ++b)
            if (tile[b] == t) tile[b] = D;
          tot[D] += add;
          longest[D] = std::max(longest[D], add_longest);
          its[D].swap(joined);
          durs[D] = after;
          tot[t] = 0.0;
          longest[t] = 0.0;
          its[t].clear();
          durs[t] = 0.0;
        }
      }
    }
  }

  // 1b. split-K.  An item longer than the mean load of a workgroup slot bounds
  // the launch's makespan from below; when a matrix's items exceed 1.25x that
  // mean, its K is cut into S slices (whole k-tiles each) of about half the
  // mean, whose FP32 partial panels go to the workspace and are summed by the
  // split-K reduction.  Never for FP32_SEQ (one sequential chain per element
  // by definition).
  hp.ws_off.assign(batch, -1);
  if (prec != VBG_PREC_FP32_SEQ && o.split_k_min >= 0) {
    const int kmin = o.split_k_min > 0 ? o.split_k_min : 256;
    double total = 0.0, tot_t[NTILES] = {};
    for (int b = 0; b < batch; ++b)
      if (M[b] > 0 && N[b] > 0) {
        total += matrix_cost(tile[b], sh[b]);
        tot_t[tile[b]] += matrix_cost(tile[b], sh[b]);
      }
    // slots of the shape carrying the most work
    const int Dt = (int)(std::max_element(tot_t, tot_t + NTILES) - tot_t);
    const double avg = total / ((double)cus * occ(fam, Dt));
    const int64_t ws_cap = ws_cap_floats;
    int64_t ws = 0;
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0) continue;
      const int ks = cdiv(K[b], VBG_KSLICE_ALIGN);
      int S = 1;
      if (o.force_ksplit > 0) {
        S = std::min(o.force_ksplit, ks);
      } else if (K[b] >= kmin) {
        const double ic = item_cost(tile[b], sh[b]);
        if (ic > 1.25 * avg) S = std::min<int>({(int)std::ceil(2.0 * ic / avg), ks / 2, 16});
        if (S >= 2) {
          // The slice count that makes the launch of this matrix's shape
          // shortest in whole rounds (the rule above can leave a last round
          // nearly empty: one 1408 x 2048 x 32636 matrix on 256x256 tiles cut
          // into 11 slices is 528 items on 256 slots, 16 slices 768): the
          // other matrices of the shape as they stand, this one cut into
          // S' = 2 .. the rule's cap.
          const int cap = std::min(16, std::max(2, ks / 2));
          std::vector<CostCount> others;
          for (int c = 0; c < batch; ++c)
            if (c != b && tile[c] == tile[b] && M[c] > 0 && N[c] > 0) {
              const int Sc = hp.desc[c].ksplit;
              others.push_back({g_cm.fixed + (item_cost(tile[c], sh[c]) - g_cm.fixed) / Sc,
                                (int64_t)num_tiles(tile[c], sh[c]) * Sc});
            }
          const int nslots = cus * occ(fam, tile[b]);
          double best = 0.0;
          int bestS = S;
          for (int s2 = 2; s2 <= cap; ++s2) {
            std::vector<CostCount> its = others;
            its.push_back({g_cm.fixed + (ic - g_cm.fixed) / s2, (int64_t)num_tiles(tile[b], sh[b]) * s2});
            const double m = launch_makespan(its, nslots);
            if (s2 == 2 || m < best - 1e-9) {
              best = m;
              bestS = s2;
            }
          }
          S = bestS;
        }
      }
      const int64_t need = (int64_t)S * M[b] * N[b];
      if (S >= 2 && ws + need <= ws_cap) {
        hp.desc[b].ksplit = S;
        hp.ws_off[b] = ws;
        ws += need;
      }
    }
    hp.ws_floats = ws;
  }
  auto slice_cost = [&](int t, int b) {
    const int S = hp.desc[b].ksplit;
    return g_cm.fixed + (item_cost(t, sh[b]) - g_cm.fixed) / S;
  };

  // 2. group matrices by kernel family/tile; 3. LPT over XCD queues.
  hp.nq = o.xcd_queues ? VBG_NUM_XCD : 1;
  // 2a. parts: the items of one matrix that run on one tile shape.  With
  // edge tiles on, a matrix on a tile with a 128-wide side whose last tile row
  // (or column) holds at most 64 rows (columns) covers that strip with
  // 64-wide tiles instead, in the launch of that shape: about half the padded
  // MFMA work of a random-size edge goes away.  Item coordinates are in units
  // of the item's own tile (128*q = 64*(2q)).
  struct Part {
    int b;
    int tm0, tm1, tn0, tn1;  // tile ranges in the part's own tile units
  };
  std::vector<std::vector<Part>> parts(NTILES);
  // Mixed launch: the parts of every shape from 64x64 to 128x128 share one
  // persistent launch, so edge strips and medium matrices fill the big tiles'
  // queues instead of running as launches of their own after them.  FP32 runs
  // mixed always: its mixed kernel runs 128x128 items as fast as the one-shape
  // kernel (256 x 512x512x2048: 134 TFLOP/s either way), so one launch saves
  // the second launch's start and tail (unbalanced B=2048 batch: 7.30-7.32 ms
  // against 7.36-7.38 with the edges and medium matrices in a launch of their
  // own after the 128x128 one).  Default (0): mixed for FP32, FP16 and FP8 and
  // for small batches; edge tiles where the launches are mixed and for the
  // FP32 and BF16 families.
  //   A batch whose launches would all be short (modelled total under
  // kSmallBatch units: every launch is start-up latency) also runs mixed: one
  // launch instead of one per shape.
  const bool split_scaled = prec == VBG_PREC_FP16_SCALED || prec == VBG_PREC_FP8_E4M3;
  bool small_batch = false;
  {
    const double kSmallBatch = 32.0;
    double tot[NTILES] = {}, longest[NTILES] = {};
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0) continue;
      tot[tile[b]] += matrix_cost(tile[b], sh[b]);
      longest[tile[b]] = std::max(longest[tile[b]], item_cost(tile[b], sh[b]));
    }
    double total = 0.0;
    for (int t = 0; t < NTILES; ++t)
      if (tot[t] > 0) total += std::max(tot[t] / ((double)cus * occ(fam, t)), longest[t]);
    small_batch = total < kSmallBatch;
  }
  //   Otherwise the plan runs mixed when more than kMixFrac of the modelled
  // cost of the 64x64..128x128 parts lies outside the largest shape: those
  // launches would run serially after the dominant one.  Measured on the
  // batching workload (log-uniform sizes, about 25% outside): mixed is +13%
  // for bf16x3; on the unbalanced B=2048 batch (about 4% outside) separate
  // launches stay faster for bf16x3.
  const bool mix_family = fam == FAM_FP32 || fam == FAM_SPLIT || fam == FAM_SPLIT_PK;
  auto build_parts = [&](bool edges) {
    for (auto& v : parts) v.clear();
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0) continue;
      const int t = tile[b];
      hp.tile_of[b] = t;
      hp.desc[b].tiles_n = cdiv(N[b], kTileBN[t]);
      const int BMt = kTileBM[t], BNt = kTileBN[t];
      const int r = M[b] % BMt, s = N[b] % BNt;
      const int t_bot = BNt == 128 ? T64x128 : T64;  // 64 x BN
      const int t_rgt = BMt == 128 ? T128x64 : T64;  // BM x 64
      // (edge strips only for tiles with both sides 64 or 128)
      const bool sq = BMt >= 64 && BMt <= 128 && BNt >= 64 && BNt <= 128;
      const bool half_m = edges && sq && BMt == 128 && r > 0 && r <= 64 && tile_allowed(prec, t_bot, o.tile_mask) &&
                          (BNt == 128 || tile_allowed(prec, T64, o.tile_mask));
      const bool half_n = edges && sq && BNt == 128 && s > 0 && s <= 64 && tile_allowed(prec, t_rgt, o.tile_mask) &&
                          (BMt == 128 || tile_allowed(prec, T64, o.tile_mask));
      const int rows = half_m ? M[b] / BMt : cdiv(M[b], BMt);
      const int cols = half_n ? N[b] / BNt : cdiv(N[b], BNt);
      if (rows > 0 && cols > 0) parts[t].push_back({b, 0, rows, 0, cols});
      if (half_m && cols > 0) parts[t_bot].push_back({b, 2 * rows, 2 * rows + 1, 0, cols});
      if (half_n && rows > 0) parts[t_rgt].push_back({b, 0, rows, 2 * cols, 2 * cols + 1});
      if (half_m && half_n) parts[T64].push_back({b, 2 * rows, 2 * rows + 1, 2 * cols, 2 * cols + 1});
    }
  };
  bool mixed = mix_family && (o.mixed_tiles > 0 ||
                              (o.mixed_tiles == 0 && (fam == FAM_FP32 || split_scaled || small_batch)));
  if (mix_family && o.mixed_tiles == 0 && !mixed) {
    const double kMixFrac = 0.12;
    build_parts(o.edge_tiles >= 0);  // the parts a mixed launch would have
    double cost[NTILES] = {};
    for (int t = 0; t < NTILES; ++t)
      if (mixable(t))
        for (const Part& x : parts[t])
          cost[t] += (double)(x.tm1 - x.tm0) * (x.tn1 - x.tn0) * item_cost(t, sh[x.b]);
    double all = 0.0, top = 0.0;
    for (int t = 0; t < NTILES; ++t) {
      all += cost[t];
      top = std::max(top, cost[t]);
    }
    mixed = all > 0.0 && all - top > kMixFrac * all;
  }
  const bool edges = fam != FAM_SEQ && (o.edge_tiles > 0 ||
                                        (o.edge_tiles == 0 && (mixed || fam == FAM_FP32 || fam == FAM_BF16)));
  build_parts(edges);

  // 2b. one persistent launch per tile shape (or one for all mixed shapes);
  // 3. LPT over XCD queues.
  struct TPart {
    Part p;
    int t;
    double cost = 0.0;  // modelled cost of the whole part (all slices)
    double item = 0.0;  // modelled cost of one of its items
  };
  std::vector<std::vector<TPart>> groups;
  std::vector<int> group_tile;
  // Partial mixing (default, when the launches are not all mixed): the shape
  // with the most modelled work keeps a launch of its own and its kernel's
  // allocation; the other 64x64 .. 128x128 shapes -- edge strips, corners,
  // medium matrices -- share one mixed launch instead of a short launch each,
  // whose start-up and tail dominate (measured on the unbalanced B=2048 FP32
  // batch: three edge launches of 1619 / 1498 / 386 items took 0.29 / 0.26 /
  // 0.10 ms at 41 / 50 / 2.5 TFLOP/s).
  int keep = -1;
  if (mix_family && !mixed && o.mixed_tiles == 0 && VBG_PARTIAL_MIX) {
    double best = 0.0;
    int others = 0;
    for (int t = 0; t < NTILES; ++t) {
      if (parts[t].empty() || !mixable(t)) continue;
      double c = 0.0;
      for (const Part& x : parts[t]) c += (double)(x.tm1 - x.tm0) * (x.tn1 - x.tn0) * item_cost(t, sh[x.b]);
      if (keep < 0 || c > best) {
        best = c;
        keep = t;
      }
      ++others;
    }
    if (others < 3) keep = -1;  // at most one short launch beside the kept one: no gain
  }
  {
    std::vector<TPart> mix;
    for (int t = 0; t < NTILES; ++t) {
      if (parts[t].empty()) continue;
      const bool into_mix = (mixed && mixable(t)) || (keep >= 0 && mixable(t) && t != keep);
      std::vector<TPart> g;
      for (const Part& x : parts[t]) {
        const double item = slice_cost(t, x.b);
        const double cost = (double)(x.tm1 - x.tm0) * (x.tn1 - x.tn0) * hp.desc[x.b].ksplit * item;
        (into_mix ? mix : g).push_back({x, t, cost, item});
      }
      if (!into_mix) {
        groups.push_back(std::move(g));
        group_tile.push_back(t);
      }
    }
    if (!mix.empty()) {
      groups.push_back(std::move(mix));
      group_tile.push_back(T_MIXED);
    }
  }
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    std::vector<TPart>& ps = groups[gi];
    Launch L;
    L.tile = group_tile[gi];
    L.family = fam;
    // Band split (off by default): a part with more work than one queue's
    // share is cut into bands of tile rows (of tile columns when it has fewer
    // rows than bands), one band per share, so each XCD works a contiguous
    // block of the matrix instead of stealing consecutive items of one queue.
    // Measured on the grouped BF16 passes, whose big group is one such part:
    // forward 0.479 vs 0.473-0.480 ms, backward 1.007 vs 1.014 ms -- no gain,
    // so L2 locality across XCDs does not bound those tiles.
    if (hp.nq > 1 && (VBG_BAND_SPLIT || o.xcd_queues >= 2)) {
      double total = 0.0;
      for (const TPart& x : ps) total += x.cost;
      const double share = total / hp.nq;
      std::vector<TPart> cut;
      for (const TPart& x : ps) {
        const int rows = x.p.tm1 - x.p.tm0, cols = x.p.tn1 - x.p.tn0;
        const int want = share > 0.0 ? (int)std::ceil(x.cost / share - 1e-9) : 1;
        const bool by_rows = rows >= want || rows >= cols;
        const int span = by_rows ? rows : cols;
        const int nb = std::min(want, span);
        if (nb < 2) {
          cut.push_back(x);
          continue;
        }
        for (int i = 0; i < nb; ++i) {
          TPart y = x;
          const int lo = (int)((int64_t)span * i / nb), hi = (int)((int64_t)span * (i + 1) / nb);
          if (by_rows) {
            y.p.tm0 = x.p.tm0 + lo;
            y.p.tm1 = x.p.tm0 + hi;
          } else {
            y.p.tn0 = x.p.tn0 + lo;
            y.p.tn1 = x.p.tn0 + hi;
          }
          y.cost = x.cost * (hi - lo) / span;
          cut.push_back(y);
        }
      }
      ps.swap(cut);
    }
    // longest parts first, ties by matrix index for determinism (costs are
    // computed once per part, not per comparison)
    std::stable_sort(ps.begin(), ps.end(), [](const TPart& x, const TPart& y) { return x.cost > y.cost; });
    std::vector<std::vector<TPart>> qm(hp.nq);
    std::vector<double> load(hp.nq, 0.0);
    for (const TPart& x : ps) {
      const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      qm[q].push_back(x);
      load[q] += x.cost;
      L.cost += x.cost;
    }
    L.q_begin.push_back(0);
    size_t n_items = 0;
    for (const TPart& x : ps)
      n_items += (size_t)(x.p.tm1 - x.p.tm0) * (x.p.tn1 - x.p.tn0) * hp.desc[x.p.b].ksplit;
    L.items.reserve(n_items);
    for (int q = 0; q < hp.nq; ++q) {
      // per-item cost decreasing; a part's tiles stay together
      std::stable_sort(qm[q].begin(), qm[q].end(), [](const TPart& x, const TPart& y) { return x.item > y.item; });
      // tiles in runs of `grp` tile rows, column by column: the ~64 items an
      // XCD runs at once form a grp x (64/grp) block of the matrix, touching
      // grp A panels and 64/grp B panels instead of 64/tn rows x tn columns
      const int grp = o.tile_group > 0 ? o.tile_group : 8;
      for (const TPart& tp : qm[q]) {
        const Part& x = tp.p;
        for (int sl = 0; sl < hp.desc[x.b].ksplit; ++sl)
          for (int i0 = x.tm0; i0 < x.tm1; i0 += grp)
            for (int j = x.tn0; j < x.tn1; ++j)
              for (int i = i0; i < std::min(x.tm1, i0 + grp); ++i)
                L.items.push_back({x.b, i, j, (short)sl, (short)tp.t});
      }
      L.q_begin.push_back((int)L.items.size());
    }
    const int per_cu = occ(fam, L.tile);
    int grid = o.grid_wgs > 0 ? o.grid_wgs : cus * per_cu;
    grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)L.items.size()));
    L.grid = grid;
    hp.launches.push_back(std::move(L));
  }
  std::stable_sort(hp.launches.begin(), hp.launches.end(),
                   [](const Launch& x, const Launch& y) { return x.cost > y.cost; });
  for (int b = 0; b < batch; ++b) {
    if (hp.desc[b].ksplit < 2) continue;
    const int64_t n = (int64_t)M[b] * N[b], target = 1 << 12;  // 16 elements per thread
    for (int64_t s0 = 0; s0 < n; s0 += target) hp.red_chunks.push_back({b, 0, s0, std::min(n, s0 + target)});
  }
  if (prec == VBG_PREC_FP16_SCALED || prec == VBG_PREC_FP8_E4M3) {
    const int64_t target = 1 << 16;  // elements per chunk
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0 || K[b] == 0) continue;
      for (int op = 0; op < 2; ++op) {
        const int rows = op == 0 ? (ta ? K[b] : M[b]) : (tb ? N[b] : K[b]);
        const int cols = op == 0 ? (ta ? M[b] : K[b]) : (tb ? K[b] : N[b]);
        const int ld = op == 0 ? lda[b] : ldb[b];
        if (ld == rows) {
          const int64_t n = (int64_t)rows * cols;
          for (int64_t s0 = 0; s0 < n; s0 += target) hp.chunks.push_back({b, op, s0, std::min(n, s0 + target)});
        } else {
          const int64_t per = std::max<int64_t>(1, target / rows);
          for (int64_t c0 = 0; c0 < cols; c0 += per)
            hp.chunks.push_back({b, op, c0, std::min<int64_t>(cols, c0 + per)});
        }
      }
    }
  }
  if (pack) {
    // Packed A (M rows) then packed B (N rows) per matrix, each of Kp = K
    // rounded up to VBG_PACK_KALIGN elements per row and 256-byte aligned;
    // one pack job per 64 rows x VBG_PACK_K k.  Matrices with K = 0 have no
    // packed operands (their items only scale C).
    const int ebp = prec == VBG_PREC_BF16X3 ? 4 : prec == VBG_PREC_FP16_SCALED ? 2 : 1;
    hp.pk_off.assign(2 * (size_t)batch, 0);
    int64_t off = 0;
    for (int b = 0; b < batch; ++b) {
      if (M[b] == 0 || N[b] == 0 || K[b] == 0) continue;
      const int64_t Kp = (int64_t)cdiv(K[b], VBG_PACK_KALIGN) * VBG_PACK_KALIGN;
      for (int op = 0; op < 2; ++op) {
        const int rows = op == 0 ? M[b] : N[b];
        const int64_t bytes = (int64_t)rows * Kp * ebp;
        if (bytes >= ((int64_t)1 << 31)) {
          err = "packed operand of matrix " + std::to_string(b) + " exceeds 2 GiB";
          return VBG_ERR_NOT_SUPPORTED;
        }
        hp.pk_off[2 * (size_t)b + op] = off;
        off += (bytes + 255) & ~(int64_t)255;
        for (int r0 = 0; r0 < rows; r0 += 64)
          for (int k0 = 0; k0 < Kp; k0 += VBG_PACK_K) hp.pack_jobs.push_back({b, op, r0, k0});
      }
    }
    hp.pk_bytes = off;
  }
  return VBG_OK;
}

}  // namespace vbg
