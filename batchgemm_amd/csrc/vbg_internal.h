// vbg_internal.h -- structures shared by the host planner and the gfx950 kernels.
#pragma once
#include <stdint.h>

// Per-matrix descriptor, uploaded once per plan (replaces the reference's
// separate dev_M / dev_N / dev_K arrays, ROCm/framework/gemm.cuh:98-105).
struct vbg_mat_desc {
  int M, N, K;
  int lda, ldb, ldc;
  int tiles_n;  // tiles along N for the matrix's tile shape
  int ksplit;   // number of K slices (1 = no split)
};

// One unit of work: a (matrix, tile_m, tile_n, k-slice) quadruple.
struct vbg_item {
  int mat;
  int tm, tn;     // tile coordinates in units of the item's tile shape
  short kslice;   // split-K slice (0 when K is not split)
  short tile;     // tile id of this item (T16 .. T128)
};

// A slice of one stored operand for the per-matrix amax pre-pass (FP16 / FP8):
// elements [begin, end) when the operand is packed (ld == rows), else columns.
struct vbg_chunk {
  int mat;
  int op;  // 0 = A, 1 = B
  int64_t begin, end;
};

// One job of the operand pack pre-pass (split precisions with packed
// operands): rows [row0, row0 + 64) x k [k0, k0 + VBG_PACK_K) of operand `op`
// (0 = A, 1 = B) of matrix `mat`, written quantised to the packed workspace.
struct vbg_pack_job {
  int mat;
  int op;
  int row0, k0;
};
// Packed operands hold K rounded up to this many elements per row (whole
// k-tiles for every split tile depth, zeros past K).
#define VBG_PACK_KALIGN 64
#define VBG_PACK_K 256

// Arguments of every persistent GEMM kernel (passed by value).
struct vbg_args {
  const vbg_mat_desc* desc;
  const vbg_item* items;
  const int* q_begin;     // [nq + 1] item ranges of the work queues
  unsigned* counters;     // [nq] queue heads, [nq] done counter
  const float* const* A;  // device arrays of device pointers
  const float* const* B;
  float* const* C;
  float* ws;              // split-K partial sums (ws_off per item slice), or null
  const int64_t* ws_off;  // per-matrix offset into ws (floats), or null
  const float* scales;    // per-matrix quantisation scales [2*batch] (FP16 / FP8), or null
  const char* pk;         // packed operands (split precisions), or null: operands from A / B
  const int64_t* pk_off;  // [2*batch] byte offsets of packed A and B in pk
  float alpha, beta;
  int nq;
  int look_all;           // 1: a workgroup whose queue is empty reads every queue's head at once
  int trans_a, trans_b;   // 0 = N, 1 = T
  int prec;               // vbg_prec
};

// Tile ids (mirror vbg_tile in include/vbgemm.h).  256x128 is built for the
// bf16x3 and BF16 families, 256x256 for the BF16 family only, 16x128 and
// 128x16 (skinny: one side of at most 16) for the FP32 family only.
enum {
  T16 = 0, T32 = 1, T64 = 2, T128x64 = 3, T64x128 = 4, T128 = 5, T256x128 = 6, T16x128 = 7, T128x16 = 8,
  T256 = 9, NTILES = 10
};
// Launch tile id of a mixed launch: one persistent kernel whose items are of
// the shapes 64x64, 128x64, 64x128 and 128x128 (and, FP32, 16x128 and
// 128x16), each carrying its own tile id.
enum { T_MIXED = 10 };
static const int kTileBM[NTILES] = {16, 32, 64, 128, 64, 128, 256, 16, 128, 256};
static const int kTileBN[NTILES] = {16, 32, 64, 64, 128, 128, 128, 128, 16, 256};

// Work-queue count of the XCD-aware scheduler: MI355X has 8 XCDs and
// dispatches consecutive workgroup ids round-robin across them.
#define VBG_NUM_XCD 8
// Granularity of split-K slices (elements of K); a multiple of every tile's BK.
#define VBG_KSLICE_ALIGN 64
