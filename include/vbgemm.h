/*
 * vbgemm.h -- C ABI of the MI355X (gfx950) load-balanced variable-size batched
 * GEMM library (libvbgemm.so).
 *
 * Plain C types only: device pointers are `const float* const*` arrays that
 * live in device memory (as in the reference's dev_A/dev_B/dev_C), dims and
 * leading dimensions are host `int` arrays, streams are `void*` holding a
 * hipStream_t (NULL = the default stream).  Every entry point returns a
 * vbg_status (0 on success, negative on error) and never exits the process;
 * vbg_status_string() names a code and vbg_last_error() gives the detail of the
 * last error raised on the calling thread.
 *
 * GEMM semantics are hipBLAS column-major (ROCm/default/gemm.cuh:82:
 * hipblasSgemm(handle, OP_N, OP_N, M, N, K, &alpha, A, lda=M, B, ldb=K,
 * &beta, C, ldc=M)):  C_b = alpha * op(A_b) * op(B_b) + beta * C_b with op(A_b)
 * M_b x K_b and op(B_b) K_b x N_b.  beta == 0 does not read C.
 *
 * Which reference interface each entry point replaces is stated at each
 * declaration (paths relative to the reference root).
 */
#ifndef VBGEMM_H
#define VBGEMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBG_ABI_VERSION 2

typedef enum vbg_status {
  VBG_OK = 0,
  VBG_ERR_INVALID_ARG = -1,   /* bad dimension, leading dimension, pointer or option */
  VBG_ERR_HIP = -2,           /* a HIP runtime call failed */
  VBG_ERR_NOT_SUPPORTED = -3, /* valid request this build does not implement */
  VBG_ERR_ALLOC = -4,         /* host or device allocation failed */
  VBG_ERR_PLAN_MISMATCH = -5, /* plan used with the wrong entry point or device */
  VBG_ERR_UNDEFINED = -6      /* the reference planner's behaviour is undefined here */
} vbg_status;

/* Arithmetic of a plan.  Paths 0-4 take FP32 operands and write FP32 C. */
typedef enum vbg_prec {
  VBG_PREC_FP32 = 0,        /* FP32 MFMA (v_mfma_f32_32x32x2_f32), FP32 accumulate */
  VBG_PREC_BF16X3 = 1,      /* x = hi + lo in bf16; hi*hi + hi*lo + lo*hi on bf16 MFMA */
  VBG_PREC_FP16_SCALED = 2, /* per-matrix power-of-two scale into FP16, FP16 MFMA */
  VBG_PREC_FP8_E4M3 = 3,    /* per-matrix amax scale (448/amax) into OCP e4m3, FP8 MFMA */
  VBG_PREC_FP32_SEQ = 4,    /* one sequential fma chain per element over ascending k
                               (VALU); with alpha = beta = 1 it is seeded from C and is
                               bit-identical to the reference's correct tiles */
  VBG_PREC_BF16 = 5,        /* BF16 operands (the pointer arrays hold BF16 data), FP32 C,
                               v_mfma_f32_32x32x16_bf16; lda/ldb multiples of 8 and
                               16-byte-aligned operand pointers */
  VBG_PREC_BF16_OUT_BF16 = 6 /* as VBG_PREC_BF16 with a BF16 C */
} vbg_prec;

/* How the planner picks each matrix's tile and orders the work map. */
typedef enum vbg_policy {
  VBG_POLICY_AUTO = 0,      /* MI355X cost model over the tile family */
  VBG_POLICY_REFERENCE = 1, /* reference tile choice + wavefront cascade
                               (ROCm/framework/gemm.cuh:141-201) mapped onto this
                               library's tiles; matrices the reference cannot tile
                               fall back to AUTO */
  VBG_POLICY_FIXED = 2,     /* every matrix uses opts.fixed_tile */
  VBG_POLICY_WANG = 3,      /* Wang's global TLP-chosen square tile
                               (ROCm/tiling/gemm_wang.cuh:115-182) */
  VBG_POLICY_LI = 4         /* Li's global TLP tile over the 6-tile set
                               (CUDA/batching/gemm.cu:99-127) */
} vbg_policy;

/* Tile shapes of the kernel family (index = tile id; BM x BN output tile). */
typedef enum vbg_tile {
  VBG_TILE_16x16 = 0,
  VBG_TILE_32x32 = 1,
  VBG_TILE_64x64 = 2,
  VBG_TILE_128x64 = 3,
  VBG_TILE_64x128 = 4,
  VBG_TILE_128x128 = 5,
  VBG_TILE_256x128 = 6,     /* bf16x3 and BF16 operands only: 4 waves of 128 x 64 each */
  VBG_TILE_16x128 = 7,      /* FP32 only, M <= 16: one item spans 128 columns */
  VBG_TILE_128x16 = 8,      /* FP32 only, N <= 16: one item spans 128 rows */
  VBG_TILE_256x256 = 9,     /* BF16 operands only: 4 waves of 128 x 128, 64-deep k-tiles */
  VBG_TILE_COUNT = 10
} vbg_tile;

typedef struct vbg_opts {
  int policy;           /* vbg_policy */
  int wavefront_thres;  /* REFERENCE: cascade threshold (reference default 6000) */
  int tlp_thres;        /* WANG / LI: TLP threshold */
  int fixed_tile;       /* FIXED: vbg_tile id */
  unsigned tile_mask;   /* AUTO: bit t enables tile t (0 = all; each precision uses the
                           shapes built for it) */
  int big_tile_min_mn;  /* AUTO: min(M,N) at which 128x128 becomes eligible (0 = default) */
  int split_k_min;      /* split K of matrices with K >= this whose one item would be the
                           launch's critical path (0 = 256, <0 = never split) */
  int grid_wgs;         /* persistent grid size per kernel (0 = CUs x occupancy) */
  int xcd_queues;       /* 1 = XCD-aware per-XCD work queues with stealing, 0 = one queue;
                           2 = per-XCD queues, and a matrix with more work than one queue's
                           share is first cut into bands of tile rows, one per share */
  int host_only;        /* 1 = plan on the host only (no device upload, cannot run);
                           for inspection and tests on machines without a GPU */
  int num_cus;          /* host_only: CU count to plan for (0 = 256) */
  int force_ksplit;     /* >0: split every matrix's K into this many slices (capped at its
                           k-tile count; tests and tuning); 0 = the planner decides */
  int tile_group;       /* tile order inside a matrix: runs of this many tile rows, column by
                           column (0 = 8; 1 = plain row-major), so the workgroups of one XCD
                           that run consecutive items share A and B panels in its L2 */
  int edge_tiles;       /* cover a last tile row / column of <= 64 rows / columns with 64-wide
                           tiles (less padded work): > 0 on, < 0 off, 0 = default (on for FP32,
                           BF16 and mixed launches; off for bf16x3) */
  int timing_slots;     /* > 0: every run records HIP events around each of its kernels on the
                           run's stream, keeping the last `timing_slots` runs for
                           vbg_plan_launch_times (all work still runs; a timing plan must not
                           be captured in a graph).  0 = off */
  int mixed_tiles;      /* > 0: FP32 and split-precision plans run every 64x64 .. 128x128 tile
                           (interior tiles, edge strips, medium matrices) in one persistent
                           launch whose items carry their own shape; < 0: one launch per tile
                           shape; 0 = default (mixed for FP16 / FP8, and for batches whose
                           launches would all be short) */
  int consolidate;      /* AUTO: >= 0 (default) moves the matrices of a tile shape whose launch
                           would be short into the dominant launch's shape when that is
                           modelled cheaper than a launch of their own; < 0 off */
  int pack_operands;    /* bf16x3 / FP16 / FP8: > 0 quantise every operand once per run into a
                           plan-owned workspace (k-contiguous rows, K padded to 64) that the
                           GEMM kernels stream without converting; < 0 quantise on the
                           HBM -> LDS path instead; 0 = default (FP16 / FP8 packed when the
                           batch's arithmetic intensity, 2*sum(MNK) / 4*sum(MK+KN), is at
                           least 64 flop/B and the workspace stays within 8 GiB; bf16x3
                           unpacked) */
  int item_fixed_x10;   /* AUTO cost model: fixed cost of one work item (queue pop, pipeline
                           fill, epilogue) in tenths of a 128x128x32 k-step (0 = 30) */
  int small_tile_eff_pct; /* AUTO cost model: scale (percent) of the modelled MFMA efficiency
                           of the tiles smaller than 128x128 (0 = 100); above 100 favours
                           them against the large tiles */
} vbg_opts;

typedef struct vbg_plan vbg_plan;

/* Fills *opts with the defaults (AUTO policy, XCD queues on, thres 6000). */
int vbg_opts_default(vbg_opts* opts);

const char* vbg_status_string(int status);
const char* vbg_last_error(void);
int vbg_abi_version(void);

/*
 * Creates a plan for `batch` column-major GEMMs of the given shapes.
 * Replaces the reference's host planner and uploads
 * (ROCm/framework/gemm.cuh:86-105 dims upload, :141-229 planner, :208-232
 * strategy upload): M, N, K, lda, ldb, ldc are host arrays; the plan copies
 * what it needs to device memory it owns, so the launch allocates nothing.
 * transA / transB: 'N' or 'T' for the whole batch.  Requirements: M,N,K >= 0;
 * lda >= max(1, rows of stored A), likewise ldb, ldc >= max(1, M).
 */
int vbg_plan_create(vbg_plan** plan, int batch, const int* M, const int* N, const int* K,
                    const int* lda, const int* ldb, const int* ldc, char transA, char transB,
                    int prec, const vbg_opts* opts);

int vbg_plan_destroy(vbg_plan* plan);

/*
 * Runs the plan: C_b = alpha*op(A_b)*op(B_b) + beta*C_b for every b.
 * Replaces the framework launch hipLaunchKernelGGL(gemm<256>, ...,
 * dev_M, dev_N, dev_K, dev_A, dev_B, dev_C, dev_T) (ROCm/framework/gemm.cuh:
 * 246-247, kernel ROCm/framework/kernel.h:9-42) and the per-matrix loop
 * hipblasSgemm(OP_N, OP_N, ...) (ROCm/default/gemm.cuh:79-83).
 * dA, dB, dC: device arrays of `batch` device pointers.  Asynchronous on
 * `stream`, allocates nothing, issues no memset: safe under hipGraph capture.
 * Launches that share a plan must not overlap in time (the plan's work-queue
 * counters are shared); same-stream launches never overlap.
 */
int vbg_sgemm_vbatched(vbg_plan* plan, float alpha, const float* const* dA,
                       const float* const* dB, float beta, float* const* dC, void* stream);

/* One-shot form: plan, run and destroy (synchronises the stream before the
 * plan's device buffers are freed).  Same semantics as above. */
int vbg_sgemm_vbatched_oneshot(int batch, const int* M, const int* N, const int* K,
                               const int* lda, const int* ldb, const int* ldc, char transA,
                               char transB, int prec, float alpha, const float* const* dA,
                               const float* const* dB, float beta, float* const* dC,
                               void* stream);

/*
 * One-shot form that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code that must not call
 * hipMalloc: serving loops, harnesses that forbid allocation).  Same semantics
 * as vbg_sgemm_vbatched_oneshot for FP32-operand paths (prec 0-4; BF16
 * operands return VBG_ERR_NOT_SUPPORTED), but the plan's device work map and
 * split-K workspace live in a static device array of
 * VBG_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
 * process ends: at most VBG_NOALLOC_IMAGE_BYTES of work map (a batch that
 * needs more returns VBG_ERR_INVALID_ARG; about 16 bytes per work item and 80
 * per matrix), the rest split-K partial panels (matrices whose panels would
 * not fit are not split).  Split precisions quantise in the k-loop (no packed
 * operands).  Uploads with hipMemcpyAsync and synchronises `stream` before it
 * returns.  Every call on one GPU shares that GPU's array: calls that overlap
 * in time on one GPU (other streams or threads) race.
 */
#define VBG_NOALLOC_SCRATCH_BYTES (64u << 20)
#define VBG_NOALLOC_IMAGE_BYTES (16u << 20)
int vbg_sgemm_vbatched_noalloc(int batch, const int* M, const int* N, const int* K,
                               const int* lda, const int* ldb, const int* ldc, char transA,
                               char transB, int prec, float alpha, const float* const* dA,
                               const float* const* dB, float beta, float* const* dC,
                               void* stream);

/* Plan inspection (works on host_only plans).  what:
 *   0 number of work items             1 number of kernel launches (families used)
 *   2 tile id per matrix (out[batch])  3 persistent grid size of launch 0
 *   4 work items as (mat, tile_m, tile_n, k_begin, k_end) x n (out[5*n])
 *   5 split-K parts per matrix (out[batch])
 *   6 work items as (mat, tile id, tile_m, tile_n, k slice) x n (out[5*n]); tile_m / tile_n
 *     count in units of that item's tile shape
 *   7 kernel table, one row of (kind, tile id, first item, item count, grid) per kernel a run
 *     issues, in issue order (out[5*rows]); tile id 10 = a mixed launch (items of the 64x64 ..
 *     128x128 shapes and, FP32, the skinny ones, see query 6); kind 0 = scale pre-pass (FP16 / FP8), 1 = FP32
 *     MFMA tiles, 2 = FP32 sequential-fma tiles, 3 = split-precision tiles, 4 = BF16 tiles,
 *     5 = split-K reduction, 6 = operand pack pre-pass (grid = pack jobs), 7 = split-precision
 *     tiles on packed operands; items index the lists of queries 4 and 6
 *   8 bytes of the packed-operand workspace (0 when the plan does not pack), as the return
 *     value, saturated at INT_MAX
 * Returns the number of ints written (or the scalar for 0, 1, 3). */
int vbg_plan_query(const vbg_plan* plan, int what, int* out, int out_len);

/*
 * Per-kernel durations of a timing plan (opts.timing_slots > 0), in the row
 * order of query 7: ms[r] = mean over the recorded runs (at most timing_slots,
 * the most recent) of the time between the HIP events before and after kernel
 * r on the run's stream.  Waits for those events.  Counterpart of the
 * reference's hipEvent timing of its launches (ROCm/framework/gemm.cuh:
 * 240-266), resolved per kernel.  Returns the number of rows written, or a
 * negative status (VBG_ERR_PLAN_MISMATCH for a plan without timing or with no
 * recorded run).
 */
int vbg_plan_launch_times(const vbg_plan* plan, float* ms, int n);

/*
 * Reference planner, restated (ROCm/framework/gemm.cuh:121-229): initial tile
 * choice over {16x16,16x32,32x32,32x64,64x64}, wavefront accounting, the
 * demotion cascade while total_wavefront <= wavefront_thres (including its
 * kernel_thread[t_strategy[i]] read, gemm.cuh:180), and the bounding-box grid.
 * Outputs t_strategy[batch], kernel_thread[batch], grid[3], wavefronts[2]
 * (initial, final).  Returns VBG_ERR_UNDEFINED where the reference is
 * undefined (no dividing tile, or the :180 read past the array).
 */
int vbg_ref_plan(int batch, const int* M, const int* N, const int* K, int wavefront_thres,
                 int* t_strategy, int* kernel_thread, int* grid, int64_t* wavefronts);

/* ------------------------------------------------------------------ grouped */
/*
 * Grouped GEMM for the differentiation path (no reference counterpart: the
 * reference is forward FP32 only; this reuses the variable-size work map).
 * Row-major BF16 operands, FP32 accumulate:
 *   forward   C[rows_g, N]  = A[rows_g, K] * B_g[K, N]
 *   backward  dA[rows_g, K] = dC[rows_g, N] * B_g^T
 *             dB_g[K, N]    = A[rows_g, K]^T * dC[rows_g, N]
 * A and C/dC/dA are stacked over groups (sum of M_g rows, group g starts at
 * row offset sum_{h<g} M_h); B / dB are [G, K, N].  dtype codes: 0 = fp32,
 * 1 = bf16 for the C / dC / dA / dB buffers.  Groups with M_g = 0 write zero dB.
 * A plan owns one side stream: the kernels of a pass's few-row groups run on
 * it beside the GEMM of the large groups, forked from and joined back to
 * `stream` by events, so a call is still ordered on `stream` (and can be
 * captured into a HIP graph).  Calls on one plan must not overlap in time:
 * use one plan per stream.
 */
typedef struct vbg_grouped_plan vbg_grouped_plan;
int vbg_grouped_plan_create(vbg_grouped_plan** plan, int G, const int* M_host, int K, int N,
                            const vbg_opts* opts);
int vbg_grouped_plan_destroy(vbg_grouped_plan* plan);
int vbg_grouped_gemm_fwd(vbg_grouped_plan* plan, const void* A_bf16, const void* B_bf16,
                         void* C, int c_dtype, void* stream);
int vbg_grouped_gemm_bwd(vbg_grouped_plan* plan, const void* A_bf16, const void* B_bf16,
                         const void* dC, int dc_dtype, void* dA, void* dB, int grad_dtype,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VBGEMM_H */
