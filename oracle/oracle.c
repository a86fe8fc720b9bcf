/*
 * oracle.c -- CPU restatement of the reference's load-balanced variable-size
 * batched SGEMM path.  TEST INFRASTRUCTURE ONLY: the product library never
 * links, loads or calls anything in this directory.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it, and only as the checker.
 *
 * Reference files restated here (paths relative to the reference root):
 *   ROCm/framework/gemm.cuh:34-48    dataset ingest (whitespace "M N K" triples)
 *   ROCm/framework/gemm.cuh:112-116  FLOP accounting sum(2MNK + 2MN)
 *   ROCm/framework/gemm.cuh:121-229  planner: tile table, initial tile choice,
 *                                    wavefront accounting, demotion cascade,
 *                                    bounding-box grid
 *   ROCm/framework/kernel_128_new.h:131-201  gemm_128_32x32 (column-major tile)
 *   ROCm/framework/kernel_256_new.h:2-72  gemm_256_32x64 (row-major tile):  
 *                                    per element a sequential fma chain over
 *                                    ascending k, seeded from C
 *   ROCm/default/gemm.cuh:82         hipblasSgemm(OP_N, OP_N, ..., lda=M,
 *                                    ldb=K, ldc=M): column-major alpha/beta GEMM
 *
 * What pins each function (tests/test_oracle.py):
 *   orc_read_dataset + orc_flops_ref : the reference's logged GFLOP totals
 *                                      (ROCm/framework/log-test-verify3)
 *   orc_ref_plan                     : hand-derived small cases (including the
 *                                      kernel_thread[t_strategy[i]] quirk at
 *                                      gemm.cuh:180), brute-force invariants and
 *                                      dataset facts (initial tile counts,
 *                                      initial wavefront totals, B=256 grid)
 *   orc_dgemm / orc_dgemm_vbatched   : numpy float64 matmul on random shapes,
 *                                      all transpose/ld/alpha/beta variants
 *   orc_sgemm_vbatched               : orc_dgemm at FP32 tolerance
 *   orc_ref_tile_seq                 : closed form C = c + K*a*b on constant
 *                                      data (the reference's own log-verify
 *                                      check, ROCm/tiling/log-verify) and an
 *                                      exact-product float64 emulation
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_IO (-2)
#define ORC_UNDEFINED (-3) /* the reference's behaviour is undefined here */

/* ------------------------------------------------------------------------ */
/* a1: dataset ingest -- ROCm/framework/gemm.cuh:34-48                        */
/* Reads up to `max` whitespace-separated "M N K" triples; returns the count. */
int orc_read_dataset(const char* path, int max, int* M, int* N, int* K) {
  FILE* f = fopen(path, "r");
  if (!f) return ORC_ERR_IO;
  int n = 0;
  while (n < max && fscanf(f, "%d %d %d", &M[n], &N[n], &K[n]) == 3) ++n;
  fclose(f);
  return n;
}

/* a2: FLOP accounting -- ROCm/framework/gemm.cuh:112-116 counts
 * sum(2MNK + 2MN) / 1e9 in double, adding per-matrix GFLOP terms in order. */
double orc_flops_ref(int batch, const int* M, const int* N, const int* K) {
  double g = 0.0;
  for (int i = 0; i < batch; ++i)
    g += ((2 * (int64_t)M[i] * (int64_t)N[i] * (int64_t)K[i]) +
          (2 * (int64_t)M[i] * (int64_t)N[i])) / 1.0e9;
  return g;
}

/* The metric's convention: 2 * sum(MNK), exact in int64. */
int64_t orc_flops(int batch, const int* M, const int* N, const int* K) {
  int64_t s = 0;
  for (int i = 0; i < batch; ++i) s += 2 * (int64_t)M[i] * N[i] * K[i];
  return s;
}

/* ------------------------------------------------------------------------ */
/* a3-a7: planner -- ROCm/framework/gemm.cuh:121-229                          */
static const int REF_TILE[5][2] = {{16, 16}, {16, 32}, {32, 32}, {32, 64}, {64, 64}};

static int ref_threads(int t) { return (t == 0 || t == 1 || t == 2) ? 128 : 256; }

/*
 * Outputs:
 *   t_strategy[batch]   tile index per matrix (0..4) after the cascade
 *   kernel_thread[batch] per-matrix thread count array after the cascade
 *   t_init[batch]       tile index from the initial choice (may be NULL)
 *   grid[3]             launch grid (x, y, z)
 *   wf[2]               total_wavefront after initial choice and after cascade
 * Returns ORC_UNDEFINED when some matrix has no dividing tile (the reference
 * leaves t_strategy[i] as uninitialised malloc memory, gemm.cuh:139-151), or
 * when the cascade would read kernel_thread[] past its end (the index is a
 * tile number, gemm.cuh:180, so batch < 5 can read out of bounds).
 */
int orc_ref_plan(int batch, const int* M, const int* N, const int* K, int thres,
                 int* t_strategy, int* kernel_thread, int* t_init, int* grid, int64_t* wf) {
  (void)K;
  if (batch <= 0) return ORC_ERR_ARG;
  /* :141-151 initial choice: largest index whose tile fits and divides. */
  for (int i = 0; i < batch; ++i) {
    int chosen = -1;
    for (int j = 4; j >= 0; --j) {
      if (REF_TILE[j][0] <= M[i] && REF_TILE[j][1] <= N[i] && M[i] % REF_TILE[j][0] == 0 &&
          N[i] % REF_TILE[j][1] == 0) {
        chosen = j;
        break;
      }
    }
    if (chosen < 0) return ORC_UNDEFINED;
    t_strategy[i] = chosen;
    if (t_init) t_init[i] = chosen;
  }
  /* :154-171 wavefront accounting (int arithmetic, 64 = AMD wavefront). */
  int total = 0;
  for (int i = 0; i < batch; ++i) {
    int t = t_strategy[i];
    kernel_thread[i] = ref_threads(t);
    total += M[i] * N[i] / (REF_TILE[t][0] * REF_TILE[t][1]) * (kernel_thread[i] / 64);
  }
  if (wf) wf[0] = total;
  /* :174-201 demotion cascade. The "old" wavefront count reads
   * kernel_thread[t_strategy[i]] -- the array indexed by the tile number,
   * not by i (gemm.cuh:180).  Restated as is. */
  for (int j = 4; j > 0 && total <= thres; --j) {
    for (int i = 0; i < batch && total <= thres; ++i) {
      if (t_strategy[i] == j) {
        int ti = t_strategy[i];
        if (ti >= batch) return ORC_UNDEFINED;
        int old_value = M[i] * N[i] / (REF_TILE[ti][0] * REF_TILE[ti][1]) * (kernel_thread[ti] / 64);
        if (t_strategy[i] != 0) t_strategy[i] = t_strategy[i] - 1;
        kernel_thread[i] = ref_threads(t_strategy[i]);
        int tn = t_strategy[i];
        int new_value = M[i] * N[i] / (REF_TILE[tn][0] * REF_TILE[tn][1]) * (kernel_thread[i] / 64);
        total += new_value - old_value;
      }
    }
  }
  if (wf) wf[1] = total;
  /* :219-229 bounding-box grid */
  int gx = M[0] / REF_TILE[t_strategy[0]][0];
  int gy = N[0] / REF_TILE[t_strategy[0]][1];
  for (int i = 1; i < batch; ++i) {
    int x = M[i] / REF_TILE[t_strategy[i]][0];
    int y = N[i] / REF_TILE[t_strategy[i]][1];
    gx = gx > x ? gx : x;
    gy = gy > y ? gy : y;
  }
  grid[0] = gx;
  grid[1] = gy;
  grid[2] = batch;
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* FP64 numeric oracle: column-major GEMM with hipBLAS conventions          */
/* (ROCm/default/gemm.cuh:82): C_out = alpha*op(A)*op(B) + beta*C_in.         */
/* op(A) is M x K, op(B) is K x N.  trans = 'N' or 'T'.  Inputs are float     */
/* (the product's operand type); every product and sum is taken in double.    */
/* beta == 0 ignores C_in entirely (NaN in C_in does not propagate), as BLAS.  */
static inline double elemA(const float* A, int lda, char t, int i, int k) {
  return (t == 'N' || t == 'n') ? (double)A[(size_t)i + (size_t)k * lda]
                                : (double)A[(size_t)k + (size_t)i * lda];
}
static inline double elemB(const float* B, int ldb, char t, int k, int j) {
  return (t == 'N' || t == 'n') ? (double)B[(size_t)k + (size_t)j * ldb]
                                : (double)B[(size_t)j + (size_t)k * ldb];
}

int orc_dgemm(char ta, char tb, int M, int N, int K, double alpha, const float* A, int lda,
              const float* B, int ldb, double beta, const float* Cin, int ldc, double* Cout,
              int ldo) {
  if (M < 0 || N < 0 || K < 0) return ORC_ERR_ARG;
  /* op(A) packed to double, column-major M x K, so the inner loop runs over
   * contiguous i; each C(i,j) is still the plain sum over k = 0..K-1 of
   * A(i,k)*B(k,j) in double, then alpha*s + beta*C. */
  double* ap = (double*)malloc(sizeof(double) * ((size_t)M * K + 1));
  double* acc = (double*)malloc(sizeof(double) * ((size_t)M + 1));
  if (!ap || !acc) {
    free(ap);
    free(acc);
    return ORC_ERR_ARG;
  }
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < M; ++i) ap[(size_t)i + (size_t)k * M] = elemA(A, lda, ta, i, k);
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < M; ++i) acc[i] = 0.0;
    for (int k = 0; k < K; ++k) {
      const double b = elemB(B, ldb, tb, k, j);
      const double* a = ap + (size_t)k * M;
      for (int i = 0; i < M; ++i) acc[i] += a[i] * b;
    }
    for (int i = 0; i < M; ++i) {
      double c = alpha * acc[i];
      if (beta != 0.0) c += beta * (double)Cin[(size_t)i + (size_t)j * ldc];
      Cout[(size_t)i + (size_t)j * ldo] = c;
    }
  }
  free(ap);
  free(acc);
  return ORC_OK;
}

/* Batched FP64 oracle over host pointer arrays; Cout[b] has ld = ldc[b]. */
int orc_dgemm_vbatched(char ta, char tb, int batch, const int* M, const int* N, const int* K,
                       double alpha, const float* const* A, const int* lda,
                       const float* const* B, const int* ldb, double beta,
                       const float* const* Cin, const int* ldc, double* const* Cout) {
  int rc = ORC_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
  for (int b = 0; b < batch; ++b) {
    int r = orc_dgemm(ta, tb, M[b], N[b], K[b], alpha, A[b], lda[b], B[b], ldb[b], beta, Cin[b],
                      ldc[b], Cout[b], ldc[b]);
    if (r < rc) rc = r;
  }
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Host per-matrix SGEMM loop (the CPU baseline): FP32 column-major GEMM,     */
/* alpha/beta, lda/ldb/ldc, transposes; cache-blocked with a register tile    */
/* the compiler vectorises; OpenMP across matrices (nthreads <= 0: all).      */
#define SG_MB 64
#define SG_NB 4
#define SG_KB 256

static void sgemm_one(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
                      const float* B, int ldb, float beta, float* C, int ldc, float* apack) {
  /* scale C by beta first (beta == 0 overwrites) */
  for (int j = 0; j < N; ++j) {
    float* c = C + (size_t)j * ldc;
    if (beta == 0.0f)
      for (int i = 0; i < M; ++i) c[i] = 0.0f;
    else if (beta != 1.0f)
      for (int i = 0; i < M; ++i) c[i] *= beta;
  }
  int an = (ta == 'N' || ta == 'n');
  int bn = (tb == 'N' || tb == 'n');
  for (int k0 = 0; k0 < K; k0 += SG_KB) {
    int kb = K - k0 < SG_KB ? K - k0 : SG_KB;
    for (int i0 = 0; i0 < M; i0 += SG_MB) {
      int mb = M - i0 < SG_MB ? M - i0 : SG_MB;
      /* pack alpha*op(A)[i0:i0+mb, k0:k0+kb] as [k][SG_MB] (zero padded) */
      for (int k = 0; k < kb; ++k) {
        float* ap = apack + (size_t)k * SG_MB;
        for (int i = 0; i < mb; ++i) {
          float a = an ? A[(size_t)(i0 + i) + (size_t)(k0 + k) * lda]
                       : A[(size_t)(k0 + k) + (size_t)(i0 + i) * lda];
          ap[i] = alpha * a;
        }
        for (int i = mb; i < SG_MB; ++i) ap[i] = 0.0f;
      }
      for (int j0 = 0; j0 < N; j0 += SG_NB) {
        int nb = N - j0 < SG_NB ? N - j0 : SG_NB;
        float acc[SG_NB][SG_MB];
        memset(acc, 0, sizeof(acc));
        for (int k = 0; k < kb; ++k) {
          const float* ap = apack + (size_t)k * SG_MB;
          float bv[SG_NB];
          for (int jj = 0; jj < SG_NB; ++jj) {
            int j = j0 + (jj < nb ? jj : 0);
            bv[jj] = bn ? B[(size_t)(k0 + k) + (size_t)j * ldb] : B[(size_t)j + (size_t)(k0 + k) * ldb];
          }
          for (int jj = 0; jj < SG_NB; ++jj) {
#pragma omp simd
            for (int i = 0; i < SG_MB; ++i) acc[jj][i] += ap[i] * bv[jj];
          }
        }
        for (int jj = 0; jj < nb; ++jj) {
          float* c = C + (size_t)(j0 + jj) * ldc + i0;
          for (int i = 0; i < mb; ++i) c[i] += acc[jj][i];
        }
      }
    }
  }
}

int orc_sgemm_vbatched(char ta, char tb, int batch, const int* M, const int* N, const int* K,
                       float alpha, const float* const* A, const int* lda, const float* const* B,
                       const int* ldb, float beta, float* const* C, const int* ldc, int nthreads) {
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif
  for (int b = 0; b < batch; ++b)
    if (M[b] < 0 || N[b] < 0 || K[b] < 0) return ORC_ERR_ARG;
#pragma omp parallel num_threads(nt)
  {
    float* apack = (float*)malloc(sizeof(float) * SG_MB * SG_KB);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < batch; ++b)
      sgemm_one(ta, tb, M[b], N[b], K[b], alpha, A[b], lda[b], B[b], ldb[b], beta, C[b], ldc[b],
                apack);
    free(apack);
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Reference tile semantics (a12 gemm_128_32x32, a13 gemm_256_32x64): each    */
/* output element is C <- fma(a_k, b_k, C) for k = 0, 1, ..., K-1, seeded     */
/* with the old C, no alpha/beta (kernel_128_new.h:160-187 and                */
/* kernel_256_new.h:28-67, fmas :43-51).  rowmajor = 0: lda=M, ldb=K, ldc=M; */
/* rowmajor = 1: lda=K, ldb=N, ldc=N.  Its only arithmetic is the explicit    */
/* fmaf chain, so the result does not depend on the compiler's contraction.  */
int orc_ref_tile_seq(int rowmajor, int M, int N, int K, const float* A, const float* B, float* C) {
  if (M < 0 || N < 0 || K < 0) return ORC_ERR_ARG;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      size_t ci = rowmajor ? (size_t)i * N + j : (size_t)i + (size_t)j * M;
      float c = C[ci];
      for (int k = 0; k < K; ++k) {
        float a = rowmajor ? A[(size_t)i * K + k] : A[(size_t)i + (size_t)k * M];
        float b = rowmajor ? B[(size_t)k * N + j] : B[(size_t)k + (size_t)j * K];
        c = fmaf(a, b, c);
      }
      C[ci] = c;
    }
  return ORC_OK;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
