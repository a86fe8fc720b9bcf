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
// Groups of at most kSmallRows rows take dB from a direct kernel instead: for
// them dB_g is an outer product of reduction length M_g <= 16, whose cost is
// writing K x N outputs, not MFMA work -- as GEMM items they would pad M_g up
// to a 64-deep k-tile in every 128 x 128 output tile.
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

// dB_g[k][n] = sum_{r < M_g} A[off_g + r][k] * dC[off_g + r][n] for the small
// groups (blockIdx.y indexes them): 8 consecutive n of one k per thread, FP32
// accumulation in row order, 16- or 32-byte stores.  BF16 products are exact
// in FP32, so only the additions round.
template <bool OUT_BF16>
__global__ void __launch_bounds__(256)
    vbg_grouped_db_small(const int* small, int n_small, const int* msize, const int64_t* rowoff,
                         int K, int N, const unsigned short* A, const unsigned short* dC, void* dB) {
  const int64_t n8 = N / 8;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)K * n8) return;
  const int k = (int)(idx / n8), n0 = (int)(idx - k * n8) * 8;
  for (int j = blockIdx.y; j < n_small; j += gridDim.y) {  // groups beyond gridDim.y loop
  const int g = small[j];
  const int64_t off = rowoff[g];
  const int mg = msize[g];
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int r = 0; r < mg; ++r) {
    const float av = bf16_to_f32(A[(off + r) * K + k]);
    const uint4 raw = *(const uint4*)(dC + (off + r) * N + n0);
    const unsigned w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      acc[2 * e] = fmaf(av, __builtin_bit_cast(float, w[e] << 16), acc[2 * e]);
      acc[2 * e + 1] = fmaf(av, __builtin_bit_cast(float, w[e] & 0xffff0000u), acc[2 * e + 1]);
    }
  }
  const int64_t o = (int64_t)g * K * N + (int64_t)k * N + n0;
  if (OUT_BF16) {
    unsigned w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const __bf16 lo = (__bf16)acc[2 * e], hi = (__bf16)acc[2 * e + 1];
      w[e] = (unsigned)__builtin_bit_cast(unsigned short, lo) |
             ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
    }
    *(uint4*)((unsigned short*)dB + o) = uint4{w[0], w[1], w[2], w[3]};
  } else {
    float4* q = (float4*)((float*)dB + o);
    q[0] = float4{acc[0], acc[1], acc[2], acc[3]};
    q[1] = float4{acc[4], acc[5], acc[6], acc[7]};
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
  int* big = nullptr;      // groups whose dB runs on the GEMM plan
  int* small = nullptr;    // groups whose dB runs on vbg_grouped_db_small
  int n_big = 0, n_small = 0;
};

static int gfail(int code, const std::string& m) {
  g_gerr = m;
  return code;
}

extern "C" {

int vbg_grouped_plan_destroy(vbg_grouped_plan* p) {
  if (!p) return VBG_OK;
  for (auto& pp : p->plans)
    for (auto* q : pp) vbg_plan_destroy(q);
  if (p->dbuf) (void)hipFree(p->dbuf);
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
  std::vector<int> m(M_host, M_host + G), kk(G, K), nn(G, N);
  std::vector<int> lN(G, N), lK(G, K), lM(G);
  for (int g = 0; g < G; ++g) lM[g] = std::max(1, M_host[g]);
  std::vector<int> big, small, bigN, bigK, bigM;
  for (int g = 0; g < G; ++g) {
    if (M_host[g] > kSmallRows) {
      big.push_back(g);
      bigN.push_back(N);
      bigK.push_back(K);
      bigM.push_back(M_host[g]);
    } else {
      small.push_back(g);
    }
  }
  // forward: (m=N, n=M_g, k=K) NN, lda=N, ldb=K, ldc=N
  // dA:      (m=K, n=M_g, k=N) TN, lda=N, ldb=N, ldc=K
  // dB:      (m=N, n=K, k=M_g) NT, lda=N, ldb=K, ldc=N
  int rc = VBG_OK;
  for (int o = 0; o < 2 && rc == VBG_OK; ++o) {
    const int prec = o ? VBG_PREC_BF16_OUT_BF16 : VBG_PREC_BF16;
    rc = vbg_plan_create(&p->plans[PASS_FWD][o], G, nn.data(), m.data(), kk.data(), lN.data(),
                         lK.data(), lN.data(), 'N', 'N', prec, opts);
    if (rc == VBG_OK)
      rc = vbg_plan_create(&p->plans[PASS_DA][o], G, kk.data(), m.data(), nn.data(), lN.data(),
                           lN.data(), lK.data(), 'T', 'N', prec, opts);
    if (rc == VBG_OK && !big.empty())
      rc = vbg_plan_create(&p->plans[PASS_DB][o], (int)bigN.size(), bigN.data(), bigK.data(),
                           bigM.data(), bigN.data(), bigK.data(), bigN.data(), 'N', 'T', prec, opts);
  }
  if (rc != VBG_OK) {
    g_gerr = vbg_last_error();
    vbg_grouped_plan_destroy(p);
    return rc;
  }
  // device layout: rowoff[G] | 9 pointer arrays [G] | msize[G] | big[G] | small[G]
  const size_t ints_off = sizeof(int64_t) * G + 9 * sizeof(void*) * (size_t)G;
  const size_t bytes = ints_off + 3 * sizeof(int) * (size_t)G;
  std::vector<char> host(bytes, 0);
  std::memcpy(host.data(), off.data(), sizeof(int64_t) * G);
  int* hints = (int*)(host.data() + ints_off);
  std::memcpy(hints, M_host, sizeof(int) * G);
  if (!big.empty()) std::memcpy(hints + G, big.data(), sizeof(int) * big.size());
  if (!small.empty()) std::memcpy(hints + 2 * G, small.data(), sizeof(int) * small.size());
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
  p->n_big = (int)big.size();
  p->n_small = (int)small.size();
  *plan = p;
  return VBG_OK;
}

static bool aligned16(const void* q) { return ((uintptr_t)q & 15) == 0; }

static int run_pass(vbg_grouped_plan* p, int pass, const void* a, const void* b, const void* c,
                    void* x, int c_es, int x_es, int out_bf16, hipStream_t s) {
  // dB: the GEMM plan covers the big groups only (entry j = group big[j])
  const int n = pass == PASS_DB ? p->n_big : p->G;
  const int* gidx = pass == PASS_DB ? p->big : nullptr;
  if (n > 0) {
    hipLaunchKernelGGL(vbg_grouped_ptrs, dim3((n + 255) / 256), dim3(256), 0, s, pass, n, gidx, p->K,
                       p->N, p->rowoff, (const char*)a, (const char*)b, (const char*)c, (char*)x,
                       c_es, x_es, (const void**)p->ptrs[pass][0], (const void**)p->ptrs[pass][1],
                       p->ptrs[pass][2]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gfail(VBG_ERR_HIP, hipGetErrorString(e));
    const int rc = vbg_sgemm_vbatched(p->plans[pass][out_bf16], 1.0f,
                                      (const float* const*)p->ptrs[pass][0],
                                      (const float* const*)p->ptrs[pass][1], 0.0f,
                                      (float* const*)p->ptrs[pass][2], s);
    if (rc != VBG_OK) return rc;
  }
  if (pass == PASS_DB && p->n_small > 0) {
    const int64_t threads = (int64_t)p->K * (p->N / 8);
    const dim3 grid((unsigned)((threads + 255) / 256), (unsigned)std::min(p->n_small, 65535));
    if (out_bf16)
      hipLaunchKernelGGL(vbg_grouped_db_small<true>, grid, dim3(256), 0, s, p->small, p->n_small,
                         p->msize, p->rowoff, p->K, p->N, (const unsigned short*)a,
                         (const unsigned short*)c, x);
    else
      hipLaunchKernelGGL(vbg_grouped_db_small<false>, grid, dim3(256), 0, s, p->small, p->n_small,
                         p->msize, p->rowoff, p->K, p->N, (const unsigned short*)a,
                         (const unsigned short*)c, x);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gfail(VBG_ERR_HIP, hipGetErrorString(e));
  }
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
