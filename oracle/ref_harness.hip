// Launch harness for the reference's framework kernels (test infrastructure).
//
// Compiled by oracle/build_ref.sh against the reference's own kernel sources
// where they lie (ROCm/framework/kernel.h, which pulls in kernel_128_new.h and
// kernel_256_new.h); nothing of them is copied here.  The result,
// oracle/_ref/libref_framework.so, is loaded only by the GPU parity tests to
// run the reference's dispatcher gemm<256> (ROCm/framework/kernel.h:9-42) on
// the same device buffers the product path uses.
//
// The launch mirrors ROCm/framework/gemm.cuh:238: grid (gx, gy, batch),
// 256 threads, `lds_bytes` of dynamic LDS.  The reference launches 4096 bytes;
// its 32x64 tile indexes 6144 bytes of LDS (ROCm/framework/kernel_256_new.h:5,
// :58), which the CUDA build provides, so callers choose the size.
//
// Callers guarantee in-bounds use: every matrix's tile id is 2 (32x32,
// column-major, ROCm/framework/kernel_128_new.h:131-201) or 3 (32x64,
// row-major, kernel_256_new.h:2-72), M and N are multiples of the tile, K is a
// positive multiple of 8, and pointers are 16-byte aligned.  Under those
// conditions every global and LDS access of both tiles stays inside its buffer.
#include <hip/hip_runtime.h>

#include "kernel.h"

extern "C" int ref_framework_gemm(int batch, int gx, int gy, int lds_bytes, int* dM, int* dN,
                                  int* dK, float** dA, float** dB, float** dC, int* dT,
                                  void* stream) {
  if (batch <= 0 || gx <= 0 || gy <= 0 || lds_bytes < 0 || lds_bytes > 65536) return -1;
  hipLaunchKernelGGL(gemm<256>, dim3(gx, gy, batch), dim3(256), lds_bytes, (hipStream_t)stream,
                     dM, dN, dK, dA, dB, dC, dT);
  return (int)hipGetLastError();
}
