// vbg_static.hip -- vbg_sgemm_vbatched_noalloc: the one-shot batched GEMM for
// callers that cannot allocate device memory.  No counterpart in the
// reference.  The plan's device image (descriptors, work items, queue starts
// and counters, scales, chunks) and its split-K workspace live in one
// file-scope device array, `scratch`, of VBG_NOALLOC_SCRATCH_BYTES per GPU,
// held by every GPU the library runs on until the process ends.
//
//   [0, VBG_NOALLOC_IMAGE_BYTES)                 the plan's image
//   [VBG_NOALLOC_IMAGE_BYTES, SCRATCH_BYTES)     split-K partial panels
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is taken once per device with hipGetSymbolAddress (the
// device current) and only that address is used, for the upload and as the
// kernels' buffer arguments.  Every call on a GPU shares that GPU's array and
// overwrites it: calls that overlap in time on one GPU (other streams or
// threads) race.  Each call synchronises its stream before returning, so
// calls from one thread never overlap.  The queue counters start at zero in
// every uploaded image and the kernels leave them at zero.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vbgemm.h"
#include "vbg_run.h"

namespace {

constexpr size_t kScratchBytes = VBG_NOALLOC_SCRATCH_BYTES;
constexpr size_t kImageBytes = VBG_NOALLOC_IMAGE_BYTES;
static_assert(kImageBytes % 256 == 0 && kImageBytes < kScratchBytes, "scratch layout");

static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (null: not yet taken)

int fail(int code, const std::string& msg) {
  vbg::set_last_error(msg);
  return code;
}

// Device address of `scratch` on device `dev` (current), taken on first use.
int scratch_address(int dev, char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(VBG_ERR_NOT_SUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(VBG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = (char*)p;
  }
  *out = g_scratch[dev];
  return VBG_OK;
}

}  // namespace

extern "C" int vbg_sgemm_vbatched_noalloc(int batch, const int* M, const int* N, const int* K,
                                          const int* lda, const int* ldb, const int* ldc, char transA,
                                          char transB, int prec, float alpha, const float* const* dA,
                                          const float* const* dB, float beta, float* const* dC,
                                          void* stream) {
  auto tr = [](char c, bool& t) {
    if (c == 'N' || c == 'n') t = false;
    else if (c == 'T' || c == 't' || c == 'C' || c == 'c') t = true;
    else return false;
    return true;
  };
  bool ta, tb;
  if (!tr(transA, ta) || !tr(transB, tb)) return fail(VBG_ERR_INVALID_ARG, "trans must be 'N' or 'T'");
  if (prec == VBG_PREC_BF16 || prec == VBG_PREC_BF16_OUT_BF16)
    return fail(VBG_ERR_NOT_SUPPORTED, "the no-allocation entry point takes FP32 operands (prec 0-4)");
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(VBG_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return fail(VBG_ERR_HIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e));
  char* base = nullptr;
  if (int rc = scratch_address(dev, &base)) return rc;

  vbg_opts o;
  vbg_opts_default(&o);
  o.pack_operands = -1;  // no packed-operand workspace: operands are quantised in the k-loop
  o.timing_slots = 0;
  vbg::HostPlan hp;
  std::string err;
  const int64_t ws_cap = (int64_t)((kScratchBytes - kImageBytes) / sizeof(float));
  int rc = vbg::build_plan(hp, batch, M, N, K, lda, ldb, ldc, ta, tb, prec, o, cus, vbg::occupancy, err,
                           ws_cap);
  if (rc != VBG_OK) return fail(rc, err);
  if (hp.batch == 0) return VBG_OK;
  if (!dA || !dB || !dC) return fail(VBG_ERR_INVALID_ARG, "null pointer array");
  vbg::ImageLayout lay;
  std::vector<char> host;
  vbg::build_image(hp, lay, host);
  if (lay.bytes > kImageBytes)
    return fail(VBG_ERR_INVALID_ARG, "batch too large for the no-allocation scratch: its work map needs " +
                                         std::to_string(lay.bytes) + " bytes, the limit is " +
                                         std::to_string(kImageBytes));
  hipStream_t s = (hipStream_t)stream;
  e = hipMemcpyAsync(base, host.data(), lay.bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail(VBG_ERR_HIP, std::string("hipMemcpyAsync(work map): ") + hipGetErrorString(e));
  vbg::DeviceView v;
  vbg::view_image(hp, lay, base, v);
  v.ws = hp.ws_floats > 0 ? (float*)(base + kImageBytes) : nullptr;
  v.pk = nullptr;
  rc = vbg::issue_run(hp, v, alpha, dA, dB, beta, dC, s, nullptr, err);
  // the host image must outlive its copy, and the next call reuses the scratch
  e = hipStreamSynchronize(s);
  if (rc != VBG_OK) return fail(rc, err);
  if (e != hipSuccess) return fail(VBG_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return VBG_OK;
}
