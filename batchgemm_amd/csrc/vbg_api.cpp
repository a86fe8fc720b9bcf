// vbg_api.cpp -- the C ABI (include/vbgemm.h): plan objects, device work-map
// upload, launches and error reporting.  Errors are returned as vbg_status
// codes; nothing here exits the process (the reference's ErrChk prints and
// exits, ROCm/include/util.h:58-64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vbgemm.h"
#include "vbg_launch.h"
#include "vbg_planner.h"
#include "vbg_run.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return VBG_ERR_HIP;
}

}  // namespace

void vbg::set_last_error(const std::string& msg) { g_err = msg; }

struct vbg_plan {
  vbg::HostPlan hp;
  vbg_opts opts;
  int device = -1;
  bool host_only = false;
  void* dbuf = nullptr;  // one device allocation: the plan's image (descs, items, queues, counters, scales)
  float* d_ws = nullptr; // split-K partial panels (own allocation)
  char* d_pk = nullptr;  // packed operands (own allocation)
  vbg::DeviceView view;
  // Kernel timing (opts.timing_slots > 0): `slots` rings of rows + 1 events.
  std::vector<hipEvent_t> ev;
  int slots = 0;
  int64_t runs = 0;
};

namespace {

bool has_prepass(const vbg::HostPlan& hp) {
  return hp.prec == VBG_PREC_FP16_SCALED || hp.prec == VBG_PREC_FP8_E4M3;
}

bool has_pack(const vbg::HostPlan& hp) { return !hp.pack_jobs.empty(); }

using vbg::kernel_rows;

int family_kind(int fam) {
  switch (fam) {
    case vbg::FAM_FP32: return 1;
    case vbg::FAM_SEQ: return 2;
    case vbg::FAM_SPLIT: return 3;
    case vbg::FAM_BF16: return 4;
    case vbg::FAM_SPLIT_PK: return 7;
    default: return -1;
  }
}

void destroy_events(vbg_plan* p) {
  for (hipEvent_t e : p->ev)
    if (e) (void)hipEventDestroy(e);
  p->ev.clear();
}

}  // namespace

extern "C" {

int vbg_abi_version(void) { return VBG_ABI_VERSION; }

const char* vbg_status_string(int s) {
  switch (s) {
    case VBG_OK: return "VBG_OK";
    case VBG_ERR_INVALID_ARG: return "VBG_ERR_INVALID_ARG";
    case VBG_ERR_HIP: return "VBG_ERR_HIP";
    case VBG_ERR_NOT_SUPPORTED: return "VBG_ERR_NOT_SUPPORTED";
    case VBG_ERR_ALLOC: return "VBG_ERR_ALLOC";
    case VBG_ERR_PLAN_MISMATCH: return "VBG_ERR_PLAN_MISMATCH";
    case VBG_ERR_UNDEFINED: return "VBG_ERR_UNDEFINED";
    default: return "VBG_ERR_UNKNOWN";
  }
}

const char* vbg_last_error(void) { return g_err.c_str(); }

int vbg_opts_default(vbg_opts* o) {
  if (!o) return fail(VBG_ERR_INVALID_ARG, "null opts");
  std::memset(o, 0, sizeof(*o));
  o->policy = VBG_POLICY_AUTO;
  o->wavefront_thres = 6000;  // ROCm/framework/run.sh:4
  o->tlp_thres = 6000;
  o->fixed_tile = VBG_TILE_128x128;
  o->xcd_queues = 1;
  return VBG_OK;
}

int vbg_ref_plan(int batch, const int* M, const int* N, const int* K, int wavefront_thres,
                 int* t_strategy, int* kernel_thread, int* grid, int64_t* wavefronts) {
  int rc = vbg::ref_plan(batch, M, N, K, wavefront_thres, t_strategy, kernel_thread, grid, wavefronts);
  if (rc == VBG_ERR_UNDEFINED) g_err = "reference planner undefined for this batch";
  else if (rc != VBG_OK) g_err = "invalid argument";
  return rc;
}

int vbg_plan_create(vbg_plan** plan, int batch, const int* M, const int* N, const int* K,
                    const int* lda, const int* ldb, const int* ldc, char transA, char transB,
                    int prec, const vbg_opts* opts) {
  if (!plan) return fail(VBG_ERR_INVALID_ARG, "null plan out-pointer");
  *plan = nullptr;
  auto tr = [](char c, bool& t) {
    if (c == 'N' || c == 'n') t = false;
    else if (c == 'T' || c == 't' || c == 'C' || c == 'c') t = true;
    else return false;
    return true;
  };
  bool ta, tb;
  if (!tr(transA, ta) || !tr(transB, tb)) return fail(VBG_ERR_INVALID_ARG, "trans must be 'N' or 'T'");
  vbg_opts o;
  vbg_opts_default(&o);
  if (opts) o = *opts;
  const bool host_only = o.host_only != 0;

  int dev = -1, cus = o.num_cus > 0 ? o.num_cus : 256;
  if (!host_only) {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    cus = prop.multiProcessorCount;
  }
  vbg_plan* p = new (std::nothrow) vbg_plan();
  if (!p) return fail(VBG_ERR_ALLOC, "host allocation");
  p->opts = o;
  p->device = dev;
  p->host_only = host_only;
  std::string err;
  int rc = vbg::build_plan(p->hp, batch, M, N, K, lda, ldb, ldc, ta, tb, prec, o, cus,
                           host_only ? nullptr : vbg::occupancy, err);
  if (rc != VBG_OK) {
    delete p;
    return fail(rc, err);
  }
  if (!host_only) {
    const vbg::HostPlan& hp = p->hp;
    vbg::ImageLayout lay;
    std::vector<char> host;
    vbg::build_image(hp, lay, host);
    hipError_t e = hipMalloc(&p->dbuf, lay.bytes);
    if (e != hipSuccess) {
      delete p;
      return hip_fail(e, "hipMalloc(plan)");
    }
    e = hipMemcpy(p->dbuf, host.data(), lay.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p->dbuf);
      delete p;
      return hip_fail(e, "hipMemcpy(plan)");
    }
    vbg::view_image(hp, lay, (char*)p->dbuf, p->view);
    if (hp.ws_floats > 0) {
      e = hipMalloc(&p->d_ws, sizeof(float) * hp.ws_floats);
      if (e != hipSuccess) {
        int rc2 = hip_fail(e, "hipMalloc(split-K workspace)");
        vbg_plan_destroy(p);
        return rc2;
      }
    }
    if (hp.pk_bytes > 0) {
      e = hipMalloc(&p->d_pk, (size_t)hp.pk_bytes);
      if (e != hipSuccess) {
        int rc2 = hip_fail(e, "hipMalloc(packed operands)");
        vbg_plan_destroy(p);
        return rc2;
      }
    }
    p->view.ws = p->d_ws;
    p->view.pk = p->d_pk;
    if (o.timing_slots > 0) {
      p->slots = o.timing_slots;
      p->ev.assign((size_t)p->slots * (kernel_rows(hp) + 1), nullptr);
      for (auto& ev : p->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
          int rc2 = hip_fail(e, "hipEventCreate");
          vbg_plan_destroy(p);
          return rc2;
        }
      }
    }
  }
  *plan = p;
  return VBG_OK;
}

int vbg_plan_destroy(vbg_plan* p) {
  if (!p) return VBG_OK;
  int rc = VBG_OK;
  destroy_events(p);
  if (p->dbuf) {
    hipError_t e = hipFree(p->dbuf);
    if (e != hipSuccess) rc = hip_fail(e, "hipFree(plan)");
  }
  if (p->d_ws) {
    hipError_t e = hipFree(p->d_ws);
    if (e != hipSuccess) rc = hip_fail(e, "hipFree(workspace)");
  }
  if (p->d_pk) {
    hipError_t e = hipFree(p->d_pk);
    if (e != hipSuccess) rc = hip_fail(e, "hipFree(packed operands)");
  }
  delete p;
  return rc;
}

int vbg_sgemm_vbatched(vbg_plan* p, float alpha, const float* const* dA, const float* const* dB,
                       float beta, float* const* dC, void* stream) {
  if (!p) return fail(VBG_ERR_INVALID_ARG, "null plan");
  if (p->host_only) return fail(VBG_ERR_PLAN_MISMATCH, "host_only plan cannot run");
  const vbg::HostPlan& hp = p->hp;
  if (hp.batch == 0) return VBG_OK;
  if (!dA || !dB || !dC) return fail(VBG_ERR_INVALID_ARG, "null pointer array");
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev != p->device) return fail(VBG_ERR_PLAN_MISMATCH, "plan was created on another device");
  hipStream_t s = (hipStream_t)stream;

  // Timing plans bracket every kernel with events of this run's ring slot.
  hipEvent_t* ev = nullptr;
  if (p->slots > 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    e = hipStreamIsCapturing(s, &cs);
    if (e != hipSuccess) return hip_fail(e, "hipStreamIsCapturing");
    if (cs != hipStreamCaptureStatusNone)
      return fail(VBG_ERR_NOT_SUPPORTED, "a timing plan cannot be captured in a graph");
    ev = &p->ev[(size_t)(p->runs % p->slots) * (kernel_rows(hp) + 1)];
  }
  std::string err;
  const int rc = vbg::issue_run(hp, p->view, alpha, dA, dB, beta, dC, s, ev, err);
  if (rc != VBG_OK) return fail(rc, err);
  if (ev) ++p->runs;
  return VBG_OK;
}

int vbg_plan_launch_times(const vbg_plan* p, float* ms, int n) {
  if (!p) return fail(VBG_ERR_INVALID_ARG, "null plan");
  if (p->slots <= 0 || p->runs == 0)
    return fail(VBG_ERR_PLAN_MISMATCH, "plan has no timing slots or no recorded run");
  const int rows = kernel_rows(p->hp);
  if (!ms || n < rows) return fail(VBG_ERR_INVALID_ARG, "out too small");
  const int64_t nrec = std::min<int64_t>(p->runs, p->slots);
  std::vector<double> acc(rows, 0.0);
  for (int64_t r = p->runs - nrec; r < p->runs; ++r) {
    const hipEvent_t* ev = &p->ev[(size_t)(r % p->slots) * (rows + 1)];
    hipError_t e = hipEventSynchronize(ev[rows]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    for (int k = 0; k < rows; ++k) {
      float t = 0.0f;
      e = hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
      acc[k] += t;
    }
  }
  for (int k = 0; k < rows; ++k) ms[k] = (float)(acc[k] / (double)nrec);
  return rows;
}

int vbg_sgemm_vbatched_oneshot(int batch, const int* M, const int* N, const int* K, const int* lda,
                               const int* ldb, const int* ldc, char transA, char transB, int prec,
                               float alpha, const float* const* dA, const float* const* dB,
                               float beta, float* const* dC, void* stream) {
  vbg_plan* p = nullptr;
  int rc = vbg_plan_create(&p, batch, M, N, K, lda, ldb, ldc, transA, transB, prec, nullptr);
  if (rc != VBG_OK) return rc;
  rc = vbg_sgemm_vbatched(p, alpha, dA, dB, beta, dC, stream);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (rc == VBG_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  int rc2 = vbg_plan_destroy(p);
  return rc != VBG_OK ? rc : rc2;
}

int vbg_plan_query(const vbg_plan* p, int what, int* out, int out_len) {
  if (!p) return fail(VBG_ERR_INVALID_ARG, "null plan");
  const vbg::HostPlan& hp = p->hp;
  int64_t n = 0;
  for (const auto& L : hp.launches) n += (int64_t)L.items.size();
  switch (what) {
    case 0: return (int)n;
    case 1: return (int)hp.launches.size();
    case 3: return hp.launches.empty() ? 0 : hp.launches[0].grid;
    case 2:
    case 5: {
      if (!out || out_len < hp.batch) return fail(VBG_ERR_INVALID_ARG, "out too small");
      for (int b = 0; b < hp.batch; ++b) out[b] = what == 2 ? hp.tile_of[b] : hp.desc[b].ksplit;
      return hp.batch;
    }
    case 6: {
      if (!out || out_len < 5 * n) return fail(VBG_ERR_INVALID_ARG, "out too small");
      int w = 0;
      for (const auto& L : hp.launches)
        for (const auto& it : L.items) {
          out[w++] = it.mat;
          out[w++] = it.tile;
          out[w++] = it.tm;
          out[w++] = it.tn;
          out[w++] = it.kslice;
        }
      return w;
    }
    case 4: {
      if (!out || out_len < 5 * n) return fail(VBG_ERR_INVALID_ARG, "out too small");
      int w = 0;
      for (const auto& L : hp.launches)
        for (const auto& it : L.items) {
          const vbg_mat_desc& d = hp.desc[it.mat];
          int kb = 0, ke = d.K;
          if (d.ksplit > 1) {  // same rule as the kernels' vbg_k_range
            const int al = VBG_KSLICE_ALIGN;
            const int kt = (d.K + al - 1) / al, per = (kt + d.ksplit - 1) / d.ksplit;
            kb = std::min(d.K, it.kslice * per * al);
            ke = std::min(d.K, kb + per * al);
          }
          out[w++] = it.mat;
          out[w++] = it.tm;
          out[w++] = it.tn;
          out[w++] = kb;
          out[w++] = ke;
        }
      return w;
    }
    case 7: {
      const int rows = kernel_rows(hp);
      if (!out || out_len < 5 * rows) return fail(VBG_ERR_INVALID_ARG, "out too small");
      int w = 0, first = 0;
      auto put = [&](int kind, int tile, int f, int cnt, int grid) {
        out[w++] = kind;
        out[w++] = tile;
        out[w++] = f;
        out[w++] = cnt;
        out[w++] = grid;
      };
      if (has_prepass(hp)) put(0, -1, 0, 0, (int)hp.chunks.size());
      if (has_pack(hp)) put(6, -1, 0, 0, (int)hp.pack_jobs.size());
      for (const auto& L : hp.launches) {
        put(family_kind(L.family), L.tile, first, (int)L.items.size(), L.grid);
        first += (int)L.items.size();
      }
      if (!hp.red_chunks.empty()) put(5, -1, 0, 0, (int)hp.red_chunks.size());
      return w;
    }
    case 8: return (int)std::min<int64_t>(hp.pk_bytes, 0x7fffffff);
    default: return fail(VBG_ERR_INVALID_ARG, "unknown query");
  }
}

}  // extern "C"
