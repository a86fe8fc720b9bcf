// vbg_run.cpp -- device image of a host plan and the kernel sequence of one
// run (see vbg_run.h).  The image replaces the reference's per-run uploads of
// dev_M / dev_N / dev_K and the strategy array (ROCm/framework/gemm.cuh:
// 86-105, 232): it is built once and uploaded in one copy.
#include "vbg_run.h"

#include <algorithm>
#include <cstring>

#include "../../include/vbgemm.h"
#include "vbg_launch.h"

namespace vbg {

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

bool has_prepass(const HostPlan& hp) {
  return hp.prec == VBG_PREC_FP16_SCALED || hp.prec == VBG_PREC_FP8_E4M3;
}

bool has_pack(const HostPlan& hp) { return !hp.pack_jobs.empty(); }

std::string last_hip() { return hipGetErrorString(hipGetLastError()); }

}  // namespace

int occupancy(int family, int prec, int tile, bool ta, bool tb) {
  switch (family) {
    case FAM_FP32: return vbg_occupancy_fp32(tile, ta, tb);
    case FAM_SEQ: return vbg_occupancy_seq(ta, tb);
    case FAM_SPLIT: return vbg_occupancy_split(prec, tile, ta, tb);
    case FAM_SPLIT_PK: return vbg_occupancy_split_pk(prec, tile);
    case FAM_BF16: return vbg_occupancy_bf16(prec, tile, ta, tb);
    default: return 1;
  }
}

int kernel_rows(const HostPlan& hp) {
  return (has_prepass(hp) ? 1 : 0) + (has_pack(hp) ? 1 : 0) + (int)hp.launches.size() +
         (hp.red_chunks.empty() ? 0 : 1);
}

void build_image(const HostPlan& hp, ImageLayout& lay, std::vector<char>& host) {
  const size_t B = (size_t)std::max(1, hp.batch);
  size_t off = 0;
  lay.desc = off;
  off = align256(off + sizeof(vbg_mat_desc) * B);
  lay.scales = off;
  off = align256(off + sizeof(float) * 4 * B);
  lay.amax = off;
  off = align256(off + sizeof(unsigned) * 2 * B);
  lay.chunks = off;
  off = align256(off + sizeof(vbg_chunk) * hp.chunks.size());
  lay.red = off;
  off = align256(off + sizeof(vbg_chunk) * hp.red_chunks.size());
  lay.wsoff = off;
  off = align256(off + sizeof(int64_t) * B);
  lay.pkoff = off;
  off = align256(off + sizeof(int64_t) * hp.pk_off.size());
  lay.jobs = off;
  off = align256(off + sizeof(vbg_pack_job) * hp.pack_jobs.size());
  lay.items.clear();
  lay.qbegin.clear();
  lay.counters.clear();
  for (const auto& L : hp.launches) {
    lay.items.push_back(off);
    off = align256(off + sizeof(vbg_item) * L.items.size());
    lay.qbegin.push_back(off);
    off = align256(off + sizeof(int) * L.q_begin.size());
    lay.counters.push_back(off);
    off = align256(off + sizeof(unsigned) * (hp.nq + 1));
  }
  lay.bytes = off;
  host.assign(off, 0);
  char* h = host.data();
  if (hp.batch) std::memcpy(h + lay.desc, hp.desc.data(), sizeof(vbg_mat_desc) * hp.batch);
  if (!hp.chunks.empty()) std::memcpy(h + lay.chunks, hp.chunks.data(), sizeof(vbg_chunk) * hp.chunks.size());
  if (!hp.red_chunks.empty())
    std::memcpy(h + lay.red, hp.red_chunks.data(), sizeof(vbg_chunk) * hp.red_chunks.size());
  if (hp.batch) std::memcpy(h + lay.wsoff, hp.ws_off.data(), sizeof(int64_t) * hp.batch);
  if (!hp.pk_off.empty()) std::memcpy(h + lay.pkoff, hp.pk_off.data(), sizeof(int64_t) * hp.pk_off.size());
  if (!hp.pack_jobs.empty())
    std::memcpy(h + lay.jobs, hp.pack_jobs.data(), sizeof(vbg_pack_job) * hp.pack_jobs.size());
  for (size_t l = 0; l < hp.launches.size(); ++l) {
    const auto& L = hp.launches[l];
    std::memcpy(h + lay.items[l], L.items.data(), sizeof(vbg_item) * L.items.size());
    std::memcpy(h + lay.qbegin[l], L.q_begin.data(), sizeof(int) * L.q_begin.size());
  }
}

void view_image(const HostPlan& hp, const ImageLayout& lay, char* base, DeviceView& v) {
  v.desc = (vbg_mat_desc*)(base + lay.desc);
  v.scales = (float*)(base + lay.scales);
  v.amax = (unsigned*)(base + lay.amax);
  v.chunks = (vbg_chunk*)(base + lay.chunks);
  v.red = (vbg_chunk*)(base + lay.red);
  v.ws_off = (int64_t*)(base + lay.wsoff);
  v.pk_off = (int64_t*)(base + lay.pkoff);
  v.jobs = (vbg_pack_job*)(base + lay.jobs);
  v.launches.clear();
  for (size_t l = 0; l < hp.launches.size(); ++l)
    v.launches.push_back({(vbg_item*)(base + lay.items[l]), (int*)(base + lay.qbegin[l]),
                          (unsigned*)(base + lay.counters[l])});
}

int issue_run(const HostPlan& hp, const DeviceView& v, float alpha, const float* const* dA,
              const float* const* dB, float beta, float* const* dC, hipStream_t s, hipEvent_t* ev,
              std::string& err) {
  int row = 0;
  if (ev) {
    hipError_t e = hipEventRecord(ev[0], s);
    if (e != hipSuccess) {
      err = std::string("hipEventRecord: ") + hipGetErrorString(e);
      return VBG_ERR_HIP;
    }
  }
  auto mark = [&]() -> int {
    if (!ev) return VBG_OK;
    hipError_t e = hipEventRecord(ev[++row], s);
    if (e == hipSuccess) return VBG_OK;
    err = std::string("hipEventRecord: ") + hipGetErrorString(e);
    return VBG_ERR_HIP;
  };
  if (has_prepass(hp)) {
    if (vbg_launch_scales(hp.prec, hp.batch, hp.ta, hp.tb, v.desc, v.chunks, (int)hp.chunks.size(), dA,
                          dB, v.amax, v.scales, s)) {
      err = "scale pre-pass launch: " + last_hip();
      return VBG_ERR_HIP;
    }
    if (int rc = mark()) return rc;
  }
  if (has_pack(hp)) {
    if (vbg_launch_pack(hp.prec, hp.ta, hp.tb, v.jobs, (int)hp.pack_jobs.size(), v.desc, dA, dB, v.scales,
                        v.pk, v.pk_off, s)) {
      err = "operand pack launch: " + last_hip();
      return VBG_ERR_HIP;
    }
    if (int rc = mark()) return rc;
  }
  for (size_t l = 0; l < hp.launches.size(); ++l) {
    const Launch& L = hp.launches[l];
    vbg_args a;
    a.desc = v.desc;
    a.items = v.launches[l].items;
    a.q_begin = v.launches[l].q_begin;
    a.counters = v.launches[l].counters;
    a.A = dA;
    a.B = dB;
    a.C = dC;
    a.ws = v.ws;
    a.ws_off = v.ws_off;
    a.scales = v.scales;
    a.pk = L.family == FAM_SPLIT_PK ? v.pk : nullptr;
    a.pk_off = v.pk_off;
    a.alpha = alpha;
    a.beta = beta;
    a.nq = hp.nq;
    // a launch with at most two items per workgroup spends a large share of
    // its time in the final walk over the queues: look at all of them at once
    a.look_all = (int64_t)L.items.size() <= 2 * (int64_t)L.grid ? 1 : 0;
    a.trans_a = hp.ta;
    a.trans_b = hp.tb;
    a.prec = hp.prec;
    int r;
    switch (L.family) {
      case FAM_FP32: r = vbg_launch_fp32(L.tile, hp.ta, hp.tb, &a, L.grid, s); break;
      case FAM_SEQ: r = vbg_launch_seq(hp.ta, hp.tb, &a, L.grid, s); break;
      case FAM_SPLIT: r = vbg_launch_split(L.tile, hp.ta, hp.tb, &a, L.grid, s); break;
      case FAM_SPLIT_PK: r = vbg_launch_split_pk(L.tile, &a, L.grid, s); break;
      case FAM_BF16: r = vbg_launch_bf16(L.tile, hp.ta, hp.tb, &a, L.grid, s); break;
      default: r = -1;
    }
    if (r == -1) {
      err = "tile/precision combination not built";
      return VBG_ERR_NOT_SUPPORTED;
    }
    if (r != 0) {
      err = "kernel launch: " + last_hip();
      return VBG_ERR_HIP;
    }
    if (int rc = mark()) return rc;
  }
  if (!hp.red_chunks.empty()) {
    if (vbg_launch_splitk_reduce(v.red, (int)hp.red_chunks.size(), v.desc, v.ws, v.ws_off, dC, alpha, beta,
                                 hp.prec == VBG_PREC_BF16_OUT_BF16, s)) {
      err = "split-K reduction launch: " + last_hip();
      return VBG_ERR_HIP;
    }
    if (int rc = mark()) return rc;
  }
  return VBG_OK;
}

}  // namespace vbg
