// vbg_run.h -- a host plan's device image (descriptors, work map, queues,
// counters, scales, chunks) and the kernel sequence of one run, shared by
// the plan objects of the C ABI (vbg_api.cpp, device memory they allocate)
// and the entry point that allocates nothing (vbg_static.hip, a static
// device array).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vbg_planner.h"

namespace vbg {

// Byte offsets of the sections of a plan's device image (256-byte aligned).
struct ImageLayout {
  size_t desc = 0, scales = 0, amax = 0, chunks = 0, red = 0, wsoff = 0, pkoff = 0, jobs = 0;
  std::vector<size_t> items, qbegin, counters;  // per launch
  size_t bytes = 0;                             // total
};

// Device addresses of a plan's image, split-K workspace and packed operands.
struct DeviceView {
  vbg_mat_desc* desc = nullptr;
  float* scales = nullptr;
  unsigned* amax = nullptr;
  vbg_chunk* chunks = nullptr;
  vbg_chunk* red = nullptr;
  int64_t* ws_off = nullptr;
  int64_t* pk_off = nullptr;
  vbg_pack_job* jobs = nullptr;
  float* ws = nullptr;
  char* pk = nullptr;
  struct Launch {
    vbg_item* items;
    int* q_begin;
    unsigned* counters;
  };
  std::vector<Launch> launches;
};

// Lays out hp's image and fills `host` with it (queue counters zero).
void build_image(const HostPlan& hp, ImageLayout& lay, std::vector<char>& host);

// Device addresses of an image uploaded at `base`; ws and pk are set apart.
void view_image(const HostPlan& hp, const ImageLayout& lay, char* base, DeviceView& v);

// Kernels one run issues, in order: [scale pre-pass], [operand pack], one per
// launch, [split-K reduction].
int kernel_rows(const HostPlan& hp);

// Issues one run on `stream`: every kernel of the plan, in order.  `ev`
// (kernel_rows + 1 events, or null) is recorded before the first kernel and
// after each.  Returns a vbg_status, with `err` set on failure.
int issue_run(const HostPlan& hp, const DeviceView& v, float alpha, const float* const* dA,
              const float* const* dB, float beta, float* const* dC, hipStream_t stream,
              hipEvent_t* ev, std::string& err);

// Resident workgroups per CU for the kernel family that runs `tile`.
int occupancy(int family, int prec, int tile, bool ta, bool tb);

// Sets the calling thread's error detail (vbg_last_error).
void set_last_error(const std::string& msg);

}  // namespace vbg
