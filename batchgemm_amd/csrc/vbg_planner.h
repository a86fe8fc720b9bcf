// vbg_planner.h -- host planner: per-matrix tile choice, split-K, and the
// XCD-aware persistent work map.  Pure C++ (no HIP), so it runs and is tested
// on machines without a GPU.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "vbg_internal.h"

struct vbg_opts;

namespace vbg {

// Kernel family a tile id runs on for a precision (one persistent launch per
// family used by a plan).
enum Family { FAM_FP32 = 0, FAM_SEQ = 1, FAM_SPLIT = 2, FAM_BF16 = 3, FAM_SPLIT_PK = 4 };

struct Launch {
  int tile = 0;                 // tile id (T16 .. T128), or T_MIXED (tile per item)
  int family = FAM_FP32;
  std::vector<vbg_item> items;  // in queue order
  std::vector<int> q_begin;     // nq + 1
  int grid = 0;                 // persistent workgroups
  double cost = 0.0;            // modelled cost (for ordering launches)
};

struct HostPlan {
  int batch = 0;
  bool ta = false, tb = false;
  int prec = 0;
  std::vector<vbg_mat_desc> desc;
  std::vector<int> tile_of;      // tile id per matrix (-1: no work)
  std::vector<Launch> launches;  // in launch order
  int nq = 1;
  int64_t ws_floats = 0;         // split-K workspace
  std::vector<int64_t> ws_off;   // per matrix (floats), -1 if not split
  std::vector<vbg_chunk> chunks; // amax pre-pass slices (FP16 / FP8)
  std::vector<vbg_chunk> red_chunks;  // split-K reduction slices
  // Packed operands (FAM_SPLIT_PK): byte offsets of packed A and B per matrix
  // (2 * batch), the workspace size, and the pack pre-pass jobs.
  std::vector<int64_t> pk_off;
  int64_t pk_bytes = 0;
  std::vector<vbg_pack_job> pack_jobs;
};

// Fills `hp` from the shapes and options.  `occupancy(tile)` gives the
// resident workgroups per CU for the family that runs `tile` (1 if unknown).
// Split-K partial panels take at most `ws_cap_floats` floats of workspace in
// total (matrices that would not fit are not split).  Returns 0 or a negative
// vbg_status, with `err` set.
constexpr int64_t kDefaultWsCapFloats = (int64_t)1 << 27;  // 512 MB
int build_plan(HostPlan& hp, int batch, const int* M, const int* N, const int* K, const int* lda,
               const int* ldb, const int* ldc, bool ta, bool tb, int prec, const vbg_opts& o,
               int num_cus, int (*occupancy)(int family, int prec, int tile, bool ta, bool tb),
               std::string& err, int64_t ws_cap_floats = kDefaultWsCapFloats);

// Reference planner restated from ROCm/framework/gemm.cuh:121-229.
int ref_plan(int batch, const int* M, const int* N, const int* K, int thres, int* t_strategy,
             int* kernel_thread, int* grid, int64_t* wavefronts);

}  // namespace vbg
