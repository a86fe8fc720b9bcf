// vbg_split.hip -- split-precision family (bf16x3, FP16 scaled, FP8 e4m3).
#include "vbg_tiles.h"
#include "vbg_launch.h"

int vbg_launch_split(int, int, int, const vbg_args*, int, void*) { return -1; }
int vbg_occupancy_split(int, int, int) { return 1; }
int vbg_launch_scales(int, int, int, int, const vbg_mat_desc*, const float* const*,
                      const float* const*, unsigned*, float*, void*) {
  return -1;
}
