// vbg_grouped.hip -- grouped BF16 GEMM forward / backward (differentiation path).
#include <hip/hip_runtime.h>

#include "../../include/vbgemm.h"

extern "C" {
int vbg_grouped_plan_create(vbg_grouped_plan** plan, int, const int*, int, int, const vbg_opts*) {
  if (plan) *plan = nullptr;
  return VBG_ERR_NOT_SUPPORTED;
}
int vbg_grouped_plan_destroy(vbg_grouped_plan*) { return VBG_OK; }
int vbg_grouped_gemm_fwd(vbg_grouped_plan*, const void*, const void*, void*, int, void*) {
  return VBG_ERR_NOT_SUPPORTED;
}
int vbg_grouped_gemm_bwd(vbg_grouped_plan*, const void*, const void*, const void*, int, void*,
                         void*, int, void*) {
  return VBG_ERR_NOT_SUPPORTED;
}
}
