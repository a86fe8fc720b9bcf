// vbg_launch.h -- host entry points of the kernel translation units, called by
// the C ABI layer (vbg_api.cpp).  Return 0 on success, -1 for an unsupported
// tile, -2 for a HIP launch error.
#pragma once
#include "vbg_internal.h"

int vbg_launch_fp32(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream);
int vbg_occupancy_fp32(int tile, int ta, int tb);
int vbg_launch_seq(int ta, int tb, const vbg_args* a, int grid, void* stream);
int vbg_occupancy_seq(int ta, int tb);

// Split-precision family (bf16x3, FP16 scaled, FP8 e4m3); a->prec selects.
int vbg_launch_split(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream);
int vbg_occupancy_split(int prec, int tile, int ta, int tb);
// The same kernels on packed operands (a->pk, a->pk_off), and the operand
// pack pre-pass that writes them (one workgroup per job; FP16 / FP8 read the
// scales of the amax pre-pass).
int vbg_launch_split_pk(int tile, const vbg_args* a, int grid, void* stream);
int vbg_occupancy_split_pk(int prec, int tile);
int vbg_launch_pack(int prec, int ta, int tb, const vbg_pack_job* jobs, int n_jobs,
                    const vbg_mat_desc* desc, const float* const* A, const float* const* B,
                    const float* scales, char* pk, const int64_t* pk_off, void* stream);
// Per-matrix amax pre-pass and scales (FP16 / FP8): zero amax, amax over the
// chunks, then scales[4b..4b+2] = {sA, sB, 1/(sA*sB)}.
int vbg_launch_scales(int prec, int batch, int ta, int tb, const vbg_mat_desc* desc,
                      const vbg_chunk* chunks, int n_chunks, const float* const* A,
                      const float* const* B, unsigned* amax, float* scales, void* stream);

// BF16-operand family (grouped passes); a->prec selects FP32 or BF16 C.
int vbg_launch_bf16(int tile, int ta, int tb, const vbg_args* a, int grid, void* stream);
int vbg_occupancy_bf16(int prec, int tile, int ta, int tb);

// Split-K reduction over the plan's reduction chunks (elements [begin, end)
// of one split matrix's M x N output in column-major order).
int vbg_launch_splitk_reduce(const vbg_chunk* chunks, int n_chunks, const vbg_mat_desc* desc,
                             const float* ws, const int64_t* ws_off, float* const* C, float alpha,
                             float beta, int out_bf16, void* stream);
