"""batchgemm_amd -- load-balanced variable-size batched GEMM for MI355X (gfx950).

Hand-written gfx950 HIP kernels behind a C ABI (include/vbgemm.h, built as
batchgemm_amd/libvbgemm.so); this package is the host-side mirror of that ABI
for Python callers.
"""
from ._lib import LIB_PATH, PREC, POLICY, TILES, TILE_SHAPES, VbgError, lib  # noqa: F401
from .api import Plan, colmajor_empty, ld_of, pointer_array, ref_plan, sgemm_vbatched  # noqa: F401

__all__ = ["Plan", "sgemm_vbatched", "ref_plan", "colmajor_empty", "ld_of", "pointer_array",
           "VbgError", "lib", "PREC", "POLICY", "TILES", "TILE_SHAPES", "LIB_PATH"]
