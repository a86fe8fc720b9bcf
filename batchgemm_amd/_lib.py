"""ctypes binding of libvbgemm.so (the C ABI declared in include/vbgemm.h).

The library is built in-tree (``make -C batchgemm_amd``) and loaded from this
directory; there is no fallback: a missing library raises at first use.
torch is imported first so that its bundled HIP runtime (same soname,
libamdhip64.so.7) is the one the library binds to -- one runtime per process.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
# VBG_LIBRARY selects another build of the same ABI (A/B comparisons of kernels).
LIB_PATH = os.environ.get("VBG_LIBRARY") or os.path.join(_HERE, "libvbgemm.so")

VBG_OK = 0
VBG_ERR_INVALID_ARG = -1
VBG_ERR_HIP = -2
VBG_ERR_NOT_SUPPORTED = -3
VBG_ERR_ALLOC = -4
VBG_ERR_PLAN_MISMATCH = -5
VBG_ERR_UNDEFINED = -6

PREC = {"fp32": 0, "bf16x3": 1, "fp16": 2, "fp8": 3, "fp32_seq": 4, "bf16": 5, "bf16_out_bf16": 6}
POLICY = {"auto": 0, "reference": 1, "fixed": 2, "wang": 3, "li": 4}
TILES = {"16x16": 0, "32x32": 1, "64x64": 2, "128x64": 3, "64x128": 4, "128x128": 5, "256x128": 6,
         "16x128": 7, "128x16": 8, "256x256": 9}
TILE_SHAPES = [(16, 16), (32, 32), (64, 64), (128, 64), (64, 128), (128, 128), (256, 128), (16, 128),
               (128, 16), (256, 256)]
TILE_MIXED = 10  # launch tile id of a mixed launch (query 7): the shape is per item

# Every symbol include/vbgemm.h declares (checked by the CPU test-suite).
EXPORTS = [
    "vbg_abi_version", "vbg_status_string", "vbg_last_error", "vbg_opts_default",
    "vbg_plan_create", "vbg_plan_destroy", "vbg_sgemm_vbatched", "vbg_sgemm_vbatched_oneshot",
    "vbg_plan_query", "vbg_ref_plan", "vbg_grouped_plan_create", "vbg_grouped_plan_destroy",
    "vbg_grouped_gemm_fwd", "vbg_grouped_gemm_bwd", "vbg_plan_launch_times",
    "vbg_sgemm_vbatched_noalloc",
]


class VbgError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str):
        self.code = code
        super().__init__(f"{where}: {code} ({detail})")


class vbg_opts(ctypes.Structure):
    _fields_ = [
        ("policy", ctypes.c_int),
        ("wavefront_thres", ctypes.c_int),
        ("tlp_thres", ctypes.c_int),
        ("fixed_tile", ctypes.c_int),
        ("tile_mask", ctypes.c_uint),
        ("big_tile_min_mn", ctypes.c_int),
        ("split_k_min", ctypes.c_int),
        ("grid_wgs", ctypes.c_int),
        ("xcd_queues", ctypes.c_int),
        ("host_only", ctypes.c_int),
        ("num_cus", ctypes.c_int),
        ("force_ksplit", ctypes.c_int),
        ("tile_group", ctypes.c_int),
        ("edge_tiles", ctypes.c_int),
        ("timing_slots", ctypes.c_int),
        ("mixed_tiles", ctypes.c_int),
        ("consolidate", ctypes.c_int),
        ("pack_operands", ctypes.c_int),
        ("item_fixed_x10", ctypes.c_int),
        ("small_tile_eff_pct", ctypes.c_int),
    ]


_lib = None


def lib():
    """Load libvbgemm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C batchgemm_amd` "
            "(or __graft_entry__.build()); there is no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    L.vbg_abi_version.restype = ctypes.c_int
    L.vbg_status_string.argtypes = [ctypes.c_int]
    L.vbg_status_string.restype = ctypes.c_char_p
    L.vbg_last_error.restype = ctypes.c_char_p
    L.vbg_opts_default.argtypes = [ctypes.POINTER(vbg_opts)]
    L.vbg_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ip, ip, ip, ip, ip,
                                  ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                  ctypes.POINTER(vbg_opts)]
    L.vbg_plan_destroy.argtypes = [vp]
    L.vbg_sgemm_vbatched.argtypes = [vp, ctypes.c_float, vp, vp, ctypes.c_float, vp, vp]
    L.vbg_sgemm_vbatched_oneshot.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip, ip, ctypes.c_char,
                                             ctypes.c_char, ctypes.c_int, ctypes.c_float, vp, vp,
                                             ctypes.c_float, vp, vp]
    L.vbg_sgemm_vbatched_noalloc.argtypes = L.vbg_sgemm_vbatched_oneshot.argtypes
    L.vbg_plan_query.argtypes = [vp, ctypes.c_int, ip, ctypes.c_int]
    L.vbg_plan_launch_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    L.vbg_ref_plan.argtypes = [ctypes.c_int, ip, ip, ip, ctypes.c_int, ip, ip, ip,
                               ctypes.POINTER(ctypes.c_int64)]
    L.vbg_grouped_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(vbg_opts)]
    L.vbg_grouped_plan_destroy.argtypes = [vp]
    L.vbg_grouped_gemm_fwd.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
    L.vbg_grouped_gemm_bwd.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
    _lib = L
    return L


def check(rc: int, where: str) -> int:
    if rc < 0:
        L = lib()
        raise VbgError(rc, where, f"{L.vbg_status_string(rc).decode()}: {L.vbg_last_error().decode()}")
    return rc


def default_opts(**kw) -> vbg_opts:
    o = vbg_opts()
    check(lib().vbg_opts_default(ctypes.byref(o)), "vbg_opts_default")
    for k, v in kw.items():
        if k == "policy" and isinstance(v, str):
            v = POLICY[v]
        if k == "fixed_tile" and isinstance(v, str):
            v = TILES[v]
        setattr(o, k, int(v))
    return o
