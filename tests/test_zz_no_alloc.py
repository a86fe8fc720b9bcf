"""The entry point that allocates nothing (vbg_sgemm_vbatched_noalloc, needs
an MI355X): through the C ABI, two calls in a row (the second on a smaller
batch, so it reuses the static scratch over the first call's contents),
against the one-shot entry point that plans and allocates
(vbg_sgemm_vbatched_oneshot) on the same inputs; and a batch whose work map
exceeds the scratch limit is refused with VBG_ERR_INVALID_ARG before any
work is issued."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(bg, fn, vb, prec, alpha, beta):
    import torch
    L = bg.lib()
    ip = ctypes.POINTER(ctypes.c_int)
    arrs = [np.ascontiguousarray(x, np.int32) for x in (vb.M, vb.N, vb.K, vb.lda, vb.ldb, vb.ldc)]
    rc = getattr(L, fn)(vb.batch, *[a.ctypes.data_as(ip) for a in arrs], vb.transA.encode(),
                        vb.transB.encode(), bg.PREC[prec], alpha, vb.dA.data_ptr(), vb.dB.data_ptr(),
                        beta, vb.dC.data_ptr(), torch.cuda.current_stream().cuda_stream)
    bg._lib.check(rc, fn)
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp8"])
def test_noalloc_matches_oneshot_twice(bg, gpu, prec):
    import torch
    from batchgemm_amd import workloads as w
    # first a larger batch that splits K (one 64 x 64 x 6000 matrix among
    # small ones: few items, so its one item would be the launch's critical
    # path), then a smaller one on the scratch the first call left behind
    cases = [
        (list(w.shapes_uniform(300, 1, 300, seed=3)), "NT"),
        ([[64, 5, 130], [64, 17, 1], [6000, 9, 40]], "NN"),
    ]
    for i, ((M, N, K), trans) in enumerate(cases):
        print(f"[noalloc] {prec} call {i}: batch {len(M)} trans {trans}", flush=True)
        vb = w.VBatch(M, N, K, trans[0], trans[1]).fill("normal", seed=40 + i, c_dist="uniform")
        C0 = vb.C.clone()
        _call(bg, "vbg_sgemm_vbatched_oneshot", vb, prec, 1.5, 0.5)
        want = vb.C.clone()
        vb.C.copy_(C0)
        _call(bg, "vbg_sgemm_vbatched_noalloc", vb, prec, 1.5, 0.5)
        got = vb.C.clone()
        if torch.equal(got.view(torch.int32), want.view(torch.int32)):
            continue
        # same arithmetic, possibly other split-K choices (the static workspace
        # is smaller than a plan's): FP32 summation-order differences only
        err = float(torch.linalg.norm((got - want).double()) / torch.linalg.norm(want.double()))
        print(f"  not bit-identical: rel diff {err:.3g}", flush=True)
        assert err <= 1e-6


def test_noalloc_rejects_oversized_work_map(bg, gpu):
    """300,000 1x1x1 matrices need about 30 MB of work map (> 16 MiB):
    refused with a clear error, nothing launched (the pointer arrays are
    never read)."""
    import torch
    L = bg.lib()
    B = 300000
    ones = np.ones(B, np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    dummy = torch.zeros(B, dtype=torch.int64, device="cuda")
    rc = L.vbg_sgemm_vbatched_noalloc(B, *[ones.ctypes.data_as(ip)] * 6, b"N", b"N", 0, 1.0,
                                      dummy.data_ptr(), dummy.data_ptr(), 0.0, dummy.data_ptr(),
                                      torch.cuda.current_stream().cuda_stream)
    assert rc == bg._lib.VBG_ERR_INVALID_ARG
    assert b"too large" in L.vbg_last_error()
