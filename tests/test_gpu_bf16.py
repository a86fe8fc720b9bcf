"""BF16-operand plans (the tile family under the grouped passes) on ragged
shapes (needs an MI355X).

M, N and K that are not multiples of 4 or 8, every transpose pair, and leading
dimensions padded past the row count (multiples of 8, as the family
requires): the loaders' whole-vector and straddling paths, for operands stored
k-contiguous and mn-contiguous, on the default tiles and on the 256x128
and 256x256 tiles.  The reference is FP64 on the same BF16
values; the kernels form exact BF16 products and accumulate in FP32, so the
error is bounded by about K * 2^-24 of sum |a||b| (checked at 1e-5 of that
sum, plus the final FP32 or BF16 rounding of C)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


M = [5, 64, 130, 67, 1, 200, 33, 129]
N = [7, 64, 129, 250, 3, 1, 90, 131]
K = [3, 64, 77, 300, 1, 129, 520, 66]


def up8(x):
    return (x + 7) // 8 * 8


@pytest.mark.parametrize("tile", ["auto", "256x128", "256x256"])
@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("prec", ["bf16", "bf16_out_bf16"])
def test_bf16_plan_ragged_vs_fp64(bg, gpu, trans, prec, tile):
    import torch
    from batchgemm_amd.api import pointer_array
    say(f"[bf16 plan] {prec} trans={trans} tile={tile}")
    ta, tb = trans[0] == "T", trans[1] == "T"
    g = torch.Generator(device="cuda").manual_seed(7)
    out_dt = torch.bfloat16 if prec == "bf16_out_bf16" else torch.float32
    As, Bs, Cs, lda, ldb, ldc = [], [], [], [], [], []
    for m, n, k in zip(M, N, K):
        ar, ac = (k, m) if ta else (m, k)
        br, bc = (n, k) if tb else (k, n)
        la, lb = up8(ar) + 8, up8(br) + 8
        As.append(torch.randn(ac, la, generator=g, device="cuda").to(torch.bfloat16))
        Bs.append(torch.randn(bc, lb, generator=g, device="cuda").to(torch.bfloat16))
        Cs.append(torch.randn(n, m, generator=g, device="cuda").to(out_dt))
        lda.append(la); ldb.append(lb); ldc.append(m)
    C0 = [c.clone() for c in Cs]
    opts = {} if tile == "auto" else {"policy": "fixed", "fixed_tile": tile}
    plan = bg.Plan(M, N, K, lda, ldb, ldc, trans[0], trans[1], prec=prec, **opts)
    alpha, beta = 1.5, 0.5
    plan.run(pointer_array(As), pointer_array(Bs), pointer_array(Cs), alpha, beta)
    torch.cuda.synchronize()
    for b, (m, n, k) in enumerate(zip(M, N, K)):
        ar = k if ta else m
        br = n if tb else k
        A = As[b][:, :ar].double().T.cpu().numpy()      # stored (ar x ac), column-major
        B = Bs[b][:, :br].double().T.cpu().numpy()
        opA = A.T if ta else A
        opB = B.T if tb else B
        c0 = C0[b].double().T.cpu().numpy()
        ref = alpha * opA @ opB + beta * c0
        bound = abs(alpha) * np.abs(opA) @ np.abs(opB) + abs(beta) * np.abs(c0)
        got = Cs[b].double().T.cpu().numpy()
        out_round = 2.0 ** -8 if prec == "bf16_out_bf16" else 2.0 ** -23
        err = np.abs(got - ref)
        ok = err <= 1e-5 * bound + out_round * np.abs(ref) + 1e-30
        assert ok.all(), (b, (m, n, k), float((err / (bound + 1e-30)).max()))
