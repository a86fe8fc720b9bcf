"""The reference-protocol benchmark driver on the GPU: constant-data known
answer C_i = (i+3) + 11 K_i (i+1)(i+2) for the reference planner's choice and
for the library's own planner (needs an MI355X)."""
import sys

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B,policy,prec", [(64, "reference", "fp32"), (256, "reference", "fp32_seq"),
                                           (256, "auto", "fp32"), (128, "wang", "fp32")])
def test_driver_known_answer(bg, gpu, capsys, B, policy, prec):
    from batchgemm_amd import driver
    print(f"[driver] B={B} {policy} {prec}", file=sys.stderr, flush=True)
    rc = driver.main([str(B), "6000", "6000", "--policy", policy, "--prec", prec, "--verify"])
    out = capsys.readouterr().out.splitlines()
    assert rc == 0
    assert out[0].startswith("time======") and out[0].endswith("========N_RUNS====10=====")
    gflop = float(out[1].split()[0].split("==")[1])
    assert abs(gflop - driver.reference_gflop(*__import__("batchgemm_amd").workloads.read_dataset(
        __import__("batchgemm_amd").workloads.dataset_path(), B))) < 1e-6
