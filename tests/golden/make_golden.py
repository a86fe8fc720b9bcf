"""Writes the golden fixtures of tests/golden from the reference tree:

- reference_logs.json: numbers read from the reference's run logs, with the
  file:line they come from (the oracle's dataset ingest and FLOP accounting
  are compared against them);
- data_MN_K_512_128 and data_MN_K_512_128_sort: the reference's dataset and
  its sorted first 256 lines, as shipped (ROCm/data/; the sorted file is the
  output of ROCm/data/sort.cpp, which pins the Wang ordering).  The dataset
  is also copied to batchgemm_amd/data/, the benchmark driver's default input.

Run in the build container (the reference tree is not present elsewhere):
    python tests/golden/make_golden.py
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_gflops(path):
    rows = []
    with open(os.path.join(REF, path)) as f:
        for ln, line in enumerate(f, 1):
            m = re.match(r"gflops==([0-9.]+)\s+([0-9.]+)", line)
            if m:
                rows.append({"line": ln, "gflop": m.group(1), "gflops_per_s": float(m.group(2))})
    return rows


def parse_verify(path):
    out = []
    with open(os.path.join(REF, path)) as f:
        for ln, line in enumerate(f, 1):
            m = re.match(r"A=([0-9.]+),B=([0-9.]+),C=([0-9.]+)", line)
            if m:
                out.append({"line": ln, "A": float(m.group(1)), "B": float(m.group(2)),
                            "C": float(m.group(3))})
    return out


def main():
    batches = [8, 16, 32, 64, 128, 256]
    doc = {
        "batches": batches,
        "rocm_framework": {"source": "ROCm/framework/log-test-verify3",
                           "rows": parse_gflops("ROCm/framework/log-test-verify3")},
        "cuda_framework": {"source": "CUDA/framework/log-test-verify3",
                           "rows": parse_gflops("CUDA/framework/log-test-verify3")},
        "rocm_hipblas_loop": {"source": "ROCm/default/log-default",
                              "rows": parse_gflops("ROCm/default/log-default")},
        "wang_log_verify": {"source": "ROCm/tiling/log-verify",
                            "runs": 1001,
                            "rows": parse_verify("ROCm/tiling/log-verify")},
    }
    with open(os.path.join(HERE, "reference_logs.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    for name in ("data_MN_K_512_128", "data_MN_K_512_128_sort"):
        shutil.copyfile(os.path.join(REF, "ROCm", "data", name), os.path.join(HERE, name))
    # the benchmark driver's default dataset, shipped with the package
    pkg_data = os.path.join(os.path.dirname(os.path.dirname(HERE)), "batchgemm_amd", "data")
    os.makedirs(pkg_data, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "ROCm", "data", "data_MN_K_512_128"),
                    os.path.join(pkg_data, "data_MN_K_512_128"))


if __name__ == "__main__":
    main()
