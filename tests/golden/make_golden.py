"""Writes the committed fixtures under tests/golden/.

1. reference_graphs/: the edge-list data files the reference ships as test
   inputs (04_test_graphs/TEST/*), copied byte for byte (data, not code).
2. c1_uniform.json: SHA-256 of every CSR array of the C1 workload (uniform
   N=65,536, E=1,048,576, seed 27491095) and sampled FP32 ranks after 20
   sweeps, as produced by the oracle, for regression of the GPU path.

Run where the reference tree is mounted (it reads /root/reference); the GPU
tests only read the committed output.
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/04_test_graphs/TEST"


def copy_reference_graphs():
    out = os.path.join(HERE, "reference_graphs")
    for d in sorted(os.listdir(REF)):
        for f in sorted(os.listdir(os.path.join(REF, d))):
            dst = os.path.join(out, d)
            os.makedirs(dst, exist_ok=True)
            shutil.copyfile(os.path.join(REF, d, f), os.path.join(dst, f))


def c1_golden():
    import accelgraph_express_amd as ag
    import oracle
    src, dst = ag.gen_uniform(27491095, 65536, 1048576)
    csr = oracle.build_csr(src, dst, 65536)
    ranks, it, err, psum, act, errs = oracle.pagerank_pull_f32(csr, 20, 0.0)
    rng = np.random.RandomState(7)
    sample = np.sort(rng.choice(65536, 64, replace=False))
    g = {
        "workload": "uniform N=65536 E=1048576 seed=27491095, sort 0, 20 sweeps, epsilon 0",
        "input_sha256": {"src": hashlib.sha256(src.tobytes()).hexdigest(),
                         "dst": hashlib.sha256(dst.tobytes()).hexdigest()},
        "csr_sha256": {k: hashlib.sha256(np.ascontiguousarray(csr[k]).tobytes()).hexdigest()
                       for k in oracle.CSR_KEYS},
        "iterations": it,
        "error_total": err,
        "pr_sum": psum,
        "rank_sample_index": sample.tolist(),
        "rank_sample_f32_hex": [float(ranks[i]).hex() for i in sample],
    }
    with open(os.path.join(HERE, "c1_uniform.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    copy_reference_graphs()
    c1_golden()
