"""Benchmark: pull PageRank on R-MAT scale 22 (N=4,194,304, E=67,108,864) on one MI355X.

A step is one pageRankPullGraphCSR-equivalent call on the device-resident
graph: init + 20 fixed FP32 sweeps (epsilon 0) + normalise, output left in
HBM.  The CSR (count-sort build, bit-exact with the oracle) and the gather
layout are built once before timing and reported separately.

    value    = GTEPS = E * 20 / step seconds / 1e9
    roofline = HBM-bound sweep kernel: algorithmic bytes per sweep
               (4E + 24N + 4) / mean sweep duration (HIP events on the
               library's stream, over the timed steps) vs 8 TB/s
    cpu_baseline = the oracle's OpenMP restatement of the reference loop
               (pageRank.c:1024-1085 + the schedule(dynamic,1024) gather of
               pageRank.c:2890-2907) on the host's cores, same graph.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--scale 22]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PR GTEPS + %HBM peak (R-MAT s22); PPR queries/s + p99 latency at L1<=1e-4"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2022)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every box here has one "
              "GPU and this path is single-device.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import accelgraph_express_amd as ag

    scale, ef, iters = args.scale, args.edge_factor, args.iterations
    t = time.time()
    src, dst = ag.gen_rmat(args.seed, scale, ef)
    N, E = 1 << scale, src.size
    log(f"R-MAT s{scale} ef{ef}: N={N} E={E} generated in {time.time() - t:.1f}s")

    ctx = ag.Context(0)
    if args.copy_probe:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 268435456 bytes read + written by one device copy")
        del a, b
    d_src = torch.from_numpy(src.view(np.int32)).to("cuda")
    d_dst = torch.from_numpy(dst.view(np.int32)).to("cuda")
    torch.cuda.synchronize()
    # CSR build from HBM-resident edges (one warm build, then a timed one)
    ag.Graph.build(ctx, None, None, N, device_ptrs=(d_src.data_ptr(), d_dst.data_ptr()),
                   num_edges=E).free()
    torch.cuda.synchronize()
    t = time.perf_counter()
    g = ag.Graph.build(ctx, None, None, N, device_ptrs=(d_src.data_ptr(), d_dst.data_ptr()),
                       num_edges=E)
    torch.cuda.synchronize()
    csr_ms = (time.perf_counter() - t) * 1e3
    plan_ms = g.prepare()
    info = g.plan_info()
    del d_src, d_dst
    log(f"CSR build {csr_ms:.1f} ms, gather layout {plan_ms:.1f} ms, plan {info}")

    out = torch.empty(N, dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        g.pagerank_pull_device(out.data_ptr(), iterations=iters, epsilon=0.0)
    torch.cuda.synchronize()
    loop_ms = 0.0
    sweeps = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = g.pagerank_pull_device(out.data_ptr(), iterations=iters, epsilon=0.0)
        loop_ms += r.loop_ms
        sweeps += r.sweeps
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    gteps = E * iters / (ms_step / 1e3) / 1e9
    sweep_ms = loop_ms / max(sweeps, 1)
    bytes_sweep = 4 * E + 24 * N + 4
    achieved = bytes_sweep / (sweep_ms / 1e3) / 1e9
    log(f"step {ms_step:.3f} ms, sweep {sweep_ms * 1e3:.1f} us, {gteps:.1f} GTEPS, "
        f"{achieved:.0f} GB/s algorithmic")

    traffic = None
    tf = os.path.join(ROOT, "profiles", f"pr_sweep_rmat_s{scale}_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_sweep")

    ranks = out.cpu().numpy()
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        threads = oracle.omp_max_threads()
        log(f"CPU baseline: oracle OpenMP CSR build + {iters}-sweep pull PR, {threads} threads")
        csr_s, csr = oracle.csr_build_omp(src, dst, N, threads)
        secs, cpu_ranks, it_cpu, _ = oracle.pagerank_pull_omp(csr, iters, 0.0, threads)
        cpu = {"value": round(E * iters / secs / 1e9, 4), "unit": "GTEPS", "cores": threads,
               "kind": "port",
               "sample": f"full {iters}-sweep FP32 pull PageRank on the same R-MAT s{scale} graph "
                         f"({secs:.2f} s loop; OpenMP CSR build {csr_s:.2f} s not included)"}
        rel = np.abs(ranks.astype(np.float64) - cpu_ranks) / cpu_ranks.astype(np.float64)
        parity = {"max_rel_vs_cpu_fp32": float(rel.max()), "tolerance": 1e-5}
        log(f"CPU {cpu['value']} GTEPS; GPU vs CPU max rel {rel.max():.2e}")

    line = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT, a=0.57 b=0.19 c=0.19, vertex ids permuted)",
        "config": {"workload": f"rmat_s{scale}_ef{ef}_pull_pagerank_fp32_{iters}sweeps",
                   "num_vertices": N, "num_edges": E, "iterations": iters, "epsilon": 0.0,
                   "csr_build_ms": round(csr_ms, 2), "gather_layout_ms": round(plan_ms, 2),
                   "plan": info},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_pr_sweep",
                     "sweep_us": round(sweep_ms * 1e3, 2),
                     "algorithmic_bytes_per_sweep": bytes_sweep},
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
