#!/usr/bin/env python3
"""Benchmark: LP edges/s per iteration on a CGR-compressed graph (MI355X).

Workload (default, "rmat22"): R-MAT scale 22 (4,194,304 vertices), edge
factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 1, symmetrised, deduplicated,
self loops removed, BFS-relabelled from vertex 0; CGR threshold 0; labels
start as the vertex id; count-min sketch (depth 4, width 1024, seed 0) for
degree > 4096.  One step = one synchronous fused decode+LP iteration (NOVA)
over every vertex, inputs resident in HBM.  The CGLP two-kernel method is
timed on the same graph and reported beside it.

Prints one JSON line.  `value` = m * steps / elapsed (arcs per second for
K = 1); the roofline object describes the dominant kernel family, timed
with HIP events on the engine's stream; cpu_baseline times the CPU oracle's
decode+LP step on the host cores over a bounded random sample of the same
graph's vertices (about 20 s of CPU work).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "rmat22": dict(gen="rmat", scale=22, ef=16, seed=1, T=0, K=1,
                   cms=dict(cms_thresh=4096, cms_depth=4, cms_width=1024, cms_seed=0)),
    "lattice1024": dict(gen="lattice", side=1024, seed=0, T=0, K=1, cms=dict()),
}


def build_graph(w):
    import lpa_amd as L
    if w["gen"] == "rmat":
        g = L.gen_rmat(w["scale"], w["ef"], 0.57, 0.19, 0.19, seed=w["seed"], bfs=True)
    else:
        g = L.gen_lattice(w["side"], 0.10, 0.005, seed=w["seed"])
    return g, g.encode(w["T"], 128)


def alg_bytes(vertices, arcs, bits, csr_arcs, K):
    """algorithmic HBM bytes of one pass: compressed stream once, 8 B offset
    per compressed vertex, per label vector 4 B gather per arc plus 4 B read
    and 4 B write per vertex; small-vertex CSR: 4 B per arc + 8 B offset"""
    n_c = vertices  # compressed or CSR vertices of this family
    return bits / 8.0 + 8.0 * n_c + K * (4.0 * arcs + 8.0 * vertices) + 4.0 * csr_arcs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cglp", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import lpa_amd as L
    w = WORKLOADS[args.workload]
    t0 = time.time()
    g, cgr = build_graph(w)
    m = int(g.m)
    print(f"[bench] graph n={g.n} m={m} bits={cgr.nbits} built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    e = L.Engine(0)
    e.upload(cgr)
    e.config(K=w["K"], **w["cms"])
    e.schedule_degree()
    stats = e.stats()
    e.labels_init()
    for _ in range(args.warmup):
        e.nova_iterate()
    t = time.perf_counter()
    ms, changed = e.time_iterations(args.steps, "nova")
    wall = time.perf_counter() - t
    kt = e.kernel_times()
    ms_per_step = float(np.mean(ms))
    value = m * w["K"] * args.steps / (float(np.sum(ms)) / 1e3)

    # dominant kernel family and its algorithmic bytes per step
    dom = max((k for k in kt if stats[k]["vertices"]), key=lambda k: kt[k])
    s = stats[dom]
    dom_bytes = alg_bytes(s["vertices"], s["arcs"], s["bits"], s["csr_arcs"], w["K"])
    achieved = dom_bytes / (kt[dom] / 1e3) / 1e9
    it_bytes = (cgr.nbits / 8 + 8 * (cgr.n_c + 1) + w["K"] * (4 * m + 8 * g.n))
    line = {
        "metric": "LP edges/sec per iteration on CGR graph; learned vs degree-binned schedule",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": int(g.n), "m": m, "cgr_bits": int(cgr.nbits),
                   "K": w["K"], "method": "nova-fused", "schedule": "degree", **w["cms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": dom, "kernel_ms_per_step": kt[dom], "alg_bytes_per_step": dom_bytes},
        "iteration_roofline": {"alg_bytes": it_bytes, "achieved_GBs": it_bytes / (ms_per_step / 1e3) / 1e9,
                               "frac": it_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": kt,
        "bins": {k: v for k, v in stats.items() if v["vertices"]},
        "changed": [int(x) for x in changed],
        "device": e.info(),
        "wall_s": wall,
    }
    if not args.no_cglp:
        e.labels_init()
        ms2, _ = e.time_iterations(max(2, args.steps // 2), "cglp")
        line["cglp_ms_per_step"] = float(np.mean(ms2))
        line["cglp_edges_per_s"] = m * w["K"] / (float(np.mean(ms2)) / 1e3)
        line["nova_over_cglp"] = line["cglp_ms_per_step"] / ms_per_step
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_run(cgr, g.degrees(), w)
    print(json.dumps(line))


def cpu_baseline_run(cgr, deg, w, budget_s=20.0):
    """times the CPU oracle's decode+LP step on the host cores over a
    bounded prefix of stream positions (whole iterations when they fit)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cores = min(len(os.sched_getaffinity(0)), 16)
    view = O.CgrView(np.frombuffer(cgr.stream_bytes(), np.uint8), cgr.nbits, cgr.offsets.copy(),
                     cgr.vertices.copy(), cgr.small_offsets.copy(), cgr.small_nbrs.copy())
    c = O.cfg(K=w["K"], nthreads=cores, **w["cms"])
    lab = np.repeat(np.arange(cgr.n, dtype=np.uint32), w["K"])
    out = lab.copy()
    rng = np.random.default_rng(0)
    pos_all = rng.permutation(cgr.n).astype(np.uint64)
    n_probe = min(cgr.n, 20000)
    t = time.perf_counter()
    O.step_cgr_subset(view, lab, pos_all[:n_probe], out, c)
    probe_s = time.perf_counter() - t
    n_take = int(min(cgr.n, max(n_probe, n_probe * budget_s / max(probe_s, 1e-3))))
    sample = np.sort(pos_all[:n_take])
    t = time.perf_counter()
    O.step_cgr_subset(view, lab, sample, out, c)
    el = time.perf_counter() - t
    arcs = int(deg[cgr.vertices[sample.astype(np.int64)]].sum())
    return {"value": arcs * w["K"] / el, "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": f"oracle decode+LP step over {n_take} of {cgr.n} random stream positions "
                      f"({arcs} arcs), {el:.1f} s"}


if __name__ == "__main__":
    main()
