#!/usr/bin/env python3
"""Benchmark: LP edges/s per iteration on a CGR graph, learned vs degree-binned schedule.

Default workload "rmat27": R-MAT scale 27 (134,217,728 vertices), edge factor
32, a/b/c/d = 0.57/0.19/0.19/0.05, seed 4, symmetrised, deduplicated, self
loops removed, BFS-relabelled from vertex 0 (8.36 G arcs); CGR threshold 0;
labels start as the vertex id; count-min sketch (depth 4, width 2048, seed 0)
for degree > 8192; K = 8 label vectors.

One step = one synchronous fused decode+LP iteration (NOVA) over every vertex
and every label vector, inputs resident in HBM, timed with HIP events on the
engine's stream.  Edges/s = K * m / t_step.  Reported:
  value            fused step with K = 8 label vectors under the learned
                   schedule as judged: the cost model is trained on work
                   recorded on the 1024x1024 lattice and on R-MAT scale 22,
                   and the degree schedule judges the learned layout on the
                   same steps and stays when it is faster (config.schedule
                   names the layout that ran; learned_vs_degree gives both)
  degree_k1        degree-binned fused step, K = 1 (the plain kernel)
  degree_k8        degree-binned fused step, K = 8
  cglp_k1          the two-kernel method, K = 1
  csr_k1           comparator: LP over the uncompressed CSR (decoded once,
                   no decode in the step), K = 1
The roofline object describes the dominant kernel family of the headline run;
cpu_baseline times the CPU oracle's decode+LP step on the host cores, at the
headline's K and at K = 1, over a bounded random sample of the graph's
vertices.
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
    # R-MAT scale 22, ef 16, seed 1; CMS 4x1024 for deg > 4096; K = 8 steps
    # replayed from a captured HIP graph
    "rmat22": dict(gen="rmat", scale=22, ef=16, seed=1, T=0, K=8, graph=True,
                   cms=dict(cms_thresh=4096, cms_depth=4, cms_width=1024, cms_seed=0)),
    # 1024x1024 lattice, 10% edges dropped, 0.5% shortcuts, seed 0; exact
    "lattice1024": dict(gen="lattice", side=1024, seed=0, T=0, K=1, cms=dict()),
    # copying model: 50M vertices, mean out-degree 30, power law 2.1 capped at
    # 100k, 70% copied within +-1000, seed 2, directed; T = 32; CMS for deg > 8192
    "copy50m": dict(gen="copying", n=50_000_000, seed=2, T=32, K=1,
                    cms=dict(cms_thresh=8192, cms_depth=4, cms_width=2048, cms_seed=0)),
    # R-MAT scale 24, ef 16, seed 3; K = 64 semi-supervised vectors (1% seeds
    # from 100 classes, clamped); exact; autotuned bins, steps replayed from
    # a captured HIP graph
    "rmat24k64": dict(gen="rmat", scale=24, ef=16, seed=3, T=0, K=64, semi=True, cms=dict(), graph=True,
                      autotune=True),
    # R-MAT scale 27, ef 32, seed 4; K = 8; CMS 4x2048 for deg > 8192; steps
    # replayed from a captured HIP graph (one launch per iteration)
    "rmat27": dict(gen="rmat", scale=27, ef=32, seed=4, T=0, K=8, graph=True,
                   cms=dict(cms_thresh=8192, cms_depth=4, cms_width=2048, cms_seed=0)),
}


def semi_labels(n, K, seed, classes=100, frac=0.01):
    """K label vectors: in each, frac of the vertices seeded uniformly with
    one of `classes` labels, the rest unlabelled (0xFFFFFFFF)"""
    rng = np.random.default_rng(seed)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, int(frac * n), replace=False)
        lab[idx, k] = rng.integers(0, classes, len(idx), dtype=np.uint32)
    return lab.reshape(-1)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_graph(w):
    import lpa_amd as L
    if w["gen"] == "rmat":
        # drawn on the host, BFS-relabelled on the device (the same graph as
        # the host relabelling, bit for bit)
        g = L.gen_rmat_device_bfs(w["scale"], w["ef"], 0.57, 0.19, 0.19, seed=w["seed"])
    elif w["gen"] == "copying":
        g = L.gen_copying(w["n"], 30.0, 2.1, 100000, 0.7, 1000, seed=w["seed"])
    else:
        g = L.gen_lattice(w["side"], 0.10, 0.005, seed=w["seed"])
    t0 = time.perf_counter()
    cgr = g.encode(w["T"], 128)
    return g, cgr, time.perf_counter() - t0


def device_encode_seconds(L, g, w):
    """wall time of CGR-encoding the same CSR on the device (CSR upload over
    PCIe, record lengths, scan, bit writing, validation pass), best of 2"""
    best = None
    for _ in range(2):
        e = L.Engine(0)
        t0 = time.perf_counter()
        e.encode_csr(g.row, g.col, w["T"])
        dt = time.perf_counter() - t0
        e.close()
        best = dt if best is None else min(best, dt)
    return best


def alg_bytes(st, K):
    """algorithmic HBM bytes of one pass over a kernel family's vertices:
    stream once, 8 B offset per vertex, per label vector a 4 B gather per
    arc plus 4 B read and 4 B write per vertex; small-CSR arcs 4 B"""
    return st["bits"] / 8.0 + 8.0 * st["vertices"] + K * (4.0 * st["arcs"] + 8.0 * st["vertices"]) \
        + 4.0 * st["csr_arcs"]


def timed(e, steps, warmup, method="nova", labels=None, use_graph=False):
    """warmup untimed steps, then `steps` steps timed with HIP events (per
    kernel family too when eager); use_graph replays captured steps"""
    e.labels_init(labels)
    e.set_graph_mode(use_graph)
    for _ in range(warmup):
        {"nova": e.nova_iterate, "cglp": e.cglp_iterate, "csr": e.csr_iterate}[method]()
    ms, changed = e.time_iterations(steps, method, use_graph=use_graph)
    kt = e.kernel_times()
    e.set_graph_mode(False)
    return ms, changed, kt


def train_model(L, cgr_main, w, K):
    """cost model trained on work recorded (K label vectors) on the lattice
    and on this graph"""
    samples = []
    lat = L.gen_lattice(1024, 0.10, 0.005, seed=0).encode(0, 128)
    r22 = cgr_main
    if cgr_main.n > (1 << 22):
        # the model learns from the lattice and R-MAT scale 22 only
        r22 = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, bfs=True).encode(0, 128)
    for cgr, cms in ((lat, {}), (r22, WORKLOADS["rmat22"]["cms"])):
        e = L.Engine(0)
        e.upload(cgr)
        e.config(K=K, **cms)
        e.schedule_degree()
        e.labels_init()
        e.nova_iterate()
        samples.append(L.schedule.record_samples(e, max_samples=200000))
        e.close()
    return L.schedule.train_cost_model(samples, epochs=40, batch=2048, lr=2e-3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat27", choices=sorted(WORKLOADS))
    ap.add_argument("--K", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cglp", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-device-encode", action="store_true",
                    help="skip timing the CGR encoding of the workload's CSR on the device")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline phase (for profiling): no K=1, two-kernel or degree-binned K runs")
    ap.add_argument("--stream-host", type=int, default=0, metavar="BATCH_BYTES",
                    help="out-of-core: keep the CGR stream in pinned host memory and copy it to the device in "
                         "batches of at most BATCH_BYTES every step (host-link-inclusive rate)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="finish with three plain device copies of 256 MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import lpa_amd as L
    w = dict(WORKLOADS[args.workload])
    K = args.K or w["K"]
    t0 = time.time()
    g, cgr, host_enc = build_graph(w)
    m = int(g.m)
    log(f"graph n={g.n} m={m} bits={cgr.nbits} built in {time.time() - t0:.1f}s (host CGR encode {host_enc:.2f}s)")
    dev_enc = None
    if not args.no_device_encode and not args.headline_only:
        dev_enc = device_encode_seconds(L, g, w)
        log(f"device CGR encode {dev_enc:.3f}s")

    e = L.Engine(0)
    if args.stream_host:
        e.stream_residence(True, args.stream_host)
    e.upload(cgr)
    res = {}
    semi = bool(w.get("semi"))
    lab1 = semi_labels(g.n, 1, w["seed"]) if semi else None
    labK = semi_labels(g.n, K, w["seed"]) if semi else None
    if not args.headline_only:
        # degree-binned plain kernel, K = 1
        e.config(K=1, semi=semi, **w["cms"])
        e.schedule_degree()
        ms, ch1, kt1 = timed(e, args.steps, args.warmup, labels=lab1)
        res["degree_k1"] = dict(edges_per_s=m / (float(np.mean(ms)) / 1e3), ms_per_step=float(np.mean(ms)),
                                kernel_ms=kt1, changed=[int(x) for x in ch1])
        log(f"degree K=1: {np.mean(ms):.3f} ms/step")
        if not args.no_cglp:
            ms2, _, _ = timed(e, max(2, args.steps // 2), 1, "cglp", labels=lab1)
            res["cglp_k1"] = dict(edges_per_s=m / (float(np.mean(ms2)) / 1e3), ms_per_step=float(np.mean(ms2)))
            log(f"cglp K=1: {np.mean(ms2):.3f} ms/step")
            # comparator: LP over the uncompressed CSR (no decode in the step)
            ms3, _, _ = timed(e, max(2, args.steps // 2), 1, "csr", labels=lab1)
            res["csr_k1"] = dict(edges_per_s=m / (float(np.mean(ms3)) / 1e3), ms_per_step=float(np.mean(ms3)),
                                 note="uncompressed CSR, same kernels; the CSR adds 4 B per arc of HBM")
            log(f"uncompressed CSR K=1: {np.mean(ms3):.3f} ms/step")
    head_ms = None
    if not args.headline_only or args.no_learned:
        # degree-binned, K vectors
        e.config(K=K, semi=semi, **w["cms"])
        e.schedule_degree()
        msk, chk, ktk = timed(e, args.steps, args.warmup, labels=labK)
        res[f"degree_k{K}"] = dict(edges_per_s=K * m / (float(np.mean(msk)) / 1e3), ms_per_step=float(np.mean(msk)))
        log(f"degree K={K}: {np.mean(msk):.3f} ms/step")
        head_ms, head_kt, schedule = msk, ktk, "degree"
    if not args.no_learned:
        model, losses = train_model(L, cgr, w, K)
        e.config(K=K, semi=semi, **w["cms"])
        # the degree schedule judges the learned one on the same steps and
        # stays if it is faster (the learned layout's fallback)
        e.schedule_degree()
        e.labels_init(labK)
        kept, judge_l, judge_d = e.schedule_learned_judged(model, iters=5)
        log(f"judge: learned layout {judge_l:.3f} ms, degree layout {judge_d:.3f} ms -> "
            f"{'learned' if kept else 'degree'} kept")
        layout = "learned" if kept else "degree"
        if w.get("autotune") and not kept:
            # the degree layout stays: its bin thresholds from the autotuner,
            # tuned on the steps the headline times (after the warm-up)
            e.labels_init(labK)
            for _ in range(args.warmup):
                e.nova_iterate()
            tuned, tuned_ms = e.autotune(iters=2)
            res["autotune"] = dict(params=tuned, median_ms=tuned_ms)
            layout = "degree (autotuned)"
            log(f"autotuned degree schedule {tuned}: {tuned_ms:.3f} ms/step")
        use_graph = bool(w.get("graph"))
        msl, chl, ktl = timed(e, args.steps, args.warmup, labels=labK, use_graph=use_graph)
        if use_graph:
            # the per-family split comes from the same steps run eagerly
            mse, _, ktl = timed(e, args.steps, args.warmup, labels=labK)
            res["eager_ms_per_step"] = float(np.mean(mse))
            log(f"same steps eager: {np.mean(mse):.3f} ms/step")
        res[f"learned_k{K}"] = dict(edges_per_s=K * m / (float(np.mean(msl)) / 1e3),
                                    ms_per_step=float(np.mean(msl)), model_loss=[float(losses[0]), float(losses[-1])],
                                    bins=e.counts(), layout=layout, use_graph=use_graph)
        res["learned_vs_degree"] = {
            "K": K, "learned_layout_ms": judge_l, "degree_layout_ms": judge_d,
            "learned_gain": judge_d / judge_l, "kept": "learned" if kept else "degree",
            "note": "median of 5 steps per layout from the same labels, measured degree, learned, degree"}
        log(f"headline K={K} ({layout} layout{', HIP-graph replay' if use_graph else ''}): "
            f"{np.mean(msl):.3f} ms/step")
        head_ms, head_kt = msl, ktl
        schedule = "learned" if kept else layout + ", the learned layout rejected by the judge"
    if args.copy_calibration:
        cms_ = e.calibration_copy(256 << 20, 3)
        log(f"copy calibration: 3 copies of 256 MiB, {cms_:.3f} ms each ({2 * 256 * 1.048576 / cms_:.0f} GB/s)")
    gather = None
    if not args.stream_host and not args.copy_calibration:
        # the ceiling of the step's dominant access: random reads of label
        # rows (the K labels of a vertex; 8 at a time beyond K = 8) from a
        # buffer the size of a label array, one per arc
        words = 1 if K == 1 else 2 if K < 4 else 4 if K < 8 else 8
        g_ms = e.calibration_gather(4 * K * g.n, words, m, 3)
        gather = {"row_bytes": 4 * words, "footprint_bytes": 4 * K * int(g.n), "gathers": m,
                  "ms": g_ms, "rows_per_s": m / (g_ms / 1e3)}
        log(f"random-gather ceiling: {m} reads of {4 * words} B rows over {4 * K * g.n / 1e9:.2f} GB: "
            f"{g_ms:.2f} ms ({gather['rows_per_s'] / 1e9:.1f} G rows/s)")
    stats = e.stats()
    ms_per_step = float(np.mean(head_ms))
    value = K * m * args.steps / (float(np.sum(head_ms)) / 1e3)
    it_bytes = cgr.nbits / 8 + 8 * (cgr.n_c + 1) + K * (4 * m + 8 * g.n)
    if any(head_kt[k] > 0 for k in head_kt):
        dom = max((k for k in head_kt if stats[k]["vertices"]), key=lambda k: head_kt[k])
        dom_bytes, dom_ms = alg_bytes(stats[dom], K), head_kt[dom]
    else:
        # a host-resident stream's batched step records no per-family times:
        # the whole step stands for the kernel
        dom, dom_bytes, dom_ms = "step", it_bytes, ms_per_step
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
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
        "config": {"workload": args.workload, "n": int(g.n), "m": m, "cgr_bits": int(cgr.nbits), "K": K,
                   "semi_supervised": semi, "threshold": w["T"], "method": "nova-fused", "schedule": schedule,
                   "hip_graph_replay": bool(w.get("graph")) and not args.no_learned, **w["cms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "kernel_ms_per_step": dom_ms, "alg_bytes_per_step": dom_bytes},
        "iteration_roofline": {"alg_bytes": it_bytes, "achieved_GBs": it_bytes / (ms_per_step / 1e3) / 1e9,
                               "frac": it_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": head_kt,
        "bins": {k: v for k, v in stats.items() if v["vertices"]},
        "gather_ceiling": gather,
        "device": e.info(),
        "encode": {"host_s": host_enc, "host_threads": L.threads(), "device_s": dev_enc,
                   "note": "CGR encode of the workload's CSR; device time includes the CSR upload"},
        **res,
    }
    if args.stream_host:
        nb, stage = e.stream_batches()
        line["config"]["stream"] = (f"host-resident (out-of-core): {nb} batches of at most {args.stream_host} bytes, "
                                    f"{stage} bytes of staging; the rate includes the host-to-device copies")
    tr = None if args.stream_host else measured_traffic(args.workload, K, dom)
    if tr:
        line["roofline"]["traffic"] = tr["bytes"]
        line["roofline"]["traffic_source"] = tr["source"]
    if gather:
        # arcs per second of the step and of the dominant family against the
        # random-gather ceiling (each arc gathers one row per 8 vectors)
        per_arc = max(1, (K + 7) // 8)
        step_rows = m * per_arc / (ms_per_step / 1e3)
        gather["step_rows_per_s"] = step_rows
        gather["step_frac"] = step_rows / gather["rows_per_s"]
        if dom in stats and dom_ms > 0:
            fam_rows = stats[dom]["arcs"] * per_arc / (dom_ms / 1e3)
            gather["family"] = dom
            gather["family_rows_per_s"] = fam_rows
            gather["family_frac"] = fam_rows / gather["rows_per_s"]
    if "degree_k1" in res:
        line["speedup_vs_degree_k1"] = value / res["degree_k1"]["edges_per_s"]
    if not args.no_cpu_baseline:
        deg = g.degrees()
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        view = O.CgrView(np.frombuffer(cgr.stream_bytes(), np.uint8), cgr.nbits, cgr.offsets.copy(),
                         cgr.vertices.copy(), cgr.small_offsets.copy(), cgr.small_nbrs.copy())
        cb = cpu_baseline_run(O, view, cgr, deg, w, K)
        if K != 1:
            cb["k1"] = cpu_baseline_run(O, view, cgr, deg, w, 1)
        line["cpu_baseline"] = cb
    print(json.dumps(line))


def measured_traffic(workload, K, fam):
    """HBM bytes per step of a kernel family from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (copy-calibrated)"""
    path = os.path.join(ROOT, "profiles", f"{workload}_k{K}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    t = d.get("traffic_per_iter", {}).get(fam)
    if not t:
        return None
    return {"bytes": t["fetch_bytes"] + t["write_bytes"], "source": os.path.relpath(path, ROOT)}


def cpu_baseline_run(O, view, cgr, deg, w, K, budget_s=20.0):
    """times the CPU oracle (module O, graph view) decode+LP step with K label
    vectors on the host cores over a bounded random sample of stream
    positions (whole iterations when they fit the budget); edges/s = K *
    arcs / s"""
    cores = min(len(os.sched_getaffinity(0)), 16)
    semi = bool(w.get("semi"))
    c = O.cfg(K=K, nthreads=cores, semi=semi, **w["cms"])
    lab = semi_labels(cgr.n, K, w["seed"]) if semi else np.repeat(np.arange(cgr.n, dtype=np.uint32), K)
    seeds = (lab != O.UNLABELLED).astype(np.uint8) if semi else None
    out = lab.copy()
    rng = np.random.default_rng(0)
    pos_all = rng.permutation(cgr.n).astype(np.uint64)
    n_probe = min(cgr.n, 20000)
    t = time.perf_counter()
    O.step_cgr_subset(view, lab, np.sort(pos_all[:n_probe]), out, c, seeds)
    probe_s = time.perf_counter() - t
    full_s = probe_s * cgr.n / n_probe
    if full_s * 3 <= budget_s:
        reps = min(10, max(1, int(budget_s / max(full_s, 1e-3))))
        t = time.perf_counter()
        for _ in range(reps):
            O.step_cgr(view, lab, c, seeds)
        el = (time.perf_counter() - t) / reps
        arcs = int(deg.sum())
        return {"value": K * arcs / el, "unit": "edges/s", "cores": cores, "kind": "port", "K": K,
                "sample": f"{reps} full synchronous oracle steps over all {cgr.n} vertices ({arcs} arcs, "
                          f"K = {K}), {el:.2f} s each"}
    n_take = int(min(cgr.n, max(n_probe, n_probe * budget_s / max(probe_s, 1e-3))))
    for attempt in range(2):
        sample = np.sort(pos_all[:n_take])
        t = time.perf_counter()
        O.step_cgr_subset(view, lab, sample, out, c, seeds)
        el = time.perf_counter() - t
        # the small probe overstates the time per vertex: when the sample
        # ran well under the budget, take one sized by its own rate
        if attempt or el >= 0.5 * budget_s or n_take >= cgr.n:
            break
        n_take = int(min(cgr.n, n_take * budget_s / max(el, 1e-3)))
    arcs = int(deg[cgr.vertices[sample.astype(np.int64)]].sum())
    return {"value": K * arcs / el, "unit": "edges/s", "cores": cores, "kind": "port", "K": K,
            "sample": f"oracle decode+LP step (K = {K}) over {n_take} of {cgr.n} random stream positions "
                      f"({arcs} arcs), {el:.1f} s"}


if __name__ == "__main__":
    main()
