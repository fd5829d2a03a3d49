#!/usr/bin/env python3
"""Write golden CGR file sets with the REFERENCE compressor.

Runs ``oracle/_ref/ref_cgr`` -- the reference's own ``File``/``CgrCompressor``
(src/file.cpp:24-95, src/compressor.cpp:22-290), compiled from its sources by
``make -C oracle ref`` -- on small deterministic edge lists, and stores the
edge list plus every file the reference writes under ``tests/golden/<case>/``.
Those files are data (inputs and expected outputs); no reference code is kept.

Cases cover: the hand-derived record of vertex 5 (N = {1,6,7,8,9,12}), a
lattice with dropped edges and shortcuts, records with many interval and
residual segments (256-bit segmentation and the merged last segment),
vertices of degree >= 128 (stream order), and small-vertex CSR splits
(threshold 16 and 32).

usage: python tests/golden/make_golden.py   (from the repository root)
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_cgr")
WORK = os.path.join(ROOT, "oracle", "_ref", "work")
SUFFIXES = [".graph", ".offset", ".vertices.bin", ".graph.small.bin",
            ".offset.small.bin", ".row", ".clm"]


def case_a4():
    # directed list of vertex 5 only: N(5) = {1,6,7,8,9,12}
    return [(5, x) for x in (1, 6, 7, 8, 9, 12)], 0


def case_lattice(side=16, seed=11):
    rng = np.random.default_rng(seed)
    edges = set()
    for r in range(side):
        for c in range(side):
            v = r * side + c
            if c + 1 < side and rng.random() >= 0.10:
                edges.add((v, v + 1))
            if r + 1 < side and rng.random() >= 0.10:
                edges.add((v, v + side))
    n = side * side
    while len(edges) < int(1.9 * n) + 8:
        u, w = (int(x) for x in rng.integers(0, n, 2))
        if u != w:
            edges.add((min(u, w), max(u, w)))
    return sorted(edges), 1


def _runs_adj(rng, v, n, deg, run_frac):
    """sorted adjacency of ~deg ids mixing +1 runs and isolated residuals"""
    out = set()
    while len(out) < deg:
        if rng.random() < run_frac:
            start = int(rng.integers(0, n - 12))
            for x in range(start, start + int(rng.integers(4, 12))):
                out.add(x)
        else:
            out.add(int(rng.integers(0, n)))
    out.discard(v)
    return sorted(out)


def case_mixed(seed=5, n=400):
    rng = np.random.default_rng(seed)
    edges = []
    for v in range(n):
        r = rng.random()
        if r < 0.15:
            deg = 0
        elif r < 0.75:
            deg = int(rng.integers(1, 40))
        elif r < 0.95:
            deg = int(rng.integers(40, 200))
        else:
            deg = int(rng.integers(200, 380))
        if deg:
            for x in _runs_adj(rng, v, n, deg, run_frac=rng.choice([0.0, 0.2, 0.6])):
                edges.append((v, x))
    return edges, 0


def case_multiseg(seed=9, n=6000):
    rng = np.random.default_rng(seed)
    edges = []
    for v in range(0, n, 400):
        deg = int(rng.integers(600, 2600))
        for x in _runs_adj(rng, v, n, deg, run_frac=0.5):
            edges.append((v, x))
    # a vertex of many short intervals separated by gaps (interval segments)
    v, x, adj = 3, 10, []
    for _ in range(220):
        ln = int(rng.integers(4, 7))
        adj.extend(range(x, x + ln))
        x += ln + int(rng.integers(1, 30))
    edges.extend((v, a) for a in adj if a < n)
    # a vertex far from its neighbours (long first codes)
    edges.extend((n - 1, a) for a in (0, 1, 2, 3, 4, 17, 2000, 4000))
    return edges, 0


def case_rmat(scale=10, ef=8, seed=21):
    rng = np.random.default_rng(seed)
    m = ef << scale
    u = np.zeros(m, np.int64)
    w = np.zeros(m, np.int64)
    for _ in range(scale):
        r = rng.random(m)
        bu = (r >= 0.76).astype(np.int64)                     # c or d
        bw = (((r >= 0.57) & (r < 0.76)) | (r >= 0.95)).astype(np.int64)  # b or d
        u = (u << 1) | bu
        w = (w << 1) | bw
    keep = u != w
    a, b = np.minimum(u[keep], w[keep]), np.maximum(u[keep], w[keep])
    pairs = np.unique(a * (1 << scale) + b)
    return [(int(p >> scale), int(p & ((1 << scale) - 1))) for p in pairs], 1


CASES = {
    "a4_vertex5": (case_a4, [0]),
    "lattice16": (case_lattice, [0, 16]),
    "mixed400": (case_mixed, [0, 32]),
    "multiseg6000": (case_multiseg, [0]),
    "rmat10": (case_rmat, [0, 16]),
}


def main():
    # the reference compressor is built here, from its sources under
    # /root/reference, into oracle/_ref (ignored by git and never shipped
    # with the tree)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if not os.path.exists(REF_BIN):
        sys.exit("the reference compressor did not build (oracle/Makefile `ref`)")
    for name, (fn, thresholds) in CASES.items():
        edges, one_way = fn()
        for t in thresholds:
            case = f"{name}_t{t}"
            wdir = os.path.join(WORK, case)
            shutil.rmtree(wdir, ignore_errors=True)
            os.makedirs(wdir)
            src = os.path.join(wdir, "edges.txt")
            with open(src, "w") as f:
                f.writelines(f"{a} {b}\n" for a, b in edges)
            subprocess.run([REF_BIN, src, str(one_way), str(t)], check=True,
                           stdout=subprocess.DEVNULL)
            out = os.path.join(HERE, case)
            shutil.rmtree(out, ignore_errors=True)
            os.makedirs(out)
            shutil.copy(src, os.path.join(out, "edges.txt"))
            for s in SUFFIXES:
                if os.path.exists(src + s):
                    shutil.copy(src + s, os.path.join(out, "edges.txt" + s))
            with open(os.path.join(out, "params.json"), "w") as f:
                json.dump({"one_way": one_way, "small_thresh": t, "large_thresh": 128,
                           "zeta_k": 3, "min_itv_len": 4, "seg_bits": 256,
                           "made_by": "oracle/_ref/ref_cgr (reference File/CgrCompressor)"},
                          f, indent=1)
            print(case, len(edges), "edges")


if __name__ == "__main__":
    main()
