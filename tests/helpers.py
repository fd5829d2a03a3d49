"""Shared test helpers: golden fixture loading and small graph builders."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_cases():
    return sorted(d for d in os.listdir(GOLDEN)
                  if os.path.isfile(os.path.join(GOLDEN, d, "params.json")))


def load_golden(case):
    """Reference-written file set of one case, parsed as the reference's
    loader does (file.cpp:197-262): .graph as bytes, .offset text."""
    d = os.path.join(GOLDEN, case)
    base = os.path.join(d, "edges.txt")
    p = json.load(open(os.path.join(d, "params.json")))
    edges = np.loadtxt(base, dtype=np.int64, ndmin=2)
    with open(base + ".offset") as f:
        vals = [int(x) for x in f.read().split()]
    nc = vals[0]
    end_offs = np.array(vals[1:1 + nc], np.uint64)
    g = dict(params=p, edges=edges, end_offs=end_offs, n_c=nc,
             stream=np.fromfile(base + ".graph", np.uint8),
             vertices=np.fromfile(base + ".vertices.bin", np.uint32),
             row=np.fromfile(base + ".row", np.uint32).astype(np.uint64),
             clm=np.fromfile(base + ".clm", np.uint32))
    if os.path.exists(base + ".graph.small.bin"):
        g["small_nbrs"] = np.fromfile(base + ".graph.small.bin", np.uint32)
        g["small_offs"] = np.fromfile(base + ".offset.small.bin", np.uint32)
    else:
        g["small_nbrs"] = np.zeros(0, np.uint32)
        g["small_offs"] = np.zeros(1, np.uint32)
    g["n"] = len(g["row"]) - 1
    return g


def csr_from_edges(n, edges, undirected):
    """sorted, deduplicated CSR (uint64 row, uint32 col)"""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    if undirected:
        e = np.concatenate([e, e[:, ::-1]])
    e = e[e[:, 0] != e[:, 1]] if len(e) else e
    key = np.unique(e[:, 0] * n + e[:, 1]) if len(e) else np.zeros(0, np.int64)
    src, dst = key // n, key % n
    row = np.zeros(n + 1, np.uint64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.uint64), dst.astype(np.uint32)


def brute_mode_step(row, col, labels, K=1, seeds=None, unl=0xFFFFFFFF):
    """pure-Python LP-SC step (dict counting) for tiny graphs"""
    n = len(row) - 1
    out = labels.copy()
    for v in range(n):
        for k in range(K):
            if seeds is not None and seeds[v * K + k]:
                continue
            cnt = {}
            for e in range(int(row[v]), int(row[v + 1])):
                lab = int(labels[int(col[e]) * K + k])
                if lab != unl:
                    cnt[lab] = cnt.get(lab, 0) + 1
            if cnt:
                best = max(cnt.values())
                out[v * K + k] = min(l for l, c in cnt.items() if c == best)
    return out
