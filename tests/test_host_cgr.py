"""The product's host encoder and file set against the reference compressor's
golden files and against the oracle's codec."""
import os

import numpy as np
import pytest

import lpa_amd as L
import pyoracle as O
from helpers import golden_cases, load_golden


def product_csr(g):
    return L.Csr.from_arrays(g["row"], g["clm"])


@pytest.mark.parametrize("case", golden_cases())
def test_encoder_matches_reference_files(case, tmp_path):
    g = load_golden(case)
    t = g["params"]["small_thresh"]
    c = product_csr(g).encode(t, 128)
    assert c.n_c == g["n_c"]
    assert np.array_equal(c.offsets[1:], g["end_offs"])
    assert c.stream_bytes() == bytes(g["stream"])
    assert np.array_equal(c.vertices, g["vertices"])
    if c.n_c < c.n:
        assert np.array_equal(c.small_offsets, g["small_offs"])
        assert np.array_equal(c.small_nbrs, g["small_nbrs"])
    # file set written by the product is byte-identical to the reference's
    p = str(tmp_path / "g")
    c.write(p, product_csr(g))
    src = os.path.join(os.path.dirname(__file__), "golden", case, "edges.txt")
    for suf in (".graph", ".offset", ".vertices.bin", ".row", ".clm",
                ".graph.small.bin", ".offset.small.bin"):
        if os.path.exists(src + suf):
            assert open(p + suf, "rb").read() == open(src + suf, "rb").read(), suf


@pytest.mark.parametrize("case", golden_cases())
def test_loader_reads_reference_files(case):
    g = load_golden(case)
    src = os.path.join(os.path.dirname(__file__), "golden", case, "edges.txt")
    t = g["params"]["small_thresh"]
    c = L.Cgr.load(src, t)
    assert c.n == g["n"] and c.n_c == g["n_c"]
    assert c.stream_bytes() == bytes(g["stream"])
    # big-endian word packing of File::load_compress_graph (file.cpp:222-238)
    raw = bytes(g["stream"]) + b"\0" * 4
    want = [int.from_bytes(raw[4 * i:4 * i + 4], "big") for i in range((len(g["stream"]) + 3) // 4)]
    assert list(c.words[:c.nwords]) == want


def test_edge_list_reader_matches_reference_csr():
    g = load_golden("lattice16_t0")
    src = os.path.join(os.path.dirname(__file__), "golden", "lattice16_t0", "edges.txt")
    c = L.Csr.read_edge_list(src, undirected=True)
    assert np.array_equal(c.row, g["row"]) and np.array_equal(c.col, g["clm"])


@pytest.mark.parametrize("gen", ["lattice", "rmat", "copying"])
def test_generated_graphs_encode_like_oracle(gen):
    if gen == "lattice":
        g = L.gen_lattice(64, 0.1, 0.005, seed=0)
    elif gen == "rmat":
        g = L.gen_rmat(12, 16, seed=1)
    else:
        g = L.gen_copying(5000, 30.0, 2.1, 2000, 0.7, 100, seed=2)
    for t in (0, 32):
        c = g.encode(t, 128)
        w = O.write_cgr(g.row, g.col, t, 128)
        assert c.stream_bytes() == bytes(w["stream"])
        assert np.array_equal(c.offsets[1:], w["end_offs"])
        assert np.array_equal(c.vertices, w["vertices"])


def test_generators_deterministic_and_clean():
    a = L.gen_rmat(11, 8, seed=5)
    b = L.gen_rmat(11, 8, seed=5)
    assert np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col)
    c = L.gen_rmat(11, 8, seed=6)
    assert not np.array_equal(a.col, c.col)
    row, col = a.row.astype(np.int64), a.col.astype(np.int64)
    src = np.repeat(np.arange(a.n), np.diff(row))
    assert (col != src).all()                                   # no self loops
    assert all((np.diff(col[row[v]:row[v + 1]]) > 0).all() for v in range(a.n))  # sorted, unique
    # symmetric
    fwd = set(zip(src.tolist(), col.tolist()))
    assert all((d, s) in fwd for s, d in list(fwd)[:5000])


def test_lattice_counts():
    g = L.gen_lattice(1024, 0.10, 0.005, seed=0)
    lattice_edges = 2 * 1024 * 1023
    und = g.m // 2
    # 10% dropped (binomial), exactly floor(0.005 * 2^20) = 5242 shortcuts
    assert abs((und - 5242) - 0.9 * lattice_edges) < 5 * np.sqrt(0.09 * lattice_edges)


def test_bfs_relabel_is_bfs_order():
    g = L.gen_rmat(10, 8, seed=3, bfs=False)
    r, perm = g.bfs_relabel(0)
    assert perm[0] == 0 and len(set(perm.tolist())) == g.n
    # relabelled graph is isomorphic: arc (u,v) -> (perm[u], perm[v])
    row, col = g.row.astype(np.int64), g.col
    src = np.repeat(np.arange(g.n), np.diff(row))
    a = set(zip(perm[src].tolist(), perm[col].tolist()))
    rr = r.row.astype(np.int64)
    b = set(zip(np.repeat(np.arange(r.n), np.diff(rr)).tolist(), r.col.tolist()))
    assert a == b
    # BFS layers are contiguous in the new ids: distances are non-decreasing
    dist = np.full(r.n, -1)
    dist[0] = 0
    q = [0]
    for u in q:
        for w in r.col[rr[u]:rr[u + 1]]:
            if dist[w] < 0:
                dist[w] = dist[u] + 1
                q.append(int(w))
    reach = dist[dist >= 0]
    assert (np.diff(reach) >= 0).all() and (dist[len(reach):] == -1).all()


def test_loader_errors(tmp_path):
    g = load_golden("mixed400_t0")
    c = product_csr(g).encode(0, 128)
    p = str(tmp_path / "x")
    c.write(p)
    with open(p + ".graph", "r+b") as f:    # truncate the stream
        f.truncate(10)
    os.remove(p + ".offset64.bin")
    with pytest.raises(L.LpaError):
        L.Cgr.load(p)
    with pytest.raises(L.LpaError):
        L.Cgr.load(str(tmp_path / "missing"))
    with open(p + ".offset", "w") as f:     # count larger than the lines present
        f.write("999999\n4\n")
    with pytest.raises(L.LpaError):
        L.Cgr.load(p)


def test_duplicate_and_self_loop_edges_are_removed():
    c = L.Csr.from_edges(4, [0, 0, 1, 2, 2], [1, 1, 1, 3, 3], undirected=True)
    assert list(c.row) == [0, 1, 2, 3, 4] and list(c.col) == [1, 0, 3, 2]


def test_parallel_bfs_equals_sequential_bfs():
    import collections
    for seed in range(4):
        g = L.gen_rmat(11, 8, seed=40 + seed, bfs=False)
        r, perm = g.bfs_relabel(0)
        row, col = g.row.astype(np.int64), g.col
        seen = np.full(g.n, -1)
        seen[0] = 0
        q = collections.deque([0])
        order = [0]
        while q:
            u = q.popleft()
            for w in col[row[u]:row[u + 1]]:
                if seen[w] < 0:
                    seen[w] = len(order)
                    order.append(int(w))
                    q.append(int(w))
        for v in range(g.n):
            if seen[v] < 0:
                seen[v] = len(order)
                order.append(v)
        assert np.array_equal(perm, seen)
