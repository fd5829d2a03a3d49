"""The synthetic workload generators are frozen: each configuration's CSR
(row and col arrays, after symmetrising, deduplication and BFS relabelling
where the configuration asks for it) hashes to a committed checksum, so the
graphs the measurements rest on cannot drift unnoticed.  Every random
decision comes from a counter-based stream of the seed, so the graphs do not
depend on the number of threads either.

Checksum: xxh3_64 over the bytes of row (uint64) then col (uint32).
"""
import os
import subprocess
import sys

import numpy as np
import pytest
import xxhash

import lpa_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def checksum(g):
    h = xxhash.xxh3_64()
    h.update(np.ascontiguousarray(g.row).tobytes())
    h.update(np.ascontiguousarray(g.col).tobytes())
    return h.hexdigest()


FROZEN = {
    # 1024 x 1024 lattice, 10 % of the edges dropped, 0.5 % shortcuts, seed 0
    "lattice1024_seed0": (lambda: L.gen_lattice(1024, 0.10, 0.005, seed=0), 1048576, 3781984, "b2af7d86b9028e67"),
    # R-MAT scale 22, edge factor 16, seed 1, symmetrised, BFS order
    "rmat22_seed1": (lambda: L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, bfs=True), 4194304, 128307786,
                     "b2ac7b4be1bac7fa"),
    # the copying model's rules (mean out-degree 30, exponent 2.1, cap 100k,
    # 70 % copied within +-1000, seed 2) at a million vertices
    "copying1m_seed2": (lambda: L.gen_copying(1_000_000, 30.0, 2.1, 100000, 0.7, 1000, seed=2), 1000000, 29858200,
                        "409e1a9964bf799b"),
    # R-MAT scale 24, edge factor 16, seed 3, symmetrised, BFS order
    "rmat24_seed3": (lambda: L.gen_rmat(24, 16, 0.57, 0.19, 0.19, seed=3, bfs=True), 16777216, 520761802,
                     "d50d1452d707001b"),
}


@pytest.mark.parametrize("name", sorted(FROZEN))
def test_frozen_generator_checksums(name):
    make, n, m, digest = FROZEN[name]
    g = make()
    assert (g.n, g.m) == (n, m)
    assert checksum(g) == digest


def test_copying_model_reaches_its_mean_out_degree():
    """copies are distinct positions of the prototype's list, so they do not
    collapse as duplicates: the mean out-degree stays at the target"""
    g = L.gen_copying(1_000_000, 30.0, 2.1, 100000, 0.7, 1000, seed=2)
    assert 29.5 < g.m / g.n < 30.5
    assert g.degrees().max() <= 100000


def test_graphs_do_not_depend_on_thread_count():
    code = ("import sys; sys.path.insert(0, %r); import lpa_amd as L, numpy as np, xxhash\n"
            "g = L.gen_rmat(16, 16, seed=9, bfs=True); c = L.gen_copying(50000, 30.0, 2.1, 5000, 0.7, 100, seed=9)\n"
            "h = xxhash.xxh3_64()\n"
            "for x in (g.row, g.col, c.row, c.col): h.update(np.ascontiguousarray(x).tobytes())\n"
            "print(h.hexdigest())\n" % ROOT)
    out = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        out.add(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                               check=True, timeout=300).stdout.strip())
    assert len(out) == 1, out
