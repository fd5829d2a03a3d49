"""BFS relabelling on the device (lpa_csr_bfs_relabel) against the host's
lpah_csr_bfs_relabel, bit for bit: the new CSR (rows and their sorted
columns) and the permutation."""
import numpy as np
import pytest

import lpa_amd as L
from helpers import csr_from_edges

pytestmark = pytest.mark.gpu


def check(e, g, root=0):
    print(f"[gpu-case] bfs relabel n={g.n} m={g.m} root={root}", flush=True)
    r, perm = g.bfs_relabel(root)
    row, col, dperm = e.bfs_relabel(g.row, g.col, root)
    assert np.array_equal(dperm, perm)
    assert np.array_equal(row, r.row)
    assert np.array_equal(col, r.col)


@pytest.fixture(scope="module")
def engine():
    e = L.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("scale,ef,seed,root", [(14, 16, 1, 0), (18, 16, 2, 0), (16, 8, 3, 777)])
def test_rmat(engine, scale, ef, seed, root):
    check(engine, L.gen_rmat(scale, ef, seed=seed, bfs=False), root)


def test_lattice_long_bfs(engine):
    """a 512 x 512 lattice: about a thousand BFS levels"""
    check(engine, L.gen_lattice(512, 0.10, 0.0, seed=4), 0)


def test_directed_copying_model(engine):
    check(engine, L.gen_copying(30000, 30.0, 2.1, 3000, 0.7, 200, seed=5), 11)


def test_disconnected_and_isolated(engine):
    """two components and isolated vertices: the unreached follow by id"""
    edges = [(0, 1), (1, 2), (2, 0), (5, 6), (6, 7), (9, 3)]
    row, col = csr_from_edges(12, edges, undirected=True)
    g = L.Csr.from_arrays(row, col)
    check(engine, g, 0)
    check(engine, g, 6)
    one = L.Csr.from_arrays(np.zeros(2, np.uint64), np.zeros(0, np.uint32))
    check(engine, one, 0)


def test_rejects_bad_ids(engine):
    row = np.array([0, 2, 3], np.uint64)
    col = np.array([1, 5, 0], np.uint32)   # 5 >= n
    with pytest.raises(L.LpaError):
        engine.bfs_relabel(row, col, 0)
    with pytest.raises(L.LpaError):
        engine.bfs_relabel(np.array([0, 1, 2], np.uint64), np.array([1, 0], np.uint32), 2)   # root >= n
