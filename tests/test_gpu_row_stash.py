"""Row stashes of the exact one-workgroup heavy kernel for degrees up to
4096 (K >= 16): every neighbour's label row is read whole once per step
into the workgroup's stash (vector-major, groups of four vectors), and the
batches of vectors read the stash instead of gathering the rows again;
workgroups take their vertices from a work counter.  Labels and change
counts equal the oracle's after every step, and equal an engine with the
stash turned off (LPA_ROW_STASH=0)."""
import numpy as np
import pytest

import lpa_amd as L
from gpu_common import make_engine, run_and_compare
from test_gpu_parity import _hub_graph

pytestmark = pytest.mark.gpu


def _semi_labels(n, K, seed, classes=50):
    rng = np.random.default_rng(seed)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, n // 50, replace=False)
        lab[idx, k] = rng.integers(0, classes, len(idx), dtype=np.uint32)
    return lab.reshape(-1)


@pytest.mark.parametrize("K,semi", [(16, False), (18, False), (64, True)])
def test_hub_graph(K, semi):
    """hubs of degree 3,000 and 6,000 (both one-workgroup instances) and
    70,000 (chunked); K = 16 reads the rows in 16-byte pieces, K = 18 word
    by word with a partial last group of four and partial batches"""
    cgr = _hub_graph().encode(0, 128)
    lab = _semi_labels(cgr.n, K, 7) if semi else None
    run_and_compare(cgr, 3, labels0=lab, K=K, semi=semi)


@pytest.mark.parametrize("K", [18, 32])
def test_every_vertex_heavy(K):
    """all 32,768 vertices of R-MAT 15 in the heavy bin: each resident
    workgroup loops over many vertices of different degrees and re-uses its
    stash region"""
    g = L.gen_rmat(15, 16, seed=16)
    bins = np.full(g.n, L.engine.BIN_GLOBAL, np.uint8)
    run_and_compare(g.encode(0, 128), 3, K=K, schedule=bins)


def test_same_as_without_stash(monkeypatch):
    """64 semi-supervised vectors, every vertex of R-MAT 16 heavy: the
    stashed engine and an engine without stashes agree step by step"""
    g = L.gen_rmat(16, 16, seed=18)
    cgr = g.encode(0, 128)
    K = 64
    bins = np.full(g.n, L.engine.BIN_GLOBAL, np.uint8)
    lab = _semi_labels(g.n, K, 9)
    a = make_engine(cgr, K=K, semi=True, schedule=bins)
    monkeypatch.setenv("LPA_ROW_STASH", "0")
    b = make_engine(cgr, K=K, semi=True, schedule=bins)
    monkeypatch.delenv("LPA_ROW_STASH")
    a.labels_init(lab)
    b.labels_init(lab)
    print(f"[gpu-case] R-MAT 16 K={K} every vertex heavy, stash on / off", flush=True)
    for _ in range(4):
        assert a.nova_iterate() == b.nova_iterate()
        assert np.array_equal(a.labels(), b.labels())
    a.close()
    b.close()
