"""Parity at the benchmark workloads' full sizes, through the C ABI.

R-MAT scale 22 (4.2 M vertices, 128 M arcs): ten steps with K = 1 and ten
with K = 8 (count-min sketch 4 x 1024 above degree 4096), bit for bit
against the oracle's own trajectory after every step.

R-MAT scale 27 (134 M vertices, 8.36 G arcs, K = 8, count-min sketch 4 x 2048
above degree 8192; ten steps), R-MAT scale 24 (16.8 M vertices, K = 64
semi-supervised vectors with 1 % seeds from 100 classes; fifteen steps) and
the copying model at 50 M vertices (1.5 G arcs, T = 32; six steps) are
too large for the oracle to follow a whole trajectory in minutes, so they
take one-step checks: after a step the oracle recomputes L_{t+1} from the
engine's own L_t.  Every vertex and every vector is recomputed after the
first and the last step; a random sample of vertices after the others.
When each step passes for every vertex, the whole trajectory is the
oracle's.
"""
import time

import numpy as np
import pytest
import xxhash

import lpa_amd as L
import pyoracle as O
from gpu_common import make_engine, oracle_view, run_and_compare

pytestmark = pytest.mark.gpu

CMS22 = dict(cms_thresh=4096, cms_depth=4, cms_width=1024, cms_seed=0)
# R-MAT scale 27, edge factor 32, seed 4, BFS order: xxh3_64 of row then col
RMAT27_CHECKSUM = "99d290a7c788d9c0"


def checksum(g):
    h = xxhash.xxh3_64()
    h.update(np.ascontiguousarray(g.row).tobytes())
    h.update(np.ascontiguousarray(g.col).tobytes())
    return h.hexdigest()


_capsys = None


@pytest.fixture(autouse=True)
def _live_log(capsys):
    """progress lines bypass pytest's capture: the full-size tests run for
    minutes, and a run that prints nothing for that long looks hung"""
    global _capsys
    _capsys = capsys
    yield
    _capsys = None


def log(msg):
    if _capsys is None:
        print(f"[scale] {msg}", flush=True)
        return
    with _capsys.disabled():
        print(f"[scale] {msg}", flush=True)


@pytest.fixture(scope="module")
def rmat22():
    g = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, bfs=True)
    return g.encode(0, 128)


def test_rmat22_ten_steps_k1(rmat22):
    run_and_compare(rmat22, 10, **CMS22)


def test_rmat22_ten_steps_k8(rmat22):
    """K = 8 into the settled regime: hubs above degree 65,535 run through
    the chunked count-min kernels"""
    run_and_compare(rmat22, 10, K=8, **CMS22)


def test_rmat20_k64_uncached_rows():
    """64 label vectors of a million vertices (268 MB of label rows, more
    than the last-level cache): the kernels gather 8 vectors per row"""
    g = L.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=6, bfs=True)
    run_and_compare(g.encode(0, 128), 2, K=64, cms_thresh=3000, cms_depth=4, cms_width=1024, cms_seed=4)


def one_step_checks(cgr, iters, full_at, K, semi=False, labels0=None, sample=200_000, **cms):
    """runs `iters` engine steps; after every step the oracle recomputes the
    step from the engine's previous labels: all vertices after the steps in
    full_at (0-based), a random sample of vertices after the others"""
    e = make_engine(cgr, K=K, semi=semi, **cms)
    e.labels_init(labels0)
    view = oracle_view(cgr)
    c = O.cfg(K=K, semi=semi, **cms)
    seeds = None
    if semi:
        seeds = (np.ascontiguousarray(labels0, np.uint32) != O.UNLABELLED).astype(np.uint8)
    rng = np.random.default_rng(7)
    npos = cgr.n
    before = e.labels()
    checked = 0
    for it in range(iters):
        ch = e.nova_iterate()
        after = e.labels()
        t0 = time.perf_counter()
        if it in full_at:
            want, och = O.step_cgr(view, before, c, seeds)
            bad = np.nonzero(after != want)[0]
            assert ch == och, (it, ch, och)
            n_lab = len(want)
        else:
            pos = np.sort(rng.choice(npos, min(sample, npos), replace=False)).astype(np.uint64)
            want = before.copy()
            O.step_cgr_subset(view, before, pos, want, c, seeds)
            verts = cgr.vertices[pos.astype(np.int64)].astype(np.int64)
            idx = (verts[:, None] * K + np.arange(K)[None, :]).reshape(-1)
            bad = idx[np.nonzero(after[idx] != want[idx])[0]]
            n_lab = len(idx)
        checked += n_lab
        log(f"step {it + 1}: {'all' if it in full_at else 'sampled'} {n_lab} labels checked in "
            f"{time.perf_counter() - t0:.1f} s, {len(bad)} differ, {ch} changed")
        assert len(bad) == 0, (it, bad[:5], after[bad[:5]], want[bad[:5]])
        before = after
    e.close()
    return checked


def test_rmat27_k8_one_step_checks():
    """R-MAT scale 27, edge factor 32, seed 4, BFS order; K = 8 from L0 = v;
    CMS 4 x 2048 above degree 8192; ten steps"""
    t0 = time.perf_counter()
    log("R-MAT 27: generating (host draws, device BFS relabelling)")
    g = L.gen_rmat_device_bfs(27, 32, 0.57, 0.19, 0.19, seed=4)
    log(f"R-MAT 27: n={g.n} m={g.m} (BFS relabelled on the device) in {time.perf_counter() - t0:.0f} s; "
        f"checksum {checksum(g)}")
    assert checksum(g) == RMAT27_CHECKSUM
    cgr = g.encode(0, 128)
    log(f"R-MAT 27: bits={cgr.nbits} built in {time.perf_counter() - t0:.0f} s")
    del g
    one_step_checks(cgr, 10, (0, 9), 8, cms_thresh=8192, cms_depth=4, cms_width=2048, cms_seed=0)


def test_rmat24_k64_semi_one_step_checks():
    """R-MAT scale 24, edge factor 16, seed 3, BFS order; K = 64 vectors, each
    with 1 % of the vertices seeded from 100 classes (clamped), the rest
    unlabelled; exact mode; fifteen steps"""
    log("R-MAT 24: generating (host draws, device BFS relabelling)")
    g = L.gen_rmat_device_bfs(24, 16, 0.57, 0.19, 0.19, seed=3)
    log(f"R-MAT 24: n={g.n} m={g.m}")
    assert checksum(g) == "d50d1452d707001b"   # the host-relabelled graph's (test_generators.py)
    cgr = g.encode(0, 128)
    n, K = g.n, 64
    del g
    rng = np.random.default_rng(3)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, n // 100, replace=False)
        lab[idx, k] = rng.integers(0, 100, len(idx), dtype=np.uint32)
    one_step_checks(cgr, 15, (0, 14), K, semi=True, labels0=lab.reshape(-1), sample=20_000)


def test_copy50m_one_step_checks():
    """the copying model at 50 M vertices (mean out-degree 30, power law 2.1
    capped at 100 k, 70 % copied within +-1000, seed 2; directed), T = 32 (a
    small-vertex CSR part), K = 1, CMS 4 x 2048 above degree 8192; six steps,
    every vertex recomputed after the first and the last"""
    t0 = time.perf_counter()
    log("copying 50M: generating")
    g = L.gen_copying(50_000_000, 30.0, 2.1, 100000, 0.7, 1000, seed=2)
    cgr = g.encode(32, 128)
    log(f"copying 50M: n={g.n} m={g.m} n_c={cgr.n_c} built in {time.perf_counter() - t0:.0f} s")
    del g
    one_step_checks(cgr, 6, (0, 5), 1, cms_thresh=8192, cms_depth=4, cms_width=2048, cms_seed=0)
