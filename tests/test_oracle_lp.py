"""Pins the oracle's LP-SC (synchronous, most frequent label, ties to the
smallest, unlabelled ignored, seeds clamped, CMS estimate for heavy vertices)
against closed-form trajectories and brute-force counting."""
import numpy as np
import pytest

import pyoracle as O
from helpers import brute_mode_step, csr_from_edges, golden_cases, load_golden

UNL = O.UNLABELLED


def ring(n):
    return csr_from_edges(n, [(v, (v + 1) % n) for v in range(n)], undirected=True)


def test_ring_first_step():
    n = 9
    row, col = ring(n)
    out, ch = O.step_csr(row, col, np.arange(n, dtype=np.uint32))
    want = [min((v - 1) % n, (v + 1) % n) for v in range(n)]
    assert list(out) == want and ch == n


def test_star_oscillates():
    s = 6
    row, col = csr_from_edges(s + 1, [(0, i) for i in range(1, s + 1)], undirected=True)
    lab = np.arange(s + 1, dtype=np.uint32)
    lab, _ = O.step_csr(row, col, lab)
    assert lab[0] == 1 and (lab[1:] == 0).all()
    for t in range(4):
        lab, _ = O.step_csr(row, col, lab)
        c, leaf = (0, 1) if t % 2 == 0 else (1, 0)
        assert lab[0] == c and (lab[1:] == leaf).all()


def test_complete_graph():
    n = 7
    row, col = csr_from_edges(n, [(a, b) for a in range(n) for b in range(a + 1, n)], True)
    lab = np.arange(n, dtype=np.uint32)
    l1, _ = O.step_csr(row, col, lab)
    assert l1[0] == 1 and (l1[1:] == 0).all()
    l2, _ = O.step_csr(row, col, l1)
    assert (l2 == 0).all()


def test_isolated_keeps_label():
    row, col = csr_from_edges(5, [(0, 1)], True)
    out, ch = O.step_csr(row, col, np.array([10, 20, 30, 40, 50], np.uint32))
    assert list(out) == [20, 10, 30, 40, 50] and ch == 2


def test_semi_supervised_path():
    n = 10
    row, col = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)], True)
    lab = np.full(n, UNL, np.uint32)
    lab[0] = 7
    seeds = (lab != UNL).astype(np.uint8)
    c = O.cfg(semi=True)
    for t in range(1, n):
        lab, _ = O.step_csr(row, col, lab, c, seeds)
        assert (lab[:t + 1] == 7).all() and (lab[t + 1:] == UNL).all()


def test_seed_stays_clamped():
    row, col = csr_from_edges(3, [(0, 1), (0, 2)], True)
    lab = np.array([5, 9, 9], np.uint32)
    seeds = np.array([1, 0, 0], np.uint8)
    out, _ = O.step_csr(row, col, lab, O.cfg(semi=True), seeds)
    assert out[0] == 5 and list(out[1:]) == [5, 5]


def random_graph(rng, n, m):
    return csr_from_edges(n, rng.integers(0, n, (m, 2)), undirected=True)


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_exact(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 120))
    row, col = random_graph(rng, n, int(rng.integers(0, 6 * n)))
    K = int(rng.integers(1, 4))
    lab = rng.integers(0, 6, n * K).astype(np.uint32)      # few labels -> many ties
    semi = bool(seed % 2)
    seeds = None
    if semi:
        lab[rng.random(n * K) < 0.5] = UNL
        seeds = (lab != UNL).astype(np.uint8)
    for _ in range(3):
        want = brute_mode_step(row, col, lab, K, seeds)
        got, ch = O.step_csr(row, col, lab, O.cfg(K=K, semi=semi), seeds)
        assert np.array_equal(got, want)
        assert ch == int((got != lab).sum())
        lab = got


def test_batched_equals_independent_runs():
    rng = np.random.default_rng(7)
    n, K = 300, 5
    row, col = random_graph(rng, n, 900)
    lab = rng.integers(0, 40, (n, K)).astype(np.uint32)
    allk, _ = O.lp_run_csr(row, col, lab.reshape(-1), 6, O.cfg(K=K))
    for k in range(K):
        one, _ = O.lp_run_csr(row, col, lab[:, k].copy(), 6, O.cfg(K=1))
        assert np.array_equal(allk.reshape(n, K)[:, k], one)


def test_cms_width_one_gives_smallest_label():
    # every label collides in every row -> equal estimates -> smallest present
    rng = np.random.default_rng(1)
    row, col = random_graph(rng, 200, 2000)
    lab = rng.integers(0, 1000, 200).astype(np.uint32)
    got, _ = O.step_csr(row, col, lab, O.cfg(cms_thresh=1, cms_depth=3, cms_width=1))
    deg = np.diff(row.astype(np.int64))
    for v in range(200):
        nb = col[row[v]:row[v + 1]]
        if deg[v] > 1:
            assert got[v] == lab[nb].min()


def test_cms_without_collisions_is_exact():
    rng = np.random.default_rng(2)
    row, col = random_graph(rng, 400, 4000)
    lab = rng.integers(0, 12, 400).astype(np.uint32)
    depth, width, seed = 4, 1 << 16, 99
    ab = O.cms_params(seed, depth)
    labels = np.arange(12, dtype=np.uint64)
    for r in range(depth):   # the 12 labels land in distinct columns of some row
        cols = (ab[2 * r] * labels + ab[2 * r + 1]) >> np.uint64(48)
        if len(set(cols.tolist())) == 12:
            break
    else:
        pytest.skip("all rows collide for this seed")
    exact, _ = O.step_csr(row, col, lab)
    cms, _ = O.step_csr(row, col, lab, O.cfg(cms_thresh=1, cms_depth=depth, cms_width=width,
                                             cms_seed=seed))
    assert np.array_equal(exact, cms)


def test_cms_params_odd_multipliers():
    ab = O.cms_params(12345, 4)
    assert all(int(a) & 1 for a in ab[0::2])
    assert len(set(ab.tolist())) == 8


@pytest.mark.parametrize("case", golden_cases())
def test_cgr_step_equals_csr_step(case):
    """decoding the reference-written file set gives the same LP step as the CSR"""
    g = load_golden(case)
    view = O.CgrView(g["stream"], int(g["end_offs"][-1]) if g["n_c"] else 0,
                     np.concatenate([[0], g["end_offs"]]), g["vertices"],
                     g["small_offs"].astype(np.uint64), g["small_nbrs"])
    rng = np.random.default_rng(0)
    for K in (1, 3):
        lab = rng.integers(0, 9, g["n"] * K).astype(np.uint32)
        a, ca = O.step_cgr(view, lab, O.cfg(K=K))
        b, cb = O.step_csr(g["row"], g["clm"], lab, O.cfg(K=K))
        assert np.array_equal(a, b) and ca == cb
