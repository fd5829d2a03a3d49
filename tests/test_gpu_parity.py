"""GPU parity: the gfx950 engine against the CPU oracle on the same inputs,
bit for bit, every iteration, through the C ABI."""
import os

import numpy as np
import pytest

import lpa_amd as L
import pyoracle as O
from gpu_common import make_engine, oracle_view, run_and_compare
from helpers import golden_cases, load_golden

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("case", golden_cases())
def test_reference_file_sets(case):
    g = load_golden(case)
    cgr = L.Cgr.load(os.path.join(GOLDEN, case, "edges.txt"), g["params"]["small_thresh"])
    run_and_compare(cgr, 5)


@pytest.mark.parametrize("sched", [dict(lane_max=32, wave_max=256), dict(lane_max=0, wave_max=0, block_max=1),
                                   dict(lane_max=0, wave_max=512), dict(lane_max=8, wave_max=32, block_max=64),
                                   dict(lane_max=64, wave_max=512), dict(bin_order=1, xcd_tiles=1),
                                   dict(bin_order=2, xcd_tiles=1)])
def test_each_kernel_family(sched):
    g = L.gen_rmat(13, 16, seed=11)
    run_and_compare(g.encode(0, 128), 4, **sched)


@pytest.mark.parametrize("K", [16, 32])
@pytest.mark.parametrize("sched", [dict(lane_max=0, wave_max=512), dict(lane_max=8, wave_max=128, quarter_max=64),
                                   dict(lane_max=8, wave_max=32, block_max=64)])
def test_group_kernels_several_batches(K, sched):
    """quarter-wave, half-wave, wave and workgroup groups counting two
    (K = 16) or four (K = 32) batches of 8 vectors per vertex"""
    g = L.gen_rmat(13, 16, seed=11)
    run_and_compare(g.encode(0, 128), 3, K=K, **sched)


def test_group_kernels_several_batches_csr_part():
    """the same over the small-vertex CSR part (T = 32) of a copying-model
    graph, 32 vectors"""
    g = L.gen_copying(20000, 30.0, 2.1, 3000, 0.7, 200, seed=2)
    run_and_compare(g.encode(32, 128), 3, K=32)


def test_device_degrees_and_decode_to_csr():
    g = L.gen_rmat(14, 16, seed=12)
    cgr = g.encode(0, 128)
    e = make_engine(cgr)
    assert np.array_equal(e.degrees(), g.degrees().astype(np.uint32))
    row, col = e.cglp_csr()
    assert np.array_equal(row, g.row)
    for v in range(0, g.n, 97):
        assert np.array_equal(np.sort(col[row[v]:row[v + 1]]), g.col[g.row[v]:g.row[v + 1]])


def test_lattice_config_full_size():
    """1024x1024 lattice, p_drop 0.10, 0.5% shortcuts, seed 0; 10 iterations"""
    g = L.gen_lattice(1024, 0.10, 0.005, seed=0)
    run_and_compare(g.encode(0, 128), 10)


def test_cglp_method_matches():
    g = L.gen_rmat(13, 16, seed=13)
    run_and_compare(g.encode(0, 128), 4, method="cglp")


def test_uncompressed_csr_method_matches():
    """LP over the uncompressed CSR (decoded once, reused every step): the
    same labels as NOVA and the oracle, exact and with CMS, K = 1 and 3"""
    g = L.gen_rmat(13, 16, seed=13)
    run_and_compare(g.encode(0, 128), 4, method="csr")
    run_and_compare(g.encode(16, 128), 3, method="csr", K=3, cms_thresh=100, cms_depth=4, cms_width=256,
                    cms_seed=3)
    run_and_compare(_hub_graph().encode(0, 128), 2, method="csr")


def test_cms_vertices():
    g = L.gen_rmat(14, 16, seed=14)
    run_and_compare(g.encode(0, 128), 4, cms_thresh=64, cms_depth=4, cms_width=256, cms_seed=7)


def test_batched_and_semi_supervised():
    g = L.gen_rmat(12, 16, seed=15)
    cgr = g.encode(0, 128)
    rng = np.random.default_rng(0)
    K = 5
    lab = np.full(g.n * K, 0xFFFFFFFF, np.uint32)
    seed_mask = rng.random(g.n * K) < 0.05
    lab[seed_mask] = rng.integers(0, 20, seed_mask.sum())
    run_and_compare(cgr, 5, labels0=lab, K=K, semi=True)
    run_and_compare(cgr, 3, K=3)
    run_and_compare(cgr, 3, K=8, lane_max=64)


def test_small_vertex_csr_part():
    g = L.gen_copying(20000, 30.0, 2.1, 3000, 0.7, 200, seed=2)
    run_and_compare(g.encode(32, 128), 4, cms_thresh=1000, cms_depth=4, cms_width=512, cms_seed=3)


def test_global_hash_bin():
    g = L.gen_rmat(15, 16, seed=16)
    cgr = g.encode(0, 128)
    bins = np.full(g.n, L.engine.BIN_GLOBAL, np.uint8)
    run_and_compare(cgr, 3, schedule=bins)


def test_malformed_stream_rejected():
    g = L.gen_rmat(10, 8, seed=17)
    cgr = g.encode(0, 128)
    words = cgr.words.copy()
    rec = np.diff(cgr.offsets.astype(np.int64))
    i = int(np.argmax(rec))
    mid = (int(cgr.offsets[i]) + int(rec[i]) // 2) // 32
    words[mid:mid + 4] = 0      # 128 zero bits inside the longest record
    e = L.Engine(0)
    with pytest.raises(L.LpaError):
        e.upload_arrays(cgr.n, cgr.n_c, words, cgr.nwords, cgr.offsets, cgr.vertices)


def test_graph_replay_matches_eager():
    """steps replayed from captured HIP graphs, with eager work and
    synchronisations in between, match the eager engine every time"""
    g = L.gen_rmat(14, 16, seed=31)
    cgr = g.encode(0, 128)
    eager = make_engine(cgr, cms_thresh=1000, cms_depth=4, cms_width=256, cms_seed=2)
    graph = make_engine(cgr, cms_thresh=1000, cms_depth=4, cms_width=256, cms_seed=2)
    eager.labels_init()
    graph.labels_init()
    graph.set_graph_mode(True)
    for it in range(8):
        a = eager.nova_iterate()
        b = graph.nova_iterate()
        assert a == b and np.array_equal(eager.labels(), graph.labels()), it
        graph.degrees()          # unrelated eager work between replays
    # the two-kernel method through graphs as well
    eager.labels_init()
    graph.labels_init()
    for it in range(4):
        assert eager.cglp_iterate() == graph.cglp_iterate()
        assert np.array_equal(eager.labels(), graph.labels()), it
    # and the uncompressed-CSR comparator
    eager.labels_init()
    graph.labels_init()
    for it in range(4):
        assert eager.csr_iterate() == graph.csr_iterate()
        assert np.array_equal(eager.labels(), graph.labels()), it


@pytest.mark.parametrize("K,cms,table_gb", [(1, False, None), (9, False, None), (8, True, None),
                                            (9, False, "0.004"), (64, False, None)])
def test_graph_replay_hubs_across_generation_wrap(K, cms, table_gb, monkeypatch):
    """steps replayed from captured HIP graphs on the hub graph (the
    one-workgroup and chunked heavy kernels: generation-tagged exact tables,
    count-min sketches and the own-label counters) match eager steps, from
    few-valued random labels, across the exact tables' generation wrap (the
    pass counter reaches 255 after 250 warm-up steps; with a 4 MiB table
    budget the hub's table holds 2 of 9 vectors, five passes a step; at
    K = 64 the degree-3,000 hub goes through the row-stash kernel, whose
    work counter is reset by its own last workgroup on every replay)"""
    cgr = _hub_graph().encode(0, 128)
    cfg = dict(cms_thresh=2500, cms_depth=4, cms_width=1024, cms_seed=4) if cms else {}
    print(f"[gpu-case] replay hubs K={K} cms={cms} table_gb={table_gb}", flush=True)
    if table_gb:
        monkeypatch.setenv("LPA_TABLE_GB", table_gb)
    eager = make_engine(cgr, K=K, **cfg)
    graph = make_engine(cgr, K=K, **cfg)
    monkeypatch.delenv("LPA_TABLE_GB", raising=False)
    eager.labels_init()
    graph.labels_init()
    for _ in range(250):
        eager.nova_iterate()
        graph.nova_iterate()
    rng = np.random.default_rng(K)
    lab = rng.integers(0, 30, cgr.n * K).astype(np.uint32)
    eager.labels_init(lab)
    graph.labels_init(lab)
    graph.set_graph_mode(True)
    for it in range(8):
        a = eager.nova_iterate()
        b = graph.nova_iterate()
        assert a == b and np.array_equal(eager.labels(), graph.labels()), it
        graph.degrees()          # unrelated eager work between replays
    eager.close()
    graph.close()


def test_graph_replay_semi_supervised_k64():
    """64 semi-supervised vectors (1 % seeds from 100 classes, clamped),
    exact mode: the batched-instance configuration, replayed"""
    g = L.gen_rmat(16, 16, seed=33, bfs=True)
    cgr = g.encode(0, 128)
    n, K = g.n, 64
    rng = np.random.default_rng(3)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, n // 100, replace=False)
        lab[idx, k] = rng.integers(0, 100, len(idx), dtype=np.uint32)
    eager = make_engine(cgr, K=K, semi=True)
    graph = make_engine(cgr, K=K, semi=True)
    eager.labels_init(lab.reshape(-1))
    graph.labels_init(lab.reshape(-1))
    graph.set_graph_mode(True)
    for it in range(10):
        a = eager.nova_iterate()
        b = graph.nova_iterate()
        assert a == b and np.array_equal(eager.labels(), graph.labels()), it
    run_and_compare(cgr, 3, labels0=lab.reshape(-1), K=K, semi=True, engine=graph)


def test_autotune_keeps_results():
    g = L.gen_rmat(14, 16, seed=32)
    cgr = g.encode(0, 128)
    e = make_engine(cgr)
    e.labels_init()
    e.nova_iterate()
    before = e.labels()
    best, ms = e.autotune(iters=2)
    assert np.array_equal(e.labels(), before) and ms > 0
    print("autotuned", best, ms)
    run_and_compare(cgr, 3, engine=e)


def _hub_graph(seed=21):
    """background of ~8 random neighbours per vertex plus hubs of degree
    3,000 / 6,000 / 70,000, so both the one-workgroup and the chunked
    heavy paths run, exact and sketched"""
    rng = np.random.default_rng(seed)
    n = 120000
    src = [rng.integers(0, n, 4 * n)]
    dst = [(src[0] + rng.integers(1, 200, 4 * n)) % n]
    for hub, d in ((0, 70000), (1, 3000), (2, 6000)):
        src.append(np.full(d, hub))
        dst.append(rng.choice(np.arange(3, n), d, replace=False))
    src = np.concatenate(src).astype(np.uint32)
    dst = np.concatenate(dst).astype(np.uint32)
    return L.Csr.from_edges(n, src, dst, True)


def test_heavy_local_and_chunked_paths():
    g = _hub_graph()
    deg = g.degrees()
    assert deg.max() > 65536 and ((deg > 2048) & (deg <= 4096)).any() and ((deg > 4096) & (deg < 65536)).any()
    cgr = g.encode(0, 128)
    run_and_compare(cgr, 4)                                  # exact: global bin above 2048
    run_and_compare(cgr, 4, cms_thresh=1000, cms_depth=4, cms_width=512, cms_seed=5)
    run_and_compare(cgr, 3, K=3, cms_thresh=2500, cms_depth=3, cms_width=1024, cms_seed=6)
    run_and_compare(cgr, 3, K=2)
    run_and_compare(g.encode(16, 128), 3, cms_thresh=1000, cms_depth=4, cms_width=512, cms_seed=8)


@pytest.mark.parametrize("K", [5, 9])
def test_partial_vector_batches(K):
    """K not a multiple of the staged batch (8 vectors in the wave groups,
    the 128 KB-table hub kernel and the CMS one-workgroup kernel, 4
    elsewhere): the last batch is partial; exact and sketched hubs"""
    cgr = _hub_graph(seed=23).encode(0, 128)
    run_and_compare(cgr, 2, K=K)
    run_and_compare(cgr, 2, K=K, cms_thresh=2500, cms_depth=4, cms_width=1024, cms_seed=4)


def test_exact_tables_across_generation_wrap():
    """exact-mode HBM tables carry an 8-bit pass generation and are cleared
    only when it wraps (every 255 passes): after 252 steps, eight steps from
    fresh random labels (few distinct values, so stale counts would change
    modes) straddle the wrap, each against the oracle"""
    g = _hub_graph()
    cgr = g.encode(0, 128)
    e = make_engine(cgr)
    e.labels_init()
    for _ in range(252):
        e.nova_iterate()
    view = oracle_view(cgr)
    c = O.cfg(K=1)
    rng = np.random.default_rng(5)
    for it in range(8):
        lab = rng.integers(0, 40, cgr.n).astype(np.uint32)
        e.labels_init(lab)
        ch = e.nova_iterate()
        exp, och = O.step_cgr(view, lab, c, None)
        assert np.array_equal(e.labels(), exp), it
        assert ch == och, it
    e.close()


def test_exact_tables_in_batches(monkeypatch):
    """a 4 MiB table budget (LPA_TABLE_GB) holds the 70,000-id hub's table
    for 2 of 9 vectors: five batches a step, each gathering the hub's rows
    again and advancing the tables' generation, which so wraps every 51
    steps; after 48 steps, eight steps from fresh random labels straddle a
    wrap, each against the oracle"""
    monkeypatch.setenv("LPA_TABLE_GB", "0.004")
    cgr = _hub_graph().encode(0, 128)
    K = 9
    e = make_engine(cgr, K=K)
    monkeypatch.delenv("LPA_TABLE_GB")
    e.labels_init()
    for _ in range(48):
        e.nova_iterate()
    view = oracle_view(cgr)
    c = O.cfg(K=K)
    rng = np.random.default_rng(6)
    for it in range(8):
        lab = rng.integers(0, 40, cgr.n * K).astype(np.uint32)
        e.labels_init(lab)
        ch = e.nova_iterate()
        exp, och = O.step_cgr(view, lab, c, None)
        assert np.array_equal(e.labels(), exp), it
        assert ch == och, it
    e.close()


def test_exact_tables_untagged(monkeypatch):
    """LPA_TABLE_TAGS=0 selects the per-pass clearing of the exact HBM tables
    (the layout used when a chunked list is 2^24 ids or longer)"""
    monkeypatch.setenv("LPA_TABLE_TAGS", "0")
    g = _hub_graph()
    cgr = g.encode(0, 128)
    run_and_compare(cgr, 3)
    run_and_compare(cgr, 2, K=3)
