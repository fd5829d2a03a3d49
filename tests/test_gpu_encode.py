"""CGR encoding on the device (lpa_graph_encode_csr): the stream, offsets
and vertex order equal the host encoder's (itself byte-identical to the
reference compressor), LP on the device-encoded graph matches the oracle,
and malformed rows are rejected."""
import numpy as np
import pytest

import lpa_amd as L
from gpu_common import run_and_compare
from helpers import golden_cases, load_golden

pytestmark = pytest.mark.gpu


def host_words64(cgr):
    b = bytes(cgr.stream_bytes())
    b += b"\0" * ((-len(b)) % 8)
    return np.frombuffer(b, ">u8").astype(np.uint64)


def check_same_stream(e, cgr):
    words, offs, vert, nbits = e.stream()
    assert nbits == cgr.nbits
    assert np.array_equal(offs, cgr.offsets)
    assert np.array_equal(vert, cgr.vertices)
    hw = host_words64(cgr)
    nw = (nbits + 63) // 64
    assert np.array_equal(words[:nw], hw[:nw])


@pytest.mark.parametrize("case", golden_cases())
def test_golden_cases(case):
    g = load_golden(case)
    csr = L.Csr.from_arrays(g["row"], g["clm"])
    cgr = csr.encode(g["params"]["small_thresh"], 128)
    e = L.Engine(0)
    e.encode_csr(g["row"], g["clm"], g["params"]["small_thresh"])
    check_same_stream(e, cgr)
    e.close()


def _long_rows_graph(seed=25):
    """background edges plus five hubs whose rows are longer than the
    per-lane encoder's limit (65,536 ids): mixes of short runs, intervals
    and runs of thousands of ids, so the workgroup kernels encode them"""
    rng = np.random.default_rng(seed)
    n = 2_000_000
    src = [rng.integers(0, n, 1_000_000)]
    dst = [rng.integers(0, n, 1_000_000)]
    for h, (d, runmax) in enumerate(((70_000, 8), (150_000, 40), (300_000, 3), (90_000, 5000), (66_000, 1))):
        ids, x = [], int(rng.integers(10, 100))
        while len(ids) < d and x < n - runmax:
            run = int(rng.integers(1, runmax + 1))
            ids.extend(range(x, x + run))
            x += run + int(rng.integers(2, 30))
        ids = np.array(ids[:d], np.int64)
        src.append(np.full(len(ids), h + 1))
        dst.append(ids)
    src = np.concatenate(src).astype(np.uint32)
    dst = np.concatenate(dst).astype(np.uint32)
    return L.Csr.from_edges(n, src, dst, False)


@pytest.mark.parametrize("gen,T", [("rmat", 0), ("lattice", 0), ("copying", 32), ("hubs", 0), ("long_rows", 0)])
def test_generated_graphs(gen, T):
    if gen == "rmat":
        g = L.gen_rmat(16, 16, seed=21)
    elif gen == "lattice":
        g = L.gen_lattice(512, 0.1, 0.005, seed=22)
    elif gen == "copying":
        g = L.gen_copying(100000, 30.0, 2.1, 20000, 0.7, 500, seed=23)
    elif gen == "long_rows":
        g = _long_rows_graph()
        assert (g.degrees() > 65536).sum() == 5
    else:
        rng = np.random.default_rng(24)
        n = 300000
        src = np.concatenate([rng.integers(0, n, 600000), np.zeros(150000, np.int64)])
        dst = np.concatenate([rng.integers(0, n, 600000), rng.choice(np.arange(1, n), 150000, replace=False)])
        g = L.Csr.from_edges(n, src.astype(np.uint32), dst.astype(np.uint32), True)
    cgr = g.encode(T, 128)
    e = L.Engine(0)
    e.encode_csr(g.row, g.col, T)
    check_same_stream(e, cgr)
    assert np.array_equal(e.degrees(), g.degrees().astype(np.uint32))
    e.close()


def test_long_rows_in_several_scratch_groups(monkeypatch):
    """the long rows' scratch limited to 200,000 entries: the five hub rows
    are encoded in groups, each measured and written in turn"""
    monkeypatch.setenv("LPA_LONG_SCRATCH", "200000")
    g = _long_rows_graph()
    cgr = g.encode(0, 128)
    e = L.Engine(0)
    e.encode_csr(g.row, g.col, 0)
    check_same_stream(e, cgr)
    e.close()


def test_lp_on_device_encoded_graph():
    g = L.gen_rmat(14, 16, seed=25)
    cgr = g.encode(0, 128)
    e = L.Engine(0)
    e.encode_csr(g.row, g.col, 0)
    e.config(K=2, cms_thresh=200, cms_depth=4, cms_width=256, cms_seed=9)
    e.schedule_degree()
    run_and_compare(cgr, 4, K=2, cms_thresh=200, cms_depth=4, cms_width=256, cms_seed=9, engine=e)


def test_malformed_rows_rejected():
    row = np.array([0, 2, 3, 3], np.uint64)
    for col in (np.array([2, 1, 0], np.uint32),     # unsorted row
                np.array([1, 1, 0], np.uint32),     # duplicate id
                np.array([1, 7, 0], np.uint32)):    # id >= n
        e = L.Engine(0)
        with pytest.raises(L.LpaError):
            e.encode_csr(row, col, 0)
        e.close()
