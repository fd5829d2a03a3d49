"""Pins the oracle's CGR codec against the reference compressor's own output
(golden file sets from src/compressor.cpp, built under oracle/_ref), the
hand-derived record of vertex 5, and the closed forms of the code lengths."""
import numpy as np
import pytest

import pyoracle as O
from helpers import csr_from_edges, golden_cases, load_golden


def test_code_lengths_closed_form():
    # gamma: 0^L then L+1 bits of x+1 -> 2L+1; zeta_3: 0^h 1 then 3(h+1) bits -> 4(h+1)
    for x in list(range(1 << 14)) + [(1 << s) - 2 for s in range(15, 40)] + [(1 << s) - 1 for s in range(15, 40)]:
        L = (x + 1).bit_length() - 1
        assert O.gamma_len(x) == 2 * L + 1
        assert O.zeta_len(x) == (L // 3 + 1) * 4


def test_hand_derived_vertex5():
    # N(5) = {1,6,7,8,9,12}: interval (6,4) and residuals {1,12}; 28-bit record
    # 1010 0111 1011 0100 1000 0100 1011 after five 4-bit empty records "1111"
    row, col = csr_from_edges(13, [(5, x) for x in (1, 6, 7, 8, 9, 12)], undirected=False)
    g = O.write_cgr(row, col, 0)
    assert g["nbits"] == 12 * 4 + 28
    assert bytes(g["stream"]) == bytes.fromhex("fffffa7b484bfffffff0")
    nb, end = O.decode_record(g["stream"], g["nbits"], 20, 5)
    assert list(nb) == [6, 7, 8, 9, 1, 12] and end == 48
    # loaded as one big-endian word at bit 20..47: 0xA7B484B0 when aligned
    bits = int.from_bytes(bytes(g["stream"]), "big") >> (80 - 48)
    assert (bits & ((1 << 28) - 1)) << 4 == 0xA7B484B0


@pytest.mark.parametrize("case", golden_cases())
def test_golden_csr_matches_edge_list(case):
    g = load_golden(case)
    row, col = csr_from_edges(g["n"], g["edges"], undirected=bool(g["params"]["one_way"]))
    assert np.array_equal(row, g["row"]) and np.array_equal(col, g["clm"])


@pytest.mark.parametrize("case", golden_cases())
def test_golden_encoder_bytes(case):
    g = load_golden(case)
    t = g["params"]["small_thresh"]
    w = O.write_cgr(g["row"], g["clm"], t, 128)
    assert w["n_c"] == g["n_c"]
    assert np.array_equal(w["end_offs"], g["end_offs"])
    assert bytes(w["stream"]) == bytes(g["stream"])
    assert np.array_equal(w["vertices"], g["vertices"])
    if g["n_c"] < g["n"]:
        assert np.array_equal(w["small_offs"], g["small_offs"])
        assert np.array_equal(w["small_nbrs"], g["small_nbrs"])


@pytest.mark.parametrize("case", golden_cases())
def test_golden_decode_every_record(case):
    """decode(.graph) == CSR row for every compressed vertex (the reference's
    only test, test/decompress_test.cu:81-160, applied to all vertices)"""
    g = load_golden(case)
    starts = np.concatenate([[0], g["end_offs"]]).astype(np.uint64)
    nbits = int(g["end_offs"][-1]) if g["n_c"] else 0
    for i in range(g["n_c"]):
        v = int(g["vertices"][i])
        nb, end = O.decode_record(g["stream"], nbits, int(starts[i]), v)
        assert end == int(starts[i + 1])
        want = g["clm"][int(g["row"][v]):int(g["row"][v + 1])]
        assert np.array_equal(np.sort(nb), want), (case, v)


def test_golden_cases_exercise_format():
    """the fixtures reach multi-segment records, merged last segments and
    deg >= 128 vertices"""
    g = load_golden("multiseg6000_t0")
    deg = np.diff(g["row"].astype(np.int64))
    assert deg.max() > 2000
    rec = np.diff(np.concatenate([[0], g["end_offs"].astype(np.int64)]))
    assert rec.max() > 10 * 256
    mixed = load_golden("mixed400_t32")
    assert mixed["n_c"] < mixed["n"] and len(mixed["small_nbrs"]) > 0


def test_random_round_trip():
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(2, 3000))
        m = int(rng.integers(0, 20 * n))
        runs = []
        for _ in range(int(rng.integers(0, n))):
            s = int(rng.integers(0, n))
            v = int(rng.integers(0, n))
            runs += [(v, x) for x in range(s, min(n, s + int(rng.integers(1, 40))))]
        e = np.concatenate([rng.integers(0, n, (m, 2)), np.array(runs, np.int64).reshape(-1, 2)])
        row, col = csr_from_edges(n, e, undirected=bool(trial % 2))
        w = O.write_cgr(row, col, int(rng.choice([0, 8, 32])))
        starts = np.concatenate([[0], w["end_offs"]])
        for i in range(w["n_c"]):
            v = int(w["vertices"][i])
            nb, end = O.decode_record(w["stream"], w["nbits"], int(starts[i]), v)
            assert end == starts[i + 1]
            assert np.array_equal(np.sort(nb), col[int(row[v]):int(row[v + 1])])


def test_truncated_stream_is_an_error():
    g = load_golden("mixed400_t0")
    starts = np.concatenate([[0], g["end_offs"]])
    i = int(np.argmax(np.diff(starts)))
    with pytest.raises(ValueError):
        O.decode_record(g["stream"], int(starts[i + 1]) - 3, int(starts[i]), int(g["vertices"][i]))
