"""The device CGR record encoder (encode_kernels.hpp), built for the host:
record lengths and bits against the reference compressor's golden file
sets, and against the host encoder on generated graphs."""
import ctypes as C
import os

import numpy as np
import pytest

import lpa_amd as L
from helpers import golden_cases, load_golden

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libdecode_check.so")


@pytest.fixture(scope="module")
def dc():
    if not os.path.exists(LIB):
        pytest.skip("tests/native not built (make -C tests/native)")
    lib = C.CDLL(LIB)
    P = C.c_void_p
    lib.dc_encode_bits.restype = C.c_uint64
    lib.dc_encode_bits.argtypes = [C.c_uint32, P, C.c_uint32]
    lib.dc_encode_record.restype = None
    lib.dc_encode_record.argtypes = [C.c_uint32, P, C.c_uint32, P, C.c_uint64, C.c_uint64]
    return lib


def encode_stream(dc, row, col, order):
    """encodes the records of the vertices in `order` back to back; returns
    (stream bytes, start offsets)"""
    col = np.ascontiguousarray(col, np.uint32)
    starts = [0]
    for v in order:
        a, b = int(row[v]), int(row[v + 1])
        starts.append(starts[-1] + dc.dc_encode_bits(int(v), col[a:].ctypes.data, b - a))
    nbits = starts[-1]
    words = np.zeros((nbits + 63) // 64 + 2, np.uint64)
    for i, v in enumerate(order):
        a, b = int(row[v]), int(row[v + 1])
        dc.dc_encode_record(int(v), col[a:].ctypes.data, b - a, words.ctypes.data, starts[i], starts[i + 1])
    raw = words.astype(">u8").tobytes()[:(nbits + 7) // 8]
    return np.frombuffer(raw, np.uint8), np.array(starts, np.uint64)


@pytest.mark.parametrize("case", golden_cases())
def test_golden_streams(dc, case):
    g = load_golden(case)
    order = g["vertices"][:g["n_c"]]
    stream, starts = encode_stream(dc, g["row"], g["clm"], order)
    assert np.array_equal(starts[1:], g["end_offs"])
    assert np.array_equal(stream, g["stream"][:len(stream)])
    assert not g["stream"][len(stream):].any()   # the reference pads the last byte only


@pytest.mark.parametrize("gen", ["rmat", "lattice", "copying"])
def test_generated_vs_host_encoder(dc, gen):
    if gen == "rmat":
        g = L.gen_rmat(12, 16, seed=5)
    elif gen == "lattice":
        g = L.gen_lattice(64, 0.1, 0.01, seed=3)
    else:
        g = L.gen_copying(4000, 20.0, 2.1, 2000, 0.7, 100, seed=4)
    T = 16 if gen == "copying" else 0
    cgr = g.encode(T, 128)
    order = cgr.vertices[:cgr.n_c]
    stream, starts = encode_stream(dc, g.row, g.col, order)
    assert np.array_equal(starts, cgr.offsets)
    host = np.frombuffer(cgr.stream_bytes(), np.uint8)
    assert np.array_equal(stream, host[:len(stream)])


# --------------------------------------------------------------- long rows

def _long_row(rng, n, d_target):
    """sorted distinct ids: runs of 1-3 ids (residual entries), 4-20, and a
    few of hundreds to thousands (interval entries), separated by gaps of
    2 to 10,000"""
    out, x = [], int(rng.integers(0, 1000))
    while len(out) < d_target:
        r = rng.random()
        run = (1 if r < 0.55 else 2 if r < 0.65 else 3 if r < 0.7 else
               int(rng.integers(4, 21)) if r < 0.97 else int(rng.integers(100, 5000)))
        gap = int(rng.integers(2, 10000)) if rng.random() < 0.9 else 2
        if x + run >= n:
            break
        out.extend(range(x, x + run))
        x += run + gap
    return np.array(out, np.uint32)


@pytest.fixture(scope="module")
def dcl(dc):
    dc.dc_encode_long.restype = C.c_uint64
    dc.dc_encode_long.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                  C.c_uint64]
    return dc


def _both(dc, v, adj, n, T, start):
    d = len(adj)
    adj = np.ascontiguousarray(adj, np.uint32) if d else np.zeros(1, np.uint32)
    bits = dc.dc_encode_bits(v, adj.ctypes.data, d)
    nw = (start + bits) // 64 + 4
    w1 = np.zeros(nw, np.uint64)
    w2 = np.zeros(nw, np.uint64)
    dc.dc_encode_record(v, adj.ctypes.data, d, w1.ctypes.data, start, start + bits)
    lbits = dc.dc_encode_long(v, adj.ctypes.data, d, n, T, w2.ctypes.data, start)
    return bits, lbits, w1, w2


@pytest.mark.parametrize("T", [1, 7, 64, 1024])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_long_row_phases_match_record_encoder(dcl, T, seed):
    """the workgroup phases (lane by lane) give the record encoder's bits
    and length, at unaligned record starts, with ids up to 2^30"""
    rng = np.random.default_rng(seed)
    n = (1 << 30) - 1
    for d_target, start in ((70000, 13), (3000, 64 * 5 + 37), (200000, 0)):
        adj = _long_row(rng, n, d_target)
        v = int(rng.integers(0, n))
        bits, lbits, w1, w2 = _both(dcl, v, adj, n, T, start)
        assert lbits == bits, (d_target, T)
        assert np.array_equal(w1, w2), (d_target, T)


@pytest.mark.parametrize("case", ["empty", "one", "one_run", "singles", "runs_of_3_and_4", "top_ids"])
def test_long_row_edge_cases(dcl, case):
    n = 1 << 30
    adj = {"empty": np.zeros(0, np.uint32), "one": np.array([5], np.uint32),
           "one_run": np.arange(100, 20100, dtype=np.uint32),
           "singles": np.arange(0, 400000, 2, dtype=np.uint32),
           "runs_of_3_and_4": np.concatenate([np.arange(10 * k, 10 * k + (3 + k % 2)) for k in range(5000)]).astype(np.uint32),
           "top_ids": np.arange(n - 3000, n, 3, dtype=np.uint32)}[case]
    for T in (1, 64, 1024):
        for v in (0, 7, n - 1):
            bits, lbits, w1, w2 = _both(dcl, v, adj, n, T, 29)
            assert lbits == bits and np.array_equal(w1, w2), (case, T, v)


def test_long_row_rejects_bad_rows(dcl):
    n = 1000
    for adj in (np.array([3, 2, 5], np.uint32), np.array([3, 3, 5], np.uint32), np.array([1, 2, 1000], np.uint32)):
        for T in (1, 2, 1024):
            assert dcl.dc_encode_long(0, adj.ctypes.data, len(adj), n, T, None, 0) == 2**64 - 1
