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
