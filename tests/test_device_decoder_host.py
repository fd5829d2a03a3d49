"""The device CGR decoders (cgr_decode.hpp), built for the host, against the
oracle: serial record decode and the segment-parallel order the group
kernels use, on the reference's golden file sets and on generated graphs."""
import ctypes as C
import os

import numpy as np
import pytest

import lpa_amd as L
import pyoracle as O
from helpers import golden_cases, load_golden

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libdecode_check.so")


@pytest.fixture(scope="module")
def dc():
    if not os.path.exists(LIB):
        pytest.skip("tests/native not built (make -C tests/native)")
    lib = C.CDLL(LIB)
    P = C.c_void_p
    lib.dc_decode_record.restype = C.c_int64
    lib.dc_decode_record.argtypes = [P, C.c_uint64, C.c_uint32, P, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.dc_decode_segments.restype = C.c_int64
    lib.dc_decode_segments.argtypes = [P, C.c_uint64, C.c_uint32, P, C.c_uint64]
    return lib


def words64(stream_bytes):
    b = bytes(stream_bytes)
    b += b"\0" * ((-len(b)) % 8) + b"\0" * 64
    return np.frombuffer(b, ">u8").astype(np.uint64)


def check_all(dc, stream, nbits, starts, vertices):
    w = words64(stream)
    cap = 1 << 16
    out = np.zeros(cap, np.uint32)
    end = C.c_uint64()
    for i in range(len(starts) - 1):
        v = int(vertices[i])
        want, want_end = O.decode_record(np.frombuffer(bytes(stream), np.uint8), nbits, int(starts[i]), v, cap=cap)
        d = dc.dc_decode_record(w.ctypes.data, int(starts[i]), v, out.ctypes.data, cap, C.byref(end))
        assert d == len(want) and np.array_equal(out[:d], want) and end.value == want_end, (i, v)
        d2 = dc.dc_decode_segments(w.ctypes.data, int(starts[i]), v, out.ctypes.data, cap)
        assert d2 == len(want) and np.array_equal(np.sort(out[:d2]), np.sort(want)), (i, v)


@pytest.mark.parametrize("case", golden_cases())
def test_golden(dc, case):
    g = load_golden(case)
    starts = np.concatenate([[0], g["end_offs"]]).astype(np.uint64)
    nbits = int(g["end_offs"][-1]) if g["n_c"] else 0
    check_all(dc, g["stream"], nbits, starts, g["vertices"])


@pytest.mark.parametrize("gen", ["rmat", "lattice", "copying"])
def test_generated(dc, gen):
    if gen == "rmat":
        g = L.gen_rmat(13, 16, seed=7)
    elif gen == "lattice":
        g = L.gen_lattice(64, 0.1, 0.005, seed=1)
    else:
        g = L.gen_copying(8000, 30.0, 2.1, 4000, 0.7, 200, seed=3)
    c = g.encode(0, 128)
    check_all(dc, c.stream_bytes(), c.nbits, c.offsets, c.vertices)
