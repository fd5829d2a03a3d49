"""The allocation-free device encoder, lpa_encode_csr_device, against the
allocating one it varies (lpa_graph_encode_csr, read back with
lpa_graph_stream): the same stream bits and offsets, bit for bit, on two
calls in a row, the second with a smaller graph (the library's scratch
keeps the first call's contents, which must not leak into the second).

Inputs and outputs are torch device tensors (device memory plumbing); the
engine's stream does not order with torch's, so the test synchronises.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import lpa_amd as L

pytestmark = pytest.mark.gpu


def stream_order(row, T):
    deg = np.diff(row.astype(np.int64))
    big = np.nonzero(deg >= 128)[0]
    mid = np.nonzero((deg >= T) & (deg < 128))[0]
    return np.concatenate([big, mid]).astype(np.uint32)


def encode_no_alloc(e, row, col, T, words_cap=None):
    n = len(row) - 1
    vid = stream_order(row, T)
    dev = torch.device("cuda:0")
    d_row = torch.from_numpy(row.view(np.int64)).to(dev)
    d_col = torch.from_numpy(col.view(np.int32) if len(col) else np.zeros(1, np.int32)).to(dev)
    d_vid = torch.from_numpy(vid.view(np.int32) if len(vid) else np.zeros(1, np.int32)).to(dev)
    d_offs = torch.full((len(vid) + 1,), -1, dtype=torch.int64, device=dev)
    if words_cap is None:
        words_cap = int(row[-1]) * 2 + len(vid) + 8   # ample: < 64 bits per arc plus headers
    d_words = torch.full((words_cap,), -1, dtype=torch.int64, device=dev)   # garbage to be cleared
    torch.cuda.synchronize()
    nbits = C.c_uint64(0)
    rc = e._L.lpa_encode_csr_device(e._h, n, len(vid), d_row.data_ptr(), d_col.data_ptr(), d_vid.data_ptr(),
                                    d_offs.data_ptr(), d_words.data_ptr(), words_cap, C.byref(nbits))
    torch.cuda.synchronize()
    if rc:
        return rc, None, None, None
    nb = nbits.value
    words = d_words.cpu().numpy().view(np.uint64)[:(nb + 63) // 64]
    return 0, d_offs.cpu().numpy().view(np.uint64), words, nb


def encode_allocating(row, col, T):
    e = L.Engine(0)
    e.encode_csr(row, col, T)
    words, offs, vert, nbits = e.stream()
    e.close()
    return offs, words, nbits


@pytest.mark.parametrize("T", [0, 32])
def test_no_alloc_encoder_matches_allocating(T):
    e = L.Engine(0)
    for g in (L.gen_rmat(15, 16, seed=41), L.gen_rmat(12, 8, seed=42)):   # second call smaller
        print(f"[gpu-case] no-alloc encode n={g.n} m={g.m} T={T}", flush=True)
        rc, offs, words, nbits = encode_no_alloc(e, g.row, g.col, T)
        assert rc == 0, L.engine.load().lpa_last_error()
        w_offs, w_words, w_nbits = encode_allocating(g.row, g.col, T)
        assert nbits == w_nbits
        assert np.array_equal(offs, w_offs)
        assert np.array_equal(words, w_words[:len(words)])
    e.close()


def test_no_alloc_encoder_rejects_small_output_and_bad_rows():
    e = L.Engine(0)
    g = L.gen_rmat(10, 8, seed=43)
    rc, *_ = encode_no_alloc(e, g.row, g.col, 0, words_cap=4)
    assert rc == -5 and b"words" in L.engine.load().lpa_last_error()   # LPA_ERR_RANGE
    col = g.col.copy()
    v = int(np.argmax(np.diff(g.row.astype(np.int64))))
    col[g.row[v]], col[g.row[v] + 1] = col[g.row[v] + 1], col[g.row[v]]   # an unsorted row
    rc, *_ = encode_no_alloc(e, g.row, col, 0)
    assert rc == -3   # LPA_ERR_FORMAT
    # the next call after a rejected one is clean
    rc, offs, words, nbits = encode_no_alloc(e, g.row, g.col, 0)
    w_offs, w_words, w_nbits = encode_allocating(g.row, g.col, 0)
    assert rc == 0 and nbits == w_nbits and np.array_equal(offs, w_offs)
    assert np.array_equal(words, w_words[:len(words)])
    e.close()
