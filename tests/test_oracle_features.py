"""Pins the oracle's cost-model features (no reference counterpart).

* The hand-derived record of vertex 5 with N = {1, 6, 7, 8, 9, 12}: one
  interval (6, length 4), residuals 1 and 12, 28 bits, one segment per
  block, mean interval length 4.
* On every reference-written golden file set: the bit length of each record
  equals the reference compressor's offsets; #intervals / #residuals equal a
  run split done here with numpy; the segment counts equal the headers read
  from the reference's .graph bytes by a bit reader written here from the
  format definition (gamma(#segments - 1) opens each block; segments are
  256 bits apart; the residual block starts where the last interval segment
  ends).
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import csr_from_edges, golden_cases, load_golden


class Bits:
    def __init__(self, stream):
        self.bits = np.unpackbits(np.asarray(stream, np.uint8))
        self.pos = 0

    def gamma(self):
        z = 0
        while self.bits[self.pos + z] == 0:
            z += 1
        v = 0
        for b in self.bits[self.pos + z:self.pos + 2 * z + 1]:
            v = 2 * v + int(b)
        self.pos += 2 * z + 1
        return v - 1


def block_segments(stream, start):
    """(#interval segments, #residual segments) of the record at bit start"""
    r = Bits(stream)
    r.pos = start
    si = r.gamma() + 1
    base = r.pos
    r.pos = base + 256 * (si - 1)   # the last interval segment
    c = r.gamma()
    for _ in range(2 * c):          # left code and length code per interval
        r.gamma()
    sr = r.gamma() + 1
    return si, sr


def runs(adj):
    """(#intervals (runs >= 4), #residuals, total interval length)"""
    adj = np.asarray(adj, np.int64)
    if len(adj) == 0:
        return 0, 0, 0
    cut = np.nonzero(np.diff(adj) != 1)[0] + 1
    lens = np.diff(np.concatenate([[0], cut, [len(adj)]]))
    itv = lens[lens >= 4]
    return len(itv), int(lens[lens < 4].sum()), int(itv.sum())


def test_hand_derived_vertex5():
    row, col = csr_from_edges(13, [(5, u) for u in (1, 6, 7, 8, 9, 12)], undirected=False)
    f = O.record_features(row, col)
    assert list(f[5][:7]) == [6, 1, 2, 28, 1, 1, 4]
    assert f[5][7] == pytest.approx(np.log2(7.0), rel=1e-6)
    # an isolated vertex is "1111": no intervals, one empty segment per block
    assert list(f[0][:7]) == [0, 0, 0, 4, 1, 1, 0]


@pytest.mark.parametrize("case", golden_cases())
def test_features_against_reference_files(case):
    g = load_golden(case)
    T = g["params"]["small_thresh"]
    f = O.record_features(g["row"], g["clm"], T)
    deg = np.diff(g["row"].astype(np.int64))
    starts = np.concatenate([[0], g["end_offs"]]).astype(np.int64)
    for i in range(g["n_c"]):
        v = int(g["vertices"][i])
        adj = g["clm"][g["row"][v]:g["row"][v + 1]]
        ni, nr, isum = runs(adj)
        si, sr = block_segments(g["stream"], int(starts[i]))
        want = [deg[v], ni, nr, starts[i + 1] - starts[i], si, sr, (isum / ni) if ni else 0.0]
        assert np.allclose(f[v][:7], want, rtol=1e-7, atol=0), (case, v, f[v], want)
    for v in g["vertices"][g["n_c"]:]:
        d = deg[v]
        assert list(f[v][:7]) == [d, 0, d, 32 * d, 0, 0, 0]
    assert np.allclose(f[:, 7], np.log2(1.0 + deg), rtol=1e-6)
