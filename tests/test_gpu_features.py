"""The engine's cost-model features (lpa_graph_features, computed on the
device while validating the stream, and lpa_model_features, the model's
inputs) against the oracle's CPU extraction, through the C ABI.

Counts (degree, intervals, residuals, bits, segments) are exact.  The mean
interval length is one IEEE division on both sides (compared to 1 ulp); the
log2 columns use the device's and the host's log2, compared to a relative
1e-6.
"""
import os

import numpy as np
import pytest

import lpa_amd as L
import pyoracle as O
from gpu_common import make_engine
from helpers import golden_cases, load_golden

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_features(e, row, col, T):
    want = O.record_features(row, col, T)
    got = e.features()
    assert np.array_equal(got[:, :6], want[:, :6])
    assert np.allclose(got[:, 6], want[:, 6], rtol=2 ** -23, atol=0)
    assert np.allclose(got[:, 7], want[:, 7], rtol=1e-6, atol=0)
    return want


@pytest.mark.parametrize("case", golden_cases())
def test_features_golden(case):
    g = load_golden(case)
    T = g["params"]["small_thresh"]
    cgr = L.Cgr.load(os.path.join(GOLDEN, case, "edges.txt"), T)
    e = make_engine(cgr)
    check_features(e, g["row"], g["clm"], T)
    e.close()


@pytest.mark.parametrize("graph", ["rmat", "copying"])
def test_features_and_model_inputs(graph):
    if graph == "rmat":
        g, T = L.gen_rmat(14, 16, seed=31), 0
    else:
        g, T = L.gen_copying(20000, 30.0, 2.1, 3000, 0.7, 200, seed=2), 32
    cgr = g.encode(T, 128)
    e = make_engine(cgr)
    want = check_features(e, g.row, g.col, T)
    e.labels_init()
    e.nova_iterate()
    x = e.model_features()
    assert np.allclose(x[:, :7], np.log2(1.0 + want[:, :7]), rtol=1e-6, atol=1e-7)
    # column 7: did vector 0's label change in the last step
    lab1 = e.labels()
    assert np.array_equal(x[:, 7], (lab1 != np.arange(g.n, dtype=np.uint32)).astype(np.float32))
    e.close()


def test_gather_calibration():
    """the random-gather measurement runs for every row width and rejects
    bad arguments"""
    e = L.Engine(0)
    for words in (1, 2, 4, 8, 16):
        ms = e.calibration_gather(1 << 24, words, 1 << 20, 2)
        assert 0 < ms < 1000
    for bad in ((1 << 24, 3, 1 << 20), (0, 8, 1 << 20), (1 << 24, 8, 0)):
        with pytest.raises(L.LpaError):
            e.calibration_gather(*bad)
    e.close()
