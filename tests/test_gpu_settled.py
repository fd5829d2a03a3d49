"""Parity in the settled regime, where a vertex's own label is usually a
strict majority of its neighbours and every kernel family settles such
vectors without a table or sketch (the own-label check).  Long runs from
L0 = v pass through the unsettled steps into the settled ones; every label
and change count is compared with the oracle after every step.

Count-min sketch vectors are settled by the own label only when no present
label shares its cell in every row; a sketch 4 cells wide and 2 rows deep
makes such collisions common, so the fallback to full estimation runs too.
"""
import numpy as np
import pytest

import lpa_amd as L
from gpu_common import run_and_compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat18():
    return L.gen_rmat(18, 16, seed=51, bfs=True).encode(0, 128)


@pytest.mark.parametrize("K", [1, 3, 8])
def test_settled_exact(rmat18, K):
    """lane / wave / workgroup kernels and the exact one-workgroup hub kernel
    (degrees up to 25,252; the exact chunks above 8,192)"""
    run_and_compare(rmat18, 12, K=K)


@pytest.mark.parametrize("K", [1, 8])
def test_settled_cms(rmat18, K):
    """count-min sketch above degree 1,000 (one-workgroup CMS kernel)"""
    run_and_compare(rmat18, 12, K=K, cms_thresh=1000, cms_depth=4, cms_width=256, cms_seed=3)


@pytest.mark.parametrize("K", [1, 8])
def test_settled_cms_full_row_collisions(rmat18, K):
    """width 4, depth 2: a label shares the own label's cells in both rows
    with probability 1/16, so settled-looking vectors often need the full
    estimate"""
    run_and_compare(rmat18, 10, K=K, cms_thresh=200, cms_depth=2, cms_width=4, cms_seed=9)


def test_settled_semi_supervised(rmat18):
    """5 vectors, 5 % seeds from 10 classes, clamped; exact and sketched"""
    n, K = rmat18.n, 5
    rng = np.random.default_rng(8)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, n // 20, replace=False)
        lab[idx, k] = rng.integers(0, 10, len(idx), dtype=np.uint32)
    run_and_compare(rmat18, 10, labels0=lab.reshape(-1), K=K, semi=True)
    run_and_compare(rmat18, 10, labels0=lab.reshape(-1), K=K, semi=True, cms_thresh=1000, cms_depth=4,
                    cms_width=256, cms_seed=3)


def test_settled_small_csr_part():
    """copying-model graph with a small-vertex CSR (T = 32) and CMS hubs"""
    g = L.gen_copying(20000, 30.0, 2.1, 3000, 0.7, 200, seed=2)
    run_and_compare(g.encode(32, 128), 10, K=3, cms_thresh=1000, cms_depth=4, cms_width=512, cms_seed=3)
