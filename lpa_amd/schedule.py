"""Learned schedule: record per-vertex work, train the cost model, bin.

The engine records, for one step with every vertex forced into one kernel
family, each vertex's share of the device time (lpa_record_family: every
bin's measured time split among its vertices in proportion to the
lane-cycles they took, in picoseconds).  The model (8 -> 32 -> 32 -> 3,
trained on the device) learns log2(1 + cost) per family from the vertex's
CGR features; the learned schedule puts every vertex in the family of
least predicted cost.  Results do not depend on the
schedule: every family computes the same labels.
"""
import numpy as np

from .engine import Mlp


def record_samples(engine, max_samples=200000, seed=0):
    """features (n x 8), targets (n x 3) and mask (n x 3) of the engine's
    current graph, from one recorded step per kernel family; vertices of
    degree 0 and CMS vertices are left out.  Subsampled to max_samples."""
    x = engine.model_features()
    t = np.zeros((engine.n, 3), np.float32)
    m = np.zeros((engine.n, 3), bool)
    for fam in range(3):
        cost, meas = engine.record_family(fam)
        t[:, fam] = np.log2(1.0 + cost.astype(np.float64)).astype(np.float32)
        m[:, fam] = meas & (cost > 0)
    keep = np.nonzero(m.any(axis=1))[0]
    if len(keep) > max_samples:
        keep = np.sort(np.random.default_rng(seed).choice(keep, max_samples, replace=False))
    return x[keep], t[keep], m[keep]


def train_cost_model(samples, epochs=40, batch=2048, lr=2e-3, seed=0, device=0):
    """samples: list of (x, t, mask) from record_samples on one or more graphs"""
    x = np.concatenate([s[0] for s in samples])
    t = np.concatenate([s[1] for s in samples])
    m = np.concatenate([s[2] for s in samples])
    model = Mlp(device, seed)
    losses = model.train(x, t, m.astype(np.uint8), epochs=epochs, batch=batch, lr=lr, seed=seed)
    return model, losses
