"""Learned schedule: the device MLP against the float64 CPU reference, and
labels that do not depend on the schedule."""
import numpy as np
import pytest

import lpa_amd as L
import mlp_ref as R
from gpu_common import run_and_compare

pytestmark = pytest.mark.gpu


def test_mlp_forward_matches_reference():
    m = L.engine.Mlp(0, seed=3)
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, R.NPARAM).astype(np.float32)
    m.params = p
    x = rng.normal(0, 2, (5000, 8)).astype(np.float32)
    y = m.forward(x)
    want, (_, _, _, _, h2) = R.forward(p.astype(np.float64), x.astype(np.float64))
    # relative error <= 1e-5, relative to the magnitude of the terms each
    # output sums (sum |W3 h2| + |b3|), so outputs that cancel to near zero
    # are held to the same float32 rounding as the others
    w = R.unpack(p.astype(np.float64))
    scale = np.abs(h2) @ np.abs(w["W3"]).T + np.abs(w["b3"])
    assert (np.abs(y - want) <= 1e-5 * scale).all(), np.max(np.abs(y - want) / scale)


def test_mlp_gradient_matches_reference():
    m = L.engine.Mlp(0, seed=4)
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, R.NPARAM).astype(np.float32)
    m.params = p
    x = rng.normal(0, 1, (3000, 8)).astype(np.float32)
    t = rng.normal(0, 1, (3000, 3)).astype(np.float32)
    mask = rng.random((3000, 3)) < 0.6
    loss, g = m.loss_grad(x, t, mask)
    want_l, want_g = R.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), mask)
    assert abs(loss - want_l) <= 1e-4 * abs(want_l)
    scale = np.abs(want_g).max()
    assert np.allclose(g, want_g, rtol=1e-3, atol=1e-5 * scale)


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 10, (20000, 8)).astype(np.float32)
    t = np.stack([x[:, 0] * 0.5 + 1, np.abs(x[:, 1] - x[:, 3]), 0.2 * x[:, 2] * x[:, 4] / 10], 1).astype(np.float32)
    mask = np.ones_like(t, np.uint8)
    m = L.engine.Mlp(0, seed=5)
    losses = m.train(x, t, mask, epochs=30, batch=512, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0]
    # the recorded loss is the reference loss of the trained parameters
    final = R.loss(m.params.astype(np.float64), x.astype(np.float64), t.astype(np.float64), mask.astype(bool))
    assert final < 0.2 * losses[0]


def test_learned_schedule_is_bit_exact():
    g = L.gen_rmat(14, 16, seed=21)
    cgr = g.encode(0, 128)
    e = L.Engine(0)
    e.upload(cgr)
    e.config(K=2, cms_thresh=2000, cms_depth=4, cms_width=512, cms_seed=1)
    e.schedule_degree()
    e.labels_init()
    samples = L.schedule.record_samples(e)
    model, losses = L.schedule.train_cost_model([samples], epochs=15, batch=1024)
    assert np.isfinite(losses).all() and losses[-1] < losses[0]
    e.schedule_learned(model)
    run_and_compare(cgr, 4, K=2, cms_thresh=2000, cms_depth=4, cms_width=512, cms_seed=1, engine=e)


def test_learned_schedule_gain_is_measured():
    """what the trained network gains over the degree-threshold schedule it
    replaces: both time the same K = 8 steps of R-MAT 18 (HIP events, median
    of 5) from the same labels and must produce the same labels; the ratio is
    printed.  The learned schedule keeps a vertex in its degree family unless
    the model is confident, so it must not lose much either (> 0.8x)."""
    g = L.gen_rmat(18, 16, 0.57, 0.19, 0.19, seed=1, bfs=True)
    cgr = g.encode(0, 128)
    cms = dict(cms_thresh=1024, cms_depth=4, cms_width=1024, cms_seed=0)
    e = L.Engine(0)
    e.upload(cgr)
    e.config(K=8, **cms)
    e.schedule_degree()
    e.labels_init()
    samples = L.schedule.record_samples(e)
    model, _ = L.schedule.train_cost_model([samples], epochs=20, batch=2048)
    out = {}
    for name in ("degree", "learned"):
        if name == "degree":
            e.schedule_degree()
        else:
            e.schedule_learned(model)
        e.labels_init()
        e.nova_iterate()
        ms, _ = e.time_iterations(5)
        out[name] = (float(np.median(ms)), e.labels(), e.counts())
    gain = out["degree"][0] / out["learned"][0]
    print(f"[learned] degree {out['degree'][0]:.3f} ms, learned {out['learned'][0]:.3f} ms, "
          f"gain {gain:.3f}x; bins degree {out['degree'][2]} learned {out['learned'][2]}", flush=True)
    assert np.array_equal(out["degree"][1], out["learned"][1])
    assert gain > 0.8
    e.close()


def test_judged_learned_schedule_keeps_labels():
    """the degree schedule judges the learned layout: both are timed from the
    current labels, which come back unchanged, and the kept layout gives the
    oracle's labels"""
    g = L.gen_rmat(14, 16, seed=22)
    cgr = g.encode(0, 128)
    e = L.Engine(0)
    e.upload(cgr)
    e.config(K=3, cms_thresh=1500, cms_depth=4, cms_width=512, cms_seed=2)
    e.schedule_degree()
    e.labels_init()
    samples = L.schedule.record_samples(e)
    model, _ = L.schedule.train_cost_model([samples], epochs=10, batch=1024)
    e.labels_init()
    e.nova_iterate()
    before = e.labels()
    kept, tl, td = e.schedule_learned_judged(model, iters=3)
    print(f"[judge] learned {tl:.3f} ms, degree {td:.3f} ms, kept {'learned' if kept else 'degree'}", flush=True)
    assert tl > 0 and td > 0 and kept == (tl <= td)
    assert np.array_equal(e.labels(), before)
    run_and_compare(cgr, 3, K=3, cms_thresh=1500, cms_depth=4, cms_width=512, cms_seed=2, engine=e)
    # an engine with labels but no schedule yet
    f = L.Engine(0)
    f.upload(cgr)
    f.config(K=3, cms_thresh=1500, cms_depth=4, cms_width=512, cms_seed=2)
    f.labels_init()
    kept, tl, td = f.schedule_learned_judged(model, iters=2)
    assert tl > 0 and td > 0
    run_and_compare(cgr, 2, K=3, cms_thresh=1500, cms_depth=4, cms_width=512, cms_seed=2, engine=f)
