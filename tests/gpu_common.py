"""Helpers of the GPU parity tests: build a graph, run the engine and the
oracle on the same input, compare bit for bit."""
import numpy as np

import lpa_amd as L
import pyoracle as O


def oracle_view(cgr):
    return O.CgrView(np.frombuffer(cgr.stream_bytes(), np.uint8), cgr.nbits, cgr.offsets.copy(),
                     cgr.vertices.copy(), cgr.small_offsets.copy(), cgr.small_nbrs.copy())


def make_engine(cgr, K=1, semi=False, cms_thresh=0, cms_depth=4, cms_width=1024, cms_seed=0,
                schedule=None, **sched):
    e = L.Engine(0)
    e.upload(cgr)
    e.config(K=K, semi=semi, cms_thresh=cms_thresh, cms_depth=cms_depth, cms_width=cms_width,
             cms_seed=cms_seed)
    if schedule is None:
        e.schedule_degree(**sched)
    else:
        e.schedule_bins(schedule)
    return e


def run_and_compare(cgr, iters, labels0=None, method="nova", K=1, semi=False, cms_thresh=0,
                    cms_depth=4, cms_width=1024, cms_seed=0, schedule=None, engine=None, **sched):
    """runs `iters` synchronous steps on the GPU and on the oracle; asserts
    identical labels and change counts after every step"""
    print(f"[gpu-case] n={cgr.n} n_c={cgr.n_c} nbits={cgr.nbits} K={K} semi={semi} "
          f"cms={cms_thresh} method={method} sched={sched or 'bins' if schedule is not None else sched}",
          flush=True)
    e = engine or make_engine(cgr, K, semi, cms_thresh, cms_depth, cms_width, cms_seed, schedule, **sched)
    e.labels_init(labels0)
    view = oracle_view(cgr)
    c = O.cfg(K=K, semi=semi, cms_thresh=cms_thresh, cms_depth=cms_depth, cms_width=cms_width,
              cms_seed=cms_seed)
    lab = (np.repeat(np.arange(cgr.n, dtype=np.uint32), K) if labels0 is None
           else np.ascontiguousarray(labels0, np.uint32).reshape(-1).copy())
    seeds = (lab != O.UNLABELLED).astype(np.uint8) if semi else None
    for it in range(iters):
        ch = {"nova": e.nova_iterate, "cglp": e.cglp_iterate, "csr": e.csr_iterate}[method]()
        lab, och = O.step_cgr(view, lab, c, seeds)
        got = e.labels()
        if not np.array_equal(got, lab):
            bad = np.nonzero(got != lab)[0]
            raise AssertionError(f"iteration {it}: {len(bad)} labels differ, first at {bad[:5]} "
                                 f"gpu={got[bad[:5]]} oracle={lab[bad[:5]]}")
        assert ch == och, (it, ch, och)
    return e, lab
