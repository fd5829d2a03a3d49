"""Out-of-core operation (lpa_stream_residence): the CGR stream in pinned
host memory, copied to the device batch by batch every step.  Every label
and change count is compared with the oracle after every step, with batches
small enough that bins, heavy records' decode tasks and the two staging
buffers' turns are split many ways; moving the stream between HBM and host
memory mid-run changes nothing."""
import numpy as np
import pytest

import lpa_amd as L
from gpu_common import make_engine, run_and_compare
from test_gpu_parity import _hub_graph

pytestmark = pytest.mark.gpu


def host_engine(cgr, batch_bytes, **kw):
    e = make_engine(cgr, **kw)
    e.stream_residence(True, batch_bytes)
    nb, stage = e.stream_batches()
    print(f"[gpu-case] host stream: {nb} batches, {stage} staging bytes", flush=True)
    assert nb >= 1 and stage > 0
    return e


@pytest.mark.parametrize("K", [1, 8])
@pytest.mark.parametrize("cms", [False, True])
def test_hub_graph_small_batches(K, cms):
    """4 KiB batches: the 70,000-id hub record is a batch of its own, the
    other heavy records' decode tasks fall into many batches"""
    cgr = _hub_graph().encode(0, 128)
    cfg = dict(cms_thresh=2500, cms_depth=4, cms_width=1024, cms_seed=4) if cms else {}
    e = host_engine(cgr, 4096, K=K, **cfg)
    assert e.stream_batches()[0] > 100
    run_and_compare(cgr, 4, K=K, engine=e, **cfg)


def test_rmat_cms_k8_one_mib_batches():
    cgr = L.gen_rmat(19, 16, seed=41, bfs=True).encode(0, 128)
    cfg = dict(cms_thresh=1000, cms_depth=4, cms_width=512, cms_seed=5)
    e = host_engine(cgr, 1 << 20, K=8, **cfg)
    assert e.stream_batches()[0] > 4
    run_and_compare(cgr, 4, K=8, engine=e, **cfg)


def test_semi_supervised_and_small_csr_part():
    """copying-model graph with a small-vertex CSR (T = 32), 5 vectors,
    semi-supervised, fused and two-kernel steps"""
    g = L.gen_copying(20000, 30.0, 2.1, 3000, 0.7, 200, seed=2)
    cgr = g.encode(32, 128)
    n, K = cgr.n, 5
    rng = np.random.default_rng(6)
    lab = np.full((n, K), 0xFFFFFFFF, np.uint32)
    for k in range(K):
        idx = rng.choice(n, n // 20, replace=False)
        lab[idx, k] = rng.integers(0, 10, len(idx), dtype=np.uint32)
    for method in ("nova", "cglp"):
        e = host_engine(cgr, 16384, K=K, semi=True)
        run_and_compare(cgr, 5, labels0=lab.reshape(-1), K=K, semi=True, method=method, engine=e)


def test_residence_set_before_upload_and_moved_back():
    """the stream goes to host memory at upload (never to HBM) and is laid
    out batch by batch; moving it between HBM and host memory mid-run, and
    re-scheduling on the host, keeps every step equal to an HBM engine's"""
    cgr = L.gen_rmat(15, 16, seed=42).encode(0, 128)
    a = L.Engine(0)
    a.stream_residence(True, 8192)
    a.upload(cgr)
    a.config(K=3)
    a.schedule_degree()
    b = make_engine(cgr, K=3)
    a.labels_init()
    b.labels_init()

    def same(steps):
        for _ in range(steps):
            assert a.nova_iterate() == b.nova_iterate()
            assert np.array_equal(a.labels(), b.labels())

    same(3)
    a.stream_residence(False)
    b.stream_residence(True, 65536)
    assert a.stream_batches() == (0, 0) and b.stream_batches()[0] > 1
    same(2)
    b.schedule_degree(lane_max=16, wave_max=256, bin_order=1)
    same(2)
    ms, ch = b.time_iterations(2)
    a.time_iterations(2)
    assert np.array_equal(a.labels(), b.labels())
    a.close()
    b.close()


def test_device_encoded_graph_on_host():
    """a graph CGR-encoded on the device, then moved to host memory"""
    g = L.gen_rmat(14, 16, seed=43)
    e = L.Engine(0)
    e.stream_residence(True, 4096)
    e.encode_csr(g.row, g.col, 0)
    e.config(K=2)
    e.schedule_degree()
    run_and_compare(g.encode(0, 128), 3, K=2, engine=e)


def test_graph_capture_refused_on_host():
    cgr = L.gen_rmat(12, 16, seed=44).encode(0, 128)
    e = host_engine(cgr, 4096)
    e.labels_init()
    with pytest.raises(L.LpaError):
        e.set_graph_mode(True)
    with pytest.raises(L.LpaError):
        e.time_iterations(1, use_graph=True)
    e.stream_residence(False)
    e.set_graph_mode(True)
    with pytest.raises(L.LpaError):
        e.stream_residence(True, 4096)
    e.close()


def test_rmat22_k8_sixteen_mib_batches():
    """R-MAT scale 22 (128 M arcs, 240 MB of stream) in 16 MiB batches, K = 8,
    CMS 4 x 1024 above degree 4096: fused and two-kernel steps"""
    cgr = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, bfs=True).encode(0, 128)
    cfg = dict(cms_thresh=4096, cms_depth=4, cms_width=1024, cms_seed=0)
    for method in ("nova", "cglp"):
        e = host_engine(cgr, 16 << 20, K=8, **cfg)
        assert e.stream_batches()[0] >= 10
        run_and_compare(cgr, 3, K=8, method=method, engine=e, **cfg)
        e.close()
