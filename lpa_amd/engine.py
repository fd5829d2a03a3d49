"""ctypes binding of liblpa_hip.so (include/lpa_amd.h): the gfx950 engine.

The product path: every call lands in the HIP library; there is no CPU
fallback.  Loading fails loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

from .host import Cgr, Csr, LpaError, gen_rmat

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_VP = C.c_void_p

BIN_SKIP, BIN_LANE, BIN_WAVE, BIN_BLOCK, BIN_GLOBAL, BIN_CMS = range(6)
BIN_NAMES = ["skip", "lane", "wave", "block", "global", "cms"]
UNLABELLED = 0xFFFFFFFF

EXPORTS = {
    "lpa_last_error": (C.c_char_p, []),
    "lpa_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "lpa_engine_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
    "lpa_engine_destroy": (C.c_int, [_VP]),
    "lpa_engine_info": (C.c_int, [_VP, _VP]),
    "lpa_graph_upload": (C.c_int, [_VP, C.c_uint64, C.c_uint64, _VP, C.c_uint64, _VP, _VP, _VP, _VP]),
    "lpa_graph_degrees": (C.c_int, [_VP, _VP]),
    "lpa_graph_features": (C.c_int, [_VP, _VP]),
    "lpa_lp_config": (C.c_int, [_VP, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
    "lpa_schedule_degree": (C.c_int, [_VP, _VP]),
    "lpa_schedule_bins": (C.c_int, [_VP, _VP, C.c_uint32]),
    "lpa_schedule_counts": (C.c_int, [_VP, _VP]),
    "lpa_schedule_stats": (C.c_int, [_VP, _VP]),
    "lpa_labels_init": (C.c_int, [_VP, _VP]),
    "lpa_labels_get": (C.c_int, [_VP, _VP]),
    "lpa_nova_iterate": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "lpa_cglp_decode": (C.c_int, [_VP]),
    "lpa_cglp_iterate": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "lpa_csr_iterate": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "lpa_cglp_csr_get": (C.c_int, [_VP, _VP, _VP, C.POINTER(C.c_uint64)]),
    "lpa_time_iterations": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _VP, _VP]),
    "lpa_set_graph_mode": (C.c_int, [_VP, C.c_int]),
    "lpa_stream_residence": (C.c_int, [_VP, C.c_int, C.c_uint64]),
    "lpa_stream_batches": (C.c_int, [_VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "lpa_autotune": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "lpa_model_features": (C.c_int, [_VP, _VP]),
    "lpa_record_family": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "lpa_mlp_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(_VP)]),
    "lpa_mlp_destroy": (C.c_int, [_VP]),
    "lpa_mlp_get_params": (C.c_int, [_VP, _VP]),
    "lpa_mlp_set_params": (C.c_int, [_VP, _VP]),
    "lpa_mlp_forward": (C.c_int, [_VP, _VP, C.c_uint64, _VP]),
    "lpa_mlp_loss_grad": (C.c_int, [_VP, _VP, _VP, _VP, C.c_uint64, _VP, _VP]),
    "lpa_mlp_train": (C.c_int, [_VP, _VP, _VP, _VP, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_uint64, _VP]),
    "lpa_schedule_learned": (C.c_int, [_VP, _VP]),
    "lpa_schedule_learned_judged": (C.c_int, [_VP, _VP, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                              C.POINTER(C.c_int)]),
    "lpa_last_kernel_times": (C.c_int, [_VP, _VP]),
    "lpa_calibration_copy": (C.c_int, [_VP, C.c_uint64, C.c_int, _VP]),
    "lpa_calibration_gather": (C.c_int, [_VP, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, _VP]),
    "lpa_graph_encode_csr": (C.c_int, [_VP, C.c_uint64, _VP, _VP, C.c_uint32]),
    "lpa_graph_stream": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "lpa_encode_csr_device": (C.c_int, [_VP, C.c_uint64, C.c_uint64, _VP, _VP, _VP, _VP, _VP, C.c_uint64,
                                        C.POINTER(C.c_uint64)]),
    "lpa_csr_bfs_relabel": (C.c_int, [_VP, C.c_uint64, _VP, _VP, C.c_uint32, _VP, _VP, _VP]),
}


class SchedParams(C.Structure):
    _fields_ = [("lane_max", C.c_uint32), ("wave_max", C.c_uint32),
                ("block_max", C.c_uint32), ("relayout", C.c_uint32), ("quarter_max", C.c_uint32),
                ("bin_order", C.c_uint32), ("xcd_tiles", C.c_uint32)]


def library_path():
    return os.path.join(_HERE, "liblpa_hip.so")


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise LpaError(f"{path} is not built: the HIP engine is required (no CPU fallback)")
    L = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _LIB = L
    return L


# lpa_time_iterations methods
METHODS = {"nova": 0, "cglp": 1, "csr": 2}


def _check(rc):
    if rc != 0:
        raise LpaError(f"lpa engine error {rc}: {load().lpa_last_error().decode()}")


def device_count():
    c = C.c_int(0)
    _check(load().lpa_device_count(C.byref(c)))
    return c.value


def _p(a):
    return None if a is None else a.ctypes.data


class Engine:
    """One engine per device: graph, schedule, label buffers, stream."""

    def __init__(self, device=0):
        self._L = load()
        h = _VP()
        _check(self._L.lpa_engine_create(device, C.byref(h)))
        self._h = h
        self.n = 0
        self.K = 1

    def close(self):
        if getattr(self, "_h", None):
            self._L.lpa_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self):
        x = np.zeros(4, np.uint64)
        _check(self._L.lpa_engine_info(self._h, _p(x)))
        return dict(zip(("compute_units", "lds_bytes", "wave_cap", "block_cap"), map(int, x)))

    # graph -------------------------------------------------------------
    def upload(self, cgr: Cgr):
        self.upload_arrays(cgr.n, cgr.n_c, cgr.words, cgr.nwords, cgr.offsets, cgr.vertices,
                           cgr.small_offsets, cgr.small_nbrs)

    def encode_csr(self, row, col, small_thresh=0):
        """CGR-encodes a CSR on the device and makes it the engine's graph
        (the stream is the reference compressor's, bit for bit)"""
        row = np.ascontiguousarray(row, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        if len(col) == 0:
            col = np.zeros(1, np.uint32)
        n = len(row) - 1
        _check(self._L.lpa_graph_encode_csr(self._h, n, _p(row), _p(col), small_thresh))
        self.n = n

    def bfs_relabel(self, row, col, root=0):
        """BFS relabelling of a CSR on the device: (new_row, new_col, perm),
        perm[v] = new id of v -- the host's Csr.bfs_relabel, bit for bit"""
        row = np.ascontiguousarray(row, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        n = len(row) - 1
        new_row = np.zeros(n + 1, np.uint64)
        new_col = np.zeros(max(len(col), 1), np.uint32)
        perm = np.zeros(max(n, 1), np.uint32)
        _check(self._L.lpa_csr_bfs_relabel(self._h, n, _p(row), _p(col if len(col) else new_col), root,
                                           _p(new_row), _p(new_col), _p(perm)))
        return new_row, new_col[:len(col)], perm[:n]

    def stream(self):
        """(words u64 [ceil(nbits/64)], offsets u64 [n_c+1], vertices u32 [n], nbits)"""
        n_c, nbits = C.c_uint64(0), C.c_uint64(0)
        _check(self._L.lpa_graph_stream(self._h, C.byref(n_c), C.byref(nbits), None, None, None))
        words = np.zeros(max((nbits.value + 63) // 64, 1), np.uint64)
        offs = np.zeros(n_c.value + 1, np.uint64)
        vert = np.zeros(self.n, np.uint32)
        _check(self._L.lpa_graph_stream(self._h, C.byref(n_c), C.byref(nbits), _p(words), _p(offs), _p(vert)))
        return words[:(nbits.value + 63) // 64], offs, vert, nbits.value

    def upload_arrays(self, n, n_c, words32, nwords, offsets, vertices, small_offsets=None,
                      small_nbrs=None):
        words32 = np.ascontiguousarray(words32, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        vertices = np.ascontiguousarray(vertices, np.uint32)
        so = None if small_offsets is None else np.ascontiguousarray(small_offsets, np.uint64)
        sn = None if small_nbrs is None else np.ascontiguousarray(small_nbrs, np.uint32)
        if sn is not None and len(sn) == 0:
            sn = np.zeros(1, np.uint32)
        _check(self._L.lpa_graph_upload(self._h, n, n_c, _p(words32), nwords, _p(offsets),
                                        _p(vertices), _p(so), _p(sn)))
        self.n = n

    def degrees(self):
        d = np.zeros(self.n, np.uint32)
        _check(self._L.lpa_graph_degrees(self._h, _p(d)))
        return d

    def features(self):
        f = np.zeros((self.n, 8), np.float32)
        _check(self._L.lpa_graph_features(self._h, _p(f)))
        return f

    # configuration -----------------------------------------------------
    def config(self, K=1, semi=False, cms_thresh=0, cms_depth=4, cms_width=1024, cms_seed=0):
        _check(self._L.lpa_lp_config(self._h, K, int(semi), cms_thresh, cms_depth, cms_width, cms_seed))
        self.K = K

    def schedule_degree(self, lane_max=32, wave_max=512, block_max=0, quarter_max=128, bin_order=0, xcd_tiles=0):
        p = SchedParams(lane_max, wave_max, block_max, 1, quarter_max, bin_order, xcd_tiles)
        _check(self._L.lpa_schedule_degree(self._h, C.byref(p)))

    def autotune(self, iters=3):
        """fastest degree schedule for this graph; labels are left unchanged"""
        p = SchedParams()
        ms = C.c_float(0)
        _check(self._L.lpa_autotune(self._h, iters, C.byref(p), C.byref(ms)))
        return dict(lane_max=p.lane_max, wave_max=p.wave_max, quarter_max=p.quarter_max, bin_order=p.bin_order,
                    xcd_tiles=p.xcd_tiles), ms.value

    def set_graph_mode(self, on=True):
        _check(self._L.lpa_set_graph_mode(self._h, int(on)))

    def stream_residence(self, host, batch_bytes=0):
        """keep the CGR stream in pinned host memory (host=True) and stream
        it to the device in batches of at most batch_bytes every step (0:
        1 GiB), or back in HBM (host=False)"""
        _check(self._L.lpa_stream_residence(self._h, 1 if host else 0, int(batch_bytes)))

    def stream_batches(self):
        """(batches of the current layout, bytes of both staging buffers)"""
        nb, sb = C.c_uint64(), C.c_uint64()
        _check(self._L.lpa_stream_batches(self._h, C.byref(nb), C.byref(sb)))
        return nb.value, sb.value

    def schedule_bins(self, bins):
        b = np.ascontiguousarray(bins, np.uint8)
        _check(self._L.lpa_schedule_bins(self._h, _p(b), 1))

    def counts(self):
        c = np.zeros(6, np.uint64)
        _check(self._L.lpa_schedule_counts(self._h, _p(c)))
        return dict(zip(BIN_NAMES, c.tolist()))

    def stats(self):
        """per bin kind: vertices, arcs, stream bits, small-CSR arcs"""
        st = np.zeros((6, 4), np.uint64)
        _check(self._L.lpa_schedule_stats(self._h, _p(st)))
        return {name: dict(zip(("vertices", "arcs", "bits", "csr_arcs"), map(int, st[i])))
                for i, name in enumerate(BIN_NAMES)}

    # labels --------------------------------------------------------------
    def labels_init(self, labels=None):
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint32).reshape(-1)
        if lab is not None and len(lab) != self.n * self.K:
            raise LpaError("labels must have n*K entries")
        _check(self._L.lpa_labels_init(self._h, _p(lab)))

    def labels(self):
        out = np.zeros(self.n * self.K, np.uint32)
        _check(self._L.lpa_labels_get(self._h, _p(out)))
        return out

    # steps ---------------------------------------------------------------
    def nova_iterate(self):
        c = C.c_uint64(0)
        _check(self._L.lpa_nova_iterate(self._h, C.byref(c)))
        return c.value

    def cglp_decode(self):
        _check(self._L.lpa_cglp_decode(self._h))

    def cglp_iterate(self):
        c = C.c_uint64(0)
        _check(self._L.lpa_cglp_iterate(self._h, C.byref(c)))
        return c.value

    def csr_iterate(self):
        """one step of LP over the uncompressed CSR (decoded once, reused)"""
        c = C.c_uint64(0)
        _check(self._L.lpa_csr_iterate(self._h, C.byref(c)))
        return c.value

    def cglp_csr(self):
        row = np.zeros(self.n + 1, np.uint64)
        m = C.c_uint64(0)
        _check(self._L.lpa_cglp_csr_get(self._h, _p(row), None, C.byref(m)))
        col = np.zeros(max(m.value, 1), np.uint32)
        _check(self._L.lpa_cglp_csr_get(self._h, _p(row), _p(col), C.byref(m)))
        return row, col[:m.value]

    def time_iterations(self, iters, method="nova", use_graph=False):
        ms = np.zeros(iters, np.float32)
        ch = np.zeros(iters, np.uint64)
        _check(self._L.lpa_time_iterations(self._h, METHODS[method], iters,
                                           int(use_graph), _p(ms), _p(ch)))
        return ms, ch

    # learned schedule ------------------------------------------------------
    def model_features(self):
        x = np.zeros((self.n, 8), np.float32)
        _check(self._L.lpa_model_features(self._h, _p(x)))
        return x

    def record_family(self, family):
        """one step with every vertex forced into kernel family 0/1/2 and the
        recorder on: (lane-cycles per vertex, measured mask)"""
        cost = np.zeros(self.n, np.uint64)
        meas = np.zeros(self.n, np.uint8)
        _check(self._L.lpa_record_family(self._h, family, _p(cost), _p(meas)))
        return cost, meas.astype(bool)

    def schedule_learned(self, mlp):
        _check(self._L.lpa_schedule_learned(self._h, mlp._h))

    def schedule_learned_judged(self, mlp, iters=5):
        """learned layout unless the degree layout's median step (iters
        steps from the current labels, which are put back) is faster;
        returns (kept_learned, learned_ms, degree_ms)"""
        tl, td, k = C.c_float(0), C.c_float(0), C.c_int(0)
        _check(self._L.lpa_schedule_learned_judged(self._h, mlp._h, iters, C.byref(tl), C.byref(td), C.byref(k)))
        return bool(k.value), tl.value, td.value

    def kernel_times(self):
        t = np.zeros(6, np.float32)
        _check(self._L.lpa_last_kernel_times(self._h, _p(t)))
        return dict(zip(BIN_NAMES, t.tolist()))

    def calibration_copy(self, nbytes=256 << 20, reps=3):
        """plain device copy of nbytes, reps times; returns ms per copy"""
        ms = C.c_float(0)
        _check(self._L.lpa_calibration_copy(self._h, nbytes, reps, C.byref(ms)))
        return ms.value

    def calibration_gather(self, nbytes, row_words, gathers, reps=3):
        """random gathers of rows of row_words 32-bit words from a buffer of
        nbytes; returns ms per `gathers` reads"""
        ms = C.c_float(0)
        _check(self._L.lpa_calibration_gather(self._h, int(nbytes), int(row_words), int(gathers), reps,
                                              C.byref(ms)))
        return ms.value


MLP_NPARAM = 1443


class Mlp:
    """8 -> 32 -> 32 -> 3 cost model with device forward / backward / Adam."""

    def __init__(self, device=0, seed=0):
        self._L = load()
        h = _VP()
        _check(self._L.lpa_mlp_create(device, seed, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.lpa_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        _check(self._L.lpa_mlp_get_params(self._h, _p(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (MLP_NPARAM,)
        _check(self._L.lpa_mlp_set_params(self._h, _p(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), 3), np.float32)
        _check(self._L.lpa_mlp_forward(self._h, _p(x), len(x), _p(y)))
        return y

    def loss_grad(self, x, t, mask):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        m = np.ascontiguousarray(mask, np.uint8)
        loss = np.zeros(1, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        _check(self._L.lpa_mlp_loss_grad(self._h, _p(x), _p(t), _p(m), len(x), _p(loss), _p(g)))
        return float(loss[0]), g

    def train(self, x, t, mask, epochs=20, batch=1024, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        m = np.ascontiguousarray(mask, np.uint8)
        losses = np.zeros(epochs, np.float32)
        _check(self._L.lpa_mlp_train(self._h, _p(x), _p(t), _p(m), len(x), epochs, batch, lr, seed, _p(losses)))
        return losses


def gen_rmat_device_bfs(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, device=0):
    """R-MAT graph drawn on the host (gen_rmat with bfs=False) and
    BFS-relabelled from vertex 0 on the device: the graph gen_rmat(...,
    bfs=True) builds, bit for bit, without the host's BFS"""
    raw = gen_rmat(scale, edge_factor, a, b, c, seed=seed, bfs=False)
    e = Engine(device)
    try:
        row, col, _ = e.bfs_relabel(raw.row, raw.col, 0)
    finally:
        e.close()
    del raw
    return Csr.from_arrays(row, col)
