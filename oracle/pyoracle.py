"""ctypes binding of the CPU oracle (liblpa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
UNLABELLED = 0xFFFFFFFF


class LPCfg(C.Structure):
    _fields_ = [("K", C.c_uint32), ("semi", C.c_int32), ("cms_thresh", C.c_uint32),
                ("cms_depth", C.c_uint32), ("cms_width", C.c_uint32),
                ("cms_seed", C.c_uint64), ("nthreads", C.c_int32)]


def cfg(K=1, semi=False, cms_thresh=0, cms_depth=4, cms_width=1024, cms_seed=0, nthreads=0):
    return LPCfg(K, int(semi), cms_thresh, cms_depth, cms_width, cms_seed, nthreads)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblpa_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(path)
        L.orc_gamma_len.restype = C.c_uint32
        L.orc_gamma_len.argtypes = [C.c_uint64]
        L.orc_zeta_len.restype = C.c_uint32
        L.orc_zeta_len.argtypes = [C.c_uint64]
        L.orc_record_bits.restype = C.c_int
        L.orc_record_bits.argtypes = [C.c_uint64, U64P, U32P, U64P]
        L.orc_record_features.restype = C.c_int
        L.orc_record_features.argtypes = [C.c_uint64, U64P, U32P, C.c_uint32,
                                          np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
        L.orc_write_cgr.restype = C.c_int
        L.orc_write_cgr.argtypes = [C.c_uint64, U64P, U32P, C.c_uint32, C.c_uint32,
                                    U8P, C.c_uint64, U64P, U32P, U32P, U32P, U64P]
        L.orc_decode_record.restype = C.c_int64
        L.orc_decode_record.argtypes = [U8P, C.c_uint64, C.c_uint64, C.c_uint32, U32P,
                                        C.c_uint64, C.POINTER(C.c_uint64)]
        L.orc_cms_params.restype = None
        L.orc_cms_params.argtypes = [C.c_uint64, C.c_uint32, U64P]
        L.orc_step_csr.restype = C.c_int64
        L.orc_step_csr.argtypes = [C.c_uint64, U64P, U32P, C.POINTER(LPCfg), C.c_void_p, U32P, U32P]
        L.orc_step_cgr.restype = C.c_int64
        L.orc_step_cgr.argtypes = [C.c_uint64, U8P, C.c_uint64, C.c_uint64, U64P, U32P,
                                   C.c_void_p, C.c_void_p, C.POINTER(LPCfg), C.c_void_p, U32P, U32P]
        L.orc_step_cgr_subset.restype = C.c_int64
        L.orc_step_cgr_subset.argtypes = [C.c_uint64, U8P, C.c_uint64, C.c_uint64, U64P, U32P,
                                          C.c_void_p, C.c_void_p, C.POINTER(LPCfg), C.c_void_p,
                                          U64P, C.c_uint64, U32P, U32P]
        L.orc_lp_run.restype = C.c_int64
        L.orc_lp_run.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                 C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(LPCfg), U32P, C.c_int32, C.c_int32, U64P]
        L.orc_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def gamma_len(x):
    return lib().orc_gamma_len(int(x))


def zeta_len(x):
    return lib().orc_zeta_len(int(x))


def record_bits(row, col):
    row = np.ascontiguousarray(row, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    n = len(row) - 1
    out = np.zeros(n, np.uint64)
    rc = lib().orc_record_bits(n, row, col, out)
    if rc:
        raise ValueError(f"orc_record_bits failed: {rc}")
    return out


def record_features(row, col, small_thresh=0):
    """(n, 8) float32 cost-model features per vertex id: degree, #intervals,
    #residuals, record bits, #interval segments, #residual segments, mean
    interval length, log2(1 + degree) (small-CSR vertices: d, 0, d, 32 d, 0,
    0, 0, log2(1 + d))"""
    row = np.ascontiguousarray(row, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    n = len(row) - 1
    out = np.zeros((n, 8), np.float32)
    rc = lib().orc_record_features(n, row, col, small_thresh, out)
    if rc:
        raise ValueError(f"orc_record_features failed: {rc}")
    return out


def write_cgr(row, col, small_thresh=0, large_thresh=128):
    """Reference file set in memory: dict with stream (bytes, MSB first),
    nbits, end_offs (n_c), vertices (n), small_offs, small_nbrs."""
    row = np.ascontiguousarray(row, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    n = len(row) - 1
    total = int(record_bits(row, col).sum())
    stream = np.zeros(total // 8 + 1, np.uint8)
    end_offs = np.zeros(n + 1, np.uint64)
    vertices = np.zeros(n, np.uint32)
    small_offs = np.zeros(n + 1, np.uint32)
    small_nbrs = np.zeros(len(col) + 1, np.uint32)
    out3 = np.zeros(3, np.uint64)
    rc = lib().orc_write_cgr(n, row, col, small_thresh, large_thresh, stream, len(stream),
                             end_offs, vertices, small_offs, small_nbrs, out3)
    if rc:
        raise ValueError(f"orc_write_cgr failed: {rc}")
    nc, nbits, ms = (int(x) for x in out3)
    return dict(stream=stream[:(nbits + 7) // 8].copy(), nbits=nbits, n_c=nc,
                end_offs=end_offs[:nc].copy(), vertices=vertices,
                small_offs=small_offs[:n - nc + 1].copy(), small_nbrs=small_nbrs[:ms].copy())


def decode_record(stream, nbits, pos, v, cap=1 << 12):
    stream = np.ascontiguousarray(stream, np.uint8)
    out = np.zeros(cap, np.uint32)
    end = C.c_uint64(0)
    d = lib().orc_decode_record(stream, nbits, pos, v, out, cap, C.byref(end))
    if d < 0:
        raise ValueError(f"malformed record at bit {pos}: {d}")
    if d > cap:
        return decode_record(stream, nbits, pos, v, cap=int(d))
    return out[:d].copy(), int(end.value)


def cms_params(seed, depth):
    ab = np.zeros(2 * depth, np.uint64)
    lib().orc_cms_params(seed, depth, ab)
    return ab


def _seeds_arg(seeds):
    if seeds is None:
        return None, None
    s = np.ascontiguousarray(seeds, np.uint8)
    return s, s.ctypes.data_as(C.c_void_p)


def step_csr(row, col, labels, c=None, seeds=None):
    c = c or cfg()
    row = np.ascontiguousarray(row, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    lab = np.ascontiguousarray(labels, np.uint32)
    out = np.empty_like(lab)
    keep, sp = _seeds_arg(seeds)
    ch = lib().orc_step_csr(len(row) - 1, row, col, C.byref(c), sp, lab, out)
    if ch < 0:
        raise ValueError(f"orc_step_csr failed: {ch}")
    return out, int(ch)


class CgrView:
    """Arrays of a CGR file set as the oracle reads them."""

    def __init__(self, stream, nbits, offs, vertices, small_offs=None, small_nbrs=None):
        self.stream = np.ascontiguousarray(stream, np.uint8)
        self.nbits = int(nbits)
        self.offs = np.ascontiguousarray(offs, np.uint64)          # n_c + 1 start offsets
        self.vertices = np.ascontiguousarray(vertices, np.uint32)
        self.n = len(self.vertices)
        self.n_c = len(self.offs) - 1
        self.small_offs = None if small_offs is None else np.ascontiguousarray(small_offs, np.uint64)
        self.small_nbrs = None if small_nbrs is None else np.ascontiguousarray(small_nbrs, np.uint32)
        if self.n_c < self.n and self.small_offs is None:
            raise ValueError("small vertices present but no small CSR")


def step_cgr(g, labels, c=None, seeds=None):
    c = c or cfg()
    lab = np.ascontiguousarray(labels, np.uint32)
    out = np.empty_like(lab)
    keep, sp = _seeds_arg(seeds)
    ch = lib().orc_step_cgr(g.n, g.stream, g.nbits, g.n_c, g.offs, g.vertices,
                            _ptr(g.small_offs), _ptr(g.small_nbrs), C.byref(c), sp, lab, out)
    if ch < 0:
        raise ValueError(f"orc_step_cgr failed: {ch}")
    return out, int(ch)


def step_cgr_subset(g, labels, positions, out, c=None, seeds=None):
    """Recompute out[v] for the vertices at the given stream positions."""
    c = c or cfg()
    lab = np.ascontiguousarray(labels, np.uint32)
    pos = np.ascontiguousarray(positions, np.uint64)
    keep, sp = _seeds_arg(seeds)
    ch = lib().orc_step_cgr_subset(g.n, g.stream, g.nbits, g.n_c, g.offs, g.vertices,
                                   _ptr(g.small_offs), _ptr(g.small_nbrs), C.byref(c), sp,
                                   pos, len(pos), lab, out)
    if ch < 0:
        raise ValueError(f"orc_step_cgr_subset failed: {ch}")
    return int(ch)


def lp_run_csr(row, col, labels, iters, c=None, until_converged=False):
    c = c or cfg()
    row = np.ascontiguousarray(row, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    lab = np.array(labels, np.uint32, copy=True)
    changed = np.zeros(max(iters, 1), np.uint64)
    it = lib().orc_lp_run(len(row) - 1, _ptr(row), _ptr(col), None, 0, 0, None, None, None,
                          None, C.byref(c), lab, iters, int(until_converged), changed)
    if it < 0:
        raise ValueError(f"orc_lp_run failed: {it}")
    return lab, changed[:it].copy()


def lp_run_cgr(g, labels, iters, c=None, until_converged=False):
    c = c or cfg()
    lab = np.array(labels, np.uint32, copy=True)
    changed = np.zeros(max(iters, 1), np.uint64)
    it = lib().orc_lp_run(g.n, None, None, _ptr(g.stream), g.nbits, g.n_c, _ptr(g.offs),
                          _ptr(g.vertices), _ptr(g.small_offs), _ptr(g.small_nbrs),
                          C.byref(c), lab, iters, int(until_converged), changed)
    if it < 0:
        raise ValueError(f"orc_lp_run failed: {it}")
    return lab, changed[:it].copy()


def max_threads():
    return lib().orc_max_threads()
