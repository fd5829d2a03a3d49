"""ctypes binding of liblpa_host.so (include/lpa_host.h).

Host-side graph preparation: synthetic generators, CSR building, BFS
relabelling, the CGR encoder and the reference's on-disk file set.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_VP = C.c_void_p


class LpaError(RuntimeError):
    pass


def _load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liblpa_host.so")
    if not os.path.exists(path):
        raise LpaError(f"{path} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    L.lpah_last_error.restype = C.c_char_p
    sigs = {
        "lpah_csr_from_edges": (C.c_int, [C.c_uint64, _VP, _VP, C.c_uint64, C.c_int, C.POINTER(_VP)]),
        "lpah_csr_from_arrays": (C.c_int, [C.c_uint64, _VP, _VP, C.POINTER(_VP)]),
        "lpah_csr_read_edge_list": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_VP)]),
        "lpah_csr_n": (C.c_uint64, [_VP]),
        "lpah_csr_m": (C.c_uint64, [_VP]),
        "lpah_csr_row": (_VP, [_VP]),
        "lpah_csr_col": (_VP, [_VP]),
        "lpah_csr_max_degree": (C.c_uint32, [_VP]),
        "lpah_csr_free": (None, [_VP]),
        "lpah_csr_bfs_relabel": (C.c_int, [_VP, C.c_uint32, _VP, C.POINTER(_VP)]),
        "lpah_gen_lattice": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.POINTER(_VP)]),
        "lpah_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                    C.c_uint64, C.c_int, C.POINTER(_VP)]),
        "lpah_gen_copying": (C.c_int, [C.c_uint64, C.c_double, C.c_double, C.c_uint32, C.c_double,
                                       C.c_uint32, C.c_uint64, C.POINTER(_VP)]),
        "lpah_cgr_encode": (C.c_int, [_VP, C.c_uint32, C.c_uint32, C.POINTER(_VP)]),
        "lpah_cgr_n": (C.c_uint64, [_VP]),
        "lpah_cgr_nc": (C.c_uint64, [_VP]),
        "lpah_cgr_nbits": (C.c_uint64, [_VP]),
        "lpah_cgr_words": (_VP, [_VP]),
        "lpah_cgr_nwords": (C.c_uint64, [_VP]),
        "lpah_cgr_offsets": (_VP, [_VP]),
        "lpah_cgr_vertices": (_VP, [_VP]),
        "lpah_cgr_small_offsets": (_VP, [_VP]),
        "lpah_cgr_small_nbrs": (_VP, [_VP]),
        "lpah_cgr_small_m": (C.c_uint64, [_VP]),
        "lpah_cgr_free": (None, [_VP]),
        "lpah_files_write": (C.c_int, [_VP, _VP, C.c_char_p]),
        "lpah_files_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_VP)]),
        "lpah_threads": (C.c_int, []),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise LpaError(f"lpa host error {rc}: {_load().lpah_last_error().decode()}")


def _view(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    ct = {np.uint32: C.c_uint32, np.uint64: C.c_uint64}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,))


class Csr:
    """Host CSR graph (rows sorted, no duplicates, no self loops)."""

    def __init__(self, handle):
        self._h = _VP(handle)
        L = _load()
        self.n = L.lpah_csr_n(self._h)
        self.m = L.lpah_csr_m(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.lpah_csr_free(self._h)
            self._h = None

    @property
    def row(self):
        return _view(_load().lpah_csr_row(self._h), self.n + 1, np.uint64)

    @property
    def col(self):
        return _view(_load().lpah_csr_col(self._h), self.m, np.uint32)

    def degrees(self):
        return np.diff(self.row.astype(np.int64))

    def max_degree(self):
        return _load().lpah_csr_max_degree(self._h)

    @staticmethod
    def from_edges(n, src, dst, undirected):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = _VP()
        _check(_load().lpah_csr_from_edges(n, src.ctypes.data, dst.ctypes.data, len(src),
                                           int(undirected), C.byref(h)))
        return Csr(h.value)

    @staticmethod
    def from_arrays(row, col):
        row = np.ascontiguousarray(row, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        h = _VP()
        _check(_load().lpah_csr_from_arrays(len(row) - 1, row.ctypes.data, col.ctypes.data, C.byref(h)))
        return Csr(h.value)

    @staticmethod
    def read_edge_list(path, undirected):
        h = _VP()
        _check(_load().lpah_csr_read_edge_list(path.encode(), int(undirected), C.byref(h)))
        return Csr(h.value)

    def bfs_relabel(self, root=0):
        perm = np.zeros(self.n, np.uint32)
        h = _VP()
        _check(_load().lpah_csr_bfs_relabel(self._h, root, perm.ctypes.data, C.byref(h)))
        return Csr(h.value), perm

    def encode(self, small_thresh=0, large_thresh=128):
        h = _VP()
        _check(_load().lpah_cgr_encode(self._h, small_thresh, large_thresh, C.byref(h)))
        return Cgr(h.value)


def gen_lattice(side=1024, p_drop=0.10, shortcut_frac=0.005, seed=0):
    h = _VP()
    _check(_load().lpah_gen_lattice(side, p_drop, shortcut_frac, seed, C.byref(h)))
    return Csr(h.value)


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, bfs=True):
    h = _VP()
    _check(_load().lpah_gen_rmat(scale, edge_factor, a, b, c, seed, int(bfs), C.byref(h)))
    return Csr(h.value)


def gen_copying(n, mean_deg=30.0, exponent=2.1, cap=100000, copy_frac=0.7, window=1000, seed=2):
    h = _VP()
    _check(_load().lpah_gen_copying(n, mean_deg, exponent, cap, copy_frac, window, seed, C.byref(h)))
    return Csr(h.value)


class Cgr:
    """CGR file set in host memory (stream, start offsets, vertex order, small CSR)."""

    def __init__(self, handle):
        self._h = _VP(handle)
        L = _load()
        self.n = L.lpah_cgr_n(self._h)
        self.n_c = L.lpah_cgr_nc(self._h)
        self.nbits = L.lpah_cgr_nbits(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.lpah_cgr_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def words(self):
        """big-endian uint32 words (File::graph) followed by 4 zero words"""
        L = _load()
        return _view(L.lpah_cgr_words(self._h), L.lpah_cgr_nwords(self._h) + 4, np.uint32)

    @property
    def nwords(self):
        return _load().lpah_cgr_nwords(self._h)

    @property
    def offsets(self):
        return _view(_load().lpah_cgr_offsets(self._h), self.n_c + 1, np.uint64)

    @property
    def vertices(self):
        return _view(_load().lpah_cgr_vertices(self._h), self.n, np.uint32)

    @property
    def small_offsets(self):
        return _view(_load().lpah_cgr_small_offsets(self._h), self.n - self.n_c + 1, np.uint64)

    @property
    def small_nbrs(self):
        L = _load()
        return _view(L.lpah_cgr_small_nbrs(self._h), L.lpah_cgr_small_m(self._h), np.uint32)

    def stream_bytes(self):
        """the .graph bytes (MSB first)"""
        w = self.words[:self.nwords].astype(">u4").tobytes()
        return w[:(self.nbits + 7) // 8]

    def write(self, path, csr=None):
        _check(_load().lpah_files_write(self._h, csr._h if csr is not None else None, path.encode()))

    @staticmethod
    def load(path, small_thresh=0):
        h = _VP()
        _check(_load().lpah_files_load(path.encode(), small_thresh, C.byref(h)))
        return Cgr(h.value)


def threads():
    return _load().lpah_threads()
