"""Python binding of liblvgraph.so (include/lvgraph.h): graph generators,
batch generation / tidying / application and file readers.  Graphs come
back as dicts of numpy arrays in the layout the Louvain library takes
(``off`` u64, ``keys`` u32, ``w`` f32, optional ``ex`` u8)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblvgraph.so")

EXPORTS = ["lvg_rmat", "lvg_planted_partition", "lvg_read_mtx", "lvg_write_mtx", "lvg_read_temporal",
           "lvg_from_edges", "lvg_free", "lvg_span", "lvg_edges", "lvg_order", "lvg_offsets", "lvg_keys",
           "lvg_weights", "lvg_exists", "lvg_random_batch", "lvg_tidy_batch", "lvg_batch_free",
           "lvg_batch_deletions", "lvg_batch_insertions", "lvg_batch_u", "lvg_batch_v", "lvg_batch_w",
           "lvg_apply_batch"]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python __graft_entry__.py build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    for n in ("lvg_rmat", "lvg_planted_partition", "lvg_read_mtx", "lvg_read_temporal", "lvg_from_edges",
              "lvg_random_batch", "lvg_tidy_batch", "lvg_apply_batch"):
        getattr(L, n).restype = vp
    L.lvg_rmat.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.lvg_planted_partition.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                        C.c_int]
    L.lvg_read_mtx.argtypes = [C.c_char_p]
    L.lvg_write_mtx.argtypes = [vp, C.c_char_p]
    L.lvg_read_temporal.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_int]
    L.lvg_from_edges.argtypes = [C.c_uint64, vp, vp, vp, C.c_uint64, vp]
    L.lvg_free.argtypes = [vp]
    L.lvg_free.restype = None
    for n in ("lvg_span", "lvg_edges", "lvg_order", "lvg_batch_deletions", "lvg_batch_insertions"):
        getattr(L, n).restype = C.c_uint64
        getattr(L, n).argtypes = [vp]
    for n in ("lvg_offsets", "lvg_keys", "lvg_weights", "lvg_exists"):
        getattr(L, n).restype = vp
        getattr(L, n).argtypes = [vp]
    L.lvg_random_batch.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_float, C.c_uint64]
    L.lvg_tidy_batch.argtypes = [vp, vp, vp, vp, C.c_uint64, vp, vp, vp, C.c_uint64]
    L.lvg_batch_free.argtypes = [vp]
    L.lvg_batch_free.restype = None
    for n in ("lvg_batch_u", "lvg_batch_v", "lvg_batch_w"):
        getattr(L, n).restype = vp
        getattr(L, n).argtypes = [vp, C.c_int]
    L.lvg_apply_batch.argtypes = [vp, vp]
    return L


_L = _load()


def _copy(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,)).copy()


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class _Handle:
    """Owns a native lvg_graph; converts to the numpy dict on demand."""

    def __init__(self, h):
        if not h:
            raise RuntimeError("lvgraph: generation/reading failed")
        self.h = C.c_void_p(h)

    @property
    def span(self):
        return _L.lvg_span(self.h)

    @property
    def edges(self):
        return _L.lvg_edges(self.h)

    def view(self):
        """Zero-copy numpy views of the native arrays (valid while self lives)."""
        S, E = self.span, self.edges
        def v(ptr, n, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,))
        g = {"off": v(_L.lvg_offsets(self.h), S + 1, np.uint64), "keys": v(_L.lvg_keys(self.h), E, np.uint32),
             "w": v(_L.lvg_weights(self.h), E, np.float32)}
        ex = _L.lvg_exists(self.h)
        if ex:
            g["ex"] = v(ex, S, np.uint8)
        return g

    def random_batch(self, n_del, n_ins, seed, weight=1.0):
        return random_batch(self, n_del, n_ins, seed, weight)

    def apply(self, dels, ins):
        """New native graph with the (tidy) batch applied."""
        d, i = _batch_arrays(dels, ins)
        bh = _tidy_handle(self, d, i)
        try:
            return _Handle(_L.lvg_apply_batch(self.h, C.c_void_p(bh)))
        finally:
            _L.lvg_batch_free(C.c_void_p(bh))

    def to_dict(self):
        S = _L.lvg_span(self.h)
        E = _L.lvg_edges(self.h)
        g = {"off": _copy(_L.lvg_offsets(self.h), S + 1, np.uint64),
             "keys": _copy(_L.lvg_keys(self.h), E, np.uint32),
             "w": _copy(_L.lvg_weights(self.h), E, np.float32)}
        ex = _L.lvg_exists(self.h)
        if ex:
            g["ex"] = _copy(ex, S, np.uint8)
        return g

    def __del__(self):
        if getattr(self, "h", None):
            _L.lvg_free(self.h)
            self.h = None


def _handle_of(g):
    """Native copy of a numpy-dict graph (needed by batch generation/application)."""
    S = len(g["off"]) - 1
    off = np.asarray(g["off"], np.int64)
    deg = np.diff(off)
    u = np.repeat(np.arange(S, dtype=np.uint32), deg)
    v = np.ascontiguousarray(g["keys"], np.uint32)
    w = np.ascontiguousarray(g["w"], np.float32)
    ex = None if g.get("ex") is None else np.ascontiguousarray(g["ex"], np.uint8)
    return _Handle(_L.lvg_from_edges(S, _p(u), _p(v), _p(w), len(u), _p(ex)))


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=7):
    return _Handle(_L.lvg_rmat(scale, edge_factor, a, b, c, seed)).to_dict()


def rmat_native(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=7):
    """R-MAT as a native handle (no copy); .view() gives numpy views."""
    return _Handle(_L.lvg_rmat(scale, edge_factor, a, b, c, seed))


def planted_partition_native(n, blocks, avg_degree, p_inter, seed, exponent=2.0, one_indexed=True):
    return _Handle(_L.lvg_planted_partition(n, blocks, exponent, avg_degree, p_inter, seed, int(one_indexed)))


def planted_partition(n, blocks, avg_degree, p_inter, seed, exponent=2.0, one_indexed=True):
    return _Handle(_L.lvg_planted_partition(n, blocks, exponent, avg_degree, p_inter, seed,
                                            int(one_indexed))).to_dict()


def read_mtx(path):
    return _Handle(_L.lvg_read_mtx(path.encode())).to_dict()


def write_mtx(g, path):
    h = _handle_of(g)
    if _L.lvg_write_mtx(h.h, path.encode()) != 0:
        raise OSError(f"cannot write {path}")


def read_temporal(path, rows, lines, symmetric=True):
    return _Handle(_L.lvg_read_temporal(path.encode(), rows, lines, int(symmetric))).to_dict()


def from_edges(span, u, v, w=None, ex=None):
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    w = np.ones(len(u), np.float32) if w is None else np.ascontiguousarray(w, np.float32)
    ex = None if ex is None else np.ascontiguousarray(ex, np.uint8)
    return _Handle(_L.lvg_from_edges(span, _p(u), _p(v), _p(w), len(u), _p(ex))).to_dict()


def _batch_dicts(bh):
    out = []
    n = (_L.lvg_batch_deletions(bh), _L.lvg_batch_insertions(bh))
    for k in (0, 1):
        out.append({"u": _copy(_L.lvg_batch_u(bh, k), n[k], np.uint32),
                    "v": _copy(_L.lvg_batch_v(bh, k), n[k], np.uint32),
                    "w": _copy(_L.lvg_batch_w(bh, k), n[k], np.float32)})
    return out


def random_batch(g, n_del, n_ins, seed, weight=1.0):
    """(deletions, insertions), tidy, both directions (inc/batch.hxx:100-209)."""
    h = g if isinstance(g, _Handle) else _handle_of(g)
    bh = _L.lvg_random_batch(h.h, n_del, n_ins, weight, seed)
    try:
        return tuple(_batch_dicts(C.c_void_p(bh)))
    finally:
        _L.lvg_batch_free(C.c_void_p(bh))


def _batch_arrays(dels, ins):
    d = [np.ascontiguousarray(dels["u"], np.uint32), np.ascontiguousarray(dels["v"], np.uint32),
         np.ascontiguousarray(dels["w"], np.float32)]
    i = [np.ascontiguousarray(ins["u"], np.uint32), np.ascontiguousarray(ins["v"], np.uint32),
         np.ascontiguousarray(ins["w"], np.float32)]
    return d, i


def _tidy_handle(h, d, i):
    return _L.lvg_tidy_batch(h.h, _p(d[0]), _p(d[1]), _p(d[2]), len(d[0]), _p(i[0]), _p(i[1]), _p(i[2]), len(i[0]))


def tidy_batch(g, dels, ins):
    """filterEdgesByExistenceU + sortEdgesByIdU + uniqueEdgesU (inc/batch.hxx:154-209)."""
    h = g if isinstance(g, _Handle) else _handle_of(g)
    d, i = _batch_arrays(dels, ins)
    bh = _tidy_handle(h, d, i)
    try:
        return tuple(_batch_dicts(C.c_void_p(bh)))
    finally:
        _L.lvg_batch_free(C.c_void_p(bh))


def apply_batch(g, dels, ins):
    """Graph after removing deletions then adding insertions (inc/batch.hxx:219-226)."""
    h = _handle_of(g)
    d, i = _batch_arrays(dels, ins)
    bh = _tidy_handle(h, d, i)
    try:
        return _Handle(_L.lvg_apply_batch(h.h, C.c_void_p(bh))).to_dict()
    finally:
        _L.lvg_batch_free(C.c_void_p(bh))
