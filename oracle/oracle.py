"""ctypes bindings to the CPU checkers.  TEST INFRASTRUCTURE, NOT THE PRODUCT.

* ``Oracle``  -> ``oracle/liboracle.so``, the C + OpenMP restatement of the
  reference's Louvain (``louvain_oracle.c``; every function there cites the
  ``inc/louvain.hxx`` / ``inc/properties.hxx`` lines it follows).
* ``RefLib``  -> ``oracle/_ref/libref.so``, the reference's own header-only
  code compiled in place by ``oracle/Makefile`` (only where the reference
  sources exist; never on the GPU box).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may use this module, and only as the checker / CPU baseline.

A graph is a dict of numpy arrays: ``off`` (u64, S+1), ``keys`` (u32),
``w`` (f32, or f64 for a super-graph), optional ``deg`` (u32, gapped CSR) and
``ex`` (u8 exists mask).  A batch is a dict with ``u``, ``v`` (u32), ``w`` (f32).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

STATIC, NAIVE_DYNAMIC, DELTA_SCREENING, DYNAMIC_FRONTIER = 0, 1, 2, 3
APPROACHES = {"static": STATIC, "naive-dynamic": NAIVE_DYNAMIC,
              "delta-screening": DELTA_SCREENING, "dynamic-frontier": DYNAMIC_FRONTIER}


class Options(C.Structure):
    """LouvainOptions (inc/louvain.hxx:30-63), defaults from :60."""
    _fields_ = [("repeat", C.c_int32), ("resolution", C.c_double), ("tolerance", C.c_double),
                ("aggregation_tolerance", C.c_double), ("tolerance_drop", C.c_double),
                ("max_iterations", C.c_int32), ("max_passes", C.c_int32)]

    def __init__(self, repeat=1, resolution=1.0, tolerance=1e-2, aggregation_tolerance=1.0,
                 tolerance_drop=10.0, max_iterations=20, max_passes=10):
        super().__init__(repeat, resolution, tolerance, aggregation_tolerance, tolerance_drop,
                         max_iterations, max_passes)


class Out(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("passes", C.c_int32), ("time", C.c_float),
                ("marking_time", C.c_float), ("initialization_time", C.c_float),
                ("first_pass_time", C.c_float), ("local_move_time", C.c_float),
                ("aggregation_time", C.c_float), ("affected_vertices", C.c_uint64)]


class _Graph(C.Structure):
    _fields_ = [("S", C.c_uint64), ("off", C.c_void_p), ("deg", C.c_void_p), ("keys", C.c_void_p),
                ("w32", C.c_void_p), ("w64", C.c_void_p), ("ex", C.c_void_p)]


class _Batch(C.Structure):
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("n", C.c_uint64)]


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def norm_graph(g):
    """Contiguous, correctly typed copy of a graph dict."""
    w = g["w"]
    out = {"off": _c(g["off"], np.uint64), "keys": _c(g["keys"], np.uint32),
           "w": _c(w, np.float64 if w.dtype == np.float64 else np.float32),
           "deg": _c(g.get("deg"), np.uint32), "ex": _c(g.get("ex"), np.uint8)}
    out["S"] = len(out["off"]) - 1
    return out


def norm_batch(b):
    if b is None:
        b = {"u": np.zeros(0, np.uint32), "v": np.zeros(0, np.uint32), "w": np.zeros(0, np.float32)}
    return {"u": _c(b["u"], np.uint32), "v": _c(b["v"], np.uint32), "w": _c(b["w"], np.float32)}


class Result:
    def __init__(self, membership, vertex_weight, community_weight, out):
        self.membership = membership
        self.vertex_weight = vertex_weight
        self.community_weight = community_weight
        for f, _ in Out._fields_:
            setattr(self, f, getattr(out, f))


class Oracle:
    """The C restatement (oracle/louvain_oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(path)
        L = self.lib
        L.oracle_edge_weight.restype = C.c_double
        L.oracle_modularity.restype = C.c_double
        L.oracle_renumber.restype = C.c_uint64
        L.oracle_max_threads.restype = C.c_int

    def set_threads(self, t):
        self.lib.oracle_set_threads(int(t))

    def max_threads(self):
        return self.lib.oracle_max_threads()

    @staticmethod
    def _g(g):
        g = norm_graph(g)
        w64 = g["w"].dtype == np.float64
        s = _Graph(g["S"], _p(g["off"]), _p(g["deg"]), _p(g["keys"]),
                   None if w64 else _p(g["w"]), _p(g["w"]) if w64 else None, _p(g["ex"]))
        return s, g

    @staticmethod
    def _b(b):
        b = norm_batch(b)
        return _Batch(_p(b["u"]), _p(b["v"]), _p(b["w"]), len(b["u"])), b

    def edge_weight(self, g):
        s, keep = self._g(g)
        return self.lib.oracle_edge_weight(C.byref(s))

    def modularity(self, g, memb, R=1.0):
        s, keep = self._g(g)
        m = _c(memb, np.uint32)
        M = self.lib.oracle_edge_weight(C.byref(s)) / 2
        return self.lib.oracle_modularity(C.byref(s), _p(m), C.c_double(M), C.c_double(R))

    def louvain(self, approach, g, dels=None, ins=None, q=None, qvtot=None, qctot=None, options=None):
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        s, keep = self._g(g)
        bd, kd = self._b(dels)
        bi, ki = self._b(ins)
        S = keep["S"]
        q = _c(q if q is not None else np.zeros(S), np.uint32)
        qv = _c(qvtot if qvtot is not None else np.zeros(S), np.float64)
        qc = _c(qctot if qctot is not None else np.zeros(S), np.float64)
        o = options or Options()
        memb = np.zeros(S, np.uint32)
        vt = np.zeros(S, np.float64)
        ct = np.zeros(S, np.float64)
        out = Out()
        rc = self.lib.oracle_louvain(approach, C.byref(s), C.byref(bd), C.byref(bi), _p(q), _p(qv), _p(qc),
                                     C.byref(o), _p(memb), _p(vt), _p(ct), C.byref(out))
        if rc != 0:
            raise RuntimeError(f"oracle_louvain failed: {rc}")
        return Result(memb, vt, ct, out)

    def update_weights(self, dels, ins, vcom, vtot, ctot):
        bd, kd = self._b(dels)
        bi, ki = self._b(ins)
        vt = np.array(vtot, np.float64)
        ct = np.array(ctot, np.float64)
        vc = _c(vcom, np.uint32)
        self.lib.oracle_update_weights(_p(vt), _p(ct), C.byref(bd), C.byref(bi), _p(vc))
        return vt, ct

    def frontier(self, S, dels, ins, vcom):
        bd, kd = self._b(dels)
        bi, ki = self._b(ins)
        vc = _c(vcom, np.uint32)
        a = np.zeros(S, np.uint8)
        self.lib.oracle_frontier(_p(a), C.c_uint64(S), C.byref(bd), C.byref(bi), _p(vc))
        return a

    def delta_screening(self, g, dels, ins, vcom, vtot, ctot, M, R=1.0):
        s, keep = self._g(g)
        bd, kd = self._b(dels)
        bi, ki = self._b(ins)
        S = keep["S"]
        a, b, c = (np.zeros(S, np.uint8) for _ in range(3))
        vc, vt, ct = _c(vcom, np.uint32), _c(vtot, np.float64), _c(ctot, np.float64)
        self.lib.oracle_delta_screening(_p(a), _p(b), _p(c), C.byref(s), C.byref(bd), C.byref(bi),
                                        _p(vc), _p(vt), _p(ct), C.c_double(M), C.c_double(R))
        return a, b, c

    def choose(self, g, vcom, vtot, ctot, M, R=1.0):
        s, keep = self._g(g)
        S = keep["S"]
        bc = np.zeros(S, np.uint32)
        be = np.zeros(S, np.float64)
        vc, vt, ct = _c(vcom, np.uint32), _c(vtot, np.float64), _c(ctot, np.float64)
        self.lib.oracle_choose(_p(bc), _p(be), C.byref(s), _p(vc), _p(vt), _p(ct), C.c_double(M), C.c_double(R))
        return bc, be

    def renumber(self, g, vcom):
        s, keep = self._g(g)
        S = keep["S"]
        vc = np.array(vcom, np.uint32)
        cext = np.zeros(S, np.uint32)
        CN = self.lib.oracle_renumber(_p(vc), C.c_uint64(len(vc)), _p(cext), C.byref(s))
        return vc, cext, CN

    def aggregate(self, g, vcom, CN):
        s, keep = self._g(g)
        S = keep["S"]
        E = int(keep["off"][-1])
        zoff = np.zeros(CN + 1, np.uint64)
        zdeg = np.zeros(max(CN, 1), np.uint32)
        zkeys = np.zeros(max(E, 1), np.uint32)
        zw = np.zeros(max(E, 1), np.float64)
        coff = np.zeros(CN + 1, np.uint32)
        cedg = np.zeros(max(S, 1), np.uint32)
        vc = _c(vcom, np.uint32)
        self.lib.oracle_aggregate(_p(zoff), _p(zdeg), _p(zkeys), _p(zw), _p(coff), _p(cedg), C.byref(s), _p(vc),
                                  C.c_uint64(CN))
        return {"off": zoff, "deg": zdeg[:CN], "keys": zkeys[:E], "w": zw[:E], "coff": coff, "cedg": cedg[:S]}


class RefLib:
    """The reference compiled in place (oracle/_ref/libref.so)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where /root/reference exists")
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_edge_weight.restype = C.c_double
        L.ref_modularity.restype = C.c_double
        L.ref_renumber.restype = C.c_uint64

    def set_threads(self, t):
        self.lib.ref_set_threads(int(t))

    @staticmethod
    def _gargs(g):
        g = norm_graph(g)
        return g, [C.c_uint64(g["S"]), _p(g["off"]), _p(g["deg"]), _p(g["keys"])]

    def edge_weight(self, g):
        g, a = self._gargs(g)
        return self.lib.ref_edge_weight(*a, _p(g["w"]), _p(g["ex"]))

    def modularity(self, g, memb, R=1.0):
        g, a = self._gargs(g)
        m = _c(memb, np.uint32)
        return self.lib.ref_modularity(*a, _p(g["w"]), _p(g["ex"]), _p(m), C.c_double(R))

    @staticmethod
    def _bargs(b):
        b = norm_batch(b)
        return b, [_p(b["u"]), _p(b["v"]), _p(b["w"]), C.c_uint64(len(b["u"]))]

    def louvain(self, approach, g, dels=None, ins=None, q=None, qvtot=None, qctot=None, options=None):
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        g, a = self._gargs(g)
        S = g["S"]
        bd, da = self._bargs(dels)
        bi, ia = self._bargs(ins)
        q = _c(q if q is not None else np.zeros(S), np.uint32)
        qv = _c(qvtot if qvtot is not None else np.zeros(S), np.float64)
        qc = _c(qctot if qctot is not None else np.zeros(S), np.float64)
        o = options or Options()
        memb = np.zeros(S, np.uint32)
        vt = np.zeros(S, np.float64)
        ct = np.zeros(S, np.float64)
        out = Out()
        rc = self.lib.ref_louvain(approach, *a, _p(g["w"]), _p(g["ex"]), *da, *ia, _p(q), _p(qv), _p(qc),
                                  C.byref(o), _p(memb), _p(vt), _p(ct), C.byref(out))
        if rc != 0:
            raise RuntimeError(f"ref_louvain failed: {rc}")
        return Result(memb, vt, ct, out)

    def tidy_update(self, g, dels, ins):
        """tidyBatchUpdateU then applyBatchUpdateOmpU (inc/batch.hxx:202-248):
        the tidy (deletions, insertions) and the updated graph as a packed CSR
        dict with its exists mask."""
        g, a = self._gargs(g)
        bd, da = self._bargs(dels)
        bi, ia = self._bargs(ins)
        nd, ni = len(bd["u"]), len(bi["u"])
        td = {"u": np.zeros(nd, np.uint32), "v": np.zeros(nd, np.uint32), "w": np.zeros(nd, np.float32)}
        ti = {"u": np.zeros(ni, np.uint32), "v": np.zeros(ni, np.uint32), "w": np.zeros(ni, np.float32)}
        tnd, tni = C.c_uint64(0), C.c_uint64(0)
        S = g["S"]
        cap = int(g["off"][-1]) + ni
        yoff = np.zeros(S + 1, np.uint64)
        yk = np.zeros(max(cap, 1), np.uint32)
        yw = np.zeros(max(cap, 1), np.float32)
        yex = np.zeros(S, np.uint8)
        self.lib.ref_tidy_update.restype = C.c_int64
        e = self.lib.ref_tidy_update(*a, _p(g["w"]), _p(g["ex"]), *da, *ia,
                                     _p(td["u"]), _p(td["v"]), _p(td["w"]), C.byref(tnd),
                                     _p(ti["u"]), _p(ti["v"]), _p(ti["w"]), C.byref(tni),
                                     _p(yoff), _p(yk), _p(yw), _p(yex), C.c_uint64(cap))
        if e < 0:
            raise RuntimeError("ref_tidy_update: edge capacity too small")
        for d, n in ((td, tnd.value), (ti, tni.value)):
            for k in d:
                d[k] = d[k][:n].copy()
        return td, ti, {"off": yoff, "keys": yk[:e].copy(), "w": yw[:e].copy(), "ex": yex}

    def update_weights(self, g, dels, ins, vcom, vtot, ctot):
        g, a = self._gargs(g)
        bd, da = self._bargs(dels)
        bi, ia = self._bargs(ins)
        vt = np.array(vtot, np.float64)
        ct = np.array(ctot, np.float64)
        vc = _c(vcom, np.uint32)
        self.lib.ref_update_weights(*a, _p(g["w"]), _p(g["ex"]), *da, *ia, _p(vc), _p(vt), _p(ct))
        return vt, ct

    def frontier(self, g, dels, ins, vcom):
        g, a = self._gargs(g)
        bd, da = self._bargs(dels)
        bi, ia = self._bargs(ins)
        vc = _c(vcom, np.uint32)
        out = np.zeros(g["S"], np.uint8)
        self.lib.ref_frontier(*a, _p(g["w"]), _p(g["ex"]), *da, *ia, _p(vc), _p(out))
        return out

    def delta_screening(self, g, dels, ins, vcom, vtot, ctot, M, R=1.0):
        g, a = self._gargs(g)
        bd, da = self._bargs(dels)
        bi, ia = self._bargs(ins)
        S = g["S"]
        x, y, z = (np.zeros(S, np.uint8) for _ in range(3))
        vc, vt, ct = _c(vcom, np.uint32), _c(vtot, np.float64), _c(ctot, np.float64)
        self.lib.ref_delta_screening(*a, _p(g["w"]), _p(g["ex"]), *da, *ia, _p(vc), _p(vt), _p(ct),
                                     C.c_double(M), C.c_double(R), _p(x), _p(y), _p(z))
        return x, y, z

    def choose(self, g, vcom, vtot, ctot, M, R=1.0):
        g, a = self._gargs(g)
        S = g["S"]
        w64 = g["w"].dtype == np.float64
        bc = np.zeros(S, np.uint32)
        be = np.zeros(S, np.float64)
        vc, vt, ct = _c(vcom, np.uint32), _c(vtot, np.float64), _c(ctot, np.float64)
        self.lib.ref_choose(*a, _p(g["w"]), C.c_int(int(w64)), _p(g["ex"]), _p(vc), _p(vt), _p(ct),
                            C.c_double(M), C.c_double(R), _p(bc), _p(be))
        return bc, be

    def renumber(self, g, vcom):
        g, a = self._gargs(g)
        S = g["S"]
        w64 = g["w"].dtype == np.float64
        vc = np.array(vcom, np.uint32)
        cext = np.zeros(S, np.uint32)
        CN = self.lib.ref_renumber(*a, _p(g["w"]), C.c_int(int(w64)), _p(g["ex"]), _p(vc), _p(cext))
        return vc, cext, CN

    def aggregate(self, g, vcom, CN):
        g, a = self._gargs(g)
        S = g["S"]
        E = int(g["off"][-1])
        w64 = g["w"].dtype == np.float64
        zoff = np.zeros(CN + 1, np.uint64)
        zdeg = np.zeros(max(CN, 1), np.uint32)
        zkeys = np.zeros(max(E, 1), np.uint32)
        zw = np.zeros(max(E, 1), np.float64)
        coff = np.zeros(CN + 1, np.uint32)
        cedg = np.zeros(max(S, 1), np.uint32)
        vc = _c(vcom, np.uint32)
        self.lib.ref_aggregate(*a, _p(g["w"]), C.c_int(int(w64)), _p(g["ex"]), _p(vc), C.c_uint64(CN),
                               _p(zoff), _p(zdeg), _p(zkeys), _p(zw), _p(coff), _p(cedg))
        return {"off": zoff, "deg": zdeg[:CN], "keys": zkeys[:E], "w": zw[:E], "coff": coff, "cedg": cedg[:S]}
