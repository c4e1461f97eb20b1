"""Python binding of liblouvain_hip.so (the MI355X Louvain library).

Mirrors the reference's interface (puzzlef/louvain-communities-openmp-dynamic):

    LouvainOptions                      inc/louvain.hxx:30-63
    LouvainResult                       inc/louvain.hxx:77-145
    louvainStaticHip(x, o)              <- louvainStaticOmp                inc/louvain.hxx:1285
    louvainNaiveDynamicHip(y, d, i, q, qvtot, qctot, o)                    inc/louvain.hxx:1349
    louvainDynamicDeltaScreeningHip(...)                                   inc/louvain.hxx:1542
    louvainDynamicFrontierHip(...)                                         inc/louvain.hxx:1674

A graph is a dict of numpy arrays (``off`` u64[S+1], ``keys`` u32, ``w`` f32
or f64, optional ``deg`` u32 for a gapped CSR, optional ``ex`` u8 exists
mask); rows sorted by neighbour id.  A batch is a dict of ``u``, ``v`` (u32)
and ``w`` (f32) arrays, tidy (sorted by (u, v), unique, both directions).

There is no CPU fallback: importing this module on a machine without the
built library raises, and every call without a HIP device raises.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblouvain_hip.so")

LV_STATIC, LV_NAIVE_DYNAMIC, LV_DELTA_SCREENING, LV_DYNAMIC_FRONTIER = 0, 1, 2, 3
APPROACHES = {"static": LV_STATIC, "naive-dynamic": LV_NAIVE_DYNAMIC,
              "delta-screening": LV_DELTA_SCREENING, "dynamic-frontier": LV_DYNAMIC_FRONTIER}
PATH_SWEEP, PATH_FRONTIER, PATH_WORKLIST = 0, 1, 2
LV_OK, LV_ERR_INVALID, LV_ERR_NOMEM, LV_ERR_HIP, LV_ERR_UNSORTED, LV_ERR_NODEVICE, LV_ERR_LIMIT, LV_ERR_PENDING = (
    0, -1, -2, -3, -4, -5, -6, -7)

# Every symbol include/louvain_hip.h declares.
EXPORTS = [
    "lv_options_default", "lv_status_string", "lv_context_create", "lv_context_destroy", "lv_last_error",
    "lv_context_set_path", "lv_context_set_schedule", "lv_context_set_worklist_schedule", "lv_context_set_pick_less", "lv_context_set_pass_schedule", "lv_context_set_validate", "lv_context_set_sweep_hubs", "lv_context_set_lead_ids", "lv_context_set_workspace_limit", "lv_context_workspace_bytes", "lv_context_set_slabs", "lv_context_set_hub_streams", "lv_context_set_hub_threshold",
    "lv_context_set_profiling", "lv_context_set_graphs", "lv_context_set_hub_tables",
    "lv_dstate_check_hub_tables", "lv_stats_get",
    "lv_stats_reset", "lv_louvain_static", "lv_louvain_naive_dynamic", "lv_louvain_delta_screening",
    "lv_louvain_dynamic_frontier", "lv_dgraph_create", "lv_dgraph_destroy", "lv_dgraph_span", "lv_dgraph_edges",
    "lv_dbatch_create", "lv_dbatch_destroy", "lv_dstate_create", "lv_dstate_destroy", "lv_dstate_upload",
    "lv_dstate_download", "lv_dstate_copy", "lv_run", "lv_dmodularity", "lv_edge_weight", "lv_modularity",
    "lv_dgraph_capacity", "lv_dgraph_update", "lv_dgraph_download", "lv_dbatch_sizes", "lv_dbatch_download",
    "lv_run_batched", "lv_dmodularity_noalloc",
    "lv_vertex_weights", "lv_update_weights", "lv_affected_frontier", "lv_affected_delta_screening",
    "lv_choose_community", "lv_renumber", "lv_aggregate",
    "lv_dgraph_device_bytes", "lv_dstate_device_bytes",
    "lv_server_options_default", "lv_server_create", "lv_server_destroy", "lv_server_add_graph",
    "lv_server_submit", "lv_server_poll", "lv_server_drain", "lv_server_graph_state", "lv_server_get_stats",
    "lv_server_last_error", "lv_server_now_ms", "lv_server_graph_tuning",
    "lv_autotune", "lv_context_set_autotune", "lv_context_add_tuning", "lv_context_apply_tuning", "lv_context_tuning",
    "lv_dstate_clone",
]


class LvError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class _Options(C.Structure):
    _fields_ = [("repeat", C.c_int32), ("resolution", C.c_double), ("tolerance", C.c_double),
                ("aggregation_tolerance", C.c_double), ("tolerance_drop", C.c_double),
                ("max_iterations", C.c_int32), ("max_passes", C.c_int32)]


class _Graph(C.Structure):
    _fields_ = [("span", C.c_uint64), ("offsets", C.c_void_p), ("degrees", C.c_void_p), ("keys", C.c_void_p),
                ("weights", C.c_void_p), ("weights_f64", C.c_void_p), ("exists", C.c_void_p)]


class _Batch(C.Structure):
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("size", C.c_uint64)]


class _Stats(C.Structure):
    _fields_ = [("move_ms", C.c_double * 5), ("move_launches", C.c_uint64 * 5), ("vertices", C.c_uint64 * 5),
                ("edges", C.c_uint64 * 5), ("candidates", C.c_uint64 * 5), ("moves", C.c_uint64 * 5),
                ("moved_edges", C.c_uint64 * 5), ("runs", C.c_uint64), ("fused_ms", C.c_double),
                ("fused_launches", C.c_uint64)]


class _Result(C.Structure):
    _fields_ = [("membership", C.c_void_p), ("vertex_weight", C.c_void_p), ("community_weight", C.c_void_p),
                ("iterations", C.c_int32), ("passes", C.c_int32), ("time", C.c_float),
                ("marking_time", C.c_float), ("initialization_time", C.c_float),
                ("first_pass_time", C.c_float), ("local_move_time", C.c_float),
                ("aggregation_time", C.c_float), ("affected_vertices", C.c_uint64)]


class _ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("memory_budget", C.c_uint64), ("options", _Options),
                ("path", C.c_int32), ("modularity", C.c_int32), ("keep_states", C.c_int32), ("autotune", C.c_int32),
                ("probe_fraction", C.c_double)]


class _TuneChoice(C.Structure):
    _fields_ = [("worklist_hub_degree", C.c_uint32), ("sweep_hub_degree", C.c_uint32),
                ("worklist_streams", C.c_uint64), ("interleave", C.c_int32), ("stream_fraction", C.c_double),
                ("ms", C.c_double), ("default_ms", C.c_double), ("modularity", C.c_double),
                ("default_modularity", C.c_double), ("candidates", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

    @classmethod
    def of(cls, d):
        return cls(*[d[f] if d and f in d else 0 for f, _ in cls._fields_])


class _Completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("graph_id", C.c_uint32), ("status", C.c_int32),
                ("submitted_ms", C.c_double), ("started_ms", C.c_double), ("done_ms", C.c_double),
                ("result", _Result), ("modularity", C.c_double), ("deletions", C.c_uint64),
                ("insertions", C.c_uint64)]


class _ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "resident_bytes", "workspace_bytes",
                                          "inflight_bytes", "peak_bytes")]


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python __graft_entry__.py build` (hipcc, gfx950)")
    lib = C.CDLL(path)
    lib.lv_status_string.restype = C.c_char_p
    lib.lv_last_error.restype = C.c_char_p
    lib.lv_dgraph_span.restype = C.c_uint64
    lib.lv_dgraph_edges.restype = C.c_uint64
    lib.lv_dgraph_capacity.restype = C.c_uint64
    lib.lv_context_workspace_bytes.restype = C.c_uint64
    lib.lv_dgraph_device_bytes.restype = C.c_uint64
    lib.lv_dstate_device_bytes.restype = C.c_uint64
    lib.lv_server_last_error.restype = C.c_char_p
    lib.lv_server_now_ms.restype = C.c_double
    for name in ("lv_context_destroy", "lv_dgraph_destroy", "lv_dbatch_destroy", "lv_dstate_destroy",
                 "lv_server_destroy", "lv_server_options_default"):
        getattr(lib, name).restype = None
    return lib


_LIB = load_library()


@dataclass
class LouvainOptions:
    """LouvainOptions (inc/louvain.hxx:30-63); defaults from inc/louvain.hxx:60."""
    repeat: int = 1
    resolution: float = 1.0
    tolerance: float = 1e-2
    aggregationTolerance: float = 1.0
    toleranceDrop: float = 10.0
    maxIterations: int = 20
    maxPasses: int = 10

    def _c(self):
        return _Options(self.repeat, self.resolution, self.tolerance, self.aggregationTolerance,
                        self.toleranceDrop, self.maxIterations, self.maxPasses)


@dataclass
class LouvainResult:
    """LouvainResult (inc/louvain.hxx:77-145)."""
    membership: np.ndarray = None
    vertexWeight: np.ndarray = None
    communityWeight: np.ndarray = None
    iterations: int = 0
    passes: int = 0
    time: float = 0.0
    markingTime: float = 0.0
    initializationTime: float = 0.0
    firstPassTime: float = 0.0
    localMoveTime: float = 0.0
    aggregationTime: float = 0.0
    affectedVertices: int = 0
    counters: dict = field(default_factory=dict)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _arr(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


class _GraphArgs:
    """Keeps the typed numpy arrays alive while the C struct points at them."""

    def __init__(self, g):
        w = np.asarray(g["w"])
        self.off = _arr(g["off"], np.uint64)
        self.keys = _arr(g["keys"], np.uint32)
        self.deg = _arr(g.get("deg"), np.uint32)
        self.ex = _arr(g.get("ex"), np.uint8)
        self.f64 = w.dtype == np.float64
        self.w = _arr(w, np.float64 if self.f64 else np.float32)
        self.S = len(self.off) - 1
        self.c = _Graph(self.S, _p(self.off), _p(self.deg), _p(self.keys),
                        None if self.f64 else _p(self.w), _p(self.w) if self.f64 else None, _p(self.ex))


class _BatchArgs:
    def __init__(self, b):
        if b is None:
            b = {"u": np.zeros(0, np.uint32), "v": np.zeros(0, np.uint32), "w": np.zeros(0, np.float32)}
        self.u = _arr(b["u"], np.uint32)
        self.v = _arr(b["v"], np.uint32)
        self.w = _arr(b["w"], np.float32)
        self.c = _Batch(_p(self.u), _p(self.v), _p(self.w), len(self.u))


class Context:
    """One HIP device + stream + grow-only device workspace (lv_context)."""

    def __init__(self, device=0, path=PATH_WORKLIST):
        self.lib = _LIB
        self.h = C.c_void_p()
        rc = self.lib.lv_context_create(int(device), C.byref(self.h))
        if rc != 0:
            raise LvError(rc, "lv_context_create: " + self.lib.lv_status_string(rc).decode())
        self.set_path(path)

    def close(self):
        if self.h:
            self.lib.lv_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_path(self, path):
        self._check(self.lib.lv_context_set_path(self.h, int(path)), "lv_context_set_path")

    def set_hub_streams(self, streams=256):
        self._check(self.lib.lv_context_set_hub_streams(self.h, C.c_uint64(streams)), "lv_context_set_hub_streams")

    def set_hub_threshold(self, sweep_degree=4096, worklist_degree=256):
        self._check(self.lib.lv_context_set_hub_threshold(self.h, C.c_uint32(sweep_degree),
                                                          C.c_uint32(worklist_degree)),
                    "lv_context_set_hub_threshold")

    def set_slabs(self, slabs=1):
        self._check(self.lib.lv_context_set_slabs(self.h, C.c_uint64(slabs)), "lv_context_set_slabs")

    def set_pick_less(self, iterations=0):
        self._check(self.lib.lv_context_set_pick_less(self.h, int(iterations)), "lv_context_set_pick_less")

    def set_lead_ids(self, ids=0):
        self._check(self.lib.lv_context_set_lead_ids(self.h, C.c_uint64(ids)), "lv_context_set_lead_ids")

    def set_sweep_hubs(self, in_sweep=False):
        self._check(self.lib.lv_context_set_sweep_hubs(self.h, int(bool(in_sweep))), "lv_context_set_sweep_hubs")

    def set_validate(self, mode=0):
        """Validated moves: 0 off (default), 1 sweeps, 2 sweeps and worklists."""
        self._check(self.lib.lv_context_set_validate(self.h, int(mode)), "lv_context_set_validate")

    def set_workspace_limit(self, nbytes=0):
        self._check(self.lib.lv_context_set_workspace_limit(self.h, C.c_uint64(nbytes)),
                    "lv_context_set_workspace_limit")

    def workspace_bytes(self):
        return int(self.lib.lv_context_workspace_bytes(self.h))

    def set_pass_schedule(self, lead_iterations=1, lead_streams=16, lead_chunk=2048, late_streams=256, late_chunk=2048):
        self._check(self.lib.lv_context_set_pass_schedule(self.h, int(lead_iterations), C.c_uint64(lead_streams),
                                                          C.c_uint64(lead_chunk), C.c_uint64(late_streams),
                                                          C.c_uint64(late_chunk)), "lv_context_set_pass_schedule")

    def set_schedule(self, stream_fraction=0.002, min_streams=1, max_streams=1 << 20, chunk=0, order=2):
        """order: 0 degree tiers lowest first, 1 highest first, 2 ascending ids, no tiers."""
        self._check(self.lib.lv_context_set_schedule(self.h, C.c_double(stream_fraction), C.c_uint64(min_streams),
                                                     C.c_uint64(max_streams), C.c_uint64(chunk), int(order)),
                    "lv_context_set_schedule")

    def set_worklist_schedule(self, streams=3072, interleave=False):
        """Worklist sub-rounds: at most `streams` groups in the id-order kernel,
        the queue dealt out round-robin (interleave) or in consecutive parts."""
        self._check(self.lib.lv_context_set_worklist_schedule(self.h, C.c_uint64(streams), int(bool(interleave))),
                    "lv_context_set_worklist_schedule")

    def autotune(self, approach, dgraph, dbatch, dstate, o=None, reps=3):
        """lv_autotune: time the candidate schedules on clones of dstate with
        batch dbatch; the choice (a dict) is cached for dgraph's shape and applied."""
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        oc = (o or LouvainOptions())._c()
        out = _TuneChoice()
        self._check(self.lib.lv_autotune(self.h, int(approach), dgraph.h, dbatch.h, dstate.h, C.byref(oc), int(reps),
                                         C.byref(out)), "lv_autotune")
        return out.as_dict()

    def set_autotune(self, on=True):
        self._check(self.lib.lv_context_set_autotune(self.h, int(bool(on))), "lv_context_set_autotune")

    def apply_tuning(self, choice=None):
        """Set the tuned knobs by hand to a choice (dict), or to the defaults."""
        k = _TuneChoice.of(choice) if choice else None
        self._check(self.lib.lv_context_apply_tuning(self.h, C.byref(k) if k else None), "lv_context_apply_tuning")

    def add_tuning(self, approach, dgraph, choice):
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        k = _TuneChoice.of(choice)
        self._check(self.lib.lv_context_add_tuning(self.h, int(approach), dgraph.h, C.byref(k)), "lv_context_add_tuning")

    def tuning(self, approach, dgraph, graphs=1):
        """The cached choice for dgraph's shape (candidates 0: none cached)."""
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        out = _TuneChoice()
        self._check(self.lib.lv_context_tuning(self.h, int(approach), dgraph.h, C.c_uint64(graphs), C.byref(out)),
                    "lv_context_tuning")
        return out.as_dict()

    def set_graphs(self, on=True):
        """Replay worklist sub-round groups from captured HIP graphs (default on)."""
        self._check(self.lib.lv_context_set_graphs(self.h, int(bool(on))), "lv_context_set_graphs")

    def set_hub_tables(self, on=True):
        """Evaluate hubs of resident states from neighbour-community tables (default on)."""
        self._check(self.lib.lv_context_set_hub_tables(self.h, int(bool(on))), "lv_context_set_hub_tables")

    def check_hub_tables(self, dgraph, dstate):
        """(tables compared, tables differing from a fresh build) of a state's hub tables."""
        t, m = C.c_uint64(), C.c_uint64()
        self._check(self.lib.lv_dstate_check_hub_tables(self.h, dgraph.h, dstate.h, C.byref(t), C.byref(m)),
                    "lv_dstate_check_hub_tables")
        return t.value, m.value

    def set_profiling(self, on=True):
        self._check(self.lib.lv_context_set_profiling(self.h, int(bool(on))), "lv_context_set_profiling")

    def stats_reset(self):
        self._check(self.lib.lv_stats_reset(self.h), "lv_stats_reset")

    def stats(self):
        """Local-moving kernel statistics per degree tier (lv_stats)."""
        st = _Stats()
        self._check(self.lib.lv_stats_get(self.h, C.byref(st)), "lv_stats_get")
        return {f: (list(getattr(st, f)) if isinstance(getattr(st, f), C.Array) else getattr(st, f))
                for f, _ in _Stats._fields_}

    def _check(self, rc, what):
        if rc != 0:
            detail = self.lib.lv_last_error(self.h).decode() if self.h else ""
            raise LvError(rc, f"{what}: {self.lib.lv_status_string(rc).decode()}: {detail}")

    # -- the four entry points ------------------------------------------------
    def louvain(self, approach, g, deletions=None, insertions=None, q=None, qvtot=None, qctot=None,
                o=None):
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        o = o or LouvainOptions()
        ga = _GraphArgs(g)
        S = ga.S
        memb = np.zeros(S, np.uint32)
        vt = np.zeros(S, np.float64)
        ct = np.zeros(S, np.float64)
        r = _Result(_p(memb), _p(vt), _p(ct))
        oc = o._c()
        if approach == LV_STATIC:
            rc = self.lib.lv_louvain_static(self.h, C.byref(ga.c), C.byref(oc), C.byref(r))
        else:
            d, i = _BatchArgs(deletions), _BatchArgs(insertions)
            qq, qv, qc = _arr(q, np.uint32), _arr(qvtot, np.float64), _arr(qctot, np.float64)
            fn = {LV_NAIVE_DYNAMIC: self.lib.lv_louvain_naive_dynamic,
                  LV_DELTA_SCREENING: self.lib.lv_louvain_delta_screening,
                  LV_DYNAMIC_FRONTIER: self.lib.lv_louvain_dynamic_frontier}[approach]
            rc = fn(self.h, C.byref(ga.c), C.byref(d.c), C.byref(i.c), _p(qq), _p(qv), _p(qc), C.byref(oc),
                    C.byref(r))
        self._check(rc, "louvain")
        return _to_result(r, memb, vt, ct)

    # -- phase entry points ---------------------------------------------------
    def edge_weight(self, g):
        ga = _GraphArgs(g)
        out = C.c_double()
        self._check(self.lib.lv_edge_weight(self.h, C.byref(ga.c), C.byref(out)), "lv_edge_weight")
        return out.value

    def modularity(self, g, membership, resolution=1.0):
        ga = _GraphArgs(g)
        m = _arr(membership, np.uint32)
        out = C.c_double()
        self._check(self.lib.lv_modularity(self.h, C.byref(ga.c), _p(m), C.c_double(resolution), C.byref(out)),
                    "lv_modularity")
        return out.value

    def vertex_weights(self, g):
        ga = _GraphArgs(g)
        out = np.zeros(ga.S, np.float64)
        self._check(self.lib.lv_vertex_weights(self.h, C.byref(ga.c), _p(out)), "lv_vertex_weights")
        return out

    def update_weights(self, span, deletions, insertions, vcom, vtot, ctot):
        d, i = _BatchArgs(deletions), _BatchArgs(insertions)
        vc = _arr(vcom, np.uint32)
        vt = np.array(vtot, np.float64)
        ct = np.array(ctot, np.float64)
        self._check(self.lib.lv_update_weights(self.h, C.c_uint64(span), C.byref(d.c), C.byref(i.c), _p(vc),
                                               _p(vt), _p(ct)), "lv_update_weights")
        return vt, ct

    def affected_frontier(self, span, deletions, insertions, vcom):
        d, i = _BatchArgs(deletions), _BatchArgs(insertions)
        vc = _arr(vcom, np.uint32)
        out = np.zeros(span, np.uint8)
        self._check(self.lib.lv_affected_frontier(self.h, C.c_uint64(span), C.byref(d.c), C.byref(i.c), _p(vc),
                                                  _p(out)), "lv_affected_frontier")
        return out

    def affected_delta_screening(self, g, deletions, insertions, vcom, vtot, ctot, M, resolution=1.0):
        ga = _GraphArgs(g)
        d, i = _BatchArgs(deletions), _BatchArgs(insertions)
        vc, vt, ct = _arr(vcom, np.uint32), _arr(vtot, np.float64), _arr(ctot, np.float64)
        a, b, c = (np.zeros(ga.S, np.uint8) for _ in range(3))
        self._check(self.lib.lv_affected_delta_screening(self.h, C.byref(ga.c), C.byref(d.c), C.byref(i.c), _p(vc),
                                                         _p(vt), _p(ct), C.c_double(M), C.c_double(resolution),
                                                         _p(a), _p(b), _p(c)), "lv_affected_delta_screening")
        return a, b, c

    def choose_community(self, g, vcom, vtot, ctot, M, resolution=1.0):
        ga = _GraphArgs(g)
        vc, vt, ct = _arr(vcom, np.uint32), _arr(vtot, np.float64), _arr(ctot, np.float64)
        bc = np.zeros(ga.S, np.uint32)
        be = np.zeros(ga.S, np.float64)
        self._check(self.lib.lv_choose_community(self.h, C.byref(ga.c), _p(vc), _p(vt), _p(ct), C.c_double(M),
                                                 C.c_double(resolution), _p(bc), _p(be)), "lv_choose_community")
        return bc, be

    def renumber(self, g, vcom):
        ga = _GraphArgs(g)
        vc = np.array(vcom, np.uint32)
        cext = np.zeros(ga.S, np.uint32)
        n = C.c_uint64()
        self._check(self.lib.lv_renumber(self.h, C.byref(ga.c), _p(vc), C.c_uint64(len(vc)), _p(cext),
                                         C.byref(n)), "lv_renumber")
        return vc, cext, n.value

    def aggregate(self, g, vcom, C_):
        ga = _GraphArgs(g)
        E = int(ga.off[-1])
        vc = _arr(vcom, np.uint32)
        zoff = np.zeros(C_ + 1, np.uint64)
        zdeg = np.zeros(max(C_, 1), np.uint32)
        zkeys = np.zeros(max(E, 1), np.uint32)
        zw = np.zeros(max(E, 1), np.float64)
        self._check(self.lib.lv_aggregate(self.h, C.byref(ga.c), _p(vc), C.c_uint64(C_), _p(zoff), _p(zdeg),
                                          _p(zkeys), _p(zw)), "lv_aggregate")
        return {"off": zoff, "deg": zdeg[:C_], "keys": zkeys[:E], "w": zw[:E]}

    # -- device-resident form -------------------------------------------------
    def device_graph(self, g):
        return DeviceGraph(self, g)

    def device_batch(self, deletions, insertions):
        return DeviceBatch(self, deletions, insertions)

    def device_state(self, span):
        return DeviceState(self, span)

    def run(self, approach, dgraph, dbatch, dstate, o=None, copy_back=False):
        """lv_run on device-resident inputs; the state is updated in place."""
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        o = o or LouvainOptions()
        S = dgraph.span
        memb = vt = ct = None
        if copy_back:
            memb, vt, ct = np.zeros(S, np.uint32), np.zeros(S, np.float64), np.zeros(S, np.float64)
        r = _Result(_p(memb), _p(vt), _p(ct))
        oc = o._c()
        self._check(self.lib.lv_run(self.h, int(approach), dgraph.h, dbatch.h if dbatch else None, dstate.h,
                                    C.byref(oc), C.byref(r)), "lv_run")
        return _to_result(r, memb, vt, ct)

    def run_batched(self, approach, dgraph, dbatch, dstate, voff, o=None, copy_back=True):
        """lv_run_batched: every graph of the union at once; returns one
        LouvainResult per graph (graph-local ids when copy_back)."""
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        o = o or LouvainOptions()
        voff = np.ascontiguousarray(voff, np.uint64)
        G = len(voff) - 1
        res = (_Result * G)()
        keep = []
        for k in range(G):
            if copy_back:
                n = int(voff[k + 1] - voff[k])
                arrs = (np.zeros(n, np.uint32), np.zeros(n, np.float64), np.zeros(n, np.float64))
                keep.append(arrs)
                res[k] = _Result(_p(arrs[0]), _p(arrs[1]), _p(arrs[2]))
            else:
                keep.append((None, None, None))
        oc = o._c()
        self._check(self.lib.lv_run_batched(self.h, int(approach), dgraph.h, dbatch.h if dbatch else None, dstate.h,
                                            _p(voff), C.c_uint64(G), C.byref(oc), res), "lv_run_batched")
        return [_to_result(res[k], *keep[k]) for k in range(G)]

    def dmodularity_noalloc(self, dgraph, dstate, resolution=1.0):
        """lv_dmodularity_noalloc: no device allocation, spans up to 2^21."""
        out = C.c_double(0)
        self._check(self.lib.lv_dmodularity_noalloc(self.h, dgraph.h, dstate.h, C.c_double(resolution),
                                                    C.byref(out)), "lv_dmodularity_noalloc")
        return out.value

    def dmodularity(self, dgraph, dstate, resolution=1.0):
        out = C.c_double()
        self._check(self.lib.lv_dmodularity(self.h, dgraph.h, dstate.h, C.c_double(resolution), C.byref(out)),
                    "lv_dmodularity")
        return out.value


class DeviceGraph:
    def __init__(self, ctx, g):
        self.ctx = ctx
        self.h = C.c_void_p()
        ga = _GraphArgs(g)
        ctx._check(ctx.lib.lv_dgraph_create(ctx.h, C.byref(ga.c), C.byref(self.h)), "lv_dgraph_create")
        self.span = ctx.lib.lv_dgraph_span(self.h)
        self.edges = ctx.lib.lv_dgraph_edges(self.h)

    def update(self, deletions, insertions, ctx=None):
        """lv_dgraph_update: tidy the raw batch on the device, apply it to this
        graph in place (on ctx's stream, default the creating context);
        returns the tidy batch as a DeviceBatch."""
        ctx = ctx or self.ctx
        d, i = _BatchArgs(deletions), _BatchArgs(insertions)
        b = DeviceBatch.__new__(DeviceBatch)
        b.ctx, b.h = ctx, C.c_void_p()
        ctx._check(ctx.lib.lv_dgraph_update(ctx.h, self.h, C.byref(d.c), C.byref(i.c), C.byref(b.h)),
                   "lv_dgraph_update")
        nd, ni = C.c_uint64(), C.c_uint64()
        ctx.lib.lv_dbatch_sizes(b.h, C.byref(nd), C.byref(ni))
        b.size = nd.value + ni.value
        self.edges = ctx.lib.lv_dgraph_edges(self.h)
        return b

    def download(self):
        """The CSR as dict(off, deg, keys, w, ex) (rows may carry slack)."""
        S, E = self.span, self.ctx.lib.lv_dgraph_capacity(self.h)
        off, deg = np.zeros(S + 1, np.uint64), np.zeros(S, np.uint32)
        keys, w, ex = np.zeros(max(E, 1), np.uint32), np.zeros(max(E, 1), np.float32), np.zeros(S, np.uint8)
        self.ctx._check(self.ctx.lib.lv_dgraph_download(self.ctx.h, self.h, _p(off), _p(deg), _p(keys), _p(w), _p(ex)),
                        "lv_dgraph_download")
        return {"off": off, "deg": deg, "keys": keys[:E], "w": w[:E], "ex": ex}

    def close(self):
        if self.h:
            self.ctx.lib.lv_dgraph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    def __init__(self, ctx, deletions, insertions):
        self.ctx = ctx
        self.h = C.c_void_p()
        d, i = _BatchArgs(deletions), _BatchArgs(insertions)
        self.size = len(d.u) + len(i.u)
        ctx._check(ctx.lib.lv_dbatch_create(ctx.h, C.byref(d.c), C.byref(i.c), C.byref(self.h)), "lv_dbatch_create")

    def download(self):
        """(deletions, insertions) as dicts of u, v, w arrays."""
        nd, ni = C.c_uint64(), C.c_uint64()
        self.ctx.lib.lv_dbatch_sizes(self.h, C.byref(nd), C.byref(ni))
        nd, ni = nd.value, ni.value
        du, dv, dw = np.zeros(max(nd, 1), np.uint32), np.zeros(max(nd, 1), np.uint32), np.zeros(max(nd, 1), np.float32)
        iu, iv, iw = np.zeros(max(ni, 1), np.uint32), np.zeros(max(ni, 1), np.uint32), np.zeros(max(ni, 1), np.float32)
        self.ctx._check(self.ctx.lib.lv_dbatch_download(self.ctx.h, self.h, _p(du), _p(dv), _p(dw), _p(iu), _p(iv),
                                                        _p(iw)), "lv_dbatch_download")
        return ({"u": du[:nd], "v": dv[:nd], "w": dw[:nd]}, {"u": iu[:ni], "v": iv[:ni], "w": iw[:ni]})

    def close(self):
        if self.h:
            self.ctx.lib.lv_dbatch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceState:
    def __init__(self, ctx, span):
        self.ctx = ctx
        self.span = span
        self.h = C.c_void_p()
        ctx._check(ctx.lib.lv_dstate_create(ctx.h, C.c_uint64(span), C.byref(self.h)), "lv_dstate_create")

    def upload(self, q, qvtot, qctot):
        qq, qv, qc = _arr(q, np.uint32), _arr(qvtot, np.float64), _arr(qctot, np.float64)
        self.ctx._check(self.ctx.lib.lv_dstate_upload(self.ctx.h, self.h, _p(qq), _p(qv), _p(qc)), "lv_dstate_upload")

    def download(self, ctx=None):
        """(membership, K, Sigma_tot), copied on ctx's stream (default the
        creating context)."""
        ctx = ctx or self.ctx
        q = np.zeros(self.span, np.uint32)
        v = np.zeros(self.span, np.float64)
        c = np.zeros(self.span, np.float64)
        ctx._check(ctx.lib.lv_dstate_download(ctx.h, self.h, _p(q), _p(v), _p(c)), "lv_dstate_download")
        return q, v, c

    def copy_from(self, other):
        self.ctx._check(self.ctx.lib.lv_dstate_copy(self.ctx.h, self.h, other.h), "lv_dstate_copy")

    def clone_from(self, other):
        """lv_dstate_clone: other's state with its hub tables."""
        self.ctx._check(self.ctx.lib.lv_dstate_clone(self.ctx.h, self.h, other.h), "lv_dstate_clone")

    def close(self):
        if self.h:
            self.ctx.lib.lv_dstate_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _to_result(r, memb, vt, ct):
    return LouvainResult(memb, vt, ct, r.iterations, r.passes, r.time, r.marking_time, r.initialization_time,
                         r.first_pass_time, r.local_move_time, r.aggregation_time, r.affected_vertices)


_DEFAULT_CTX = None


def default_context():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


# ---- the reference's four entry points, by name -------------------------------
def louvainStaticHip(x, o=None):
    """louvainStaticOmp (inc/louvain.hxx:1285-1296) on the GPU."""
    return default_context().louvain(LV_STATIC, x, o=o)


def louvainNaiveDynamicHip(y, deletions, insertions, q, qvtot, qctot, o=None):
    """louvainNaiveDynamicOmp (inc/louvain.hxx:1349-1366) on the GPU."""
    return default_context().louvain(LV_NAIVE_DYNAMIC, y, deletions, insertions, q, qvtot, qctot, o)


def louvainDynamicDeltaScreeningHip(y, deletions, insertions, q, qvtot, qctot, o=None):
    """louvainDynamicDeltaScreeningOmp (inc/louvain.hxx:1542-1565) on the GPU."""
    return default_context().louvain(LV_DELTA_SCREENING, y, deletions, insertions, q, qvtot, qctot, o)


def louvainDynamicFrontierHip(y, deletions, insertions, q, qvtot, qctot, o=None):
    """louvainDynamicFrontierOmp (inc/louvain.hxx:1674-1690) on the GPU."""
    return default_context().louvain(LV_DYNAMIC_FRONTIER, y, deletions, insertions, q, qvtot, qctot, o)


# ---------------------------------------------------------------- batched form
def union_graphs(graphs):
    """Graphs (numpy CSR dicts) side by side: (union dict, voff), graph k
    owning ids [voff[k], voff[k+1])."""
    spans = [len(g["off"]) - 1 for g in graphs]
    voff = np.concatenate([[0], np.cumsum(spans)]).astype(np.uint64)
    eoff = np.concatenate([[0], np.cumsum([int(g["off"][-1]) for g in graphs])]).astype(np.uint64)
    off = np.concatenate([np.asarray(g["off"][:-1], np.uint64) + eoff[k] for k, g in enumerate(graphs)]
                         + [eoff[-1:]]).astype(np.uint64)
    keys = np.concatenate([np.asarray(g["keys"], np.uint32) + np.uint32(voff[k]) for k, g in enumerate(graphs)])
    w = np.concatenate([np.asarray(g["w"], np.float32) for g in graphs])
    u = {"off": off, "keys": keys, "w": w}
    if any("ex" in g for g in graphs):
        u["ex"] = np.concatenate([np.asarray(g.get("ex", np.ones(n, np.uint8)), np.uint8)
                                  for g, n in zip(graphs, spans)])
    return u, voff


def union_batch(batches, voff):
    """Per-graph batches (dicts u, v, w) as one batch over the union."""
    return {"u": np.concatenate([np.asarray(b["u"], np.uint32) + np.uint32(voff[k]) for k, b in enumerate(batches)]),
            "v": np.concatenate([np.asarray(b["v"], np.uint32) + np.uint32(voff[k]) for k, b in enumerate(batches)]),
            "w": np.concatenate([np.asarray(b["w"], np.float32) for b in batches])}


def union_state(states, voff):
    """Per-graph (membership, K, Sigma_tot) as one union state: graph k's
    community ids shifted into its range."""
    S = int(voff[-1])
    q, v, c = np.zeros(S, np.uint32), np.zeros(S, np.float64), np.zeros(S, np.float64)
    for k, (qk, vk, ck) in enumerate(states):
        a, b = int(voff[k]), int(voff[k + 1])
        q[a:b] = np.asarray(qk, np.uint32) + np.uint32(a)
        v[a:b] = vk
        c[a:b] = ck
    return q, v, c
