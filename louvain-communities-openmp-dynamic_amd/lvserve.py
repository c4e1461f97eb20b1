"""Serving loop: resident graphs, DF on update batches as they arrive.

A thin host-side layer over the library's lv_server_* C ABI
(include/louvain_hip.h).  Graphs, memberships, K, Sigma_tot and the hub
neighbour-community tables stay in HBM.  Each request is a raw batch
(deletions, insertions) for one graph; the slot that takes it tidies and
applies it on the device (lv_dgraph_update) and runs Dynamic Frontier
Louvain from the graph's current state (lv_run), the per-batch step of
main.cxx:104-135.

Continuous batching happens inside the library: a fixed set of slots (one
lv_context, hence one HIP stream and workspace, per slot, each on a worker
thread) take requests in submission order as they become free, so requests
for different graphs run on the device at the same time; requests for one
graph run one at a time, in submission order (each batch updates the graph
the next one applies to).  Admission is by device memory: a request starts
only while the resident graphs and states, the slots' workspaces and the
batches in flight fit the budget, or when nothing else runs.  Adding a
graph runs a warm-up on every slot, so workspaces are sized before serving.
"""
import ctypes as C
import time

import numpy as np

import lvhip


class Request:
    __slots__ = ("rid", "graph", "deletions", "insertions", "arrival", "start", "done", "result", "modularity",
                 "bytes", "error", "state", "ticket")

    def __init__(self, rid, graph, deletions, insertions, arrival):
        self.rid, self.graph = rid, graph
        self.deletions, self.insertions = deletions, insertions
        self.arrival = arrival
        self.start = self.done = self.ticket = None
        self.result = self.modularity = self.error = self.state = None
        # host bytes of the raw batch (u, v, w per entry)
        self.bytes = 12 * (len(deletions["u"]) + len(insertions["u"]))

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class Server:
    """Resident graphs served by `slots` workers of an lv_server.  graphs
    maps a name to a host graph (numpy CSR dict or anything with view());
    each starts from a Static Louvain run on the device unless `states`
    gives (q, K, Sigma_tot).  memory_budget: device bytes for graphs,
    states, workspaces and batches in flight (None: no limit)."""

    def __init__(self, graphs, slots=4, memory_budget=None, options=None, device=0, path=None,
                 modularity=False, states=None, keep_states=False, autotune=True, probe_fraction=1e-4):
        self.lib = lvhip._LIB
        o = lvhip._ServerOptions()
        self.lib.lv_server_options_default(C.byref(o))
        o.device, o.slots = int(device), int(slots)
        o.memory_budget = int(memory_budget or 0)
        o.options = (options or lvhip.LouvainOptions())._c()
        if path is not None:
            o.path = int(path)
        o.modularity, o.keep_states = int(bool(modularity)), int(bool(keep_states))
        o.autotune, o.probe_fraction = int(bool(autotune)), float(probe_fraction)
        self.keep_states = keep_states
        self.h = C.c_void_p()
        rc = self.lib.lv_server_create(C.byref(o), C.byref(self.h))
        if rc != 0:
            raise lvhip.LvError(rc, "lv_server_create: " + self.lib.lv_status_string(rc).decode())
        self.ids, self.spans = {}, {}
        for name, g in graphs.items():
            view = g.view() if hasattr(g, "view") else g
            ga = lvhip._GraphArgs(view)
            gid = C.c_uint32()
            st = states.get(name) if states else None
            if st is not None:
                q, v, c = (lvhip._arr(st[0], np.uint32), lvhip._arr(st[1], np.float64), lvhip._arr(st[2], np.float64))
                rc = self.lib.lv_server_add_graph(self.h, C.byref(ga.c), lvhip._p(q), lvhip._p(v), lvhip._p(c),
                                                  C.byref(gid))
            else:
                rc = self.lib.lv_server_add_graph(self.h, C.byref(ga.c), None, None, None, C.byref(gid))
            self._check(rc, f"lv_server_add_graph({name})")
            self.ids[name], self.spans[name] = gid.value, ga.S

    def _check(self, rc, what):
        if rc != 0:
            raise lvhip.LvError(rc, f"{what}: {self.lib.lv_server_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.lv_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        """Seconds on the server's clock."""
        return self.lib.lv_server_now_ms(self.h) * 1e-3

    def submit(self, r):
        """Queue one request (its batch is copied); returns its ticket."""
        d, i = lvhip._BatchArgs(r.deletions), lvhip._BatchArgs(r.insertions)
        t = C.c_uint64()
        self._check(self.lib.lv_server_submit(self.h, self.ids[r.graph], C.byref(d.c), C.byref(i.c), C.byref(t)),
                    "lv_server_submit")
        r.ticket = t.value
        return r.ticket

    def poll(self, r, timeout_ms=-1, t0=0.0):
        """Wait up to timeout_ms (< 0: until done) for a submitted request and
        fill its start / done times (seconds since t0 on the server's clock),
        result, modularity, state and error.  False if still pending."""
        out = lvhip._Completion()
        state = None
        if self.keep_states:
            S = self.spans[r.graph]
            state = (np.zeros(S, np.uint32), np.zeros(S, np.float64), np.zeros(S, np.float64))
            out.result.membership, out.result.vertex_weight, out.result.community_weight = (
                state[0].ctypes.data, state[1].ctypes.data, state[2].ctypes.data)
        rc = self.lib.lv_server_poll(self.h, C.c_uint64(r.ticket), int(timeout_ms), C.byref(out))
        if rc == lvhip.LV_ERR_PENDING:
            return False
        self._check(rc, "lv_server_poll")
        r.start, r.done = out.started_ms * 1e-3 - t0, out.done_ms * 1e-3 - t0
        if out.status != 0:
            r.error = lvhip.LvError(out.status, self.lib.lv_server_last_error(self.h).decode())
            return True
        r.result = lvhip._to_result(out.result, *(state or (None, None, None)))
        r.modularity = out.modularity if not np.isnan(out.modularity) else None
        r.state = state
        return True

    def run(self, requests, realtime=True):
        """Serve the requests (arrival times in seconds from the start, in
        order).  With realtime, a request is submitted at its arrival time;
        otherwise all are submitted at once.  Returns the requests, each with
        start / done times (seconds from the start), result and, if asked,
        modularity and state."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = self.now()
        for r in pending:
            if realtime:
                delay = t0 + r.arrival - self.now()
                if delay > 0:
                    time.sleep(delay)
            else:
                r.arrival = self.now() - t0
            self.submit(r)
        for r in pending:
            self.poll(r, -1, t0)
        return pending

    def graph_state(self, name):
        """(membership, K, Sigma_tot) of a graph once its requests are done."""
        S = self.spans[name]
        q, v, c = np.zeros(S, np.uint32), np.zeros(S, np.float64), np.zeros(S, np.float64)
        self._check(self.lib.lv_server_graph_state(self.h, self.ids[name], lvhip._p(q), lvhip._p(v), lvhip._p(c)),
                    "lv_server_graph_state")
        return q, v, c

    def tuning(self, name):
        """The DF schedule chosen for a graph when it was added (a dict;
        candidates 0: not tuned)."""
        k = lvhip._TuneChoice()
        self._check(self.lib.lv_server_graph_tuning(self.h, self.ids[name], C.byref(k)), "lv_server_graph_tuning")
        return k.as_dict()

    def stats(self):
        st = lvhip._ServerStats()
        self._check(self.lib.lv_server_get_stats(self.h, C.byref(st)), "lv_server_get_stats")
        return {f: getattr(st, f) for f, _ in lvhip._ServerStats._fields_}


def poisson_arrivals(n, rate, seed):
    """n arrival times (seconds) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def summary(requests):
    lat = np.array([r.latency for r in requests if r.error is None], dtype=np.float64) * 1e3
    errors = sum(r.error is not None for r in requests)
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"requests": len(requests), "errors": errors, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "throughput_rps": len(requests) / span if span > 0 else None}
