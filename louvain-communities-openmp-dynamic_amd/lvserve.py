"""Serving loop: resident graphs, DF on update batches as they arrive.

Graphs, memberships, K and Sigma_tot stay in HBM (lv_dgraph / lv_dstate).
Each request is a raw batch (deletions, insertions) for one graph; the
worker that takes it tidies and applies it on the device (lv_dgraph_update)
and runs Dynamic Frontier Louvain from the graph's current state (lv_run),
the per-batch step of main.cxx:104-135.

Continuous batching: a fixed set of slots (one lv_context, hence one HIP
stream, per slot) take requests in arrival order as they become free, so
requests for different graphs run on the device at the same time.  Requests
for one graph run one at a time, in arrival order (each batch updates the
graph the next one applies to).  Admission is by device-memory footprint: a
request enters a slot only while the bytes of the batches in flight stay
within the budget.
"""
import heapq
import threading
import time

import numpy as np

import lvhip


class Request:
    __slots__ = ("rid", "graph", "deletions", "insertions", "arrival", "start", "done", "result", "modularity",
                 "bytes", "error", "state")

    def __init__(self, rid, graph, deletions, insertions, arrival):
        self.rid, self.graph = rid, graph
        self.deletions, self.insertions = deletions, insertions
        self.arrival = arrival
        self.start = self.done = None
        self.result = self.modularity = self.error = self.state = None
        # device bytes while in flight: the raw batch and its tidy copy
        n = len(deletions["u"]) + len(insertions["u"])
        self.bytes = 2 * 12 * n

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class Server:
    """Resident graphs served by `slots` workers.  graphs maps a name to a
    host graph (numpy CSR dict or anything with view()); each starts from a
    Static Louvain run on the device unless `states` gives (q, K, Sigma_tot)."""

    def __init__(self, graphs, slots=4, memory_budget=1 << 30, options=None, device=0, path=None,
                 modularity=False, states=None, keep_states=False):
        self.slots = [lvhip.Context(device) for _ in range(slots)]
        if path is not None:
            for c in self.slots:
                c.set_path(path)
        self.options = options or lvhip.LouvainOptions()
        self.budget = memory_budget
        self.want_modularity = modularity
        self.keep_states = keep_states
        ctx = self.slots[0]
        self.graphs, self.states = {}, {}
        for name, g in graphs.items():
            view = g.view() if hasattr(g, "view") else g
            dg = ctx.device_graph(view)
            st = ctx.device_state(dg.span)
            if states and name in states:
                st.upload(*states[name])
            else:
                ctx.run(lvhip.LV_STATIC, dg, None, st, self.options)
            self.graphs[name], self.states[name] = dg, st
        self._lock = threading.Condition()

    def close(self):
        for d in list(self.graphs.values()) + list(self.states.values()):
            d.close()
        for c in self.slots:
            c.close()

    def run(self, requests, realtime=True):
        """Serve the requests (arrival times in seconds from the start, in
        order).  With realtime, a request is released at its arrival time;
        otherwise all are released at once.  Returns the requests, each with
        start / done times (seconds from the start), result and, if asked,
        modularity."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        waiting = []             # released, not started: (arrival, rid, request)
        busy_graphs = set()
        in_flight = [0]
        state = {"released": 0}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        cv = self._lock

        def take():
            # oldest released request whose graph is idle and whose batch fits the budget
            skipped = []
            got = None
            while waiting:
                item = heapq.heappop(waiting)
                r = item[2]
                if r.graph in busy_graphs:
                    skipped.append(item)
                    continue
                if in_flight[0] and in_flight[0] + r.bytes > self.budget:
                    skipped.append(item)
                    break
                got = r
                break
            for it in skipped:
                heapq.heappush(waiting, it)
            return got

        def worker(ctx):
            while True:
                with cv:
                    while True:
                        r = take()
                        if r is not None:
                            break
                        if state["released"] == len(pending) and not waiting:
                            return  # everything is taken; the rest is running elsewhere
                        cv.wait(0.05)
                    busy_graphs.add(r.graph)
                    in_flight[0] += r.bytes
                r.start = now()
                try:
                    dg, st = self.graphs[r.graph], self.states[r.graph]
                    tb = dg.update(r.deletions, r.insertions, ctx=ctx)
                    r.result = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, tb, st, self.options)
                    tb.close()
                    if self.want_modularity:
                        r.modularity = ctx.dmodularity(dg, st)
                    if self.keep_states:
                        r.state = st.download(ctx)
                except Exception as e:  # reported per request
                    r.error = e
                r.done = now()
                with cv:
                    busy_graphs.discard(r.graph)
                    in_flight[0] -= r.bytes
                    cv.notify_all()

        threads = [threading.Thread(target=worker, args=(c,), daemon=True) for c in self.slots]
        for t in threads:
            t.start()
        for r in pending:
            if realtime:
                delay = r.arrival - now()
                if delay > 0:
                    time.sleep(delay)
            else:
                r.arrival = now()
            with cv:
                heapq.heappush(waiting, (r.arrival, r.rid, r))
                state["released"] += 1
                cv.notify_all()
        for t in threads:
            t.join()
        return pending


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
