"""Serving loop (lvserve): resident graphs, requests admitted into slots as
they free up, work of different graphs overlapping on the device.  Each
request's result in the mix equals its result served alone (graphs of at
most 2048 vertices run as one stream, so both are deterministic)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs():
    import lvgraph
    out = {}
    for k, n in enumerate([1500, 1800, 2000, 1200, 1700, 1900]):
        out[f"pp{k}"] = lvgraph.planted_partition(n, 12, 16, 0.25, 100 + k)
    return out


def _requests(graphs, per_graph, seed):
    import lvserve
    from test_gpu_update import _raw_batch
    rng = np.random.default_rng(seed)
    reqs, rid = [], 0
    for _ in range(per_graph):
        for name, g in graphs.items():
            n = int(rng.integers(4, 40))
            d, i = _raw_batch(g, rng, n)
            reqs.append(lvserve.Request(rid, name, d, i, 0.0))
            rid += 1
    return reqs


def test_results_alone_equal_results_in_the_mix(oracle):
    import lvserve
    graphs = _graphs()
    alone_srv = lvserve.Server(graphs, slots=1, keep_states=True, modularity=True)
    mix_srv = lvserve.Server(graphs, slots=4, keep_states=True, modularity=True)
    try:
        reqs_a = _requests(graphs, 4, 7)
        reqs_m = _requests(graphs, 4, 7)
        # alone: one at a time, each finished before the next is released
        for r in reqs_a:
            alone_srv.run([r], realtime=False)
        # mix: all released at once into four slots
        mix_srv.run(reqs_m, realtime=False)
        overlap = sum(1 for a in reqs_m for b in reqs_m
                      if a.graph != b.graph and a.start < b.done and b.start < a.done)
        print("overlapping request pairs:", overlap, flush=True)
        for a, m in zip(reqs_a, reqs_m):
            assert a.error is None and m.error is None, (a.error, m.error)
            assert np.array_equal(a.state[0], m.state[0]), a.rid
            assert np.array_equal(a.state[1], m.state[1]), a.rid
            assert np.allclose(a.state[2], m.state[2], rtol=1e-12, atol=1e-9), a.rid
            assert a.modularity == m.modularity
        # requests of one graph ran in arrival order, one at a time
        for name in graphs:
            mine = sorted((r for r in reqs_m if r.graph == name), key=lambda r: r.rid)
            for x, y in zip(mine, mine[1:]):
                assert x.done <= y.start
    finally:
        alone_srv.close()
        mix_srv.close()


def test_poisson_serving_reports_latency():
    import lvserve
    graphs = _graphs()
    srv = lvserve.Server(graphs, slots=3)
    try:
        reqs = _requests(graphs, 10, 3)
        for r, t in zip(reqs, lvserve.poisson_arrivals(len(reqs), 400.0, 9)):
            r.arrival = float(t)
        done = srv.run(reqs, realtime=True)
        s = lvserve.summary(done)
        print(s, flush=True)
        assert s["errors"] == 0 and s["requests"] == len(reqs)
        assert all(r.start >= r.arrival for r in done)
    finally:
        srv.close()
