"""Hub neighbour-community tables kept with a resident state, through the
C ABI: after every DF run of a chain the tables equal a fresh build from the
graph and the state's membership (counts exact, weights within 1e-9
relative), and each batch's modularity is within 1e-3 of the oracle's DF
from the same carry-in state, with K exact.

A chain starts from the reference's Static result (oracle, one thread).  The
first run of an uploaded state has no tables; the second builds them; later
runs bring them to their graph with the batch's edges and keep them through
every move.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q_TOL = 1e-3


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import lvhip
    c = lvhip.Context(0)
    yield c
    c.close()


def _chain(scale, seed, n, n_del, n_ins):
    import lvgraph
    graphs, batches = [lvgraph.rmat(scale, 16, seed=seed)], []
    for k in range(n):
        d, i = lvgraph.random_batch(graphs[-1], n_del, n_ins, seed * 100 + k)
        batches.append((d, i))
        graphs.append(lvgraph.apply_batch(graphs[-1], d, i))
    return graphs, batches


def _check_run(ctx, oracle, g, dg, db, d, i, st, k, expect_tables):
    import lvhip
    prev = st.download()
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db, st)
    t, m = ctx.check_hub_tables(dg, st)
    q = ctx.dmodularity(dg, st)
    o = oracle.louvain("dynamic-frontier", g, d, i, *prev)
    q_ref = oracle.modularity(g, o.membership)
    _say(f"  run {k}: tables {t} mismatched {m}; Q gpu={q:.6f} ref={q_ref:.6f}")
    assert m == 0, (k, t, m)
    if expect_tables:
        assert t > 0, k
    assert abs(q - q_ref) <= Q_TOL, (k, q, q_ref)
    assert np.array_equal(st.download()[1], o.vertex_weight), k


def test_tables_exact_through_a_chain(ctx, oracle):
    """R-MAT scale 15, six batches, each graph its own resident copy (the
    reference's driver makes a new graph per batch, main.cxx:105-109): no
    graph is its predecessor updated in place, so the tables are rebuilt on
    every run from the second on."""
    graphs, batches = _chain(15, 11, 6, 150, 600)
    s0 = oracle.louvain("static", graphs[0])
    S = len(graphs[0]["off"]) - 1
    st = ctx.device_state(S)
    st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
    assert ctx.check_hub_tables(ctx.device_graph(graphs[0]), st) == (0, 0)  # none after an upload
    for k, ((d, i), g) in enumerate(zip(batches, graphs[1:])):
        _check_run(ctx, oracle, g, ctx.device_graph(g), ctx.device_batch(d, i), d, i, st, k, expect_tables=k >= 1)


def test_tables_exact_with_graph_updated_in_place(ctx, oracle):
    """The serving pattern: one resident graph updated in place by each raw
    batch (device tidy and update), the state carried."""
    graphs, batches = _chain(14, 12, 5, 100, 400)
    s0 = oracle.louvain("static", graphs[0])
    S = len(graphs[0]["off"]) - 1
    dg = ctx.device_graph(graphs[0])
    st = ctx.device_state(S)
    st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
    for k, ((d, i), g) in enumerate(zip(batches, graphs[1:])):
        db = dg.update(d, i)
        _check_run(ctx, oracle, g, dg, db, d, i, st, k, expect_tables=k >= 1)


def test_tables_off_matches_on(ctx, oracle):
    """The same chain with the tables switched off: modularity within 1e-3."""
    import lvhip
    graphs, batches = _chain(15, 13, 4, 150, 600)
    s0 = oracle.louvain("static", graphs[0])
    S = len(graphs[0]["off"]) - 1
    dgs = [ctx.device_graph(g) for g in graphs[1:]]
    dbs = [ctx.device_batch(d, i) for d, i in batches]
    qs = {}
    try:
        for on in (True, False):
            ctx.set_hub_tables(on)
            st = ctx.device_state(S)
            st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
            qs[on] = []
            for dg, db in zip(dgs, dbs):
                ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db, st)
                qs[on].append(ctx.dmodularity(dg, st))
            if not on:
                assert ctx.check_hub_tables(dgs[-1], st) == (0, 0)
    finally:
        ctx.set_hub_tables(True)
    _say("  Q with tables", [round(q, 6) for q in qs[True]], "without", [round(q, 6) for q in qs[False]])
    assert max(abs(a - b) for a, b in zip(qs[True], qs[False])) <= Q_TOL


def test_table_dropped_by_a_batch(ctx, oracle):
    """A hub of 300 edges whose table (256 slots) a batch overflows with 300
    new neighbour communities: the batch drops the table, the run starts
    without the edge-scanning hub kernels, and the hub, queued for them, is
    evaluated from its edges a group later.  Tables of the other vertices
    stay exact; modularity within 1e-3 of the oracle's, K exact."""
    import lvgraph
    import lvhip
    S, hub = 1600, 0
    # hub - 1..300, those in pairs; 1000..1599 in 300 disconnected pairs
    near = np.arange(1, 301, dtype=np.uint32)
    far = np.arange(1000, 1600, dtype=np.uint32)
    pu = np.concatenate([near[0::2], far[0::2]])
    pv = np.concatenate([near[1::2], far[1::2]])
    u = np.concatenate([np.full(300, hub, np.uint32), near, pu, pv])
    v = np.concatenate([near, np.full(300, hub, np.uint32), pv, pu])
    x = lvgraph.from_edges(S, u, v)

    def ins(a, b):
        a, b = np.asarray(a, np.uint32), np.asarray(b, np.uint32)
        return {"u": np.concatenate([a, b]), "v": np.concatenate([b, a]), "w": np.ones(2 * len(a), np.float32)}

    none = {"u": np.zeros(0, np.uint32), "v": np.zeros(0, np.uint32), "w": np.zeros(0, np.float32)}
    batches = [(none, ins([1], [4])), (none, ins([5], [8])),  # small batches: runs 0 (no tables) and 1 (build)
               (none, ins(np.full(300, hub), far[0::2]))]  # the hub meets 300 singleton-pair communities
    graphs = [x]
    for d, i in batches:
        graphs.append(lvgraph.apply_batch(graphs[-1], d, i))
    s0 = oracle.louvain("static", x)
    st = ctx.device_state(S)
    st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
    dg = ctx.device_graph(x)  # updated in place: the tables are carried from run to run
    for k, ((d, i), g) in enumerate(zip(batches, graphs[1:])):
        db = dg.update(d, i)
        prev = st.download()
        ctx.stats_reset()
        ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db, st)
        stats = ctx.stats()
        t, m = ctx.check_hub_tables(dg, st)
        q = ctx.dmodularity(dg, st)
        o = oracle.louvain("dynamic-frontier", g, d, i, *prev)
        q_ref = oracle.modularity(g, o.membership)
        _say(f"  run {k}: tables {t} mismatched {m}; small hubs {stats['vertices'][2]}; Q gpu={q:.6f} ref={q_ref:.6f}")
        assert m == 0, (k, t, m)
        assert abs(q - q_ref) <= Q_TOL, (k, q, q_ref)
        assert np.array_equal(st.download()[1], o.vertex_weight), k
        if k == 1:
            assert t >= 1  # the hub's table was built
        if k == 2:
            assert t == 0  # ... and dropped by the batch
            assert stats["vertices"][2] >= 1  # the hub went through the block-per-hub kernel


def test_tables_rebuilt_when_the_batch_does_not_chain(ctx, oracle):
    """Tables are carried forward only when the run's batch was made by
    lv_dgraph_update from the graph they describe into the run's graph.  The
    same update batch run a second time, and a host-made batch of the same
    edges, do not chain: the run rebuilds the tables, which equal a fresh
    build afterwards."""
    import lvgraph
    import lvhip
    graphs, batches = _chain(14, 17, 2, 100, 400)
    s0 = oracle.louvain("static", graphs[0])
    S = len(graphs[0]["off"]) - 1
    dg = ctx.device_graph(graphs[0])
    st = ctx.device_state(S)
    st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
    db0 = dg.update(*batches[0])
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db0, st)  # first run of the state: no tables
    db1 = dg.update(*batches[1])
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db1, st)  # second: tables built
    t, m = ctx.check_hub_tables(dg, st)
    _say(f"  built: tables {t} mismatched {m}")
    assert t > 0 and m == 0
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db1, st)  # the same batch again: rebuilt
    t, m = ctx.check_hub_tables(dg, st)
    _say(f"  same batch again: tables {t} mismatched {m}")
    assert t > 0 and m == 0
    hb = ctx.device_batch(*lvgraph.tidy_batch(graphs[1], *batches[1]))
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, hb, st)  # a host batch: rebuilt
    t, m = ctx.check_hub_tables(dg, st)
    _say(f"  host batch: tables {t} mismatched {m}")
    assert t > 0 and m == 0
