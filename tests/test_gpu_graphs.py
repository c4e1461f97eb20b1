"""Worklist sub-round groups replayed from captured HIP graphs, against the
same launches enqueued one by one, through the C ABI.

One stream and no hub kernels make a run deterministic, so every replay --
within a run and across runs on the same resident graph, with other device
work and host synchronisations between them -- must give the eager result
bit for bit.  With the default schedule (hub kernels on side streams beside
the low-degree kernel) runs are not deterministic; there the bar is the
usual one: K exact and modularity within 1e-3 of the eager run.
"""
import numpy as np
import pytest

import fixtures as F

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import lvhip
    c = lvhip.Context(0)
    yield c
    c.close()


def _runs(ctx, dg, db, st, state0, n):
    import lvhip
    out = []
    for _ in range(n):
        st.upload(*state0)
        r = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, db, st)
        q = ctx.dmodularity(dg, st)  # device work between replays
        m, vt, ct = st.download()
        out.append((m, vt, ct, r.iterations, r.affectedVertices, q))
    return out


@pytest.mark.parametrize("name", ["pp2k", "rmat10"])
def test_graph_replays_equal_eager_one_stream(ctx, name):
    _say("case", name)
    f = F.load(name)
    y, d, i = F.graph(f, "y"), F.batch(f, "del"), F.batch(f, "ins")
    dg, db = ctx.device_graph(y), ctx.device_batch(d, i)
    st = ctx.device_state(len(y["off"]) - 1)
    state0 = (f["static_memb"], f["static_vtot"], f["static_ctot"])
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_hub_threshold(1 << 30, 4096)
    try:
        ctx.set_graphs(False)
        eager = _runs(ctx, dg, db, st, state0, 1)[0]
        ctx.set_graphs(True)
        for k, got in enumerate(_runs(ctx, dg, db, st, state0, 4)):
            _say(f"  replay run {k}: it={got[3]} aff={got[4]} Q={got[5]:.6f}")
            for a, b in zip(got[:3], eager[:3]):
                assert np.array_equal(a, b), k
            assert got[3:] == eager[3:], k
    finally:
        ctx.set_graphs(True)
        ctx.set_schedule()
        ctx.set_hub_threshold()


def test_graph_replays_default_schedule(ctx, oracle):
    """R-MAT scale 15 (hub kernels on the side streams): consecutive batches
    on one resident graph, replayed and eager."""
    import lvgraph
    x = lvgraph.rmat(15, 16, seed=5)
    s = oracle.louvain("static", x)
    d, i = lvgraph.random_batch(x, 300, 1200, 77)
    y = lvgraph.apply_batch(x, d, i)
    dg, db = ctx.device_graph(y), ctx.device_batch(d, i)
    st = ctx.device_state(len(y["off"]) - 1)
    state0 = (s.membership, s.vertex_weight, s.community_weight)
    try:
        ctx.set_graphs(False)
        eager = _runs(ctx, dg, db, st, state0, 1)[0]
        ctx.set_graphs(True)
        for k, got in enumerate(_runs(ctx, dg, db, st, state0, 3)):
            _say(f"  run {k}: Q={got[5]:.6f} eager {eager[5]:.6f}")
            assert abs(got[5] - eager[5]) <= 1e-3
            assert np.array_equal(got[1], eager[1])  # K
    finally:
        ctx.set_graphs(True)
