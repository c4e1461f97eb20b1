"""Hub paths of the worklist local-moving (the DF / DS first pass), through
the C ABI: hubs of moderate degree (one block each, table in LDS), large
hubs (multi-block batch, HBM tables capped at 8192 slots) and a hub whose
capped table overflows (redone with a full table in the next sub-round).

Bar, as for every multi-stream run: K exact (unit weights), modularity
within 1e-3 of the oracle's DF from the same state, Sigma_tot consistent
with the membership; the per-tier counters show which hub paths ran.
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


def _check(ctx, oracle, y, d, i, memb, vtot, ctot):
    ctx.stats_reset()
    r = ctx.louvain("dynamic-frontier", y, d, i, memb, vtot, ctot)
    st = ctx.stats()
    o = oracle.louvain("dynamic-frontier", y, d, i, memb, vtot, ctot)
    q, q_ref = oracle.modularity(y, r.membership), oracle.modularity(y, o.membership)
    _say(f"  Q gpu={q:.6f} ref={q_ref:.6f} it={r.iterations} p={r.passes} "
         f"mid hubs {st['vertices'][3]} big hubs {st['vertices'][4]}")
    assert abs(q - q_ref) <= Q_TOL, (q, q_ref)
    assert np.array_equal(r.vertexWeight, o.vertex_weight)
    S = len(y["off"]) - 1
    ids = np.unique(r.membership)
    tot = np.bincount(r.membership, weights=r.vertexWeight, minlength=S)
    assert np.allclose(r.communityWeight[ids], tot[ids], rtol=1e-9, atol=1e-9)
    return st


def test_rmat15_hubs(ctx, oracle):
    """R-MAT scale 15: 576 vertices above degree 256 (one block each), 16
    above 2048 (multi-block batch), one above 4096 (its table capped)."""
    import lvgraph
    x = lvgraph.rmat(15, 16, seed=3)
    s = oracle.louvain("static", x)
    for k in range(3):
        d, i = lvgraph.random_batch(x, 200, 800, 1000 + k)
        y = lvgraph.apply_batch(x, d, i)
        _say("case rmat15 batch", k)
        st = _check(ctx, oracle, y, d, i, s.membership, s.vertex_weight, s.community_weight)
        assert st["vertices"][3] > 0  # mid-degree hubs ran


def test_star_hub_table_overflow(ctx, oracle):
    """A hub whose 20,000 neighbours are all singleton communities: more
    distinct communities than its capped table holds, so the batch redoes it
    with a full table.  Leaves are also joined in pairs so they can move."""
    import lvgraph
    S = 20001
    hub = S - 1  # above the flagged leaves: the reference reaches it in the same sweep
    leaves = np.array([v for v in range(S) if v != hub], np.uint32)
    u = np.concatenate([np.full(len(leaves), hub, np.uint32), leaves])
    v = np.concatenate([leaves, np.full(len(leaves), hub, np.uint32)])
    a, b = leaves[0::2][:5000], leaves[1::2][:5000]
    u = np.concatenate([u, a, b])
    v = np.concatenate([v, b, a])
    x = lvgraph.from_edges(S, u, v)
    memb = np.arange(S, dtype=np.uint32)
    vtot = np.diff(x["off"].astype(np.int64)).astype(np.float64)  # K of x (unit weights)
    # one new leaf pair flags its endpoints; their moves flag the hub
    q, p = int(leaves[-2]), int(leaves[-1])
    ins = {"u": np.array([q, p], np.uint32), "v": np.array([p, q], np.uint32), "w": np.ones(2, np.float32)}
    dels = {"u": np.zeros(0, np.uint32), "v": np.zeros(0, np.uint32), "w": np.zeros(0, np.float32)}
    y = lvgraph.apply_batch(x, dels, ins)
    # the second run starts the hub with a full table (marked by the first)
    for k in (0, 1):
        _say("case star overflow, run", k)
        st = _check(ctx, oracle, y, dels, ins, memb, vtot, vtot.copy())
        assert st["vertices"][4] >= 1  # the hub went through the batch
