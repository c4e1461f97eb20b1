"""HIP path vs the oracle, through the C ABI (liblouvain_hip.so).

Bar: integer / flag / index work bit-exact; K and Sigma_tot within 1e-6
relative (exact for integer weights); modularity of a final partition within
1e-3 of the oracle's on the same inputs (the reference itself is only
reproducible at one thread, so trajectories are compared through modularity).
"""
import numpy as np
import pytest

import fixtures as F
from conftest import FIXTURES

pytestmark = pytest.mark.gpu

Q_TOL = 1e-3
INTEGER_FIXTURES = [n for n in FIXTURES if n != "ppw"]


@pytest.fixture(scope="module")
def ctx():
    import lvhip
    c = lvhip.Context(0)
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


def _g(f):
    return F.graph(f, "x"), F.graph(f, "y"), F.batch(f, "del"), F.batch(f, "ins")


def _exact(f):
    return f is not None and True


@pytest.mark.parametrize("name", FIXTURES)
def test_update_weights_and_frontier(ctx, name):
    _say("case", name)
    f = F.load(name)
    x, y, d, i = _g(f)
    S = len(y["off"]) - 1
    vt, ct = ctx.update_weights(S, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"])
    if name in INTEGER_FIXTURES:
        assert np.array_equal(vt, f["upd_vtot"]) and np.array_equal(ct, f["upd_ctot"])
    else:
        assert np.allclose(vt, f["upd_vtot"], rtol=1e-6, atol=1e-9)
        assert np.allclose(ct, f["upd_ctot"], rtol=1e-6, atol=1e-9)
    assert np.array_equal(ctx.affected_frontier(S, d, i, f["static_memb"]), f["frontier"])


@pytest.mark.parametrize("name", FIXTURES)
def test_delta_screening_flags(ctx, name):
    _say("case", name)
    f = F.load(name)
    x, y, d, i = _g(f)
    M = float(f["y_edge_weight"]) / 2
    R = float(f["opts"][1])
    a, b, c = ctx.affected_delta_screening(y, d, i, f["static_memb"], f["upd_vtot"], f["upd_ctot"], M, R)
    assert np.array_equal(a, f["ds_vertices"])
    assert np.array_equal(b, f["ds_neighbors"])
    assert np.array_equal(c, f["ds_communities"])


@pytest.mark.parametrize("name", FIXTURES)
def test_choose_community_bit_exact(ctx, name):
    """The local-moving kernel's decision on a fixed snapshot: the same
    community and the same delta-Q bits as the reference."""
    _say("case", name)
    f = F.load(name)
    x, y, d, i = _g(f)
    R = float(f["opts"][1])
    M = float(f["y_edge_weight"]) / 2
    bc, be = ctx.choose_community(y, f["static_memb"], f["upd_vtot"], f["upd_ctot"], M, R)
    if name in INTEGER_FIXTURES:
        assert np.array_equal(bc, f["choose_y_c"]) and np.array_equal(be, f["choose_y_e"])
    else:
        assert np.allclose(be, f["choose_y_e"], rtol=1e-9, atol=1e-15)
        assert (bc == f["choose_y_c"]).mean() > 0.99
    Mx = float(f["x_edge_weight"]) / 2
    S = len(x["off"]) - 1
    k = ctx.vertex_weights(x)
    bc, be = ctx.choose_community(x, np.arange(S, dtype=np.uint32), k, k, Mx, R)
    if name in INTEGER_FIXTURES:
        assert np.array_equal(bc, f["choose_x_c"]) and np.array_equal(be, f["choose_x_e"])
    sg = F.super_graph(f)
    CN = int(f["renum_CN"])
    Msg = f["sg_K"].sum() / 2
    bc, be = ctx.choose_community(sg, np.arange(CN, dtype=np.uint32), f["sg_K"], f["sg_K"], Msg, R)
    if name in INTEGER_FIXTURES:
        assert np.array_equal(bc, f["sg_choose_c"]) and np.array_equal(be, f["sg_choose_e"])


def _same_row_order(a, f, prefix):
    """Each super-graph row lists its communities in the reference's order:
    first seen over the members in ascending id order and their edges
    (inc/louvain.hxx:800-812, :897-912 at one thread)."""
    off, deg = f[prefix + "off"], f[prefix + "deg"]
    for c in range(len(deg)):
        o, d = int(off[c]), int(deg[c])
        assert np.array_equal(a["keys"][o:o + d], f[prefix + "keys"][o:o + d]), c


@pytest.mark.parametrize("name", FIXTURES)
def test_renumber_and_aggregate(ctx, name):
    _say("case", name)
    f = F.load(name)
    x = F.graph(f, "x")
    vc, cext, CN = ctx.renumber(x, f["static_memb"])
    assert CN == int(f["renum_CN"])
    assert np.array_equal(vc, f["renum_vcom"]) and np.array_equal(cext, f["renum_cext"])
    a = ctx.aggregate(x, vc, CN)
    assert np.array_equal(a["off"], f["agg_off"]) and np.array_equal(a["deg"], f["agg_deg"])
    _same_row_order(a, f, "agg_")
    got = F.rows_as_maps(a["off"], a["deg"], a["keys"], a["w"])
    want = F.rows_as_maps(f["agg_off"], f["agg_deg"], f["agg_keys"], f["agg_w"])
    if name in INTEGER_FIXTURES:
        assert got == want
    else:
        for g_, w_ in zip(got, want):
            assert g_.keys() == w_.keys()
            assert np.allclose([g_[k] for k in w_], [w_[k] for k in w_], rtol=1e-12)
    # second level: f64 super-graph input
    sg = F.super_graph(f)
    vc2, _, CN2 = ctx.renumber(sg, f["sg_memb"])
    assert CN2 == int(f["sg_renum_CN"]) and np.array_equal(vc2, f["sg_renum_vcom"])
    a2 = ctx.aggregate(sg, vc2, CN2)
    assert np.array_equal(a2["off"], f["sg_agg_off"]) and np.array_equal(a2["deg"], f["sg_agg_deg"])
    _same_row_order(a2, f, "sg_agg_")
    got = F.rows_as_maps(a2["off"], a2["deg"], a2["keys"], a2["w"])
    want = F.rows_as_maps(f["sg_agg_off"], f["sg_agg_deg"], f["sg_agg_keys"], f["sg_agg_w"])
    for g_, w_ in zip(got, want):
        assert g_.keys() == w_.keys()
        assert np.allclose([g_[k] for k in w_], [w_[k] for k in w_], rtol=1e-12)


@pytest.mark.parametrize("name", FIXTURES)
def test_weights_and_modularity(ctx, oracle, name):
    _say("case", name)
    f = F.load(name)
    x = F.graph(f, "x")
    k = ctx.vertex_weights(x)
    ex = x.get("ex")
    live = np.ones(len(k), bool) if ex is None else ex.astype(bool)
    if name in INTEGER_FIXTURES:
        assert np.array_equal(k[live], f["static_vtot"][live])
        assert ctx.edge_weight(x) == f["x_edge_weight"]
    else:
        assert np.allclose(k[live], f["static_vtot"][live], rtol=1e-12)
    q = ctx.modularity(x, f["static_memb"])
    assert abs(q - float(f["static_Q"])) < 1e-12


@pytest.mark.parametrize("path", [0, 1, 2])
@pytest.mark.parametrize("name", FIXTURES)
def test_dynamic_end_to_end(ctx, oracle, name, path):
    """ND / DS / DF from the reference's static state: K exact, modularity
    within 1e-3 of the reference's result on the same inputs."""
    _say("case", name, "path", path)
    f = F.load(name)
    x, y, d, i = _g(f)
    ctx.set_path(path)
    import lvhip
    o = lvhip.LouvainOptions(*[t(v) for t, v in zip((int, float, float, float, float, int, int), f["opts"])])
    try:
        for approach in ("naive-dynamic", "delta-screening", "dynamic-frontier"):
            key = approach.replace("-", "_")
            r = ctx.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], o)
            q = oracle.modularity(y, r.membership)
            two = approach == "dynamic-frontier"
            q_ref = float(f[key + "_Q"]) if two else _reference_band(oracle, approach, y, d, i, f, o)
            _say(f"  {approach}: Q gpu={q:.6f} ref={q_ref:.6f} (1-thread {float(f[key + '_Q']):.6f}) "
                 f"it={r.iterations} p={r.passes}")
            _check_q(q, q_ref, two_sided=two)
            if name in INTEGER_FIXTURES:
                assert np.array_equal(r.vertexWeight, f[key + "_vtot"])
            else:
                assert np.allclose(r.vertexWeight, f[key + "_vtot"], rtol=1e-6, atol=1e-9)
            _check_state(y, r)
    finally:
        ctx.set_path(2)


def _reference_band(oracle, approach, y, d, i, f, o):
    """Lowest modularity the reference's OpenMP algorithm reaches on these
    inputs at 1, 2, 4 and 8 threads (the oracle restates it; at one thread it
    is bit-exact with the compiled reference)."""
    from oracle import Options
    oo = Options(o.repeat, o.resolution, o.tolerance, o.aggregationTolerance, o.toleranceDrop, o.maxIterations,
                 o.maxPasses)
    qs = []
    try:
        for t in (1, 2, 4, 8):
            oracle.set_threads(t)
            r = oracle.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], oo)
            qs.append(oracle.modularity(y, r.membership))
    finally:
        oracle.set_threads(1)
    return min(qs)


def _check_q(q_gpu, q_ref, two_sided):
    """DF (one pass over a frontier from a given state; ascending ids, a
    vertex flagged by a lower one visited in the same iteration, as in the
    reference) must match the reference within 1e-3 both ways.  ND and DS
    re-optimise the whole graph, where the reference's own result moves by
    more than 1e-3 with its thread count (Static on a 100k planted
    partition: Q 0.518 at one thread, 0.574 at eight); for those the bar is
    "not worse than the lowest of the reference's 1/2/4/8-thread results by
    more than 1e-3"."""
    if two_sided:
        assert abs(q_gpu - q_ref) <= Q_TOL, (q_gpu, q_ref)
    else:
        assert q_gpu >= q_ref - Q_TOL, (q_gpu, q_ref)


def _check_state(g, r):
    """Sigma_tot at live ids sums to 2M; for one-pass results every live
    community's Sigma_tot equals the sum of its members' K."""
    S = len(g["off"]) - 1
    ex = g.get("ex")
    live = np.ones(S, bool) if ex is None else ex.astype(bool)
    m = r.membership[live]
    assert (m < S).all()
    m2 = float(g["w"].astype(np.float64).sum())
    ids = np.unique(m)
    assert np.isclose(r.communityWeight[ids].sum(), m2, rtol=1e-9)
    if r.passes == 1:
        tot = np.bincount(m, weights=r.vertexWeight[live], minlength=S)
        assert np.allclose(r.communityWeight[ids], tot[ids], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("name", FIXTURES)
def test_static_end_to_end(ctx, oracle, name):
    _say("case", name)
    f = F.load(name)
    x = F.graph(f, "x")
    import lvhip
    o = lvhip.LouvainOptions(*[t(v) for t, v in zip((int, float, float, float, float, int, int), f["opts"])])
    r = ctx.louvain("static", x, o=o)
    q = oracle.modularity(x, r.membership)
    _say(f"  static: Q gpu={q:.6f} ref={float(f['static_Q']):.6f} it={r.iterations} p={r.passes}")
    # the fixtures have at most 2048 ids, one reference chunk: each level runs
    # as one stream, so the result is the reference's at one thread
    assert q >= float(f["static_Q"]) - Q_TOL
    assert np.array_equal(r.vertexWeight, f["static_vtot"]) or name not in INTEGER_FIXTURES
    _check_state(x, r)


def test_unsorted_batch_rejected(ctx):
    import lvhip
    f = F.load("tiny1")
    y = F.graph(f, "y")
    d, i = F.batch(f, "del"), F.batch(f, "ins")
    bad = {k: v[::-1].copy() for k, v in i.items()}
    with pytest.raises(lvhip.LvError) as e:
        ctx.louvain("dynamic-frontier", y, d, bad, f["static_memb"], f["static_vtot"], f["static_ctot"])
    assert e.value.status == -4


@pytest.mark.parametrize("defect", ["unsorted", "repeated", "out_of_range", "offsets"])
def test_bad_graph_rows_rejected(ctx, defect):
    """Input rows must be DiGraph's: neighbour ids < span, strictly ascending,
    inside the edge arrays."""
    import lvhip
    f = F.load("pp2k")
    x = {k: v.copy() for k, v in F.graph(f, "x").items()}
    off, keys = x["off"], x["keys"]
    u = int(np.argmax((np.diff(off) >= 3) & (off[:-1] > 0)))  # a row of at least three entries, not the first
    o = int(off[u])
    if defect == "unsorted":
        keys[o], keys[o + 1] = keys[o + 1], keys[o]
    elif defect == "repeated":
        keys[o + 1] = keys[o]
    elif defect == "offsets":
        off[u + 1] = off[u] - 1
    else:
        keys[o + 2] = len(off) - 1 + 5
    with pytest.raises(lvhip.LvError) as e:
        ctx.louvain("static", x)
    assert e.value.status == (-4 if defect in ("unsorted", "repeated") else -1)


def test_out_of_range_batch_rejected(ctx):
    import lvhip
    f = F.load("tiny1")
    y = F.graph(f, "y")
    d = F.batch(f, "del")
    bad = {"u": np.array([1, 200], np.uint32), "v": np.array([200, 1], np.uint32), "w": np.ones(2, np.float32)}
    with pytest.raises(lvhip.LvError) as e:
        ctx.louvain("dynamic-frontier", y, d, bad, f["static_memb"], f["static_vtot"], f["static_ctot"])
    assert e.value.status == -1


def test_device_batch_checked_against_each_graph(ctx):
    """A device batch is checked once, when it is made; each run compares its
    largest id with the run's graph (out of range: -1) and reports an
    unsorted batch (-4), as the host entry points do."""
    import lvhip
    f = F.load("tiny1")
    y = F.graph(f, "y")
    S = len(y["off"]) - 1
    d, i = F.batch(f, "del"), F.batch(f, "ins")
    dg = ctx.device_graph(y)
    st = ctx.device_state(S)
    st.upload(f["static_memb"], f["static_vtot"], f["static_ctot"])
    far = {"u": np.array([1, S], np.uint32), "v": np.array([S, 1], np.uint32), "w": np.ones(2, np.float32)}
    with pytest.raises(lvhip.LvError) as e:
        ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, ctx.device_batch(d, far), st)
    assert e.value.status == -1
    rev = {k: v[::-1].copy() for k, v in i.items()}
    with pytest.raises(lvhip.LvError) as e:
        ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, ctx.device_batch(d, rev), st)
    assert e.value.status == -4
    ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, ctx.device_batch(d, i), st)  # the good batch still runs


@pytest.mark.parametrize("name", INTEGER_FIXTURES)
def test_single_stream_all_passes_bit_exact(ctx, name):
    """One stream, every pass: super-graph rows in the reference's order
    (members ascending, communities first seen), so each later pass sweeps
    the same rows in the same order as the reference at one thread, and the
    final membership, K, iterations and passes are equal bit for bit, for
    Static, ND, DS and DF."""
    import lvhip
    _say("case", name)
    f = F.load(name)
    x, y, d, i = _g(f)
    o = lvhip.LouvainOptions(*[t(v) for t, v in zip((int, float, float, float, float, int, int), f["opts"])])
    from oracle import Options, Oracle
    oo = Options(o.repeat, o.resolution, o.tolerance, o.aggregationTolerance, o.toleranceDrop, o.maxIterations,
                 o.maxPasses)
    orc = Oracle()
    orc.set_threads(1)
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_hub_threshold(1 << 30, 4096)
    ctx.set_path(0)
    try:
        for approach in ("static", "naive-dynamic", "delta-screening", "dynamic-frontier"):
            if approach == "static":
                r = ctx.louvain(approach, x, o=o)
                ref = orc.louvain(approach, x, options=oo)
            else:
                r = ctx.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], o)
                ref = orc.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], oo)
            _say(f"  {approach}: it={r.iterations}/{ref.iterations} passes={r.passes}/{ref.passes}")
            assert np.array_equal(r.membership, ref.membership), approach
            assert np.array_equal(r.vertexWeight, ref.vertex_weight), approach
            live = np.unique(r.membership)
            assert np.array_equal(r.communityWeight[live], ref.community_weight[live]), approach
            assert (r.iterations, r.passes) == (ref.iterations, ref.passes), approach
    finally:
        ctx.set_schedule()
        ctx.set_hub_threshold()
        ctx.set_path(2)


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("name", INTEGER_FIXTURES)
def test_single_stream_reproduces_reference_bit_exact(ctx, name, path):
    """One stream, ascending ids, every vertex in the id-order kernel: the
    GPU's first local-moving pass is the reference's sequential sweep
    (inc/louvain.hxx:594-605 at one thread), so membership, K, Sigma_tot,
    iterations and affected count equal the golden vectors bit for bit --
    on the plain path (every id checked) and on the frontier path (flags
    read 64 ids at a time, re-read after every visit)."""
    import lvhip
    _say("case", name, "path", path)
    f = F.load(name)
    x, y, d, i = _g(f)
    o = lvhip.LouvainOptions(*[t(v) for t, v in zip((int, float, float, float, float, int, int), f["opts"])])
    o.maxPasses = 1
    from oracle import Options
    oo = Options(o.repeat, o.resolution, o.tolerance, o.aggregationTolerance, o.toleranceDrop, o.maxIterations, 1)
    from oracle import Oracle
    orc = Oracle()
    orc.set_threads(1)
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_hub_threshold(1 << 30, 4096)
    ctx.set_path(path)
    try:
        for approach in ("static", "naive-dynamic", "delta-screening", "dynamic-frontier"):
            if approach == "static":
                r = ctx.louvain(approach, x, o=o)
                ref = orc.louvain(approach, x, options=oo)
            else:
                r = ctx.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], o)
                ref = orc.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], oo)
            _say(f"  {approach}: it={r.iterations}/{ref.iterations} aff={r.affectedVertices}/{ref.affected_vertices}")
            assert np.array_equal(r.membership, ref.membership), approach
            assert np.array_equal(r.vertexWeight, ref.vertex_weight), approach
            assert np.array_equal(r.communityWeight, ref.community_weight), approach
            assert (r.iterations, r.passes, r.affectedVertices) == (ref.iterations, ref.passes,
                                                                      ref.affected_vertices), approach
    finally:
        ctx.set_schedule()
        ctx.set_hub_threshold()
        ctx.set_path(2)


def test_static_planted_partition_100k_mtx(ctx, tmp_path, oracle):
    """The static parity workload: a 100k-vertex planted partition (100
    power-law blocks, average degree 16, 30% inter-block edges, unit f32
    weights, seed 42) written as a 1-indexed MTX and read back; Static with
    tol 1e-2, drop 10, L=20, P=10.  Modularity not worse than the lowest of
    the reference's 1/2/4/8-thread results by more than 1e-3; K exact."""
    import lvgraph
    import lvhip
    from oracle import Options
    g = lvgraph.planted_partition(100000, 100, 16, 0.3, 42, one_indexed=True)
    p = str(tmp_path / "pp100k.mtx")
    lvgraph.write_mtx(g, p)
    x = lvgraph.read_mtx(p)
    assert np.array_equal(x["off"], g["off"]) and np.array_equal(x["keys"], g["keys"])
    r = ctx.louvain("static", x, o=lvhip.LouvainOptions())
    q = oracle.modularity(x, r.membership)
    qs = []
    try:
        for t in (1, 2, 4, 8):
            oracle.set_threads(t)
            ref = oracle.louvain("static", x, options=Options())
            qs.append(oracle.modularity(x, ref.membership))
    finally:
        oracle.set_threads(1)
    _say(f"  static pp100k: Q gpu={q:.6f} ref(1/2/4/8)={['%.6f' % v for v in qs]} it={r.iterations} p={r.passes} "
         f"t={r.time:.1f} ms")
    assert q >= min(qs) - Q_TOL
    assert np.array_equal(r.vertexWeight, ref.vertex_weight)
    _check_state(x, r)


def test_single_stream_hubs_in_the_sweep_bit_exact(ctx, oracle):
    """One stream with the hubs in the sweep (lv_context_set_sweep_hubs):
    every vertex, those of more than 4096 edges through a table of their own,
    is visited in ascending id order, the reference's sequential order, so
    Static through every pass equals the reference at one thread bit for bit
    on an R-MAT graph with such hubs."""
    import lvgraph
    x = lvgraph.rmat(15, 16, seed=5)
    deg = np.diff(x["off"].astype(np.int64))
    assert (deg > 4096).sum() >= 1, int(deg.max())
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_sweep_hubs(True)
    ctx.set_path(0)
    try:
        r = ctx.louvain("static", x)
    finally:
        ctx.set_schedule()
        ctx.set_sweep_hubs(False)
        ctx.set_path(2)
    ref = oracle.louvain("static", x)
    _say(f"  {int((deg > 4096).sum())} hubs above 4096 edges; it={r.iterations}/{ref.iterations} "
         f"passes={r.passes}/{ref.passes}")
    assert np.array_equal(r.membership, ref.membership)
    assert np.array_equal(r.vertexWeight, ref.vertex_weight)
    assert (r.iterations, r.passes) == (ref.iterations, ref.passes)


def test_single_stream_repeat_bit_exact(ctx):
    """repeat = 3 (inc/louvain.hxx:1240-1249: every repeat starts from the
    caller's state; the result is the last repeat's, the times the means):
    at one stream each approach gives the membership, K, Sigma_tot,
    iterations and passes of the reference run with repeat = 3, and those of
    a repeat = 1 run."""
    import lvhip
    from oracle import Options, Oracle
    f = F.load("pp2k")
    x, y, d, i = _g(f)
    orc = Oracle()
    orc.set_threads(1)
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_hub_threshold(1 << 30, 4096)
    ctx.set_path(0)
    try:
        for approach in ("static", "naive-dynamic", "delta-screening", "dynamic-frontier"):
            got = {}
            for rep in (1, 3):
                o = lvhip.LouvainOptions(repeat=rep)
                oo = Options(repeat=rep)
                if approach == "static":
                    r = ctx.louvain(approach, x, o=o)
                    ref = orc.louvain(approach, x, options=oo)
                else:
                    r = ctx.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], o)
                    ref = orc.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], oo)
                assert np.array_equal(r.membership, ref.membership), (approach, rep)
                assert np.array_equal(r.vertexWeight, ref.vertex_weight), (approach, rep)
                live = np.unique(r.membership)
                assert np.array_equal(r.communityWeight[live], ref.community_weight[live]), (approach, rep)
                assert (r.iterations, r.passes) == (ref.iterations, ref.passes), (approach, rep)
                assert r.time > 0
                got[rep] = r
            assert np.array_equal(got[1].membership, got[3].membership), approach
            _say(f"  {approach}: repeat 3 it={got[3].iterations} passes={got[3].passes} "
                 f"time {got[3].time:.3f} ms (repeat 1: {got[1].time:.3f} ms)")
    finally:
        ctx.set_schedule()
        ctx.set_hub_threshold()
        ctx.set_path(2)
