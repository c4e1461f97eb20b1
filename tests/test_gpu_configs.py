"""The benchmark configurations at their own sizes, through the C ABI, each
against the oracle (the C restatement of the reference, bit-exact with it at
one thread) run on the same inputs from the same carry-in state:

  config 2  DF on R-MAT 22 (edge factor 16), 5 consecutive batches of 1e-4|E|,
            80/20, seed 7, applied in place on the device
  config 3  R-MAT 24: DF at 1e-4|E|, ND and DS at 1e-5|E|
  config 4  the batched form: 256 planted-partition graphs, one DF update of
            1e-3|E| each, in one lv_run_batched
  config 5  the serving loop: resident R-MAT 16..18 graphs, every request's
            result alone and in a 4-slot mix under the default schedule
  Static    R-MAT 18 and 20 against the reference's 1/2/4/8/16-thread band

Bars: DF |dQ| <= 1e-3 both ways; ND / DS not below the reference by more
than 1e-3 and Static's mean over three runs not below the lowest of the
reference's results by more than 1e-3 (those re-optimise the whole graph,
where the reference's own result moves with its thread count); K exact
(unit weights: the incremental update is exact in f64); Sigma_tot of the live
communities sums to 2M."""
import os

import numpy as np
import pytest

import fixtures as F

pytestmark = pytest.mark.gpu

Q_TOL = 1e-3
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import lvhip
    c = lvhip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    from oracle import Oracle
    o = Oracle()
    o.set_threads(THREADS)
    return o


def _sigma_total(g_edges_w_sum, memb, ctot):
    ids = np.unique(memb)
    assert np.isclose(ctot[ids].sum(), g_edges_w_sum, rtol=1e-9), (ctot[ids].sum(), g_edges_w_sum)


def _static_state(ctx, dg, S):
    import lvhip
    st = ctx.device_state(S)
    ctx.run(lvhip.LV_STATIC, dg, None, st)
    return st


def test_config2_rmat22_df_chain(ctx, orc):
    import lvgraph
    import lvhip
    H = lvgraph.rmat_native(22, 16, 0.57, 0.19, 0.19, 7)
    S = H.span
    B = int(round(1e-4 * (16 << 22)))
    n_ins = int(round(0.8 * B))
    dg = ctx.device_graph(H.view())
    st = _static_state(ctx, dg, S)
    _say(f"R-MAT 22: S={S} E={H.edges}; batches of {B} updates")
    for k in range(5):
        d, i = H.random_batch(B - n_ins, n_ins, 7 * 1000 + k + 1)
        prev = st.download()
        tb = dg.update(d, i)
        r = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, tb, st)
        tb.close()
        q = ctx.dmodularity(dg, st)
        H = H.apply(d, i)
        y = H.view()
        assert dg.edges == H.edges
        ref = orc.louvain("dynamic-frontier", y, d, i, *prev)
        q_ref = orc.modularity(y, ref.membership)
        memb, vt, ct = st.download()
        _say(f"  batch {k}: Q gpu={q:.6f} ref={q_ref:.6f} dQ={q - q_ref:+.2e} it={r.iterations} "
             f"aff={r.affectedVertices} t={r.time:.3f} ms")
        assert abs(q - q_ref) <= Q_TOL, k
        assert np.array_equal(vt, ref.vertex_weight), k
        _sigma_total(float(H.edges), memb, ct)
    dg.close()
    st.close()


def test_config3_rmat24_df_nd_ds(ctx, orc):
    import lvgraph
    import lvhip
    H = lvgraph.rmat_native(24, 16, 0.57, 0.19, 0.19, 24)
    S = H.span
    dg0 = ctx.device_graph(H.view())
    st0 = _static_state(ctx, dg0, S)
    dg0.close()
    s0 = st0.download()
    _say(f"R-MAT 24: S={S} E={H.edges}")
    und = H.edges // 2
    for frac, approaches in ((1e-4, ("dynamic-frontier",)), (1e-5, ("naive-dynamic", "delta-screening"))):
        b = int(round(frac * und))
        n_ins = int(round(0.8 * b))
        d, i = H.random_batch(b - n_ins, n_ins, 24 * 100 + int(-np.log10(frac)))
        Y = H.apply(d, i)
        y = Y.view()
        td, ti = d, i  # random_batch's lists are tidy against the graph before the batch
        dy = ctx.device_graph(y)
        db = ctx.device_batch(td, ti)
        for name in approaches:
            st = ctx.device_state(S)
            st.upload(*s0)
            r = ctx.run(lvhip.APPROACHES[name], dy, db, st)
            q = ctx.dmodularity(dy, st)
            ref = orc.louvain(name, y, td, ti, *s0)
            q_ref = orc.modularity(y, ref.membership)
            memb, vt, ct = st.download()
            _say(f"  {frac:g}|E| {name}: Q gpu={q:.6f} ref={q_ref:.6f} dQ={q - q_ref:+.2e} it={r.iterations} "
                 f"p={r.passes} t={r.time:.2f} ms")
            if name == "dynamic-frontier":
                assert abs(q - q_ref) <= Q_TOL
            else:
                assert q >= q_ref - Q_TOL
            assert np.array_equal(vt, ref.vertex_weight), name
            _sigma_total(float(Y.edges), memb, ct)
            st.close()
        dy.close()
        db.close()
        del Y, y


@pytest.mark.parametrize("scale", [18, 20])
def test_static_rmat_against_reference_band(ctx, scale):
    """Static re-optimises the whole graph from singletons, and concurrent
    moves change its outcome from run to run: the reference's own result
    moves with its thread count (R-MAT 20: 0.0751 at one thread, 0.0759-0.0770
    at 2-16), and the GPU's many streams move it by about +-7e-4 between runs,
    with an occasional run up to 3e-3 lower.  Bar: the mean modularity of
    three GPU runs not below the lowest of the reference's 1/2/4/8/16-thread
    results by more than 1e-3, and no run below it by more than 3e-3; K
    exact.  (R-MAT 22 and 24 are measured by bench.py against the reference
    on the same host.)"""
    import lvgraph
    import lvhip
    from oracle import Oracle
    H = lvgraph.rmat_native(scale, 16, 0.57, 0.19, 0.19, 7)
    x = H.view()
    o = Oracle()
    runs = []
    for _ in range(3):
        r = ctx.louvain("static", x, o=lvhip.LouvainOptions())
        runs.append((o.modularity(x, r.membership), r))
    qs = {}
    for t in sorted({1, 2, 4, 8, THREADS}):
        o.set_threads(t)
        ref = o.louvain("static", x)
        qs[t] = o.modularity(x, ref.membership)
    gq = [q for q, _ in runs]
    _say(f"R-MAT {scale} Static: Q gpu " + " ".join(f"{q:.6f}" for q in gq) + f" (mean {np.mean(gq):.6f}); "
         "reference " + ", ".join(f"{t} threads {v:.6f}" for t, v in qs.items()) +
         f"; gpu it={runs[0][1].iterations} p={runs[0][1].passes} t={runs[0][1].time:.1f} ms")
    low = min(qs.values())
    assert np.mean(gq) >= low - Q_TOL
    assert min(gq) >= low - 3 * Q_TOL
    for _, r in runs:
        assert np.array_equal(r.vertexWeight, ref.vertex_weight)


def test_config4_batched_256_graphs(ctx):
    """256 planted-partition graphs (n log-uniform 1k-200k, average degree
    U[8, 32], inter-block fraction U[0.2, 0.4], seed 256), one DF update of
    1e-3|E| each, all in one lv_run_batched from the reference's Static
    result (main.cxx:89 then :127-135); K exact.  Modularity against the
    reference's 1- and 16-thread results from the same carry-in state
    (fixtures.within_band): no graph below them by more than 1e-3, and at
    most 1% of the graphs above them by more than 1e-3.  (On graphs of a few
    thousand vertices the visit order alone -- a few streams where the
    reference runs one or two threads, and sub-rounds that visit the vertices
    flagged during a sweep after the rest of the sweep rather than in their
    id turn -- has moved single graphs' results up by 1.6e-3 and 2e-3.)  The
    batched Static itself is checked by test_gpu_batched."""
    import lvgraph
    import lvhip
    from oracle import Oracle
    o = Oracle()
    o.set_threads(THREADS)
    rng = np.random.default_rng(256)
    xs, ys, dels, inss = [], [], [], []
    for k in range(256):
        n = int(round(10 ** rng.uniform(3, np.log10(200000))))
        deg = float(rng.uniform(8, 32))
        p_inter = float(rng.uniform(0.2, 0.4))
        g = lvgraph.planted_partition(n, max(2, n // 1000), deg, p_inter, 256 * 1000 + k)
        b = max(1, int(round(1e-3 * len(g["keys"]) / 2)))
        n_ins = int(round(0.8 * b))
        d, i = lvgraph.random_batch(g, b - n_ins, n_ins, 256 * 7919 + k)
        td, ti = lvgraph.tidy_batch(g, d, i)
        xs.append(g)
        ys.append(lvgraph.apply_batch(g, d, i))
        dels.append(td)
        inss.append(ti)
    _, voff = lvhip.union_graphs(xs)
    states0 = []
    for g in xs:
        r = o.louvain("static", g)
        states0.append((r.membership, r.vertex_weight, r.community_weight))
    uy, _ = lvhip.union_graphs(ys)
    dy = ctx.device_graph(uy)
    db = ctx.device_batch(lvhip.union_batch(dels, voff), lvhip.union_batch(inss, voff))
    ds = ctx.device_state(int(voff[-1]))
    ds.upload(*lvhip.union_state(states0, voff))
    res = ctx.run_batched("dynamic-frontier", dy, db, ds, voff)
    _say(f"256 graphs, {int(voff[-1])} vertices, {len(uy['keys'])} directed edges: batched DF {res[0].time:.2f} ms")
    worst = 0.0
    above = []
    for k in range(256):
        band = []
        for t in (1, THREADS):
            o.set_threads(t)
            ref = o.louvain("dynamic-frontier", ys[k], dels[k], inss[k], *states0[k])
            band.append(o.modularity(ys[k], ref.membership))
        q = o.modularity(ys[k], res[k].membership)
        dist = max(min(band) - q, q - max(band), 0.0)
        worst = max(worst, dist)
        assert q >= min(band) - Q_TOL, (k, q, band)
        if not F.within_band(q, band, Q_TOL):
            above.append((k, q, band))
        assert np.array_equal(res[k].vertexWeight, ref.vertex_weight), k
    _say(f"  worst distance from the reference band: {worst:.2e}; above it by more than 1e-3: {above}")
    assert len(above) <= 256 // 100, above
    for h in (dy, db, ds):
        h.close()


def test_config5_serving_alone_equals_mix(orc):
    """Resident R-MAT 16, 17 and 18 graphs (two each) under the default
    schedule -- several streams per run, hub tables, side streams and graph
    replay, four contexts at once in the mix -- both servers starting from
    the same states (the reference's Static result).  Every request (a raw
    batch of 1e-6..1e-3 |E|, 80/20) served alone and in the mix: the states
    after it give modularity within 1e-3 of each other, K equal.  (DF against
    the oracle from the same carry-in state is configs 2 and 3.)"""
    import lvgraph
    import lvserve
    graphs = {}
    for scale in (16, 17, 18):
        for k in range(2):
            graphs[f"rmat{scale}_{k}"] = lvgraph.rmat_native(scale, 16, 0.57, 0.19, 0.19, 500 + scale * 10 + k)
    names = list(graphs)
    states = {}
    for n, g in graphs.items():
        r = orc.louvain("static", g.view())
        states[n] = (r.membership, r.vertex_weight, r.community_weight)
    rng = np.random.default_rng(55)

    def requests():
        reqs, host = [], {n: g for n, g in graphs.items()}
        for k in range(36):
            name = names[k % len(names)]
            g = host[name]
            b = max(4, int(round(10 ** rng.uniform(-6, -3) * (g.edges // 2))))
            n_ins = int(round(0.8 * b))
            d, i = g.random_batch(b - n_ins, n_ins, 9000 + k)
            reqs.append(lvserve.Request(k, name, d, i, 0.0))
            host[name] = g.apply(d, i)
        return reqs

    reqs_a = requests()
    reqs_m = [lvserve.Request(r.rid, r.graph, r.deletions, r.insertions, 0.0) for r in reqs_a]
    alone = lvserve.Server(graphs, slots=1, keep_states=True, modularity=True, states=states)
    mix = lvserve.Server(graphs, slots=4, keep_states=True, modularity=True, states=states)
    try:
        for r in reqs_a:
            alone.run([r], realtime=False)
        mix.run(reqs_m, realtime=False)
        overlap = sum(1 for a in reqs_m for b in reqs_m
                      if a.graph != b.graph and a.start < b.done and b.start < a.done)
        _say(f"36 requests over {len(graphs)} graphs; overlapping request pairs in the mix: {overlap}")
        assert overlap > 0
        worst = 0.0
        for a, m in zip(reqs_a, reqs_m):
            assert a.error is None and m.error is None, (a.error, m.error)
            worst = max(worst, abs(a.modularity - m.modularity))
            assert abs(a.modularity - m.modularity) <= Q_TOL, (a.rid, a.modularity, m.modularity)
            assert np.array_equal(a.state[1], m.state[1]), a.rid
        _say(f"  worst |Q alone - Q mix| {worst:.2e}")
    finally:
        alone.close()
        mix.close()
