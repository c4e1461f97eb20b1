"""The oracle (oracle/louvain_oracle.c) pinned against the reference.

1. Bit-exact, at one OpenMP thread, against the golden vectors the compiled
   reference wrote (tests/golden/make_golden.py): every hot-path function.
2. Closed forms and invariants that do not depend on the reference:
   K = row sums, sum(Sigma_tot) = 2M, renumbering = order-preserving
   compaction, super-graph weight = 2M, modularity against an independent
   numpy formula and a brute-force optimum on the tiny graphs.
3. Mutants: the same checks must FAIL for an oracle built with a plausible
   mistake planted (tie-break, self-loop, community-0 rule).
"""
import itertools
import os
import subprocess
import tempfile

import numpy as np
import pytest

import fixtures as F
from conftest import FIXTURES, ROOT


def _graphs(f):
    return F.graph(f, "x"), F.graph(f, "y"), F.batch(f, "del"), F.batch(f, "ins")


@pytest.mark.parametrize("name", FIXTURES)
def test_static_matches_reference(oracle, name):
    f = F.load(name)
    x, y, d, i = _graphs(f)
    r = oracle.louvain("static", x, options=F.options(f))
    assert np.array_equal(r.membership, f["static_memb"])
    assert np.array_equal(r.vertex_weight, f["static_vtot"])
    assert np.array_equal(r.community_weight, f["static_ctot"])
    assert [r.iterations, r.passes, r.affected_vertices] == f["static_meta"].tolist()
    assert oracle.modularity(x, r.membership) == f["static_Q"]


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("approach", ["naive-dynamic", "delta-screening", "dynamic-frontier"])
def test_dynamic_matches_reference(oracle, name, approach):
    f = F.load(name)
    x, y, d, i = _graphs(f)
    r = oracle.louvain(approach, y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"], F.options(f))
    key = approach.replace("-", "_")
    assert np.array_equal(r.membership, f[f"{key}_memb"])
    assert np.array_equal(r.vertex_weight, f[f"{key}_vtot"])
    assert np.array_equal(r.community_weight, f[f"{key}_ctot"])
    assert [r.iterations, r.passes, r.affected_vertices] == f[f"{key}_meta"].tolist()


@pytest.mark.parametrize("name", FIXTURES)
def test_phases_match_reference(oracle, name):
    f = F.load(name)
    x, y, d, i = _graphs(f)
    R = float(f["opts"][1])
    vt, ct = oracle.update_weights(d, i, f["static_memb"], f["static_vtot"], f["static_ctot"])
    assert np.array_equal(vt, f["upd_vtot"]) and np.array_equal(ct, f["upd_ctot"])
    assert np.array_equal(oracle.frontier(len(y["off"]) - 1, d, i, f["static_memb"]), f["frontier"])
    M = oracle.edge_weight(y) / 2
    assert 2 * M == f["y_edge_weight"]
    dv, dn, dc = oracle.delta_screening(y, d, i, f["static_memb"], vt, ct, M, R)
    assert np.array_equal(dv, f["ds_vertices"])
    assert np.array_equal(dn, f["ds_neighbors"])
    assert np.array_equal(dc, f["ds_communities"])
    bc, be = oracle.choose(y, f["static_memb"], vt, ct, M, R)
    assert np.array_equal(bc, f["choose_y_c"]) and np.array_equal(be, f["choose_y_e"])
    S = len(x["off"]) - 1
    Mx = oracle.edge_weight(x) / 2
    ids = np.arange(S, dtype=np.uint32)
    bc, be = oracle.choose(x, ids, f["static_vtot"] * 0 + _row_sums(x), _row_sums(x), Mx, R)
    assert np.array_equal(bc, f["choose_x_c"]) and np.array_equal(be, f["choose_x_e"])


def _row_sums(g):
    S = len(g["off"]) - 1
    k = np.zeros(S)
    off = g["off"].astype(np.int64)
    deg = np.diff(off) if g.get("deg") is None else g["deg"].astype(np.int64)
    for u in range(S):
        k[u] = g["w"][off[u]:off[u] + deg[u]].astype(np.float64).sum()
    return k


@pytest.mark.parametrize("name", FIXTURES)
def test_renumber_aggregate_match_reference(oracle, name):
    f = F.load(name)
    x = F.graph(f, "x")
    vc, cext, CN = oracle.renumber(x, f["static_memb"])
    assert CN == f["renum_CN"]
    assert np.array_equal(vc, f["renum_vcom"]) and np.array_equal(cext, f["renum_cext"])
    a = oracle.aggregate(x, vc, CN)
    for k in ("off", "deg", "coff", "cedg"):
        assert np.array_equal(a[k], f[f"agg_{k}"]), k
    # one thread: row order (first-seen) is deterministic, so keys/weights match as laid out
    for c in range(CN):
        o, n = int(a["off"][c]), int(a["deg"][c])
        assert np.array_equal(a["keys"][o:o + n], f["agg_keys"][o:o + n])
        assert np.array_equal(a["w"][o:o + n], f["agg_w"][o:o + n])
    # second level: super-graph input with f64 weights
    sg = F.super_graph(f)
    vc2, _, CN2 = oracle.renumber(sg, f["sg_memb"])
    assert CN2 == f["sg_renum_CN"] and np.array_equal(vc2, f["sg_renum_vcom"])
    a2 = oracle.aggregate(sg, vc2, CN2)
    for k in ("off", "deg", "coff", "cedg"):
        assert np.array_equal(a2[k], f[f"sg_agg_{k}"]), k
    assert F.rows_as_maps(a2["off"], a2["deg"], a2["keys"], a2["w"]) == \
        F.rows_as_maps(f["sg_agg_off"], f["sg_agg_deg"], f["sg_agg_keys"], f["sg_agg_w"])
    Msg = f["sg_K"].sum() / 2
    bc, be = oracle.choose(sg, np.arange(CN, dtype=np.uint32), f["sg_K"], f["sg_K"], Msg, float(f["opts"][1]))
    assert np.array_equal(bc, f["sg_choose_c"]) and np.array_equal(be, f["sg_choose_e"])


# ---- closed forms and invariants (independent of the reference) ----

def _np_modularity(g, memb, R=1.0):
    S = len(g["off"]) - 1
    ex = g.get("ex")
    off = g["off"].astype(np.int64)
    u = np.repeat(np.arange(S), np.diff(off))
    if ex is not None:
        keep = ex[u] != 0
    else:
        keep = np.ones(len(u), bool)
    u, v, w = u[keep], g["keys"].astype(np.int64)[keep], g["w"].astype(np.float64)[keep]
    m2 = w.sum()
    cu, cv = memb[u], memb[v]
    cin = np.bincount(cu[cu == cv], weights=w[cu == cv], minlength=S)
    ctot = np.bincount(cu, weights=w, minlength=S)
    return float((cin / m2 - R * (ctot / m2) ** 2).sum())


@pytest.mark.parametrize("name", FIXTURES)
def test_invariants(oracle, name):
    f = F.load(name)
    x, y, d, i = _graphs(f)
    r = oracle.louvain("static", x, options=F.options(f))
    k = _row_sums(x)
    ex = x.get("ex")
    live = np.ones(len(k), bool) if ex is None else ex.astype(bool)
    assert np.allclose(r.vertex_weight[live], k[live], rtol=0, atol=1e-9)
    M2 = oracle.edge_weight(x)
    assert np.isclose(k[live].sum(), M2)
    # Sigma_tot at the live final community ids sums to 2M
    ids = np.unique(r.membership[live])
    assert np.isclose(r.community_weight[ids].sum(), M2, rtol=1e-12)
    # each live community's Sigma_tot is the sum of its members' K (passes==1:
    # ids index first-level communities directly)
    # modularity against an independent formula
    assert np.isclose(oracle.modularity(x, r.membership), _np_modularity(x, r.membership), atol=1e-12)
    # renumbering = order-preserving compaction of the used ids
    vc, cext, CN = oracle.renumber(x, r.membership)
    used = np.unique(r.membership[live])
    assert CN == len(used)
    assert np.array_equal(vc[live], np.searchsorted(used, r.membership[live]))
    # super-graph: total weight is 2M, offsets are scans of community degree sums
    a = oracle.aggregate(x, vc, CN)
    assert np.isclose(a["w"][:int(a["off"][-1])][_live_slots(a)].sum(), M2)
    degsum = np.bincount(vc[live], weights=np.diff(x["off"].astype(np.int64))[live], minlength=CN)
    assert np.array_equal(np.diff(a["off"].astype(np.int64)), degsum.astype(np.int64))
    assert (a["deg"].astype(np.int64) <= degsum).all()
    # K / Sigma_tot after the batch = before +- the batch sums
    vt, ct = oracle.update_weights(d, i, f["static_memb"], f["static_vtot"], f["static_ctot"])
    dk = np.zeros(len(k))
    np.add.at(dk, i["u"].astype(np.int64), i["w"].astype(np.float64))
    np.subtract.at(dk, d["u"].astype(np.int64), d["w"].astype(np.float64))
    assert np.allclose(vt, f["static_vtot"] + dk, atol=1e-9)
    assert np.allclose(vt[live], _row_sums(y)[live], atol=1e-9)
    # DF frontier: the set formula
    memb = f["static_memb"]
    want = np.zeros(len(k), np.uint8)
    want[d["u"][memb[d["u"]] == memb[d["v"]]]] = 1
    want[i["u"][memb[i["u"]] != memb[i["v"]]]] = 1
    assert np.array_equal(oracle.frontier(len(k), d, i, memb), want)


def _live_slots(a):
    m = np.zeros(int(a["off"][-1]), bool)
    for c in range(len(a["deg"])):
        m[int(a["off"][c]):int(a["off"][c]) + int(a["deg"][c])] = True
    return m


def _partitions(n):
    """All set partitions of range(n) as label arrays (restricted growth strings)."""
    def rec(i, labels, k):
        if i == n:
            yield list(labels)
            return
        for c in range(k + 1):
            labels.append(c)
            yield from rec(i + 1, labels, max(k, c + 1))
            labels.pop()
    yield from rec(0, [], 0)


@pytest.mark.parametrize("name", ["tiny1", "tiny0"])
def test_brute_force_bound(oracle, name):
    """Louvain's modularity never exceeds the exhaustive optimum, and reaches
    it on these two-cluster toy graphs."""
    f = F.load(name)
    x = F.graph(f, "x")
    S = len(x["off"]) - 1
    ex = x.get("ex")
    verts = [u for u in range(S) if ex is None or ex[u]]
    best = -1.0
    for lab in _partitions(len(verts)):
        memb = np.zeros(S, np.uint32)
        memb[verts] = lab
        best = max(best, _np_modularity(x, memb))
    q = float(f["static_Q"])
    assert q <= best + 1e-12
    assert q >= best - 0.05


def test_threads_do_not_change_integer_phases():
    """Multi-threaded oracle: thread-independent outputs stay bit-exact."""
    from oracle import Oracle
    o = Oracle()
    o.set_threads(4)
    try:
        f = F.load("pp2k")
        x, y, d, i = _graphs(f)
        vt, ct = o.update_weights(d, i, f["static_memb"], f["static_vtot"], f["static_ctot"])
        assert np.array_equal(vt, f["upd_vtot"]) and np.array_equal(ct, f["upd_ctot"])
        assert np.array_equal(o.frontier(len(y["off"]) - 1, d, i, f["static_memb"]), f["frontier"])
        M = o.edge_weight(y) / 2
        dv, dn, dc = o.delta_screening(y, d, i, f["static_memb"], vt, ct, M)
        assert np.array_equal(dv, f["ds_vertices"]) and np.array_equal(dc, f["ds_communities"])
        vc, cext, CN = o.renumber(x, f["static_memb"])
        assert np.array_equal(vc, f["renum_vcom"])
        a = o.aggregate(x, vc, CN)
        assert np.array_equal(a["off"], f["agg_off"]) and np.array_equal(a["deg"], f["agg_deg"])
        assert F.rows_as_maps(a["off"], a["deg"], a["keys"], a["w"]) == \
            F.rows_as_maps(f["agg_off"], f["agg_deg"], f["agg_keys"], f["agg_w"])
        r = o.louvain("static", x)
        assert abs(o.modularity(x, r.membership) - float(f["static_Q"])) < 0.02
    finally:
        o.set_threads(1)


MUTANTS = ["ORACLE_MUTANT_TIEBREAK", "ORACLE_MUTANT_SELFLOOP", "ORACLE_MUTANT_COMMUNITY0"]


@pytest.mark.parametrize("mutant", MUTANTS)
def test_planted_mistake_is_caught(mutant):
    """Build the oracle with one plausible mistake planted and check that the
    golden comparison rejects it."""
    from oracle import Oracle
    src = os.path.join(ROOT, "oracle", "louvain_oracle.c")
    with tempfile.TemporaryDirectory() as tmp:
        so = os.path.join(tmp, "libmutant.so")
        subprocess.run(["gcc", "-std=c11", "-O1", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
                        f"-D{mutant}=1", "-o", so, src], check=True)
        o = Oracle(so)
        o.set_threads(1)
        caught = False
        for name in FIXTURES:
            f = F.load(name)
            x, y, d, i = _graphs(f)
            r = o.louvain("static", x, options=F.options(f))
            if not np.array_equal(r.membership, f["static_memb"]):
                caught = True
                break
            r = o.louvain("naive-dynamic", y, d, i, f["static_memb"], f["static_vtot"], f["static_ctot"],
                          F.options(f))
            if not np.array_equal(r.membership, f["naive_dynamic_memb"]):
                caught = True
                break
        assert caught, f"{mutant} was not detected by the golden vectors"


@pytest.mark.parametrize("approach", ["static", "naive-dynamic", "delta-screening", "dynamic-frontier"])
def test_repeat_matches_compiled_reference(oracle, approach):
    """repeat = 3 (inc/louvain.hxx:1240-1249) against the reference compiled
    in place (oracle/_ref), one thread: same membership, K, Sigma_tot,
    iterations and passes."""
    import oracle as O
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref is not built (the reference sources are not here)")
    ref = O.RefLib()
    ref.set_threads(1)
    f = F.load("pp2k")
    x, y, d, i = _graphs(f)
    o = O.Options(repeat=3)
    if approach == "static":
        a, b = oracle.louvain(approach, x, options=o), ref.louvain(approach, x, options=o)
    else:
        st = (f["static_memb"], f["static_vtot"], f["static_ctot"])
        a, b = oracle.louvain(approach, y, d, i, *st, o), ref.louvain(approach, y, d, i, *st, o)
    assert np.array_equal(a.membership, b.membership)
    assert np.array_equal(a.vertex_weight, b.vertex_weight)
    assert np.array_equal(a.community_weight, b.community_weight)
    assert (a.iterations, a.passes) == (b.iterations, b.passes)
