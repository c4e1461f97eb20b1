"""Device-side batch tidy and graph update (lv_dgraph_update) against the
host implementation (lvgraph.tidy_batch / apply_batch, inc/batch.hxx:154-236
semantics), through the C ABI.  Bar: bit-exact CSR rows, weights, existence
flags and tidy batches; DF on the updated device graph equals DF on a fresh
upload of the host result (one stream, so both runs are deterministic)."""
import numpy as np
import pytest

import fixtures as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import lvhip
    c = lvhip.Context(0)
    yield c
    c.close()


def _raw_batch(g, rng, n):
    """An untidy batch: existing and absent deletions, absent and existing
    insertions, both directions, repeats (with other weights), shuffled."""
    off, keys = np.asarray(g["off"], np.int64), np.asarray(g["keys"], np.uint32)
    S, E = len(off) - 1, int(off[-1])
    src = (np.searchsorted(off, np.arange(E), side="right") - 1).astype(np.uint32)
    pick = rng.integers(0, E, n)
    du = np.concatenate([src[pick], keys[pick], rng.integers(0, S, n // 4), src[pick[: n // 8]]])
    dv = np.concatenate([keys[pick], src[pick], rng.integers(0, S, n // 4), keys[pick[: n // 8]]])
    iu0, iv0 = rng.integers(0, S, n), rng.integers(0, S, n)
    iu = np.concatenate([iu0, iv0, src[pick[: n // 4]], iu0[: n // 8]])
    iv = np.concatenate([iv0, iu0, keys[pick[: n // 4]], iv0[: n // 8]])
    pd, pi = rng.permutation(len(du)), rng.permutation(len(iu))
    dels = {"u": du[pd].astype(np.uint32), "v": dv[pd].astype(np.uint32),
            "w": rng.random(len(du)).astype(np.float32)}
    ins = {"u": iu[pi].astype(np.uint32), "v": iv[pi].astype(np.uint32),
           "w": (rng.random(len(iu)) + 0.5).astype(np.float32)}
    return dels, ins


def _live(d):
    """A downloaded gapped CSR as a packed one."""
    off, deg = d["off"].astype(np.int64), d["deg"].astype(np.int64)
    idx = np.concatenate([np.arange(o, o + k) for o, k in zip(off[:-1], deg)]) if deg.sum() else np.zeros(0, np.int64)
    return {"off": np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64), "keys": d["keys"][idx],
            "w": d["w"][idx], "ex": d["ex"]}


def _same_batch(a, b):
    for k in ("u", "v", "w"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["pp2k", "rmat10", "tiny1"])
def test_device_update_matches_host(ctx, oracle, name):
    import lvgraph
    print("case", name, flush=True)
    f = F.load(name)
    host = F.graph(f, "x")
    host = {"off": host["off"], "keys": host["keys"], "w": host["w"], **({"ex": host["ex"]} if "ex" in host else {})}
    S = len(host["off"]) - 1
    dg = ctx.device_graph(host)
    rng = np.random.default_rng(11)
    # the first update outgrows the packed upload (a full rebuild with
    # slack); the small ones then fit in place; the large one outgrows again
    for k, n in enumerate([max(S // 20, 4), max(S // 50, 2), max(S // 50, 2), S, max(S // 50, 2)]):
        dels, ins = _raw_batch(host, rng, n)
        tb = dg.update(dels, ins)
        want_d, want_i = lvgraph.tidy_batch(host, dels, ins)
        got_d, got_i = tb.download()
        _same_batch(got_d, want_d)
        _same_batch(got_i, want_i)
        host = lvgraph.apply_batch(host, dels, ins)
        got = _live(dg.download())
        assert np.array_equal(got["off"], np.asarray(host["off"], np.uint64)), k
        assert np.array_equal(got["keys"], host["keys"]), k
        assert np.array_equal(got["w"], host["w"]), k
        assert np.array_equal(got["ex"], host.get("ex", np.ones(S, np.uint8))), k
        assert dg.edges == len(host["keys"])
        # 2M through the device modularity of the singleton partition
        st = ctx.device_state(S)
        memb = np.arange(S, dtype=np.uint32)
        st.upload(memb, np.zeros(S), np.zeros(S))
        q = ctx.dmodularity(dg, st)
        assert abs(q - oracle.modularity(host, memb)) < 1e-9, k


def test_dynamic_frontier_on_updated_graph(ctx, oracle):
    """DF from the reference's static state on the device-updated graph is
    bit-identical to DF on the host-updated graph uploaded afresh."""
    import lvgraph
    import lvhip
    f = F.load("rmat10")
    x = F.graph(f, "x")
    x = {"off": x["off"], "keys": x["keys"], "w": x["w"], **({"ex": x["ex"]} if "ex" in x else {})}
    S = len(x["off"]) - 1
    rng = np.random.default_rng(5)
    dels, ins = _raw_batch(x, rng, 60)
    dg = ctx.device_graph(x)
    tb = dg.update(dels, ins)
    y = lvgraph.apply_batch(x, dels, ins)
    td, ti = lvgraph.tidy_batch(x, dels, ins)
    fresh, fb = ctx.device_graph(y), ctx.device_batch(td, ti)
    ctx.set_schedule(1e-12, 1, 1, 16, 2)
    ctx.set_hub_threshold(1 << 30, 4096)
    try:
        out = []
        for g, b in ((dg, tb), (fresh, fb)):
            st = ctx.device_state(S)
            st.upload(f["static_memb"], f["static_vtot"], f["static_ctot"])
            r = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, g, b, st)
            out.append((st.download(), r.iterations, r.passes))
        (m1, v1, c1), it1, p1 = out[0]
        (m2, v2, c2), it2, p2 = out[1]
        assert np.array_equal(m1, m2) and np.array_equal(v1, v2) and np.array_equal(c1, c2)
        assert (it1, p1) == (it2, p2)
    finally:
        ctx.set_schedule()
        ctx.set_hub_threshold()


def test_update_rejects_ids_beyond_span(ctx):
    import lvhip
    f = F.load("tiny1")
    x = F.graph(f, "x")
    S = len(x["off"]) - 1
    dg = ctx.device_graph(x)
    ins = {"u": np.array([1, S + 3], np.uint32), "v": np.array([S + 3, 1], np.uint32), "w": np.ones(2, np.float32)}
    with pytest.raises(lvhip.LvError) as e:
        dg.update(None, ins)
    assert e.value.status == -1


def test_device_update_matches_reference_vectors(ctx):
    """lv_dgraph_update against the reference's own tidyBatchUpdateU +
    applyBatchUpdateOmpU (tests/golden/update_raw.npz; the weight rule for
    repeated entries is in update_golden.py), on the packed upload and again
    on its rebuilt copy."""
    import update_golden as U
    for name, g, (rd, ri), (td, ti), y in U.load():
        print("case", name, flush=True)
        dg = ctx.device_graph(g)
        tb = dg.update(rd, ri)
        got_d, got_i = tb.download()
        rep = U.check_tidy(got_d, td, rd, f"{name} deletions") + U.check_tidy(got_i, ti, ri, f"{name} insertions")
        U.check_graph(_live(dg.download()), y, ri, name)
        print(f"  {len(td['u'])} deletions, {len(ti['u'])} insertions, {rep} repeated pairs", flush=True)
        tb.close()
        dg.close()
