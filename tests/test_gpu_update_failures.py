"""A device graph update (lv_dgraph_update) that fails part way leaves the
graph as it was and the context ready for the next update.

The failures are forced from the host: a context workspace limit
(lv_context_set_workspace_limit) that grows step by step, so that each
attempt fails at a later buffer of the tidy and the update, until one
succeeds.  After every failed attempt the device graph must equal the graph
before the batch; the batch that finally goes through, and the unlimited
batches after it, must give the host update's CSR bit for bit (the host
update restates inc/batch.hxx:202-248).  No kernel is made to fail."""
import numpy as np
import pytest

from test_gpu_update import _live, _raw_batch

pytestmark = pytest.mark.gpu


def _same_graph(dg, host, S, what):
    got = _live(dg.download())
    assert np.array_equal(got["off"], np.asarray(host["off"], np.uint64)), what
    assert np.array_equal(got["keys"], host["keys"]), what
    assert np.array_equal(got["w"], host["w"]), what
    assert np.array_equal(got["ex"], host.get("ex", np.ones(S, np.uint8))), what


def _update_under_growing_limits(ctx, dg, host, S, dels, ins, failed_at):
    """Attempt the update under limits from 4 KiB up (x1.25 a step) until it
    succeeds; returns the number of failed attempts."""
    import lvhip
    limit, fails = 4096, 0
    while True:
        ctx.set_workspace_limit(max(limit, ctx.workspace_bytes() // 2))
        try:
            tb = dg.update(dels, ins, ctx=ctx)
            tb.close()
            break
        except lvhip.LvError as e:
            assert e.status == -2, e  # LV_ERR_NOMEM
            msg = str(e)
            at = msg.find("(buffer ")
            failed_at.add(msg[at + 8:msg.find(")", at)] if at >= 0 else msg)
            fails += 1
            _same_graph(dg, host, S, f"after failed attempt {fails}")
        limit = int(limit * 1.25) + 256
        assert fails < 400
    ctx.set_workspace_limit(0)
    return fails


def test_failed_update_leaves_graph_and_context_intact():
    import lvgraph
    import lvhip
    host = lvgraph.rmat(12, 8, seed=21)
    S = len(host["off"]) - 1
    rng = np.random.default_rng(5)
    ctx = lvhip.Context(0)
    try:
        dg = ctx.device_graph(host)
        failed_at = set()
        # first update: the packed upload outgrows its rows (a rebuild with
        # slack); second: in place; then unlimited updates
        for k, n in enumerate([300, 40, 40, 600, 40]):
            dels, ins = _raw_batch(host, rng, n)
            print(f"batch {k}: {len(dels['u'])} raw deletions, {len(ins['u'])} raw insertions", flush=True)
            if k < 2:
                # a fresh context for the limited attempts: every buffer of the
                # tidy and the update is grown under the limit
                lim = lvhip.Context(0)
                try:
                    fails = _update_under_growing_limits(lim, dg, host, S, dels, ins, failed_at)
                finally:
                    lim.close()
                print(f"  {fails} failed attempts before the update went through", flush=True)
                assert fails >= 5
            else:
                dg.update(dels, ins).close()
            host = lvgraph.apply_batch(host, dels, ins)
            _same_graph(dg, host, S, f"after batch {k}")
            assert dg.edges == len(host["keys"])
        print("failed at buffers:", sorted(failed_at), flush=True)
        # failures inside the update itself, after the per-row degree changes
        # were written, are among them
        assert failed_at & {"up.T", "up.so", "up.keep", "up.yk", "up.cap", "up.order"}, failed_at
        dg.close()
    finally:
        ctx.close()


def test_failed_update_then_dynamic_frontier_matches_oracle(oracle):
    """The state and hub tables stay usable: DF after a failed and then a
    successful update equals the oracle's DF from the same carry-in state
    within 1e-3, K exact."""
    import lvgraph
    import lvhip
    host = lvgraph.rmat(13, 16, seed=4)
    S = len(host["off"]) - 1
    s0 = oracle.louvain("static", host)
    ctx = lvhip.Context(0)
    try:
        dg = ctx.device_graph(host)
        st = ctx.device_state(S)
        st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
        rng = np.random.default_rng(9)
        for k in range(4):
            d, i = lvgraph.random_batch(host, 20, 80, 300 + k)
            if k == 2:
                ctx.set_workspace_limit(ctx.workspace_bytes() + 1)  # no buffer may grow
                with pytest.raises(lvhip.LvError):
                    dg.update(*_raw_batch(host, rng, 5000))
                ctx.set_workspace_limit(0)
                _same_graph(dg, host, S, "after the refused batch")
            prev = st.download()
            tb = dg.update(d, i)
            y = lvgraph.apply_batch(host, d, i)
            ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, tb, st)
            q = ctx.dmodularity(dg, st)
            td, ti = lvgraph.tidy_batch(host, d, i)
            o = oracle.louvain("dynamic-frontier", y, td, ti, *prev)
            q_ref = oracle.modularity(y, o.membership)
            t, m = ctx.check_hub_tables(dg, st)
            print(f"run {k}: Q gpu={q:.6f} ref={q_ref:.6f} tables {t} mismatched {m}", flush=True)
            assert abs(q - q_ref) <= 1e-3
            assert np.array_equal(st.download()[1], o.vertex_weight)
            assert m == 0
            host = y
            tb.close()
        dg.close()
        st.close()
    finally:
        ctx.close()
