"""lv_dmodularity_noalloc, the entry point that allocates nothing, against
lv_dmodularity (the same kernels, which add with FP atomics, so within 1e-12
relative) through the C ABI: two calls in a row, the second on a smaller
graph, then the first graph again (the shared scratch keeps the last call's
contents, and each call clears what it uses); the context's workspace does
not grow; a span above the limit is refused with LV_ERR_LIMIT."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 21


def _close(a, b):
    return abs(a - b) <= 1e-12 * max(1.0, abs(b))


def _free_device_bytes(ctx):
    # hipMemGetInfo through the HIP runtime the library already loaded
    free, total = C.c_size_t(0), C.c_size_t(0)
    rt = C.CDLL("libamdhip64.so")
    assert rt.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_modularity_matches_and_allocates_nothing(oracle):
    import fixtures as F
    import lvgraph
    import lvhip
    ctx = lvhip.Context(0)
    try:
        big = lvgraph.rmat(17, 16, seed=3)
        f = F.load("pp2k")
        small = F.graph(f, "x")
        cases = []
        for name, g in (("rmat17", big), ("pp2k", small)):
            s0 = oracle.louvain("static", g)
            dg = ctx.device_graph(g)
            st = ctx.device_state(len(g["off"]) - 1)
            st.upload(s0.membership, s0.vertex_weight, s0.community_weight)
            cases.append((name, dg, st, oracle.modularity(g, s0.membership)))
        for name, dg, st, _ in cases:  # warm the existing entry point's workspace
            ctx.dmodularity(dg, st)
        ws0 = ctx.workspace_bytes()
        free0 = _free_device_bytes(ctx)
        for name, dg, st, q_ref in cases + cases[:1]:  # big, smaller, big again
            for R in (1.0, 0.5):
                q = ctx.dmodularity_noalloc(dg, st, R)
                q_alloc = ctx.dmodularity(dg, st, R)
                print(f"{name} R={R}: noalloc {q!r} lv_dmodularity {q_alloc!r}", flush=True)
                assert _close(q, q_alloc), (name, R, q, q_alloc)
                if R == 1.0:
                    assert abs(q - q_ref) <= 1e-9, (name, q, q_ref)
        assert ctx.workspace_bytes() == ws0
        assert _free_device_bytes(ctx) >= free0
        for _, dg, st, _ in cases:
            dg.close()
            st.close()
    finally:
        ctx.close()


def test_noalloc_modularity_refuses_a_span_above_its_limit():
    import lvgraph
    import lvhip
    ctx = lvhip.Context(0)
    try:
        S = LIMIT + 1
        u = np.array([0, 1, LIMIT - 1, LIMIT], np.uint32)
        v = np.array([1, 0, LIMIT, LIMIT - 1], np.uint32)
        g = lvgraph.from_edges(S, u, v)
        dg = ctx.device_graph(g)
        st = ctx.device_state(S)
        memb = np.arange(S, dtype=np.uint32)
        st.upload(memb, np.zeros(S), np.zeros(S))
        with pytest.raises(lvhip.LvError) as e:
            ctx.dmodularity_noalloc(dg, st)
        assert e.value.status == -6  # LV_ERR_LIMIT
        print(str(e.value), flush=True)
        assert np.isfinite(ctx.dmodularity(dg, st))  # the allocating entry point still serves it
        dg.close()
        st.close()
    finally:
        ctx.close()
