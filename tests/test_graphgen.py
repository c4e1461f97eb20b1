"""Host graph tools (liblvgraph.so) against independent numpy restatements
of the reference's batch semantics (inc/batch.hxx:154-248) and file formats
(inc/mtx.hxx, inc/snap.hxx)."""
import os
import tempfile

import numpy as np

import npgraph as G


def _check_csr(g, symmetric=True, loops_ok=True):
    S = len(g["off"]) - 1
    off = g["off"].astype(np.int64)
    assert off[0] == 0 and (np.diff(off) >= 0).all() and off[-1] == len(g["keys"])
    u = np.repeat(np.arange(S), np.diff(off))
    v = g["keys"].astype(np.int64)
    assert (v < S).all()
    key = u * S + v
    assert (np.diff(key) > 0).all(), "rows must be sorted and unique"
    if symmetric:
        assert np.array_equal(np.sort(key), np.sort(v * S + u))
    if not loops_ok:
        assert (u != v).all()


def test_rmat_structure():
    import lvgraph
    g = lvgraph.rmat(10, 8, seed=3)
    _check_csr(g)
    assert len(g["off"]) == 1025 and "ex" not in g
    g2 = lvgraph.rmat(10, 8, seed=3)
    assert np.array_equal(g["keys"], g2["keys"]), "deterministic for a seed"
    # skew: the R-MAT hub (vertex 0) has the largest degree
    deg = np.diff(g["off"].astype(np.int64))
    assert deg.argmax() == 0


def test_rmat_thread_independent():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import lvgraph, numpy as np; "
            "g = lvgraph.rmat(11, 4, seed=5); print(int(g['keys'].astype(np.int64).sum()), len(g['keys']))")
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "louvain-communities-openmp-dynamic_amd")
    outs = set()
    for t in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code % pkg], env=env, capture_output=True, text=True,
                                check=True).stdout)
    assert len(outs) == 1


def test_planted_partition_structure():
    import lvgraph
    g = lvgraph.planted_partition(5000, 20, 12, 0.3, seed=42, one_indexed=True)
    _check_csr(g, loops_ok=False)
    assert g["ex"][0] == 0 and g["ex"][1:].all()
    assert len(g["off"]) == 5002
    assert abs(len(g["keys"]) / 5000 - 12) < 1.0


def test_batch_tidy_apply_match_numpy():
    import lvgraph
    g = lvgraph.rmat(9, 8, seed=1)
    dels, ins = lvgraph.random_batch(g, 40, 160, seed=9)
    # tidy: deletions exist, insertions do not, sorted unique, both directions
    for b, exist in ((dels, True), (ins, False)):
        assert (G.has_edge(g, b["u"], b["v"]) == exist).all()
        key = b["u"].astype(np.int64) * 512 + b["v"]
        assert (np.diff(key) > 0).all()
        rev = b["v"].astype(np.int64) * 512 + b["u"]
        assert np.array_equal(np.sort(key), np.sort(rev))
    assert len(dels["u"]) > 0 and len(ins["u"]) > 0
    # re-tidying through the numpy restatement changes nothing
    t = G.tidy(g, dels["u"], dels["v"], dels["w"], True)
    assert all(np.array_equal(t[k], dels[k]) for k in "uvw")
    t = G.tidy(g, ins["u"], ins["v"], ins["w"], False)
    assert all(np.array_equal(t[k], ins[k]) for k in "uvw")
    y = lvgraph.apply_batch(g, dels, ins)
    y2 = G.apply_batch(g, dels, ins)
    assert np.array_equal(y["off"], y2["off"]) and np.array_equal(y["keys"], y2["keys"])
    assert np.array_equal(y["w"], y2["w"])


def test_tidy_batch_filters_sorts_dedups():
    import lvgraph
    g = G.symmetric_csr(6, np.array([0, 1, 2]), np.array([1, 2, 3]))
    dels = {"u": np.array([2, 0, 0, 4], np.uint32), "v": np.array([1, 1, 1, 5], np.uint32),
            "w": np.ones(4, np.float32)}
    ins = {"u": np.array([5, 0, 0, 3], np.uint32), "v": np.array([4, 1, 5, 2], np.uint32),
           "w": np.ones(4, np.float32)}
    d, i = lvgraph.tidy_batch(g, dels, ins)
    assert list(zip(d["u"], d["v"])) == [(0, 1), (2, 1)]
    assert list(zip(i["u"], i["v"])) == [(0, 5), (5, 4)]


def test_mtx_roundtrip_one_indexed():
    import lvgraph
    g = lvgraph.planted_partition(300, 5, 8, 0.3, seed=2, one_indexed=True)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.mtx")
        lvgraph.write_mtx(g, p)
        h = lvgraph.read_mtx(p)
    for k in ("off", "keys", "w", "ex"):
        assert np.array_equal(g[k], h[k]), k


def test_mtx_symmetric_and_temporal_readers():
    import lvgraph
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "s.mtx")
        with open(p, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real symmetric\n% comment\n4 4 3\n1 2 1.5\n2 3 2\n4 4 1\n")
        g = lvgraph.read_mtx(p)
        assert len(g["off"]) == 6 and g["ex"][0] == 0
        pairs = set(zip(np.repeat(np.arange(5), np.diff(g["off"].astype(np.int64))).tolist(), g["keys"].tolist()))
        assert pairs == {(1, 2), (2, 1), (2, 3), (3, 2), (4, 4)}
        t = os.path.join(d, "t.txt")
        with open(t, "w") as f:
            f.write("1 2 100\n2 3 101\n3 1 102\n4 5 103\n")
        g = lvgraph.read_temporal(t, 5, 3, True)
        assert len(g["off"]) == 7
        assert g["ex"].tolist() == [0, 1, 1, 1, 1, 1]
        assert int(g["off"][-1]) == 6
