"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs the reference's own OpenMP Louvain (inc/louvain.hxx), compiled in place
into oracle/_ref/libref.so by ``make -C oracle ref``, with ONE OpenMP thread
(the only deterministic setting: louvain.hxx:594 schedules vertices
dynamically).  Each fixture holds a small graph, one tidy batch, and the
reference's outputs for every function on the hot path:

  static / naive-dynamic / delta-screening / dynamic-frontier end to end
  (louvain.hxx:1285, :1349, :1542, :1674), the K/Sigma_tot update (:374-388),
  the DF frontier (:1606-1624), the DS affected sets (:1441-1486), the
  per-vertex best community on a snapshot (:455-464 via :599-601),
  renumbering (:668-682, :941-945) and aggregation (:800-812, :987-992) on
  the input graph and on its first super-graph, edge weight and modularity
  (properties.hxx:96-106, :178-192).

It also writes update_raw.npz: raw (untidy) batches -- deletions of absent
edges, insertions of present ones, repeated entries with different weights,
both directions, shuffled -- on four of the graphs, with the reference's
tidyBatchUpdateU result (inc/batch.hxx:202-209) and the graph after its
applyBatchUpdateOmpU (inc/batch.hxx:241-248).

Usage:  python tests/golden/make_golden.py [--update-only]   (needs /root/reference)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import npgraph as G  # noqa: E402
from oracle import Options, RefLib  # noqa: E402


def tiny_graphs():
    # 1-indexed: vertex 0 absent, a self-loop on 3, two triangles joined by one edge.
    e1 = [(1, 2), (2, 3), (1, 3), (3, 4), (4, 5), (5, 6), (4, 6), (6, 7), (7, 8), (3, 3)]
    u, v = zip(*e1)
    ex = np.ones(9, np.uint8)
    ex[0] = 0
    g1 = G.symmetric_csr(9, np.array(u), np.array(v), None, ex)
    # 0-indexed: exercises "no move into community 0" (louvain.hxx:602).
    e0 = [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (4, 5), (3, 5), (5, 6), (6, 7), (7, 7), (1, 6)]
    u, v = zip(*e0)
    g0 = G.symmetric_csr(8, np.array(u), np.array(v))
    return g1, g0


def record(ref, g, dels, ins, opts, prefix, out):
    """Every hot-path output on graph g (pre-update x) and its update y."""
    for k, a in g.items():
        out[f"{prefix}x_{k}"] = a
    for k, a in dels.items():
        out[f"{prefix}del_{k}"] = a
    for k, a in ins.items():
        out[f"{prefix}ins_{k}"] = a
    y = G.apply_batch(g, dels, ins)
    for k, a in y.items():
        out[f"{prefix}y_{k}"] = a
    out[f"{prefix}opts"] = np.array([opts.repeat, opts.resolution, opts.tolerance, opts.aggregation_tolerance,
                                     opts.tolerance_drop, opts.max_iterations, opts.max_passes], np.float64)
    out[f"{prefix}x_edge_weight"] = np.float64(ref.edge_weight(g))
    out[f"{prefix}y_edge_weight"] = np.float64(ref.edge_weight(y))
    s = ref.louvain("static", g, options=opts)
    out[f"{prefix}static_memb"] = s.membership
    out[f"{prefix}static_vtot"] = s.vertex_weight
    out[f"{prefix}static_ctot"] = s.community_weight
    out[f"{prefix}static_meta"] = np.array([s.iterations, s.passes, s.affected_vertices], np.int64)
    out[f"{prefix}static_Q"] = np.float64(ref.modularity(g, s.membership))
    for name in ("naive-dynamic", "delta-screening", "dynamic-frontier"):
        r = ref.louvain(name, y, dels, ins, s.membership, s.vertex_weight, s.community_weight, opts)
        key = name.replace("-", "_")
        out[f"{prefix}{key}_memb"] = r.membership
        out[f"{prefix}{key}_vtot"] = r.vertex_weight
        out[f"{prefix}{key}_ctot"] = r.community_weight
        out[f"{prefix}{key}_meta"] = np.array([r.iterations, r.passes, r.affected_vertices], np.int64)
        out[f"{prefix}{key}_Q"] = np.float64(ref.modularity(y, r.membership))
    # phases on y, starting from the static state
    vt, ct = ref.update_weights(y, dels, ins, s.membership, s.vertex_weight, s.community_weight)
    out[f"{prefix}upd_vtot"] = vt
    out[f"{prefix}upd_ctot"] = ct
    out[f"{prefix}frontier"] = ref.frontier(y, dels, ins, s.membership)
    M = ref.edge_weight(y) / 2
    R = opts.resolution
    dv, dn, dc = ref.delta_screening(y, dels, ins, s.membership, vt, ct, M, R)
    out[f"{prefix}ds_vertices"], out[f"{prefix}ds_neighbors"], out[f"{prefix}ds_communities"] = dv, dn, dc
    # best community per vertex on the singleton snapshot of x and on the
    # updated-state snapshot of y
    Mx = ref.edge_weight(g) / 2
    S = len(g["off"]) - 1
    ids = np.arange(S, dtype=np.uint32)
    kx = np.zeros(S)
    u, _, w = G.edges_of(g)
    np.add.at(kx, u, w.astype(np.float64))
    out[f"{prefix}choose_x_c"], out[f"{prefix}choose_x_e"] = ref.choose(g, ids, kx, kx, Mx, R)
    out[f"{prefix}choose_y_c"], out[f"{prefix}choose_y_e"] = ref.choose(y, s.membership, vt, ct, M, R)
    # renumber + aggregate the static membership on x
    vc, cext, CN = ref.renumber(g, s.membership)
    out[f"{prefix}renum_vcom"], out[f"{prefix}renum_cext"], out[f"{prefix}renum_CN"] = vc, cext, np.int64(CN)
    agg = ref.aggregate(g, vc, CN)
    for k, a in agg.items():
        out[f"{prefix}agg_{k}"] = a
    # one level up: the super-graph (f64 weights) with a coarser membership
    sg = {"off": agg["off"], "deg": agg["deg"], "keys": agg["keys"], "w": agg["w"]}
    sg_memb = (np.arange(CN) // 2).astype(np.uint32)  # pair up super-vertices
    vc2, cext2, CN2 = ref.renumber(sg, sg_memb)
    out[f"{prefix}sg_memb"] = sg_memb
    out[f"{prefix}sg_renum_vcom"], out[f"{prefix}sg_renum_CN"] = vc2, np.int64(CN2)
    agg2 = ref.aggregate(sg, vc2, CN2)
    for k, a in agg2.items():
        out[f"{prefix}sg_agg_{k}"] = a
    ksg = np.zeros(CN)
    for c in range(CN):
        o, d = int(agg["off"][c]), int(agg["deg"][c])
        ksg[c] = agg["w"][o:o + d].sum()
    Msg = ksg.sum() / 2
    out[f"{prefix}sg_choose_c"], out[f"{prefix}sg_choose_e"] = ref.choose(
        sg, np.arange(CN, dtype=np.uint32), ksg, ksg, Msg, R)
    out[f"{prefix}sg_K"] = ksg


def raw_batch(g, rng, n):
    """An untidy batch on g: present and absent deletions, absent and present
    insertions, both directions, repeats carrying other weights, shuffled."""
    u, v, _w = G.edges_of(g)
    S, E = len(g["off"]) - 1, len(u)
    pick = rng.integers(0, E, n)
    du = np.concatenate([u[pick], v[pick], rng.integers(0, S, n // 4), u[pick[: n // 8]]])
    dv = np.concatenate([v[pick], u[pick], rng.integers(0, S, n // 4), v[pick[: n // 8]]])
    iu0, iv0 = rng.integers(0, S, n), rng.integers(0, S, n)
    iu = np.concatenate([iu0, iv0, u[pick[: n // 4]], iu0[: n // 8], iu0[: n // 16]])
    iv = np.concatenate([iv0, iu0, v[pick[: n // 4]], iv0[: n // 8], iv0[: n // 16]])
    pd, pi = rng.permutation(len(du)), rng.permutation(len(iu))
    dels = {"u": du[pd].astype(np.uint32), "v": dv[pd].astype(np.uint32),
            "w": rng.uniform(0.5, 2.0, len(du)).astype(np.float32)}
    ins = {"u": iu[pi].astype(np.uint32), "v": iv[pi].astype(np.uint32),
           "w": rng.uniform(0.5, 2.0, len(iu)).astype(np.float32)}
    return dels, ins


def update_fixtures(ref, graphs):
    """update_raw.npz: per case the graph, a raw batch, the reference's tidy
    lists and its updated graph (packed CSR, rows sorted, exists mask)."""
    out = {}
    rng = np.random.default_rng(2024)
    for name, g, n in graphs:
        dels, ins = raw_batch(g, rng, n)
        td, ti, y = ref.tidy_update(g, dels, ins)
        for k, a in g.items():
            out[f"{name}_x_{k}"] = a
        for pre, b in (("raw_del", dels), ("raw_ins", ins), ("tidy_del", td), ("tidy_ins", ti)):
            for k, a in b.items():
                out[f"{name}_{pre}_{k}"] = a
        for k, a in y.items():
            out[f"{name}_y_{k}"] = a
        print(f"update {name}: raw {len(dels['u'])}/{len(ins['u'])} -> tidy {len(td['u'])}/{len(ti['u'])}, "
              f"E {int(g['off'][-1])} -> {len(y['keys'])}")
    out["cases"] = np.array([name for name, _, _ in graphs])
    path = os.path.join(HERE, "update_raw.npz")
    np.savez_compressed(path, **out)
    print(f"update_raw: {os.path.getsize(path)} B")


def main():
    ref = RefLib()
    ref.set_threads(1)
    g1, g0 = tiny_graphs()
    cases = []
    # tiny graphs: hand-picked batches
    d1 = G.tidy(g1, [4, 3], [3, 4], [1, 1], True)
    i1 = G.tidy(g1, [2, 8, 1, 5], [8, 2, 5, 1], [1, 1, 1, 1], False)
    cases.append(("tiny1", g1, d1, i1, Options()))
    d0 = G.tidy(g0, [2, 3], [3, 2], [1, 1], True)
    i0 = G.tidy(g0, [0, 4, 7, 2], [4, 0, 2, 7], [1, 1, 1, 1], False)
    cases.append(("tiny0", g0, d0, i0, Options()))
    # planted partition, 1-indexed, unit weights
    pp, _ = G.planted_partition(2000, 16, 12, 0.3, 42, one_indexed=True)
    dp, ip = G.random_batch(pp, 40, 160, 7)
    cases.append(("pp2k", pp, dp, ip, Options()))
    # same graph, non-default options
    cases.append(("pp2k_opts", pp, dp, ip, Options(repeat=1, resolution=0.8, tolerance=1e-3,
                                                    aggregation_tolerance=0.8, tolerance_drop=2,
                                                    max_iterations=6, max_passes=3)))
    # R-MAT scale 10 with self-loops, 0-indexed
    rm = G.rmat(10, 8, 3)
    dr, ir = G.random_batch(rm, 30, 120, 11)
    cases.append(("rmat10", rm, dr, ir, Options()))
    # weighted planted partition (non-integer f32 weights)
    rng = np.random.default_rng(5)
    ppw, _ = G.planted_partition(1500, 12, 10, 0.3, 9)
    u, v, _w = G.edges_of(ppw)
    half = u < v
    ww = rng.uniform(0.25, 4.0, half.sum()).astype(np.float32)
    ppw = G.symmetric_csr(1500, u[half], v[half], ww)
    dw, iw = G.random_batch(ppw, 25, 100, 13, weight=1.5)
    cases.append(("ppw", ppw, dw, iw, Options()))
    update_fixtures(ref, [("tiny1", g1, 8), ("pp2k", pp, 200), ("rmat10", rm, 300), ("ppw", ppw, 150)])
    if "--update-only" in sys.argv:
        return
    for name, g, d, i, o in cases:
        out = {}
        record(ref, g, d, i, o, "", out)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: S={len(g['off']) - 1} E={int(g['off'][-1])} dels={len(d['u'])} ins={len(i['u'])} "
              f"static passes={out['static_meta'][1]} Q={out['static_Q']:.6f} -> {os.path.getsize(path)} B")


if __name__ == "__main__":
    main()
