"""Time the oracle (louvain_oracle.c) beside the reference compiled in place
(oracle/_ref/libref.so, built by ``make -C oracle ref``) on the same inputs,
at the same thread count.  CPU only; needs the reference build.

Inputs: R-MAT (a, b, c) = (.57, .19, .19), edge factor 16, symmetrised, unit
weights (lvgraph.rmat_native), then batches of 1e-4 |E| undirected updates,
80% insertions / 20% deletions.  Static runs on the graph; ND, DS and DF run
on each updated graph from the Static result of the graph before it
(main.cxx:101-135).  Each approach reports the median of the times the two
libraries measure themselves (the reference's timed region,
inc/louvain.hxx:1167-1217), and the ratio oracle / reference.

Usage:  python oracle/calibrate.py [--scale 20] [--threads 8] [--batches 3]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "louvain-communities-openmp-dynamic_amd"))

import lvgraph  # noqa: E402
from oracle import Oracle, RefLib  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--batches", type=int, default=3)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()
    g = lvgraph.rmat_native(args.scale, 16, 0.57, 0.19, 0.19, args.seed)
    x = g.view()
    print(f"R-MAT scale {args.scale}: {len(x['off']) - 1} vertices, {len(x['keys'])} directed edges, "
          f"{args.threads} threads", flush=True)
    libs = {"oracle": Oracle(), "reference": RefLib()}
    for lib in libs.values():
        lib.set_threads(args.threads)
    times = {(n, a): [] for n in libs for a in ("static", "naive-dynamic", "delta-screening", "dynamic-frontier")}
    quality = {k: [] for k in times}
    und = len(x["keys"]) // 2
    b = max(1, int(round(1e-4 * und)))
    n_ins = int(round(0.8 * b))
    for k in range(args.batches):
        for name, lib in libs.items():
            r = lib.louvain("static", x)
            times[(name, "static")].append(r.time)
            quality[(name, "static")].append(lib.modularity(x, r.membership))
        d, i = g.random_batch(b - n_ins, n_ins, args.seed * 1000 + k)
        td, ti = lvgraph.tidy_batch(x, d, i)
        y = lvgraph.apply_batch(x, d, i)
        s0 = libs["reference"].louvain("static", x)
        for name, lib in libs.items():
            for a in ("naive-dynamic", "delta-screening", "dynamic-frontier"):
                r = lib.louvain(a, y, td, ti, s0.membership, s0.vertex_weight, s0.community_weight)
                times[(name, a)].append(r.time)
                quality[(name, a)].append(lib.modularity(y, r.membership))
        print(f"batch {k} done", flush=True)
    print("| Approach | Reference ms | Oracle ms | Oracle / reference | Q reference | Q oracle |")
    print("|---|---|---|---|---|---|")
    for a in ("static", "naive-dynamic", "delta-screening", "dynamic-frontier"):
        tr, to = np.median(times[("reference", a)]), np.median(times[("oracle", a)])
        qr, qo = np.median(quality[("reference", a)]), np.median(quality[("oracle", a)])
        print(f"| {a} | {tr:.1f} | {to:.1f} | {to / tr:.2f} | {qr:.5f} | {qo:.5f} |", flush=True)


if __name__ == "__main__":
    main()
