"""Benchmark: DF Louvain on R-MAT scale 24 (edge factor 16, a=.57 b=.19 c=.19,
symmetrised, unit weights), consecutive random batches of 1e-4 |E| (80%
insertions / 20% deletions), seed 7 -- the workload of the headline metric
"DF-Louvain edges/s & p99 batch latency at dQ<=1e-3 vs plain GPU path".

A step is one Dynamic Frontier run (louvainDynamicFrontierOmp,
inc/louvain.hxx:1674) on one batch, with the graph resident in HBM and
already updated in place by that batch (lv_dgraph_update, untimed like the
reference driver's update, main.cxx:104-109) when the timed region starts;
the state carries from batch to batch as in main.cxx:127-135.  value =
directed edges of the updated graphs processed per second of the host
clock around the K timed runs.  The same chain is then run on the frontier
and the plain GPU path (one sweep launch over all vertices per iteration),
a few of the timed batches (--cpu-batches, evenly spread) are checked
against the reference's DF from the same carry-in state (|dQ| <= 1e-3), and
the reference's own DF on those batches is timed on the host cores
(cpu_baseline): its code compiled in place (oracle/_ref) where that was
built, else the oracle restatement.  A bounded serving segment (resident
R-MAT 20..23 graphs, Poisson arrivals) reports p50 / p99 latency.

    python bench.py [--steps K] [--warmup W] [--scale 24] [--no-cpu] [--no-plain] [--no-serve]
    python bench.py --serve | --batched | --sweep     (the other forms)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "louvain-communities-openmp-dynamic_amd"))
# CPUs this process may use, read before any OpenMP runtime loads.  The
# CPU-baseline threads stay unbound: OMP_PROC_BIND would pin the process's
# main thread, which also drives the GPU, to one core.
HOST_CORES = len(os.sched_getaffinity(0))


def cpu_checker():
    """The CPU side of the bench: the reference's own code compiled in place
    (oracle/_ref/libref.so, kind "reference") where it was built, else the
    oracle restatement (kind "port").  Both time the reference's timed
    region (inc/louvain.hxx:1167-1217) themselves."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if os.path.exists(oracle.REF_SO):
        return oracle.RefLib(), "reference"
    return oracle.Oracle(), "port"


def host_info():
    """CPU model, sockets and CPUs of the host, and the CPUs this process may use."""
    model, sockets = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name") and model is None:
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("physical id"):
                    sockets.add(ln.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"cpu": model, "sockets": len(sockets) or None, "nproc": os.cpu_count(), "affinity": HOST_CORES}

HBM_PEAK_GBPS = 8000.0
KERNEL_NAME = "move_kernel<256, 64, 512, float, 2>"
FUSED_KERNEL = "wl_fused_kernel<float>"
SMALL_HUB_KERNEL = "hub_mid_kernel<256, 2048u, 1024u, float>"
MID_HUB_KERNEL = "hub_mid_kernel<512, 2048u, 16384u, float>"
def pmc_summary(kernel, scale):
    """The committed FETCH_SIZE / WRITE_SIZE summary of the DF chain at this
    R-MAT scale for this kernel (named by the kernel's template name)."""
    return os.path.join(ROOT, "profiles", f"{kernel.split('<')[0]}_df_rmat{scale}_pmc.txt")
METRIC = "DF-Louvain edges/s & p99 batch latency at dQ<=1e-3 vs plain GPU path"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else None


def algorithmic_bytes(st, tier):
    """Local-moving bytes of one tier, first pass (f32 weights): 12 B per
    visited edge (key, weight, community gather), 29 B per processed vertex
    (offset, degree, community, K, own Sigma_tot, flag), 8 B per candidate
    community (Sigma_tot gather), 20 B per move (two Sigma_tot updates and
    the membership store) and 1 B per flagged neighbour.  The hubs evaluated
    from their neighbour-community tables (slot 1) read table slots instead
    of edges: 16 B each (key, count, weight), counted in their "edges"."""
    per_edge = 16 if tier == 1 else 12
    return (per_edge * st["edges"][tier] + 29 * st["vertices"][tier] + 8 * st["candidates"][tier]
            + 20 * st["moves"][tier] + st["moved_edges"][tier])


def pmc_traffic(kernel, scale):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE summary for this scale (two rocprofv3 --pmc runs
    of this script, each scaled by the counter's reading for a copy of known
    size); None if there is no summary for the scale."""
    path = pmc_summary(kernel, scale)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split("\t")
        if len(parts) >= 2 and parts[0] == "traffic_bytes_per_launch " + kernel:
            return {"bytes_per_launch": float(parts[1]), "source": os.path.relpath(path, ROOT)}
    return None


def serve_main(args):
    """Serving form: resident R-MAT graphs of scales 20, 21, 22 and 23 (two of
    each), update batches arriving as a Poisson process, each of a size
    log-uniform in [1e-6, 1e-3] |E| (80/20 insert/delete), served by lvserve
    (device tidy + update + DF per request, several requests in flight)."""
    import numpy as np

    import lvgraph
    import lvhip
    import lvserve

    rng = np.random.default_rng(args.serve_seed)
    graphs = {}
    t = time.time()
    for scale in args.serve_scales:
        for k in range(2):
            graphs[f"rmat{scale}_{k}"] = lvgraph.rmat_native(scale, args.edge_factor, 0.57, 0.19, 0.19,
                                                             args.serve_seed * 100 + scale * 10 + k)
    log(f"{len(graphs)} graphs generated ({time.time() - t:.1f}s): "
        + ", ".join(f"{n} E={g.edges}" for n, g in graphs.items()))
    names = list(graphs)
    n_req = args.serve_requests
    arrivals = lvserve.poisson_arrivals(n_req, args.serve_rate, args.serve_seed)
    t = time.time()
    reqs = []
    for k in range(n_req):
        name = names[int(rng.integers(len(names)))]
        g = graphs[name]
        und = g.edges // 2
        frac = 10 ** rng.uniform(-6, -3)
        b = max(1, int(round(frac * und)))
        n_ins = int(round(0.8 * b))
        d, i = g.random_batch(b - n_ins, n_ins, args.serve_seed * 1000003 + k)
        reqs.append(lvserve.Request(k, name, d, i, float(arrivals[k])))
    log(f"{n_req} requests generated ({time.time() - t:.1f}s), mean batch "
        f"{np.mean([len(r.insertions['u']) + len(r.deletions['u']) for r in reqs]):.0f} directed entries")
    t = time.time()
    srv = lvserve.Server(graphs, slots=args.serve_slots, memory_budget=args.serve_budget_mib << 20)
    log(f"resident: {len(graphs)} graphs uploaded + Static ({time.time() - t:.1f}s)")
    # warm-up: one request per graph, not counted
    warm = [lvserve.Request(-1 - j, n, *graphs[n].random_batch(10, 40, 77 + j), 0.0) for j, n in enumerate(names)]
    srv.run(warm, realtime=False)
    t = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t
    s = lvserve.summary(done)
    edges = sum(graphs[r.graph].edges for r in done if r.error is None)
    log(f"served {s['requests']} requests in {wall:.1f}s: p50 {s['p50_ms']:.2f} ms, p99 {s['p99_ms']:.2f} ms, "
        f"max {s['max_ms']:.2f} ms, {s['errors']} errors")
    out = {"metric": "DF serving latency (p50/p99) under Poisson arrivals", "value": s["p99_ms"], "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
           "mean_ms": s["mean_ms"], "max_ms": s["max_ms"], "requests": s["requests"], "errors": s["errors"],
           "throughput_rps": s["throughput_rps"], "edges_per_s": edges / wall, "wall_s": wall,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{len(graphs)} resident R-MAT graphs, scales {args.serve_scales} x2, edge factor "
                                  f"{args.edge_factor}; Poisson arrivals {args.serve_rate}/s; batch size log-uniform "
                                  f"1e-6..1e-3 |E|, 80/20 insert/delete; seed {args.serve_seed}",
                      "slots": args.serve_slots, "budget_mib": args.serve_budget_mib}}
    srv.close()
    print(json.dumps(out), flush=True)


def batched_main(args):
    """Batched form: 256 planted-partition graphs (n log-uniform in
    [1k, 200k], average degree U[8, 32], inter-block fraction U[0.2, 0.4]),
    one DF update of 1e-3 |E| each (80/20), run together by lv_run_batched
    over the union of the graphs: every launch covers all of them, each
    graph keeping its own M, tolerance, iterations, passes and streams.
    The previous states are the reference's Static result of each graph.
    Graphs, batches and the previous states are resident in HBM when the
    timed region starts; each step re-uploads the previous states first
    (untimed), since a run updates them."""
    import numpy as np

    import lvgraph
    import lvhip

    rng = np.random.default_rng(args.batched_seed)
    t = time.time()
    xs, ys, dels, inss = [], [], [], []
    for k in range(args.batched_graphs):
        n = int(round(10 ** rng.uniform(3, np.log10(200000))))
        deg = float(rng.uniform(8, 32))
        p_inter = float(rng.uniform(0.2, 0.4))
        blocks = max(2, n // 1000)
        g = lvgraph.planted_partition(n, blocks, deg, p_inter, args.batched_seed * 1000 + k)
        b = max(1, int(round(1e-3 * len(g["keys"]) / 2)))
        n_ins = int(round(0.8 * b))
        d, i = lvgraph.random_batch(g, b - n_ins, n_ins, args.batched_seed * 7919 + k)
        td, ti = lvgraph.tidy_batch(g, d, i)
        xs.append(g)
        ys.append(lvgraph.apply_batch(g, d, i))
        dels.append(td)
        inss.append(ti)
    G = len(xs)
    S_tot = sum(len(g["off"]) - 1 for g in ys)
    E_tot = sum(len(g["keys"]) for g in ys)
    log(f"{G} graphs ({S_tot} vertices, {E_tot} directed edges) and their batches generated "
        f"({time.time() - t:.1f}s)")
    ctx = lvhip.Context(0)
    t = time.time()
    ux, voff = lvhip.union_graphs(xs)
    dx = ctx.device_graph(ux)
    dsx = ctx.device_state(int(voff[-1]))
    static = ctx.run_batched("static", dx, None, dsx, voff)
    dx.close()
    dsx.close()
    log(f"Static of every graph on the GPU, batched ({static[0].time:.1f} ms; {time.time() - t:.1f}s with upload)")
    # the DF carry-in: the reference's own Static result of each graph, as
    # in main.cxx:89 (louvainStaticOmp) on the host cores
    orc, kind = cpu_checker()
    orc.set_threads(HOST_CORES)
    t = time.time()
    states0 = []
    for g in xs:
        r = orc.louvain("static", g)
        states0.append((r.membership, r.vertex_weight, r.community_weight))
    log(f"reference Static of every graph ({kind}, {HOST_CORES} threads, {time.time() - t:.1f}s)")
    if args.no_tune:
        ctx.set_autotune(False)  # the default schedule (lv_run_batched tunes a new shape otherwise)
    uy, voff_y = lvhip.union_graphs(ys)
    assert np.array_equal(voff, voff_y)
    dy = ctx.device_graph(uy)
    db = ctx.device_batch(lvhip.union_batch(dels, voff), lvhip.union_batch(inss, voff))
    ds = ctx.device_state(int(voff[-1]))
    u0 = lvhip.union_state(states0, voff)
    walls, gpu_ms = [], []
    for step in range(args.warmup + args.steps):
        ds.upload(*u0)  # synchronous
        t0 = time.perf_counter()
        res = ctx.run_batched("dynamic-frontier", dy, db, ds, voff, copy_back=False)
        wall = time.perf_counter() - t0
        if step >= args.warmup:
            walls.append(wall)
            gpu_ms.append(res[0].time)
    wall = float(np.median(walls))
    it = [r.iterations for r in res]
    ps = [r.passes for r in res]
    log(f"batched DF over {G} graphs: {wall * 1e3:.2f} ms per run (median of {len(walls)}; device events "
        f"{np.median(gpu_ms):.2f} ms), iterations {min(it)}..{max(it)}, passes {min(ps)}..{max(ps)}")
    # the same runs with every local-moving launch timed by events: the
    # roofline of the dominant kernel (the id-order kernel over the union,
    # slot 0; the block-per-hub kernels and the hub batch beside it)
    ctx.stats_reset()
    ctx.set_profiling(True)
    for step in range(args.steps):
        ds.upload(*u0)
        ctx.run_batched("dynamic-frontier", dy, db, ds, voff, copy_back=False)
    ctx.set_profiling(False)
    st = ctx.stats()
    dom = max(range(5), key=lambda k: st["move_ms"][k])
    launches = max(st["move_launches"][dom], 1)
    b = algorithmic_bytes(st, dom)
    ms = st["move_ms"][dom]
    gbps = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    names = {0: "move_kernel<256, 64, 512, float, 2>", 2: SMALL_HUB_KERNEL, 3: MID_HUB_KERNEL,
             4: "hub batch (hub_prep/insert/best/apply/flag)"}
    roofline = {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
                "traffic": None, "kernel": names.get(dom, str(dom)), "avg_launch_us": ms * 1e3 / launches,
                "bytes_per_launch": b / launches, "launches": launches,
                "per_slot_ms": st["move_ms"], "per_slot_launches": st["move_launches"]}
    log(f"roofline: {roofline['kernel']}: {b / launches / 1e6:.2f} MB algorithmic per launch, "
        f"{roofline['avg_launch_us']:.1f} us -> {gbps:.1f} GB/s; per-slot ms {[round(x, 2) for x in st['move_ms']]}")
    ds.upload(*u0)
    res = ctx.run_batched("dynamic-frontier", dy, db, ds, voff)
    cpu = None
    dq = over = None
    if not args.no_cpu:
        cores = HOST_CORES
        # The reference's own DF result moves with its thread count (on a
        # weak start, one thread may run several iterations and passes where
        # sixteen stop after one), so dQ is the distance from the GPU's
        # modularity to the band of the reference at 1 and `cores` threads.
        ctime, dqs = 0.0, []
        for k in range(G):
            q0, v0, c0 = states0[k]
            orc.set_threads(cores)
            ref = orc.louvain("dynamic-frontier", ys[k], dels[k], inss[k], q0, v0, c0)
            ctime += ref.time
            orc.set_threads(1)
            ref1 = orc.louvain("dynamic-frontier", ys[k], dels[k], inss[k], q0, v0, c0)
            qa, qb = orc.modularity(ys[k], ref.membership), orc.modularity(ys[k], ref1.membership)
            q = orc.modularity(ys[k], res[k].membership)
            dqs.append(max(min(qa, qb) - q, q - max(qa, qb), 0.0))
            if dqs[-1] == max(dqs):
                worst = (k, len(ys[k]["off"]) - 1, q, qb, qa, res[k].passes, ref1.passes, ref.passes)
        dq = max(dqs)
        over = sum(d > 1e-3 for d in dqs)
        log(f"worst graph {worst[0]} (n {worst[1]}): Q {worst[2]:.5f}, reference 1 thread {worst[3]:.5f} / "
            f"{cores} threads {worst[4]:.5f}; passes {worst[5]} vs {worst[6]} / {worst[7]}; "
            f"{over} of {G} graphs beyond 1e-3 of the band")
        cpu = {"value": E_tot / (ctime * 1e-3), "unit": "edges/s", "cores": cores, "kind": kind,
               "sample": f"{kind} DF on all {G} graphs, one after another, {ctime:.1f} ms in all",
               "host": host_info()}
        log(f"{kind}: {ctime:.1f} ms for all graphs on {cores} threads; max distance of the GPU's Q from the "
            f"reference's 1/{cores}-thread band {dq:.3g}")
    out = {"metric": "batched DF: graph edges/s over 256 planted-partition graphs", "value": E_tot / wall,
           "unit": "edges/s", "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall * 1e3, "device_ms": float(np.median(gpu_ms)), "dq_max": dq,
           "dq_over_1e-3": over,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{G} planted-partition graphs in one union, n log-uniform 1k-200k, avg degree "
                                  f"U[8,32], inter-block U[0.2,0.4], one DF update of 1e-3|E| each; seed "
                                  f"{args.batched_seed}", "vertices": S_tot, "edges": E_tot},
           "roofline": roofline, "cpu_baseline": cpu}
    for d in (dy, db, ds):
        d.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def sweep_main(args):
    """R-MAT scale 24 sweep (a=.57 b=.19 c=.19, edge factor 16, symmetrised,
    unit weights, seed 24): Static on the graph, then for each batch fraction
    1e-7 .. 1e-1 of |E| undirected (80% insertions / 20% deletions, each
    applied to the original graph) ND, DS and DF from the reference's Static
    state, on the GPU and with the reference on the host cores, on the same
    inputs.  Per fraction and approach: time (the reference's timed region)
    and modularity; GPU runs are timed on their second call (the first
    allocates the context's workspace)."""
    import numpy as np

    import lvgraph
    import lvhip

    t = time.time()
    g = lvgraph.rmat_native(args.sweep_scale, args.edge_factor, 0.57, 0.19, 0.19, args.sweep_seed)
    x = g.view()
    S, E = len(x["off"]) - 1, g.edges
    log(f"R-MAT scale {args.sweep_scale}: {S} vertices, {E} directed edges ({time.time() - t:.1f}s)")
    ctx = lvhip.Context(0)
    dx = ctx.device_graph(x)
    st = ctx.device_state(S)
    ctx.run(lvhip.LV_STATIC, dx, None, st)
    rs = ctx.run(lvhip.LV_STATIC, dx, None, st)
    q_static = ctx.dmodularity(dx, st)
    orc, kind = cpu_checker()
    cores = HOST_CORES
    orc.set_threads(cores)
    r0 = orc.louvain("static", x)
    q0 = orc.modularity(x, r0.membership)
    log(f"Static: GPU {rs.time:.1f} ms Q {q_static:.6f}; {kind} ({cores} threads) {r0.time:.1f} ms Q {q0:.6f}")
    static = {"gpu_ms": rs.time, "gpu_modularity": q_static, "cpu_ms": r0.time, "cpu_modularity": q0}
    dx.close()
    rows = []
    und = E // 2
    names = {"naive-dynamic": lvhip.LV_NAIVE_DYNAMIC, "delta-screening": lvhip.LV_DELTA_SCREENING,
             "dynamic-frontier": lvhip.LV_DYNAMIC_FRONTIER}
    gpu_df_ms, df_edges = 0.0, 0
    for e in range(-7, 0):
        frac = 10.0 ** e
        b = max(1, int(round(frac * und)))
        n_ins = int(round(0.8 * b))
        t = time.time()
        d, i = g.random_batch(b - n_ins, n_ins, args.sweep_seed * 100 + 7 + e)
        td, ti = lvgraph.tidy_batch(x, d, i)
        y = lvgraph.apply_batch(x, d, i)
        Ey = len(y["keys"]) if "deg" not in y else int(np.sum(y["deg"]))
        dy = ctx.device_graph(y)
        db = ctx.device_batch(td, ti)
        row = {"fraction": frac, "updates": b, "edges": Ey}
        prep = time.time() - t
        for name, a in names.items():
            for _ in range(2):
                st.upload(r0.membership, r0.vertex_weight, r0.community_weight)
                r = ctx.run(a, dy, db, st)
            q = ctx.dmodularity(dy, st)
            ref = orc.louvain(name, y, td, ti, r0.membership, r0.vertex_weight, r0.community_weight)
            qr = orc.modularity(y, ref.membership)
            row[name] = {"gpu_ms": r.time, "cpu_ms": ref.time, "speedup": ref.time / r.time if r.time else None,
                         "gpu_modularity": q, "cpu_modularity": qr, "dq": q - qr,
                         "gpu_iterations": r.iterations, "gpu_passes": r.passes,
                         "cpu_iterations": ref.iterations, "cpu_passes": ref.passes}
            if name == "dynamic-frontier":
                gpu_df_ms += r.time
                df_edges += Ey
        dy.close()
        db.close()
        rows.append(row)
        log(f"1e{e}: {b} updates (prep {prep:.1f}s); " + "; ".join(
            f"{n.split('-')[0][:5]} GPU {row[n]['gpu_ms']:.2f} ms / {kind} {row[n]['cpu_ms']:.1f} ms, "
            f"dQ {row[n]['dq']:+.1e}" for n in names))
    st.close()
    ctx.close()
    dq = {n: max(abs(r[n]["dq"]) for r in rows) for n in names}
    out = {"metric": "R-MAT 24 sweep: ND / DS / DF time and dQ vs the reference per batch fraction",
           "value": df_edges / (gpu_df_ms * 1e-3), "unit": "edges/s", "higher_is_better": True, "n_gpus": 1,
           "dtype": "f64", "data": "synthetic", "dq_max": dq, "static": static, "fractions": rows,
           "config": {"workload": f"R-MAT scale {args.sweep_scale} edge factor {args.edge_factor} (a=.57 b=.19 "
                                  f"c=.19) symmetrised unit weights; batch fractions 1e-7..1e-1 of |E| undirected, "
                                  f"80/20, each on the original graph; seed {args.sweep_seed}",
                      "vertices": S, "edges": E},
           "cpu_baseline": {"kind": kind, "cores": cores, "host": host_info(), "unit": "ms",
                            "sample": "the reference on every fraction and approach, same inputs"}}
    print(json.dumps(out), flush=True)


def run_chain(ctx, lvhip, g0, batches, state0, W, K, path, profile=False):
    """The DF chain the way a resident graph serves it: graph 0 uploaded,
    then per batch the device tidy and in-place update (lv_dgraph_update,
    untimed, as in the reference's driver, main.cxx:104-109) and one DF run
    (louvainDynamicFrontierOmp, inc/louvain.hxx:1674) timed by the host
    clock around lv_run, with the state carried (main.cxx:127-135).  Batches
    [0, W) warm up.  Returns the run walls (s), the runs' own timed regions
    (ms), the edges of each updated graph, the update walls (s) and the
    kernel statistics.  With profile, every local-moving launch is bracketed
    by HIP events on its stream (launched one by one, not replayed)."""
    ctx.set_path(path)
    dg = ctx.device_graph(g0)
    state = ctx.device_state(dg.span)
    state.copy_from(state0)
    walls, lat, edges, upd = [], [], [], []
    for k in range(W + K):
        if k == W:
            ctx.stats_reset()
            ctx.set_profiling(profile)
        t0 = time.perf_counter()
        tb = dg.update(*batches[k])
        t1 = time.perf_counter()
        r = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dg, tb, state)
        t2 = time.perf_counter()
        tb.close()
        if k >= W:
            walls.append(t2 - t1)
            upd.append(t1 - t0)
            lat.append(r.time)
            edges.append(dg.edges)
    ctx.set_profiling(False)
    st = ctx.stats()
    dg.close()
    state.close()
    return walls, lat, edges, upd, st


def serve_segment(args, log_prefix="serve"):
    """A bounded run of the serving form (config 5's shape): resident R-MAT
    graphs of scales args.serve_scales, two each, Poisson arrivals at
    args.serve_rate, batches log-uniform in 1e-6..1e-3 |E| (80/20), served by
    lvserve; returns its summary."""
    import numpy as np

    import lvgraph
    import lvserve
    rng = np.random.default_rng(args.serve_seed)
    graphs = {}
    for scale in args.serve_scales:
        for k in range(2):
            graphs[f"rmat{scale}_{k}"] = lvgraph.rmat_native(scale, args.edge_factor, 0.57, 0.19, 0.19,
                                                             args.serve_seed * 100 + scale * 10 + k)
    names = list(graphs)
    arrivals = lvserve.poisson_arrivals(args.serve_requests, args.serve_rate, args.serve_seed)
    reqs = []
    for k in range(args.serve_requests):
        name = names[int(rng.integers(len(names)))]
        g = graphs[name]
        b = max(1, int(round(10 ** rng.uniform(-6, -3) * (g.edges // 2))))
        n_ins = int(round(0.8 * b))
        d, i = g.random_batch(b - n_ins, n_ins, args.serve_seed * 1000003 + k)
        reqs.append(lvserve.Request(k, name, d, i, float(arrivals[k])))
    srv = lvserve.Server(graphs, slots=args.serve_slots, memory_budget=args.serve_budget_mib << 20)
    warm = [lvserve.Request(-1 - j, n, *graphs[n].random_batch(10, 40, 77 + j), 0.0) for j, n in enumerate(names)]
    srv.run(warm, realtime=False)
    t = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t
    s = lvserve.summary(done)
    s["edges_per_s"] = sum(graphs[r.graph].edges for r in done if r.error is None) / wall
    s["graphs"] = {n: g.edges for n, g in graphs.items()}
    log(f"{log_prefix}: {s['requests']} requests over {len(graphs)} resident graphs in {wall:.1f}s: "
        f"p50 {s['p50_ms']:.2f} ms, p99 {s['p99_ms']:.2f} ms, max {s['max_ms']:.2f} ms, {s['errors']} errors")
    srv.close()
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch-fraction", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference check and CPU baseline")
    ap.add_argument("--cpu-batches", type=int, default=3,
                    help="timed batches the reference also runs (dQ and the CPU baseline), evenly spread")
    ap.add_argument("--no-plain", action="store_true", help="skip the frontier and plain paths")
    ap.add_argument("--no-static", action="store_true", help="skip the (reported) GPU Static run")
    ap.add_argument("--static-runs", type=int, default=5,
                    help="GPU and reference Static runs reported (their mean; the DF chain starts from the "
                         "reference's first)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving segment")
    ap.add_argument("--no-tune", action="store_true", help="run the default schedule (no lv_autotune)")
    ap.add_argument("--stream-fraction", type=float, default=None,
                    help="streams per live vertex of a level (lv_context_set_schedule)")
    ap.add_argument("--worklist-hub-degree", type=int, default=None,
                    help="degree above which worklist vertices run in the hub batch (256..4096)")
    ap.add_argument("--log-lines", default=None,
                    help="also write the DF chain's results as the reference driver's log lines to this file")
    ap.add_argument("--path", choices=("frontier", "worklist", "plain"), default="worklist",
                    help="local-moving path of the timed chain (the others are timed beside it)")
    ap.add_argument("--serve", action="store_true", help="run the serving form instead of the DF chain")
    ap.add_argument("--serve-scales", type=int, nargs="+", default=[20, 21, 22, 23])
    ap.add_argument("--serve-requests", type=int, default=None,
                    help="requests (default 10000 with --serve, 2000 for the segment of the default run)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-mib", type=int, default=65536,
                    help="device memory of the server: graphs, states, workspaces and batches in flight")
    ap.add_argument("--serve-seed", type=int, default=8)
    ap.add_argument("--batched", action="store_true", help="run the batched form (many graphs at once)")
    ap.add_argument("--batched-graphs", type=int, default=256)
    ap.add_argument("--batched-seed", type=int, default=256)
    ap.add_argument("--sweep", action="store_true", help="run the R-MAT 24 batch-fraction sweep")
    ap.add_argument("--sweep-scale", type=int, default=24)
    ap.add_argument("--sweep-seed", type=int, default=24)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        args.serve_requests = args.serve_requests or 10000
        return serve_main(args)
    if args.batched:
        return batched_main(args)
    if args.sweep:
        return sweep_main(args)
    import numpy as np

    import lvgraph
    import lvhip

    W, K = args.warmup, args.steps
    n_und = args.edge_factor << args.scale
    B = int(round(args.batch_fraction * n_und))
    n_ins = int(round(0.8 * B))
    n_del = B - n_ins
    t = time.time()
    H0 = lvgraph.rmat_native(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed)
    g0 = H0.view()
    S = H0.span
    log(f"rmat scale {args.scale}: S={S} E={H0.edges} ({time.time() - t:.1f}s)")

    ctx = lvhip.Context(0)
    if args.stream_fraction is not None:
        ctx.set_schedule(stream_fraction=args.stream_fraction)
    if args.worklist_hub_degree is not None:
        ctx.set_hub_threshold(worklist_degree=args.worklist_hub_degree)
    dg0 = ctx.device_graph(g0)
    state0 = ctx.device_state(S)
    # GPU Static Louvain on the original graph (reported), and the starting
    # state of the DF chain: the reference's own Static result, as in
    # main.cxx:89 (louvainStaticOmp), on the host cores.
    # (Static's outcome moves from run to run on the GPU with its concurrent
    # streams, and in the reference with its threads' schedule: both are run
    # --static-runs times and reported by their mean, with every run listed)
    rs, q_static, static_runs = None, None, []
    if not args.no_static:
        for k in range(max(1, args.static_runs)):
            t = time.time()
            rs = ctx.run(lvhip.LV_STATIC, dg0, None, state0)
            static_runs.append((ctx.dmodularity(dg0, state0), rs.time))
            log(f"GPU static: Q={static_runs[-1][0]:.6f} iters={rs.iterations} passes={rs.passes} "
                f"time={rs.time:.1f} ms (wall {time.time() - t:.1f}s)")
        q_static = sum(q for q, _ in static_runs) / len(static_runs)
    orc, kind = cpu_checker()
    cores = HOST_CORES
    orc.set_threads(cores)
    ref_runs = []
    for k in range(max(1, args.static_runs if not args.no_static else 1)):
        t = time.time()
        r = orc.louvain("static", g0)
        state0.upload(r.membership, r.vertex_weight, r.community_weight)
        ref_runs.append((ctx.dmodularity(dg0, state0), r.time))
        log(f"reference static ({kind}, {cores} threads): Q={ref_runs[-1][0]:.6f} iters={r.iterations} "
            f"passes={r.passes} time={r.time:.1f} ms (wall {time.time() - t:.1f}s)")
        if k == 0:
            r0 = r  # the DF chain's carry-in state
    state0.upload(r0.membership, r0.vertex_weight, r0.community_weight)
    q_static_ref = sum(q for q, _ in ref_runs) / len(ref_runs)
    dg0.close()

    # The batches, and the quality chain (untimed): each batch's device
    # update and DF run, then the reference's DF on the same updated graph
    # from the same carry-in state (dQ per batch, and the CPU baseline).
    paths = {"frontier": lvhip.PATH_FRONTIER, "worklist": lvhip.PATH_WORKLIST, "plain": lvhip.PATH_SWEEP}
    ctx.set_path(paths[args.path])
    t = time.time()
    batches, dqs, ctime, cedges = [], [], 0.0, 0
    logf = open(args.log_lines, "w") if args.log_lines else None
    if logf:
        # the reference driver's log format (main.cxx:78-84), readable by its process.js
        import lvlog
        logf.write(lvlog.graph_line(f"synthetic/rmat-{args.scale}.txt") + "\n")
        logf.write(lvlog.order_line(S, H0.edges) + "\n")
        if rs is not None:
            logf.write(lvlog.result_line(rs, "louvainStaticHip", 1, static_runs[-1][0]) + "\n")
    dq_state = ctx.device_state(S)
    dq_state.copy_from(state0)
    dqg = ctx.device_graph(g0)
    Hk, H_warm = H0, H0
    n_chk = max(1, min(args.cpu_batches, K))
    checked = {W + int(round(j * (K - 1) / max(n_chk - 1, 1))) for j in range(n_chk)}
    for k in range(W + K):
        d, i = Hk.random_batch(n_del, n_ins, args.seed * 1000 + k + 1)
        batches.append((d, i))
        check = not args.no_cpu and k in checked
        prev = dq_state.download() if check else None
        tb = dqg.update(d, i)
        r = ctx.run(lvhip.LV_DYNAMIC_FRONTIER, dqg, tb, dq_state)
        tb.close()
        q = ctx.dmodularity(dqg, dq_state) if (check or logf) else None
        if check or logf:
            Hn = Hk.apply(d, i)
            if check:
                ref = orc.louvain("dynamic-frontier", Hn.view(), d, i, *prev)
                dqs.append(q - orc.modularity(Hn.view(), ref.membership))
                ctime += ref.time
                cedges += Hn.edges
            if logf:
                logf.write(lvlog.order_line(S, Hn.edges, insertions=len(i["u"])) + "\n")
                logf.write(lvlog.result_line(r, "louvainDynamicFrontierHip", 1, q, 0.0, args.batch_fraction) + "\n")
            Hk = Hn
        else:
            Hk = Hk.apply(d, i)
        if k == W - 1:
            H_warm = Hk
    dqg.close()
    dq_state.close()
    if logf:
        logf.close()
        log(f"log lines written to {args.log_lines}")
    del Hk
    log(f"{W + K} batches of {B} updates ({n_del} del / {n_ins} ins undirected), quality chain and reference "
        f"runs ({time.time() - t:.1f}s)")

    # Schedule tuning (lv_autotune, untimed): the graph and state after the
    # warm-up batches (the second builds the hub tables), a probe batch of the
    # same kind beyond the chain; the choice is cached for the graph's shape
    # and taken by every later DF run of the worklist path.
    tuning = None
    if not args.no_tune and args.path == "worklist":
        t = time.time()
        ctx.set_path(paths["worklist"])
        tg = ctx.device_graph(g0)
        ts = ctx.device_state(S)
        ts.copy_from(state0)
        for k in range(W):
            tb = tg.update(*batches[k])
            ctx.run(lvhip.LV_DYNAMIC_FRONTIER, tg, tb, ts)
            tb.close()
        pd, pi = H_warm.random_batch(n_del, n_ins, args.seed * 1000 + 999)
        tb = tg.update(pd, pi)
        tuning = ctx.autotune(lvhip.LV_DYNAMIC_FRONTIER, tg, tb, ts, reps=5)
        for h in (tb, tg, ts):
            h.close()
        del pd, pi
        log(f"tuned ({tuning['candidates']} schedules, {time.time() - t:.1f}s): worklist hub degree "
            f"{tuning['worklist_hub_degree']}, streams {tuning['worklist_streams']}, interleave "
            f"{tuning['interleave']}: {tuning['ms']:.3f} ms vs defaults {tuning['default_ms']:.3f} ms on the probe")

    del H_warm
    # Timed: the chosen path, then the others beside it, each over the same chain.
    walls, lat, edges, upd, _ = run_chain(ctx, lvhip, g0, batches, state0, W, K, paths[args.path])
    value = sum(edges) / sum(walls)
    log(f"DF {args.path}: {K} runs {sum(walls) * 1e3:.2f} ms wall, {value:.4g} edges/s, "
        f"p50 {percentile(walls, 50) * 1e3:.3f} ms p99 {percentile(walls, 99) * 1e3:.3f} ms")
    # the same chain with every local-moving launch timed by events: the
    # per-kernel durations of the roofline
    pw, _, _, _, st = run_chain(ctx, lvhip, g0, batches, state0, W, K, paths[args.path], profile=True)
    log(f"DF {args.path}, profiled (per-launch events, no graph replay): {sum(pw) * 1e3:.2f} ms wall")
    others = {}
    if not args.no_plain:
        for name, path in paths.items():
            if name == args.path:
                continue
            ow, olat, oedges, _, _ = run_chain(ctx, lvhip, g0, batches, state0, W, K, path)
            ov = sum(oedges) / sum(ow)
            others[name] = {"value": ov, "ms_per_step": sum(ow) * 1e3 / K, "p50_ms": percentile(ow, 50) * 1e3,
                            "p99_ms": percentile(ow, 99) * 1e3, "speedup": value / ov}
            log(f"DF {name}: {sum(ow) * 1e3:.2f} ms wall, {ov:.4g} edges/s, {args.path} path "
                f"{others[name]['speedup']:.2f}x faster")
    ctx.set_path(paths[args.path])
    e2e = [(a + b) * 1e3 for a, b in zip(walls, upd)]
    end_to_end = {"p50_ms": percentile(e2e, 50), "p99_ms": percentile(e2e, 99),
                  "update_p50_ms": percentile(upd, 50) * 1e3, "update_p99_ms": percentile(upd, 99) * 1e3}
    log(f"end to end (H2D + device tidy + graph update + DF): p50 {end_to_end['p50_ms']:.3f} ms, "
        f"p99 {end_to_end['p99_ms']:.3f} ms; update alone p50 {end_to_end['update_p50_ms']:.3f} ms")

    # Roofline of the dominant local-moving kernel (largest total time).
    # Stats slots on the worklist path: 0 = the id-order kernel (one launch
    # per sub-round), 1 = hubs evaluated from their tables, 2 / 3 = the
    # block-per-hub kernels (small / mid hubs), 4 = the hub batch.
    def kernel_line(slot, name):
        launches = max(st["move_launches"][slot], 1)
        ms = st["move_ms"][slot]
        b = algorithmic_bytes(st, slot)
        gbps = (b / (ms * 1e-3) / 1e9) if ms > 0 else 0.0
        return {"kernel": name, "achieved": gbps, "avg_launch_us": ms * 1e3 / launches,
                "bytes_per_launch": b / launches, "launches": launches,
                "per_launch": {k: st[k][slot] / launches for k in ("vertices", "edges", "candidates", "moves")}}
    lines = {0: kernel_line(0, KERNEL_NAME), 2: kernel_line(2, SMALL_HUB_KERNEL),
             3: kernel_line(3, MID_HUB_KERNEL),
             4: kernel_line(4, "hub batch (hub_prep/insert/best/apply/flag)")}
    dom = max((0, 2, 3), key=lambda k: st["move_ms"][k])
    main_k, hub_k = lines[dom], lines[4]
    if st["fused_launches"] and st["fused_ms"] > max(st["move_ms"][k] for k in (0, 2, 3)):
        # sub-rounds whose id-order kernel and tabled hubs ran as one launch
        # (wl_fused_kernel): its bytes are those of slots 0 and 1 (which also
        # count any sub-round that ran them apart: move_launches of slot 0)
        fl, fms = st["fused_launches"], st["fused_ms"]
        b = algorithmic_bytes(st, 0) + algorithmic_bytes(st, 1)
        main_k = {"kernel": FUSED_KERNEL, "achieved": b / (fms * 1e-3) / 1e9, "avg_launch_us": fms * 1e3 / fl,
                  "bytes_per_launch": b / fl, "launches": fl, "unfused_id_order_launches": st["move_launches"][0],
                  "per_launch": {k: (st[k][0] + st[k][1]) / fl for k in ("vertices", "edges", "candidates", "moves")}}
        dom = None
    traffic = pmc_traffic(main_k["kernel"], args.scale)
    roofline = {"bound": "hbm", "achieved": main_k["achieved"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": main_k["achieved"] / HBM_PEAK_GBPS, "traffic": traffic and traffic["bytes_per_launch"],
                "kernel": main_k["kernel"], "avg_launch_us": main_k["avg_launch_us"],
                "bytes_per_launch": main_k["bytes_per_launch"], "launches": main_k["launches"],
                "traffic_source": traffic and traffic["source"], "hub_batch": hub_k,
                "other_kernels": [lines[k] for k in (0, 2, 3) if k != dom and st["move_launches"][k]]}
    # edges visited by local-moving over the K timed runs: every kernel but
    # the tabled hubs, which read their tables instead
    visits = sum(st["edges"]) - st["edges"][1]
    log("kernel stats per slot: " + "; ".join(
        f"{k}: launches {st['move_launches'][k]} ms {st['move_ms'][k]:.3f} vertices {st['vertices'][k]} "
        f"edges {st['edges'][k]} candidates {st['candidates'][k]} moves {st['moves'][k]}" for k in range(5)) +
        f"; fused launches {st['fused_launches']} ms {st['fused_ms']:.3f}")
    log(f"roofline: {main_k['kernel']}: {main_k['bytes_per_launch'] / 1e6:.2f} MB algorithmic per launch, "
        f"{main_k['avg_launch_us']:.1f} us -> {main_k['achieved']:.1f} GB/s; per launch " + ", ".join(
            f"{k} {main_k['per_launch'][k]:.0f}" for k in main_k["per_launch"]))

    cpu, dq_max = None, None
    if not args.no_cpu:
        dq_max = max(abs(x) for x in dqs)
        log("dQ per batch (GPU - reference): " + " ".join(f"{x:+.2e}" for x in dqs))
        what = "the reference's DF compiled in place" if kind == "reference" else "oracle DF (C + OpenMP)"
        cpu = {"value": cedges / (ctime * 1e-3), "unit": "edges/s", "cores": cores, "kind": kind,
               "sample": f"{what} on {len(dqs)} of the {K} timed batches (evenly spread), same updated "
                         f"graphs and GPU carry-in states, reference timed region; {ctime:.1f} ms total",
               "host": host_info()}
        log(f"{kind}: max |dQ| = {dq_max:.3g}; cpu {cpu['value']:.4g} edges/s on {cores} threads")

    serving = None
    if not args.no_serve:
        args.serve_requests = args.serve_requests or 2000
        serving = serve_segment(args)

    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": sum(walls) * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"DF Louvain, R-MAT scale {args.scale} edge factor {args.edge_factor} "
                               f"(a=.57 b=.19 c=.19) symmetrised unit weights, consecutive batches of "
                               f"{args.batch_fraction:g}|E| 80/20 insert/delete applied in place, seed {args.seed}",
                   "vertices": S, "edges": H0.edges, "batch_updates": B, "parallelism": "1 GPU"},
        "p50_batch_ms": percentile(walls, 50) * 1e3, "p99_batch_ms": percentile(walls, 99) * 1e3,
        "algorithm_p50_ms": percentile(lat, 50),
        "edge_visits_per_s": visits / (sum(pw) if sum(pw) > 0 else 1.0),
        "end_to_end": end_to_end,
        "dq_max": dq_max, "static": {"gpu_modularity": q_static,
                                     "gpu_ms": static_runs and sum(t for _, t in static_runs) / len(static_runs),
                                     "gpu_runs": [q for q, _ in static_runs],
                                     "reference_modularity": q_static_ref,
                                     "reference_ms": sum(t for _, t in ref_runs) / len(ref_runs),
                                     "reference_runs": [q for q, _ in ref_runs], "reference_threads": cores,
                                     "statistic": "mean of the runs"},
        "path": args.path, "plain_gpu_path": others.get("plain"),
        "other_gpu_paths": {k: v for k, v in others.items() if k != "plain"},
        "roofline": roofline, "cpu_baseline": cpu, "serving": serving, "tuning": tuning,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
