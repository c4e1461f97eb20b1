"""Loading helpers for the golden vectors in tests/golden/*.npz, and the
bar of the DF tests against the reference at several thread counts."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))  # allow_pickle stays False
    return {k: z[k] for k in z.files}


def graph(f, which):
    g = {"off": f[f"{which}_off"], "keys": f[f"{which}_keys"], "w": f[f"{which}_w"]}
    if f"{which}_ex" in f:
        g["ex"] = f[f"{which}_ex"]
    return g


def batch(f, which):
    return {"u": f[f"{which}_u"], "v": f[f"{which}_v"], "w": f[f"{which}_w"]}


def options(f):
    from oracle import Options
    o = f["opts"]
    return Options(int(o[0]), float(o[1]), float(o[2]), float(o[3]), float(o[4]), int(o[5]), int(o[6]))


def super_graph(f, prefix="agg"):
    return {"off": f[f"{prefix}_off"], "deg": f[f"{prefix}_deg"], "keys": f[f"{prefix}_keys"], "w": f[f"{prefix}_w"]}


def rows_as_maps(off, deg, keys, w):
    """Gapped CSR rows as {key: weight} dicts (row order of a super-graph is
    the hashtable's first-seen order, which depends on thread timing)."""
    out = []
    for c in range(len(deg)):
        o, d = int(off[c]), int(deg[c])
        out.append(dict(zip(keys[o:o + d].tolist(), w[o:o + d].tolist())))
    return out


def within_band(q, band, tol):
    """DF against the reference's results at several thread counts from the
    same carry-in state: not below the lowest by more than tol, and, where
    those results agree with each other within tol, not above the highest by
    more than tol.  (Where they do not -- a small graph on which one thread
    count runs a second pass and another stops after one -- the reference
    gives no single result to hold the GPU's above.)"""
    lo, hi = min(band), max(band)
    if q < lo - tol:
        return False
    return hi - lo > tol or q <= hi + tol
