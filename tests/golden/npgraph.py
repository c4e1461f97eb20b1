"""Small numpy graph builders for fixtures and CPU tests.

Graphs are dicts of CSR arrays (``off`` u64[S+1], ``keys`` u32, ``w`` f32,
optional ``ex`` u8) with rows sorted by neighbour id and one entry per
(u, v), the layout the reference's DiGraph keeps (inc/_bitset.hxx:40,
:245-262).  Batches follow the reference's tidy form: existence-filtered,
sorted by (u, v), unique, both directions present (inc/batch.hxx:202-209).
"""
import numpy as np


def csr_from_edges(S, u, v, w=None, ex=None):
    """Sorted, de-duplicated CSR from directed (u, v[, w]) lists (last write wins)."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.ones(len(u), np.float32) if w is None else np.asarray(w, np.float32)
    key = u * S + v
    # keep the last occurrence of each (u, v), like set_union_last_inplace
    rev = np.arange(len(key))[::-1]
    _, first_in_rev = np.unique(key[::-1], return_index=True)
    keep = rev[first_in_rev]
    keep.sort()
    key, w = key[keep], w[keep]
    order = np.argsort(key, kind="stable")
    key, w = key[order], w[order]
    uu = (key // S).astype(np.int64)
    vv = (key % S).astype(np.uint32)
    off = np.zeros(S + 1, np.uint64)
    np.add.at(off, uu + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    g = {"off": off, "keys": vv, "w": w.astype(np.float32)}
    if ex is not None:
        g["ex"] = np.asarray(ex, np.uint8)
    return g


def symmetric_csr(S, u, v, w=None, ex=None):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.ones(len(u), np.float32) if w is None else np.asarray(w, np.float32)
    return csr_from_edges(S, np.concatenate([u, v]), np.concatenate([v, u]), np.concatenate([w, w]), ex)


def edges_of(g):
    S = len(g["off"]) - 1
    off = g["off"].astype(np.int64)
    deg = np.diff(off) if g.get("deg") is None else g["deg"].astype(np.int64)
    u = np.repeat(np.arange(S), deg)
    if g.get("deg") is None:
        return u, g["keys"].astype(np.int64), g["w"]
    idx = np.concatenate([np.arange(off[i], off[i] + deg[i]) for i in range(S)]) if S else np.zeros(0, np.int64)
    return u, g["keys"][idx].astype(np.int64), g["w"][idx]


def planted_partition(n, blocks, avg_deg, p_inter, seed, exponent=2.0, one_indexed=False):
    """Planted partition with power-law block sizes; unit weights, no self-loops."""
    rng = np.random.default_rng(seed)
    raw = 1.0 / np.arange(1, blocks + 1) ** (exponent - 1.0)
    sizes = np.maximum(2, np.floor(raw / raw.sum() * n).astype(np.int64))
    sizes[0] += n - sizes.sum()
    bid = np.repeat(np.arange(blocks), sizes)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    m = n * avg_deg // 2
    n_inter = int(m * p_inter)
    n_intra = m - n_inter
    u1 = rng.integers(0, n, n_inter)
    v1 = rng.integers(0, n, n_inter)
    u2 = rng.integers(0, n, n_intra)
    b = bid[u2]
    v2 = starts[b] + rng.integers(0, sizes[b])
    u = np.concatenate([u1, u2])
    v = np.concatenate([v1, v2])
    keep = u != v
    u, v = u[keep], v[keep]
    base = 1 if one_indexed else 0
    S = n + base
    ex = None
    if one_indexed:
        ex = np.ones(S, np.uint8)
        ex[0] = 0
    return symmetric_csr(S, u + base, v + base, None, ex), bid


def rmat(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19):
    """R-MAT edges, symmetrised and de-duplicated; self-loops kept."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    m = edge_factor * n
    u = np.zeros(m, np.int64)
    v = np.zeros(m, np.int64)
    for bit in range(scale):
        r = rng.random(m)
        ubit = r >= a + b
        vbit = ((r >= a) & (r < a + b)) | (r >= a + b + c)
        u |= ubit.astype(np.int64) << bit
        v |= vbit.astype(np.int64) << bit
    return symmetric_csr(n, u, v)


def has_edge(g, u, v):
    off = g["off"]
    out = np.zeros(len(u), bool)
    for i, (a, b) in enumerate(zip(u, v)):
        row = g["keys"][off[a]:off[a + 1]]
        j = np.searchsorted(row, b)
        out[i] = j < len(row) and row[j] == b
    return out


def tidy(g, u, v, w, exists):
    """filterEdgesByExistenceU + sortEdgesByIdU + uniqueEdgesU (inc/batch.hxx:154-209)."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.asarray(w, np.float32)
    keep = has_edge(g, u, v) == exists
    u, v, w = u[keep], v[keep], w[keep]
    order = np.lexsort((v, u))
    u, v, w = u[order], v[order], w[order]
    if len(u):
        first = np.ones(len(u), bool)
        first[1:] = (u[1:] != u[:-1]) | (v[1:] != v[:-1])
        u, v, w = u[first], v[first], w[first]
    return {"u": u.astype(np.uint32), "v": v.astype(np.uint32), "w": w.astype(np.float32)}


def random_batch(g, n_del, n_ins, seed, weight=1.0):
    """Random undirected deletions (an edge of a uniform vertex) and insertions
    (uniform vertex pairs), both directions, then tidied against g."""
    rng = np.random.default_rng(seed)
    S = len(g["off"]) - 1
    off = g["off"].astype(np.int64)
    ex = g.get("ex")
    lo = 1 if ex is not None and not ex[0] else 0
    du, dv, dw = [], [], []
    for _ in range(n_del):
        for _retry in range(5):
            a = int(rng.integers(lo, S))
            d = off[a + 1] - off[a]
            if d == 0:
                continue
            j = off[a] + int(rng.integers(0, d))
            b = int(g["keys"][j])
            ww = float(g["w"][j])
            du += [a, b]
            dv += [b, a]
            dw += [ww, ww]
            break
    iu = rng.integers(lo, S, n_ins)
    iv = rng.integers(lo, S, n_ins)
    dels = tidy(g, du, dv, dw, True)
    ins = tidy(g, np.concatenate([iu, iv]), np.concatenate([iv, iu]), np.full(2 * n_ins, weight, np.float32), False)
    return dels, ins


def apply_batch(g, dels, ins):
    """applyBatchUpdateU (inc/batch.hxx:219-226) on a CSR: remove then add."""
    S = len(g["off"]) - 1
    u, v, w = edges_of(g)
    key = u * S + v
    dkey = dels["u"].astype(np.int64) * S + dels["v"].astype(np.int64)
    keep = ~np.isin(key, dkey)
    u, v, w = u[keep], v[keep], w[keep]
    u = np.concatenate([u, ins["u"].astype(np.int64)])
    v = np.concatenate([v, ins["v"].astype(np.int64)])
    w = np.concatenate([w, ins["w"]])
    return csr_from_edges(S, u, v, w, g.get("ex"))
