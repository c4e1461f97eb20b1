"""Checks of a batch tidy and graph update against the reference's own
(tests/golden/update_raw.npz, written by make_golden.py from
tidyBatchUpdateU, inc/batch.hxx:202-209, and applyBatchUpdateOmpU,
inc/batch.hxx:241-248).

The reference sorts each list with std::sort, which is not stable, and then
keeps the first entry of every run of equal (u, v) (inc/batch.hxx:169-192):
of repeated entries that carry different weights it keeps an unspecified
one.  This build keeps the first in input order.  So the kept (u, v) pairs,
the updated rows and the existence flags must be identical; a weight must be
identical wherever the raw list gives its pair a single weight, and at a
pair repeated with different weights it must be the first in input order
(while the reference's is one of the pair's raw weights)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "update_raw.npz")


def load():
    f = np.load(GOLDEN)
    cases = []
    for name in f["cases"]:
        name = str(name)
        g = {k: f[f"{name}_x_{k}"] for k in ("off", "keys", "w", "ex") if f"{name}_x_{k}" in f}
        raw = [{k: f[f"{name}_raw_{p}_{k}"] for k in "uvw"} for p in ("del", "ins")]
        tidy = [{k: f[f"{name}_tidy_{p}_{k}"] for k in "uvw"} for p in ("del", "ins")]
        y = {k: f[f"{name}_y_{k}"] for k in ("off", "keys", "w", "ex")}
        cases.append((name, g, raw, tidy, y))
    return cases


def _weights_by_pair(raw):
    """(u, v) -> (first weight in input order, set of weights)."""
    first, every = {}, {}
    for u, v, w in zip(raw["u"].tolist(), raw["v"].tolist(), raw["w"].tolist()):
        first.setdefault((u, v), w)
        every.setdefault((u, v), set()).add(w)
    return first, every


def check_tidy(got, want, raw, what):
    """got: this build's tidy list; want: the reference's."""
    assert np.array_equal(got["u"], want["u"]) and np.array_equal(got["v"], want["v"]), what
    first, every = _weights_by_pair(raw)
    repeats = 0
    for u, v, gw, rw in zip(got["u"].tolist(), got["v"].tolist(), got["w"].tolist(), want["w"].tolist()):
        ws = every[(u, v)]
        if len(ws) == 1:
            assert gw == rw, (what, u, v, gw, rw)
        else:
            repeats += 1
            assert gw == first[(u, v)], (what, u, v, gw, first[(u, v)])
            assert rw in ws, (what, u, v, rw, ws)
    return repeats


def check_graph(got, want, ins_raw, what):
    """got / want: packed CSR dicts (rows sorted), exists masks."""
    assert np.array_equal(np.asarray(got["off"], np.uint64), want["off"]), what
    assert np.array_equal(got["keys"], want["keys"]), what
    S = len(want["off"]) - 1
    assert np.array_equal(np.asarray(got.get("ex", np.ones(S, np.uint8)), np.uint8), want["ex"]), what
    first, every = _weights_by_pair(ins_raw)
    src = np.repeat(np.arange(S), np.diff(want["off"].astype(np.int64)))
    for i in np.nonzero(got["w"] != want["w"])[0].tolist():
        pair = (int(src[i]), int(want["keys"][i]))
        assert pair in every and len(every[pair]) > 1, (what, pair)  # only an inserted, repeated pair
        assert got["w"][i] == first[pair] and want["w"][i] in every[pair], (what, pair)
