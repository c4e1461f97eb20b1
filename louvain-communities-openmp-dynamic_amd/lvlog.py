"""Result log lines and CSV in the reference's formats, so its
post-processing reads this library's results unchanged.

    result_line   the per-run line printed by main.cxx:78-84 (glog)
    graph_line    "Loading graph <path> ..." (main.cxx's loader banner)
    order_line    "order: N size: M [directed] {}" (+ " (insertions=K)")
    read_log      process.js:43-101 (readLogLine / readLog): rows per graph
    write_csv     process.js:29-35 (writeCsv): header line, quoted values

Numbers in the CSV are written as JavaScript's Number#toString writes them
(shortest round-trip digits; exponent form below 1e-6 and from 1e21 up).
"""
import re
from decimal import Decimal

# main.cxx:79-81
_RESULT_FMT = ("{-%.3e/+%.3e batchf, %03d threads} -> "
               "{%09.1fms, %09.1fms mark, %09.1fms init, %09.1fms firstpass, %09.1fms locmove, %09.1fms aggr, "
               "%.3e aff, %04d iters, %03d passes, %01.9f modularity} %s")

# process.js:5-9
_ROMPTH = re.compile(r"^OMP_NUM_THREADS=(\d+)")
_RGRAPH = re.compile(r"^Loading graph .*/(.*?)\.txt \.\.\.", re.M)
_RORDER = re.compile(r"^order: (\d+) size: (\d+) (?:\[\w+\] )?\{\}", re.M)
_RBATCH = re.compile(r"^order: (\d+) size: (\d+) \[directed\] \{\} \(insertions=(\d+)\)", re.M)
_RRESLT = re.compile(r"^\{-(.+?)/\+(.+?) batchf, (.+?) threads\} -> \{(.+?)ms, (.+?)ms mark, (.+?)ms init, "
                     r"(.+?)ms firstpass, (.+?)ms locmove, (.+?)ms aggr, (.+?) aff, (.+?) iters, (.+?) passes, "
                     r"(.+?) modularity\} (.+)", re.M)
_RSTAMP = re.compile(r"^\d+-\d+-\d+ \d+:\d+:\d+ ")

_RESULT_FIELDS = ("time", "marking_time", "initialization_time", "first_pass_time", "local_moving_phase_time",
                  "aggregation_phase_time", "affected_vertices", "iterations", "passes", "modularity")


def result_line(r, technique, num_threads, modularity, deletions_fraction=0.0, insertions_fraction=0.0):
    """The line main.cxx's glog prints for one result (a LouvainResult-like
    object with the reference's field names)."""
    return _RESULT_FMT % (float(deletions_fraction), float(insertions_fraction), int(num_threads), r.time,
                          r.markingTime, r.initializationTime, r.firstPassTime, r.localMoveTime, r.aggregationTime,
                          float(r.affectedVertices), int(r.iterations), int(r.passes), float(modularity), technique)


def graph_line(path):
    return f"Loading graph {path} ..."


def order_line(order, size, insertions=None):
    s = f"order: {int(order)} size: {int(size)} [directed] {{}}"
    return s if insertions is None else s + f" (insertions={int(insertions)})"


def _parse(x):
    """JavaScript parseFloat of a captured field: the longest leading
    decimal literal, NaN when there is none."""
    m = re.match(r"\s*[+-]?(?:Infinity|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)", x)
    if not m:
        return float("nan")
    t = m.group(0).strip()
    return float(t.replace("Infinity", "inf"))


def _line(ln, data, state):
    ln = _RSTAMP.sub("", ln, count=1)
    if _ROMPTH.search(ln):
        state["omp_num_threads"] = _parse(_ROMPTH.search(ln).group(1))
    elif _RGRAPH.search(ln):
        g = _RGRAPH.search(ln).group(1)
        data.setdefault(g, [])
        state["graph"] = g
    elif _RBATCH.search(ln):
        m = _RBATCH.search(ln)
        state["order"], state["size"] = _parse(m.group(1)), _parse(m.group(2))
        state["batch_deletions_fraction"] = 0
        state["batch_insertions_fraction"] = 0
        state["batch_index"] = (state.get("batch_index") or 0) + 1
    elif _RORDER.search(ln):
        m = _RORDER.search(ln)
        state["order"], state["size"] = _parse(m.group(1)), _parse(m.group(2))
        state["batch_deletions_fraction"] = 0
        state["batch_insertions_fraction"] = 0
        state["batch_index"] = -1
    elif _RRESLT.search(ln):
        m = _RRESLT.search(ln)
        row = dict(state)
        row["batch_deletions_fraction"] = _parse(m.group(1))
        row["batch_insertions_fraction"] = _parse(m.group(2))
        row["num_threads"] = _parse(m.group(3))
        for k, v in zip(_RESULT_FIELDS, m.groups()[3:13]):
            row[k] = _parse(v)
        row["technique"] = m.group(14)
        data[state["graph"]].append(row)
    return state


def read_log(text):
    """Rows per graph, in log order (process.js readLog)."""
    data, state = {}, {}
    for ln in text.replace("\r\n", "\n").split("\n"):
        state = _line(ln, data, state)
    return data


def js_number(x):
    """Number#toString (ECMAScript Number::toString, radix 10)."""
    if isinstance(x, str):
        return x
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0:
        return "0"
    sign = "-" if x < 0 else ""
    t = Decimal(repr(abs(x))).normalize().as_tuple()
    digits = "".join(map(str, t.digits))
    k = len(digits)
    n = k + t.exponent  # value = 0.digits * 10^n
    if k <= n <= 21:
        s = digits + "0" * (n - k)
    elif 0 < n <= 21:
        s = digits[:n] + "." + digits[n:]
    elif -6 < n <= 0:
        s = "0." + "0" * (-n) + digits
    else:
        e = n - 1
        s = digits[0] + ("." + digits[1:] if k > 1 else "") + "e" + ("+" if e >= 0 else "-") + str(abs(e))
    return sign + s


def csv_text(rows):
    """process.js writeCsv: the first row's keys, then every row's values,
    each in double quotes."""
    cols = list(rows[0].keys())
    out = ",".join(cols) + "\n"
    for r in rows:
        out += ",".join(f'"{js_number(v)}"' for v in r.values()) + "\n"
    return out


def write_csv(path, rows):
    with open(path, "w", newline="") as f:
        f.write(csv_text(rows))
