"""Log-line and CSV compatibility with the reference's driver output
(main.cxx:78-84) and its post-processing (process.js)."""
from types import SimpleNamespace

import lvlog


def _result(**kw):
    base = dict(time=12.34, markingTime=0.05, initializationTime=0.25, firstPassTime=10.0, localMoveTime=9.876,
                aggregationTime=1.5, affectedVertices=27159, iterations=3, passes=2)
    base.update(kw)
    return SimpleNamespace(**base)


def test_result_line_is_the_reference_printf():
    # C printf of main.cxx:79-81 on these values, written out by hand
    want = ("{-2.000e-05/+8.000e-05 batchf, 016 threads} -> {0000012.3ms, 0000000.1ms mark, 0000000.2ms init, "
            "0000010.0ms firstpass, 0000009.9ms locmove, 0000001.5ms aggr, 2.716e+04 aff, 0003 iters, 002 passes, "
            "0.068123457 modularity} louvainDynamicFrontierHip")
    got = lvlog.result_line(_result(), "louvainDynamicFrontierHip", 16, 0.0681234567, 2e-5, 8e-5)
    assert got == want


def test_read_log_follows_the_reference_parser():
    lines = [
        "OMP_NUM_THREADS=16",
        "2024-01-02 03:04:05 " + lvlog.graph_line("/data/graphs/rmat-22.txt"),
        lvlog.order_line(4194304, 128307572),
        lvlog.result_line(_result(), "louvainStaticHip", 16, 0.0665),
        lvlog.order_line(4194304, 128310000, insertions=5369),
        lvlog.result_line(_result(passes=1), "louvainDynamicFrontierHip", 16, 0.0701, 0.0, 1e-4),
        lvlog.order_line(4194304, 128320000, insertions=5300),
        lvlog.result_line(_result(iterations=1), "louvainDynamicFrontierHip", 16, 0.0702, 0.0, 1e-4),
    ]
    data = lvlog.read_log("\n".join(lines))
    rows = data["rmat-22"]
    # -1 after the original graph, then the pre-increment of process.js:59-60 (-1 is truthy there)
    assert [r["batch_index"] for r in rows] == [-1, 0, 1]
    assert [r["technique"] for r in rows] == ["louvainStaticHip", "louvainDynamicFrontierHip",
                                              "louvainDynamicFrontierHip"]
    r = rows[1]
    assert list(r) == ["omp_num_threads", "graph", "order", "size", "batch_deletions_fraction",
                       "batch_insertions_fraction", "batch_index", "num_threads", "time", "marking_time",
                       "initialization_time", "first_pass_time", "local_moving_phase_time",
                       "aggregation_phase_time", "affected_vertices", "iterations", "passes", "modularity",
                       "technique"]
    assert (r["order"], r["size"], r["batch_insertions_fraction"], r["num_threads"]) == (4194304, 128310000, 1e-4, 16)
    assert (r["time"], r["local_moving_phase_time"], r["affected_vertices"]) == (12.3, 9.9, 27160.0)
    assert (r["iterations"], r["passes"], r["modularity"]) == (3, 1, 0.0701)


def test_js_number_matches_number_tostring():
    cases = {0.0001: "0.0001", 1e-7: "1e-7", 1.5e-6: "0.0000015", 1e-6: "0.000001", 100.0: "100", 2.5: "2.5",
             -0.5: "-0.5", 1e21: "1e+21", 1.2345e22: "1.2345e+22", 123456789012345680000.0: "123456789012345680000",
             0.1 + 0.2: "0.30000000000000004", 27160.0: "27160", 0: "0", float("nan"): "NaN", -1: "-1"}
    for x, want in cases.items():
        assert lvlog.js_number(x) == want, (x, want)


def test_csv_text_is_write_csv_of_process_js():
    rows = [{"graph": "g", "time": 12.3, "technique": "a"}, {"graph": "g", "time": 1e-7, "technique": "b"}]
    assert lvlog.csv_text(rows) == 'graph,time,technique\n"g","12.3","a"\n"g","1e-7","b"\n'
