"""CPU-side checks of the native libraries (no GPU compute).

* liblouvain_hip.so loads and exports every function include/louvain_hip.h
  declares; option defaults equal LouvainOptions' (inc/louvain.hxx:60);
  without a device the context reports LV_ERR_NODEVICE instead of crashing.
* liblvgraph.so exports every function include/lvgraph.h declares.
"""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(lvg?_[a-z0-9_]+)\s*\(", src)
    return sorted(set(n for n in names if not n.endswith("_t")))


def test_louvain_hip_exports_every_declared_symbol():
    import lvhip
    declared = _declared("louvain_hip.h")
    assert len(declared) >= 30
    missing = [n for n in declared if not hasattr(lvhip._LIB, n)]
    assert not missing, missing
    assert sorted(lvhip.EXPORTS) == declared


def test_lvgraph_exports_every_declared_symbol():
    import lvgraph
    declared = _declared("lvgraph.h")
    missing = [n for n in declared if not hasattr(lvgraph._L, n)]
    assert not missing, missing
    assert sorted(lvgraph.EXPORTS) == declared


def test_options_default_match_reference():
    import lvhip
    o = lvhip._Options()
    lvhip._LIB.lv_options_default(C.byref(o))
    ref = lvhip.LouvainOptions()
    assert (o.repeat, o.resolution, o.tolerance, o.aggregation_tolerance, o.tolerance_drop, o.max_iterations,
            o.max_passes) == (1, 1.0, 1e-2, 1.0, 10.0, 20, 10)
    assert (ref.repeat, ref.resolution, ref.tolerance, ref.aggregationTolerance, ref.toleranceDrop,
            ref.maxIterations, ref.maxPasses) == (1, 1.0, 1e-2, 1.0, 10.0, 20, 10)


def test_status_strings():
    import lvhip
    for s in (0, -1, -2, -3, -4, -5, -6):
        assert lvhip._LIB.lv_status_string(s)
    assert lvhip._LIB.lv_status_string(-4).decode() == "batch not sorted/unique"


def test_null_arguments_rejected_without_device():
    import lvhip
    L = lvhip._LIB
    assert L.lv_context_create(0, None) == -1
    assert L.lv_louvain_static(None, None, None, None) == -1
    assert L.lv_run(None, 0, None, None, None, None, None) == -1
    assert L.lv_context_set_path(None, 0) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_context_without_device_reports_nodevice():
    import lvhip
    h = C.c_void_p()
    assert lvhip._LIB.lv_context_create(0, C.byref(h)) == -5
    with pytest.raises(lvhip.LvError):
        lvhip.Context(0)
