"""explain() over GPU node trees: the reference's annotations followed by
what the device does (hash or tuple-code tables, string keys, the device
scan feed)."""
import numpy as np
import pytest

from helpers import rng

pytestmark = pytest.mark.gpu


def test_gpu_plan_notes(gpu, tmp_path):
    r = rng(5)
    p = tmp_path / "e.vtr"
    gpu.vtr_write(p, {"k": r.integers(0, 9, 100), "x": r.normal(size=100)}, rg_rows=40)
    f = gpu.filter_node("gpu", gpu.scan_vtr(p), ("cmp", ">", ("col", "x"), ("dbl", 0.0)))
    g = gpu.group_agg_node("gpu", f, ["k"], [("n", "n", None, 0)], None)
    lines = gpu.explain(g)
    assert lines[0].startswith("GroupAggNode [materializes, 1 keys; gpu: hash group-by, first-seen order")
    assert lines[1] == "  FilterNode [streaming; gpu: predicate on the device, selection vector]"
    assert lines[2] == "    ScanNode [streaming, 2 cols, v3 stats]"
    gpu.core().vx_node_free(g)
    t = gpu.Table({"s": np.array(["a", "b"]), "x": np.ones(2)})
    g = gpu.group_agg_node("gpu", t.scan(), ["s"], [("n", "n", None, 0)], "/tmp")
    assert gpu.explain(g)[0] == ("GroupAggNode [materializes, 1 keys; gpu: tuple-code group-by (string keys by "
                                 "exact device dictionary codes), key order]")
    gpu.core().vx_node_free(g)
    L, R = gpu.Table({"k": np.arange(3)}), gpu.Table({"k": np.arange(2)})
    j = gpu.join_node("gpu", L.scan(), R.scan(), "left", [(0, 0)])
    assert gpu.explain(j)[0] == ("JoinNode [build right + stream left, left, 1 keys; gpu: hash table built on the "
                                 "device, probe per batch]")
    gpu.core().vx_node_free(j)
