"""explain() plan text (src/r_bridge.c:575-698) over oracle node trees: the
reference's line format and annotations, on the CPU.  The GPU nodes' notes
are checked in tests/test_gpu_explain.py."""
import numpy as np

from helpers import rng


def test_filter_group_plan(native, oracle):
    t = native.Table({"k": np.arange(10), "x": np.ones(10), "s": np.array(["a"] * 10)})
    f = native.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 0.0)))
    g = native.group_agg_node("oracle", f, ["k", "s"], [("n", "n", None, 0)])
    assert native.explain(g) == ["GroupAggNode [materializes, 2 keys]", "  FilterNode [streaming]",
                                 "    MemScanNode [in memory, 3 cols]"]
    native.core().vx_node_free(g)


def test_join_plan(native, oracle):
    left = native.Table({"k": np.arange(4), "x": np.ones(4)})
    right = native.Table({"k": np.arange(3)})
    for kind in ("inner", "left", "full", "semi", "anti"):
        j = native.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)])
        assert native.explain(j) == [f"JoinNode [build right + stream left, {kind}, 1 keys]",
                                     "  MemScanNode [in memory, 2 cols]", "  MemScanNode [in memory, 1 cols]"]
        native.core().vx_node_free(j)


def test_scan_plan(native, oracle, tmp_path):
    r = rng(3)
    p = tmp_path / "e.vtr"
    native.vtr_write(p, {"k": r.integers(0, 9, 100), "x": r.normal(size=100)}, rg_rows=40)
    s = native.scan_vtr(p)
    assert native.explain(s) == ["ScanNode [streaming, 2 cols, v3 stats]"]
    native.core().vx_node_free(s)
    s = native.scan_vtr(p, [1])
    assert native.explain(s) == ["ScanNode [streaming, 1/2 cols (pruned), v3 stats]"]
    native.core().vx_node_free(s)
    f = native.pushdown(native.filter_node("oracle", native.scan_vtr(p), ("cmp", ">", ("col", "x"), ("dbl", 0.0))))
    assert native.explain(f) == ["FilterNode [streaming]", "  ScanNode [streaming, 2 cols, predicate pushdown, v3 stats]"]
    native.core().vx_node_free(f)


def test_length_and_truncation(native, oracle):
    t = native.Table({"k": np.arange(3)})
    s = t.scan()
    lib = native.nodes()
    import ctypes as C
    n = lib.vx_explain(s, None, 0)
    assert n == len("MemScanNode [in memory, 1 cols]\n")
    buf = C.create_string_buffer(8)
    assert lib.vx_explain(s, buf, 8) == n and buf.value == b"MemScan"
    native.core().vx_node_free(s)
