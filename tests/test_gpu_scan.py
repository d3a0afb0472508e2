"""GPU group-by nodes over a .vtr scan (or a filter over one) fed straight
from the file to the device: pinned staging, copies on the copy stream, the
filter evaluated inside the group-by feed, row groups pruned with the
scan's pushed-down predicate (src/scan.c:10-138).  Every result must equal
the oracle nodes pulling the same file through the CPU scan: keys, order,
counts and integer aggregates and min / max exact, double sums and means
within 1e-12 * sum|x|."""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_tables_equal, dbl_with_specials, rng, with_na
import vtr_pack as vp

pytestmark = pytest.mark.gpu

AGGS = [("n", "n", None, 0), ("cx", "count", "x", 0), ("sx", "sum", "x", 1), ("mx", "mean", "x", 1),
        ("lo", "min", "x", 1), ("hi", "max", "v", 0), ("sv", "sum", "v", 0)]


def table(seed, n):
    r = rng(seed)
    k = r.integers(-20, 280, n)
    x = dbl_with_specials(r, n, 0.01, 0.02)
    x[0] = np.nan                                   # the first row group's statistics are NaN
    v = r.integers(-2**40, 2**40, n)
    b = r.random(n) < 0.4
    return {"k": (k, with_na(r, n, 0.03)), "x": (x, with_na(r, n, 0.02)), "v": v, "b": b,
            "ax": np.abs(np.nan_to_num(x, nan=0.0))}


def scan_groups(gpu):
    return gpu.nodes().vx_scan_feed_groups()


def run(gpu, engine, path, pred, keys, temp_dir, pushdown, aggs=AGGS):
    node = gpu.scan_vtr(path)
    if pred is not None:
        node = gpu.filter_node(engine, node, pred)
        if pushdown:
            gpu.pushdown(node)
    return gpu.collect(gpu.group_agg_node(engine, node, keys, aggs, temp_dir))


def closes(exp):
    sa = exp.pop("__sa")[0]
    cnt = exp.pop("__cx")[0]
    return {"sx": sa, "mx": sa / np.maximum(cnt, 1)}


@pytest.fixture(scope="module")
def feed_symbol(gpu):
    lib = gpu.nodes()
    lib.vx_scan_feed_groups.restype = C.c_int64
    lib.vx_scan_feed_groups.argtypes = []
    return lib


@pytest.mark.parametrize("temp_dir,pushdown,pred", [
    ("/tmp", True, ("cmp", ">", ("col", "x"), ("dbl", 0.25))),
    (None, True, ("cmp", ">", ("col", "x"), ("dbl", 0.25))),
    ("/tmp", False, ("cmp", "<=", ("col", "k"), ("int", 100))),
    ("/tmp", True, ("and", ("cmp", ">=", ("col", "k"), ("int", 0)), ("cmp", "==", ("col", "b"), ("lgl", True)))),
    ("/tmp", True, None),
])
def test_group_over_vtr_scan(gpu, oracle, feed_symbol, tmp_path, temp_dir, pushdown, pred):
    t = table(41, 90_000)
    p = tmp_path / "t.vtr"
    gpu.vtr_write(p, t, rg_rows=16_384)
    before = feed_symbol.vx_scan_feed_groups()
    got = run(gpu, "gpu", p, pred, ["k"], temp_dir, pushdown)
    assert feed_symbol.vx_scan_feed_groups() > before          # the device path ran
    exp = run(gpu, "oracle", p, pred, ["k"], temp_dir, pushdown,
              AGGS + [("__sa", "sum", "ax", 1), ("__cx", "count", "x", 0)])
    close = closes(exp)
    assert_tables_equal(got, exp, close)


def test_device_feed_off_gives_the_same(gpu, oracle, feed_symbol, tmp_path):
    """knob scan.device_feed = 1: the node pulls the scan's batches instead"""
    t = table(43, 50_000)
    p = tmp_path / "t.vtr"
    gpu.vtr_write(p, t, rg_rows=10_000)
    pred = ("cmp", "<", ("col", "x"), ("dbl", 0.0))
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"scan.device_feed")
    lib.vgpu_tune_set(b"scan.device_feed", 1)
    try:
        before = feed_symbol.vx_scan_feed_groups()
        off = run(gpu, "gpu", p, pred, ["k"], "/tmp", True)
        assert feed_symbol.vx_scan_feed_groups() == before
    finally:
        lib.vgpu_tune_set(b"scan.device_feed", old)
    on = run(gpu, "gpu", p, pred, ["k"], "/tmp", True)
    exp = run(gpu, "oracle", p, pred, ["k"], "/tmp", True, AGGS + [("__sa", "sum", "ax", 1), ("__cx", "count", "x", 0)])
    close = closes(exp)
    assert_tables_equal(off, exp, close)
    assert_tables_equal(on, exp, close)


def test_ungrouped_and_strings_and_empty_groups(gpu, oracle, feed_symbol, tmp_path):
    """summarise without keys over a file with a string column (not read)
    and empty row groups, written in the layout's version 2 (no statistics)
    and 3"""
    r = rng(47)
    n = 30_000
    x = r.normal(size=n)
    xv = r.random(n) > 0.05
    k = r.integers(0, 50, n)
    s = [f"r{i % 13}" for i in range(n)]
    cols = [("s", "str", s, None, None), ("k", "i64", k, None, None), ("x", "f64", x, xv, "x"),
            ("ax", "f64", np.abs(x), None, None)]
    aggs = [("n", "n", None, 0), ("sx", "sum", "x", 1), ("lo", "min", "x", 1)]
    for version in (2, 3):
        p = tmp_path / f"s{version}.vtr"
        p.write_bytes(vp.pack(cols, 7_000, version, empty_groups=(0, 3)))
        pred = ("cmp", ">", ("col", "k"), ("int", 10))
        for keys in ([], ["k"]):
            before = feed_symbol.vx_scan_feed_groups()
            got = run(gpu, "gpu", p, pred, keys, "/tmp", True, aggs)
            assert feed_symbol.vx_scan_feed_groups() > before
            exp = run(gpu, "oracle", p, pred, keys, "/tmp", True, aggs + [("__sa", "sum", "ax", 1)])
            sa = exp.pop("__sa")[0]
            assert_tables_equal(got, exp, {"sx": sa})


def test_non_boolean_filter_is_an_error(gpu, oracle, tmp_path):
    t = table(53, 1000)
    p = tmp_path / "t.vtr"
    gpu.vtr_write(p, t, rg_rows=300)
    with pytest.raises(gpu.VectraError, match="filter predicate must evaluate to boolean"):
        run(gpu, "gpu", p, ("arith", "+", ("col", "x"), ("dbl", 1.0)), ["k"], "/tmp", False)
