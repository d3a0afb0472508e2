"""The .vtr file format and scan node (src/vtr1.c, src/scan.c), on the CPU.

Pinned by an independent Python implementation of the layout (tests/vtr_pack.py)
in both directions (its files read by the C scan; the C writer's bytes equal
its bytes), by hand-worked statistics of small row groups, and by row-group
pruning answers worked out from each group's value range."""
import numpy as np
import pytest

from helpers import rng
import vtr_pack as vp


def write_bytes(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return p


def cols_of(r, n):
    k = r.integers(-50, 50, n)
    x = r.normal(size=n)
    x[r.random(n) < 0.05] = np.nan
    b = r.random(n) < 0.3
    s = [f"s{i % 7}" * (i % 3) for i in range(n)]
    return [("k", "i64", k, r.random(n) > 0.1, "key"), ("x", "f64", x, r.random(n) > 0.05, None),
            ("b", "bool", b, None, None), ("s", "str", s, r.random(n) > 0.2, "label")]


@pytest.mark.parametrize("version", [1, 2, 3])
def test_scan_reads_every_version(native, tmp_path, version):
    """files of versions 1-3 with int64 / double / bool / string columns, NAs,
    annotations and an empty row group read back exactly"""
    r = rng(5 + version)
    n = 2500
    cols = cols_of(r, n)
    p = write_bytes(tmp_path, "t.vtr", vp.pack(cols, 700, version, empty_groups=(2,)))
    got = native.collect(native.scan_vtr(p))
    assert list(got) == [c[0] for c in cols]
    for name, kind, vals, valid, _ in cols:
        gv, gm = got[name]
        em = np.ones(n, bool) if valid is None else valid
        assert np.array_equal(gm, em), name
        if kind == "str":
            assert list(gv[em]) == [v for v, ok in zip(vals, em) if ok]
        elif kind == "f64":
            assert np.array_equal(np.asarray(gv[em]).view(np.int64), np.asarray(vals)[em].view(np.int64))
        else:
            assert np.array_equal(np.asarray(gv[em]).astype(np.int64), np.asarray(vals)[em].astype(np.int64))


def test_scan_selected_columns(native, tmp_path):
    """a column subset comes out in file order, skipping string columns"""
    r = rng(11)
    cols = cols_of(r, 1000)
    p = write_bytes(tmp_path, "t.vtr", vp.pack(cols, 300))
    got = native.collect(native.scan_vtr(p, [2, 0]))
    assert list(got) == ["k", "b"]
    assert np.array_equal(got["k"][0][got["k"][1]], cols[0][2][cols[0][3]])


def test_c_writer_bytes_equal_the_layout(native, tmp_path):
    """the C writer's file is byte-for-byte the layout's version-3 file,
    statistics included (ragged last row group)"""
    r = rng(13)
    n = 1234
    k = r.integers(-9, 9, n)
    kv = r.random(n) > 0.2
    x = r.normal(size=n)
    x[::97] = np.nan
    xv = r.random(n) > 0.1
    b = r.random(n) < 0.5
    p = tmp_path / "c.vtr"
    native.vtr_write(p, {"k": (k, kv), "x": (x, xv), "b": b}, rg_rows=500)
    exp = vp.pack([("k", "i64", k, kv, None), ("x", "f64", x, xv, None), ("b", "bool", b, None, None)], 500)
    assert p.read_bytes() == exp


def test_statistics_known_answers():
    """hand-worked statistics of one row group per case"""
    NA = False
    # int64 [NA, 3, -2, NA, 7] -> -2 .. 7
    has, lo, hi = vp.rowgroup_stats("i64", [0, 3, -2, 0, 7], [NA, True, True, NA, True])
    assert (has, lo, hi) == (1, (-2).to_bytes(8, "little", signed=True), (7).to_bytes(8, "little", signed=True))
    # double whose first valid value is NaN: both bounds stay NaN
    has, lo, hi = vp.rowgroup_stats("f64", [np.nan, 1.0, 5.0], None)
    assert has == 1 and np.isnan(np.frombuffer(lo, "<f8")[0]) and np.isnan(np.frombuffer(hi, "<f8")[0])
    # a later NaN never replaces a bound
    has, lo, hi = vp.rowgroup_stats("f64", [1.0, np.nan, 5.0], None)
    assert np.frombuffer(lo + hi, "<f8").tolist() == [1.0, 5.0]
    # all NA: no statistics; bool: (some FALSE, some TRUE) even when all NA
    assert vp.rowgroup_stats("f64", [1.0, 2.0], [NA, NA])[0] == 0
    assert vp.rowgroup_stats("bool", [1, 0, 1], [True, NA, True]) == (1, (0).to_bytes(8, "little"),
                                                                     (1).to_bytes(8, "little"))
    assert vp.rowgroup_stats("bool", [1], [NA])[0] == 1


def test_c_writer_statistics_known_answers(native, tmp_path):
    """the same cases through the C writer, read back from its file"""
    p = tmp_path / "s.vtr"
    x = np.array([np.nan, 1.0, 5.0, 1.0, np.nan, 5.0, 2.0, 2.0])
    k = np.array([0, 3, -2, 0, 9, 9, 9, 9])
    kv = np.array([False, True, True, False, False, False, False, False])
    native.vtr_write(p, {"x": x, "k": (k, kv)}, rg_rows=3)
    h = vp.header(p.read_bytes())
    sx = [g["stats"][0] for g in h["groups"]]
    sk = [g["stats"][1] for g in h["groups"]]
    assert vp.is_nan_bits(sx[0][1]) and vp.is_nan_bits(sx[0][2])          # NaN first
    assert (sx[1][1], sx[1][2]) == (vp.f64_bits(1.0), vp.f64_bits(5.0))   # 1, NaN, 5
    assert (sx[2][1], sx[2][2]) == (vp.f64_bits(2.0), vp.f64_bits(2.0))
    assert sk[0][0] == 1 and np.int64(np.uint64(sk[0][1])) == -2 and sk[0][2] == 3
    assert sk[1][0] == 0 and sk[2][0] == 0                               # all NA


def pruned_groups(native, path, pred):
    """row groups (their g value) with rows in the filter's output when the
    predicate is pushed down onto the scan"""
    f = native.filter_node("oracle", native.scan_vtr(path), pred)
    native.pushdown(f)
    return sorted({int(v) for v in native.collect(f)["g"][0]})


@pytest.fixture
def zone_file(native, tmp_path, oracle):
    """five row groups of 4 rows: x ranges [0,3] [10,13] [-5,-2] [7,7] and a
    NaN-first group [NaN, 20, 21, 22]; g = the group number"""
    x = np.array([0, 1, 2, 3, 10, 11, 12, 13, -5, -4, -3, -2, 7, 7, 7, 7, np.nan, 20, 21, 22], float)
    g = np.repeat(np.arange(5), 4)
    k = (np.nan_to_num(x) * 2).astype(np.int64)
    p = tmp_path / "z.vtr"
    native.vtr_write(p, {"x": x, "g": g, "k": k}, rg_rows=4)
    return p


@pytest.mark.parametrize("pred,groups", [
    (("cmp", ">", ("col", "x"), ("dbl", 5.0)), [1, 3]),                 # group 4 pruned: NaN bounds
    (("cmp", ">=", ("col", "x"), ("dbl", 7.0)), [1, 3]),
    (("cmp", "<", ("col", "x"), ("dbl", 0.5)), [0, 2]),
    (("cmp", "<=", ("col", "x"), ("int", -2)), [2]),
    (("cmp", "==", ("col", "x"), ("dbl", 7.0)), [3]),
    (("cmp", "!=", ("col", "x"), ("dbl", 7.0)), [0, 1, 2, 4]),         # only [7,7] pruned
    (("cmp", ">", ("dbl", 5.0), ("col", "x")), [0, 2]),                  # literal on the left: x < 5
    (("cmp", ">", ("col", "k"), ("dbl", 21.0)), [1, 4]),                 # int64 column, double literal
    (("and", ("cmp", ">", ("col", "x"), ("dbl", 1.0)), ("cmp", "<", ("col", "x"), ("dbl", 8.0))), [0, 3]),
    (("or", ("cmp", "<", ("col", "x"), ("dbl", -4.5)), ("cmp", ">", ("col", "x"), ("dbl", 12.5))), [1, 2]),
])
def test_zone_map_pruning(native, zone_file, pred, groups):
    """row groups whose value range cannot satisfy the comparison are skipped
    (with the reference's NaN-first rule: a group whose statistics are NaN
    fails every range test and is skipped, src/scan.c:77-101)"""
    assert pruned_groups(native, zone_file, pred) == groups


def test_pruning_never_drops_a_matching_row(native, tmp_path, oracle):
    """random int64 / double data without NaN: the filter's result is the
    same with and without pushdown"""
    r = rng(17)
    n = 20000
    x = np.round(np.cumsum(r.normal(size=n)), 2)
    k = r.integers(0, 1000, n) + np.arange(n) // 50
    p = tmp_path / "r.vtr"
    native.vtr_write(p, {"x": x, "k": k}, rg_rows=997)
    for pred in [("cmp", ">", ("col", "x"), ("dbl", float(np.median(x)))),
                 ("cmp", "<=", ("col", "k"), ("int", 150)),
                 ("cmp", "==", ("col", "k"), ("int", 333)),
                 ("or", ("cmp", "<", ("col", "x"), ("dbl", -3.0)), ("cmp", ">=", ("col", "k"), ("int", 390)))]:
        plain = native.collect(native.filter_node("oracle", native.scan_vtr(p), pred))
        pushed = native.collect(native.pushdown(native.filter_node("oracle", native.scan_vtr(p), pred)))
        for c in ("x", "k"):
            assert np.array_equal(plain[c][0], pushed[c][0])


@pytest.mark.parametrize("data,msg", [
    (b"VTR2\x03\x00", "bad magic"),
    (b"VTR1\x04\x00\x00\x00\x00\x00\x00\x00", "unsupported .vtr version: 4"),
    (b"VTR1\x03\x00\x01\x00\x05\x00ab", "unexpected end of file"),
])
def test_bad_files(native, tmp_path, data, msg):
    p = write_bytes(tmp_path, "bad.vtr", data)
    with pytest.raises(native.VectraError, match=msg):
        native.scan_vtr(p)
