"""group_by over several columns through the GPU GroupAggNode against the
oracle node (src/group_agg.c with n_keys > 1, key equality per column as
src/hash.c:48-78): int64 / double / bool key columns with NA, NaN (every
NaN tuple its own group), -0 / +0 (one group), INT64_MIN and values near
2^63 (exact int64 keys), first-occurrence and key order, host row groups
with and without a filter.  Keys, order, counts, int64 sums and min / max
exact; double sums and means within 1e-12 * sum|x|."""
import numpy as np
import pytest

from helpers import assert_tables_equal, dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu

AGGS = [("n", "n", None, 0), ("sx", "sum", "x", 1), ("mx", "mean", "x", 1), ("lo", "min", "x", 1),
        ("hv", "max", "v", 0), ("cv", "count", "v", 0)]


def table(seed, n, kinds, nan=True):
    r = rng(seed)
    cols = {}
    for i, kind in enumerate(kinds):
        name = f"k{i}"
        if kind == "i64":
            k = r.integers(0, 6, n) * (2**61) + r.integers(-3, 3, n)     # values near +-2^62, 2^63
            k[r.random(n) < 0.01] = np.iinfo(np.int64).min
            cols[name] = (k, with_na(r, n, 0.03))
        elif kind == "f64":
            k = np.round(r.normal(size=n) * 2) / 2
            u = r.random(n)
            if nan:
                k[u < 0.004] = np.nan
            z = (u > 0.5) & (u < 0.53)
            k[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
            cols[name] = (k, with_na(r, n, 0.03))
        else:
            cols[name] = (r.random(n) < 0.3, with_na(r, n, 0.05))
    x = dbl_with_specials(r, n, 0.005, 0.01)
    cols["x"] = (x, with_na(r, n, 0.02))
    cols["v"] = (r.integers(-2**40, 2**40, n), with_na(r, n, 0.02))
    cols["ax"] = np.abs(np.nan_to_num(x, nan=0.0))
    return cols


def run(nv, engine, cols, keys, temp_dir, rg_rows, pred, aggs):
    node = nv.Table(cols, rg_rows=rg_rows).scan()
    if pred is not None:
        node = nv.filter_node(engine, node, pred)
    return nv.collect(nv.group_agg_node(engine, node, keys, aggs, temp_dir))


@pytest.mark.parametrize("kinds,temp_dir,pred", [
    (("i64", "f64"), "/tmp", None),
    (("i64", "f64"), None, None),
    (("f64", "bool", "i64"), "/tmp", ("cmp", ">", ("col", "x"), ("dbl", -0.5))),
    (("bool", "bool"), None, ("cmp", "<", ("col", "v"), ("int", 0))),
    (("i64", "i64"), "/tmp", None),
])
def test_multikey_matches_oracle(gpu, oracle, kinds, temp_dir, pred):
    # NaN key values only on the hash path: with temp_dir the reference's
    # merge sort compares NaN as equal to every key, which is not an order
    cols = table(61 + len(kinds), 120_000, kinds, nan=temp_dir is None)
    keys = [f"k{i}" for i in range(len(kinds))]
    got = run(gpu, "gpu", cols, keys, temp_dir, 25_000, pred, AGGS)
    exp = run(gpu, "oracle", cols, keys, temp_dir, 25_000, pred, AGGS + [("__sa", "sum", "ax", 1),
                                                                        ("__cx", "count", "x", 0)])
    sa = exp.pop("__sa")[0]
    cnt = exp.pop("__cx")[0]
    assert_tables_equal(got, exp, {"sx": sa, "mx": sa / np.maximum(cnt, 1)})


def test_multikey_too_many_aggregates(gpu):
    cols = table(71, 100, ("i64", "i64", "i64", "i64"))
    aggs = [(f"a{i}", "sum", "x", 0) for i in range(4)]
    with pytest.raises(gpu.VectraError, match="aggregates over 4 grouping columns"):
        run(gpu, "gpu", cols, ["k0", "k1", "k2", "k3"], "/tmp", 0, None, aggs)


def join_tables(nv, seed, n_left, n_right, card, ltypes, rtypes):
    """left (a, b, x) and right (y, a, b, w) with two key columns each;
    duplicates on both sides, NA keys, NaN and -0 / +0 in double keys"""
    def mk():
        r = rng(seed)

        def key(n, t):
            k = r.integers(0, card, n).astype(t)
            if t == np.float64:
                k[r.random(n) < 0.02] = np.nan
                z = k == 0
                k[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
            return k
        la, lb = key(n_left, ltypes[0]), key(n_left, ltypes[1])
        ra, rb = key(n_right, rtypes[0]), key(n_right, rtypes[1])
        left = nv.Table({"a": (la, with_na(r, n_left, 0.04)), "b": (lb, with_na(r, n_left, 0.04)),
                         "x": r.normal(size=n_left)}, rg_rows=900)
        right = nv.Table({"y": (r.integers(0, 10**6, n_right), with_na(r, n_right, 0.05)),
                          "a": (ra, with_na(r, n_right, 0.04)), "b": (rb, with_na(r, n_right, 0.04)),
                          "w": r.random(n_right)}, rg_rows=777)
        return left, right
    return mk


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
@pytest.mark.parametrize("types", [((np.int64, np.float64), (np.int64, np.float64)),
                                   ((np.int64, np.int64), (np.float64, np.int64))])
def test_multikey_join_matches_oracle(gpu, oracle, kind, types):
    """joins on two key columns (a, b): rows and order as the oracle's, the
    second case coercing the left int64 `a` to the right's double"""
    mk = join_tables(gpu, 83, 6000, 2500, 12, types[0], types[1])
    if kind in ("left", "full") and types[0][0] != types[1][0]:
        # the reference's chain walk compares the uncoerced probe key's bytes
        # (src/join.c:398-415); the GPU path declines such joins explicitly
        left, right = mk()
        node = gpu.join_node("gpu", left.scan(), right.scan(), kind, [(0, 1), (1, 2)])
        with pytest.raises(gpu.VectraError, match="coerced to double is not evaluated on the GPU"):
            gpu.next_batches(node)
        return
    out = {}
    for engine in ("gpu", "oracle"):
        left, right = mk()
        node = gpu.join_node(engine, left.scan(), right.scan(), kind, [(0, 1), (1, 2)])
        out[engine] = gpu.next_batches(node)
    got, exp = out["gpu"], out["oracle"]
    assert len(got) == len(exp), f"{len(got)} batches vs {len(exp)}"
    for g, e in zip(got, exp):
        assert g["sel"] is None and e["sel"] is None
        assert_tables_equal(g["cols"], e["cols"])


def test_multikey_key_order_many_groups(gpu, oracle):
    """key order over about 150,000 groups of (int64, double, bool) tuples
    with NA in every column: the group order is sorted on the device (a
    stable sort per column and class), against the oracle's merge sort"""
    r = rng(97)
    n = 400_000
    a = r.integers(-5000, 5000, n) * 3
    b = np.round(r.normal(size=n) * 8) / 4
    z = r.random(n) < 0.05
    b[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
    cols = {"a": (a, with_na(r, n, 0.02)), "b": (b, with_na(r, n, 0.02)), "c": (r.random(n) < 0.5, with_na(r, n, 0.1)),
            "x": r.normal(size=n), "v": r.integers(-2**40, 2**40, n)}
    cols["ax"] = np.abs(cols["x"])
    aggs = [("n", "n", None, 0), ("sx", "sum", "x", 0), ("hv", "max", "v", 0)]
    got = run(gpu, "gpu", cols, ["a", "b", "c"], "/tmp", 100_000, None, aggs)
    exp = run(gpu, "oracle", cols, ["a", "b", "c"], "/tmp", 100_000, None, aggs + [("__sa", "sum", "ax", 0)])
    assert len(got["n"][0]) > 100_000
    sa = exp.pop("__sa")[0]
    assert_tables_equal(got, exp, {"sx": sa})


def test_pairs_unequal_counts_differing_tuples(gpu):
    """the device pair check of joins on tuple codes: a pair counts when a
    column is NA on either side or its values differ (double by ==: -0
    equals +0, NaN equals nothing); pairs without a build row are skipped"""
    import ctypes as C
    lib = gpu.vgpu()
    pa = gpu.DeviceColumn(np.array([1, 2, 3, 4, 5, 6]), np.array([1, 1, 1, 0, 1, 1], bool))
    pb = gpu.DeviceColumn(np.array([0.0, 1.5, -0.0, 2.0, np.nan, 7.0]))
    ba = gpu.DeviceColumn(np.array([1, 2, 3, 4, 5, 9]))
    bb = gpu.DeviceColumn(np.array([0.0, 1.5, 0.0, 2.0, np.nan, 7.0]))
    pr = np.array([0, 1, 2, 3, 4, 5, 5], np.int32)
    br = np.array([0, 1, 2, 3, 4, 5, -1], np.int64)
    dpr, dbr = lib.vgpu_malloc(pr.nbytes), lib.vgpu_malloc(br.nbytes)
    assert lib.vgpu_memcpy_h2d(dpr, pr.ctypes.data, pr.nbytes) == 0
    assert lib.vgpu_memcpy_h2d(dbr, br.ctypes.data, br.nbytes) == 0
    P = (gpu.VgpuCol * 2)(pa.col, pb.col)
    B = (gpu.VgpuCol * 2)(ba.col, bb.col)
    out = C.c_int64(-1)
    assert lib.vgpu_pairs_unequal(P, B, 2, dpr, dbr, len(pr), C.byref(out)) == 0, lib.vgpu_last_error()
    # pair 3: probe NA; pair 4: NaN; pair 5: 6 vs 9; pair 6: no build row; pair 2: -0 == +0
    assert out.value == 3
    assert lib.vgpu_pairs_unequal(P, B, 1, dpr, dbr, 3, C.byref(out)) == 0 and out.value == 0
    for c in (pa, pb, ba, bb):
        c.free()
    lib.vgpu_free(dpr)
    lib.vgpu_free(dbr)
