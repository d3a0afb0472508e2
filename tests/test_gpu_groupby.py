"""GPU GroupAggNode vs the oracle (src/group_agg.c, src/agg_ops.c, src/hash.c).

Keys, group order, counts, int64 sum/min/max and double min/max (including
NaN-first and signed-zero cases) must be bit-exact; double sum and mean
within 1e-12 * sum|x| of the group.
"""
import numpy as np
import pytest

from helpers import assert_tables_equal, dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu

ALL_AGGS = [
    ("n", "n", None, 0),
    ("cx", "count", "x", 0),
    ("sx", "sum", "x", 0), ("sx_rm", "sum", "x", 1),
    ("mx", "mean", "x", 0), ("mx_rm", "mean", "x", 1),
    ("lo", "min", "x", 0), ("lo_rm", "min", "x", 1),
    ("hi", "max", "x", 0), ("hi_rm", "max", "x", 1),
    ("sv", "sum", "v", 1), ("mv", "mean", "v", 1), ("lov", "min", "v", 1), ("hiv", "max", "v", 0),
    ("sb", "sum", "b", 1), ("lob", "min", "b", 1),
]
CLOSE = {"sx", "sx_rm", "mx", "mx_rm"}


def table(nv, n, seed, key_kind, card, rg, p_na=0.03, p_nan=0.02):
    r = rng(seed)
    if key_kind == "int":
        kv = r.integers(-card, card, n)
        kv[r.random(n) < 0.01] = np.iinfo(np.int64).min
        key = (kv, with_na(r, n, p_na))
    elif key_kind == "dbl":
        kv = np.round(r.normal(size=n) * card) / 4.0
        u = r.random(n)
        kv[u < p_nan] = np.nan
        z = (u > 0.5) & (u < 0.52)
        kv[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
        key = (kv, with_na(r, n, p_na))
    elif key_kind == "bool":
        key = (r.random(n) < 0.3, with_na(r, n, p_na))
    else:
        key = np.zeros(n, np.int64)
    x = dbl_with_specials(r, n, p_nan, 0.03)
    cols = {
        "g": key,
        "x": (x, with_na(r, n, p_na)),
        "v": (r.integers(-10**12, 10**12, n), with_na(r, n, p_na)),
        "b": (r.random(n) < 0.5, with_na(r, n, p_na)),
        "ax": np.abs(np.nan_to_num(x, nan=0.0)),
    }
    return nv.Table(cols, rg_rows=rg)


def run(nv, engine, t, keys, aggs, temp_dir, pred=None):
    node = t.scan()
    if pred is not None:
        node = nv.filter_node(engine, node, pred)
    return nv.collect(nv.group_agg_node(engine, node, keys, aggs, temp_dir))


def check(nv, mk, keys, aggs, temp_dir, pred=None):
    got = run(nv, "gpu", mk(), keys, aggs, temp_dir, pred)
    extra = [("__sa", "sum", "ax", 1), ("__cnt", "count", "x", 0)]
    exp = run(nv, "oracle", mk(), keys, list(aggs) + extra, temp_dir, pred)
    sa = exp.pop("__sa")[0]
    cnt = exp.pop("__cnt")[0]
    close = {}
    for name, kind, col, _ in aggs:
        if name in CLOSE:
            close[name] = sa if kind == "sum" else sa / np.maximum(cnt, 1)
    assert_tables_equal(got, exp, close)
    return got


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
@pytest.mark.parametrize("key_kind,card", [("int", 40), ("int", 30000), ("bool", 0), ("dbl", 8)])
@pytest.mark.parametrize("rg", [0, 3001])
def test_groupby_matches_oracle(gpu, oracle, temp_dir, key_kind, card, rg):
    if key_kind == "dbl" and temp_dir is not None:
        mk = lambda: table(gpu, 40000, 11, key_kind, card, rg, p_nan=0.0)
    else:
        mk = lambda: table(gpu, 40000, 11, key_kind, card, rg)
    check(gpu, mk, ["g"], ALL_AGGS, temp_dir)


@pytest.mark.parametrize("prefetch,partials", [(1, 2), (2, 2), (2, 1)])
@pytest.mark.parametrize("key_kind,card,n", [("int", 40, 60000), ("dbl", 8, 60000), ("int", 300, 2_000_000)])
def test_groupby_feed_prefetch(gpu, oracle, prefetch, partials, key_kind, card, n):
    """the general feed with a tile's words loaded before its rows are
    applied (gb.feed.prefetch 2) and row by row (1), its blocks merged
    through folded partials (gb.merge.partials 2) and by HBM atomics (1),
    against the oracle, after a filter (NA keys and inputs, NaN and signed
    zeros)"""
    lib = gpu.vgpu()
    knobs = (b"gb.feed.prefetch", b"gb.merge.partials")
    old = [lib.vgpu_tune_get(k) for k in knobs]
    for k, v in zip(knobs, (prefetch, partials)):
        assert lib.vgpu_tune_set(k, v) == 0
    try:
        mk = lambda: table(gpu, n, 17, key_kind, card, 0 if n > 100000 else 7001)
        check(gpu, mk, ["g"], ALL_AGGS, None, pred=("cmp", ">", ("col", "ax"), ("dbl", 0.2)))
    finally:
        for k, v in zip(knobs, old):
            lib.vgpu_tune_set(k, v)


@pytest.mark.parametrize("dense", [2, 1])
@pytest.mark.parametrize("key_kind,card", [("int", 8), ("int", 40), ("bool", 0)])
@pytest.mark.parametrize("filtered", [False, True])
def test_groupby_feed_dense(gpu, oracle, dense, key_kind, card, filtered):
    """the general feed with a dense LDS table (gb.feed.dense 2: the keys of
    a short sketched range own record key - min, no hashing) and with a
    hashed one (1), against the oracle: one feed of 2M rows (sketched), NA
    keys and the int64 key equal to the hash tables' empty marker (reserved
    records), NA / NaN / signed-zero inputs, every aggregate, with and
    without a filter"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"gb.feed.dense")
    assert lib.vgpu_tune_set(b"gb.feed.dense", dense) == 0
    try:
        mk = lambda: table(gpu, 2_000_000, 41 + card, key_kind, card, 0)
        pred = ("cmp", ">", ("col", "ax"), ("dbl", 0.2)) if filtered else None
        check(gpu, mk, ["g"], ALL_AGGS, None, pred=pred)
    finally:
        lib.vgpu_tune_set(b"gb.feed.dense", old)


@pytest.mark.parametrize("partials", [1, 2])
@pytest.mark.parametrize("key_kind,card", [("int", 40), ("int", 3000), ("dbl", 16)])
def test_groupby_narrow_partials(gpu, oracle, partials, key_kind, card):
    """the narrow feed (dense and hashed LDS tables) with its blocks merged
    through folded partials and by HBM atomics, against the oracle, one
    feed of 2M rows (enough blocks for the partials)"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"gb.merge.partials")
    assert lib.vgpu_tune_set(b"gb.merge.partials", partials) == 0
    try:
        mk = lambda: table(gpu, 2_000_000, 23, key_kind, card, 0)
        check(gpu, mk, ["g"], NARROW_AGGS, None)
    finally:
        lib.vgpu_tune_set(b"gb.merge.partials", old)


# n / count / sum / mean of at most two inputs: the narrow feed kernel
NARROW_AGGS = [("n", "n", None, 0), ("cx", "count", "x", 0), ("sx", "sum", "x", 0), ("sx_rm", "sum", "x", 1),
               ("mx_rm", "mean", "x", 1), ("sv", "sum", "v", 1), ("mv", "mean", "v", 0)]


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
@pytest.mark.parametrize("key_kind,card", [("int", 40), ("int", 30000), ("bool", 0), ("dbl", 8), ("none", 0)])
@pytest.mark.parametrize("rg", [0, 3001])
def test_groupby_narrow(gpu, oracle, temp_dir, key_kind, card, rg):
    """the narrow kernel: NA / NaN / +-0 / INT64_MIN keys, NA inputs, one
    batch or many, with and without a filter in front"""
    p_nan = 0.0 if key_kind == "dbl" and temp_dir is not None else 0.02
    mk = lambda: table(gpu, 40000, 31, key_kind, card, rg, p_nan=p_nan)
    keys = [] if key_kind == "none" else ["g"]
    check(gpu, mk, keys, NARROW_AGGS, temp_dir)
    check(gpu, mk, keys, NARROW_AGGS, temp_dir, ("cmp", ">", ("col", "x"), ("dbl", 0.1)))


@pytest.mark.parametrize("card,temp_dir", [(300, "/tmp"), (300, None), (3000, "/tmp"), (10**9, None)])
def test_groupby_narrow_large_feed(gpu, oracle, card, temp_dir):
    """one feed of 1.2M rows: the sketch's key range makes a dense LDS table
    for a few hundred or thousand int64 keys (NA and INT64_MIN keys take
    their reserved records), a hashed one otherwise"""
    mk = lambda: table(gpu, 1_200_000, 37, "int", card, 0, p_na=0.01, p_nan=0.01)
    check(gpu, mk, ["g"], NARROW_AGGS, temp_dir)
    check(gpu, mk, ["g"], NARROW_AGGS, temp_dir, ("cmp", ">", ("col", "x"), ("dbl", -0.2)))


@pytest.mark.parametrize("rg", [0, 5000])
def test_ungrouped_summarise(gpu, oracle, rg):
    mk = lambda: table(gpu, 30000, 5, "none", 0, rg)
    check(gpu, mk, [], ALL_AGGS, "/tmp")


def test_ungrouped_empty_input(gpu, oracle):
    mk = lambda: gpu.Table({"x": np.zeros(0), "ax": np.zeros(0)})
    got = check(gpu, mk, [], [("n", "n", None, 0), ("s", "sum", "x", 0)], None)
    assert len(got["n"][0]) == 0


def test_groupby_after_filter(gpu, oracle):
    mk = lambda: table(gpu, 50000, 21, "int", 500, 4096)
    pred = ("cmp", ">", ("col", "x"), ("dbl", 0.0))
    check(gpu, mk, ["g"], ALL_AGGS, "/tmp", pred)
    check(gpu, mk, ["g"], ALL_AGGS, None, pred)


def test_groupby_nan_first_and_signed_zero(gpu, oracle):
    """min/max order rules of src/agg_ops.c:118-150: a NaN fed first sticks,
    later NaNs are ignored, and a zero result keeps the first zero's sign."""
    x = np.array([np.nan, 1.0, -2.0, 3.0, 0.0, -0.0, 5.0, -0.0, 0.0, 2.0, np.nan, 7.0])
    g = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5])
    mk = lambda: gpu.Table({"g": g, "x": x, "ax": np.abs(np.nan_to_num(x))}, rg_rows=5)
    aggs = [("lo", "min", "x", 0), ("hi", "max", "x", 0), ("n", "n", None, 0)]
    for td in (None, "/tmp"):
        got = check(gpu, mk, ["g"], aggs, td)
        lo = got["lo"][0]
        assert np.isnan(lo[0]) and lo[1] == -2.0
        assert np.signbit(lo[2]) is np.False_ and np.signbit(lo[3]) is np.True_
        hi = got["hi"][0]
        assert hi[4] == 2.0 and np.signbit(hi[2]) is np.False_


def test_groupby_config1(gpu, oracle):
    """filter(x > 0) |> group_by(key) |> summarise(total = sum(value), n = n(), avg = mean(value)) on 1M rows / 10 RGs."""
    n = 1_000_000
    r = rng(7)
    cols = {"id": np.arange(n), "key": r.integers(0, 1000, n), "x": r.normal(size=n), "value": r.uniform(0, 100, n)}
    mk = lambda: gpu.Table(cols, rg_rows=100_000)
    aggs = [("total", "sum", "value", 0), ("n", "n", None, 0), ("avg", "mean", "value", 0)]
    pred = ("cmp", ">", ("col", "x"), ("dbl", 0.0))
    got = run(gpu, "gpu", mk(), ["key"], aggs, "/tmp", pred)
    exp = run(gpu, "oracle", mk(), ["key"], aggs + [("sa", "sum", "value", 0)], "/tmp", pred)
    sa = exp.pop("sa")[0]
    assert_tables_equal(got, exp, {"total": sa, "avg": sa / exp["n"][0]})
    assert list(got["key"][0]) == list(range(1000))
    # fsum oracle: exact group sums
    import math
    keep = cols["x"] > 0
    for k in (0, 1, 499, 999):
        sel = keep & (cols["key"] == k)
        ex = math.fsum(cols["value"][sel])
        assert abs(got["total"][0][k] - ex) <= 1e-12 * ex


def test_groupby_known_answers(gpu):
    """test-groupby.R:1-99 and test-agg.R:1-139, string keys recoded as
    integers in the same first-seen / sorted order."""
    nv = gpu

    def q(cols, aggs, td=None, keys=("g",)):
        return nv.collect(nv.group_agg_node("gpu", nv.Table(cols).scan(), list(keys), aggs, td))

    out = q({"g": np.array([0, 1, 0, 1, 0]), "x": np.arange(1, 6)}, [("cnt", "n", None, 0)])
    assert list(out["g"][0]) == [0, 1] and list(out["cnt"][0]) == [3, 2]
    out = q({"g": np.array([0, 0, 1, 1]), "x": np.array([1., 2., 3., 4.])}, [("total", "sum", "x", 0)])
    assert list(out["total"][0]) == [3, 7]
    out = q({"g": np.array([0, 0, 1, 1]), "x": np.array([10., 20., 30., 40.])}, [("avg", "mean", "x", 0)])
    assert list(out["avg"][0]) == [15, 35]
    out = q({"g": np.array([0, 0, 1, 1]), "x": np.array([5., 1., 8., 3.])},
            [("lo", "min", "x", 0), ("hi", "max", "x", 0)])
    assert list(out["lo"][0]) == [1, 3] and list(out["hi"][0]) == [5, 8]
    out = q({"g": (np.array([0, 0, 0, 0]), np.array([1, 0, 1, 0], bool)), "x": np.array([1., 2., 3., 4.])},
            [("total", "sum", "x", 0)], td="/tmp")
    assert list(out["g"][1]) == [True, False] and list(out["total"][0]) == [4, 6]
    out = q({"g": np.array([0, 0, 1]), "x": (np.array([1., 0., 3.]), np.array([1, 0, 1], bool))},
            [("total", "sum", "x", 1), ("s", "sum", "x", 0)])
    assert list(out["total"][0]) == [1, 3] and list(out["s"][1]) == [False, True]
    out = q({"g": np.array([0, 0]), "x": (np.zeros(2), np.zeros(2, bool))},
            [("s", "sum", "x", 1), ("m", "mean", "x", 1), ("s0", "sum", "x", 0)])
    assert out["s"][0][0] == 0 and np.isnan(out["m"][0][0]) and not out["s0"][1][0]
    nv.core().vx_clear_warnings()
    out = q({"g": np.array([0, 0]), "x": (np.zeros(2), np.zeros(2, bool))},
            [("lo", "min", "x", 1), ("hi", "max", "x", 1)])
    assert out["lo"][0][0] == np.inf and out["hi"][0][0] == -np.inf
    assert nv.core().vx_warning_count() == 2
    out = q({"g": np.array([0, 0, 0]), "x": (np.array([1., 0., 3.]), np.array([1, 0, 1], bool))},
            [("cnt", "n", None, 0), ("m", "mean", "x", 1)])
    assert out["cnt"][0][0] == 3 and out["m"][0][0] == 2
    out = q({"g": np.array([0, 0, 1]), "x": np.array([100, 200, 50])}, [("s", "sum", "x", 0)])
    assert list(out["s"][0]) == [300, 50]
    with pytest.raises(nv.VectraError, match="summarise: column not found: zz"):
        q({"g": np.array([0]), "x": np.array([1.])}, [("s", "sum", "zz", 0)])


def test_groupby_high_cardinality_growth(gpu, oracle):
    """many more groups than the initial table: exercises on-device rehash"""
    mk = lambda: table(gpu, 300000, 3, "int", 10**9, 20000, p_na=0.001, p_nan=0.001)
    check(gpu, mk, ["g"], [("n", "n", None, 0), ("sv", "sum", "v", 0), ("lo", "min", "x", 0),
                           ("sx", "sum", "x", 0)], None)


@pytest.mark.parametrize("key_kind,card,temp_dir", [("int", 8, "/tmp"), ("int", 10**9, "/tmp"),
                                                    ("int", 3000, None), ("dbl", 64, None),
                                                    ("int", 300, "/tmp"), ("int", 1000, None)])
def test_groupby_large_single_feed(gpu, oracle, key_kind, card, temp_dir):
    """one batch of more than 2^20 rows and no group-count hint: a key sketch
    sizes the table and picks the path -- replicated LDS tables for a few
    keys, a denser LDS table when the tuned load does not fit, the
    partitioned path when no LDS table holds the wide records"""
    mk = lambda: table(gpu, 1_200_000, 11, key_kind, card, 0, p_na=0.01, p_nan=0.01)
    check(gpu, mk, ["g"], ALL_AGGS, temp_dir)


@pytest.fixture
def part_threshold(gpu):
    """lowers the group count from which a large first feed is partitioned"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"gb.part.min_groups")
    assert lib.vgpu_tune_set(b"gb.part.min_groups", 1000) == 0
    yield
    lib.vgpu_tune_set(b"gb.part.min_groups", old)


@pytest.mark.parametrize("key_kind,card,temp_dir", [("int", 3000, "/tmp"), ("int", 3000, None),
                                                    ("int", 10**9, None), ("dbl", 64, "/tmp"),
                                                    ("dbl", 4000, None)])
def test_groupby_partitioned(gpu, oracle, part_threshold, key_kind, card, temp_dir):
    """one batch of 1.2M rows hash-partitioned into LDS-sized partitions
    (NA keys, INT64_MIN keys, NaN keys, signed zeros, NA / NaN inputs)"""
    # NaN keys only on the hash path: the reference's sorted path splits
    # groups around NaN keys (its merge sort ranks NaN equal to every key)
    p_nan = 0.0 if key_kind == "dbl" and temp_dir is not None else 0.01
    mk = lambda: table(gpu, 1_200_000, 17, key_kind, card, 0, p_na=0.02, p_nan=p_nan)
    check(gpu, mk, ["g"], ALL_AGGS, temp_dir)


TWO_INPUT_AGGS = [a for a in ALL_AGGS if a[2] in (None, "x", "v")]


@pytest.mark.parametrize("key_kind,card,temp_dir", [("int", 3000, "/tmp"), ("int", 10**9, None),
                                                    ("dbl", 4000, None)])
def test_groupby_partitioned_two_inputs(gpu, oracle, part_threshold, key_kind, card, temp_dir):
    """the two-input staged row (four words), whose passes batch four rows
    per thread: NA / INT64_MIN / NaN keys, signed zeros, NA and NaN inputs"""
    p_nan = 0.0 if key_kind == "dbl" and temp_dir is not None else 0.01
    mk = lambda: table(gpu, 1_200_000, 19, key_kind, card, 0, p_na=0.02, p_nan=p_nan)
    check(gpu, mk, ["g"], TWO_INPUT_AGGS, temp_dir)
    check(gpu, mk, ["g"], TWO_INPUT_AGGS, temp_dir, ("cmp", ">", ("col", "x"), ("dbl", -0.3)))


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
def test_groupby_partitioned_hot_keys(gpu, oracle, part_threshold, temp_dir):
    """partitions far above one item (a hot key, a 30% NA key): partial
    records of several blocks are combined"""
    n = 1_300_000
    r = rng(23)
    kv = r.integers(-10**12, 10**12, n)
    u = r.random(n)
    kv[u < 0.25] = 7
    kv[(u >= 0.25) & (u < 0.27)] = np.iinfo(np.int64).min
    x = dbl_with_specials(r, n, 0.01, 0.02)
    cols = {"g": (kv, r.random(n) >= 0.3), "x": (x, with_na(r, n, 0.02)),
            "v": (r.integers(-10**12, 10**12, n), with_na(r, n, 0.02)),
            "b": (r.random(n) < 0.5, with_na(r, n, 0.02)), "ax": np.abs(np.nan_to_num(x, nan=0.0))}
    check(gpu, lambda: gpu.Table(cols, rg_rows=0), ["g"], ALL_AGGS, temp_dir)


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
def test_groupby_partitioned_then_more_batches(gpu, oracle, part_threshold, temp_dir):
    """a partitioned first batch folded into the hash table by later batches"""
    mk = lambda: table(gpu, 1_600_000, 29, "int", 200_000, 1_100_000, p_na=0.01, p_nan=0.01)
    check(gpu, mk, ["g"], ALL_AGGS, temp_dir)


RANGE_AGGS = [("n", -1, 0), ("count", 0, 0), ("sum", 0, 1), ("min", 0, 1), ("max", 0, 0), ("sum", 1, 1),
              ("min", 1, 1)]


def range_keys(dist, n, r):
    """key values of a range-partition case and the path its balance picks
    (2 key ranges, 1 back to hash partitions)"""
    k = r.integers(-2**62, 2**62, n)
    if dist == "dense":
        return r.integers(0, 300_000, n), 2
    if dist == "cluster":
        # 120K rows over 50 keys: within one first-pass bucket's slack, but
        # one partition of 30 items -> hash partitions
        k[:120_000] = r.integers(0, 50, 120_000)
        return r.permutation(k), 1
    if dist == "hot":
        k[r.random(n) < 0.25] = 7       # a quarter of the rows in one bucket
        return k, 1
    if dist == "empty_key":
        k[r.random(n) < 0.01] = np.iinfo(np.int64).min   # its record joins last: finish sorts
        return k, 2
    return k, 2


@pytest.mark.parametrize("dist,nulls", [("uniform", True), ("dense", True), ("cluster", True), ("hot", True),
                                        ("empty_key", True), ("uniform", False), ("empty_key", False),
                                        ("hot", False)])
def test_groupby_range_partitions(gpu, part_threshold, dist, nulls):
    """key order over int64 keys: records of key-range partitions come out of
    the partitioned feed in key order (finish sorts nothing unless an
    INT64_MIN key is present); keys, counts, int64 sums, min and max are
    bit-exact against numpy and against hash partitions (gb.part.range 1),
    double sums within 1e-12 * sum |x|; badly balanced ranges fall back.
    Without NA keys or inputs the first pass takes its plain form."""
    n = 1_500_000
    r = rng(1000 + len(dist))
    k, want_path = range_keys(dist, n, r)
    kval = r.random(n) >= 0.02 if nulls else np.ones(n, bool)
    iv = r.integers(-10**9, 10**9, n)
    ival = r.random(n) >= 0.03 if nulls else np.ones(n, bool)
    dv = r.normal(size=n) * 1e6
    cols = [gpu.DeviceColumn(k, kval if nulls else None), gpu.DeviceColumn(iv, ival if nulls else None),
            gpu.DeviceColumn(dv)]
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"gb.part.range")
    res = {}
    try:
        for mode in (2, 1):
            assert lib.vgpu_tune_set(b"gb.part.range", mode) == 0
            gb = gpu.GroupBy(gpu.VEC_INT64, [gpu.VEC_INT64, gpu.VEC_DOUBLE], RANGE_AGGS, gpu.VGPU_ORDER_KEY, 0)
            gb.feed(cols[0], cols[1:], n)
            path = gb.last_path()
            gb.finish()
            res[mode] = gb.fetch()
            gb.free()
            assert path == (want_path if mode == 2 else 1), (mode, path)
    finally:
        lib.vgpu_tune_set(b"gb.part.range", old)
        for c in cols:
            c.free()
    # numpy: regular keys ascending, then the NA group (if any)
    kk = k[kval]
    uk, inv = np.unique(kk, return_inverse=True)
    ng = len(uk)
    got = res[2]
    assert np.array_equal(got["key"][0][:ng], uk) and got["key"][1][:ng].all()
    assert len(got["key"][0]) == ng + int(nulls)
    if nulls:
        assert not got["key"][1][ng]
    m = ng + int(nulls)
    g = np.concatenate([inv, np.full((~kval).sum(), ng)])
    order = np.concatenate([np.flatnonzero(kval), np.flatnonzero(~kval)])
    ivo, ivalo, dvo = iv[order], ival[order], dv[order]
    cnt = np.bincount(g, minlength=m)
    assert np.array_equal(got["a0"][0], cnt)
    assert np.array_equal(got["a1"][0], np.bincount(g, weights=ivalo, minlength=m))
    # every result is a double: int64 sums (small enough to be exact), min
    # with na_rm (+Inf for a group with no value), max without (NA when any
    # input is NA)
    s_int = np.zeros(m, np.int64)
    np.add.at(s_int, g[ivalo], ivo[ivalo])
    assert got["a2"][1].all() and np.array_equal(got["a2"][0], s_int.astype(np.float64))
    lo = np.full(m, np.inf)
    np.minimum.at(lo, g[ivalo], ivo[ivalo].astype(np.float64))
    assert got["a3"][1].all() and np.array_equal(got["a3"][0], lo)
    hi = np.full(m, -np.inf)
    np.maximum.at(hi, g[ivalo], ivo[ivalo].astype(np.float64))
    all_valid = np.bincount(g[~ivalo], minlength=m) == 0
    assert np.array_equal(got["a4"][1], all_valid)
    assert np.array_equal(got["a4"][0][all_valid], hi[all_valid])
    sd = np.bincount(g, weights=dvo, minlength=m)
    sa = np.bincount(g, weights=np.abs(dvo), minlength=m)
    assert got["a5"][1].all() and np.all(np.abs(got["a5"][0] - sd) <= 1e-12 * sa + 1e-300)
    mn = np.full(m, np.inf)
    np.minimum.at(mn, g, dvo)
    assert got["a6"][1].all() and np.array_equal(got["a6"][0], mn)
    for col in got:
        assert np.array_equal(got[col][1], res[1][col][1]), col
        if col == "a5":
            assert np.all(np.abs(got[col][0] - res[1][col][0]) <= 2e-12 * sa + 1e-300)
        else:
            assert np.array_equal(got[col][0], res[1][col][0]), col


def test_groupby_partitioned_after_filter(gpu, oracle, part_threshold):
    """a selection vector from the filter feeds the partitioned path"""
    mk = lambda: table(gpu, 2_400_000, 31, "int", 50_000, 0, p_na=0.01, p_nan=0.01)
    check(gpu, mk, ["g"], ALL_AGGS, "/tmp", pred=("cmp", ">", ("col", "ax"), ("dbl", 0.3)))


@pytest.mark.parametrize("shape", ["general", "narrow", "multi", "partitioned"])
def test_nan_keys_in_key_order_are_an_error(gpu, oracle, shape):
    """NaN group keys with key order: the reference's merge sort ranks NaN
    equal to every key (src/sort.c:32-47, 97-116), so NaN groups land where
    its merge pattern puts them; the GPU node raises an error instead of
    returning another order, on every feed path.  The hash path (temp_dir
    None) groups NaN keys one row each, as the other tests check."""
    nv = gpu
    n = 1_200_000 if shape == "partitioned" else 40_000
    mk = lambda: table(nv, n, 41, "dbl", 4000 if shape == "partitioned" else 8, 0, p_na=0.02, p_nan=0.01)
    keys = ["g", "k2"] if shape == "multi" else ["g"]
    aggs = NARROW_AGGS if shape in ("narrow", "multi") else ALL_AGGS

    def tbl():
        t = mk()
        if shape == "multi":
            t.cols["k2"] = nv.Column(np.arange(n) % 3)
        return t
    old = nv.vgpu().vgpu_tune_get(b"gb.part.min_groups")
    if shape == "partitioned":
        assert nv.vgpu().vgpu_tune_set(b"gb.part.min_groups", 1000) == 0
    try:
        with pytest.raises(nv.VectraError, match="NaN values in a double group key with key order"):
            run(nv, "gpu", tbl(), keys, aggs, "/tmp")
        # the same rows on the hash path are computed and match the oracle
        got = run(nv, "gpu", tbl(), keys, aggs, None)
        exp = run(nv, "oracle", tbl(), keys, aggs, None)
        assert np.array_equal(got["n"][0], exp["n"][0])
    finally:
        nv.vgpu().vgpu_tune_set(b"gb.part.min_groups", old)


STAGE_AGGS = {0: [("n", -1, 0)],
              1: [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0)],
              2: [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("sum", 1, 0), ("max", 1, 0)],
              3: [("n", -1, 0), ("sum", 0, 0), ("min", 1, 0), ("sum", 2, 0), ("max", 2, 0)],
              4: [("n", -1, 0), ("max", 0, 0), ("sum", 1, 0), ("min", 2, 0), ("sum", 3, 0)]}


@pytest.mark.parametrize("stage,agg", [(2, 3), (3, 3), (3, 2)])
@pytest.mark.parametrize("range_mode", [2, 1])
@pytest.mark.parametrize("n_inputs", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("valid", [False, True])
def test_groupby_partition_staging(gpu, part_threshold, stage, agg, range_mode, n_inputs, valid):
    """the LDS-staged partition passes and resident aggregation blocks
    (gb.part.stage 2 and 3: tiles of 1,024 and 2,048 rows; gb.part.agg 3:
    items at a fixed stride with the next item's rows loaded ahead, 2: items
    by ticket) against the row-wise ones (1): staged rows of 2 to 6 words,
    the plain first pass (no validity) and the general one, key ranges and
    hashes; keys, order, counts, min and max bit-exact, sums equal or within
    1e-9 of max(1, |sum|) (summation order), counts and keys against numpy"""
    n = 1_300_000
    r = rng(4000 + 10 * n_inputs + range_mode)
    k = r.integers(-10**12, 10**12, 300_000)[r.integers(0, 300_000, n)]
    kval = r.random(n) >= 0.01 if valid else None
    types, datas = [], []
    for i in range(n_inputs):
        if i % 2 == 0:
            types.append(gpu.VEC_DOUBLE)
            datas.append(r.normal(size=n) * 1e3)
        else:
            types.append(gpu.VEC_INT64)
            datas.append(r.integers(-10**9, 10**9, n))
    cols = [gpu.DeviceColumn(k, kval)] + [gpu.DeviceColumn(d) for d in datas]
    lib = gpu.vgpu()
    old = (lib.vgpu_tune_get(b"gb.part.stage"), lib.vgpu_tune_get(b"gb.part.range"), lib.vgpu_tune_get(b"gb.part.agg"))
    res = {}
    try:
        assert lib.vgpu_tune_set(b"gb.part.range", range_mode) == 0
        assert lib.vgpu_tune_set(b"gb.part.agg", agg) == 0
        for mode in (stage, 1):
            assert lib.vgpu_tune_set(b"gb.part.stage", mode) == 0
            gb = gpu.GroupBy(gpu.VEC_INT64, types, STAGE_AGGS[n_inputs], gpu.VGPU_ORDER_KEY, 0)
            gb.feed(cols[0], cols[1:], n)
            path = gb.last_path()
            gb.finish()
            res[mode] = gb.fetch()
            gb.free()
            assert path == (2 if range_mode == 2 else 1), (mode, path)
    finally:
        lib.vgpu_tune_set(b"gb.part.stage", old[0])
        lib.vgpu_tune_set(b"gb.part.range", old[1])
        lib.vgpu_tune_set(b"gb.part.agg", old[2])
        for c in cols:
            c.free()
    kk = k if kval is None else k[kval]
    uk, cnt = np.unique(kk, return_counts=True)
    got, ref = res[stage], res[1]
    ng = len(uk)
    assert np.array_equal(got["key"][0][:ng], uk)
    assert np.array_equal(got["a0"][0][:ng], cnt)
    for col in got:
        assert np.array_equal(got[col][1], ref[col][1]), col
        a, b = got[col][0], ref[col][0]
        if a.dtype == np.float64 and col != "key":
            same = np.array_equal(a, b) or np.all(np.abs(a - b) <= 1e-9 * np.maximum(1.0, np.abs(b)))
            assert same, col
        else:
            assert np.array_equal(a, b), col


@pytest.mark.parametrize("special", ["none", "zero", "negzero", "nan"])
@pytest.mark.parametrize("range_mode", [2, 1])
@pytest.mark.parametrize("n_inputs", [1, 2])
def test_groupby_partition_rows_without_meta(gpu, part_threshold, special, range_mode, n_inputs):
    """plain key-ordered feeds stage rows without their meta word (key and
    inputs only) unless a double min / max input holds a NaN or a zero,
    which needs positions: then the feed restarts with meta words.  Both
    against the row-wise passes (gb.part.stage 1) and the oracle-free
    closed forms: keys, order, counts, min and max bit-exact (the NaN-first
    and signed-zero rules included), sums within 1e-9 of max(1, |sum|)"""
    n = 1_300_000
    r = rng(4100 + 10 * n_inputs + range_mode)
    k = r.integers(-10**12, 10**12, 300_000)[r.integers(0, 300_000, n)]
    types, datas = [], []
    for i in range(n_inputs):
        if i % 2 == 0:
            d = r.normal(size=n) * 1e3
            if special == "zero":
                d[r.integers(0, n, 50)] = 0.0
            elif special == "negzero":
                d[r.integers(0, n, 50)] = -0.0
            elif special == "nan":
                d[r.integers(0, n, 50)] = np.nan
            types.append(gpu.VEC_DOUBLE)
            datas.append(d)
        else:
            types.append(gpu.VEC_INT64)
            datas.append(r.integers(-10**9, 10**9, n))
    cols = [gpu.DeviceColumn(k)] + [gpu.DeviceColumn(d) for d in datas]
    lib = gpu.vgpu()
    knobs = (b"gb.part.stage", b"gb.part.range", b"gb.part.agg")
    old = [lib.vgpu_tune_get(x) for x in knobs]
    res = {}
    try:
        assert lib.vgpu_tune_set(b"gb.part.range", range_mode) == 0
        assert lib.vgpu_tune_set(b"gb.part.agg", 3) == 0
        for mode in (3, 1):
            assert lib.vgpu_tune_set(b"gb.part.stage", mode) == 0
            gb = gpu.GroupBy(gpu.VEC_INT64, types, STAGE_AGGS[n_inputs], gpu.VGPU_ORDER_KEY, 0)
            gb.feed(cols[0], cols[1:], n)
            gb.finish()
            res[mode] = gb.fetch()
            gb.free()
    finally:
        for x, v in zip(knobs, old):
            lib.vgpu_tune_set(x, v)
        for c in cols:
            c.free()
    uk, cnt = np.unique(k, return_counts=True)
    got, ref = res[3], res[1]
    ng = len(uk)
    assert np.array_equal(got["key"][0][:ng], uk)
    assert np.array_equal(got["a0"][0][:ng], cnt)
    for col in got:
        assert np.array_equal(got[col][1], ref[col][1]), col
        a, b = got[col][0], ref[col][0]
        if a.dtype == np.float64 and col != "key":
            bits_equal = np.array_equal(a.view(np.uint64), b.view(np.uint64))
            if col in ("a2", "a3"):   # min and max of the double input: bit for bit
                assert bits_equal, col
                continue
            close = np.all((np.isnan(a) & np.isnan(b)) | (np.abs(a - b) <= 1e-9 * np.maximum(1.0, np.abs(b))))
            assert bits_equal or close, col
        else:
            assert np.array_equal(a, b), col
