"""String keys through the GPU nodes against the oracle nodes: group_by on
a string column alone or with numeric columns (src/group_agg.c with
VEC_STRING keys: src/hash.c:35-41 hashing, :58-66 equality; src/sort.c:58-66
order in key order) and joins on string keys (src/join.c:96-108, 335-341),
every join kind, with NA strings, empty strings, shared prefixes, multi-byte
UTF-8, host row groups and a filter's selection vectors.  Keys, order,
counts and join rows exact; double sums and means within 1e-12 * sum|x|."""
import numpy as np
import pytest

from helpers import assert_tables_equal, rng, with_na

pytestmark = pytest.mark.gpu


def strings(r, n, n_distinct):
    pool = ["", "a", "ab", "abc", "b", "ba", "Z", "é", "ü1", "x" * 40, "x" * 41] + [
        f"key-{i:06d}" + "#" * (i % 9) for i in range(max(0, n_distinct - 11))]
    pool = pool[:n_distinct]
    return np.array([pool[i] for i in r.integers(0, len(pool), n)], dtype=object)


def group_table(seed, n, n_distinct):
    r = rng(seed)
    x = r.normal(size=n)
    return {"g": (strings(r, n, n_distinct), with_na(r, n, 0.05)), "k": (r.integers(0, 4, n), with_na(r, n, 0.03)),
            "x": (x, with_na(r, n, 0.02)), "v": r.integers(-2**40, 2**40, n), "ax": np.abs(x)}


AGGS = [("n", "n", None, 0), ("sx", "sum", "x", 1), ("mx", "mean", "x", 1), ("lo", "min", "v", 0),
        ("hi", "max", "x", 1)]


def group(nv, engine, cols, keys, temp_dir, rg_rows, pred, aggs):
    node = nv.Table(cols, rg_rows=rg_rows).scan()
    if pred is not None:
        node = nv.filter_node(engine, node, pred)
    return nv.collect(nv.group_agg_node(engine, node, keys, aggs, temp_dir))


@pytest.mark.parametrize("keys,temp_dir,pred,n_distinct", [
    (["g"], None, None, 50),
    (["g"], "/tmp", None, 50),
    (["g"], None, ("cmp", ">", ("col", "x"), ("dbl", 0.3)), 5000),
    (["g", "k"], "/tmp", None, 300),
    (["k", "g"], None, ("cmp", "<", ("col", "v"), ("int", 0)), 300),
])
def test_string_groupby_matches_oracle(gpu, oracle, keys, temp_dir, pred, n_distinct):
    cols = group_table(101 + n_distinct, 60_000, n_distinct)
    got = group(gpu, "gpu", cols, keys, temp_dir, 7_000, pred, AGGS)
    exp = group(gpu, "oracle", cols, keys, temp_dir, 7_000, pred,
                AGGS + [("__sa", "sum", "ax", 1), ("__cx", "count", "x", 0)])
    sa = exp.pop("__sa")[0]
    cnt = exp.pop("__cx")[0]
    assert_tables_equal(got, exp, {"sx": sa, "mx": sa / np.maximum(cnt, 1)})


def test_string_groupby_known_answers(gpu):
    """test-groupby.R:1-60 through the GPU node"""
    def q(cols, aggs, td=None):
        return gpu.collect(gpu.group_agg_node("gpu", gpu.Table(cols).scan(), ["g"], aggs, td))

    out = q({"g": np.array(["a", "b", "a", "b", "a"]), "x": np.arange(1, 6)}, [("cnt", "n", None, 0)])
    assert list(out["g"][0]) == ["a", "b"] and list(out["cnt"][0]) == [3, 2]
    out = q({"g": np.array(["a", "a", "b", "b"]), "x": np.array([10., 20., 30., 40.])}, [("avg", "mean", "x", 0)])
    assert list(out["avg"][0]) == [15, 35]
    out = q({"g": np.array(["a", "a", "b", "b"]), "x": np.array([5., 1., 8., 3.])},
            [("lo", "min", "x", 0), ("hi", "max", "x", 0)])
    assert list(out["lo"][0]) == [1, 3] and list(out["hi"][0]) == [5, 8]
    for td in (None, "/tmp"):
        out = q({"g": (np.array(["a", "", "a", ""]), np.array([1, 0, 1, 0], bool)), "x": np.array([1., 2., 3., 4.])},
                [("total", "sum", "x", 0)], td)
        assert list(out["g"][1]) == [True, False] and list(out["total"][0]) == [4, 6]


def test_string_groupby_empty_input(gpu, oracle):
    cols = {"g": np.array(["a", "b"], dtype=object), "x": np.array([1., 2.])}
    pred = ("cmp", ">", ("col", "x"), ("dbl", 5.0))
    for td in (None, "/tmp"):
        got = group(gpu, "gpu", cols, ["g"], td, 0, pred, [("n", "n", None, 0)])
        exp = group(gpu, "oracle", cols, ["g"], td, 0, pred, [("n", "n", None, 0)])
        assert_tables_equal(got, exp)


def test_string_key_order_on_the_device(gpu, oracle):
    """key order of string groups sorted on the device (length, then 8-byte
    chunks, then NA last): prefixes at and across chunk boundaries,
    embedded NUL bytes, bytes above 0x7F (unsigned), the empty string and
    NA, alone and with a number key"""
    pool = ["", "\x00", "a", "a\x00", "a\x00b", "ab", "b", "é", "z", "zz", "a" * 8, "a" * 8 + "\x00", "a" * 9,
            "a" * 16, "a" * 17, "a" * 7 + "b", "a" * 8 + "b", "ÿ", "y"]
    r = rng(811)
    n = 6_000
    g = np.array([pool[i] for i in r.integers(0, len(pool), n)], dtype=object)
    cols = {"g": (g, with_na(r, n, 0.05)), "k": r.integers(0, 3, n), "x": r.normal(size=n)}
    for keys in (["g"], ["k", "g"], ["g", "k"]):
        got = group(gpu, "gpu", cols, keys, "/tmp", 1_000, None, [("n", "n", None, 0)])
        exp = group(gpu, "oracle", cols, keys, "/tmp", 1_000, None, [("n", "n", None, 0)])
        assert_tables_equal(got, exp)


def dict_pool(r, n_distinct):
    """distinct strings with empty strings, shared prefixes, multi-byte
    UTF-8 and strings that differ only in their last byte or length"""
    base = ["", "a", "ab", "abc", "b", "ba", "é", "ü1", "x" * 40, "x" * 41, "x" * 40 + "y"]
    return base + [f"s{i:07d}" + "z" * int(r.integers(0, 5)) for i in range(n_distinct - len(base))]


@pytest.mark.parametrize("bits", [64, 1])
def test_strdict_codes_are_exact(gpu, bits):
    """codes equal iff bytes equal, over three appended batches and the
    table's growth (6,000 distinct strings, first table 1,024 slots); with
    the hash cut to one bit every string shares two slots and one tag, so
    only the byte comparisons tell strings apart; lookups return the
    appended codes and -1 for strings never appended"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"strdict.hash_bits")
    lib.vgpu_tune_set(b"strdict.hash_bits", bits)
    try:
        d = gpu.StrDict()
    finally:
        lib.vgpu_tune_set(b"strdict.hash_bits", old)
    r = rng(601 + bits)
    pool = dict_pool(r, 6_000)
    seen = {}
    batches = []
    for size in (5_000, 1, 20_000):
        vals = [None if r.random() < 0.03 else pool[int(r.integers(0, len(pool)))] for _ in range(size)]
        codes = d.append(vals)
        batches.append((vals, codes))
        for v, c in zip(vals, codes):
            if v is None:
                continue
            assert seen.setdefault(v, int(c)) == int(c), f"{v!r} got two codes"
    assert len(set(seen.values())) == len(seen), "two strings share a code"
    distinct, rows = d.size()
    assert distinct == len(seen) and rows == 25_001
    for vals, codes in batches:
        got = d.lookup(vals)
        ok = np.array([v is not None for v in vals])
        assert np.array_equal(got[ok], codes[ok])
    unseen = ["never", "x" * 42, "s9999999", "abcd", "é" * 2]
    assert list(d.lookup(unseen)) == [-1] * len(unseen)
    d.free()


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
def test_string_groupby_exact_under_forced_collisions(gpu, oracle, temp_dir):
    """the group-by node with the dictionary's hash cut to one bit: every
    string collides with every other in slot and tag, groups are still the
    oracle's (one per distinct byte string), in first-seen and key order"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"strdict.hash_bits")
    lib.vgpu_tune_set(b"strdict.hash_bits", 1)
    try:
        cols = group_table(701, 30_000, 400)
        got = group(gpu, "gpu", cols, ["g", "k"], temp_dir, 4_000, None, AGGS)
    finally:
        lib.vgpu_tune_set(b"strdict.hash_bits", old)
    exp = group(gpu, "oracle", cols, ["g", "k"], temp_dir, 4_000, None,
                AGGS + [("__sa", "sum", "ax", 1), ("__cx", "count", "x", 0)])
    sa = exp.pop("__sa")[0]
    cnt = exp.pop("__cx")[0]
    assert_tables_equal(got, exp, {"sx": sa, "mx": sa / np.maximum(cnt, 1)})


def join_tables(nv, seed, n_left, n_right, n_distinct, composite):
    def mk():
        r = rng(seed)
        lk, rk = strings(r, n_left, n_distinct), strings(r, n_right, n_distinct)
        lcols = {"name": (lk, with_na(r, n_left, 0.05))}
        rcols = {"w": r.random(n_right), "name": (rk, with_na(r, n_right, 0.05))}
        if composite:
            lcols["a"] = r.integers(0, 3, n_left)
            rcols["a"] = r.integers(0, 3, n_right).astype(np.float64)
        lcols["x"] = r.normal(size=n_left)
        rcols["label"] = strings(r, n_right, 20)
        return nv.Table(lcols, rg_rows=700), nv.Table(rcols, rg_rows=500)
    return mk


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
@pytest.mark.parametrize("composite", [False, True])
def test_string_join_matches_oracle(gpu, oracle, kind, composite):
    """a string key, or a number and a string key (the left number is
    coerced to the right's double: inner / semi / anti only, as the GPU
    path declines coerced left / full joins on several keys)"""
    mk = join_tables(gpu, 151, 5000, 1800, 400, composite)
    keys = [(0, 1)] + ([(1, 2)] if composite else [])
    if composite and kind in ("left", "full"):
        left, right = mk()
        node = gpu.join_node("gpu", left.scan(), right.scan(), kind, keys)
        with pytest.raises(gpu.VectraError, match="coerced to double is not evaluated on the GPU"):
            gpu.next_batches(node)
        return
    out = {}
    for engine in ("gpu", "oracle"):
        left, right = mk()
        lnode = gpu.filter_node(engine, left.scan(), ("cmp", ">", ("col", "x"), ("dbl", -0.8)))
        out[engine] = gpu.next_batches(gpu.join_node(engine, lnode, right.scan(), kind, keys))
    got, exp = out["gpu"], out["oracle"]
    assert len(got) == len(exp), f"{len(got)} batches vs {len(exp)}"
    for g, e in zip(got, exp):
        assert_tables_equal(g["cols"], e["cols"])


@pytest.mark.parametrize("kind,composite", [("inner", False), ("left", False), ("semi", False), ("anti", False),
                                            ("full", False), ("inner", True), ("semi", True), ("anti", True)])
def test_string_join_exact_under_forced_collisions(gpu, oracle, kind, composite):
    """the join's string dictionary with its hash cut to one bit: every
    build string shares two slots and one tag, probe strings absent from the
    build side included; rows and order are still the oracle's"""
    mk = join_tables(gpu, 157, 4000, 1500, 600, composite)
    keys = [(0, 1)] + ([(1, 2)] if composite else [])
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"strdict.hash_bits")
    out = {}
    for engine in ("gpu", "oracle"):
        left, right = mk()
        lnode = gpu.filter_node(engine, left.scan(), ("cmp", ">", ("col", "x"), ("dbl", -0.8)))
        node = gpu.join_node(engine, lnode, right.scan(), kind, keys)
        lib.vgpu_tune_set(b"strdict.hash_bits", 1)
        try:
            out[engine] = gpu.next_batches(node)
        finally:
            lib.vgpu_tune_set(b"strdict.hash_bits", old)
    got, exp = out["gpu"], out["oracle"]
    assert len(got) == len(exp), f"{len(got)} batches vs {len(exp)}"
    for g, e in zip(got, exp):
        assert_tables_equal(g["cols"], e["cols"])


def test_string_join_known_answers(gpu):
    """test-joins.R:135-163 through the GPU node"""
    L = gpu.Table({"name": np.array(["alice", "bob"]), "x": np.array([1., 2.])})
    R = gpu.Table({"name": np.array(["bob", "carol"]), "y": np.array([20., 30.])})
    out = gpu.collect(gpu.join_node("gpu", L.scan(), R.scan(), "inner", [(0, 0)]))
    assert list(out["name"][0]) == ["bob"] and list(out["y"][0]) == [20]
    L = gpu.Table({"a": np.array([1., 1., 2.]), "b": np.array(["x", "y", "x"]), "v1": np.array([10., 20., 30.])})
    R = gpu.Table({"a": np.array([1., 2.]), "b": np.array(["x", "x"]), "v2": np.array([100., 300.])})
    out = gpu.collect(gpu.join_node("gpu", L.scan(), R.scan(), "inner", [(0, 0), (1, 1)]))
    assert sorted(out["v1"][0]) == [10, 30]


def test_string_key_against_number_is_an_error(gpu):
    L = gpu.Table({"name": np.array(["a"]), "x": np.array([1.])})
    R = gpu.Table({"name": np.array([1.]), "y": np.array([2.])})
    with pytest.raises(gpu.VectraError, match="key type mismatch"):
        gpu.join_node("gpu", L.scan(), R.scan(), "inner", [(0, 0)])


def test_string_keys_long_and_shared_prefix(gpu, oracle):
    """strings up to 6,000 bytes that differ only in their last byte or
    their length, in both orders"""
    r = rng(171)
    base = "p" * 5999
    pool = np.array([base + c for c in "abc"] + [base, base[:3000], "", "q"], dtype=object)
    n = 20_000
    cols = {"g": (pool[r.integers(0, len(pool), n)], with_na(r, n, 0.05)), "x": r.normal(size=n)}
    cols["ax"] = np.abs(cols["x"])
    for td in (None, "/tmp"):
        got = group(gpu, "gpu", cols, ["g"], td, 3_000, None, [("n", "n", None, 0), ("s", "sum", "x", 0)])
        exp = group(gpu, "oracle", cols, ["g"], td, 3_000, None, [("n", "n", None, 0), ("s", "sum", "x", 0),
                                                                  ("__sa", "sum", "ax", 0)])
        sa = exp.pop("__sa")[0]
        assert_tables_equal(got, exp, {"s": sa})


@pytest.mark.parametrize("kind", ["inner", "left", "full", "semi", "anti"])
def test_string_join_empty_sides(gpu, oracle, kind):
    """an empty build side and an empty probe side"""
    full = {"name": np.array(["a", "b", "a"]), "x": np.array([1., 2., 3.])}
    empty = {"name": np.array([], dtype=object), "y": np.array([], dtype=np.float64)}
    for lcols, rcols in ((full, empty), ({"name": np.array([], dtype=object), "x": np.array([], np.float64)},
                                         {"name": np.array(["a"]), "y": np.array([5.])})):
        out = {}
        for engine in ("gpu", "oracle"):
            node = gpu.join_node(engine, gpu.Table(lcols).scan(), gpu.Table(rcols).scan(), kind, [(0, 0)])
            out[engine] = gpu.next_batches(node)
        assert len(out["gpu"]) == len(out["oracle"])
        for g, e in zip(out["gpu"], out["oracle"]):
            assert_tables_equal(g["cols"], e["cols"])
