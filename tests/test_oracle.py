"""Pins the oracle (CPU restatement of the reference) without a GPU:
published FNV-1a vectors, the reference's own known-answer tests
(transcribed from tests/testthat), brute-force restatements on small inputs
(numpy unique / first occurrence / nested-loop join), and math.fsum for the
exact sum."""
import ctypes as C
import math

import numpy as np
import pytest

from helpers import dbl_with_specials, rng, with_na


# ---------------------------------------------------------------- hashing

def str_array(nv, strings):
    data = b"".join(s.encode() for s in strings)
    offs = np.zeros(len(strings) + 1, np.int64)
    offs[1:] = np.cumsum([len(s.encode()) for s in strings])
    a = nv.VecArray()
    a.type = nv.VEC_STRING
    a.length = len(strings)
    a.validity = None
    a.buf.str.offsets = offs.ctypes.data_as(C.POINTER(C.c_int64))
    buf = C.create_string_buffer(data, max(1, len(data)))
    a.buf.str.data = C.addressof(buf)
    return a, (offs, buf)


def test_fnv1a_published_vectors(oracle, native):
    # FNV-1a 64-bit test vectors (Fowler / Noll / Vo reference set)
    vec = {"": 0xcbf29ce484222325, "a": 0xaf63dc4c8601ec8c, "foobar": 0x85944171f73967e8}
    a, keep = str_array(native, list(vec))
    for i, h in enumerate(vec.values()):
        assert oracle.oracle_hash_value(C.byref(a), i) == h


def test_hash_na_and_signed_zero(oracle, native):
    col = native.Column(np.array([0.0, -0.0, 1.5]), np.array([1, 1, 0], bool))
    a = col.as_vecarray()
    h = [oracle.oracle_hash_value(C.byref(a), i) for i in range(3)]
    assert h[0] == h[1]                                 # -0 hashes as +0 (src/hash.c:23-24)
    assert h[2] == 0xcbf29ce484222325 ^ 0xFF            # NA (src/hash.c:15)
    # int64 keys hash their 8 little-endian bytes: the bytes of 0x61 are "a\0\0\0\0\0\0\0"
    b = native.Column(np.array([0x61]))
    sa, keep = str_array(native, ["a\0\0\0\0\0\0\0"])
    assert oracle.oracle_hash_value(C.byref(b.as_vecarray()), 0) == oracle.oracle_hash_value(C.byref(sa), 0)


# ---------------------------------------------------------------- exact sum

@pytest.mark.parametrize("seed", range(6))
def test_exact_sum_matches_fsum(oracle, seed):
    r = rng(seed)
    x = np.concatenate([r.normal(size=500) * 10.0 ** r.integers(-300, 300, 500),
                        [1e308, -1e308, 5e-324, -5e-324, 1.0, 1e-16, -1.0]])
    r.shuffle(x)
    got = oracle.oracle_exact_sum(x.ctypes.data_as(C.POINTER(C.c_double)), len(x))
    assert got == math.fsum(x)


def test_exact_sum_cancellation(oracle):
    x = np.array([1e100, 1.0, -1e100, 1e-20, 3.0])
    assert oracle.oracle_exact_sum(x.ctypes.data_as(C.POINTER(C.c_double)), 5) == 4.00000000000000000001
    y = np.array([0.1] * 10)
    assert oracle.oracle_exact_sum(y.ctypes.data_as(C.POINTER(C.c_double)), 10) == math.fsum(y)


# ---------------------------------------------------------------- filter

def test_filter_known_answers(native, oracle):
    """test-filter.R:1-97 through the oracle."""
    nv = native
    t = nv.Table({"x": np.array([1, 2, 3, 4, 5]), "y": np.array([10., 20., 30., 40., 50.])})
    out = nv.collect(nv.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 3))))
    assert list(out["x"][0]) == [4, 5] and list(out["y"][0]) == [40, 50]
    t = nv.Table({"x": np.arange(1, 7)})
    out = nv.collect(nv.filter_node("oracle", t.scan(), ("or", ("cmp", "<", ("col", "x"), ("dbl", 3)),
                                                        ("cmp", ">", ("col", "x"), ("dbl", 4)))))
    assert list(out["x"][0]) == [1, 2, 5, 6]
    t = nv.Table({"x": (np.array([1, 0, 3, 0, 5]), np.array([1, 0, 1, 0, 1], bool))})
    out = nv.collect(nv.filter_node("oracle", t.scan(), ("not", ("is_na", ("col", "x")))))
    assert list(out["x"][0]) == [1, 3, 5]
    t = nv.Table({"x": (np.array([1, 0, 3]), np.array([1, 0, 1], bool)), "y": np.array([10., 20., 30.])})
    out = nv.collect(nv.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 1))))
    assert list(out["x"][0]) == [3]
    t = nv.Table({"x": np.arange(1., 11.)})
    node = nv.filter_node("oracle", nv.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 3))),
                          ("cmp", "<", ("col", "x"), ("dbl", 8)))
    assert list(nv.collect(node)["x"][0]) == [4, 5, 6, 7]


def cmp_sign(a, b):
    with np.errstate(invalid="ignore"):
        return (a > b).astype(int) - (a < b).astype(int)


@pytest.mark.parametrize("op", ["==", "!=", "<", "<=", ">", ">="])
def test_filter_brute_force(native, oracle, op):
    """sel = ascending rows whose predicate is valid and TRUE, where the
    comparison is decided on sign((a>b)-(a<b)) -- so NaN satisfies ==, <=, >=."""
    nv = native
    r = rng(1)
    n = 3000
    x = dbl_with_specials(r, n, 0.1, 0.05)
    valid = with_na(r, n, 0.1)
    rg = 512
    t = nv.Table({"x": (x, valid)}, rg_rows=rg)
    batches = nv.next_batches(nv.filter_node("oracle", t.scan(), ("cmp", op, ("col", "x"), ("dbl", 0.0))))
    c = cmp_sign(x, 0.0)
    keep = {"==": c == 0, "!=": c != 0, "<": c < 0, "<=": c <= 0, ">": c > 0, ">=": c >= 0}[op] & valid
    exp = [np.flatnonzero(keep[i:i + rg]) for i in range(0, n, rg)]
    exp = [e for e in exp if len(e)]
    assert len(batches) == len(exp)
    for b, e in zip(batches, exp):
        assert np.array_equal(b["sel"], e)


# ---------------------------------------------------------------- group-by

def brute_groupby(keys, kvalid, vals, vvalid, order):
    """reference grouping rules: NA keys form one group, NaN keys never
    merge, -0 == +0; order = 'first' (first occurrence) or 'sorted'
    (ascending, NA last)."""
    groups = {}
    firsts = []
    for i, (k, kv) in enumerate(zip(keys, kvalid)):
        if not kv:
            g = ("NA",)
        elif isinstance(k, float) and math.isnan(k):
            g = ("NaN", i)
        else:
            g = ("v", k + 0.0 if isinstance(k, float) else k)
        if g not in groups:
            groups[g] = []
            firsts.append(g)
        groups[g].append(i)
    if order == "sorted":
        firsts = sorted([g for g in firsts if g[0] == "v"], key=lambda g: g[1]) + [g for g in firsts if g[0] == "NA"]
    out = []
    for g in firsts:
        rows = groups[g]
        vs = [vals[i] for i in rows if vvalid[i]]
        out.append((g, len(rows), len(vs), math.fsum(vs) if len(vs) == len(rows) else None,
                    min(vs) if vs and len(vs) == len(rows) else None))
    return out


@pytest.mark.parametrize("order", ["first", "sorted"])
@pytest.mark.parametrize("kt", ["int", "dbl"])
def test_groupby_brute_force(native, oracle, order, kt):
    nv = native
    r = rng(2)
    n = 2000
    if kt == "int":
        keys = r.integers(-20, 20, n)
    else:
        keys = np.round(r.normal(size=n) * 3) / 2
    kvalid = with_na(r, n, 0.05)
    vals = r.normal(size=n)
    vvalid = with_na(r, n, 0.01)
    t = nv.Table({"g": (keys, kvalid), "v": (vals, vvalid)}, rg_rows=300)
    out = nv.collect(nv.group_agg_node("oracle", t.scan(), ["g"],
                                       [("n", "n", None, 0), ("c", "count", "v", 0), ("s", "sum", "v", 0),
                                        ("lo", "min", "v", 0)], "/tmp" if order == "sorted" else None))
    exp = brute_groupby(list(keys), list(kvalid), list(vals), list(vvalid), order)
    assert len(out["g"][0]) == len(exp)
    for i, (g, nrows, nvalid, s, lo) in enumerate(exp):
        if g[0] == "NA":
            assert not out["g"][1][i]
        else:
            assert out["g"][1][i] and out["g"][0][i] == g[1]
        assert out["n"][0][i] == nrows and out["c"][0][i] == nvalid
        if s is None:
            assert not out["s"][1][i] and not out["lo"][1][i]
        else:
            assert abs(out["s"][0][i] - s) <= 1e-12 * sum(abs(vals[j]) for j in range(n)
                                                          if (kvalid[j] and g[0] == "v" and keys[j] + 0.0 == g[1])
                                                          or (not kvalid[j] and g[0] == "NA"))
            assert out["lo"][0][i] == lo


def test_groupby_known_answers(native, oracle):
    """test-groupby.R:1-99 and test-agg.R:1-139 through the oracle (string
    keys recoded as integers in the same order)."""
    nv = native

    def q(cols, aggs, td=None):
        return nv.collect(nv.group_agg_node("oracle", nv.Table(cols).scan(), ["g"], aggs, td))

    out = q({"g": np.array([0, 1, 0, 1, 0]), "x": np.arange(1, 6)}, [("cnt", "n", None, 0)])
    assert list(out["cnt"][0]) == [3, 2]
    out = q({"g": np.array([0, 0, 1, 1]), "x": np.array([5., 1., 8., 3.])},
            [("lo", "min", "x", 0), ("hi", "max", "x", 0), ("m", "mean", "x", 0)])
    assert list(out["lo"][0]) == [1, 3] and list(out["hi"][0]) == [5, 8] and list(out["m"][0]) == [3, 5.5]
    out = q({"g": (np.zeros(4, np.int64), np.array([1, 0, 1, 0], bool)), "x": np.array([1., 2., 3., 4.])},
            [("total", "sum", "x", 0)], "/tmp")
    assert list(out["total"][0]) == [4, 6] and list(out["g"][1]) == [True, False]
    out = q({"g": np.array([0, 0, 1]), "x": (np.array([1., 0., 3.]), np.array([1, 0, 1], bool))},
            [("s", "sum", "x", 0), ("m", "mean", "x", 0), ("srm", "sum", "x", 1)])
    assert list(out["s"][1]) == [False, True] and out["srm"][0][0] == 1
    native.core().vx_clear_warnings()
    out = q({"g": np.array([0, 0]), "x": (np.zeros(2), np.zeros(2, bool))},
            [("s", "sum", "x", 1), ("m", "mean", "x", 1), ("lo", "min", "x", 1), ("hi", "max", "x", 1)])
    assert out["s"][0][0] == 0 and np.isnan(out["m"][0][0])
    assert out["lo"][0][0] == np.inf and out["hi"][0][0] == -np.inf
    assert native.core().vx_warning_count() == 2
    out = q({"g": np.array([0, 0, 1]), "x": np.array([100, 200, 50])}, [("s", "sum", "x", 0)])
    assert list(out["s"][0]) == [300, 50]


def test_groupby_nan_and_zero_order_rules(native, oracle):
    nv = native
    x = np.array([np.nan, 1.0, 2.0, np.nan, 0.0, -0.0, -0.0, 0.0])
    g = np.array([0, 0, 1, 1, 2, 2, 3, 3])
    out = nv.collect(nv.group_agg_node("oracle", nv.Table({"g": g, "x": x}).scan(), ["g"],
                                       [("lo", "min", "x", 0), ("hi", "max", "x", 0)], None))
    lo, hi = out["lo"][0], out["hi"][0]
    assert np.isnan(lo[0]) and np.isnan(hi[0])          # NaN first: sticks
    assert lo[1] == 2.0 and hi[1] == 2.0                 # NaN later: ignored
    assert not np.signbit(lo[2]) and np.signbit(lo[3])   # first zero wins
    assert not np.signbit(hi[2]) and np.signbit(hi[3])


# ---------------------------------------------------------------- join

def brute_join(lk, lv, rk, rv, kind, rg):
    """nested loop with the reference's order rules"""
    out = []
    n = len(lk)
    matched_build = set()
    for b0 in range(0, n, rg if rg else max(n, 1)):
        rows = range(b0, min(n, b0 + (rg if rg else n)))
        tail = []
        for i in rows:
            ms = [j for j in range(len(rk) - 1, -1, -1)
                  if lv[i] and rv[j] and lk[i] == rk[j]]
            if kind == "semi":
                if ms:
                    out.append((i, None))
            elif kind == "anti":
                if not ms:
                    out.append((i, None))
            else:
                for j in ms:
                    out.append((i, j))
                    matched_build.add(j)
                if not ms and kind in ("left", "full"):
                    tail.append((i, None))
        out.extend(tail)
    if kind == "full":
        out.extend((None, j) for j in range(len(rk)) if j not in matched_build)
    return out


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
def test_join_brute_force(native, oracle, kind):
    nv = native
    r = rng(4)
    nl, nr = 400, 150
    lk, rk = r.integers(0, 60, nl), r.integers(0, 60, nr)
    lv, rv = with_na(r, nl, 0.1), with_na(r, nr, 0.1)
    left = nv.Table({"k": (lk, lv), "i": np.arange(nl)}, rg_rows=64)
    right = nv.Table({"k": (rk, rv), "j": np.arange(nr)})
    out = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)]))
    exp = brute_join(lk, lv, rk, rv, kind, 64)
    i_vals, i_valid = out["i"]
    assert len(i_vals) == len(exp)
    for row, (i, j) in enumerate(exp):
        if i is None:
            assert not i_valid[row]
        else:
            assert i_valid[row] and i_vals[row] == i
        if kind in ("inner", "left", "full"):
            jv, jm = out["j"][0][row], out["j"][1][row]
            assert (j is None and not jm) or (jm and jv == j)


def test_join_known_answers(native, oracle):
    nv = native
    L = nv.Table({"id": np.array([1., 2.]), "x": np.array([10., 20.])})
    R = nv.Table({"id": np.array([1., 1., 2.]), "y": np.array([100., 101., 200.])})
    out = nv.collect(nv.join_node("oracle", L.scan(), R.scan(), "inner", [(0, 0)]))
    assert list(out["y"][0]) == [101, 100, 200]
    L = nv.Table({"id": np.array([1, 2]), "val": np.array([10., 20.])})
    R = nv.Table({"id": np.array([1, 2]), "val": np.array([100., 200.])})
    out = nv.collect(nv.join_node("oracle", L.scan(), R.scan(), "left", [(0, 0)]))
    assert list(out) == ["id", "val.x", "val.y"]


# ---------------------------------------------------------------- string keys

def str_keys(r, n, n_distinct, p_na=0.1):
    """random strings from a pool of n_distinct values: empty, one byte,
    shared prefixes, multi-byte UTF-8 and lengths up to 40 bytes"""
    pool = ["", "a", "ab", "abc", "b", "ba", "Z", "é", "ü1", "x" * 40] + [
        f"k{i:04d}" + "-" * (i % 7) for i in range(max(0, n_distinct - 10))]
    pool = pool[:n_distinct]
    vals = np.array([pool[i] for i in r.integers(0, len(pool), n)], dtype=object)
    return vals, r.random(n) >= p_na


def test_groupby_string_key_known_answers(native, oracle):
    """test-groupby.R:1-60 with their string keys"""
    nv = native

    def q(cols, aggs, td=None):
        return nv.collect(nv.group_agg_node("oracle", nv.Table(cols).scan(), ["g"], aggs, td))

    out = q({"g": np.array(["a", "b", "a", "b", "a"]), "x": np.arange(1, 6)}, [("cnt", "n", None, 0)])
    assert list(out["g"][0]) == ["a", "b"] and list(out["cnt"][0]) == [3, 2]
    out = q({"g": np.array(["a", "a", "b", "b"]), "x": np.array([1., 2., 3., 4.])}, [("total", "sum", "x", 0)])
    assert list(out["g"][0]) == ["a", "b"] and list(out["total"][0]) == [3, 7]
    out = q({"g": np.array(["a", "a", "b", "b"]), "x": np.array([5., 1., 8., 3.])},
            [("lo", "min", "x", 0), ("hi", "max", "x", 0)])
    assert list(out["lo"][0]) == [1, 3] and list(out["hi"][0]) == [5, 8]
    for td in (None, "/tmp"):
        out = q({"g": (np.array(["a", "", "a", ""]), np.array([1, 0, 1, 0], bool)), "x": np.array([1., 2., 3., 4.])},
                [("total", "sum", "x", 0)], td)
        assert list(out["g"][1]) == [True, False] and list(out["total"][0]) == [4, 6]


@pytest.mark.parametrize("order", ["first", "sorted"])
def test_groupby_string_key_brute_force(native, oracle, order):
    """first-occurrence order (hash path) or byte order with shorter
    prefixes first and NA last (sort path, src/sort.c:58-66)"""
    nv = native
    r = rng(31)
    n = 3000
    g, gv = str_keys(r, n, 40)
    x = r.normal(size=n)
    out = nv.collect(nv.group_agg_node("oracle", nv.Table({"g": (g, gv), "x": x}, rg_rows=700).scan(), ["g"],
                                       [("n", "n", None, 0), ("s", "sum", "x", 0)],
                                       "/tmp" if order == "sorted" else None))
    keys = [g[i].encode() if gv[i] else None for i in range(n)]
    seen = list(dict.fromkeys(keys))
    if order == "sorted":
        seen = sorted(k for k in seen if k is not None) + ([None] if None in seen else [])
    got = [v.encode() if ok else None for v, ok in zip(*out["g"])]
    assert got == seen
    for i, k in enumerate(seen):
        rows = [j for j in range(n) if keys[j] == k]
        assert out["n"][0][i] == len(rows)
        assert math.isclose(out["s"][0][i], math.fsum(x[rows]), rel_tol=1e-12, abs_tol=1e-12)


def test_groupby_string_and_int_keys(native, oracle):
    """two grouping columns, one of them strings (test-groupby.R:62-73: four groups)"""
    nv = native
    out = nv.collect(nv.group_agg_node("oracle", nv.Table({"a": np.array(["x", "x", "y", "y"]),
                                                           "b": np.array([1, 2, 1, 2]),
                                                           "v": np.array([10., 20., 30., 40.])}).scan(),
                                       ["a", "b"], [("s", "sum", "v", 0)], "/tmp"))
    assert list(out["a"][0]) == ["x", "x", "y", "y"] and list(out["b"][0]) == [1, 2, 1, 2]
    assert list(out["s"][0]) == [10, 20, 30, 40]


def test_join_string_key_known_answers(native, oracle):
    """test-joins.R:135-163 (a string key; a number and a string key)"""
    nv = native
    L = nv.Table({"name": np.array(["alice", "bob"]), "x": np.array([1., 2.])})
    R = nv.Table({"name": np.array(["bob", "carol"]), "y": np.array([20., 30.])})
    out = nv.collect(nv.join_node("oracle", L.scan(), R.scan(), "inner", [(0, 0)]))
    assert list(out["name"][0]) == ["bob"] and list(out["y"][0]) == [20]
    L = nv.Table({"a": np.array([1., 1., 2.]), "b": np.array(["x", "y", "x"]), "v1": np.array([10., 20., 30.])})
    R = nv.Table({"a": np.array([1., 2.]), "b": np.array(["x", "x"]), "v2": np.array([100., 300.])})
    out = nv.collect(nv.join_node("oracle", L.scan(), R.scan(), "inner", [(0, 0), (1, 1)]))
    assert sorted(out["v1"][0]) == [10, 30]


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
def test_join_string_key_brute_force(native, oracle, kind):
    nv = native
    r = rng(33)
    nl, nr = 300, 120
    lk, lv = str_keys(r, nl, 30)
    rk, rv = str_keys(r, nr, 30)
    left = nv.Table({"k": (lk, lv), "i": np.arange(nl)}, rg_rows=64)
    right = nv.Table({"k": (rk, rv), "j": np.arange(nr)})
    out = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)]))
    exp = brute_join(list(lk), lv, list(rk), rv, kind, 64)
    assert len(out["i"][0]) == len(exp)
    for row, (i, j) in enumerate(exp):
        assert (i is None and not out["i"][1][row]) or out["i"][0][row] == i
        if kind in ("inner", "left", "full"):
            assert (j is None and not out["j"][1][row]) or out["j"][0][row] == j
        if kind != "full" or i is not None:
            assert out["k"][1][row] == lv[i] and (not lv[i] or out["k"][0][row] == lk[i])
        else:
            assert out["k"][1][row] == rv[j] and (not rv[j] or out["k"][0][row] == rk[j])
