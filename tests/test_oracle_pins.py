"""Pins the oracle's arithmetic, coercion, comparison, aggregate and join
chain rules against independent restatements, without a GPU.

Each expected value here is computed from the C semantics the reference
relies on, not from the oracle's code: Python integers reduced modulo 2^64
for int64 wrap, C's truncating division and remainder (sign of the
dividend), numpy's IEEE double arithmetic and C's fmod, and the three-way
sign rule of src/scalar_ops.c:125-141.  A dropped NA rule, a floor instead
of a truncation, a swapped operand or a wrong coercion fails one of them.
"""
import ctypes as C
import math
import zlib

import numpy as np
import pytest

from helpers import rng

I64_MIN, I64_MAX = -(1 << 63), (1 << 63) - 1


def seed_of(*parts) -> int:
    return zlib.crc32(repr(parts).encode()) & 0xFFFF


def wrap64(x: int) -> int:
    return (x + (1 << 63)) % (1 << 64) - (1 << 63)


def c_div(a: int, b: int) -> int:
    q = abs(a) // abs(b)
    return q if (a < 0) == (b < 0) else -q


def c_mod(a: int, b: int) -> int:
    return a - b * c_div(a, b)


def eval_oracle(nv, oracle, cols: dict, tree):
    """Evaluates one expression over a one-batch table through the oracle:
    (values, valid, result type)."""
    t = nv.Table(cols)
    scan = t.scan()
    b = nv.BATCH_P()
    assert nv.core().vx_try_next(scan, C.byref(b)) == 0
    e = nv.expr(tree)
    out = C.POINTER(nv.VecArray)()
    rc = oracle.oracle_try_expr_eval(e, b, C.byref(out))
    nv.core().vec_expr_free(e)
    try:
        if rc != 0:
            raise nv.VectraError(nv.core().vx_last_error().decode())
        a = out.contents
        n = a.length
        if a.validity:
            bits = np.ctypeslib.as_array(a.validity, shape=((n + 7) // 8,))
            valid = np.unpackbits(bits, bitorder="little")[:n].astype(bool)
        else:
            valid = np.ones(n, bool)
        if a.type == nv.VEC_INT64:
            vals = np.ctypeslib.as_array(a.buf.i64, shape=(n,)).copy()
        elif a.type == nv.VEC_DOUBLE:
            vals = np.ctypeslib.as_array(a.buf.dbl, shape=(n,)).copy()
        else:
            vals = np.ctypeslib.as_array(a.buf.bln, shape=(n,)).copy()
        return vals, valid, a.type
    finally:
        if out:
            oracle.oracle_array_free(out)
        nv.core().vec_batch_free(b)
        nv.core().vx_node_free(scan)


def column(r, kind, n, p_na=0.15):
    """values with the edge cases of each type, and an NA mask"""
    if kind == "int":
        v = r.integers(-50, 50, n)
        edge = np.array([0, 1, -1, 7, -7, I64_MIN, I64_MAX, 1 << 53, (1 << 53) + 1, -(1 << 62)], np.int64)
    elif kind == "dbl":
        v = r.normal(size=n) * 20
        edge = np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 1e308, -1e308, 5e-324, 2.0 ** 53, 7.0])
    else:
        v = r.random(n) < 0.5
        edge = np.array([True, False])
    idx = r.integers(0, n, n // 3)
    v = v.copy()
    v[idx] = edge[r.integers(0, len(edge), len(idx))]
    valid = r.random(n) >= p_na
    return v, valid


def pyval(kind, v):
    if kind == "int":
        return int(v)
    if kind == "bool":
        return int(bool(v))
    return float(v)


TYPE_PAIRS = [("int", "int"), ("int", "dbl"), ("dbl", "int"), ("dbl", "dbl"), ("bool", "bool"),
              ("bool", "int"), ("int", "bool"), ("bool", "dbl"), ("dbl", "bool")]


def common(ta, tb):
    if "dbl" in (ta, tb):
        return "dbl"
    return "int"   # int64, or bool promoted to int64 for arithmetic and comparison


# ---------------------------------------------------------------- arithmetic

def arith_expected(op, ct, a, b):
    """one value of vec_arith (src/scalar_ops.c:11-66): None = NA"""
    if ct == "int":
        if op == "+":
            return wrap64(a + b)
        if op == "-":
            return wrap64(a - b)
        if op == "*":
            return wrap64(a * b)
        if b == 0:
            return None
        return wrap64(c_div(a, b)) if op == "/" else wrap64(c_mod(a, b))
    x, y = np.float64(a), np.float64(b)
    with np.errstate(all="ignore"):
        return float({"+": x + y, "-": x - y, "*": x * y, "/": x / y, "%": np.fmod(x, y)}[op])


@pytest.mark.parametrize("op", ["+", "-", "*", "/", "%"])
@pytest.mark.parametrize("types", TYPE_PAIRS)
def test_arith_brute_force(native, oracle, op, types):
    nv = native
    ta, tb = types
    r = rng(seed_of(op, ta, tb))
    n = 600
    a, av = column(r, ta, n)
    b, bv = column(r, tb, n)
    vals, valid, rtype = eval_oracle(nv, oracle, {"a": (a, av), "b": (b, bv)}, ("arith", op, ("col", "a"), ("col", "b")))
    ct = common(ta, tb)
    assert rtype == (nv.VEC_DOUBLE if ct == "dbl" else nv.VEC_INT64)
    for i in range(n):
        if not (av[i] and bv[i]):
            assert not valid[i], f"row {i}: NA operand must give NA"
            continue
        exp = arith_expected(op, ct, pyval(ta, a[i]) if ct == "int" else float(pyval(ta, a[i])),
                             pyval(tb, b[i]) if ct == "int" else float(pyval(tb, b[i])))
        if exp is None:
            assert not valid[i], f"row {i}: integer {op} by zero must give NA"
            continue
        assert valid[i]
        got = vals[i]
        if ct == "int":
            assert int(got) == exp, f"row {i}: {a[i]} {op} {b[i]} = {got}, expected {exp}"
        elif math.isnan(exp):
            assert math.isnan(got), f"row {i}: {a[i]} {op} {b[i]} = {got}, expected NaN"
        else:
            assert np.float64(got).tobytes() == np.float64(exp).tobytes(), \
                f"row {i}: {a[i]} {op} {b[i]} = {got!r}, expected {exp!r}"


def test_arith_int64_edges(native, oracle):
    """wrap on overflow, truncating division, remainder with the dividend's
    sign, INT64_MIN / -1 wrapping and % -1 giving 0"""
    nv = native
    a = np.array([7, -7, 7, -7, I64_MAX, I64_MIN, I64_MIN, I64_MIN, 5, 0], np.int64)
    b = np.array([2, 2, -2, -2, 1, -1, -1, 1, 0, 0], np.int64)
    exp = {
        "+": [9, -5, 5, -9, I64_MIN, I64_MAX, I64_MAX, I64_MIN + 1, 5, 0],
        "-": [5, -9, 9, -5, I64_MAX - 1, I64_MIN + 1, I64_MIN + 1, I64_MAX, 5, 0],
        "*": [14, -14, -14, 14, I64_MAX, I64_MIN, I64_MIN, I64_MIN, 0, 0],
        "/": [3, -3, -3, 3, I64_MAX, I64_MIN, I64_MIN, I64_MIN, None, None],
        "%": [1, -1, 1, -1, 0, 0, 0, 0, None, None],
    }
    for op, e in exp.items():
        vals, valid, _ = eval_oracle(nv, oracle, {"a": a, "b": b}, ("arith", op, ("col", "a"), ("col", "b")))
        for i, x in enumerate(e):
            if x is None:
                assert not valid[i]
            else:
                assert valid[i] and int(vals[i]) == x, f"{a[i]} {op} {b[i]}: {vals[i]} != {x}"


def test_negate_and_is_na(native, oracle):
    nv = native
    a = np.array([3, -3, 0, I64_MIN, I64_MAX], np.int64)
    av = np.array([1, 1, 1, 1, 0], bool)
    vals, valid, t = eval_oracle(nv, oracle, {"a": (a, av)}, ("neg", ("col", "a")))
    assert t == nv.VEC_INT64 and list(valid) == [True, True, True, True, False]
    assert [int(v) for v in vals[:4]] == [-3, 3, 0, I64_MIN]
    x = np.array([1.5, -0.0, 0.0, np.nan, np.inf])
    vals, valid, t = eval_oracle(nv, oracle, {"x": x}, ("neg", ("col", "x")))
    assert t == nv.VEC_DOUBLE
    assert np.array_equal(np.signbit(vals), [True, False, True, np.signbit(-np.float64(np.nan)), True])
    assert vals[0] == -1.5 and np.isnan(vals[3]) and vals[4] == -np.inf
    vals, valid, t = eval_oracle(nv, oracle, {"a": (a, av)}, ("is_na", ("col", "a")))
    assert t == nv.VEC_BOOL and valid.all() and list(vals) == [0, 0, 0, 0, 1]
    with pytest.raises(nv.VectraError, match="negate: unsupported type"):
        eval_oracle(nv, oracle, {"b": np.array([True])}, ("neg", ("col", "b")))


# ---------------------------------------------------------------- comparison

def sign3(a, b):
    """(a > b) - (a < b) as C computes it: 0 when either side is NaN"""
    return int(a > b) - int(a < b)


OPS = {"==": lambda c: c == 0, "!=": lambda c: c != 0, "<": lambda c: c < 0, "<=": lambda c: c <= 0,
       ">": lambda c: c > 0, ">=": lambda c: c >= 0}


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("types", TYPE_PAIRS)
def test_compare_brute_force(native, oracle, op, types):
    """int64 against double compares in double (2^53 + 1 equals 2^53), bool
    against bool as integers, NaN satisfies ==, <= and >="""
    nv = native
    ta, tb = types
    r = rng(seed_of(op, ta, tb, 7))
    n = 600
    a, av = column(r, ta, n)
    b, bv = column(r, tb, n)
    vals, valid, rtype = eval_oracle(nv, oracle, {"a": (a, av), "b": (b, bv)}, ("cmp", op, ("col", "a"), ("col", "b")))
    assert rtype == nv.VEC_BOOL
    ct = common(ta, tb)
    for i in range(n):
        if not (av[i] and bv[i]):
            assert not valid[i]
            continue
        x, y = pyval(ta, a[i]), pyval(tb, b[i])
        if ct == "dbl":
            x, y = float(x), float(y)
        assert valid[i] and bool(vals[i]) == OPS[op](sign3(x, y)), f"row {i}: {a[i]} {op} {b[i]}"


STR_POOL = ["", "\x00", "a", "a\x00", "ab", "abc", "b", "é", "ÿ", "z", "zz", "a" * 9, "Z"]


@pytest.mark.parametrize("op", list(OPS))
def test_compare_strings_against_bytes_order(native, oracle, op):
    """two string columns, and a column against a literal on either side:
    Python orders bytes as the reference does (memcmp over the shorter
    length, then the shorter first, src/scalar_ops.c:77-106), so each row
    is checked against bytes comparison; NA on either side is NA"""
    nv = native
    r = rng(seed_of(op, "str", 3))
    n = 400
    a = np.array([STR_POOL[i] for i in r.integers(0, len(STR_POOL), n)], dtype=object)
    b = np.array([STR_POOL[i] for i in r.integers(0, len(STR_POOL), n)], dtype=object)
    av, bv = r.random(n) > 0.1, r.random(n) > 0.1
    cols = {"a": (a, av), "b": (b, bv)}
    key = lambda x: x.encode()
    for tree, rhs in ((("cmp", op, ("col", "a"), ("col", "b")), None),
                      (("cmp", op, ("col", "a"), ("str", "ab")), "ab"),
                      (("cmp", op, ("str", "a\x00"), ("col", "a")), "a\x00")):
        if rhs == "a\x00":   # literals end at their first NUL byte, as C strings do
            rhs = "a"
        vals, valid, rtype = eval_oracle(nv, oracle, cols, tree)
        assert rtype == nv.VEC_BOOL
        for i in range(n):
            if tree[2][0] == "str":
                ok, c = av[i], sign3(key(rhs), key(a[i]))
            elif rhs is None:
                ok, c = av[i] and bv[i], sign3(key(a[i]), key(b[i]))
            else:
                ok, c = av[i], sign3(key(a[i]), key(rhs))
            assert bool(valid[i]) == bool(ok), f"row {i}"
            if ok:
                assert bool(vals[i]) == OPS[op](c), f"row {i}: {tree}"


def test_compare_int64_double_rounding(native, oracle):
    """an int64 operand is converted to double first: 2^53 + 1 == 2^53.0"""
    nv = native
    a = np.array([(1 << 53) + 1, (1 << 53) + 1, 3], np.int64)
    b = np.array([2.0 ** 53, 2.0 ** 53 + 2, 3.0])
    vals, valid, _ = eval_oracle(nv, oracle, {"a": a, "b": b}, ("cmp", "==", ("col", "a"), ("col", "b")))
    assert list(vals) == [1, 0, 1]
    vals, _, _ = eval_oracle(nv, oracle, {"a": a, "b": a}, ("cmp", "==", ("col", "a"), ("int", 1 << 53)))
    assert list(vals) == [0, 0, 0]   # int64 against int64 stays exact


def test_three_valued_logic_truth_table(native, oracle):
    """R's & and | over TRUE / FALSE / NA (src/scalar_ops.c:151-185)"""
    nv = native
    T, F, N = 1, 0, None
    vals = [T, F, N]
    a = [x for x in vals for _ in vals]
    b = [y for _ in vals for y in vals]
    col = lambda xs: (np.array([bool(x) if x is not None else False for x in xs]), np.array([x is not None for x in xs]))
    exp_and = [F if (x == 0 or y == 0) else (T if (x == 1 and y == 1) else N) for x, y in zip(a, b)]
    exp_or = [T if (x == 1 or y == 1) else (F if (x == 0 and y == 0) else N) for x, y in zip(a, b)]
    for op, exp in (("and", exp_and), ("or", exp_or)):
        v, ok, _ = eval_oracle(nv, oracle, {"a": col(a), "b": col(b)}, (op, ("col", "a"), ("col", "b")))
        got = [int(x) if o else None for x, o in zip(v, ok)]
        assert got == exp, op
    v, ok, _ = eval_oracle(nv, oracle, {"a": col(a)}, ("not", ("col", "a")))
    assert [int(x) if o else None for x, o in zip(v, ok)] == [(None if x is None else 1 - x) for x in a]


# ---------------------------------------------------------------- aggregates

def agg_expected(kind, itype, vals, na_rm):
    """one group's result of src/agg_ops.c: (value or None for NA, warned).
    vals: list of python values or None (NA) in feed order."""
    has_na = any(v is None for v in vals)
    xs = [v for v in vals if v is not None]
    if kind == "n":
        return float(len(vals)), False
    if kind == "count":
        return float(len(xs)), False
    if has_na and not na_rm:
        return None, False
    if kind in ("sum", "mean"):
        if itype == "int":
            s = float(wrap64(sum(xs)))
        else:
            s = 0.0
            with np.errstate(all="ignore"):
                for x in xs:   # sequential double sum in feed order
                    s = float(np.float64(s) + np.float64(float(x)))
        if kind == "sum":
            return s, False
        return (float("nan") if not xs else float(np.float64(s) / np.float64(len(xs)))), False
    # min / max: bool inputs never update (no branch for VEC_BOOL)
    if itype == "bool" or not xs:
        return (math.inf if kind == "min" else -math.inf), True
    best = None
    for x in xs:   # strict < / >: the first of equal values wins, NaN never replaces
        if best is None or (x < best if kind == "min" else x > best):
            best = x
    return float(best), False


@pytest.mark.parametrize("itype", ["int", "bool", "dbl"])
@pytest.mark.parametrize("na_rm", [0, 1])
def test_aggregates_brute_force(native, oracle, itype, na_rm):
    """n, count, sum, mean, min, max of int64 (exact, wrapping, mean =
    double(sum) / count), bool (double sum; min / max never take a value,
    giving +Inf / -Inf with a warning) and double inputs, in first-seen
    group order"""
    nv = native
    r = rng(91 + na_rm)
    n = 1500
    g = r.integers(0, 12, n)
    g[r.random(n) < 0.03] = 12   # one group that is mostly NA below
    x, xv = column(r, itype, n, p_na=0.05)
    xv[g == 12] = r.random((g == 12).sum()) < 0.2
    if itype == "int":
        x[r.random(n) < 0.02] = I64_MAX   # sums wrap
    kinds = ["n", "count", "sum", "mean", "min", "max"]
    aggs = [(f"o_{k}", k, None if k == "n" else "x", na_rm) for k in kinds]
    nv.core().vx_clear_warnings()
    out = nv.collect(nv.group_agg_node("oracle", nv.Table({"g": g, "x": (x, xv)}, rg_rows=400).scan(), ["g"], aggs,
                                       None))
    warnings = nv.core().vx_warning_count()
    order = list(dict.fromkeys(g.tolist()))
    assert list(out["g"][0]) == order
    exp_warn = 0
    for gi, key in enumerate(order):
        rows = np.flatnonzero(g == key)
        vals = [pyval(itype, x[i]) if xv[i] else None for i in rows]
        for k in kinds:
            exp, warned = agg_expected(k, itype, vals, na_rm)
            exp_warn += warned
            got, ok = out[f"o_{k}"][0][gi], out[f"o_{k}"][1][gi]
            if exp is None:
                assert not ok, f"group {key} {k}: expected NA"
                continue
            assert ok, f"group {key} {k}: unexpected NA"
            if k in ("sum", "mean") and itype == "dbl":
                # the oracle adds in the same feed order: bit-exact
                assert np.float64(got).tobytes() == np.float64(exp).tobytes() or (math.isnan(got) and math.isnan(exp)), \
                    f"group {key} {k}: {got!r} vs {exp!r}"
            elif math.isnan(exp):
                assert math.isnan(got)
            else:
                assert got == exp and np.signbit(got) == np.signbit(exp), f"group {key} {k}: {got!r} vs {exp!r}"
    assert warnings == exp_warn


def test_int64_sum_mean_against_numpy(native, oracle):
    """int64 sum / min / max / mean against numpy's int64 reductions (which
    wrap the same way) on groups of 10^4 values near the int64 range"""
    nv = native
    r = rng(5)
    n = 40_000
    g = r.integers(0, 4, n)
    x = r.integers(-(1 << 62), 1 << 62, n)
    out = nv.collect(nv.group_agg_node("oracle", nv.Table({"g": g, "x": x}, rg_rows=7_000).scan(), ["g"],
                                       [("s", "sum", "x", 0), ("lo", "min", "x", 0), ("hi", "max", "x", 0),
                                        ("m", "mean", "x", 0)], "/tmp"))
    for gi, key in enumerate(sorted(set(g.tolist()))):
        xs = x[g == key]
        with np.errstate(over="ignore"):
            s = np.add.reduce(xs, dtype=np.int64)
        assert out["g"][0][gi] == key
        assert out["s"][0][gi] == float(s)
        assert out["lo"][0][gi] == float(xs.min()) and out["hi"][0][gi] == float(xs.max())
        assert out["m"][0][gi] == float(s) / len(xs)


# ---------------------------------------------------------------- coerced join chains

def brute_coerced_join(lk, lv, ltype, rk, rv, rtype, kind, rg):
    """nested loop with the reference's chain rule for a probe key coerced to
    the build key's type (src/join.c:268-289, 385-432): the first match is
    the highest build row whose key equals the coerced probe key; inner
    joins keep every equal row; left / full joins keep a later (lower) equal
    row only when the ORIGINAL probe key equals the build key array's bytes:
    an int64 probe against the build double's 64 bits, a bool probe's byte
    against byte `row` of the build key array."""
    ct = "dbl" if "dbl" in (ltype, rtype) else "int"
    conv = (lambda v: float(v)) if ct == "dbl" else (lambda v: int(v))
    bk = np.array([conv(v) for v in rk], dtype=np.float64 if ct == "dbl" else np.int64)
    kbytes = bk.tobytes()
    kbits = np.frombuffer(kbytes, np.int64)
    out, matched_build = [], set()
    n = len(lk)
    for b0 in range(0, n, rg):
        tail = []
        for i in range(b0, min(n, b0 + rg)):
            ms = []
            if lv[i]:
                p = conv(int(lk[i]) if ltype != "dbl" else lk[i])
                ms = [j for j in range(len(rk) - 1, -1, -1) if rv[j] and bk[j] == p]
            if ms and kind in ("left", "full") and ltype != ct:
                keep = [ms[0]]
                for j in ms[1:]:
                    if ltype == "bool":
                        ok = kbytes[j] == int(bool(lk[i]))
                    else:
                        ok = int(kbits[j]) == int(lk[i])
                    if ok:
                        keep.append(j)
                ms = keep
            if kind == "semi":
                if ms:
                    out.append((i, None))
            elif kind == "anti":
                if not ms:
                    out.append((i, None))
            else:
                out.extend((i, j) for j in ms)
                matched_build.update(ms)
                if not ms and kind in ("left", "full"):
                    tail.append((i, None))
        out.extend(tail)
    if kind == "full":
        out.extend((None, j) for j in range(len(rk)) if j not in matched_build)
    return out


def coerced_inputs(r, ltype, rtype, nl, nr):
    if ltype == "int":
        lk = r.integers(-3, 4, nl)
    else:
        lk = r.random(nl) < 0.5
    if rtype == "dbl":
        rk = r.choice(np.array([0.0, -0.0, 1.0, -1.0, 2.0, 3.0, 0.5, np.nan]), nr)
    else:
        rk = r.integers(-1, 3, nr)
    lv = r.random(nl) >= 0.1
    rv = np.ones(nr, bool)   # NA build rows' value bytes are unspecified in the reference
    return lk, lv, rk, rv


@pytest.mark.parametrize("kind", ["inner", "left", "full", "semi", "anti"])
@pytest.mark.parametrize("types", [("int", "dbl"), ("bool", "int"), ("bool", "dbl")])
def test_coerced_join_chain_rule(native, oracle, kind, types):
    nv = native
    ltype, rtype = types
    r = rng(seed_of(kind, types))
    nl, nr = 300, 90
    lk, lv, rk, rv = coerced_inputs(r, ltype, rtype, nl, nr)
    left = nv.Table({"k": (lk, lv), "i": np.arange(nl)}, rg_rows=64)
    right = nv.Table({"k": (rk, rv), "j": np.arange(nr)})
    out = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)]))
    exp = brute_coerced_join(lk, lv, ltype, rk, rv, rtype, kind, 64)
    got_i = [int(v) if ok else None for v, ok in zip(*out["i"])]
    assert got_i == [i for i, _ in exp]
    if kind in ("inner", "left", "full"):
        got_j = [int(v) if ok else None for v, ok in zip(*out["j"])]
        assert got_j == [j for _, j in exp]


def test_coerced_left_join_zero_rule(native, oracle):
    """int64 0 against build keys +0.0 and -0.0: both match first (==), a
    later row survives only if its bits are all zero (+0.0)"""
    nv = native
    left = nv.Table({"k": np.array([0, 0], np.int64), "i": np.arange(2)})
    for rk, exp_j in ((np.array([0.0, -0.0]), [1, 0]), (np.array([-0.0, 0.0]), [1])):
        right = nv.Table({"k": rk, "j": np.arange(2)})
        out = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), "left", [(0, 0)]))
        assert list(out["j"][0]) == exp_j * 2
        right = nv.Table({"k": rk, "j": np.arange(2)})
        out = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), "inner", [(0, 0)]))
        assert list(out["j"][0]) == [1, 0] * 2


# ---------------------------------------------------------------- hidden mutate

def test_project_mul_against_numpy(native, oracle):
    """out = a * b appended after the child's columns, NA if either side is
    NA, IEEE product otherwise (src/project.c:51-252, src/scalar_ops.c:11-66)"""
    nv = native
    r = rng(8)
    n = 5000
    a, av = column(r, "dbl", n)
    b, bv = column(r, "dbl", n)
    scan = nv.Table({"a": (a, av), "b": (b, bv)}, rg_rows=999).scan()
    node = oracle.oracle_project_mul_node_create(scan, b"tmp", b"a", b"b")
    node._keep = scan   # the scan borrows the table's buffers
    out = nv.collect(node)
    assert list(out) == ["a", "b", "tmp"]
    vals, ok = out["tmp"]
    assert np.array_equal(ok, av & bv)
    with np.errstate(all="ignore"):
        exp = a * b
    m = ok
    assert np.array_equal(np.isnan(vals[m]), np.isnan(exp[m]))
    nn = m & ~np.isnan(exp)
    assert np.array_equal(vals[nn].view(np.int64), exp[nn].view(np.int64))
