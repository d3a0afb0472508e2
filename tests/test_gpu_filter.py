"""GPU FilterNode vs the oracle: selection vectors bit-exact, batch by batch
(src/filter.c:13-84 semantics: NA predicate drops the row, the NaN compare
quirk of src/scalar_ops.c:125-141, sel composition, empty batches skipped)."""
import numpy as np
import pytest

from helpers import assert_col_exact, dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu

PREDS = [
    ("cmp", ">", ("col", "x"), ("dbl", 0.0)),
    ("cmp", "<=", ("col", "k"), ("int", 500)),
    ("cmp", ">", ("col", "k"), ("dbl", 499.5)),
    ("cmp", "==", ("col", "x"), ("dbl", 0.0)),
    ("cmp", "!=", ("col", "x"), ("dbl", 0.0)),
    ("cmp", ">=", ("dbl", 0.5), ("col", "x")),
    ("cmp", "<", ("int", 300), ("col", "k")),
    ("and", ("cmp", ">", ("col", "x"), ("dbl", -1)), ("cmp", "<", ("col", "x"), ("dbl", 1))),
    ("or", ("cmp", "<", ("col", "k"), ("int", 100)), ("is_na", ("col", "x"))),
    ("not", ("col", "b")),
    ("not", ("is_na", ("col", "k"))),
    ("cmp", ">", ("arith", "*", ("col", "x"), ("col", "k")), ("dbl", 10.0)),
    ("cmp", "==", ("arith", "%", ("col", "k"), ("int", 7)), ("int", 3)),
    ("cmp", "==", ("arith", "/", ("col", "k"), ("arith", "-", ("col", "k"), ("int", 500))), ("int", 0)),
    ("cmp", "<", ("neg", ("col", "x")), ("dbl", 0.0)),
    ("cmp", "==", ("col", "b"), ("lgl", True)),
    ("cmp", ">", ("col", "b"), ("col", "x")),
    ("or", ("cmp", ">", ("col", "x"), ("na",)), ("lgl", True)),
]


STR_POOL = ["", "a", "a\x00", "ab", "abc", "abd", "b", "é", "ÿ", "z", "key-001", "key-010", "key-1", "x" * 40]
STR_PREDS = [
    ("cmp", ">", ("col", "s"), ("str", "ab")),
    ("cmp", "<=", ("str", "key-010"), ("col", "s")),
    ("cmp", "==", ("col", "s"), ("col", "t")),
    ("cmp", "<", ("col", "s"), ("col", "t")),
    ("cmp", "!=", ("col", "s"), ("str", "")),
    ("cmp", ">=", ("col", "t"), ("str", "é")),
    ("cmp", "==", ("col", "s"), ("na_str",)),
    ("and", ("cmp", ">", ("col", "s"), ("str", "a")), ("cmp", "<", ("col", "k"), ("int", 500))),
    ("or", ("cmp", "==", ("col", "s"), ("str", "z")), ("cmp", "<", ("col", "t"), ("col", "s"))),
    ("not", ("cmp", "<", ("col", "s"), ("str", "key-1"))),
]


def make_str_table(nv, n, seed, rg):
    def mk():
        r = rng(seed)
        s = np.array([STR_POOL[i] for i in r.integers(0, len(STR_POOL), n)], dtype=object)
        t = np.array([STR_POOL[i] for i in r.integers(0, len(STR_POOL), n)], dtype=object)
        return nv.Table({"s": (s, with_na(r, n, 0.05)), "t": (t, with_na(r, n, 0.03)),
                         "k": r.integers(0, 1000, n), "id": np.arange(n)}, rg_rows=rg)
    return mk


def make_table(nv, n, seed, rg):
    r = rng(seed)
    return lambda: nv.Table({
        "k": (r2(seed).integers(0, 1000, n), with_na(r2(seed + 1), n, 0.05)),
        "x": (dbl_with_specials(r2(seed + 2), n), with_na(r2(seed + 3), n, 0.05)),
        "b": (r2(seed + 4).random(n) < 0.5, with_na(r2(seed + 5), n, 0.05)),
        "id": np.arange(n),
    }, rg_rows=rg)


def r2(seed):
    return rng(seed)


def run(nv, engine, table, pred, chain=None):
    node = nv.filter_node(engine, table.scan(), pred)
    if chain is not None:
        node = nv.filter_node(engine, node, chain)
    return nv.next_batches(node)


def same_batches(got, exp):
    assert len(got) == len(exp), f"{len(got)} batches vs {len(exp)}"
    for g, e in zip(got, exp):
        assert g["n_rows"] == e["n_rows"]
        assert (g["sel"] is None) == (e["sel"] is None)
        assert np.array_equal(g["sel"], e["sel"]), "selection vectors differ"
        for k in e["cols"]:
            assert_col_exact(k, g["cols"][k], e["cols"][k])


@pytest.mark.parametrize("pi", range(len(PREDS)))
@pytest.mark.parametrize("rg", [0, 997, 4096])
def test_filter_matches_oracle(gpu, oracle, pi, rg):
    mk = make_table(gpu, 20011, 100 + pi, rg)
    same_batches(run(gpu, "gpu", mk(), PREDS[pi]), run(gpu, "oracle", mk(), PREDS[pi]))


@pytest.mark.parametrize("rg", [0, 1000])
def test_filter_composes_selection(gpu, oracle, rg):
    mk = make_table(gpu, 9001, 7, rg)
    p1 = ("cmp", ">", ("col", "x"), ("dbl", -0.5))
    p2 = ("or", ("cmp", "<", ("col", "k"), ("int", 700)), ("col", "b"))
    same_batches(run(gpu, "gpu", mk(), p1, p2), run(gpu, "oracle", mk(), p1, p2))


def test_filter_all_and_none(gpu, oracle):
    mk = make_table(gpu, 5000, 3, 700)
    for p in [("cmp", ">", ("col", "id"), ("int", -1)), ("cmp", ">", ("col", "id"), ("int", 10**9)),
              ("cmp", "<", ("col", "id"), ("int", 1))]:
        same_batches(run(gpu, "gpu", mk(), p), run(gpu, "oracle", mk(), p))


def test_filter_config1_size(gpu, oracle):
    """1M rows in 10 row groups of 100k, x > 0 (the plain CPU-runnable query's filter)."""
    n = 1_000_000
    r = rng(2024)
    x = r.normal(size=n)
    key = r.integers(0, 1000, n)
    t = lambda: gpu.Table({"id": np.arange(n), "key": key, "x": x}, rg_rows=100_000)
    p = ("cmp", ">", ("col", "x"), ("dbl", 0.0))
    got = run(gpu, "gpu", t(), p)
    exp = run(gpu, "oracle", t(), p)
    assert len(got) == 10
    same_batches(got, exp)
    # closed form: the ascending indices with x > 0 in each row group
    for i, g in enumerate(got):
        assert np.array_equal(g["sel"], np.flatnonzero(x[i * 100_000:(i + 1) * 100_000] > 0))


def test_filter_known_answers(gpu):
    """test-filter.R:1-69 transcribed."""
    nv = gpu
    t = nv.Table({"x": np.array([1, 2, 3, 4, 5]), "y": np.array([10., 20., 30., 40., 50.])})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 3))))
    assert list(out["x"][0]) == [4, 5] and list(out["y"][0]) == [40, 50]
    t = nv.Table({"x": np.arange(1, 11)})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("and", ("cmp", ">", ("col", "x"), ("dbl", 3)),
                                                     ("cmp", "<", ("col", "x"), ("dbl", 8)))))
    assert list(out["x"][0]) == [4, 5, 6, 7]
    t = nv.Table({"x": np.arange(1, 7)})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("or", ("cmp", "<", ("col", "x"), ("dbl", 3)),
                                                     ("cmp", ">", ("col", "x"), ("dbl", 4)))))
    assert list(out["x"][0]) == [1, 2, 5, 6]
    t = nv.Table({"x": np.array([True, False, True, False])})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("not", ("col", "x"))))
    assert list(out["x"][0]) == [0, 0]
    t = nv.Table({"x": (np.array([1, 0, 3, 0, 5]), np.array([1, 0, 1, 0, 1], bool))})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("not", ("is_na", ("col", "x")))))
    assert list(out["x"][0]) == [1, 3, 5]
    t = nv.Table({"x": (np.array([1, 0, 3]), np.array([1, 0, 1], bool)), "y": np.array([10., 20., 30.])})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 1))))
    assert list(out["x"][0]) == [3]
    t = nv.Table({"x": np.array([1., 2., 3.])})
    out = nv.collect(nv.filter_node("gpu", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", 100))))
    assert len(out["x"][0]) == 0


def test_filter_type_errors(gpu):
    nv = gpu
    t = nv.Table({"x": np.array([1., 2.])})
    with pytest.raises(nv.VectraError, match="filter predicate must evaluate to boolean"):
        nv.collect(nv.filter_node("gpu", t.scan(), ("col", "x")))
    t = nv.Table({"x": np.array([1., 2.])})
    with pytest.raises(nv.VectraError, match="column not found: zz"):
        nv.collect(nv.filter_node("gpu", t.scan(), ("cmp", ">", ("col", "zz"), ("dbl", 1))))
    t = nv.Table({"x": np.array([1., 2.])})
    with pytest.raises(nv.VectraError, match="bool: operands must be boolean"):
        nv.collect(nv.filter_node("gpu", t.scan(), ("and", ("col", "x"), ("lgl", True))))


@pytest.mark.parametrize("pi", range(len(STR_PREDS)))
def test_string_predicates_match_oracle(gpu, oracle, pi):
    """string comparisons (src/scalar_ops.c:77-106) against literals, the
    NA string and other string columns, inside & / | / !: a -1 / 0 / 1
    column per comparison computed on the device, the program compares it
    with 0; selection vectors bit-exact over row groups"""
    mk = make_str_table(gpu, 9001, 300 + pi, 1500)
    same_batches(run(gpu, "gpu", mk(), STR_PREDS[pi]), run(gpu, "oracle", mk(), STR_PREDS[pi]))


def test_string_predicate_below_group_by_and_join(gpu, oracle):
    """a string filter under a GPU group-by (the device scan feed leaves it
    to the FilterNode) and under a GPU join's probe side"""
    mk = make_str_table(gpu, 20_000, 911, 3000)
    pred = ("cmp", ">=", ("col", "s"), ("str", "b"))
    out = {}
    for engine in ("gpu", "oracle"):
        node = gpu.filter_node(engine, mk().scan(), pred)
        out[engine] = gpu.collect(gpu.group_agg_node(engine, node, ["t"], [("n", "n", None, 0),
                                                                         ("sk", "sum", "k", 0)], "/tmp"))
    assert_col_exact("t", out["gpu"]["t"], out["oracle"]["t"])
    assert_col_exact("n", out["gpu"]["n"], out["oracle"]["n"])
    assert_col_exact("sk", out["gpu"]["sk"], out["oracle"]["sk"])
    dim = {"dk": np.arange(0, 1000, 3), "w": np.arange(0, 1000, 3) * 0.5}
    for kind in ("inner", "left"):
        got = gpu.next_batches(gpu.join_node("gpu", gpu.filter_node("gpu", mk().scan(), pred),
                                             gpu.Table(dim).scan(), kind, [(2, 0)]))
        exp = gpu.next_batches(gpu.join_node("oracle", gpu.filter_node("oracle", mk().scan(), pred),
                                             gpu.Table(dim).scan(), kind, [(2, 0)]))
        assert len(got) == len(exp)
        for g, e in zip(got, exp):
            for c in e["cols"]:
                assert_col_exact(c, g["cols"][c], e["cols"][c])


def test_string_against_number_is_an_error(gpu):
    mk = make_str_table(gpu, 100, 3, 0)
    node = gpu.filter_node("gpu", mk().scan(), ("cmp", ">", ("col", "s"), ("int", 3)))
    with pytest.raises(gpu.VectraError, match="cannot coerce string in arithmetic/comparison"):
        gpu.next_batches(node)
