"""GPU JoinNode vs the oracle (src/join.c): output rows and their order must
match exactly -- probe order, duplicate build keys in descending build row
order, left/full unmatched rows after each batch's matches, full join's
never-matched build rows at the end."""
import numpy as np
import pytest

from helpers import assert_tables_equal, rng, with_na

pytestmark = pytest.mark.gpu


def tables(nv, seed, n_left, n_right, key_card, rg_left, ltype=np.int64, rtype=np.int64, p_na=0.05, sel=False):
    def mk():
        r = rng(seed)
        lk = r.integers(0, key_card, n_left).astype(ltype)
        rk = r.integers(0, key_card, n_right).astype(rtype)
        if ltype == np.float64:
            lk[r.random(n_left) < 0.02] = np.nan
            lk[lk == 0] = np.where(r.random((lk == 0).sum()) < 0.5, 0.0, -0.0)
        if rtype == np.float64:
            rk[r.random(n_right) < 0.02] = np.nan
        left = nv.Table({"id": (lk, with_na(r, n_left, p_na)), "x": r.normal(size=n_left),
                         "flag": r.random(n_left) < 0.5}, rg_rows=rg_left)
        right = nv.Table({"y": (r.integers(0, 10**6, n_right), with_na(r, n_right, p_na)),
                          "id": (rk, with_na(r, n_right, p_na)), "w": r.random(n_right)}, rg_rows=777)
        return left, right
    return mk


def run(nv, engine, mk, kind, sel_pred=None):
    left, right = mk()
    ln = left.scan()
    if sel_pred is not None:
        ln = nv.filter_node(engine, ln, sel_pred)
    node = nv.join_node(engine, ln, right.scan(), kind, [(0, 1)])
    return nv.next_batches(node)


def same(got, exp):
    assert len(got) == len(exp), f"{len(got)} batches vs {len(exp)}"
    for g, e in zip(got, exp):
        assert g["sel"] is None and e["sel"] is None
        assert_tables_equal(g["cols"], e["cols"])


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
@pytest.mark.parametrize("card", [50, 5000])
@pytest.mark.parametrize("rg", [0, 1000])
def test_join_matches_oracle(gpu, oracle, kind, card, rg):
    mk = tables(gpu, 31 + card, 6000, 2500, card, rg)
    same(run(gpu, "gpu", mk, kind), run(gpu, "oracle", mk, kind))


@pytest.mark.parametrize("kind", ["inner", "left", "full"])
def test_join_with_probe_selection(gpu, oracle, kind):
    mk = tables(gpu, 5, 8000, 3000, 900, 1500)
    pred = ("cmp", ">", ("col", "x"), ("dbl", 0.2))
    same(run(gpu, "gpu", mk, kind, pred), run(gpu, "oracle", mk, kind, pred))


@pytest.mark.parametrize("kind", ["inner", "left", "semi", "anti", "full"])
@pytest.mark.parametrize("lt,rt", [(np.float64, np.float64), (np.int64, np.float64), (np.float64, np.int64)])
def test_join_key_types(gpu, oracle, kind, lt, rt):
    """double keys (NaN never matches, -0 == +0) and int64 / double
    coercion, including the left/full first-match rule for a coerced probe."""
    mk = tables(gpu, 77, 4000, 1500, 300, 900, lt, rt)
    same(run(gpu, "gpu", mk, kind), run(gpu, "oracle", mk, kind))


@pytest.mark.parametrize("kind", ["inner", "left", "full", "semi", "anti"])
@pytest.mark.parametrize("types", [("bool", "int"), ("bool", "dbl"), ("int", "dbl")])
def test_join_coerced_chain_rule(gpu, oracle, kind, types):
    """a probe key coerced to the build key's type: left / full joins keep
    a later duplicate build row only when the uncoerced probe key equals the
    build key array's bytes -- a bool probe against byte `row` of the build
    array, an int64 probe against the build double's bits (src/join.c:398-411)"""
    lt, rt = types

    def mk():
        r = rng({("bool", "int"): 1, ("bool", "dbl"): 2, ("int", "dbl"): 3}[types] + 17)
        nl, nr = 5000, 700
        lk = (r.random(nl) < 0.5) if lt == "bool" else r.integers(-2, 4, nl)
        if rt == "dbl":
            rk = r.choice(np.array([0.0, -0.0, 1.0, 2.0, -1.0, 0.5, np.nan]), nr)
        else:
            rk = r.integers(-1, 3, nr)
        left = gpu.Table({"id": (lk, with_na(r, nl, 0.05)), "x": r.normal(size=nl)}, rg_rows=1024)
        # build keys all valid: the value bytes of NA build rows are not defined by the reference
        right = gpu.Table({"y": r.integers(0, 10**6, nr), "id": rk, "w": r.random(nr)})
        return left, right
    pred = ("cmp", ">", ("col", "x"), ("dbl", -0.5))
    same(run(gpu, "gpu", mk, kind), run(gpu, "oracle", mk, kind))
    same(run(gpu, "gpu", mk, kind, pred), run(gpu, "oracle", mk, kind, pred))


def test_join_unique_dimension(gpu, oracle):
    """fact -> unique-key dimension (the join-then-aggregate shape)."""
    def mk():
        r = rng(3)
        dim_keys = r.permutation(200000)[:50000].astype(np.int64) * 7 + 11
        fk = np.where(r.random(100000) < 0.9, dim_keys[r.integers(0, 50000, 100000)], -5 - r.integers(0, 10**6, 100000))
        left = gpu.Table({"fkey": fk, "value": r.lognormal(size=100000)}, rg_rows=16384)
        right = gpu.Table({"attr": r.integers(0, 100, 50000), "key": dim_keys, "weight": r.random(50000)})
        return left, right
    for kind in ("inner", "left"):
        same(run(gpu, "gpu", mk, kind), run(gpu, "oracle", mk, kind))


def test_join_known_answers(gpu):
    """test-joins.R:12-130, 314-407 transcribed."""
    nv = gpu

    def j(lcols, rcols, kind, keys=((0, 0),), rg=0):
        node = nv.join_node("gpu", nv.Table(lcols, rg_rows=rg).scan(), nv.Table(rcols).scan(), kind, list(keys))
        return nv.collect(node)

    L = {"id": np.array([1., 2., 3.]), "x": np.array([10., 20., 30.])}
    R = {"id": np.array([1., 2., 4.]), "y": np.array([100., 200., 400.])}
    out = j(L, R, "inner")
    assert list(out["id"][0]) == [1, 2] and list(out["y"][0]) == [100, 200]
    out = j(L, R, "left")
    assert list(out["id"][0]) == [1, 2, 3] and list(out["y"][1]) == [True, True, False]
    out = j(L, R, "semi")
    assert list(out) == ["id", "x"] and list(out["id"][0]) == [1, 2]
    out = j(L, R, "anti")
    assert list(out["id"][0]) == [3] and list(out["x"][0]) == [30]
    out = j({"id": np.array([1., 2.]), "x": np.array([10., 20.])},
            {"id": np.array([1., 1., 2.]), "y": np.array([100., 101., 200.])}, "inner")
    assert list(out["id"][0]) == [1, 1, 2] and list(out["y"][0]) == [101, 100, 200]
    out = j({"id": (np.array([1., 0., 3.]), np.array([1, 0, 1], bool)), "x": np.array([10., 20., 30.])},
            {"id": (np.array([1., 0., 4.]), np.array([1, 0, 1], bool)), "y": np.array([100., 200., 400.])}, "inner")
    assert list(out["id"][0]) == [1]
    out = j({"id": np.array([1., 2.]), "val": np.array([10., 20.])},
            {"id": np.array([1., 2.]), "val": np.array([100., 200.])}, "left")
    assert list(out) == ["id", "val.x", "val.y"]
    out = j({"id": np.repeat(np.arange(1, 11), 10), "x": np.arange(1, 101)},
            {"id": np.array([1, 5, 10]), "y": np.array([10, 50, 100])}, "inner", rg=10)
    assert len(out["id"][0]) == 30
    out = j({"id": np.arange(1, 51), "x": np.arange(1, 51)},
            {"id": np.array([1, 25, 50]), "y": np.array([100, 250, 500])}, "left", rg=10)
    assert len(out["id"][0]) == 50 and out["y"][1].sum() == 3
    out = j({"id": np.arange(1, 21), "x": np.arange(1, 21)},
            {"id": np.array([5, 15, 99, 100]), "y": np.array([50, 150, 990, 1000])}, "full", rg=10)
    assert len(out["id"][0]) == 22 and list(out["id"][0][-2:]) == [99, 100] and not out["x"][1][-2:].any()
    out = j({"id": np.repeat(np.arange(1, 4), 10), "x": np.arange(1, 31)},
            {"id": np.repeat(np.arange(1, 4), 5), "y": np.arange(1, 16)}, "inner", rg=10)
    assert len(out["id"][0]) == 150
    out = j({"id": np.array([1, 2, 3]), "x": np.array([10., 20., 30.])},
            {"id": np.array([1., 2., 4.]), "y": np.array([100., 200., 400.])}, "inner")
    assert list(out["id"][0]) == [1, 2]
