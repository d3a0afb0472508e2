"""The harness's ProjectNode (the stand-in for the reference's
src/project.c where R is absent) against the oracle's hidden mutate and
numpy, and its explain() line (src/r_bridge.c:613-624).  CPU only."""
import numpy as np
import pytest

from helpers import dbl_with_specials, rng, with_na


def test_hidden_mutate_matches_oracle(native, oracle):
    nv = native
    r = rng(3)
    n = 20_000
    a, av = dbl_with_specials(r, n), with_na(r, n, 0.05)
    b, bv = r.normal(size=n) * 3, with_na(r, n, 0.05)
    k = r.integers(0, 50, n)
    mk = lambda: nv.filter_node("oracle", nv.Table({"k": k, "a": (a, av), "b": (b, bv)}, rg_rows=3000).scan(),
                                ("cmp", ">", ("col", "k"), ("int", 10)))
    got = nv.collect(nv.hidden_mutate(mk(), ".vectra_tmp_s", "a", "b"))
    child = mk()
    p = oracle.oracle_project_mul_node_create(child, b".vectra_tmp_s", b"a", b"b")
    p._keep = child   # the scan below borrows the table's buffers
    exp = nv.collect(p)
    assert list(got) == list(exp) == ["k", "a", "b", ".vectra_tmp_s"]
    for c in got:
        gv, gm = got[c]
        ev, em = exp[c]
        assert np.array_equal(gm, em)
        assert np.array_equal(gv[gm].view(np.int64), ev[em].view(np.int64)) or \
            np.array_equal(np.isnan(gv[gm]), np.isnan(ev[em]))
    keep = k > 10
    with np.errstate(all="ignore"):
        prod = a[keep] * b[keep]
    m = (av & bv)[keep]
    v = got[".vectra_tmp_s"][0]
    assert np.array_equal(got[".vectra_tmp_s"][1], m)
    assert np.array_equal(np.isnan(v[m]), np.isnan(prod[m]))
    fin = m & ~np.isnan(prod)
    assert np.array_equal(v[fin], prod[fin])


def test_project_int64_arith_and_literals(native):
    """int64 arithmetic wraps, / and % by zero give NA, literals broadcast"""
    nv = native
    x = np.array([7, -7, 5, (1 << 63) - 1, 3], np.int64)
    y = np.array([2, 2, 0, 1, -1], np.int64)
    node = nv.project_node(nv.Table({"x": x, "y": y}).scan(),
                           [("x", None), ("q", ("arith", "/", ("col", "x"), ("col", "y"))),
                            ("r", ("arith", "%", ("col", "x"), ("col", "y"))),
                            ("w", ("arith", "+", ("col", "x"), ("int", 1))),
                            ("h", ("arith", "*", ("col", "x"), ("dbl", 0.5)))])
    out = nv.collect(node)
    assert list(out["q"][0][out["q"][1]]) == [3, -3, (1 << 63) - 1, -3]
    assert list(out["q"][1]) == [True, True, False, True, True]
    assert list(out["r"][0][out["r"][1]]) == [1, -1, 0, 0]
    assert out["w"][0][3] == -(1 << 63)
    assert list(out["h"][0]) == [3.5, -3.5, 2.5, float((1 << 63) - 1) * 0.5, 1.5]


def test_project_errors_and_explain(native):
    nv = native
    with pytest.raises(nv.VectraError, match="select: column not found: zz"):
        nv.project_node(nv.Table({"x": np.arange(3)}).scan(), [("zz", None)])
    node = nv.project_node(nv.Table({"x": np.arange(3)}).scan(),
                           [("x", None), (".vectra_tmp_s", ("arith", "*", ("col", "x"), ("col", "x")))])
    lines = nv.explain(node)
    assert lines[0] == "ProjectNode [streaming, hidden mutate]"
    assert lines[1].startswith("  MemScanNode")
    nv.core().vx_node_free(node)
