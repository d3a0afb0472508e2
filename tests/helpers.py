"""Shared test helpers: seeded inputs with NA / NaN / signed zeros, and the
parity comparisons (bit-exact for integers, keys, counts, selection vectors,
join rows and min/max; |a - b| <= 1e-12 * sum|x| for double sums and means).
"""
from __future__ import annotations

import numpy as np

REL_TOL = 1e-12   # double sum / mean tolerance, scaled by sum |x_i| of the group


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def with_na(r: np.random.Generator, n: int, p: float) -> np.ndarray | None:
    if p <= 0:
        return None
    return r.random(n) >= p


def dbl_with_specials(r: np.random.Generator, n: int, p_nan: float = 0.02, p_zero: float = 0.02) -> np.ndarray:
    x = r.normal(size=n)
    u = r.random(n)
    x[u < p_nan] = np.nan
    z = (u >= p_nan) & (u < p_nan + p_zero)
    x[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
    return x


def bits(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    b = a.view(np.int64).copy()
    b[np.isnan(a)] = 0x7FF8000000000000   # one NaN
    return b


def assert_col_exact(name, got, exp):
    gv, gm = got
    ev, em = exp
    assert len(gv) == len(ev), f"{name}: {len(gv)} rows vs {len(ev)}"
    assert np.array_equal(gm, em), f"{name}: validity differs at {np.flatnonzero(gm != em)[:10]}"
    if gv.dtype.kind == "f" or ev.dtype.kind == "f":
        gb, eb = bits(gv[gm]), bits(ev[em])
        bad = np.flatnonzero(gb != eb)
        assert bad.size == 0, f"{name}: {bad.size} values differ, first {bad[:5]}: {gv[gm][bad[:5]]} vs {ev[em][bad[:5]]}"
    elif gv.dtype == object or ev.dtype == object:
        assert list(gv[gm]) == list(ev[em]), f"{name}: string values differ"
    else:
        bad = np.flatnonzero(gv[gm].astype(np.int64) != ev[em].astype(np.int64))
        assert bad.size == 0, f"{name}: {bad.size} values differ, first {bad[:5]}"


def assert_col_close(name, got, exp, scale):
    """|got - exp| <= REL_TOL * scale elementwise; NaN / Inf must agree."""
    gv, gm = got
    ev, em = exp
    assert len(gv) == len(ev), f"{name}: {len(gv)} rows vs {len(ev)}"
    assert np.array_equal(gm, em), f"{name}: validity differs"
    g, e, s = gv[gm], ev[em], np.asarray(scale)[em]
    fin = np.isfinite(e)
    assert np.array_equal(np.isnan(g), np.isnan(e)), f"{name}: NaN pattern differs"
    assert np.array_equal(g[~fin & ~np.isnan(e)], e[~fin & ~np.isnan(e)]), f"{name}: infinities differ"
    err = np.abs(g[fin] - e[fin])
    lim = REL_TOL * np.maximum(s[fin], np.finfo(float).tiny)
    bad = np.flatnonzero(err > lim)
    assert bad.size == 0, f"{name}: {bad.size} values off, worst rel {np.max(err / lim) * REL_TOL:.3g}"


def assert_tables_equal(got: dict, exp: dict, close: dict | None = None):
    """close: {column: per-row tolerance scale} for double sums / means."""
    close = close or {}
    assert list(got) == list(exp), f"columns {list(got)} vs {list(exp)}"
    for k in got:
        if k in close:
            assert_col_close(k, got[k], exp[k], close[k])
        else:
            assert_col_exact(k, got[k], exp[k])
