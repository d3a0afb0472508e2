"""Batches whose queries differ only in the constant of `p op t_q` (op one
of > >= < <=) run as ONE bucketed pass (k_bt_chain_feed): every row is
aggregated once under (group, the number of queries it passes).  Each query
must still equal

    tbl |> filter(p op t_q) |> group_by(g) |> summarise(...)

run alone through the reference's nodes (src/filter.c, src/group_agg.c,
src/agg_ops.c; NaN comparisons per src/scalar_ops.c:125-141, where a NaN
operand compares "equal", so >= and <= keep it and > and < drop it):
keys, order, counts and int64 sums exact, double sums and means within
1e-12 * sum|x| of the group."""
import numpy as np
import pytest

from helpers import assert_tables_equal, dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu

# n, count(x), sum(x), mean(x), sum(v) (int64), count(v)
AGGS = [("n", "n", None, 0), ("cx", "count", "x", 0), ("sx", "sum", "x", 0), ("mx", "mean", "x", 0),
        ("sv", "sum", "v", 0), ("cv", "count", "v", 1)]
BATCH_AGGS = [("n", -1, 0), ("count", 0, 0), ("sum", 0, 0), ("mean", 0, 0), ("sum", 1, 0), ("count", 1, 1)]
CLOSE = {"sx", "mx"}
OPS = {">": "GT", ">=": "GE", "<": "LT", "<=": "LE"}
DBL_T = [0.0, -0.0, 0.5, 0.5, float("nan"), float("inf"), -float("inf"), -1.2, 1.7, 2.0]


def make_table(seed, n, key_lo=-40, key_hi=250):
    r = rng(seed)
    g = r.integers(key_lo, key_hi, n)
    p = dbl_with_specials(r, n, 0.02, 0.03)
    u = r.random(n)
    p[u < 0.01] = np.inf
    p[(u >= 0.01) & (u < 0.02)] = -np.inf
    p[(u >= 0.02) & (u < 0.03)] = 0.5          # rows equal to a threshold
    x = dbl_with_specials(r, n, 0.001, 0.0)
    v = r.integers(-2**40, 2**40, n)
    return {"g": (g, with_na(r, n, 0.03)), "p": (p, with_na(r, n, 0.02)), "x": x, "v": v,
            "ax": np.abs(np.nan_to_num(x, nan=0.0))}


def oracle_query(nv, tbl, op, t, temp_dir="/tmp", key="g", join_dim=None):
    lit = ("dbl", t) if isinstance(t, float) else ("int", t)
    cols = {k: tbl[k] for k in ("p", "x", "v", "ax")}
    if join_dim is None:
        cols["g"] = tbl["g"]
    else:
        cols["fk"] = tbl["fk"]
    node = nv.filter_node("oracle", nv.Table(cols, rg_rows=0).scan(), ("cmp", op, ("col", "p"), lit))
    if join_dim is not None:
        node = nv.join_node("oracle", node, nv.Table(join_dim, rg_rows=0).scan(), "inner", [(4, 0)])
    aggs = list(AGGS) + [("__sa", "sum", "ax", 1), ("__cnt", "count", "x", 0)]
    out = nv.collect(nv.group_agg_node("oracle", node, [key], aggs, temp_dir))
    sa = out.pop("__sa")[0]
    cnt = out.pop("__cnt")[0]
    close = {"sx": sa, "mx": sa / np.maximum(cnt, 1)}
    return {("g" if k == key else k): v for k, v in out.items()}, close


def renamed(got):
    out = {"g": got["key"]}
    for i, (name, _, _, _) in enumerate(AGGS):
        out[name] = got[f"a{i}"]
    return out


def check_all(nv, b, tbl, op, ts, **kw):
    n_groups, offs = b.counts()
    assert np.all(offs[1:len(ts)] == np.cumsum(n_groups[:len(ts)])[:-1])
    for q, t in enumerate(ts):
        exp, close = oracle_query(nv, tbl, op, t, **kw)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)


def device_cols(nv, tbl):
    return {"g": nv.DeviceColumn(*tbl["g"]), "p": nv.DeviceColumn(*tbl["p"]), "x": nv.DeviceColumn(tbl["x"]),
            "v": nv.DeviceColumn(tbl["v"])}


def plan(nv, dc, n, op, ts, ptype=None):
    ptype = nv.VEC_DOUBLE if ptype is None else ptype   # VEC_INT64 is 0
    pred = nv.cmp_const_prog(0, ptype, OPS[op], 0.0 if ptype == nv.VEC_DOUBLE else 0)
    b = nv.Batch(nv.VEC_INT64, [nv.VEC_DOUBLE, nv.VEC_INT64], BATCH_AGGS, nv.VGPU_ORDER_KEY, pred, 1, 1)
    b.plan([{"key": dc["g"], "inputs": [dc["x"], dc["v"]], "pred_cols": [dc["p"]], "n_rows": n, "param": t}
            for t in ts])
    return b


@pytest.mark.parametrize("op", list(OPS))
def test_bucketed_matches_oracle(gpu, oracle, op):
    """thresholds with duplicates, +-0, +-inf and NaN; predicate values with
    NaN, NA, +-inf and values equal to a threshold; NA group keys"""
    tbl = make_table(61, 300_000)
    dc = device_cols(gpu, tbl)
    b = plan(gpu, dc, 300_000, op, DBL_T)
    assert b.bucketed()
    check_all(gpu, b, tbl, op, DBL_T)
    b.free()
    for c in dc.values():
        c.free()


def test_bucketed_rerun_capture_and_fewer_queries(gpu, oracle):
    """a rerun with new constants on fewer queries, then the captured graph
    replayed with eager work between the replays"""
    tbl = make_table(67, 200_000)
    dc = device_cols(gpu, tbl)
    b = plan(gpu, dc, 200_000, ">", DBL_T)
    assert b.bucketed()
    new = [float("nan"), 0.25, -0.75, 0.25]
    b.rerun(new + [0.0] * (len(DBL_T) - len(new)), n_active=len(new))
    n_groups, _ = b.counts()
    assert np.all(n_groups[len(new):] == 0)
    check_all(gpu, b, tbl, ">", new)
    b.capture()
    for _ in range(3):
        b.run()
        gpu.vgpu().vgpu_synchronize()
        scratch = gpu.DeviceColumn(np.arange(1000, dtype=np.int64))
        scratch.free()
        check_all(gpu, b, tbl, ">", new)
    b.free()
    for c in dc.values():
        c.free()


def test_bucketed_same_as_per_query_feed(gpu, oracle):
    """the same batch with the bucketed pass switched off (knob
    batch.bucketed = 1) gives the same keys, counts and int64 sums, and
    double sums within the tolerance"""
    tbl = make_table(71, 250_000)
    dc = device_cols(gpu, tbl)
    ts = [-0.5, 0.0, 0.9, 1.4, 0.3, float("nan")]
    lib = gpu.vgpu()
    b1 = plan(gpu, dc, 250_000, ">=", ts)
    old = lib.vgpu_tune_get(b"batch.bucketed")
    lib.vgpu_tune_set(b"batch.bucketed", 1)
    try:
        b0 = plan(gpu, dc, 250_000, ">=", ts)
    finally:
        lib.vgpu_tune_set(b"batch.bucketed", old)
    assert b1.bucketed() and not b0.bucketed()
    for q, t in enumerate(ts):
        _, close = oracle_query(gpu, tbl, ">=", t)
        assert_tables_equal(renamed(b1.fetch(q)), renamed(b0.fetch(q)), close)
    b0.free()
    b1.free()
    for c in dc.values():
        c.free()


def test_bucketed_int64_predicate(gpu, oracle):
    """an int64 column compared with int64 constants"""
    r = rng(731)
    n = 150_000
    tbl = make_table(73, n)
    pv = r.integers(-1000, 1000, n)
    tbl["p"] = (pv, with_na(r, n, 0.02))
    dc = device_cols(gpu, tbl)
    ts = [0, -1000, 999, 1000, 5, 5, -3, 250]
    b = plan(gpu, dc, n, "<", ts, ptype=gpu.VEC_INT64)
    assert b.bucketed()
    check_all(gpu, b, tbl, "<", ts)
    b.rerun([7, -2000, 7], n_active=3)
    check_all(gpu, b, tbl, "<", [7, -2000, 7])
    b.free()
    for c in dc.values():
        c.free()


def test_bucketed_edge_sizes(gpu, oracle):
    """no rows, one row, fewer rows than one block step, and all keys NA"""
    for n in (0, 1, 100, 5000):
        tbl = make_table(79 + n, n)
        dc = device_cols(gpu, tbl)
        b = plan(gpu, dc, n, "<=", [0.1, -0.3, 0.1])
        check_all(gpu, b, tbl, "<=", [0.1, -0.3, 0.1])
        b.free()
        for c in dc.values():
            c.free()
    tbl = make_table(83, 3000)
    tbl["g"] = (tbl["g"][0], np.zeros(3000, bool))
    dc = device_cols(gpu, tbl)
    b = plan(gpu, dc, 3000, ">", [0.0, 1.0])
    assert b.bucketed()
    check_all(gpu, b, tbl, ">", [0.0, 1.0])
    b.free()
    for c in dc.values():
        c.free()


def test_bucketed_falls_back_outside_its_shape(gpu, oracle):
    """a key range wider than any LDS table (INT64_MIN among the keys): the
    batch runs per query, with the same results"""
    tbl = make_table(89, 100_000)
    g = tbl["g"][0].copy()
    g[::997] = np.iinfo(np.int64).min
    tbl["g"] = (g, tbl["g"][1])
    dc = device_cols(gpu, tbl)
    b = plan(gpu, dc, 100_000, ">", [0.0, 0.5])
    assert not b.bucketed()
    check_all(gpu, b, tbl, ">", [0.0, 0.5])
    b.free()
    for c in dc.values():
        c.free()


def join_tables(seed, n_fact, n_dim, dup):
    r = rng(seed)
    dk = r.permutation(np.arange(n_dim, dtype=np.int64) * 3 + 5)
    if dup:
        dk[: n_dim // 40] = dk[n_dim // 40: 2 * (n_dim // 40)]
    attr = r.integers(0, 200, n_dim)
    fk = np.where(r.random(n_fact) < 0.9, dk[r.integers(0, n_dim, n_fact)], -9)
    tbl = make_table(seed + 1, n_fact)
    del tbl["g"]
    tbl["fk"] = (fk, with_na(r, n_fact, 0.01))
    return tbl, {"dk": dk, "attr": (attr, with_na(r, n_dim, 0.02))}


@pytest.mark.parametrize("dense", [True, False])
@pytest.mark.parametrize("dup", [False, True])
def test_bucketed_join(gpu, oracle, dup, dense):
    """filter(p > t_q) |> inner_join(dim, fk = dk) |> group_by(attr): one
    bucketed pass probes the resident table once per selected row, through
    the direct-addressed map of unique keys (dense; keys 3 apart) or the
    hash table (duplicate keys, or the map switched off); probe keys at the
    range's ends, just past it, far past it and below it"""
    tbl, dim = join_tables(97, 300_000, 30_000, dup)
    fk = tbl["fk"][0].copy()
    dk = dim["dk"]
    fk[::997], fk[1::997], fk[2::997], fk[3::997] = dk.max() + 1, dk.min(), dk.max(), 2**62
    tbl["fk"] = (fk, tbl["fk"][1])
    dkc = gpu.DeviceColumn(dk)
    atc = gpu.DeviceColumn(*dim["attr"])
    jt = gpu.JoinTable(dkc, [atc])
    fkc = gpu.DeviceColumn(*tbl["fk"])
    pc, xc, vc = gpu.DeviceColumn(*tbl["p"]), gpu.DeviceColumn(tbl["x"]), gpu.DeviceColumn(tbl["v"])
    ts = [0.0, 1.0, -1.0, 0.5, float("inf")]
    pred = gpu.cmp_const_prog(0, gpu.VEC_DOUBLE, "GT", 0.0)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE, gpu.VEC_INT64], BATCH_AGGS, gpu.VGPU_ORDER_KEY, pred, 1, 1)
    b.join(jt, 0)
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"join.dense_map")
    lib.vgpu_tune_set(b"join.dense_map", 2 if dense else 1)   # read when the first plan builds the map
    try:
        b.plan([{"probe_key": fkc, "inputs": [xc, vc], "pred_cols": [pc], "n_rows": 300_000, "param": t}
                for t in ts])
    finally:
        lib.vgpu_tune_set(b"join.dense_map", old)
    assert b.bucketed()
    check_all(gpu, b, tbl, ">", ts, key="attr", join_dim=dim)
    b.rerun([0.2, -0.4], n_active=2)
    check_all(gpu, b, tbl, ">", [0.2, -0.4], key="attr", join_dim=dim)
    b.free()
    jt.free()
    for c in (dkc, atc, fkc, pc, xc, vc):
        c.free()


@pytest.mark.parametrize("plain,n", [(False, 260_000)] + [(True, n) for n in (260_000, 0, 1, 1_000, 8_193, 1_048_577)])
@pytest.mark.parametrize("join", [False, True])
def test_bucketed_sums_the_compared_column(gpu, oracle, join, plain, n):
    """filter(p > t_q) |> [inner_join(dim) |>] group_by(...) |> summarise(
    n = n(), c = count(p), s = sum(p), m = mean(p)): the summed input is the
    compared column itself (the scan keeps one column per row, eight rows
    per thread and step).  plain: the group / probe key has no NAs either,
    so the scan loads both columns without validity checks (rows past the
    end of a block's range masked out): row counts below one step, one past
    a step and one past a power of two leave partial last steps"""
    if join:
        tbl, dim = join_tables(101 + n, n, 20_000, False)
        key = "attr"
        if plain:
            tbl["fk"] = (tbl["fk"][0], None)
    else:
        tbl, dim = make_table(103 + n, n), None
        key = "g"
        if plain:
            tbl["g"] = (tbl["g"][0], None)
    r = rng(107 + n)
    p = r.normal(0.0, 1.0, n)
    p[::1013] = 0.5                          # rows equal to a threshold
    tbl["p"] = p
    ts = [0.0, 1.0, -1.0, 0.5, 0.5, 2.5, -0.25, 1.75]
    lib_aggs = [("n", -1, 0), ("count", 0, 0), ("sum", 0, 0), ("mean", 0, 0)]
    pc = gpu.DeviceColumn(p)
    pred = gpu.cmp_const_prog(0, gpu.VEC_DOUBLE, "GT", 0.0)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE], lib_aggs, gpu.VGPU_ORDER_KEY, pred, 1, 1)
    owned = [pc]
    if join:
        dkc, atc = gpu.DeviceColumn(dim["dk"]), gpu.DeviceColumn(*dim["attr"])
        jt = gpu.JoinTable(dkc, [atc])
        fkc = gpu.DeviceColumn(*tbl["fk"])
        owned += [dkc, atc, fkc]
        b.join(jt, 0)
        qs = [{"probe_key": fkc, "inputs": [pc], "pred_cols": [pc], "n_rows": n, "param": t} for t in ts]
    else:
        gc = gpu.DeviceColumn(*tbl["g"])
        owned.append(gc)
        qs = [{"key": gc, "inputs": [pc], "pred_cols": [pc], "n_rows": n, "param": t} for t in ts]
    b.plan(qs)
    assert b.bucketed() or n == 0
    cols = {"p": p, "ap": np.abs(p)}
    cols.update({"fk": tbl["fk"]} if join else {"g": tbl["g"]})
    for q, t in enumerate(ts):
        node = gpu.filter_node("oracle", gpu.Table(cols, rg_rows=0).scan(), ("cmp", ">", ("col", "p"), ("dbl", t)))
        if join:
            node = gpu.join_node("oracle", node, gpu.Table(dim, rg_rows=0).scan(), "inner", [(2, 0)])
        out = gpu.collect(gpu.group_agg_node(
            "oracle", node, [key], [("n", "n", None, 0), ("c", "count", "p", 0), ("s", "sum", "p", 0),
                                    ("m", "mean", "p", 0), ("__sa", "sum", "ap", 1)], "/tmp"))
        sa = out.pop("__sa")[0]
        exp = {("g" if k == key else k): v for k, v in out.items()}
        got = b.fetch(q)
        got = {"g": got["key"], "n": got["a0"], "c": got["a1"], "s": got["a2"], "m": got["a3"]}
        assert_tables_equal(got, exp, {"s": sa, "m": sa / np.maximum(exp["c"][0], 1)})
    b.free()
    if join:
        jt.free()
    for c in owned:
        c.free()


I64_TK = [0, 5, 5, -300, 999, 2**63 - 1, -(2**63) + 1, 42, -7, 1, -1, 500, -500, 6, 4, 1000, -999]
I64_TK += [(-1) ** i * 37 * i for i in range(23)]
DBL_TK = [0.0, -0.0, 0.5, 0.5, float("nan"), float("inf"), -float("inf"), -1.2, 1.7,
         -0.5, 0.25, float("nan"), 3.0, -3.0, 0.0, 1e-300, -2.2]
DBL_TK += [(-1) ** i * 0.13 * i for i in range(23)]


@pytest.mark.parametrize("k", [8, 9, 13, 16, 17, 40])
@pytest.mark.parametrize("dtype", ["double", "int64"])
@pytest.mark.parametrize("op", list(OPS))
def test_bucketed_plain_scan_every_operator(gpu, oracle, op, dtype, k):
    """the plain scan (no NAs in the compared column or the key, one sum of
    the compared column) with up to sixteen queries counts each row's passes
    against the constants in registers (8 compare slots up to eight queries,
    16 up to sixteen); seventeen and forty take a branch-free search of the
    constants with the comparison fixed at compile time: every
    operator on double and int64 columns, NaN / +-0 / +-inf and extreme
    int64 constants (INT64_MAX, INT64_MIN + 1), duplicate constants,
    values equal to a constant"""
    n = 200_003
    r = rng(211 + 7 * k + len(op) + (dtype == "int64"))
    g = r.integers(-40, 250, n)
    if dtype == "double":
        p = dbl_with_specials(r, n, 0.02, 0.03)
        u = r.random(n)
        p[u < 0.01] = np.inf
        p[(u >= 0.01) & (u < 0.02)] = -np.inf
        p[(u >= 0.02) & (u < 0.04)] = 0.5
        ts = DBL_TK[:k]
        ptype, lit = gpu.VEC_DOUBLE, "dbl"
    else:
        p = r.integers(-1000, 1000, n)
        u = r.random(n)
        p[u < 0.01] = 10**12                  # large values, no int64 sum overflows
        p[(u >= 0.01) & (u < 0.02)] = -10**12
        p[(u >= 0.02) & (u < 0.05)] = 5
        ts = I64_TK[:k]
        ptype, lit = gpu.VEC_INT64, "int"
    pc, gc = gpu.DeviceColumn(p), gpu.DeviceColumn(g)
    pred = gpu.cmp_const_prog(0, ptype, OPS[op], 0.0 if dtype == "double" else 0)
    lib_aggs = [("n", -1, 0), ("count", 0, 0), ("sum", 0, 0), ("mean", 0, 0)]
    b = gpu.Batch(gpu.VEC_INT64, [ptype], lib_aggs, gpu.VGPU_ORDER_KEY, pred, 1, 1)
    b.plan([{"key": gc, "inputs": [pc], "pred_cols": [pc], "n_rows": n, "param": t} for t in ts])
    assert b.bucketed()
    ap = np.abs(np.nan_to_num(p.astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0))
    cols = {"p": p, "ap": ap, "g": g}
    for q, t in enumerate(ts):
        node = gpu.filter_node("oracle", gpu.Table(cols, rg_rows=0).scan(), ("cmp", op, ("col", "p"), (lit, t)))
        out = gpu.collect(gpu.group_agg_node(
            "oracle", node, ["g"], [("n", "n", None, 0), ("c", "count", "p", 0), ("s", "sum", "p", 0),
                                    ("m", "mean", "p", 0), ("__sa", "sum", "ap", 1)], "/tmp"))
        sa = out.pop("__sa")[0]
        got = b.fetch(q)
        got = {"g": got["key"], "n": got["a0"], "c": got["a1"], "s": got["a2"], "m": got["a3"]}
        # int64 sums are exact, and so are their means (sum / n)
        close = {"s": sa, "m": sa / np.maximum(out["c"][0], 1)} if dtype == "double" else {}
        assert_tables_equal(got, out, close)
    b.free()
    pc.free()
    gc.free()
