"""Batched filter -> group-by queries (vgpu_batch_*) vs the oracle, query by
query: each query must equal tbl |> filter(x > t_q) |> group_by(g) |>
summarise(...) run alone through the reference's nodes (src/filter.c,
src/group_agg.c, src/agg_ops.c).  Keys, order, counts and min/max exact;
double sums and means within 1e-12 * sum|x| of the group."""
import numpy as np
import pytest

from helpers import assert_tables_equal, dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu

AGGS = [("n", "n", None, 0), ("cx", "count", "x", 0), ("sx", "sum", "x", 1), ("mx", "mean", "x", 1),
        ("lo", "min", "x", 0), ("hi", "max", "x", 1), ("s0", "sum", "x", 0)]
CLOSE = {"sx", "mx", "s0"}
SIZES = [0, 1, 50, 1000, 32767, 32768, 32769, 100_000, 250_000]
CARDS = [1, 3, 40, 500, 3000, 20000]


def make_queries(seed, n_queries, key_kind="int"):
    r = rng(seed)
    out = []
    for q in range(n_queries):
        n = SIZES[q] if q < len(SIZES) else int(10 ** r.uniform(0, 5.3))
        card = CARDS[q % len(CARDS)]
        if key_kind == "int":
            kv = r.integers(0, card, n) * 7919 - 5000
            kv[r.random(n) < 0.01] = np.iinfo(np.int64).min
            key = (kv, with_na(r, n, 0.02))
        else:
            kv = np.round(r.normal(size=n) * card) / 4.0
            u = r.random(n)
            kv[u < 0.01] = np.nan
            z = (u > 0.5) & (u < 0.52)
            kv[z] = np.where(r.random(z.sum()) < 0.5, 0.0, -0.0)
            key = (kv, with_na(r, n, 0.02))
        x = dbl_with_specials(r, n, 0.01, 0.02)
        out.append({"g": key, "x": (x, with_na(r, n, 0.02)), "t": float(r.uniform(-2.0, 1.0)),
                    "ax": np.abs(np.nan_to_num(x, nan=0.0))})
    return out


def oracle_query(nv, qd, temp_dir, keys=True, pred=None):
    cols = {"x": qd["x"], "ax": qd["ax"]}
    if keys:
        cols["g"] = qd["g"]
    t = nv.Table(cols, rg_rows=0)
    node = nv.filter_node("oracle", t.scan(), pred or ("cmp", ">", ("col", "x"), ("dbl", qd["t"])))
    aggs = list(AGGS) + [("__sa", "sum", "ax", 1), ("__cnt", "count", "x", 0)]
    out = nv.collect(nv.group_agg_node("oracle", node, ["g"] if keys else [], aggs, temp_dir))
    sa = out.pop("__sa")[0]
    cnt = out.pop("__cnt")[0]
    close = {name: (sa if kind == "sum" else sa / np.maximum(cnt, 1)) for name, kind, _, _ in AGGS if name in CLOSE}
    return out, close


def batch_aggs():
    names = {"n": "n", "count": "count", "sum": "sum", "mean": "mean", "min": "min", "max": "max"}
    return [(names[kind], -1 if col is None else 0, rm) for _, kind, col, rm in AGGS]


def renamed(got, keys=True):
    out = {}
    if keys:
        out["g"] = got["key"]
    for i, (name, _, _, _) in enumerate(AGGS):
        out[name] = got[f"a{i}"]
    return out


def device_queries(nv, qs, key_type, keys=True):
    cols, queries = [], []
    for qd in qs:
        n = len(qd["x"][0])
        x = nv.DeviceColumn(*qd["x"])
        k = nv.DeviceColumn(*qd["g"]) if keys else None
        cols += [x] + ([k] if k else [])
        queries.append({"key": k, "inputs": [x], "pred_cols": [x], "n_rows": n, "param": qd["t"]})
    return cols, queries


def run_batch(nv, qs, order, key_type=0, keys=True):
    pred = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.0)   # x > <param>, the constant is instruction 1
    b = nv.Batch(key_type if keys else nv.VGPU_KEY_NONE, [nv.VEC_DOUBLE], batch_aggs(), order, pred, 1, 1)
    cols, queries = device_queries(nv, qs, key_type, keys)
    b.plan(queries)
    return b, cols


@pytest.mark.parametrize("temp_dir", [None, "/tmp"])
def test_batch_matches_oracle(gpu, oracle, temp_dir):
    qs = make_queries(5, 40)
    order = gpu.VGPU_ORDER_KEY if temp_dir else gpu.VGPU_ORDER_FIRST_SEEN
    b, cols = run_batch(gpu, qs, order)
    n_groups, offs = b.counts()
    assert np.all(offs[1:] == np.cumsum(n_groups)[:-1])
    for q, qd in enumerate(qs):
        exp, close = oracle_query(gpu, qd, temp_dir)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_double_keys(gpu, oracle):
    """NaN keys (one group per row), signed zeros, first-seen order"""
    qs = make_queries(7, 16, key_kind="dbl")
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_FIRST_SEEN, key_type=gpu.VEC_DOUBLE)
    for q, qd in enumerate(qs):
        exp, close = oracle_query(gpu, qd, None)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_ungrouped(gpu, oracle):
    qs = make_queries(9, 12)
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_KEY, keys=False)
    for q, qd in enumerate(qs):
        exp, close = oracle_query(gpu, qd, "/tmp", keys=False)
        assert_tables_equal(renamed(b.fetch(q), keys=False), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_vm_predicate(gpu, oracle):
    """a predicate the VM evaluates: (x > t_q) & (g != -5000), t_q per query"""
    qs = make_queries(11, 20)
    ins = [("COL", gpu.VEC_DOUBLE, 0, 0, 0, 0), ("CONST", gpu.VEC_DOUBLE, 0, 0, 0, 0.0),
           ("GT", gpu.VEC_DOUBLE, 0, 1, 0, 0), ("COL", gpu.VEC_INT64, 0, 0, 1, 0),
           ("CONST", gpu.VEC_INT64, 0, 0, 0, -5000), ("NE", gpu.VEC_INT64, 3, 4, 0, 0),
           ("AND", gpu.VEC_BOOL, 2, 5, 0, 0)]
    pred = gpu.compile_prog(ins)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE], batch_aggs(), gpu.VGPU_ORDER_KEY, pred, 2, 1)
    cols, queries = device_queries(gpu, qs, gpu.VEC_INT64)
    for q, qd in zip(queries, qs):
        q["pred_cols"] = [q["inputs"][0], q["key"]]
    b.plan(queries)
    for q, qd in enumerate(qs):
        p = ("and", ("cmp", ">", ("col", "x"), ("dbl", qd["t"])), ("cmp", "!=", ("col", "g"), ("int", -5000)))
        exp, close = oracle_query(gpu, qd, "/tmp", pred=p)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_graph_replays(gpu, oracle):
    """the captured graph replayed three times, the program's own eager work
    in between: every replay matches the oracle"""
    qs = make_queries(13, 24)
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_KEY)
    b.capture()
    expected = [oracle_query(gpu, qd, "/tmp") for qd in qs]
    for rep in range(3):
        b.run()
        gpu.vgpu().vgpu_synchronize()
        scratch = gpu.DeviceColumn(np.arange(1000, dtype=np.int64))   # eager work between replays
        scratch.free()
        for q in (0, 5, 8, 17, 23):
            exp, close = expected[q]
            assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_same_alone_and_in_the_mix(gpu, oracle):
    """a query's result does not depend on the queries batched with it"""
    qs = make_queries(17, 30)
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_KEY)
    for q in (3, 7, 29):
        alone, cols1 = run_batch(gpu, [qs[q]], gpu.VGPU_ORDER_KEY)
        exp, close = oracle_query(gpu, qs[q], "/tmp")
        assert_tables_equal(renamed(b.fetch(q)), renamed(alone.fetch(0)), close)
        assert_tables_equal(renamed(alone.fetch(0)), exp, close)
        alone.free()
        for c in cols1:
            c.free()
    b.free()
    for c in cols:
        c.free()


def test_batch_flags_changed_inputs(gpu, oracle):
    """keys rewritten after plan beyond the planned sort range: the run is
    reported as invalid instead of returning misordered groups"""
    qs = make_queries(19, 6)
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_KEY)
    key_col = cols[2 * 4 + 1]     # query 4's key column (x, g per query)
    big = np.full(key_col.n, 2**61, np.int64)
    gpu._vcheck(gpu.vgpu().vgpu_memcpy_h2d(key_col.data, big.ctypes.data, big.nbytes))
    b.run()
    with pytest.raises(gpu.VectraError, match="inputs changed since plan"):
        b.counts()
    b.free()
    for c in cols:
        c.free()


def shared_table(seed, n, card):
    r = rng(seed)
    kv = r.integers(0, card, n) * 3 - 100
    kv[r.random(n) < 0.01] = np.iinfo(np.int64).min
    x = dbl_with_specials(r, n, 0.01, 0.02)
    return {"g": (kv, with_na(r, n, 0.02)), "x": (x, with_na(r, n, 0.02)), "ax": np.abs(np.nan_to_num(x, nan=0.0))}


@pytest.mark.parametrize("shared", [1, 2])
def test_batch_shared_scan_and_rerun(gpu, oracle, shared):
    """queries over the same columns (one scan of them shared when the
    batch.shared_scan knob is 2); a rerun with new constants on fewer
    queries matches the oracle too"""
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"batch.shared_scan")
    lib.vgpu_tune_set(b"batch.shared_scan", shared)
    try:
        _shared_scan_and_rerun(gpu)
    finally:
        lib.vgpu_tune_set(b"batch.shared_scan", old)


def _shared_scan_and_rerun(gpu):
    base = shared_table(23, 700_000, 900)
    thr = [-1.5, -0.5, 0.0, 0.3, 0.9, 1.7, -3.0, 2.5, 0.1, -0.2, 0.6, 1.1]
    qs = [dict(base, t=t) for t in thr]
    pred = gpu.cmp_const_prog(0, gpu.VEC_DOUBLE, "GT", 0.0)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE], batch_aggs(), gpu.VGPU_ORDER_KEY, pred, 1, 1)
    x = gpu.DeviceColumn(*base["x"])
    k = gpu.DeviceColumn(*base["g"])
    b.plan([{"key": k, "inputs": [x], "pred_cols": [x], "n_rows": len(base["ax"]), "param": t} for t in thr])
    for q, qd in enumerate(qs):
        exp, close = oracle_query(gpu, qd, "/tmp")
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    new = [0.5, -1.0, 1.9, -0.1, 0.05]
    b.rerun(new + [0.0] * (len(thr) - len(new)), n_active=len(new))
    n_groups, _ = b.counts()
    assert np.all(n_groups[len(new):] == 0)
    for q, t in enumerate(new):
        exp, close = oracle_query(gpu, dict(base, t=t), "/tmp")
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    x.free()
    k.free()


def join_tables(seed, n_fact, n_dim, dup=True):
    r = rng(seed)
    dk = r.permutation(np.arange(n_dim, dtype=np.int64) * 5 + 11)
    if dup:   # a few keys repeat on the build side
        dk[: n_dim // 50] = dk[n_dim // 50: 2 * (n_dim // 50)]
    attr = r.integers(0, 40, n_dim)
    av = with_na(r, n_dim, 0.02)
    fk = np.where(r.random(n_fact) < 0.9, dk[r.integers(0, n_dim, n_fact)], -7)
    x = dbl_with_specials(r, n_fact, 0.01, 0.02)
    fact = {"fk": (fk, with_na(r, n_fact, 0.01)), "x": (x, with_na(r, n_fact, 0.02)),
            "ax": np.abs(np.nan_to_num(x, nan=0.0))}
    dim = {"dk": dk, "attr": (attr, av)}
    return fact, dim


def oracle_join_query(nv, fact, dim, t, temp_dir):
    f = nv.filter_node("oracle", nv.Table(fact, rg_rows=0).scan(), ("cmp", ">", ("col", "x"), ("dbl", t)))
    j = nv.join_node("oracle", f, nv.Table(dim, rg_rows=0).scan(), "inner", [(0, 0)])
    aggs = list(AGGS) + [("__sa", "sum", "ax", 1), ("__cnt", "count", "x", 0)]
    out = nv.collect(nv.group_agg_node("oracle", j, ["attr"], aggs, temp_dir))
    sa = out.pop("__sa")[0]
    cnt = out.pop("__cnt")[0]
    close = {name: (sa if kind == "sum" else sa / np.maximum(cnt, 1)) for name, kind, _, _ in AGGS if name in CLOSE}
    out = {("g" if k == "attr" else k): v for k, v in out.items()}
    return out, close


@pytest.mark.parametrize("temp_dir,dup,shared", [("/tmp", True, 1), (None, True, 1), ("/tmp", False, 1),
                                                 ("/tmp", True, 2)])
def test_batch_join(gpu, oracle, temp_dir, dup, shared):
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"batch.shared_scan")
    lib.vgpu_tune_set(b"batch.shared_scan", shared)
    try:
        _batch_join(gpu, temp_dir, dup)
    finally:
        lib.vgpu_tune_set(b"batch.shared_scan", old)


def _batch_join(gpu, temp_dir, dup):
    """filter(x > t_q) |> inner_join(dim, fk = dk) |> group_by(attr): the
    table stays resident across the queries; build keys that repeat"""
    fact, dim = join_tables(29, 400_000, 20_000, dup)
    thr = [-1.0, 0.0, 0.7, 1.5, -2.5, 0.2]
    dkc = gpu.DeviceColumn(dim["dk"])
    atc = gpu.DeviceColumn(*dim["attr"])
    jt = gpu.JoinTable(dkc, [atc])
    fkc = gpu.DeviceColumn(*fact["fk"])
    xc = gpu.DeviceColumn(*fact["x"])
    order = gpu.VGPU_ORDER_KEY if temp_dir else gpu.VGPU_ORDER_FIRST_SEEN
    pred = gpu.cmp_const_prog(0, gpu.VEC_DOUBLE, "GT", 0.0)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE], batch_aggs(), order, pred, 1, 1)
    b.join(jt, 0)
    b.plan([{"probe_key": fkc, "inputs": [xc], "pred_cols": [xc], "n_rows": len(fact["ax"]), "param": t}
            for t in thr])
    for q, t in enumerate(thr):
        exp, close = oracle_join_query(gpu, fact, dim, t, temp_dir)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.rerun([0.4, -0.6], n_active=2)
    for q, t in enumerate([0.4, -0.6]):
        exp, close = oracle_join_query(gpu, fact, dim, t, temp_dir)
        assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    jt.free()
    for c in (dkc, atc, fkc, xc):
        c.free()


def make_dense_queries(seed, n_queries):
    """int64 keys from a short range lo .. lo + card - 1 (lo negative or
    positive) with INT64_MIN and NA rows: the batch gives such queries dense
    tables (record k - lo), in LDS and in HBM"""
    r = rng(seed)
    cards = [1, 2, 40, 300, 2500, 20000, 70000]
    out = []
    for q in range(n_queries):
        n = int(10 ** r.uniform(3, 5.5)) if q >= 3 else [0, 1, 700][q]
        card = cards[q % len(cards)]
        lo = int(r.integers(-10**6, 10**6))
        kv = lo + r.integers(0, card, n)
        kv[r.random(n) < 0.01] = np.iinfo(np.int64).min
        x = dbl_with_specials(r, n, 0.01, 0.02)
        out.append({"g": (kv, with_na(r, n, 0.02)), "x": (x, with_na(r, n, 0.02)),
                    "t": float(r.uniform(-2.0, 1.0)), "ax": np.abs(np.nan_to_num(x, nan=0.0))})
    return out


@pytest.mark.parametrize("temp_dir,dense", [("/tmp", 2), (None, 2), ("/tmp", 1)])
def test_batch_dense_key_ranges(gpu, oracle, temp_dir, dense):
    """queries whose keys span a short range, with the dense tables on (2)
    and off (1): the same groups, order and values as the oracle"""
    qs = make_dense_queries(23, 24)
    order = gpu.VGPU_ORDER_KEY if temp_dir else gpu.VGPU_ORDER_FIRST_SEEN
    lib = gpu.vgpu()
    old = lib.vgpu_tune_get(b"batch.dense")
    lib.vgpu_tune_set(b"batch.dense", dense)
    try:
        b, cols = run_batch(gpu, qs, order)
    finally:
        lib.vgpu_tune_set(b"batch.dense", old)
    for rep in range(2):
        if rep:
            b.run()
        for q, qd in enumerate(qs):
            exp, close = oracle_query(gpu, qd, temp_dir)
            assert_tables_equal(renamed(b.fetch(q)), exp, close)
    b.free()
    for c in cols:
        c.free()


def test_batch_dense_flags_keys_outside_range(gpu, oracle):
    """a dense query's keys rewritten after plan to values outside the
    planned range: the run is reported as invalid"""
    qs = make_dense_queries(29, 6)
    b, cols = run_batch(gpu, qs, gpu.VGPU_ORDER_KEY)
    key_col = cols[2 * 5 + 1]     # query 5 (20,000-key range): x, g per query
    lo = int(qs[5]["g"][0][qs[5]["g"][0] != np.iinfo(np.int64).min].min())
    vals = np.full(key_col.n, lo - 3, np.int64)
    gpu._vcheck(gpu.vgpu().vgpu_memcpy_h2d(key_col.data, vals.ctypes.data, vals.nbytes))
    b.run()
    with pytest.raises(gpu.VectraError, match="inputs changed since plan"):
        b.counts()
    b.free()
    for c in cols:
        c.free()


@pytest.mark.parametrize("temp_dir", ["/tmp", None])
def test_batch_dense_tables_of_every_size(gpu, oracle, temp_dir):
    """sum / n / mean of a double without NAs (three-word records) over key
    ranges of 100 to 30,000 keys: LDS tables of up to 256 records, mid-size
    and large dense LDS tables, a dense LDS table covering the first keys of
    a range too wide for all of it (the rest in HBM), and HBM alone"""
    r = rng(31)
    aggs = [("s", "sum", "x", 0), ("n", "n", None, 0), ("m", "mean", "x", 0)]
    qs = []
    for q, card in enumerate([100, 700, 3000, 6000, 9000, 12000, 30000, 9000]):
        n = 150_000 if q < 7 else 40_000
        lo = int(r.integers(-5000, 5000))
        kv = lo + r.integers(0, card, n)
        kv[r.random(n) < 0.005] = np.iinfo(np.int64).min
        qs.append({"g": (kv, with_na(r, n, 0.01)), "x": r.normal(size=n), "t": float(r.uniform(-1.5, 0.5))})
    order = gpu.VGPU_ORDER_KEY if temp_dir else gpu.VGPU_ORDER_FIRST_SEEN
    pred = gpu.cmp_const_prog(0, gpu.VEC_DOUBLE, "GT", 0.0)
    b = gpu.Batch(gpu.VEC_INT64, [gpu.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0), ("mean", 0, 0)], order, pred, 1, 1)
    cols, queries = [], []
    for qd in qs:
        x, k = gpu.DeviceColumn(qd["x"]), gpu.DeviceColumn(*qd["g"])
        cols += [x, k]
        queries.append({"key": k, "inputs": [x], "pred_cols": [x], "n_rows": len(qd["x"]), "param": qd["t"]})
    b.plan(queries)
    for rep in range(2):
        if rep:
            b.run()
        for q, qd in enumerate(qs):
            t = gpu.Table({"x": qd["x"], "ax": np.abs(qd["x"]), "g": qd["g"]}, rg_rows=0)
            node = gpu.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "x"), ("dbl", qd["t"])))
            exp = gpu.collect(gpu.group_agg_node("oracle", node, ["g"], aggs + [("__sa", "sum", "ax", 0)], temp_dir))
            sa = exp.pop("__sa")[0]
            got = b.fetch(q)
            got = {"g": got["key"], "s": got["a0"], "n": got["a1"], "m": got["a2"]}
            assert_tables_equal(got, exp, {"s": sa, "m": sa / np.maximum(exp["n"][0], 1)})
    b.free()
    for c in cols:
        c.free()
