"""Parity at the workloads' full sizes, checked through closed forms and
exact arithmetic the full data admits (the oracle pipeline itself would
take minutes per query at these sizes):

* join then aggregate, fact 50,000,000 x dimension 1,000,000, inner and
  left, through the kernel API and through the R node tree: every group's
  key, order and count exact; every sum within 1e-12 * sum |x| of the
  correctly rounded sum of its products (oracle_exact_sum);
* group-by over 200,000,000 rows at 16 and at 50,000,000 key values: every
  group's key, order, count, int64 sum / min / max and double min / max
  exact; double sums within 1e-12 * sum |x| of the exact sum (16 keys) or
  of numpy's per-group sum, whose own error is below (n - 1) u sum |x|
  for a group of n rows, subtracted from the bound (50M keys); means =
  sum / n bit for bit;
* 4,096 batched filter -> group-by queries of uneven sizes (about 4.4e9
  rows): every 64th query against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-12


def exact_sums(oracle, values, group_ids, n_groups):
    """correctly rounded per-group sums (groups 0..n_groups-1)"""
    order = np.argsort(group_ids.astype(np.int16 if n_groups < 32768 else np.int64), kind="stable")
    v = np.ascontiguousarray(values[order])
    counts = np.bincount(group_ids, minlength=n_groups)
    out = np.zeros(n_groups)
    at = 0
    for g in range(n_groups):
        c = int(counts[g])
        out[g] = oracle.oracle_exact_sum(v[at:].ctypes.data_as(C.POINTER(C.c_double)), c) if c else 0.0
        at += c
    return out, counts


def zipf_ranks(r, n, s, n_keys):
    w = np.arange(1, n_keys + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, r.random(n), side="right"), n_keys - 1)


@pytest.fixture(scope="module")
def c2_full():
    """fact 50M rows (fkey Zipf(1.1) over 1M dimension keys, 10% absent,
    value lognormal(0,1)), dimension 1M rows (unique shuffled int64 key,
    attr U[0,100), weight U(0,1)), with each fact row's dimension row"""
    r = np.random.default_rng(0x76656374726102)
    n_fact, n_dim = 50_000_000, 1_000_000
    keys = np.unique(r.integers(1, 2**62, int(n_dim * 2.1)))
    r.shuffle(keys)
    dim_key, absent = keys[:n_dim], keys[n_dim:2 * n_dim]
    attr = r.integers(0, 100, n_dim)
    weight = r.random(n_dim)
    ranks = zipf_ranks(r, n_fact, 1.1, n_dim)
    is_absent = r.random(n_fact) < 0.1
    fkey = np.where(is_absent, absent[ranks], dim_key[ranks])
    value = r.lognormal(0.0, 1.0, n_fact)
    return {"fkey": fkey, "value": value}, {"key": dim_key, "attr": attr, "weight": weight}, ranks, is_absent


def c2_expected(oracle, c2, kind):
    fact, dim, ranks, is_absent = c2
    hit = ~is_absent
    gid = dim["attr"][ranks[hit]]
    prod = fact["value"][hit] * dim["weight"][ranks[hit]]
    sums, counts = exact_sums(oracle, prod, gid, 100)
    keys = np.flatnonzero(counts)
    exp = {"key": keys, "n": counts[keys].astype(np.float64), "s": sums[keys]}
    if kind == "left":
        exp["n_na"] = float(is_absent.sum())
    return exp


def check_c2(got_key, got_key_ok, got_s, got_s_ok, got_n, exp, kind):
    k = len(exp["key"])
    if kind == "left":
        assert len(got_key) == k + 1 and not got_key_ok[-1], "the unmatched rows' NA group comes last"
        assert got_n[-1] == exp["n_na"] and not got_s_ok[-1], "NA group: all unmatched rows, sum NA"
    else:
        assert len(got_key) == k and got_key_ok.all()
    assert np.array_equal(got_key[:k], exp["key"]), "group keys / order differ"
    assert np.array_equal(got_n[:k], exp["n"]), "counts differ"
    assert got_s_ok[:k].all()
    err = np.abs(got_s[:k] - exp["s"])
    # every product is positive: sum |x| of a group is its exact sum
    assert np.all(err <= REL * exp["s"]), f"worst relative error {np.max(err / exp['s']):.3g}"


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_full_kernel_api(gpu, oracle, c2_full, kind):
    nv = gpu
    fact, dim, _, _ = c2_full
    cols = [nv.DeviceColumn(a) for a in (fact["fkey"], fact["value"], dim["key"], dim["attr"], dim["weight"])]
    fk, val, dk, at, wt = cols
    try:
        jt = nv.JoinTable(dk, [at, wt])
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, 101)
        gb.feed_join(len(fact["fkey"]), [fk, val], [dk, at, wt], [(0, 1, 1, 2)], key=(1, 1), join=jt,
                     join_kind=nv.JOIN_NAMES[kind], probe_key_col=0)
        gb.finish()
        got = gb.fetch()
        gb.free()
        jt.free()
    finally:
        for c in cols:
            c.free()
    exp = c2_expected(oracle, c2_full, kind)
    check_c2(got["key"][0], got["key"][1], got["a0"][0], got["a0"][1], got["a1"][0], exp, kind)


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_full_node_tree(gpu, oracle, c2_full, kind):
    """the tree R builds: GroupAgg <- Project(.vectra_tmp_s = value * weight)
    <- Join, fact in row groups of 1,048,576, fused on the device"""
    nv = gpu
    fact, dim, _, _ = c2_full
    j = nv.join_node("gpu", nv.Table(fact, rg_rows=1 << 20).scan(), nv.Table(dim).scan(), kind, [(0, 0)])
    p = nv.hidden_mutate(j, ".vectra_tmp_s", "value", "weight")
    node = nv.group_agg_node("gpu", p, ["attr"], [("s", "sum", ".vectra_tmp_s", 0), ("n", "n", None, 0)], "/tmp")
    got = nv.collect(node, free=False)
    note = "\n".join(nv.explain(node))
    nv.core().vx_node_free(node)
    assert "fused with the Project and Join below" in note
    exp = c2_expected(oracle, c2_full, kind)
    check_c2(got["attr"][0], got["attr"][1], got["s"][0], got["s"][1], got["n"][0], exp, kind)


C3_AGGS = [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("mean", 0, 0),
           ("sum", 1, 0), ("min", 1, 0), ("max", 1, 0), ("mean", 1, 0)]
C3_ROWS = 200_000_000


@pytest.fixture(scope="module")
def c3_values():
    r = np.random.default_rng(0x7665637472610300)
    v1 = r.lognormal(0.0, 1.0, C3_ROWS)
    v2 = r.integers(-10**9, 10**9, C3_ROWS, endpoint=True)
    return v1, v2


def c3_run(nv, key, v1, v2):
    cols = [nv.DeviceColumn(a) for a in (key, v1, v2)]
    try:
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE, nv.VEC_INT64], C3_AGGS, nv.VGPU_ORDER_KEY, 0)
        gb.feed(cols[0], cols[1:], C3_ROWS)
        gb.finish()
        out = gb.fetch()
        gb.free()
    finally:
        for c in cols:
            c.free()
    return out


def test_c3_full_low_cardinality(gpu, oracle, c3_values):
    """16 keys, 12.5M rows per group: exact int64 sums, min / max, and the
    exact double sum of every group"""
    nv = gpu
    v1, v2 = c3_values
    key = np.random.default_rng(0x7665637472610310).integers(0, 16, C3_ROWS)
    out = c3_run(nv, key, v1, v2)
    assert list(out["key"][0]) == list(range(16)) and out["key"][1].all()
    for k in range(16):
        m = key == k
        x1, x2 = np.ascontiguousarray(v1[m]), v2[m]
        n = len(x1)
        assert out["a0"][0][k] == n
        s1 = oracle.oracle_exact_sum(x1.ctypes.data_as(C.POINTER(C.c_double)), n)
        assert abs(out["a1"][0][k] - s1) <= REL * s1, f"key {k}: v1 sum off by {abs(out['a1'][0][k] - s1) / s1:.3g}"
        assert out["a2"][0][k] == x1.min() and out["a3"][0][k] == x1.max()
        assert out["a4"][0][k] == out["a1"][0][k] / n
        s2 = int(x2.sum())
        assert out["a5"][0][k] == float(s2)
        assert out["a6"][0][k] == float(x2.min()) and out["a7"][0][k] == float(x2.max())
        assert out["a8"][0][k] == float(s2) / n


def test_c3_full_high_cardinality(gpu, oracle, c3_values):
    """50,000,000 key values (about 49.1M present), key order"""
    nv = gpu
    v1, v2 = c3_values
    card = 50_000_000
    r = np.random.default_rng(0x7665637472610320)
    idx = r.integers(0, card, C3_ROWS)
    ks = np.unique(r.integers(-2**62, 2**62, int(card * 1.01) + 1024))[:card]
    r.shuffle(ks)
    key = ks[idx]
    out = c3_run(nv, key, v1, v2)
    cnt = np.bincount(idx, minlength=card)
    order = np.argsort(ks)
    present = order[cnt[order] > 0]                # key indices in ascending key order
    assert np.array_equal(out["key"][0], ks[present]), "keys / order differ"
    n = cnt[present].astype(np.float64)
    assert np.array_equal(out["a0"][0], n), "counts differ"
    del key
    # int64 sums: at most a few dozen values of magnitude <= 1e9 per group,
    # so the double partial sums are exact integers
    s2 = np.bincount(idx, weights=v2.astype(np.float64), minlength=card)[present]
    assert np.array_equal(out["a5"][0], s2)
    assert np.array_equal(out["a8"][0], s2 / n)
    for col, v in ((2, v1), (6, v2)):
        lo = np.full(card, np.inf) if v.dtype.kind == "f" else np.full(card, np.iinfo(np.int64).max)
        np.minimum.at(lo, idx, v)
        assert np.array_equal(out[f"a{col}"][0], lo[present].astype(np.float64)), f"min of a{col} differs"
        hi = np.full(card, -np.inf) if v.dtype.kind == "f" else np.full(card, np.iinfo(np.int64).min)
        np.maximum.at(hi, idx, v)
        assert np.array_equal(out[f"a{col + 1}"][0], hi[present].astype(np.float64)), f"max of a{col + 1} differs"
    # numpy adds each group's values in row order: its error is at most
    # (n - 1) u sum |x| (all v1 > 0), which the bound subtracts
    s1 = np.bincount(idx, weights=v1, minlength=card)[present]
    big = int(cnt.max())
    assert big < 100
    err = np.abs(out["a1"][0] - s1)
    assert np.all(err <= (REL - big * 2.0 ** -53) * s1), f"worst relative error {np.max(err / s1):.3g}"
    assert np.array_equal(out["a4"][0], out["a1"][0] / n)


def test_c4_full_batch(gpu, oracle):
    """4,096 queries of 1e3..1e7 rows and 1..1e5 keys, filter(value > t_q),
    one batch (persistent feed, planned once), every 64th query compared
    with the oracle: keys, order and counts exact, sums within 1e-12 sum |x|"""
    torch = pytest.importorskip("torch")
    from scipy.special import ndtri
    nv = gpu
    Q = 4096
    r = np.random.default_rng(0x7665637472610400)
    rows = np.floor(10 ** r.uniform(3, 7, Q)).astype(np.int64)
    card = np.floor(10 ** r.uniform(0, 5, Q)).astype(np.int64)
    thr = ndtri(1.0 - r.uniform(0.01, 0.5, Q))
    gen = torch.Generator(device="cuda")
    keys, vals = [], []
    for q in range(Q):
        gen.manual_seed(0x7665637472610400 + q)
        keys.append(torch.randint(0, int(card[q]), (int(rows[q]),), dtype=torch.int64, device="cuda", generator=gen))
        vals.append(torch.randn(int(rows[q]), dtype=torch.float64, device="cuda", generator=gen))
    torch.cuda.synchronize()
    kc = [nv.DeviceColumn.wrap(k.data_ptr(), k.numel(), nv.VEC_INT64, keep=k) for k in keys]
    vc = [nv.DeviceColumn.wrap(v.data_ptr(), v.numel(), nv.VEC_DOUBLE, keep=v) for v in vals]
    pred = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.0)
    b = nv.Batch(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0), ("mean", 0, 0)], nv.VGPU_ORDER_KEY,
                 pred, 1, 1)
    b.plan([{"key": kc[q], "inputs": [vc[q]], "pred_cols": [vc[q]], "n_rows": int(rows[q]), "param": float(thr[q])}
            for q in range(Q)])
    b.run()
    for q in range(0, Q, 64):
        kh, vh = keys[q].cpu().numpy(), vals[q].cpu().numpy()
        t = nv.Table({"key": kh, "value": vh, "ax": np.abs(vh)}, rg_rows=1 << 20)
        node = nv.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "value"), ("dbl", float(thr[q]))))
        exp = nv.collect(nv.group_agg_node("oracle", node, ["key"], [("total", "sum", "value", 0), ("n", "n", None, 0),
                                                                      ("sa", "sum", "ax", 0)], "/tmp"))
        got = b.fetch(q)
        assert np.array_equal(got["key"][0], exp["key"][0]), f"query {q}: groups differ"
        assert np.array_equal(got["a1"][0], exp["n"][0]), f"query {q}: counts differ"
        err = np.abs(got["a0"][0] - exp["total"][0])
        assert np.all(err <= REL * exp["sa"][0]), f"query {q}: sums off"
        assert np.array_equal(got["a2"][0], got["a0"][0] / got["a1"][0]), f"query {q}: mean != sum / n"
    b.free()
    del kc, vc, keys, vals
    torch.cuda.empty_cache()
