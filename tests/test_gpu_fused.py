"""Device-resident kernel API: fused probe -> aggregate (the Join ->
Project(value * weight) -> GroupAgg shape of the join-then-aggregate query)
and the filtered group-by, against the oracle nodes on the same rows; plus
the device radix sort against numpy's stable argsort."""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_tables_equal, rng

pytestmark = pytest.mark.gpu


def star_schema(seed, n_fact, n_dim, p_absent=0.1, dup=False):
    r = rng(seed)
    dim_key = (r.permutation(n_dim * 3)[:n_dim].astype(np.int64) * 13 - 7)
    if dup:
        # duplicate build keys on low-frequency Zipf ranks: a few matches each
        for k in range(200, 260):
            dim_key[k * 7 + 1: k * 7 + 4] = dim_key[k * 7]
    attr = r.integers(0, 100, n_dim)
    weight = r.random(n_dim)
    ranks = np.minimum(r.zipf(1.1, n_fact) - 1, n_dim - 1)
    fkey = np.where(r.random(n_fact) < p_absent, -(r.integers(1, 10**9, n_fact)), dim_key[ranks])
    value = r.lognormal(size=n_fact)
    return fkey, value, dim_key, attr, weight


def oracle_join_agg(nv, fkey, value, dim_key, attr, weight, kind, rg=1 << 20):
    left = nv.Table({"fkey": fkey, "value": value}, rg_rows=rg)
    right = nv.Table({"key": dim_key, "attr": attr, "weight": weight})
    j = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)]))
    v = j["value"][0]
    wv, wm = j["weight"]
    prod = np.where(wm, v * wv, 0.0)
    t = nv.Table({"attr": j["attr"], "s_in": (prod, wm), "ab": np.abs(prod)})
    out = nv.collect(nv.group_agg_node("oracle", t.scan(), ["attr"],
                                       [("s", "sum", "s_in", 0), ("n", "n", None, 0), ("sa", "sum", "ab", 0)],
                                       "/tmp"))
    return out


@pytest.mark.parametrize("kind", ["inner", "left"])
@pytest.mark.parametrize("dup", [False, True])
@pytest.mark.parametrize("payload", [0, 1, 2])
def test_fused_join_aggregate(gpu, oracle, kind, dup, payload):
    """payload = build columns embedded in the join table: none, weight only
    (attr read through rows[]), or attr and weight"""
    nv = gpu
    fkey, value, dim_key, attr, weight = star_schema(5, 200_000, 20_000, dup=dup)
    exp = oracle_join_agg(nv, fkey, value, dim_key, attr, weight, kind)
    sa = exp.pop("sa")[0]
    cols = [nv.DeviceColumn(a) for a in (fkey, value, dim_key, attr, weight)]
    fk, val, dk, at, wt = cols
    jt = nv.JoinTable(dk, [[], [wt], [at, wt]][payload])
    gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, 101)
    gb.feed_join(len(fkey), [fk, val], [dk, at, wt], [(0, 1, 1, 2)], key=(1, 1), join=jt,
                 join_kind=nv.JOIN_NAMES[kind], probe_key_col=0)
    gb.finish()
    got = gb.fetch()
    got = {"attr": got["key"], "s": got["a0"], "n": got["a1"]}
    assert_tables_equal(got, exp, {"s": sa})
    gb.free()
    jt.free()
    for c in cols:
        c.free()


@pytest.mark.parametrize("kind,temp_dir,n_fact", [("inner", "/tmp", 150_000), ("left", "/tmp", 150_000),
                                                   ("inner", None, 150_000), ("left", None, 150_000),
                                                   ("inner", "/tmp", 1_200_000), ("left", None, 1_200_000)])
def test_fused_dense_group_table(gpu, oracle, kind, temp_dir, n_fact):
    """the dense-LDS star kernel (unique build keys, int64 payload group key
    of a short range) against the oracle and against the general kernel
    (knob gb.star.dense = 1), with and without the join table's key filter
    (knob join.bloom), its blocks merged through partials and by HBM
    atomics (knob gb.merge.partials): NA group keys and NA weights in the dimension,
    NA and INT64_MIN probe keys, first-seen and key order, a feed of more
    than 2^20 rows"""
    nv = gpu
    r = rng(11)
    fkey, value, dim_key, attr, weight = star_schema(7, n_fact, 12_000)
    dim_key[5] = np.iinfo(np.int64).min
    fkey[r.random(len(fkey)) < 0.01] = np.iinfo(np.int64).min
    fvalid = r.random(len(fkey)) >= 0.02
    avalid = r.random(len(attr)) >= 0.05
    wvalid = r.random(len(weight)) >= 0.03
    left = nv.Table({"fkey": (fkey, fvalid), "value": value}, rg_rows=1 << 20)
    right = nv.Table({"key": dim_key, "attr": (attr, avalid), "weight": (weight, wvalid)})
    j = nv.collect(nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)]))
    wv, wm = j["weight"]
    prod = np.where(wm, j["value"][0] * wv, 0.0)
    t = nv.Table({"attr": j["attr"], "s_in": (prod, wm), "ab": np.abs(prod)})
    exp = nv.collect(nv.group_agg_node("oracle", t.scan(), ["attr"],
                                       [("s", "sum", "s_in", 0), ("n", "n", None, 0), ("c", "count", "s_in", 0),
                                        ("m", "mean", "s_in", 1), ("sa", "sum", "ab", 0)], temp_dir))
    sa = exp.pop("sa")[0]
    fk, val = nv.DeviceColumn(fkey, fvalid), nv.DeviceColumn(value)
    dk, at, wt = nv.DeviceColumn(dim_key), nv.DeviceColumn(attr, avalid), nv.DeviceColumn(weight, wvalid)
    jt = nv.JoinTable(dk, [at, wt])
    order = nv.VGPU_ORDER_KEY if temp_dir else nv.VGPU_ORDER_FIRST_SEEN
    lib = nv.vgpu()
    outs = []
    knobs = (b"gb.star.dense", b"join.bloom", b"gb.merge.partials")
    for setting in ((2, 2, 2), (2, 1, 2), (2, 1, 1), (1, 2, 2)):
        old = [lib.vgpu_tune_get(k) for k in knobs]
        for k, v in zip(knobs, setting):
            assert lib.vgpu_tune_set(k, v) == 0
        try:
            gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0), ("count", 0, 0),
                                                              ("mean", 0, 1)], order, 101)
            gb.feed_join(len(fkey), [fk, val], [dk, at, wt], [(0, 1, 1, 2)], key=(1, 1), join=jt,
                         join_kind=nv.JOIN_NAMES[kind], probe_key_col=0)
            gb.finish()
            got = gb.fetch()
            gb.free()
        finally:
            for k, v in zip(knobs, old):
                lib.vgpu_tune_set(k, v)
        got = {"attr": got["key"], "s": got["a0"], "n": got["a1"], "c": got["a2"], "m": got["a3"]}
        cnt = np.maximum(exp["c"][0], 1)
        assert_tables_equal(got, exp, {"s": sa, "m": sa / cnt})
        outs.append(got)
    jt.free()
    for c in (fk, val, dk, at, wt):
        c.free()


def test_fused_filter_groupby(gpu, oracle):
    nv = gpu
    r = rng(9)
    n = 300_000
    key = r.integers(0, 5000, n)
    x = r.normal(size=n)
    dk, dx = nv.DeviceColumn(key), nv.DeviceColumn(x)
    prog = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.25)
    for hint in (0, 5000):
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE],
                        [("sum", 0, 0), ("n", -1, 0), ("mean", 0, 0), ("min", 0, 0), ("max", 0, 0)],
                        nv.VGPU_ORDER_FIRST_SEEN, hint)
        gb.feed(dk, [dx], n, 0, prog, [dx])
        gb.finish()
        got = gb.fetch()
        gb.free()
        t = nv.Table({"key": key, "x": x, "ax": np.abs(x)}, rg_rows=65536)
        exp = nv.collect(nv.group_agg_node("oracle", nv.filter_node("oracle", t.scan(),
                                                                     ("cmp", ">", ("col", "x"), ("dbl", 0.25))),
                                           ["key"], [("a0", "sum", "x", 0), ("a1", "n", None, 0), ("a2", "mean", "x", 0),
                                                     ("a3", "min", "x", 0), ("a4", "max", "x", 0),
                                                     ("sa", "sum", "ax", 0)], None))
        sa = exp.pop("sa")[0]
        assert_tables_equal({"key": got.pop("key"), **got}, exp, {"a0": sa, "a2": sa / exp["a1"][0]})
    dk.free()
    dx.free()


@pytest.mark.parametrize("n,bits", [(1, 64), (1000, 64), (300_001, 40), (70_000, 13)])
def test_radix_sort_pairs(gpu, n, bits):
    nv = gpu
    r = rng(n)
    keys = r.integers(0, 2**63 - 1, n, dtype=np.int64).astype(np.uint64)
    if bits < 64:
        keys &= np.uint64((1 << bits) - 1)
    keys[: n // 3] = keys[n // 2]          # ties: stability matters
    vals = np.arange(n, dtype=np.uint32)
    lib = nv.vgpu()
    dk, dv = lib.vgpu_malloc(n * 8), lib.vgpu_malloc(n * 4)
    assert lib.vgpu_memcpy_h2d(dk, keys.ctypes.data, n * 8) == 0
    assert lib.vgpu_memcpy_h2d(dv, vals.ctypes.data, n * 4) == 0
    assert lib.vgpu_sort_pairs_u64(dk, dv, n, 0, bits) == 0, lib.vgpu_last_error()
    ok, ov = np.zeros(n, np.uint64), np.zeros(n, np.uint32)
    lib.vgpu_memcpy_d2h(ok.ctypes.data, dk, n * 8)
    lib.vgpu_memcpy_d2h(ov.ctypes.data, dv, n * 4)
    lib.vgpu_free(dk)
    lib.vgpu_free(dv)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok, keys[order])
    assert np.array_equal(ov, vals[order])


def test_fused_too_many_matches_is_an_error(gpu):
    """a build key with more than 4,096 rows: the fused feed cannot keep the
    pairs' output order and says so instead of merging positions"""
    nv = gpu
    n_dim = 5000
    dim_key = np.zeros(n_dim, np.int64)          # one key, 5,000 rows
    dim_key[-1] = 1
    fkey = np.array([0, 1, 2], np.int64)
    value = np.ones(3)
    cols = [nv.DeviceColumn(a) for a in (fkey, value, dim_key, np.arange(n_dim))]
    fk, val, dk, at = cols
    jt = nv.JoinTable(dk)
    gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0)], nv.VGPU_ORDER_FIRST_SEEN)
    with pytest.raises(nv.VectraError, match="at most 4096 matches per probe row"):
        gb.feed_join(3, [fk, val], [dk, at], [(0, 1)], key=(1, 1), join=jt, join_kind=nv.JOIN_INNER)
    gb.free()
    jt.free()
    for c in cols:
        c.free()
