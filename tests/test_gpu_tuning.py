"""Results do not depend on the tuning table.  The same group-bys (four
cardinalities from 16 to 1.5M keys, key order and first-seen order) and
joins (the fused join -> aggregate pass and a JoinNode's pairs) run under
the built-in defaults, under the autotuner's table (vectra_amd/tuning.txt)
and under an extreme table (every load at its limit, small launch shapes,
partitioning from 1,024 groups).  Keys, group order, counts, int64 sums,
min and max and join rows are identical; double sums and means agree
within 1e-12 * sum |x| of the group."""
import numpy as np
import pytest

from helpers import REL_TOL, rng

pytestmark = pytest.mark.gpu

EXTREME = """0 gb.glb.load_pct 90
0 join.load_pct 95
0 gb.lds.slots_per_group 8
0 gb.lds.copies 4
0 gb.lds.rows_per_block 512
0 gb.lds.max_blocks 64
0 gb.glb.rows_per_block 128
0 gb.glb.max_blocks 256
0 gb.part.min_groups 1024
0 gb.star.items 1
"""

AGGS = [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("sum", 1, 0), ("mean", 1, 0)]
EXACT = [0, 1, 2, 3]     # n and the int64 aggregates
CLOSE = [4, 5]           # sum / mean of the double


@pytest.fixture
def tables(gpu, tmp_path):
    lib = gpu.vgpu()
    ext = tmp_path / "extreme.txt"
    ext.write_text(EXTREME)
    tuned = gpu.PKG / "tuning.txt"

    def use(name):
        lib.vgpu_tune_reset()
        if name == "tuned":
            assert lib.vgpu_tune_load(str(tuned).encode()) > 0
        elif name == "extreme":
            assert lib.vgpu_tune_load(str(ext).encode()) == 10
    yield use
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_load(str(tuned).encode()) > 0


def run_groupby(gpu, key, ints, dbls, n, order):
    gb = gpu.GroupBy(gpu.VEC_INT64, [gpu.VEC_INT64, gpu.VEC_DOUBLE], AGGS, order, 0)
    half = n // 2     # two feeds: the second grows the table the first sized
    gb.feed(key, [ints, dbls], half)
    tail = [gpu.DeviceColumn.wrap(c.data + half * 8, n - half, c.type) for c in (key, ints, dbls)]
    gb.feed(tail[0], tail[1:], n - half, pos_base=half)
    gb.finish()
    out = gb.fetch()
    gb.free()
    return out


@pytest.mark.parametrize("card", [16, 3_000, 70_000, 1_500_000])
def test_groupby_same_under_every_table(gpu, tables, card):
    r = rng(card)
    n = 3_000_000
    k = r.integers(0, card, n)
    iv = r.integers(-2**40, 2**40, n)
    dv = r.normal(size=n)
    key, ints, dbls = gpu.DeviceColumn(k), gpu.DeviceColumn(iv), gpu.DeviceColumn(dv)
    absx = {}
    try:
        for order in (gpu.VGPU_ORDER_KEY, gpu.VGPU_ORDER_FIRST_SEEN):
            res = {}
            for name in ("defaults", "tuned", "extreme"):
                tables(name)
                res[name] = run_groupby(gpu, key, ints, dbls, n, order)
            base = res["defaults"]
            if order not in absx:
                # sum |x| per group in the result's key order (the tolerance scale)
                keys = base["key"][0]
                absx[order] = np.bincount(np.searchsorted(np.sort(keys), k), weights=np.abs(dv),
                                          minlength=len(keys))[np.argsort(np.argsort(keys))]
            scale = absx[order]
            for name in ("tuned", "extreme"):
                got = res[name]
                assert np.array_equal(got["key"][0], base["key"][0]), f"{name}: keys or order differ"
                for a in EXACT:
                    assert np.array_equal(got[f"a{a}"][0], base[f"a{a}"][0]), f"{name}: aggregate {a} differs"
                cnt = np.maximum(base["a0"][0], 1)
                for a, sc in ((4, scale), (5, scale / cnt)):
                    err = np.abs(got[f"a{a}"][0] - base[f"a{a}"][0])
                    assert np.all(err <= 2 * REL_TOL * np.maximum(sc, 1e-300)), f"{name}: aggregate {a} off"
    finally:
        for c in (key, ints, dbls):
            c.free()


def test_join_same_under_every_table(gpu, oracle, tables):
    """the fused join -> aggregate pass and a JoinNode's pairs (inner and
    left) under the three tables: the join table's load differs, rows and
    order do not"""
    r = rng(91)
    n_dim, n_fact = 200_000, 2_000_000
    dk = r.permutation(n_dim * 3)[:n_dim].astype(np.int64) * 5 + 1
    attr = r.integers(0, 300, n_dim)
    fk = np.where(r.random(n_fact) < 0.9, dk[r.integers(0, n_dim, n_fact)], -7)
    v = r.integers(-1000, 1000, n_fact)
    dkc, atc, fkc, vc = (gpu.DeviceColumn(a) for a in (dk, attr, fk, v))
    fused, pairs = {}, {}
    try:
        for name in ("defaults", "tuned", "extreme"):
            tables(name)
            jt = gpu.JoinTable(dkc, [atc])
            gb = gpu.GroupBy(gpu.VEC_INT64, [gpu.VEC_INT64], [("sum", 0, 0), ("n", -1, 0)], gpu.VGPU_ORDER_KEY, 301)
            gb.feed_join(n_fact, [fkc, vc], [dkc, atc], [(0, 1)], key=(1, 1), join=jt, join_kind=gpu.JOIN_INNER,
                         probe_key_col=0)
            gb.finish()
            fused[name] = gb.fetch()
            gb.free()
            jt.free()
            sub = 300_000
            for kind in ("inner", "left"):
                left = gpu.Table({"fk": fk[:sub], "v": v[:sub]}, rg_rows=70_000)
                right = gpu.Table({"dk": dk, "attr": attr})
                pairs[(name, kind)] = gpu.next_batches(gpu.join_node("gpu", left.scan(), right.scan(), kind,
                                                                     [(0, 0)]))
        for name in ("tuned", "extreme"):
            for col in ("key", "a0", "a1"):
                assert np.array_equal(fused[name][col][0], fused["defaults"][col][0]), f"{name}: {col} differs"
            for kind in ("inner", "left"):
                got, base = pairs[(name, kind)], pairs[("defaults", kind)]
                assert len(got) == len(base)
                for g, b in zip(got, base):
                    for c in b["cols"]:
                        assert np.array_equal(g["cols"][c][0], b["cols"][c][0]), f"{name} {kind}: {c} differs"
                        assert np.array_equal(g["cols"][c][1], b["cols"][c][1])
    finally:
        for c in (dkc, atc, fkc, vc):
            c.free()
