"""The serving loop (include/vx_serve.h): every request's result is the same
whether it ran alone (plain mode) or in a continuous batch with others, and
both equal the reference pipeline run on the host (the oracle):

  class 0: filter(value > t) |> group_by(group) |> summarise(s = sum(value), n = n())
  class 1: filter(value > t) |> inner_join(dim, key) |> group_by(attr) |> summarise(...)

Keys and counts exact; sums within 1e-12 of the group's sum (values are
positive, so the sum is the sum of magnitudes)."""
import numpy as np
import pytest

from helpers import REL_TOL, rng

pytestmark = pytest.mark.gpu


def make_tables(nv, seed, n_fact, n_dim):
    r = rng(seed)
    dk = r.permutation(n_dim).astype(np.int64) * 3 + 17
    attr = r.integers(0, 64, n_dim)
    ranks = np.minimum((r.pareto(1.0, n_fact)).astype(np.int64), n_dim - 1)
    fk = np.where(r.random(n_fact) < 0.95, dk[ranks], -5)
    value = r.uniform(0, 1000, n_fact)
    group = r.integers(0, 300, n_fact)
    host = {"key": fk, "value": value, "group": group, "dk": dk, "attr": attr}
    fact = {"key": nv.DeviceColumn(fk), "value": nv.DeviceColumn(value), "group": nv.DeviceColumn(group)}
    dim = {"key": nv.DeviceColumn(dk), "attr": nv.DeviceColumn(attr)}
    return host, fact, dim


def requests(seed, n, burst):
    r = rng(seed)
    out = []
    for i in range(n):
        cls = 0 if r.random() < 0.7 else 1
        t = 1000.0 * (1.0 - r.uniform(0.01, 0.5))
        at = 0.0 if i < burst else 0.002 * (i - burst)
        out.append((cls, float(t), at))
    return out


def same(a, b):
    assert np.array_equal(a["key"][0], b["key"][0]) and np.array_equal(a["key"][1], b["key"][1])
    assert np.array_equal(a["n"][0], b["n"][0])
    assert np.array_equal(a["s"][1], b["s"][1])
    s1, s2 = a["s"][0], b["s"][0]
    assert np.all(np.abs(s1 - s2) <= REL_TOL * np.abs(s2) + 1e-300)


def oracle_request(nv, host, cls, t):
    f = nv.filter_node("oracle", nv.Table({"key": host["key"], "value": host["value"], "group": host["group"]},
                                          rg_rows=0).scan(), ("cmp", ">", ("col", "value"), ("dbl", t)))
    if cls == 0:
        g = nv.group_agg_node("oracle", f, ["group"], [("s", "sum", "value", 0), ("n", "n", None, 0)], "/tmp")
        out = nv.collect(g)
        return {"key": out["group"], "s": out["s"], "n": out["n"]}
    j = nv.join_node("oracle", f, nv.Table({"dk": host["dk"], "attr": host["attr"]}, rg_rows=0).scan(), "inner",
                     [(0, 0)])
    out = nv.collect(nv.group_agg_node("oracle", j, ["attr"], [("s", "sum", "value", 0), ("n", "n", None, 0)], "/tmp"))
    return {"key": out["attr"], "s": out["s"], "n": out["n"]}


def test_serve_same_alone_and_in_the_mix(gpu, oracle):
    host, fact, dim = make_tables(gpu, 41, 2_000_000, 20_000)
    reqs = requests(43, 40, burst=24)
    mixed, st = gpu.serve(fact, dim, reqs, batched=True, max_batch=8, in_flight=2)
    alone, st_plain = gpu.serve(fact, dim, reqs, batched=False)
    assert st["n_done"] == len(reqs) and st_plain["n_done"] == len(reqs)
    assert st["max_pass"] > 1, "the burst should have been served in shared passes"
    assert max(r["batch"] for r in mixed) > 1
    assert st["slots"] == 2 and st["max_in_flight"] == 2, "two passes should have been on the device at once"
    # passes overlap in time: some pass started before another one ended
    spans = sorted({(r["pass"], r["start_s"], r["end_s"]) for r in mixed}, key=lambda x: x[1])
    assert any(b[1] < a[2] for a, b in zip(spans, spans[1:])), "no two passes overlapped"
    for m, a in zip(mixed, alone):
        same(m, a)
    for i in (0, 1, 5, 23, 39):
        cls, t, _ = reqs[i]
        same(mixed[i], oracle_request(gpu, host, cls, t))
    for c in list(fact.values()) + list(dim.values()):
        c.free()


def test_serve_budget_limits_the_pass(gpu, oracle):
    """a query-state budget of two requests caps every pass at two"""
    host, fact, dim = make_tables(gpu, 47, 300_000, 5_000)
    reqs = requests(53, 12, burst=12)
    probe, st0 = gpu.serve(fact, dim, reqs[:1], batched=True, max_batch=8, keep_results=False)
    res, st = gpu.serve(fact, dim, reqs, batched=True, max_batch=8, state_budget=2 * st0["state_bytes"])
    assert st["max_pass"] <= 2 and st["n_done"] == len(reqs)
    for i, (cls, t, _) in enumerate(reqs):
        same(res[i], oracle_request(gpu, host, cls, t))
    for c in list(fact.values()) + list(dim.values()):
        c.free()


def test_serve_many_in_flight_same_results(gpu, oracle):
    """four pass slots over a dense arrival stream of both classes: every
    result equals the plain one-at-a-time result"""
    host, fact, dim = make_tables(gpu, 59, 1_000_000, 10_000)
    reqs = requests(61, 64, burst=40)
    mixed, st = gpu.serve(fact, dim, reqs, batched=True, max_batch=6, in_flight=4)
    alone, _ = gpu.serve(fact, dim, reqs, batched=False)
    assert st["n_done"] == len(reqs) and st["slots"] == 4 and st["max_in_flight"] > 1
    for m, a in zip(mixed, alone):
        same(m, a)
    for i in (2, 17, 63):
        cls, t, _ = reqs[i]
        same(mixed[i], oracle_request(gpu, host, cls, t))
    for c in list(fact.values()) + list(dim.values()):
        c.free()


def test_serve_hbm_budget_admission(gpu, oracle):
    """the budget counts resident columns, the join table and every pass
    slot's state: a budget that holds one slot runs one pass at a time, and
    one below the resident columns is refused"""
    host, fact, dim = make_tables(gpu, 67, 300_000, 5_000)
    reqs = requests(71, 16, burst=16)
    _, st1 = gpu.serve(fact, dim, reqs[:1], batched=True, max_batch=4, in_flight=1, keep_results=False)
    one_slot = st1["resident_bytes"]
    res, st = gpu.serve(fact, dim, reqs, batched=True, max_batch=4, in_flight=4, hbm_budget=one_slot)
    assert st["slots"] == 1 and st["max_in_flight"] == 1 and st["hbm_budget"] == one_slot
    assert st["resident_bytes"] <= one_slot
    for i, (cls, t, _) in enumerate(reqs):
        same(res[i], oracle_request(gpu, host, cls, t))
    with pytest.raises(gpu.VectraError, match="exceed the budget"):
        gpu.serve(fact, dim, reqs[:1], batched=True, hbm_budget=1 << 20)
    for c in list(fact.values()) + list(dim.values()):
        c.free()


def test_serve_pass_size_per_class(gpu, oracle):
    """a class whose bucketed scan keeps a table of groups x requests in LDS
    takes passes only as large as that table allows (class 0 here: 301
    groups, at most 44 requests), the other class the full max_batch; a
    burst of 240 requests forms full passes of both, and every result equals
    the plain one-at-a-time result and, for a sample, the oracle"""
    host, fact, dim = make_tables(gpu, 67, 1_000_000, 10_000)
    reqs = requests(71, 240, burst=240)
    mixed, st = gpu.serve(fact, dim, reqs, batched=True, max_batch=64, in_flight=2)
    alone, _ = gpu.serve(fact, dim, reqs, batched=False)
    assert st["n_done"] == len(reqs)
    b0 = max(m["batch"] for m, (c, _, _) in zip(mixed, reqs) if c == 0)
    b1 = max(m["batch"] for m, (c, _, _) in zip(mixed, reqs) if c == 1)
    assert 1 < b0 <= 44 < b1 <= 64, (b0, b1)
    for m, a in zip(mixed, alone):
        same(m, a)
    for i in (0, 41, 239):
        cls, t, _ = reqs[i]
        same(mixed[i], oracle_request(gpu, host, cls, t))
    for c in list(fact.values()) + list(dim.values()):
        c.free()
