#!/usr/bin/env python3
"""Benchmark of vectra's relational hot path on one MI355X.

Default workload (c5, the largest configuration that fits one GPU): the
serving loop over device-resident tables.  Fact 1,000,000,000 rows (key
int64 Zipf(1.0) over the 10,000,000 keys of a dimension table, value
U(0,1000), category U{0..999}; about 24 GB), dimension 10,000,000 rows
(unique shuffled int64 key, attr U{0..255}).  Poisson arrivals at a swept
offered load: 70% of the requests are

    tbl |> filter(value > t) |> group_by(category) |> summarise(s = sum(value), n = n())

and 30% filter(value > t) |> inner_join(dim, by = c(key = "key")) |>
group_by(attr) |> summarise(...), with t drawn for 1-50% selectivity.  Each
load point runs twice over the same trace shape: one query at a time
(plain) and continuously batched (several requests of a class share one
scan; two passes on the device at once; slots admitted by device memory).
`value` is the batched mode's highest achieved queries/s with p99 latency
within the SLO (100 ms); the plain mode's figure at the same SLO and their
ratio are in the line.  The roofline is that of the serving pass's scan
kernel (k_bt_chain_feed, 8 class-0 queries over the 1B rows), timed with
HIP events over --steps passes after --warmup.

Other workloads (--workload): c2 (join-then-aggregate, 50M x 1M rows,
inner and left, through the kernel API), c1 (filter -> group-by through
the drop-in nodes, 1M rows in 10 row groups), c1-vtr (the same from a .vtr
file fed straight to the device), c3-high / c3-low (group-by over 200M rows
with 50M / 16 keys, 9 aggregates), c4 (4096 uneven filter -> group-by
queries in one persistent launch, replayed as a hipGraph), strings.

Prints ONE JSON line.  Data is synthetic (seeded generators).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "rows/s + HBM GB/s for filter+group-by/join; queries/s at p99 vs plain GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def pinned(fn):
    """Runs fn() on one core (the calling thread's affinity set to the last
    CPU it may use, as `taskset -c N` does), then restores the affinity.
    Returns (fn's result, a description of the host)."""
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[-1]
    os.sched_setaffinity(0, {core})
    try:
        out = fn()
    finally:
        os.sched_setaffinity(0, allowed)
    return out, {"pinned_core": core, "cpu_model": cpu_model(), "nproc": len(allowed),
                 "host_cpus": os.cpu_count()}


ORACLE_NOTE = ("the reference's C engine needs R's headers (R.h, Rinternals.h) to build, which this image lacks, "
               "so the baseline runs the oracle: a single-threaded C restatement of the reference's filter, "
               "hash join and SortNode-fed group-by (oracle/oracle.c)")


# ---------------------------------------------------------------- data

def zipf_ranks(r, n, s, n_keys):
    w = np.arange(1, n_keys + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, r.random(n), side="right"), n_keys - 1)


def c2_data(n_fact, n_dim, seed=0x76656374726102):
    r = np.random.default_rng(seed)
    keys = np.unique(r.integers(1, 2**62, int(n_dim * 2.1)))
    r.shuffle(keys)
    dim_key, absent = keys[:n_dim], keys[n_dim:2 * n_dim]
    r.shuffle(dim_key)
    attr = r.integers(0, 100, n_dim)
    weight = r.random(n_dim)
    ranks = zipf_ranks(r, n_fact, 1.1, n_dim)
    is_absent = r.random(n_fact) < 0.1
    fkey = np.where(is_absent, absent[ranks], dim_key[ranks])
    value = r.lognormal(0.0, 1.0, n_fact)
    return {"fkey": fkey, "value": value}, {"key": dim_key, "attr": attr, "weight": weight}


# ---------------------------------------------------------------- c2 on the GPU

class C2:
    def __init__(self, nv, fact, dim):
        self.nv = nv
        self.n_fact = len(fact["fkey"])
        self.n_dim = len(dim["key"])
        self.fk = nv.DeviceColumn(fact["fkey"])
        self.val = nv.DeviceColumn(fact["value"])
        self.dk = nv.DeviceColumn(dim["key"])
        self.at = nv.DeviceColumn(dim["attr"])
        self.wt = nv.DeviceColumn(dim["weight"])

    def free(self):
        for c in (self.fk, self.val, self.dk, self.at, self.wt):
            c.free()

    def query(self, jt, kind):
        nv = self.nv
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, 101)
        gb.feed_join(self.n_fact, [self.fk, self.val], [self.dk, self.at, self.wt], [(0, 1, 1, 2)], key=(1, 1),
                     join=jt, join_kind=nv.JOIN_NAMES[kind], probe_key_col=0)
        gb.finish()
        out = gb.fetch()
        gb.free()
        return out

    def step(self):
        # the dimension's attr and weight ride in the table entries
        jt = self.nv.JoinTable(self.dk, [self.at, self.wt])
        res = {k: self.query(jt, k) for k in ("inner", "left")}
        jt.free()
        return res

    def bytes_per_probe_launch(self):
        # fact fkey + value read once; dimension attr + weight as stored (8 B each)
        return self.n_fact * 16 + self.n_dim * 16


def cpu_baseline_c2(nv, fact, dim, sample_rows):
    """The oracle (C restatement of the reference engine: hash join, hidden
    mutate, SortNode-fed group-by) on the first sample_rows fact rows, one
    thread, both queries."""
    nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
    orc = nv._libs["oracle"]
    f = {k: v[:sample_rows] for k, v in fact.items()}
    t0 = time.perf_counter()
    for kind in ("inner", "left"):
        left = nv.Table(f, rg_rows=1 << 20)
        right = nv.Table(dim)
        j = nv.join_node("oracle", left.scan(), right.scan(), kind, [(0, 0)])
        p = orc.oracle_project_mul_node_create(j, b".vectra_tmp_s", b"value", b"weight")
        p._keep = j
        g = nv.group_agg_node("oracle", p, ["attr"], [("s", "sum", ".vectra_tmp_s", 0), ("n", "n", None, 0)], "/tmp")
        nv.collect(g)
    return time.perf_counter() - t0


def run_c2(args, nv):
    n_fact, n_dim = args.fact_rows, args.dim_rows
    log(f"[bench] generating c2 data: fact {n_fact:,} rows, dim {n_dim:,} rows")
    t = time.perf_counter()
    fact, dim = c2_data(n_fact, n_dim)
    log(f"[bench] data ready in {time.perf_counter() - t:.1f}s; uploading")
    w = C2(nv, fact, dim)
    lib = nv.vgpu()
    for i in range(args.warmup):
        res = w.step()
    log("[bench] warmup done")
    # sanity: counts of the two queries
    inner_n = res["inner"]["a1"][0].sum()
    left_n = res["left"]["a1"][0].sum()
    assert left_n == n_fact, f"left join row count {left_n} != {n_fact}"
    assert inner_n == int((np.isin(fact["fkey"], dim["key"])).sum()) if n_fact <= 5_000_000 else inner_n > 0
    lib.vgpu_synchronize()
    lib.vgpu_timer_start()
    t0 = time.perf_counter()
    for i in range(args.steps):
        w.step()
    lib.vgpu_synchronize()
    wall = time.perf_counter() - t0
    ev_ms = lib.vgpu_timer_stop()
    ms = wall * 1e3 / args.steps
    rows_s = 2 * n_fact / (ms / 1e3)
    log(f"[bench] {args.steps} steps: {ms:.3f} ms/step wall, {ev_ms / args.steps:.3f} ms/step events")
    # dominant kernel: fused probe -> aggregate, timed per launch with events
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    for i in range(3):
        w.step()
    import ctypes as C
    nl = C.c_int64()
    kms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    avg_ms = kms / max(nl.value, 1)
    per_launch = w.bytes_per_probe_launch()
    achieved = per_launch / (avg_ms / 1e3) / 1e9
    log(f"[bench] k_gb_feed_join: {nl.value} launches, avg {avg_ms:.3f} ms, {achieved:.0f} GB/s")
    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n_fact)
        log(f"[bench] cpu baseline on {sample:,} fact rows")
        secs, hostinfo = pinned(lambda: cpu_baseline_c2(nv, fact, dim, sample))
        cpu = {"value": 2 * sample / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample:,} fact rows of the same tables, inner + left join -> group_by(attr) "
                         f"({secs:.1f} s); one thread pinned to core {hostinfo['pinned_core']} of "
                         f"{hostinfo['nproc']} ({hostinfo['cpu_model']}); {ORACLE_NOTE}",
               **hostinfo}
    w.free()
    gbs = (2 * per_launch) / (ms / 1e3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "k_gb_star_join_aggregate_50M.json"
    if tfile.exists() and n_fact == 50_000_000 and n_dim == 1_000_000:
        traffic = json.loads(tfile.read_text())["traffic_bytes_per_launch"]
    return {
        "metric": METRIC, "value": rows_s, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators)",
        "config": {"workload": "join-then-aggregate: fact 50M rows (fkey Zipf(1.1) over 1M dim keys, 10% absent; "
                               "value lognormal(0,1)) x dim 1M rows (unique shuffled int64 key, attr U[0,100), "
                               "weight U(0,1)); inner_join and left_join on fkey = key, then group_by(attr) |> "
                               "summarise(s = sum(value * weight), n = n()); device-resident inputs",
                   "fact_rows": n_fact, "dim_rows": n_dim, "queries_per_step": 2},
        "gb_per_s": gbs,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/k_gb_star_join_aggregate_50M.md"
                     if traffic else None,
                     "kernel": "k_gb_star (fused probe -> aggregate)", "bytes_per_launch": per_launch,
                     "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- c1: drop-in nodes

def run_c2_nodes(args, nv):
    """The join-then-aggregate query as R builds it and collect() runs it:
    GroupAgg([attr], s = sum(.vectra_tmp_s), n = n()) <- Project(...,
    .vectra_tmp_s = value * weight) <- Join(fact, dim, fkey = key), inner and
    left, through the drop-in nodes (the GPU GroupAggNode fuses the three
    into one device pass).  The fact table is host-resident and pulled in
    row groups of --vtr-rg rows, as a scan hands them over, so every step
    includes the copies of the fact columns to the device."""
    import ctypes as C
    n_fact, n_dim = args.fact_rows, args.dim_rows
    log(f"[bench] c2-nodes: generating fact {n_fact:,} rows, dim {n_dim:,} rows")
    fact, dim = c2_data(n_fact, n_dim)
    ft = nv.Table({"fkey": fact["fkey"], "value": fact["value"]}, rg_rows=args.vtr_rg)
    dt = nv.Table(dim, rg_rows=0)
    aggs = [("s", "sum", ".vectra_tmp_s", 0), ("n", "n", None, 0)]

    def tree(kind):
        j = nv.join_node("gpu", ft.scan(), dt.scan(), kind, [(0, 0)])
        return nv.group_agg_node("gpu", nv.hidden_mutate(j, ".vectra_tmp_s", "value", "weight"), ["attr"], aggs,
                                 "/tmp")

    def step():
        return {kind: nv.collect(tree(kind)) for kind in ("inner", "left")}

    node = tree("inner")
    nv.collect(node, free=False)
    note = "\n".join(nv.explain(node))
    nv.core().vx_node_free(node)
    fused = "fused with the Project and Join below" in note
    log(f"[bench] c2-nodes: fused pass {'taken' if fused else 'NOT taken'}")
    for _ in range(args.warmup):
        res = step()
    assert int(res["left"]["n"][0].sum()) == n_fact, "left join keeps every fact row"
    lib = nv.vgpu()
    lib.vgpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    lib.vgpu_synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    rows_s = 2 * n_fact / (ms / 1e3)
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    step()
    nl = C.c_int64()
    kms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    avg_ms = kms / max(nl.value, 1)
    rows_per_launch = 2 * n_fact / max(nl.value, 1)
    per_launch = rows_per_launch * 16      # fact fkey + value read once per launch
    achieved = per_launch / (avg_ms / 1e3) / 1e9
    log(f"[bench] c2-nodes: {ms:.2f} ms/step ({rows_s:.3e} fact rows/s); fused kernel {nl.value} launches, "
        f"avg {avg_ms:.3f} ms, {achieved:.0f} GB/s; kernels {kms:.2f} ms of the step")
    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n_fact)
        secs, hostinfo = pinned(lambda: cpu_baseline_c2(nv, fact, dim, sample))
        cpu = {"value": 2 * sample / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample:,} fact rows, inner + left join -> group_by(attr) ({secs:.1f} s); one "
                         f"thread pinned to core {hostinfo['pinned_core']} of {hostinfo['nproc']} "
                         f"({hostinfo['cpu_model']}); {ORACLE_NOTE}", **hostinfo}
    return {
        "metric": METRIC, "value": rows_s, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators)",
        "config": {"workload": "join-then-aggregate through the drop-in node tree R builds (GroupAgg <- "
                               "Project(value * weight) <- Join), inner + left; fact 50M rows host-resident in "
                               f"row groups of {args.vtr_rg:,} (host-to-device copies inside every step), dim 1M",
                   "fact_rows": n_fact, "dim_rows": n_dim, "row_group": args.vtr_rg, "fused": fused},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_gb_star (fused probe -> product -> aggregate, one launch per row group)",
                     "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches": nl.value,
                     "kernel_ms_per_step": kms},
        "cpu_baseline": cpu,
    }


def run_strings(args, nv):
    """group_by(s, k) |> summarise(n = n(), total = sum(x)) and
    inner_join(dim, by = "s") |> group_by(w) |> summarise(n = n()) over
    string keys: 10M fact rows in host row groups of 1M (s from 20,000
    strings of 15-23 bytes, k U{0..9}, x N(0,1)), a 20,000-row dim table with
    an int64 w; through the drop-in GroupAggNode and JoinNode (key strings
    staged over PCIe and hashed on the device)."""
    import ctypes as C
    n = args.str_rows
    r = np.random.default_rng(0x7665637472610600)
    pool = np.array([f"customer/{i:05d}/" + "abcdefgh"[: i % 9] for i in range(20_000)], dtype=object)
    s = pool[r.integers(0, len(pool), n)]
    cols = {"s": s, "k": r.integers(0, 10, n), "x": r.normal(size=n)}
    fact = nv.Table(cols, rg_rows=1_000_000)
    dim = nv.Table({"s": pool[r.permutation(len(pool))], "w": r.integers(0, 100, len(pool))})
    str_bytes = int(sum(len(v) for v in pool[r.integers(0, len(pool), 10_000)]) / 10_000 * n)

    def group(engine, tbl):
        return nv.collect(nv.group_agg_node(engine, tbl.scan(), ["s", "k"], [("n", "n", None, 0),
                                                                             ("total", "sum", "x", 0)], None))

    def join(engine, tbl):
        j = nv.join_node(engine, tbl.scan(), dim.scan(), "inner", [(0, 0)])
        return nv.collect(nv.group_agg_node(engine, j, ["w"], [("n", "n", None, 0)], "/tmp"))

    lib = nv.vgpu()
    for _ in range(args.warmup):
        g, jn = group("gpu", fact), join("gpu", fact)
    assert int(g["n"][0].sum()) == n and int(jn["n"][0].sum()) == n
    res = {}
    for name, fn in (("group", group), ("join", join)):
        lib.vgpu_synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn("gpu", fact)
        lib.vgpu_synchronize()
        res[name] = (time.perf_counter() - t0) * 1e3 / args.steps
        lib.vgpu_kernel_timing(1)
        lib.vgpu_kernel_time_ms(None, 1)
        fn("gpu", fact)
        nl = C.c_int64()
        res[name + "_k"] = (lib.vgpu_kernel_time_ms(C.byref(nl), 1), nl.value)
        lib.vgpu_kernel_timing(0)
    log(f"[bench] strings: group_by(s, k) {res['group']:.1f} ms ({n / res['group'] / 1e3:.0f} M rows/s, kernels "
        f"{res['group_k'][0]:.2f} ms in {res['group_k'][1]} launches); join on s + group_by(w) {res['join']:.1f} ms "
        f"({n / res['join'] / 1e3:.0f} M rows/s, kernels {res['join_k'][0]:.2f} ms in {res['join_k'][1]} launches); "
        f"about {str_bytes / 1e6:.0f} MB of key strings")
    cpu = None
    if not args.no_cpu:
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        m = min(n, 2_000_000)
        sample = nv.Table({k: v[:m] for k, v in cols.items()}, rg_rows=1_000_000)
        t0 = time.perf_counter()
        group("oracle", sample)
        join("oracle", sample)
        secs = time.perf_counter() - t0
        cpu = {"value": m / secs * 2, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the first {m:,} fact rows through both queries (oracle restatement of the reference "
                         f"engine, {secs:.2f} s); rows/s counts each query's rows"}
    ms = res["group"] + res["join"]
    kms = res["group_k"][0] + res["join_k"][0]
    # algorithmic bytes of both queries: the key strings and their offsets
    # read once per query, k and x (group) and the dim table (join)
    alg = 2 * (str_bytes + 8 * n) + 16 * n + 20_000 * 32
    achieved = alg / (kms / 1e3) / 1e9
    return {
        "metric": METRIC, "value": 2 * n / (ms / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators)",
        "config": {"workload": "string keys: group_by(s, k) |> summarise(n, sum(x)) and inner_join(dim, by = s) |> "
                               "group_by(w) |> summarise(n); 10M fact rows in host row groups of 1M, 20,000 distinct "
                               "strings; drop-in nodes, key strings staged over PCIe",
                   "rows": n, "distinct_strings": 20_000},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_string_keys + k_key_codes + group-by / join kernels, all launches of both queries",
                     "bytes_per_launch": alg / max(res["group_k"][1] + res["join_k"][1], 1),
                     "avg_launch_ms": kms / max(res["group_k"][1] + res["join_k"][1], 1)},
        "cpu_baseline": cpu,
    }


def run_c1(args, nv):
    """filter(x > 0) |> group_by(key) |> summarise(total = sum(value), n = n(),
    avg = mean(value)) over 1M rows in 10 row groups of 100,000, through the
    drop-in FilterNode and GroupAggNode (host batches staged per row group)."""
    n = 1_000_000
    r = np.random.default_rng(0x7665637472610100)
    cols = {"id": np.arange(n, dtype=np.int64), "key": r.integers(0, 1000, n), "x": r.normal(size=n),
            "value": r.uniform(0, 100, n)}
    t = nv.Table(cols, rg_rows=100_000)
    aggs = [("total", "sum", "value", 0), ("n", "n", None, 0), ("avg", "mean", "value", 0)]
    pred = ("cmp", ">", ("col", "x"), ("dbl", 0.0))

    def query(engine):
        node = nv.filter_node(engine, t.scan(), pred)
        return nv.collect(nv.group_agg_node(engine, node, ["key"], aggs, "/tmp"))

    lib = nv.vgpu()
    for _ in range(args.warmup):
        res = query("gpu")
    keep = cols["x"] > 0
    assert int(res["n"][0].sum()) == int(keep.sum()) and list(res["key"][0]) == list(range(1000))
    lib.vgpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        query("gpu")
    lib.vgpu_synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    import ctypes as C
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    query("gpu")
    nl = C.c_int64()
    kms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    # algorithmic bytes of the query: x, key, value read once; 1,000 x 4 outputs
    alg = n * 24 + 1000 * 4 * 8
    achieved = alg / (kms / 1e3) / 1e9
    log(f"[bench] c1: {ms:.3f} ms/query wall; {nl.value} timed kernels, {kms:.3f} ms total")
    cpu = None
    if not args.no_cpu:
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            query("oracle")
        secs = (time.perf_counter() - t0) / reps
        cpu = {"value": n / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the full 1M-row query, {reps} runs (oracle restatement of the reference engine, "
                         f"{secs * 1e3:.0f} ms per query)"}
    return {
        "metric": METRIC, "value": n / (ms / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators)",
        "config": {"workload": "filter(x > 0) |> group_by(key) |> summarise(total = sum(value), n = n(), "
                               "avg = mean(value)); 1M rows as 10 row groups of 100,000 (key U{0..999}, x N(0,1), "
                               "value U(0,100)); drop-in FilterNode + GroupAggNode, host batches staged over PCIe",
                   "rows": n, "row_groups": 10},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_filter_mask/k_filter_write + k_gb_feed, all launches of one query",
                     "bytes_per_launch": alg / max(nl.value, 1), "avg_launch_ms": kms / max(nl.value, 1)},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- c3: group-by over 200M rows

C3_AGGS = [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("mean", 0, 0),
           ("sum", 1, 0), ("min", 1, 0), ("max", 1, 0), ("mean", 1, 0)]


def c3_data(n, card, seed=0x7665637472610300):
    """key: U over `card` distinct int64 (card > 16: random distinct values;
    else 0..card-1); v1 lognormal(0,1); v2 int64 U[-1e9, 1e9]."""
    r = np.random.default_rng(seed + card)
    idx = r.integers(0, card, n)
    present = int(np.count_nonzero(np.bincount(idx, minlength=card)))
    if card > 16:
        ks = np.unique(r.integers(-2**62, 2**62, int(card * 1.01) + 1024))
        r.shuffle(ks)
        key = ks[:card][idx]
    else:
        key = idx
    del idx
    v1 = r.lognormal(0.0, 1.0, n)
    v2 = r.integers(-10**9, 10**9, n, endpoint=True)
    return {"key": key, "v1": v1, "v2": v2}, present


def run_c1_vtr(args, nv):
    """The C1 query read from a .vtr file: filter(x > 0) |> group_by(key) |>
    summarise(total = sum(value), n = n(), avg = mean(value)) through the
    drop-in nodes over scan_node_create, the filter's predicate pushed onto
    the scan.  With the device scan feed (knob scan.device_feed) the GPU
    group-by reads each row group into pinned staging and copies it on the
    copy stream while the previous group's kernels run; without it the
    nodes pull host batches (malloc'd by the scan, uploaded by the filter,
    selection vector back to the host, uploaded again by the group-by).
    The file is written first and read from the page cache, so the rates
    include PCIe and the file reads but not a disk."""
    import ctypes as C
    n, rg = args.vtr_rows, args.vtr_rg
    path = Path(args.vtr_path)
    r = np.random.default_rng(0x7665637472610101)
    cols = {"id": np.arange(n, dtype=np.int64), "key": r.integers(0, 1000, n), "x": r.normal(size=n),
            "value": r.uniform(0, 100, n)}
    t0 = time.perf_counter()
    nv.vtr_write(path, cols, rg_rows=rg)
    log(f"[bench] c1-vtr: wrote {n:,} rows in row groups of {rg:,} to {path} "
        f"({path.stat().st_size / 1e9:.2f} GB) in {time.perf_counter() - t0:.1f} s")
    aggs = [("total", "sum", "value", 0), ("n", "n", None, 0), ("avg", "mean", "value", 0)]
    pred = ("cmp", ">", ("col", "x"), ("dbl", 0.0))

    def query(engine):
        node = nv.pushdown(nv.filter_node(engine, nv.scan_vtr(path), pred))
        return nv.collect(nv.group_agg_node(engine, node, ["key"], aggs, "/tmp"))

    lib = nv.vgpu()

    def timed(feed):
        old = lib.vgpu_tune_get(b"scan.device_feed")
        lib.vgpu_tune_set(b"scan.device_feed", feed)
        try:
            for _ in range(args.warmup):
                res = query("gpu")
            lib.vgpu_synchronize()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                query("gpu")
            lib.vgpu_synchronize()
            return res, (time.perf_counter() - t1) * 1e3 / args.steps
        finally:
            lib.vgpu_tune_set(b"scan.device_feed", old)

    res_on, ms_on = timed(2)
    res_off, ms_off = timed(1)
    keep = cols["x"] > 0
    for res in (res_on, res_off):
        assert int(res["n"][0].sum()) == int(keep.sum()) and list(res["key"][0]) == list(range(1000))
    # kernels of one device-fed query
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    query("gpu")
    nl = C.c_int64()
    kms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    alg = n * 24 + 1000 * 4 * 8       # key, x, value read once; 1,000 x 4 outputs
    log(f"[bench] c1-vtr: device feed {ms_on:.1f} ms/query ({n / ms_on / 1e3:.0f} M rows/s), host batches "
        f"{ms_off:.1f} ms/query ({n / ms_off / 1e3:.0f} M rows/s); kernels of one query {kms:.1f} ms")
    cpu = None
    if not args.no_cpu:
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        t1 = time.perf_counter()
        exp = query("oracle")
        secs = time.perf_counter() - t1
        assert list(exp["key"][0]) == list(res_on["key"][0]) and np.array_equal(exp["n"][0], res_on["n"][0])
        assert np.allclose(exp["total"][0], res_on["total"][0], rtol=1e-12, atol=0)
        cpu = {"value": n / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the full query over the same file ({n:,} rows), oracle nodes over the CPU scan "
                         f"({secs:.1f} s)"}
    path.unlink()
    return {
        "metric": METRIC, "value": n / (ms_on / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_on, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators), written to a .vtr file",
        "config": {"workload": f"filter(x > 0) |> group_by(key) |> summarise(total = sum(value), n = n(), "
                               f"avg = mean(value)) over a .vtr file of {n:,} rows in row groups of {rg:,} "
                               f"(id, key U{{0..999}}, x N(0,1), value U(0,100)); drop-in ScanNode -> FilterNode "
                               f"-> GroupAggNode, predicate pushed onto the scan; file in the page cache, rates "
                               f"include the file reads and PCIe",
                   "rows": n, "row_groups": (n + rg - 1) // rg},
        "host_batches": {"ms_per_query": ms_off, "rows_per_s": n / (ms_off / 1e3)},
        "device_feed_vs_host_batches": ms_off / ms_on,
        "roofline": {"bound": "hbm", "achieved": alg / (kms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "all kernels of one device-fed query", "bytes_per_launch": alg / max(nl.value, 1),
                     "avg_launch_ms": kms / max(nl.value, 1)},
        "cpu_baseline": cpu,
    }


def c3_traffic(n, card):
    """HBM bytes per feed from the committed rocprofv3 summary of the same
    workload (FETCH_SIZE / WRITE_SIZE, scaled), or None off that workload"""
    name = "k_gb_feed_groupby_200M_rows_16_keys.json" if card == 16 else \
        "k_gbp_groupby_200M_rows_50M_keys.json" if card == 50_000_000 else None
    tfile = ROOT / "profiles" / name if name else None
    if n != 200_000_000 or tfile is None or not tfile.exists():
        return None
    return json.loads(tfile.read_text())["traffic_bytes_per_launch"]


def run_c3(args, nv, card):
    n = args.c3_rows
    log(f"[bench] generating c3 data: {n:,} rows, {card:,} key values")
    t = time.perf_counter()
    data, present = c3_data(n, card)
    log(f"[bench] data ready in {time.perf_counter() - t:.1f}s ({present:,} keys present); uploading")
    kc, c1, c2 = (nv.DeviceColumn(data[c]) for c in ("key", "v1", "v2"))
    lib = nv.vgpu()

    def step(fetch=False):
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE, nv.VEC_INT64], C3_AGGS, nv.VGPU_ORDER_KEY, 0)
        gb.feed(kc, [c1, c2], n)
        ng = gb.finish()
        out = gb.fetch() if fetch else None
        gb.free()
        return ng, out

    rel_err = None
    for i in range(max(args.warmup, 1)):
        ng, out = step(fetch=(i == 0))
        if i == 0:
            # size-independent checks of the full result
            k = out["key"][0]
            assert ng == present, (ng, present)
            assert np.all(k[1:] > k[:-1]), "keys not in ascending order"
            assert int(out["a0"][0].sum()) == n
            assert int(np.round(out["a5"][0]).astype(np.int64).sum()) == int(data["v2"].sum())
            assert out["a2"][0].min() == data["v1"].min() and out["a3"][0].max() == data["v1"].max()
            assert out["a6"][0].min() == data["v2"].min() and out["a7"][0].max() == data["v2"].max()
            import math
            tot = math.fsum(out["a1"][0])
            ex = math.fsum(data["v1"])
            rel_err = abs(tot - ex) / ex
            log(f"[bench] sum of group sums of v1 vs fsum of v1: relative error {rel_err:.2e}")
            del out
    log("[bench] warmup done, result checked")
    lib.vgpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    lib.vgpu_synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    import ctypes as C
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    for _ in range(3):
        step()
    nl = C.c_int64()
    kms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    avg_ms = kms / max(nl.value, 1)
    per_launch = n * 24
    achieved = per_launch / (avg_ms / 1e3) / 1e9
    log(f"[bench] c3 card {card}: {ms:.3f} ms/step; feed avg {avg_ms:.3f} ms, {achieved:.0f} GB/s")
    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        tb = nv.Table({k: v[:sample] for k, v in data.items()}, rg_rows=1 << 20)
        aggs = [("n", "n", None, 0)] + [(f"{f}_{c}", f, c, 0) for c in ("v1", "v2")
                                        for f in ("sum", "min", "max", "mean")]
        t0 = time.perf_counter()
        nv.collect(nv.group_agg_node("oracle", tb.scan(), ["key"], aggs, "/tmp"))
        secs = time.perf_counter() - t0
        cpu = {"value": sample / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample:,} rows, sort-based group-by (oracle restatement of the reference "
                         f"engine's GroupAgg over SortNode, {secs:.1f} s)"}
    for c in (kc, c1, c2):
        c.free()
    label = "high" if card > 16 else "low"
    return {
        "metric": METRIC, "value": n / (ms / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded numpy generators)",
        "config": {"workload": f"group_by(key) |> summarise(n, sum/min/max/mean of v1 and of v2) over {n:,} rows, "
                               f"{label} cardinality ({card:,} key values, {present:,} present); v1 lognormal(0,1), "
                               "v2 int64 U[-1e9,1e9]; key order, result left in device memory; device-resident inputs",
                   "rows": n, "key_values": card, "groups": present},
        "result_check": {"groups": present, "sum_v1_rel_err_vs_fsum": rel_err},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": c3_traffic(n, card),
                     "kernel": "k_gb_feed" if card <= 16 else
                     "partitioned feed k_gbp_hist1 -> k_gbp_agg (timed as one feed)",
                     "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- c4: batched uneven queries

C4_AGGS = [("sum", 0, 0), ("n", -1, 0), ("mean", 0, 0)]


def c4_plan(n_queries, seed=0x7665637472610400):
    """rows log-uniform in [1e3, 1e7], key cardinality log-uniform in [1, 1e5],
    selectivity U[1%, 50%] -> threshold t = Phi^-1(1 - s) on a N(0,1) value."""
    from scipy.special import ndtri
    r = np.random.default_rng(seed)
    rows = np.floor(10 ** r.uniform(3, 7, n_queries)).astype(np.int64)
    card = np.floor(10 ** r.uniform(0, 5, n_queries)).astype(np.int64)
    sel = r.uniform(0.01, 0.5, n_queries)
    return rows, card, ndtri(1.0 - sel)


def run_c4(args, nv):
    import ctypes as C
    import torch
    Q = args.c4_queries
    rows, card, thr = c4_plan(Q)
    if args.c4_card > 0:     # every query with the same key cardinality (a sweep point, not the workload)
        card[:] = args.c4_card
    total_rows = int(rows.sum())
    log(f"[bench] c4: {Q} queries, {total_rows:,} rows ({total_rows * 16 / 1e9:.1f} GB), generating on the device")
    keys, vals = [], []
    gen = torch.Generator(device="cuda")
    for q in range(Q):
        gen.manual_seed(0x7665637472610400 + q)
        keys.append(torch.randint(0, int(card[q]), (int(rows[q]),), dtype=torch.int64, device="cuda", generator=gen))
        vals.append(torch.randn(int(rows[q]), dtype=torch.float64, device="cuda", generator=gen))
    torch.cuda.synchronize()
    kc = [nv.DeviceColumn.wrap(k.data_ptr(), k.numel(), nv.VEC_INT64, keep=k) for k in keys]
    vc = [nv.DeviceColumn.wrap(v.data_ptr(), v.numel(), nv.VEC_DOUBLE, keep=v) for v in vals]
    lib = nv.vgpu()

    # batched: plan (sketch + first run), one timed kernel-level run, graph capture, replays
    pred = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.0)
    b = nv.Batch(nv.VEC_INT64, [nv.VEC_DOUBLE], C4_AGGS, nv.VGPU_ORDER_KEY, pred, 1, 1)
    t0 = time.perf_counter()
    b.plan([{"key": kc[q], "inputs": [vc[q]], "pred_cols": [vc[q]], "n_rows": int(rows[q]), "param": float(thr[q])}
            for q in range(Q)])
    lib.vgpu_synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    lib.vgpu_kernel_timing(1)
    lib.vgpu_kernel_time_ms(None, 1)
    b.run()
    nl = C.c_int64()
    feed_ms = lib.vgpu_kernel_time_ms(C.byref(nl), 1)
    lib.vgpu_kernel_timing(0)
    b.capture()
    for _ in range(args.warmup):
        b.run()
    lib.vgpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    lib.vgpu_synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    n_groups, _ = b.counts()
    log(f"[bench] c4 batched: plan {plan_ms:.1f} ms, {ms:.3f} ms per batch of {Q} queries "
        f"(feed kernel {feed_ms:.3f} ms), {int(n_groups.sum()):,} groups")

    # sampled queries: exact against the oracle (also the CPU baseline)
    sample = list(range(0, Q, 64))
    cpu = None
    check = {"queries": len(sample), "groups_equal": True, "n_equal": True}
    if not args.no_cpu:
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        secs, srows = 0.0, 0
        for q in sample:
            kh, vh = keys[q].cpu().numpy(), vals[q].cpu().numpy()
            t = nv.Table({"key": kh, "value": vh}, rg_rows=1 << 20)
            t1 = time.perf_counter()
            node = nv.filter_node("oracle", t.scan(), ("cmp", ">", ("col", "value"), ("dbl", float(thr[q]))))
            exp = nv.collect(nv.group_agg_node("oracle", node, ["key"], [("total", "sum", "value", 0),
                                                                          ("n", "n", None, 0),
                                                                          ("avg", "mean", "value", 0)], "/tmp"))
            secs += time.perf_counter() - t1
            srows += len(kh)
            got = b.fetch(q)
            check["groups_equal"] &= bool(np.array_equal(got["key"][0], exp["key"][0]))
            check["n_equal"] &= bool(np.array_equal(got["a1"][0], exp["n"][0]))
        assert check["groups_equal"] and check["n_equal"], check
        cpu = {"value": srows / secs, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"every 64th query ({len(sample)} queries, {srows:,} rows), filter -> sort-based group-by "
                         f"(oracle restatement of the reference engine, {secs:.1f} s), extrapolated by rows"}

    # plain GPU form: one query per launch sequence, no persistence, no graph
    lat = []
    aggs = C4_AGGS
    t0 = time.perf_counter()
    for q in range(Q):
        t1 = time.perf_counter()
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], aggs, nv.VGPU_ORDER_KEY, 0)
        gb.feed(kc[q], [vc[q]], int(rows[q]), pred=nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", float(thr[q])),
                pred_cols=[vc[q]])
        gb.finish()
        gb.free()
        lib.vgpu_synchronize()
        lat.append(time.perf_counter() - t1)
    plain_s = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    log(f"[bench] c4 plain: {plain_s * 1e3:.1f} ms for {Q} queries (p50 {np.percentile(lat_ms, 50):.3f} ms, "
        f"p99 {np.percentile(lat_ms, 99):.3f} ms)")
    b.free()
    alg = total_rows * 16
    achieved = alg / (feed_ms / 1e3) / 1e9
    return {
        "metric": METRIC, "value": total_rows / (ms / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (torch generators on the device, seeded per query)",
        "config": {"workload": f"{Q} independent queries filter(value > t_q) |> group_by(key) |> "
                               "summarise(total = sum(value), n = n(), avg = mean(value)); rows log-uniform "
                               "[1e3, 1e7], key cardinality log-uniform [1, 1e5] (int64), value N(0,1), selectivity "
                               "U[1%, 50%]; one persistent feed launch per batch, replayed as a hipGraph; "
                               "device-resident inputs",
                   "queries": Q, "rows": total_rows, "groups": int(n_groups.sum()),
                   "fixed_key_cardinality": args.c4_card or None, "knobs": args.tune or None},
        "batched": {"queries_per_s": Q / (ms / 1e3), "ms_per_batch": ms, "plan_ms": plan_ms},
        "plain_gpu": {"queries_per_s": Q / plain_s, "total_ms": plain_s * 1e3,
                      "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99))},
        "batched_vs_plain": (Q / (ms / 1e3)) / (Q / plain_s),
        "result_check": check,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_bt_feed (all queries)",
                     "bytes_per_launch": alg, "avg_launch_ms": feed_ms},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- c5: serving loop

def c5_tables(n_fact, n_dim, seed=0x7665637472610500):
    """On the device: dim keys unique shuffled int64, attr U{0..255}; fact key
    = dim_keys[Zipf(1.0) rank over n_dim], value U(0,1000), category
    U{0..999} (int32 in the workload, stored as int64 like the reference)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dk = torch.randperm(n_dim, device="cuda", generator=g).to(torch.int64) * 7 + 1_000_003
    attr = torch.randint(0, 256, (n_dim,), dtype=torch.int64, device="cuda", generator=g)
    w = 1.0 / torch.arange(1, n_dim + 1, dtype=torch.float64, device="cuda")
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    key = torch.empty(n_fact, dtype=torch.int64, device="cuda")
    value = torch.empty(n_fact, dtype=torch.float64, device="cuda")
    group = torch.empty(n_fact, dtype=torch.int64, device="cuda")
    chunk = 100_000_000
    for a in range(0, n_fact, chunk):
        b = min(n_fact, a + chunk)
        u = torch.rand(b - a, dtype=torch.float64, device="cuda", generator=g)
        ranks = torch.clamp(torch.searchsorted(cdf, u), max=n_dim - 1)
        key[a:b] = dk[ranks]
        value[a:b] = torch.rand(b - a, dtype=torch.float64, device="cuda", generator=g) * 1000.0
        group[a:b] = torch.randint(0, 1000, (b - a,), dtype=torch.int64, device="cuda", generator=g)
        del u, ranks
    torch.cuda.synchronize()
    return {"key": key, "value": value, "group": group}, {"key": dk, "attr": attr}


def c5_requests(rate, n, seed):
    r = np.random.default_rng(seed)
    at = np.cumsum(r.exponential(1.0 / rate, n))
    at -= at[0]
    cls = (r.random(n) >= 0.7).astype(int)           # 70% group-by, 30% join -> group-by
    t = 1000.0 * (1.0 - r.uniform(0.01, 0.5, n))     # selectivity U[1%, 50%] of value U(0,1000)
    return [(int(c), float(x), float(a)) for c, x, a in zip(cls, t, at)]


def run_c5(args, nv):
    import ctypes as C
    n_fact, n_dim = args.c5_rows, args.c5_dim
    log(f"[bench] c5: fact {n_fact:,} rows, dim {n_dim:,} rows, generating on the device")
    fact_t, dim_t = c5_tables(n_fact, n_dim)
    wrap = lambda t, ty: nv.DeviceColumn.wrap(t.data_ptr(), t.numel(), ty, keep=t)
    fact = {"key": wrap(fact_t["key"], nv.VEC_INT64), "value": wrap(fact_t["value"], nv.VEC_DOUBLE),
            "group": wrap(fact_t["group"], nv.VEC_INT64)}
    dim = {"key": wrap(dim_t["key"], nv.VEC_INT64), "attr": wrap(dim_t["attr"], nv.VEC_INT64)}
    slo = args.c5_slo_ms
    sweep = []
    for mode, rates in (("plain", args.c5_plain_rates), ("batched", args.c5_batched_rates)):
        for rate in rates:
            n = int(min(max(60, rate * args.c5_seconds), 2000))
            reqs = c5_requests(rate, n, seed=int(rate * 1000) + (7 if mode == "batched" else 3))
            res, st = nv.serve(fact, dim, reqs, batched=(mode == "batched"), max_batch=args.c5_max_batch,
                               keep_results=False, in_flight=args.c5_in_flight, hbm_budget=args.c5_hbm_budget)
            lat = np.array([r["end_s"] - a for r, (_, _, a) in zip(res, reqs)]) * 1e3
            qps = n / max(st["wall_s"], 1e-9)
            pt = {"mode": mode, "offered_qps": rate, "requests": n, "achieved_qps": qps,
                  "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                  "passes": st["n_passes"], "max_pass": st["max_pass"], "max_in_flight": st["max_in_flight"],
                  "slots": st["slots"], "hbm_budget": st["hbm_budget"], "resident_bytes": st["resident_bytes"]}
            sweep.append(pt)
            log(f"[bench] c5 {mode:7s} offered {rate:6.0f}/s: achieved {qps:7.1f}/s p50 {pt['p50_ms']:8.2f} ms "
                f"p99 {pt['p99_ms']:8.2f} ms (passes {st['n_passes']}, largest {st['max_pass']}, "
                f"in flight {st['max_in_flight']} of {st['slots']})")

    def best(mode):
        ok = [p for p in sweep if p["mode"] == mode and p["p99_ms"] <= slo]
        return max(ok, key=lambda p: p["achieved_qps"]) if ok else None

    bb, bp = best("batched"), best("plain")
    ratio = bb["achieved_qps"] / bp["achieved_qps"] if bb and bp else None

    # dominant kernel: one bucketed pass of class-0 queries (and, for the
    # record, of class-1 queries), timed with events; a pass takes the
    # serving loop's size for its class: max_batch, or the largest size
    # whose shared scan is still bucketed (its groups x queries table fits
    # a CU's LDS), found by the same binary search
    lib = nv.vgpu()
    jt = nv.JoinTable(dim["key"], [dim["attr"]])
    pass_ms, pass_k = {}, {}

    def plan(cls, k):
        thr = [1000.0 * (1.0 - s) for s in np.linspace(0.01, 0.5, k)]
        pred = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.0)
        b = nv.Batch(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, pred, 1, 1)
        if cls == 1:
            b.join(jt, 0)
        side = {"key": fact["group"]} if cls == 0 else {"probe_key": fact["key"]}
        b.plan([dict(side, inputs=[fact["value"]], pred_cols=[fact["value"]], n_rows=n_fact, param=t) for t in thr])
        return b, thr

    for cls in (0, 1):
        K = args.c5_max_batch
        b, thr = plan(cls, K)
        if not b.bucketed() and K > 1:
            lo, hi, best = 0, K - 1, None
            while lo < hi:
                mid = (lo + hi + 1) // 2
                t, tt = plan(cls, mid)
                if t.bucketed():
                    if best:
                        best[0].free()
                    best, lo = (t, tt), mid
                else:
                    t.free()
                    hi = mid - 1
            if best:
                b.free()
                (b, thr), K = best, lo
        bucketed = b.bucketed()
        for _ in range(args.warmup):
            b.rerun(thr, K)
        lib.vgpu_kernel_timing(1)
        lib.vgpu_kernel_time_ms(None, 1)
        for _ in range(args.steps):
            b.rerun(thr, K)
        nl = C.c_int64()
        pass_ms[cls] = lib.vgpu_kernel_time_ms(C.byref(nl), 1) / max(nl.value, 1)
        pass_k[cls] = K
        lib.vgpu_kernel_timing(0)
        b.free()
        log(f"[bench] c5 class-{cls} pass of {K} queries (bucketed {bucketed}): k_bt_chain_feed "
            f"{pass_ms[cls]:.3f} ms avg over {nl.value} launches")
    jt.free()
    K = pass_k[0]
    kms = pass_ms[0]
    alg = n_fact * 16   # value + group, read once per pass
    achieved = alg / (kms / 1e3) / 1e9
    log(f"[bench] c5 class-0 pass: {achieved:.0f} GB/s of algorithmic bytes")

    cpu = None
    if not args.no_cpu:
        nv.load_oracle(ROOT / "oracle" / "_build" / "liboracle.so")
        m = min(args.c5_cpu_rows, n_fact)
        host = {k: v[:m].cpu().numpy() for k, v in fact_t.items()}
        dh = {k: v.cpu().numpy() for k, v in dim_t.items()}
        fact_tbl = nv.Table(host, rg_rows=1 << 20)
        dim_tbl = nv.Table({"dk": dh["key"], "attr": dh["attr"]}, rg_rows=1 << 20)

        def both():
            secs = {}
            for cls in (0, 1):
                t1 = time.perf_counter()
                f = nv.filter_node("oracle", fact_tbl.scan(), ("cmp", ">", ("col", "value"), ("dbl", 750.0)))
                if cls == 1:
                    f = nv.join_node("oracle", f, dim_tbl.scan(), "inner", [(0, 0)])
                nv.collect(nv.group_agg_node("oracle", f, ["group" if cls == 0 else "attr"],
                                             [("s", "sum", "value", 0), ("n", "n", None, 0)], "/tmp"))
                secs[cls] = time.perf_counter() - t1
            return secs

        secs, hostinfo = pinned(both)
        per_query = 0.7 * secs[0] * n_fact / m + 0.3 * secs[1] * n_fact / m
        cpu = {"value": 1.0 / per_query, "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"one class-0 and one class-1 query (25% selectivity) on the first {m:,} fact rows "
                         f"({secs[0]:.1f} s / {secs[1]:.1f} s), extrapolated to {n_fact:,} rows and the 70/30 "
                         f"mix; one thread pinned to core {hostinfo['pinned_core']} of {hostinfo['nproc']} "
                         f"({hostinfo['cpu_model']}); {ORACLE_NOTE}",
               **hostinfo}
    traffic = None
    tfile = ROOT / "profiles" / "k_bt_chain_feed_serving_1B_rows.json"
    if tfile.exists() and n_fact == 1_000_000_000:
        prof = json.loads(tfile.read_text())
        if prof.get("queries_per_pass") == K:
            traffic = prof["traffic_bytes_per_launch"]
    # a step is one request: steps = the requests served at the reported load
    return {
        "metric": METRIC, "value": bb["achieved_qps"] if bb else 0.0, "unit": "queries/s", "n_gpus": 1,
        "steps": bb["requests"] if bb else 0, "warmup": args.warmup,
        "ms_per_step": (1e3 / bb["achieved_qps"]) if bb else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (torch generators on the device)",
        "config": {"workload": f"serving loop over device-resident tables: fact {n_fact:,} rows (key Zipf(1.0) over "
                               f"{n_dim:,} dim keys, value U(0,1000), category U{{0..999}}), dim {n_dim:,} rows "
                               "(unique key, attr U{0..255}); Poisson arrivals, 70% filter(value > t) |> "
                               "group_by(category) |> summarise(s = sum(value), n = n()), 30% filter |> "
                               "inner_join(dim) |> group_by(attr) |> summarise(...); selectivity U[1%, 50%]; "
                               f"value = queries/s at p99 <= {slo:.0f} ms (continuous batching, shared scans)",
                   "fact_rows": n_fact, "dim_rows": n_dim, "max_batch": args.c5_max_batch, "slo_p99_ms": slo},
        "batched_at_slo": bb, "plain_at_slo": bp, "batched_vs_plain_at_slo": ratio, "sweep": sweep,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/k_bt_chain_feed_serving_1B_rows.md"
                     if traffic else None,
                     "kernel": f"k_bt_chain_feed<true> (the scan of a bucketed pass of {K} class-0 queries)",
                     "bytes_per_launch": alg, "avg_launch_ms": kms, "launches_timed": args.steps},
        "pass_ms": {"class0_group_by": pass_ms[0], "class1_join_group_by": pass_ms[1],
                    "queries_per_pass": {"class0": pass_k[0], "class1": pass_k[1]}},
        "cpu_baseline": cpu,
    }


def run_extras(args, nv):
    """Short runs of the other device-resident workloads after the serving
    sweep (their own figures; the serving loop's tables are released
    first): the join-then-aggregate at 50M x 1M rows and the 200M-row
    group-by at 50M and 16 key values, five timed steps each, no CPU
    baseline (each has its own in its full run)."""
    import copy
    import gc
    try:
        import torch
        gc.collect()
        torch.cuda.empty_cache()
    except Exception:   # noqa: BLE001 -- torch only holds the serving tables
        pass
    sub = copy.copy(args)
    sub.steps, sub.warmup, sub.no_cpu = 5, 1, True
    sub.fact_rows, sub.dim_rows, sub.c3_rows = 50_000_000, 1_000_000, 200_000_000
    extras = {}
    for name, fn in (("c2_join_aggregate", lambda: run_c2(sub, nv)),
                     ("c3_high_groupby", lambda: run_c3(sub, nv, 50_000_000)),
                     ("c3_low_groupby", lambda: run_c3(sub, nv, 16))):
        t = time.perf_counter()
        o = fn()
        rf = o.get("roofline") or {}
        extras[name] = {"value": o["value"], "unit": o["unit"], "ms_per_step": o["ms_per_step"],
                        "steps": o["steps"], "workload": o["config"]["workload"],
                        "roofline": {k: rf.get(k) for k in ("kernel", "achieved", "peak", "unit", "frac",
                                                            "bytes_per_launch", "avg_launch_ms", "traffic")},
                        "result_check": o.get("result_check"), "wall_s": time.perf_counter() - t}
        log(f"[bench] extra {name}: {o['value']:.4g} {o['unit']}, {o['ms_per_step']:.3f} ms/step, "
            f"roofline {rf.get('frac', 0):.3f}")
        gc.collect()
    return extras


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=["c1", "c1-vtr", "c2", "c2-nodes", "c3-high", "c3-low",
                                                             "c4", "c5", "strings"])
    ap.add_argument("--c4-queries", type=int, default=4096)
    ap.add_argument("--c4-card", type=int, default=0, help="fix every query's key cardinality (0: log-uniform)")
    ap.add_argument("--c5-rows", type=int, default=1_000_000_000)
    ap.add_argument("--c5-dim", type=int, default=10_000_000)
    ap.add_argument("--c5-max-batch", type=int, default=64)
    ap.add_argument("--c5-cpu-rows", type=int, default=100_000_000, help="fact rows of the CPU baseline's sample")
    ap.add_argument("--c5-in-flight", type=int, default=2, help="passes on the device at once (batched mode)")
    ap.add_argument("--c5-hbm-budget", type=int, default=0, help="device bytes the serving run may hold (0: all)")
    ap.add_argument("--c5-seconds", type=float, default=3.0)
    ap.add_argument("--c5-slo-ms", type=float, default=100.0)
    ap.add_argument("--c5-plain-rates", type=float, nargs="*", default=[10, 25, 40, 55, 70, 85, 100])
    ap.add_argument("--c5-batched-rates", type=float, nargs="*",
                    default=[50, 100, 200, 400, 800, 1600, 2400, 3200, 3600, 4000, 4400, 4800, 5200])
    ap.add_argument("--c3-rows", type=int, default=200_000_000)
    ap.add_argument("--vtr-rows", type=int, default=50_000_000)
    ap.add_argument("--str-rows", type=int, default=10_000_000)
    ap.add_argument("--vtr-rg", type=int, default=1_000_000)
    ap.add_argument("--vtr-path", default="/tmp/vectra_bench_c1.vtr")
    ap.add_argument("--fact-rows", type=int, default=50_000_000)
    ap.add_argument("--dim-rows", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="c5: skip the short join-aggregate and 200M-row group-by runs after the serving sweep")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="set a launch knob (vgpu_tune_set) before the run, e.g. batch.dense=1")
    ap.add_argument("--tune-reset", action="store_true",
                    help="drop the tuning table's buckets first (so --tune values apply everywhere)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy 1 GiB three times (reference kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 1 << 30, "repeats": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(see the sharding section of the design notes)", "n_gpus": args.gpus}))
        sys.exit(2)
    from vectra_amd import native as nv
    lib = nv.vgpu()
    if lib.vgpu_init(0) != 0:
        raise SystemExit("bench: " + lib.vgpu_last_error().decode())
    if args.tune_reset:
        lib.vgpu_tune_reset()
        log("[bench] tuning table dropped: built-in defaults")
    for kv in args.tune:
        name, _, val = kv.partition("=")
        if lib.vgpu_tune_set(name.encode(), int(val)) != 0:
            raise SystemExit(f"bench: unknown knob or value {kv!r}")
        log(f"[bench] knob {name} = {val}")
    if args.workload == "c1":
        out = run_c1(args, nv)
    elif args.workload == "strings":
        out = run_strings(args, nv)
    elif args.workload == "c1-vtr":
        out = run_c1_vtr(args, nv)
    elif args.workload == "c3-high":
        out = run_c3(args, nv, 50_000_000)
    elif args.workload == "c3-low":
        out = run_c3(args, nv, 16)
    elif args.workload == "c4":
        out = run_c4(args, nv)
    elif args.workload == "c2":
        out = run_c2(args, nv)
    elif args.workload == "c2-nodes":
        out = run_c2_nodes(args, nv)
    else:
        out = run_c5(args, nv)
        if not args.no_extras:
            out["other_workloads"] = run_extras(args, nv)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
