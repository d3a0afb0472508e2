"""Autotuner of the launch shapes and table loads.

A tuning-table bucket covers work by three axes: the groups it expects,
the width of its records (8-byte words of a group-by's hot + cold record,
of a join table's entry) and the rows of the launch.  For each cardinality
bucket the tuner runs the group-by the bucket stands for on the GPU (int64
keys uniform over the bucket's cardinality, key order) in two record
widths -- sum and n of one double input (narrow), and n with sum / min /
max / mean of a double and an int64 input (wide) -- at two row counts -- a
row group of 4M rows and a 50M-row column -- under every setting of the
knobs that decide its shape: LDS table slots per group (the LDS table's
inverse load factor), copies of a small LDS table, blocks and rows per
block of the LDS-privatized feed, the group count from which a large feed
is hash-partitioned, rows per block of the HBM-table feed and the HBM
table's load factor.  It times the feed kernels with HIP events and keeps
the fastest setting per (width, rows, cardinality) bucket.  Join tables:
the load factor of a 1M-key and a 10M-key build, timed over build and a
20M / 50M-row fused probe -> aggregate whose keys follow a star schema's
skewed foreign keys (Zipf(1.1) over the dimension, 10% absent); and the
rows in flight per thread and blocks per CU of that probe.

    python -m vectra_amd.autotune --out vectra_amd/tuning.txt

writes "max_groups name value [w<=W] [rows<=R]" lines that libvgpu loads
at start-up (vgpu_tune_load; vectra_amd.native loads
vectra_amd/tuning.txt, libvgpu the file named by VECTRA_GPU_TUNING).
"""
from __future__ import annotations

import argparse
import ctypes as C
import itertools
import sys
import time
from pathlib import Path

import numpy as np

from . import native as nv

BUCKETS = [(64, 16), (1024, 256), (16384, 4096), (1 << 20, 65536), (1 << 40, 4_000_000)]
ROWS = [(8_000_000, 4_000_000), (None, 50_000_000)]   # (rows bound of the bucket, rows swept)
ANY = 1 << 40


def _feed_ms(run, reps: int = 3) -> float:
    lib = nv.vgpu()
    times = []
    for _ in range(reps):
        lib.vgpu_synchronize()
        lib.vgpu_kernel_timing(1)
        lib.vgpu_kernel_time_ms(None, 1)
        run()
        lib.vgpu_synchronize()
        n = C.c_int64()
        times.append(lib.vgpu_kernel_time_ms(C.byref(n), 1))
        lib.vgpu_kernel_timing(0)
    return float(np.median(times))


# two record widths: one sum and a count, and n with sum / min / max / mean
# of a double and an int64 input; each gets its own best setting
SHAPES = {"narrow": [("sum", 0, 0), ("n", -1, 0)],
          "wide": [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("mean", 0, 0),
                   ("sum", 1, 0), ("min", 1, 0), ("max", 1, 0), ("mean", 1, 0)]}


def shape_types(shape):
    return [nv.VEC_DOUBLE] if shape == "narrow" else [nv.VEC_DOUBLE, nv.VEC_INT64]


def record_words(shape) -> int:
    gb = nv.GroupBy(nv.VEC_INT64, shape_types(shape), SHAPES[shape], nv.VGPU_ORDER_KEY, 0)
    w = nv.vgpu().vgpu_groupby_record_words(gb.h)
    gb.free()
    return int(w)


def tune_groupby(card: int, rows: int, grid: dict, log) -> dict:
    """{shape: (best knobs, ms)} for one cardinality and row count"""
    import torch
    lib = nv.vgpu()
    g = torch.Generator(device="cuda")
    g.manual_seed(card)
    key = torch.randint(0, card, (rows,), dtype=torch.int64, device="cuda", generator=g)
    val = torch.randn(rows, dtype=torch.float64, device="cuda", generator=g)
    ival = torch.randint(-10**9, 10**9, (rows,), dtype=torch.int64, device="cuda", generator=g)
    torch.cuda.synchronize()
    kc = nv.DeviceColumn.wrap(key.data_ptr(), rows, nv.VEC_INT64, keep=key)
    vc = nv.DeviceColumn.wrap(val.data_ptr(), rows, nv.VEC_DOUBLE, keep=val)
    ic = nv.DeviceColumn.wrap(ival.data_ptr(), rows, nv.VEC_INT64, keep=ival)

    def runner(shape):
        cols = [vc] if shape == "narrow" else [vc, ic]

        def run():
            gb = nv.GroupBy(nv.VEC_INT64, shape_types(shape), SHAPES[shape], nv.VGPU_ORDER_KEY, 0)
            gb.feed(kc, cols, rows)
            gb.finish()
            gb.free()
        return run

    runs = {s: runner(s) for s in SHAPES}
    names = list(grid)
    results = []
    for values in itertools.product(*(grid[n] for n in names)):
        for n, v in zip(names, values):
            lib.vgpu_tune_set(n.encode(), v)
        ms = {s: _feed_ms(r) for s, r in runs.items()}
        results.append((dict(zip(names, values)), ms))
        log(f"  card {card:>9,} rows {rows:>11,}: {dict(zip(names, values))} -> "
            + ", ".join(f"{s} {t:.3f} ms" for s, t in ms.items()))
    lib.vgpu_tune_reset()
    del key, val, ival
    torch.cuda.empty_cache()
    return {s: min(((k, ms[s]) for k, ms in results), key=lambda r: r[1]) for s in SHAPES}


def star_probe_keys(dk, n_fact: int, g, s: float = 1.1, absent: float = 0.1):
    """probe keys of a star schema's fact: Zipf(s) over the dimension's keys
    in a random rank order (inverse CDF by search), a fraction `absent`
    replaced by keys the dimension does not hold (dk * 2 + 1 never equals an
    even dimension key: the callers' keys are all even)"""
    import torch
    n_dim = dk.numel()
    w = torch.arange(1, n_dim + 1, dtype=torch.float64, device="cuda").pow(-s)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    ranks = torch.searchsorted(cdf, torch.rand(n_fact, dtype=torch.float64, device="cuda", generator=g))
    ranks.clamp_(max=n_dim - 1)
    order = torch.randperm(n_dim, device="cuda", generator=g)
    fk = dk[order[ranks]]
    miss = torch.rand(n_fact, device="cuda", generator=g) < absent
    fk[miss] = fk[miss] * 2 + 1
    del w, cdf, ranks, order, miss
    return fk


def tune_join_load(log) -> list[tuple[int | None, int, float]]:
    """(rows bound, best join.load_pct, ms) for a 1M-key and a 10M-key build"""
    import torch
    lib = nv.vgpu()
    out = []
    for bound, n_dim, n_fact in ((2_000_000, 1_000_000, 20_000_000), (None, 10_000_000, 50_000_000)):
        g = torch.Generator(device="cuda")
        g.manual_seed(n_dim)
        dk = torch.randperm(n_dim, device="cuda", generator=g).to(torch.int64) * 14 + 6
        attr = torch.randint(0, 256, (n_dim,), dtype=torch.int64, device="cuda", generator=g)
        fk = star_probe_keys(dk, n_fact, g)
        v = torch.rand(n_fact, dtype=torch.float64, device="cuda", generator=g)
        torch.cuda.synchronize()
        wrap = lambda t, ty: nv.DeviceColumn.wrap(t.data_ptr(), t.numel(), ty, keep=t)
        dkc, atc, fkc, vc = wrap(dk, nv.VEC_INT64), wrap(attr, nv.VEC_INT64), wrap(fk, nv.VEC_INT64), wrap(v, nv.VEC_DOUBLE)
        best = None
        for pct in (25, 30, 50, 70):
            lib.vgpu_tune_set(b"join.load_pct", pct)
            times = []
            for _ in range(3):
                lib.vgpu_synchronize()
                lib.vgpu_timer_start()
                jt = nv.JoinTable(dkc, [atc])
                gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, 257)
                gb.feed_join(n_fact, [fkc, vc], [dkc, atc], [(0, 1)], key=(1, 1), join=jt,
                             join_kind=nv.JOIN_INNER, probe_key_col=0)
                gb.finish()
                times.append(lib.vgpu_timer_stop())
                gb.free()
                jt.free()
            ms = float(np.median(times))
            log(f"  join build {n_dim:,} keys + probe {n_fact:,}: join.load_pct={pct} -> {ms:.3f} ms")
            if best is None or ms < best[1]:
                best = (pct, ms)
        lib.vgpu_tune_reset()
        out.append((bound, best[0], best[1]))
        del dk, attr, fk, v
        torch.cuda.empty_cache()
    return out


def tune_star(log) -> dict:
    import torch
    lib = nv.vgpu()
    n_dim, n_fact = 1_000_000, 20_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    dk = torch.randperm(n_dim, device="cuda", generator=g).to(torch.int64) * 6 + 2
    attr = torch.randint(0, 100, (n_dim,), dtype=torch.int64, device="cuda", generator=g)
    w = torch.rand(n_dim, dtype=torch.float64, device="cuda", generator=g)
    fk = star_probe_keys(dk, n_fact, g)
    v = torch.rand(n_fact, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()
    wrap = lambda t, ty: nv.DeviceColumn.wrap(t.data_ptr(), t.numel(), ty, keep=t)
    dkc, atc, wc = wrap(dk, nv.VEC_INT64), wrap(attr, nv.VEC_INT64), wrap(w, nv.VEC_DOUBLE)
    fkc, vc = wrap(fk, nv.VEC_INT64), wrap(v, nv.VEC_DOUBLE)
    jt = nv.JoinTable(dkc, [atc, wc])

    def run():
        gb = nv.GroupBy(nv.VEC_INT64, [nv.VEC_DOUBLE], [("sum", 0, 0), ("n", -1, 0)], nv.VGPU_ORDER_KEY, 101)
        gb.feed_join(n_fact, [fkc, vc], [dkc, atc, wc], [(0, 1, 1, 2)], key=(1, 1), join=jt,
                     join_kind=nv.JOIN_INNER, probe_key_col=0)
        gb.finish()
        gb.free()

    best = None
    for items, bpc in itertools.product((1, 2, 4, 8), (2, 4, 8)):
        lib.vgpu_tune_set(b"gb.star.items", items)
        lib.vgpu_tune_set(b"gb.star.blocks_per_cu", bpc)
        ms = _feed_ms(run)
        log(f"  star: gb.star.items={items} gb.star.blocks_per_cu={bpc} -> {ms:.3f} ms")
        if best is None or ms < best[0]:
            best = (ms, items, bpc)
    lib.vgpu_tune_reset()
    jt.free()
    return {"ms": best[0], "knobs": {"gb.star.items": best[1], "gb.star.blocks_per_cu": best[2]}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "tuning.txt"))
    ap.add_argument("--rows", type=int, default=50_000_000)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    lib = nv.vgpu()
    lib.vgpu_tune_reset()
    if lib.vgpu_init(0) != 0:
        raise SystemExit("autotune: " + lib.vgpu_last_error().decode())
    lines = []
    t0 = time.time()
    words = {sh: record_words(sh) for sh in SHAPES}
    narrow_w = words["narrow"]
    for max_groups, card in BUCKETS:
        if card <= 256:
            grid = {"gb.lds.slots_per_group": [2, 4], "gb.lds.copies": [1, 2, 4],
                    "gb.lds.max_blocks": [1024, 2048], "gb.lds.rows_per_block": [4096, 32768]}
        elif card <= 4096:
            # past the LDS table's reach the choice is HBM atomics or the partitioned path
            grid = {"gb.part.min_groups": [card // 2, 1 << 40], "gb.lds.slots_per_group": [2, 4],
                    "gb.lds.max_blocks": [512, 1024, 2048], "gb.lds.rows_per_block": [8192, 32768]}
        else:
            grid = {"gb.part.min_groups": [16384, 1 << 40], "gb.glb.rows_per_block": [512, 1024, 4096],
                    "gb.glb.load_pct": [35, 50, 70]}
        for bound, rows in ROWS:
            if rows > args.rows:
                rows = args.rows
            best = tune_groupby(card, rows, grid, log)
            for sh, (knobs, ms) in best.items():
                tail = (f" w<={narrow_w}" if sh == "narrow" else "") + (f" rows<={bound}" if bound else "")
                log(f"bucket <= {max_groups:,} groups, {sh} ({words[sh]} words), {rows:,} rows: {knobs} ({ms:.3f} ms)")
                for name, value in knobs.items():
                    lines.append(f"{max_groups} {name} {value}{tail}   # {card:,} keys, {sh}, {rows:,} rows: {ms:.3f} ms")
    for bound, pct, ms in tune_join_load(log):
        tail = f" rows<={bound}" if bound else ""
        lines.append(f"{ANY} join.load_pct {pct}{tail}   # build + probe: {ms:.3f} ms")
    r = tune_star(log)
    for name, value in r["knobs"].items():
        lines.append(f"128 {name} {value}   # join -> group_by(attr): {r['ms']:.3f} ms")
    head = [f"# launch shapes and table loads per (width, rows, cardinality) bucket, measured by "
            f"vectra_amd/autotune.py",
            f"# command: python -m vectra_amd.autotune --out {Path(args.out).name} --rows {args.rows}",
            f"# {time.strftime('%Y-%m-%d')}, {time.time() - t0:.0f} s of sweeps; record words: narrow {narrow_w}, "
            f"wide {words['wide']}", "# max_groups name value [w<=W] [rows<=R]"]
    Path(args.out).write_text("\n".join(head + lines) + "\n")
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
