"""Autotuner of the group-by launch shapes.

For each cardinality bucket it runs the group-by the bucket stands for on
the GPU (50M rows, int64 keys uniform over the bucket's cardinality, key
order) with two record widths -- sum and n of one double input, and n with
sum / min / max / mean of a double and an int64 input -- under every
setting of the knobs that decide its shape -- LDS table slots per group
(the table's inverse load factor), copies of a small LDS table, blocks and
rows per block of the LDS-privatized feed, and the group count from which
a large feed is hash-partitioned -- times the feed kernels with HIP
events, and keeps per bucket the setting whose worse relative slowdown
over the two widths is smallest.  The
fused join -> aggregate shape (a unique-key dimension of 1M rows, 20M
probe rows) picks the rows in flight per thread of its probe.

    python -m vectra_amd.autotune --out vectra_amd/tuning.txt

writes "max_groups name value" lines that libvgpu loads at start-up
(vgpu_tune_load; vectra_amd.native loads vectra_amd/tuning.txt, libvgpu
the file named by VECTRA_GPU_TUNING).
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


# two record widths: one sum and a count (3 words), and n with sum / min /
# max / mean of a double and an int64 input (about 14 words); a setting is
# scored by its worse relative slowdown over the two
SHAPES = {"narrow": [("sum", 0, 0), ("n", -1, 0)],
          "wide": [("n", -1, 0), ("sum", 0, 0), ("min", 0, 0), ("max", 0, 0), ("mean", 0, 0),
                   ("sum", 1, 0), ("min", 1, 0), ("max", 1, 0), ("mean", 1, 0)]}


def tune_groupby(card: int, rows: int, grid: dict, log) -> dict:
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
        aggs = SHAPES[shape]
        types, cols = ([nv.VEC_DOUBLE], [vc]) if shape == "narrow" else ([nv.VEC_DOUBLE, nv.VEC_INT64], [vc, ic])

        def run():
            gb = nv.GroupBy(nv.VEC_INT64, types, aggs, nv.VGPU_ORDER_KEY, 0)
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
        log(f"  card {card:>9,}: {dict(zip(names, values))} -> "
            + ", ".join(f"{s} {t:.3f} ms" for s, t in ms.items()))
    lib.vgpu_tune_reset()
    best_of = {s: min(r[1][s] for r in results) for s in SHAPES}
    knobs, ms = min(results, key=lambda r: max(r[1][s] / best_of[s] for s in SHAPES))
    del key, val, ival
    torch.cuda.empty_cache()
    return {"ms": ms, "knobs": knobs}


def tune_star(log) -> dict:
    import torch
    lib = nv.vgpu()
    n_dim, n_fact = 1_000_000, 20_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    dk = torch.randperm(n_dim, device="cuda", generator=g).to(torch.int64) * 3 + 1
    attr = torch.randint(0, 100, (n_dim,), dtype=torch.int64, device="cuda", generator=g)
    w = torch.rand(n_dim, dtype=torch.float64, device="cuda", generator=g)
    fk = dk[torch.randint(0, n_dim, (n_fact,), device="cuda", generator=g)]
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
    for items in (1, 2, 4, 8):
        lib.vgpu_tune_set(b"gb.star.items", items)
        ms = _feed_ms(run)
        log(f"  star: gb.star.items={items} -> {ms:.3f} ms")
        if best is None or ms < best[0]:
            best = (ms, items)
    lib.vgpu_tune_reset()
    jt.free()
    return {"ms": best[0], "knobs": {"gb.star.items": best[1]}}


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
    for max_groups, card in BUCKETS:
        if card <= 256:
            grid = {"gb.lds.slots_per_group": [2, 4], "gb.lds.copies": [1, 2, 4],
                    "gb.lds.max_blocks": [1024, 2048], "gb.lds.rows_per_block": [4096, 32768]}
        elif card <= 4096:
            # past the LDS table's reach the choice is HBM atomics or the partitioned path
            grid = {"gb.part.min_groups": [card // 2, 1 << 40], "gb.lds.slots_per_group": [2, 4],
                    "gb.lds.max_blocks": [512, 1024, 2048], "gb.lds.rows_per_block": [8192, 32768]}
        else:
            grid = {"gb.part.min_groups": [16384, 1 << 40], "gb.glb.rows_per_block": [512, 1024, 4096]}
        r = tune_groupby(card, args.rows, grid, log)
        log(f"bucket <= {max_groups:,} groups: best {r['knobs']} ({r['ms']})")
        for name, value in r["knobs"].items():
            t = ", ".join(f"{s} {v:.3f} ms" for s, v in r["ms"].items())
            lines.append(f"{max_groups} {name} {value}   # {card:,} keys: {t}")
    r = tune_star(log)
    lines.append(f"128 gb.star.items {r['knobs']['gb.star.items']}   # join -> group_by(attr): {r['ms']:.3f} ms")
    head = [f"# group-by launch shapes per cardinality bucket, measured by vectra_amd/autotune.py",
            f"# command: python -m vectra_amd.autotune --out {Path(args.out).name} --rows {args.rows}",
            f"# {time.strftime('%Y-%m-%d')}, {time.time() - t0:.0f} s of sweeps", "# max_groups name value"]
    Path(args.out).write_text("\n".join(head + lines) + "\n")
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
