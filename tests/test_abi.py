"""The C-ABI libraries load and export every function include/*.h declares;
struct layouts seen from ctypes match the C compiler's (CPU only, no GPU
calls)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", re.M)


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
    names = set()
    for m in DECL.finditer(text):
        name = m.group(1)
        line = text[m.start():text.find("\n", m.start())]
        if line.startswith(("typedef", "#", "struct")) or name in ("sizeof",):
            continue
        names.add(name)
    return names


def test_vgpu_exports(native):
    lib = C.CDLL(str(ROOT / "vectra_amd" / "lib" / "libvgpu.so"))
    names = declared("vgpu.h")
    assert "vgpu_filter" in names and "vgpu_groupby_feed_join" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_node_exports(native):
    lib = C.CDLL(str(ROOT / "vectra_amd" / "lib" / "libvectra_gpu.so"))
    names = declared("vectra_abi.h")
    assert {"filter_node_create", "group_agg_node_create", "join_node_create"} <= names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_serve_exports(native):
    lib = C.CDLL(str(ROOT / "vectra_amd" / "lib" / "libvectra_gpu.so"))
    names = declared("vx_serve.h")
    assert {"vx_serve_run", "vx_result_free"} <= names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_serve_struct_layouts(native, tmp_path):
    """the serving loop's structs as ctypes sees them match the C compiler"""
    from vectra_amd import native as nv
    src = tmp_path / "l.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "vx_serve.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(vx_request), sizeof(vx_result), sizeof(vx_serve_cfg), sizeof(vx_serve_stats));
  printf("%zu %zu %zu %zu\n", offsetof(vx_result, n_groups), offsetof(vx_serve_cfg, n_dim),
         offsetof(vx_serve_cfg, state_budget), offsetof(vx_serve_stats, resident_bytes));
  printf("%zu %zu\n", sizeof(vgpu_query), offsetof(vgpu_query, n_rows));
  printf("%zu %zu %zu %zu\n", offsetof(vx_serve_cfg, in_flight), offsetof(vx_serve_cfg, hbm_budget),
         offsetof(vx_serve_stats, max_in_flight), offsetof(vx_serve_stats, hbm_budget));
  return 0;
}
""")
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [list(map(int, line.split())) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                   check=True).stdout.split("\n") if line]
    assert got[0] == [C.sizeof(nv.VxRequest), C.sizeof(nv.VxResult), C.sizeof(nv.VxServeCfg),
                      C.sizeof(nv.VxServeStats)]
    assert got[1] == [nv.VxResult.n_groups.offset, nv.VxServeCfg.n_dim.offset, nv.VxServeCfg.state_budget.offset,
                      nv.VxServeStats.resident_bytes.offset]
    assert got[2] == [C.sizeof(nv.VgpuQuery), nv.VgpuQuery.n_rows.offset]
    assert got[3] == [nv.VxServeCfg.in_flight.offset, nv.VxServeCfg.hbm_budget.offset,
                      nv.VxServeStats.max_in_flight.offset, nv.VxServeStats.hbm_budget.offset]


def test_struct_layouts(native, tmp_path):
    """offsets of the fields the optimizer / explain read, compiled by gcc"""
    src = tmp_path / "lay.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "vectra_abi.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(VecArray), sizeof(VecBatch), sizeof(VecNode), sizeof(VecExpr));
  printf("%zu %zu %zu\n", offsetof(FilterNode, child), offsetof(FilterNode, predicate), sizeof(AggSpec));
  printf("%zu %zu %zu %zu %zu\n", offsetof(GroupAggNode, n_keys), offsetof(GroupAggNode, key_names),
         offsetof(GroupAggNode, n_aggs), offsetof(GroupAggNode, agg_specs), offsetof(GroupAggNode, use_sorted));
  printf("%zu %zu %zu %zu\n", offsetof(JoinNode, left), offsetof(JoinNode, right), offsetof(JoinNode, kind),
         offsetof(JoinNode, n_keys));
  return 0;
}''')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    nv = native
    assert lines[0].split() == [str(C.sizeof(t)) for t in (nv.VecArray, nv.VecBatch, nv.VecNode, nv.VecExpr)]
    # reference prefixes on LP64: VecNode is 2 pointers + schema (4 words) + kind = 56 bytes
    assert lines[1].split()[:2] == ["56", "64"]
    assert lines[2].split() == ["64", "72", "80", "88", "100"]
    assert lines[3].split() == ["56", "64", "72", "76"]


def test_vtr_struct_layouts(native, tmp_path):
    """the .vtr handle and scan node layouts the GPU nodes read from a scan
    child (src/vtr1.h:7-37, src/scan.h:9-16 on LP64, worked out by hand:
    Vtr1ColStat = u8 + 16-byte union at 8; Vtr1Header = u16, VecSchema at 8
    (32 bytes), u32 at 40; Vtr1File = FILE *, header at 8, rowgroups at 56;
    ScanNode = VecNode (56), file, col_mask, next_rg at 72, predicate at 80,
    pred_borrowed at 88)"""
    src = tmp_path / "v.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "vectra_abi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(Vtr1ColStat), offsetof(Vtr1ColStat, i64.min), sizeof(Vtr1RowGroup),
         offsetof(Vtr1RowGroup, col_stats), sizeof(Vtr1Header));
  printf("%zu %zu %zu %zu\n", offsetof(Vtr1Header, schema), offsetof(Vtr1Header, n_rowgroups),
         offsetof(Vtr1File, header), offsetof(Vtr1File, rowgroups));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(ScanNode, file), offsetof(ScanNode, col_mask),
         offsetof(ScanNode, next_rg), offsetof(ScanNode, predicate), offsetof(ScanNode, pred_borrowed),
         sizeof(ScanNode));
  return 0;
}''')
    exe = tmp_path / "v"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert lines[0].split() == ["24", "8", "24", "16", "48"]
    assert lines[1].split() == ["8", "40", "8", "56"]
    assert lines[2].split() == ["56", "64", "72", "80", "88", "96"]
