"""Builds the native libraries of vectra_amd in-tree.

    libvgpu.so        hand-written gfx950 kernels behind the C ABI of include/vgpu.h
    libvx_core.so     engine plumbing (arrays, batches, errors, collect)
    libvectra_gpu.so  the drop-in C11 operator nodes of include/vectra_abi.h

Everything is compiled from source on every call (hipcc cross-compiles for
gfx950 without a GPU); outputs land in vectra_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have
    # given silently wrong results on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 atomic add as one hardware instruction instead of a CAS loop
    "-munsafe-fp-atomics",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
C_FLAGS = ["-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
           f"-I{INCLUDE}", f"-I{HOST}"]

HIP_SOURCES = ["vgpu_runtime.hip", "vgpu_filter.hip", "vgpu_groupby.hip", "vgpu_join.hip", "vgpu_batch.hip",
               "vgpu_strdict.hip"]
NODE_SOURCES = ["gpu_common.c", "devscan.c", "filter_node.c", "group_agg_node.c", "join_node.c", "serve.c",
                "explain.c"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(map(str, cmd)), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    err = "\n".join(line for line in res.stderr.splitlines()
                    if "is not a recognized feature" not in line)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(map(str, cmd))}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> dict[str, Path]:
    LIB.mkdir(exist_ok=True)
    obj_dir = LIB / "obj"
    obj_dir.mkdir(exist_ok=True)
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src in HIP_SOURCES:
            obj = obj_dir / (Path(src).stem + ".o")
            objs.append(obj)
            jobs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], verbose))
        for j in jobs:
            j.result()
    libvgpu = LIB / "libvgpu.so"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(libvgpu)], verbose)

    libcore, libnodes = build_host(LIB, verbose=verbose)
    return {"vgpu": libvgpu, "core": libcore, "nodes": libnodes}


def build_host(out: Path, sanitize: bool = False, verbose: bool = False) -> tuple[Path, Path]:
    """The C11 libraries (libvx_core, libvectra_gpu) into out, which must
    hold libvgpu.so (or a link to it).  sanitize: -O1 with AddressSanitizer
    and UndefinedBehaviorSanitizer, for CPU test runs (tests/run_sanitized.py)."""
    flags = list(C_FLAGS)
    if sanitize:
        flags = [f for f in flags if f != "-O2"] + ["-O1", "-g", "-fno-omit-frame-pointer",
                                                   "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    libcore = out / "libvx_core.so"
    _run([CC, *flags, "-shared", "-Wl,-soname,libvx_core.so", str(HOST / "vx_core.c"), str(HOST / "vx_vtr.c"), "-lm",
          "-o", str(libcore)], verbose)
    libnodes = out / "libvectra_gpu.so"
    _run([CC, *flags, "-shared", *[str(HOST / s) for s in NODE_SOURCES],
          f"-L{out}", "-lvgpu", "-lvx_core", "-lpthread", "-Wl,-rpath,$ORIGIN", "-o", str(libnodes)], verbose)
    return libcore, libnodes


if __name__ == "__main__":
    out = build(verbose="-v" in sys.argv)
    for k, v in out.items():
        print(k, v)
