"""Builds the oracle (test infrastructure) into oracle/_build/liboracle.so.

The oracle is the C restatement of the reference algorithms in oracle.c; it
links the engine plumbing of vectra_amd (libvx_core.so) so that oracle nodes
and GPU nodes share one error channel and can be stacked in one tree.

The reference engine itself cannot be compiled here: every source file of
its hot path includes <R.h> (src/error.h:4, src/agg_ops.c:4), and R is not
installed, so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "_build"
CC = os.environ.get("CC", "gcc")


def build(verbose: bool = False, out: Path = OUT, core_dir: Path | None = None, sanitize: bool = False) -> Path:
    """liboracle.so into out, linked to the libvx_core.so in core_dir
    (default vectra_amd/lib); sanitize adds AddressSanitizer and
    UndefinedBehaviorSanitizer at -O1 (tests/run_sanitized.py)."""
    out.mkdir(parents=True, exist_ok=True)
    lib = out / "liboracle.so"
    core_dir = core_dir or ROOT / "vectra_amd" / "lib"
    opt = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined"] if sanitize else ["-O2"]
    rpath = "-Wl,-rpath," + (str(core_dir) if sanitize else "$ORIGIN/../../vectra_amd/lib")
    cmd = [CC, "-std=c11", *opt, "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'vectra_amd' / 'host'}",
           str(HERE / "oracle.c"), f"-L{core_dir}", "-lvx_core", "-lm", rpath, "-o", str(lib)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
