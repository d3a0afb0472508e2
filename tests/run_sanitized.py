"""Runs the CPU test suite (pytest -m "not gpu") against copies of the C11
libraries -- libvx_core, libvectra_gpu and the oracle -- built at -O1 with
AddressSanitizer and UndefinedBehaviorSanitizer, the way the reference's CI
runs its C code under ASan and UBSan (.github/workflows/sanitizers.yml).
CPU only: libvgpu.so (the HIP kernels) is linked as built, and no test that
touches a GPU runs.

    python tests/run_sanitized.py [extra pytest args]

Builds into vectra_amd/lib/sanitized/ and oracle/_build/sanitized/, preloads
the gcc sanitizer runtimes into the Python process, and exits with pytest's
status; a sanitizer report aborts the run."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def main() -> int:
    from vectra_amd import build as vb
    lib = ROOT / "vectra_amd" / "lib"
    if not (lib / "libvgpu.so").exists():
        vb.build()
    out = lib / "sanitized"
    out.mkdir(exist_ok=True)
    link = out / "libvgpu.so"
    if link.is_symlink() or link.exists():
        link.unlink()
    link.symlink_to(lib / "libvgpu.so")
    vb.build_host(out, sanitize=True)
    sys.path.insert(0, str(ROOT / "oracle"))
    import importlib
    oracle_lib = importlib.import_module("build").build(out=ROOT / "oracle" / "_build" / "sanitized", core_dir=out,
                                                          sanitize=True)
    runtimes = [subprocess.run(["gcc", f"-print-file-name={n}"], capture_output=True, text=True,
                               check=True).stdout.strip() for n in ("libasan.so", "libubsan.so")]
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": ":".join(runtimes),
        # Python itself is not instrumented and keeps memory until exit
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "VECTRA_AMD_LIB_DIR": str(out),
        "VECTRA_ORACLE_LIB": str(oracle_lib),
    })
    # --capture=sys: a sanitizer report written to fd 2 reaches the terminal before the abort
    cmd = [sys.executable, "-m", "pytest", str(ROOT / "tests"), "-q", "-m", "not gpu", "-p", "no:cacheprovider", "--capture=sys",
           *sys.argv[1:]]
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


if __name__ == "__main__":
    sys.exit(main())
