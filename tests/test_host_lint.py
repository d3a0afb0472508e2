"""Every host C source (the nodes, the harness core, the oracle) passes
clang's diagnostics that gcc's -Wall lacks, as errors: a function whose
every path calls itself (an -O2 build turns it into a loop that never
returns), and self-assignment.  CPU only; skipped where clang is absent."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CLANG = shutil.which("clang") or ("/opt/rocm/llvm/bin/clang" if Path("/opt/rocm/llvm/bin/clang").exists() else None)
SOURCES = sorted((ROOT / "vectra_amd" / "host").glob("*.c")) + [ROOT / "oracle" / "oracle.c"]


@pytest.mark.skipif(CLANG is None, reason="clang not found")
@pytest.mark.parametrize("src", SOURCES, ids=lambda p: p.name)
def test_host_source_clean(src):
    cmd = [CLANG, "-std=c11", "-fsyntax-only", "-Wall", "-Wno-unused-function",
           "-Werror=infinite-recursion", "-Werror=self-assign", "-Werror=sometimes-uninitialized",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'vectra_amd' / 'host'}", f"-I{ROOT / 'oracle'}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
