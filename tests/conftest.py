import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

ORACLE_LIB = Path(os.environ.get("VECTRA_ORACLE_LIB", ROOT / "oracle" / "_build" / "liboracle.so"))
NATIVE = [ROOT / "vectra_amd" / "lib" / n for n in ("libvgpu.so", "libvx_core.so", "libvectra_gpu.so")]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    # load what is built; build only when something is missing
    if not all(p.exists() for p in NATIVE):
        from vectra_amd import build as b
        b.build()
    if not ORACLE_LIB.exists():
        sys.path.insert(0, str(ROOT / "oracle"))
        import importlib
        ob = importlib.import_module("build")
        ob.build()


@pytest.fixture(scope="session")
def native():
    _ensure_built()
    from vectra_amd import native as nv
    nv.core()
    return nv


@pytest.fixture(scope="session")
def oracle(native):
    return native.load_oracle(ORACLE_LIB)


@pytest.fixture(scope="session")
def gpu(native):
    lib = native.vgpu()
    dev = int(os.environ.get("VECTRA_GPU_DEVICE", "0"))
    if lib.vgpu_init(dev) != 0:
        pytest.fail("GPU tests need a gfx950 device: " + lib.vgpu_last_error().decode())
    native.nodes()
    return native
