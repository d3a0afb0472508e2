"""vectra_amd: MI355X-native (gfx950) hot path of the vectra columnar engine.

Filter -> selection vector, hash group-by aggregation and hash join build /
probe, as drop-in C11 operator nodes (include/vectra_abi.h) over hand-written
HIP kernels (include/vgpu.h).  `vectra_amd.native` loads the built libraries.
"""
from . import native  # noqa: F401

__all__ = ["native"]
