"""chasqui_amd — MI355X-native per-RK-stage MHD pipeline (HIP kernels for gfx950 behind a
C ABI, include/chasqui_amd.h), with a C++ host layer that keeps the reference's
`block.key=value` decks, problem setups and VTK/.hst outputs.

The native library is required: nothing here falls back to a CPU or Python path.
"""
from ._capi import LIB_PATH, NativeLibraryMissing, lib  # noqa: F401
from .mesh import Mesh, make_config, make_mesh_desc  # noqa: F401
from .batch import Batch  # noqa: F401

__all__ = ["Mesh", "Batch", "make_config", "make_mesh_desc", "lib", "LIB_PATH", "NativeLibraryMissing"]
