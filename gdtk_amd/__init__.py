"""gdtk_amd: MI355X-native structured-block explicit update of Eilmer (gdtk-uq/gdtk).

The compute path is the HIP library libgdtkhip.so (built in-tree from gdtk_amd/csrc by
gdtk_amd.build).  Python here is only a thin ctypes host over its C ABI
(include/gdtk_hip.h); there is no CPU fallback: if the library is missing or no GPU is
present, the calls fail loudly.
"""
from .hip import HipError, Simulation, lib_path  # noqa: F401
