"""The C-ABI library loads and exports every entry point include/gdtk_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(gdtk_hip_\w+)\s*\(", txt))
    return names


def test_library_exports_header_symbols():
    from gdtk_amd import hip
    assert os.path.exists(hip.lib_path()), "libgdtkhip.so not built"
    lib = ctypes.CDLL(hip.lib_path())
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(hip.EXPORTS) <= declared


def test_no_gpu_is_an_error_not_a_fallback():
    """Without a GPU the library refuses to create a simulation (no CPU fallback)."""
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        return
    from gdtk_amd import HipError, Simulation
    try:
        Simulation(2)
    except HipError as e:
        assert e.code in (-2,)
    else:
        raise AssertionError("simulation created without a GPU")
