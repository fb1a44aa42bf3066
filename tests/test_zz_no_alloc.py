"""The allocation-free chemistry step (gdtk_hip_chem_step_device_noalloc) against the
allocating one (gdtk_hip_chem_step_device), through the C ABI, on the same inputs: every
output equal bit for bit (same kernels, same per-cell arithmetic), on two calls in a row
with the second smaller than the first, the statistics of each call read back, and a cell
count beyond the fixed work array refused without running."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind():
    from gdtk_amd import hip
    L = hip.lib()
    L.gdtk_hip_chem_step_device_noalloc.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_double] + \
        [C.c_void_p] * 7
    L.gdtk_hip_chem_noalloc_max_cells.argtypes = [C.c_int]
    return L


def _run(L, ch, cells, noalloc, dt=1.0e-7):
    import torch
    dev = torch.device("cuda:0")
    n = cells["rho"].size
    t = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in cells.items()}
    dtc = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    out = {k: torch.full_like(t[k], np.nan) for k in ("T", "u", "p", "massf")}
    dto = torch.full_like(dtc, np.nan)
    st = torch.full((n,), 7, dtype=torch.int32, device=dev)
    args = (ch.h, n, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(),
            t["massf"].data_ptr(), dtc.data_ptr(), dt, out["T"].data_ptr(), out["u"].data_ptr(),
            out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(),
            torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = L.gdtk_hip_chem_step_device_noalloc(*args)
    else:
        rc = L.gdtk_hip_chem_step_device(*args)
    assert rc == 0, (rc, L.gdtk_hip_last_error())
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["dt_chem"] = dto.cpu().numpy()
    res["status"] = st.cpu().numpy()
    res["stats"] = ch.stats()
    res["substeps"] = ch.substeps(n)
    return res


def test_noalloc_step_equals_allocating_step():
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    L = _bind()
    ch = Chemistry(gas_record(gdata.load("tpg-air-5sp")), mech_record(gdata.load("park-air-5sp-1T")))
    nmax = L.gdtk_hip_chem_noalloc_max_cells(ch.h)
    print("noalloc: at most", nmax, "cells per call for 5-species air", flush=True)
    assert nmax >= 262144
    for n, seed in ((40000, 3), (9000, 4)):
        cells = air_cells(n, seed=seed)
        a = _run(L, ch, cells, noalloc=False)
        b = _run(L, ch, cells, noalloc=True)
        for k in ("T", "u", "p", "massf", "dt_chem", "status", "substeps"):
            assert np.array_equal(a[k], b[k], equal_nan=True), (n, k)
        for k in ("substeps_total", "iters_total", "substeps_max", "attempts_total", "cells"):
            assert a["stats"][k] == b["stats"][k], (n, k)
        assert np.all(np.isfinite(b["T"])) and b["stats"]["substeps_total"] > n
        print(f"noalloc: {n} cells equal, {b['stats']['substeps_total']} substeps, "
              f"{int((b['status'] != 0).sum())} failed cells", flush=True)
    # beyond the fixed array: refused before anything runs (the arrays are never touched)
    import torch
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda:0").data_ptr()
    rc = L.gdtk_hip_chem_step_device_noalloc(ch.h, nmax + 1, *([dummy] * 6), 1.0e-7, *([dummy] * 6), None)
    assert rc == -1, rc
    assert b"work space" in L.gdtk_hip_last_error()
