"""The two-temperature air reactor of the C ABI (gdtk_hip_tt_*): TwoTemperatureAirKinetics
(src/kinetics/two_temperature_air_kinetics.d) on the 5-species neutral TwoTemperatureAir gas
(src/gas/two_temperature_air.d), applied per cell as FVCell.thermochemical_increment does
(src/eilmer/fvcell.d:858-972).

The gas state of a cell is (rho, p, T, Tv, u, uv, massf): Tv and uv are the reference's
T_modes[0] and u_modes[0].  The reactor is built from the same gas record as the thermally
perfect gas (masses and CEA curves), a mechanism record whose forward rates may be 'Park'
rate constants, and per-species relaxation data (the molecules and their Millikan-White A).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .chem import MAX_SPECIES, Mechanism, TPGGas, _arr, _check

# src/kinetics/two_temperature_air_kinetics.d:714-720 (Gnoffo et al. Table 1) and
# src/gas/two_temperature_air.d:25 (the species with a vibrational mode)
A_MW = {"N2": 220.0, "O2": 129.0, "NO": 168.0}


class TTParams(C.Structure):
    _fields_ = [("molecular", C.c_int * MAX_SPECIES), ("A_MW", C.c_double * MAX_SPECIES),
                ("abs_tol", C.c_double), ("rel_tol", C.c_double)]


def tt_params(species, abs_tol=1.0e-9, rel_tol=1.0e-9) -> TTParams:
    """Relaxation data for the named species (neutral air: N2, O2, NO, N, O)."""
    prm = TTParams()
    for i, s in enumerate(species):
        if s not in ("N2", "O2", "NO", "N", "O"):
            raise ValueError(f"species {s!r} is not part of the 5-species neutral two-temperature air model")
        prm.molecular[i] = 1 if s in A_MW else 0
        prm.A_MW[i] = A_MW.get(s, 0.0)
    prm.abs_tol, prm.rel_tol = abs_tol, rel_tol
    return prm


_bound = False


def _lib():
    global _bound
    from . import hip
    L = hip.lib()
    if not _bound:
        L.gdtk_hip_tt_create.argtypes = [C.POINTER(TPGGas), C.POINTER(Mechanism), C.POINTER(TTParams),
                                         C.POINTER(C.c_int)]
        L.gdtk_hip_tt_destroy.argtypes = [C.c_int]
        L.gdtk_hip_tt_step_device.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 8 + [C.c_double] + \
            [C.c_void_p] * 8 + [C.c_void_p]
        L.gdtk_hip_tt_stats.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int]
        _bound = True
    return L


class TwoTemperatureReactor:
    """gdtk_hip_tt_create: the reactor on the GPU."""

    def __init__(self, gas: TPGGas, mech: Mechanism, params: TTParams):
        self.gas, self.mech, self.params = gas, mech, params
        self.nsp, self.nr = gas.nsp, mech.nr
        h = C.c_int(-1)
        _check(_lib().gdtk_hip_tt_create(C.byref(gas), C.byref(mech), C.byref(params), C.byref(h)))
        self.h = h.value

    def close(self):
        if getattr(self, "h", -1) >= 0:
            _lib().gdtk_hip_tt_destroy(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_device(self, n, rho, T, Tv, u, uv, p, massf, dt_chem, dt, T_out, Tv_out, u_out, uv_out, p_out,
                    massf_out, dt_chem_out, status, stream=None):
        """Asynchronous step on device pointers (ints, e.g. torch tensor.data_ptr())."""
        _check(_lib().gdtk_hip_tt_step_device(self.h, n, rho, T, Tv, u, uv, p, massf, dt_chem, dt, T_out, Tv_out,
                                              u_out, uv_out, p_out, massf_out, dt_chem_out, status, stream))

    def step(self, rho, T, Tv, u, uv, p, massf, dt_chem, dt):
        """One increment of n cells from host arrays (through torch device tensors); returns
        a dict of host arrays."""
        import torch
        dev = torch.device("cuda:0")
        n = len(rho)
        t = {k: torch.tensor(_arr(v), dtype=torch.float64, device=dev)
             for k, v in (("rho", rho), ("T", T), ("Tv", Tv), ("u", u), ("uv", uv), ("p", p), ("massf", massf),
                          ("dt_chem", dt_chem))}
        o = {k: torch.empty_like(t[k]) for k in ("T", "Tv", "u", "uv", "p", "massf", "dt_chem")}
        st = torch.empty(n, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream()
        self.step_device(n, *(t[k].data_ptr() for k in ("rho", "T", "Tv", "u", "uv", "p", "massf", "dt_chem")), dt,
                         *(o[k].data_ptr() for k in ("T", "Tv", "u", "uv", "p", "massf", "dt_chem")), st.data_ptr(),
                         stream.cuda_stream)
        torch.cuda.synchronize()
        out = {k: v.cpu().numpy() for k, v in o.items()}
        out["status"] = st.cpu().numpy()
        return out

    def stats(self):
        v = np.zeros(5)
        _check(_lib().gdtk_hip_tt_stats(self.h, v.ctypes.data_as(C.POINTER(C.c_double)), 5))
        return dict(substeps_total=int(v[0]), substeps_max=int(v[1]), attempts_total=int(v[2]), ms=float(v[3]),
                    cells=int(v[4]))
