"""ctypes host over libgdtkhip.so, mirroring the reference's simulation/block API.

The class layout follows the reference's control flow: a Simulation owns the
GlobalConfig subset and SimState (src/eilmer/simcore.d), blocks are added from a grid
and an initial flow state (src/eilmer/sfluidblock.d, simcore.d:326-330), and run()
is the time loop (simcore.d:1014-1377).  Every call goes to the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "libgdtkhip.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

GDTK_OK = 0
ERR_NAMES = {-1: "GDTK_ERR_ARG", -2: "GDTK_ERR_HIP", -3: "GDTK_ERR_BAD_CELL", -4: "GDTK_ERR_TOO_MANY_BAD_CELLS",
             -5: "GDTK_ERR_STEP_FAILED", -6: "GDTK_ERR_UNSUPPORTED"}

FIELDS = dict(rho=0, vx=1, vy=2, vz=3, u=4, p=5, T=6, a=7, volume=8, iLength=9, jLength=10, kLength=11, S=12,
              dt_chem=13)
# Flux calculators (src/eilmer/globalconfig.d:270-333 names; "adaptive" is the reference's alias).
FLUX_CALCULATORS = {"ausmdv": 0, "efm": 1, "adaptive_efm_ausmdv": 2, "adaptive": 2, "hanel": 3,
                    "adaptive_hanel_ausmdv": 4}
# Numerics of the flux path: Eilmer 4 (src/eilmer) or Eilmer 5 (src/lmr: dimensionalised
# van Albada epsilon, smooth-maximum AUSMDV sound speed).
NUMERICS = {"eilmer4": 0, "lmr": 1}
# Shock-detector options and defaults (src/eilmer/globalconfig.d:1061-1131).
SHOCK_DETECTOR_DEFAULTS = dict(compression_tolerance=-0.30, shear_tolerance=0.20, strict_shock_detector=True,
                               damped_outflow=False, frozen_shock_detector=False,
                               shock_detector_freeze_step=1_000_000_000)

# Every symbol the C ABI exports (include/gdtk_hip.h).
EXPORTS = [
    "gdtk_hip_sim_create", "gdtk_hip_sim_destroy", "gdtk_hip_sim_add_block", "gdtk_hip_block_set_initial",
    "gdtk_hip_block_get_field", "gdtk_hip_stage", "gdtk_hip_block_dt", "gdtk_hip_sim_run",
    "gdtk_hip_sim_step_fixed_dt", "gdtk_hip_sim_set_dt", "gdtk_hip_sim_profile", "gdtk_hip_sim_kernel_stats",
    "gdtk_hip_sim_set_tuning", "gdtk_hip_sim_autotune", "gdtk_hip_last_error", "gdtk_hip_exchange_source",
    "gdtk_hip_sim_set_gas_tpg", "gdtk_hip_block_set_initial_massf", "gdtk_hip_sim_set_chemistry",
    "gdtk_hip_sim_chem_stats", "gdtk_hip_sim_set_shock_detector", "gdtk_hip_sim_set_numerics", "gdtk_hip_block_write_flow",
    "gdtk_hip_block_read_flow", "gdtk_hip_block_set_transport", "gdtk_hip_sim_set_time", "gdtk_hip_sim_get_state", "gdtk_hip_flow_nvars",
    "gdtk_hip_flow_write_records", "gdtk_hip_flow_read_records",
    "gdtk_hip_chem_create", "gdtk_hip_chem_destroy", "gdtk_hip_chem_step", "gdtk_hip_chem_step_device",
    "gdtk_hip_chem_stats", "gdtk_hip_chem_substeps", "gdtk_hip_chem_set_tuning", "gdtk_hip_chem_autotune",
    "gdtk_hip_tpg_update",
    "gdtk_hip_mlp_create", "gdtk_hip_mlp_destroy", "gdtk_hip_mlp_nparams", "gdtk_hip_mlp_get_params",
    "gdtk_hip_mlp_set_params", "gdtk_hip_mlp_set_bf16_training", "gdtk_hip_mlp_forward", "gdtk_hip_mlp_loss_grad", "gdtk_hip_mlp_train",
    "gdtk_hip_mlp_infer", "gdtk_hip_mfma_probe", "gdtk_hip_surrogate_create", "gdtk_hip_surrogate_destroy",
    "gdtk_hip_surrogate_step_device", "gdtk_hip_surrogate_stats", "gdtk_hip_surrogate_project",
    "gdtk_hip_chem_step_device_noalloc", "gdtk_hip_chem_noalloc_max_cells",
    "gdtk_hip_tt_create", "gdtk_hip_tt_destroy", "gdtk_hip_tt_step_device", "gdtk_hip_tt_stats",
]


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class SimConfig(C.Structure):
    _fields_ = [("dimensions", C.c_int), ("flux_calculator", C.c_int), ("interpolation_order", C.c_int),
                ("cfl_value", C.c_double), ("dt_init", C.c_double), ("dt_max", C.c_double),
                ("cfl_count", C.c_int), ("max_attempts_for_step", C.c_int), ("max_invalid_cells", C.c_int),
                ("mMass", C.c_double), ("gamma", C.c_double), ("use_fused", C.c_int)]


class ShockDetector(C.Structure):
    _fields_ = [("compression_tolerance", C.c_double), ("shear_tolerance", C.c_double),
                ("strict_shock_detector", C.c_int), ("damped_outflow", C.c_int),
                ("frozen_shock_detector", C.c_int), ("shock_detector_freeze_step", C.c_int)]


class BlockDesc(C.Structure):
    _fields_ = [("nic", C.c_int), ("njc", C.c_int), ("nkc", C.c_int), ("vertices", dp),
                ("bc_kind", C.c_int * 6), ("inflow", (C.c_double * 5) * 6),
                ("other_block", C.c_int * 6), ("other_face", C.c_int * 6), ("other_orientation", C.c_int * 6),
                ("inflow_massf", (C.c_double * 8) * 6)]


def lib_path() -> str:
    return _LIB_PATH


def _bind_torch_runtime():
    """A process can hold one HIP runtime per device.  PyTorch ships its own libamdhip64;
    importing torch before the library makes the library's libamdhip64 dependency resolve
    (by SONAME) to that already-loaded runtime, so torch tensors and this library share
    one runtime, one device context and one address space.  Without torch the library uses
    the system ROCm runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise HipError(-2, f"{_LIB_PATH} is not built; run gdtk_amd.build (hipcc, gfx950)")
        _bind_torch_runtime()
        L = C.CDLL(_LIB_PATH)
        L.gdtk_hip_last_error.restype = C.c_char_p
        L.gdtk_hip_sim_create.argtypes = [C.POINTER(SimConfig), ip]
        L.gdtk_hip_sim_destroy.argtypes = [C.c_int]
        L.gdtk_hip_sim_add_block.argtypes = [C.c_int, C.POINTER(BlockDesc), ip]
        L.gdtk_hip_block_set_initial.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.gdtk_hip_block_get_field.argtypes = [C.c_int, C.c_int, C.c_int, dp]
        L.gdtk_hip_stage.argtypes = [C.c_int, C.c_int, C.c_double]
        L.gdtk_hip_block_dt.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, dp]
        L.gdtk_hip_sim_run.argtypes = [C.c_int, C.c_int, C.c_double, dp]
        L.gdtk_hip_sim_step_fixed_dt.argtypes = [C.c_int, C.c_double, dp]
        L.gdtk_hip_sim_set_dt.argtypes = [C.c_int, C.c_double, C.c_int]
        L.gdtk_hip_sim_profile.argtypes = [C.c_int, C.c_int]
        L.gdtk_hip_sim_kernel_stats.argtypes = [C.c_int, dp, C.c_int]
        L.gdtk_hip_sim_set_tuning.argtypes = [C.c_int, C.c_int, C.c_int]
        L.gdtk_hip_sim_autotune.argtypes = [C.c_int, C.c_double, C.c_int, dp, C.c_int]
        L.gdtk_hip_exchange_source.argtypes = [C.c_int] * 10 + [ip]
        L.gdtk_hip_sim_set_gas_tpg.argtypes = [C.c_int, C.c_void_p]
        L.gdtk_hip_block_set_initial_massf.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, dp]
        L.gdtk_hip_sim_set_chemistry.argtypes = [C.c_int, C.c_int, C.c_int]
        L.gdtk_hip_sim_chem_stats.argtypes = [C.c_int, dp, C.c_int]
        L.gdtk_hip_sim_set_shock_detector.argtypes = [C.c_int, C.POINTER(ShockDetector)]
        L.gdtk_hip_sim_set_numerics.argtypes = [C.c_int, C.c_int]
        L.gdtk_hip_block_write_flow.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_double,
                                                C.c_char_p]
        L.gdtk_hip_block_read_flow.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_char_p, dp]
        L.gdtk_hip_block_set_transport.argtypes = [C.c_int, C.c_int, dp, dp]
        L.gdtk_hip_sim_set_time.argtypes = [C.c_int, C.c_double]
        L.gdtk_hip_sim_get_state.argtypes = [C.c_int, dp]
        L.gdtk_hip_flow_nvars.argtypes = [C.c_int]
        L.gdtk_hip_flow_write_records.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_double, C.c_int, C.c_int,
                                                  C.c_int, C.c_int, C.c_int, C.c_char_p, dp]
        L.gdtk_hip_flow_read_records.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_char_p, dp, dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != GDTK_OK:
        raise HipError(rc, lib().gdtk_hip_last_error().decode())
    return rc


def _check_n(rc):
    if rc < 0:
        raise HipError(rc, lib().gdtk_hip_last_error().decode())
    return rc


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


class Simulation:
    """A set of structured blocks advanced by the HIP explicit update."""

    def __init__(self, dims, flux="ausmdv", cfl=0.5, dt_init=1.0e-3, dt_max=1.0e-3, cfl_count=10,
                 max_attempts_for_step=3, max_invalid_cells=0, mMass=0.02896, gamma=1.4, fused=True, tpg=None,
                 shock_detector=None, numerics="eilmer4"):
        """flux: "ausmdv", "efm" or "adaptive_efm_ausmdv" ("adaptive": EFM on faces the shock
        detector marks, AUSMDV elsewhere).  tpg: a gdtk_amd.chem.TPGGas record for a
        multi-species thermally perfect gas (replaces the ideal gas given by mMass and gamma).
        shock_detector: dict of SHOCK_DETECTOR_DEFAULTS keys to override.  numerics:
        "eilmer4" or "lmr" (NUMERICS)."""
        self.dims = dims
        self.nsp = int(tpg.nsp) if tpg is not None else 1
        self.ncq = (5 if dims == 3 else 4) + (self.nsp if self.nsp > 1 else 0)
        cfg = SimConfig(dims, FLUX_CALCULATORS[flux], 2, cfl, dt_init, dt_max, cfl_count,
                        max_attempts_for_step, max_invalid_cells, mMass, gamma, 1 if fused else 0)
        h = C.c_int(-1)
        _check(lib().gdtk_hip_sim_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self.cases = []
        if tpg is not None:
            self._tpg = tpg
            _check(lib().gdtk_hip_sim_set_gas_tpg(self.h, C.byref(tpg)))
        if shock_detector:
            self.set_shock_detector(**shock_detector)
        if numerics != "eilmer4":
            self.set_numerics(numerics)

    def set_numerics(self, numerics):
        """"eilmer4" or "lmr" (Eilmer 5's limiter smoothing and AUSMDV sound speed)."""
        if numerics not in NUMERICS:
            raise ValueError(f"unknown numerics {numerics!r}")
        _check(lib().gdtk_hip_sim_set_numerics(self.h, NUMERICS[numerics]))

    def set_shock_detector(self, **opts):
        """Shock-detector options of the adaptive flux calculator (SHOCK_DETECTOR_DEFAULTS keys)."""
        unknown = set(opts) - set(SHOCK_DETECTOR_DEFAULTS)
        if unknown:
            raise ValueError(f"unknown shock-detector options {sorted(unknown)}")
        o = dict(SHOCK_DETECTOR_DEFAULTS)
        o.update(opts)
        sd = ShockDetector(float(o["compression_tolerance"]), float(o["shear_tolerance"]),
                           int(o["strict_shock_detector"]), int(o["damped_outflow"]), int(o["frozen_shock_detector"]),
                           int(o["shock_detector_freeze_step"]))
        _check(lib().gdtk_hip_sim_set_shock_detector(self.h, C.byref(sd)))

    def close(self):
        if getattr(self, "h", -1) >= 0 and _lib is not None:
            _lib.gdtk_hip_sim_destroy(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_block(self, case) -> int:
        v, vp = _d(case.vertices.reshape(-1))
        d = BlockDesc()
        d.nic, d.njc, d.nkc = case.nic, case.njc, case.nkc
        d.vertices = vp
        for s in range(6):
            d.bc_kind[s] = int(case.bc[s])
            d.other_block[s] = int(case.other_block[s])
            d.other_face[s] = int(case.other_face[s])
            d.other_orientation[s] = int(getattr(case, "other_orientation", [0] * 6)[s])
            for q in range(5):
                d.inflow[s][q] = float(case.inflow[s][q])
            if self.nsp > 1 and getattr(case, "inflow_massf", None) is not None:
                for q in range(self.nsp):
                    d.inflow_massf[s][q] = float(case.inflow_massf[s][q])
        bid = C.c_int(-1)
        _check(lib().gdtk_hip_sim_add_block(self.h, C.byref(d), C.byref(bid)))
        arrs = [_d(case.p), _d(case.T), _d(case.vel[:, 0]), _d(case.vel[:, 1]), _d(case.vel[:, 2])]
        if self.nsp > 1:
            mf = _d(np.asarray(case.massf).reshape(-1, self.nsp).T)  # species-major
            _check(lib().gdtk_hip_block_set_initial_massf(self.h, bid.value, *[a[1] for a in arrs], mf[1]))
        else:
            _check(lib().gdtk_hip_block_set_initial(self.h, bid.value, *[a[1] for a in arrs]))
        self.cases.append(case)
        return bid.value

    def get(self, b, name):
        f = FIELDS[name] if isinstance(name, str) else int(name)
        out = np.empty(self.cases[b].ncells)
        _check(lib().gdtk_hip_block_get_field(self.h, b, f, out.ctypes.data_as(dp)))
        return out

    def get_massf(self, b):
        """Mass fractions (ncells, nsp) of a multi-species gas."""
        return np.stack([self.get(b, 300 + s) for s in range(self.nsp)], axis=1)

    def get_U(self, b):
        return np.stack([self.get(b, 100 + j) for j in range(self.ncq)], axis=1)

    def stage(self, stage, dt):
        return lib().gdtk_hip_stage(self.h, stage, dt)

    def block_dt(self, b, dt_current, cfl, check_cfl):
        out = np.zeros(3)
        _check(lib().gdtk_hip_block_dt(self.h, b, dt_current, cfl, 1 if check_cfl else 0, out.ctypes.data_as(dp)))
        return out

    @staticmethod
    def _state(st):
        return dict(time=st[0], dt=st[1], step=int(st[2]), cfl_max=st[3], dt_allow=st[4], attempts=int(st[5]))

    def run(self, max_steps, max_time=1.0e30):
        st = np.zeros(6)
        rc = lib().gdtk_hip_sim_run(self.h, max_steps, max_time, st.ctypes.data_as(dp))
        if rc < 0:
            _check(rc)
        return rc, self._state(st)

    def state(self):
        st = np.zeros(6)
        _check(lib().gdtk_hip_sim_get_state(self.h, st.ctypes.data_as(dp)))
        return self._state(st)

    def step_fixed_dt(self, dt):
        st = np.zeros(6)
        _check(lib().gdtk_hip_sim_step_fixed_dt(self.h, dt, st.ctypes.data_as(dp)))
        return self._state(st)

    def set_dt(self, dt_global, step):
        _check(lib().gdtk_hip_sim_set_dt(self.h, dt_global, step))

    def profile(self, enable=True):
        _check(lib().gdtk_hip_sim_profile(self.h, 1 if enable else 0))

    def set_kchunk(self, kc):
        _check(lib().gdtk_hip_sim_set_tuning(self.h, 1, int(kc)))

    def set_chemistry(self, chem, surrogate=None):
        """Operator-split chemistry after every step: a gdtk_amd.chem.Chemistry, optionally
        served by a gdtk_amd.surrogate.Surrogate built on it."""
        self._chem = (chem, surrogate)
        _check(lib().gdtk_hip_sim_set_chemistry(self.h, chem.h, surrogate.h if surrogate is not None else -1))

    def write_flow(self, b, filename, fmt="gziptext", label="", sim_time=None, species=None):
        """write_solution of block b: the reference's legacy flow file ("gziptext" or "rawbinary").
        species: the species names of the mass-fraction variables (default "air" for the ideal
        gas); sim_time defaults to the simulation's time."""
        if sim_time is None:
            sim_time = self.state()["time"]
        _check(lib().gdtk_hip_block_write_flow(self.h, b, os.fsencode(filename), FLOW_FORMATS[fmt],
                                               label.encode(), float(sim_time), _names(species)))

    def read_flow(self, b, filename, fmt="gziptext", species=None):
        """read_legacy_solution + cold start of block b from a flow file; returns its sim_time."""
        t = C.c_double(0.0)
        _check(lib().gdtk_hip_block_read_flow(self.h, b, os.fsencode(filename), FLOW_FORMATS[fmt], _names(species),
                                              C.byref(t)))
        return t.value

    def set_transport(self, b, mu, k):
        """Per-cell mu and k of block b, carried to the flow files."""
        m, mp = _d(mu)
        kk, kp = _d(k)
        _check(lib().gdtk_hip_block_set_transport(self.h, b, mp, kp))

    def set_time(self, t):
        _check(lib().gdtk_hip_sim_set_time(self.h, float(t)))

    def chem_stats(self):
        out = np.zeros(3)
        _check_n(lib().gdtk_hip_sim_chem_stats(self.h, out.ctypes.data_as(dp), 3))
        return dict(steps=int(out[0]), ms=out[1], surrogate_queued=int(out[2]))

    def set_graph(self, enable):
        """Run each step's attempt as one captured HIP graph (default) or kernel by kernel."""
        _check(lib().gdtk_hip_sim_set_tuning(self.h, 2, 1 if enable else 0))

    def set_wg_per_cu(self, n):
        """Workgroups per CU of the persistent stage launch (0: occupancy of the kernel)."""
        _check(lib().gdtk_hip_sim_set_tuning(self.h, 3, int(n)))

    def set_tile_x(self, tx):
        """Tile width of the fused stage kernel: 64 (x 4), 32 (x 8) or 16 (x 16) cells."""
        _check(lib().gdtk_hip_sim_set_tuning(self.h, 5, int(tx)))

    def set_variant(self, variant):
        """-1: from the configuration, 0: face-per-thread kernels, 1: fused stage kernel."""
        _check(lib().gdtk_hip_sim_set_tuning(self.h, 4, int(variant)))

    KERNEL_KINDS = ("face_flux", "cell_update", "fused_stage", "ghost_fill", "shock_detect")

    def kernel_stats(self):
        n = 3 * len(self.KERNEL_KINDS)
        out = np.zeros(n)
        _check(lib().gdtk_hip_sim_kernel_stats(self.h, out.ctypes.data_as(dp), n))
        return {k: dict(launches=int(out[3 * q]), ms=out[3 * q + 1], items_per_launch=out[3 * q + 2])
                for q, k in enumerate(self.KERNEL_KINDS)}

    def autotune(self, dt=None, reps=5):
        """Time the stage-kernel candidates on the current state and keep the fastest."""
        if dt is None:
            dt = 0.25 * self.block_dt(0, 1.0, 0.5, False)[0]
        out = np.zeros(64)
        n = lib().gdtk_hip_sim_autotune(self.h, dt, reps, out.ctypes.data_as(dp), 64)
        if n < 0:
            _check(n)
        cands = [dict(variant="fused" if out[q] == 1 else "face", kchunk=int(out[q + 1]), ms=out[q + 2],
                      tile_x=int(out[q + 3])) for q in range(4, n, 4)]
        return dict(best="fused" if out[0] == 1 else "face", kchunk=int(out[1]), ms=out[2], tile_x=int(out[3]),
                    candidates=cands)


FLOW_FORMATS = {"gziptext": 0, "rawbinary": 1}


def _names(species):
    if species is None:
        return None
    if isinstance(species, str):
        return species.encode()
    return " ".join(species).encode()


def flow_nvars(nsp):
    """Doubles per cell record of a flow file for a gas of nsp species."""
    return _check_n(lib().gdtk_hip_flow_nvars(nsp))


def write_flow_records(filename, records, dims, nic, njc, nkc, fmt="gziptext", label="", sim_time=0.0, species=None):
    """The library's flow-file writer on host records (ncell, nvars); needs no GPU."""
    r, rp = _d(records)
    nsp = r.shape[1] - 17 - (1 if r.shape[1] > 18 else 0)
    _check(lib().gdtk_hip_flow_write_records(os.fsencode(filename), FLOW_FORMATS[fmt], label.encode(), float(sim_time),
                                             dims, nic, njc, nkc, nsp, _names(species), rp))


def read_flow_records(filename, dims, nic, njc, nkc, nsp=1, fmt="gziptext", species=None):
    """The library's flow-file reader: (records (ncell, nvars), sim_time); needs no GPU."""
    out = np.zeros((nic * njc * nkc, flow_nvars(nsp)))
    t = C.c_double(0.0)
    _check(lib().gdtk_hip_flow_read_records(os.fsencode(filename), FLOW_FORMATS[fmt], dims, nic, njc, nkc, nsp,
                                            _names(species), out.ctypes.data_as(dp), C.byref(t)))
    return out, t.value


def exchange_source(dims, face, other_face, orientation, a, b, n, nic, njc, nkc):
    """(i, j, k) of the source cell of ghost layer n at in-face position (a, b) of an
    exchange face, in the other block of nic x njc x nkc cells (the mapping the exchange
    kernel uses; runs on the host)."""
    out = (C.c_int * 3)()
    _check(lib().gdtk_hip_exchange_source(dims, face, other_face, orientation, a, b, n, nic, njc, nkc,
                                          C.cast(out, ip)))
    return tuple(out)
