"""ctypes wrapper over oracle/_build/liboracle.so (test infrastructure only).

The oracle restates the D reference in C++ (oracle/eilmer_oracle.cpp).  This module
is loaded by tests, smoke() and the CPU-baseline leg of bench.py only.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_last_error.restype = C.c_char_p
        L.orc_gas_create_ideal.restype = C.c_void_p
        L.orc_gas_create_ideal.argtypes = [C.c_double, C.c_double]
        L.orc_gas_destroy.argtypes = [C.c_void_p]
        L.orc_gas_update.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int]
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_set_shock_detector.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                                 C.c_int]
        L.orc_block_face_S.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_block_set_fs.argtypes = [C.c_void_p, C.c_int] + [dp] * 10
        L.orc_sim_set_time.argtypes = [C.c_void_p, C.c_double]
        L.orc_pj_detector.restype = C.c_double
        L.orc_pj_detector.argtypes = [dp, C.c_double, dp, C.c_double, dp, dp, dp, C.c_double, C.c_double]
        L.orc_ffc_source.argtypes = [C.c_int] * 10 + [ip]
        L.orc_sim_add_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, ip, dp, ip, ip, ip]
        L.orc_block_set_initial.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, dp, dp]
        L.orc_block_get.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.orc_block_faces.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_block_face_flux.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_sim_run.argtypes = [C.c_void_p, C.c_int, C.c_double, dp]
        L.orc_sim_step_fixed_dt.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_sim_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_block_dt.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, dp]
        L.orc_sim_set_dt.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_flux.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp, dp, C.c_int]
        L.orc_sim_set_numerics.argtypes = [C.c_void_p, C.c_int]
        L.orc_exxef.argtypes = [C.c_double, dp, dp]
        L.orc_l2r2_scalar.argtypes = [dp, dp, C.c_int, C.c_int, C.c_double, dp, C.c_int]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ip)


def err():
    return lib().orc_last_error().decode()


class IdealGas:
    def __init__(self, mMass=0.02896, gamma=1.4):
        self.h = lib().orc_gas_create_ideal(mMass, gamma)
        self.nsp = 1

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_gas_destroy(self.h)
            self.h = None

    def update(self, mode, rho=0.0, p=0.0, T=0.0, u=0.0, massf=None):
        st, stp = _d([rho, p, T, u, 0.0])
        mf = None
        if massf is not None:
            mfa, mf = _d(massf)
        rc = lib().orc_gas_update(self.h, mode, stp, mf)
        if rc != 0:
            raise RuntimeError(err())
        return dict(rho=st[0], p=st[1], T=st[2], u=st[3], a=st[4])


FIELDS = dict(rho=0, vx=1, vy=2, vz=3, u=4, p=5, T=6, a=7, volume=8, iLength=9, jLength=10, kLength=11, S=12, pos_x=13, pos_y=14,
              pos_z=15, dt_chem=16)


FLUX_CALCULATORS = {"ausmdv": 0, "efm": 1, "adaptive_efm_ausmdv": 2, "adaptive": 2, "hanel": 3,
                    "adaptive_hanel_ausmdv": 4}
# globalconfig.d:1061-1131
SHOCK_DETECTOR_DEFAULTS = dict(compression_tolerance=-0.30, shear_tolerance=0.20, strict_shock_detector=True,
                               damped_outflow=False, frozen_shock_detector=False,
                               shock_detector_freeze_step=1_000_000_000)


class Sim:
    """Whole-simulation oracle: blocks, BCs, the pc explicit update and the dt controller."""

    def __init__(self, gas, dims, flux="ausmdv", cfl=0.5, dt_init=1.0e-3, dt_max=1.0e-3, cfl_count=10,
                 mutate_cell_velocities=True, shock_detector=None, numerics="eilmer4"):
        """flux: "ausmdv", "efm" or "adaptive_efm_ausmdv" (EFM at detected shocks);
        shock_detector: dict of the detector options compression_tolerance (-0.30),
        shear_tolerance (0.20), strict_shock_detector (True), damped_outflow (False),
        frozen_shock_detector (False), shock_detector_freeze_step (1e9); numerics: "eilmer4"
        or "lmr" (Eilmer 5's limiter smoothing and AUSMDV sound speed)."""
        self.gas = gas
        self.dims = dims
        fc = FLUX_CALCULATORS[flux]
        self.h = lib().orc_sim_create(gas.h, dims, fc, cfl, dt_init, dt_max, cfl_count,
                                      1 if mutate_cell_velocities else 0)
        sd = dict(SHOCK_DETECTOR_DEFAULTS)
        sd.update(shock_detector or {})
        lib().orc_sim_set_shock_detector(self.h, sd["compression_tolerance"], sd["shear_tolerance"],
                                         int(sd["strict_shock_detector"]), int(sd["damped_outflow"]),
                                         int(sd["frozen_shock_detector"]), int(sd["shock_detector_freeze_step"]))
        if numerics not in ("eilmer4", "lmr"):
            raise ValueError(f"unknown numerics {numerics!r}")
        lib().orc_sim_set_numerics(self.h, 1 if numerics == "lmr" else 0)
        self.blocks = []

    def close(self):
        if self.h:
            lib().orc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_block(self, case):
        v, vp = _d(case.vertices.reshape(-1))
        bc, bcp = _i(case.bc)
        nsp = self.gas.nsp
        inflow = np.zeros((6, 5 + nsp))
        inflow[:, :5] = case.inflow
        if nsp == 1:
            inflow[:, 5] = 1.0
        elif getattr(case, "inflow_massf", None) is not None:
            inflow[:, 5:] = case.inflow_massf
        inf, infp = _d(inflow)
        ob, obp = _i(case.other_block)
        of, ofp = _i(case.other_face)
        oo, oop = _i(case.other_orientation)
        b = lib().orc_sim_add_block(self.h, case.nic, case.njc, case.nkc, vp, bcp, infp, obp, ofp, oop)
        if b < 0:
            raise RuntimeError(err())
        arrs = [_d(case.p), _d(case.T), _d(case.vel[:, 0]), _d(case.vel[:, 1]), _d(case.vel[:, 2])]
        mf = None
        if nsp > 1:
            mf = _d(np.asarray(case.massf).reshape(-1, nsp))  # cell-major
        rc = lib().orc_block_set_initial(self.h, b, *[a[1] for a in arrs], mf[1] if mf is not None else None)
        if rc != 0:
            raise RuntimeError(err())
        self.blocks.append(case)
        return b

    def get(self, b, name):
        case = self.blocks[b]
        out = np.empty(case.ncells)
        if isinstance(name, str):
            f = FIELDS[name]
        else:
            f = name
        rc = lib().orc_block_get(self.h, b, f, out.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError(err())
        return out

    def get_U(self, b, ncq):
        return np.stack([self.get(b, 100 + j) for j in range(ncq)], axis=1)

    def faces(self, b):
        case = self.blocks[b]
        niv, njv = case.nic + 1, case.njc + 1
        nkv = case.nkc + 1
        n = niv * case.njc * case.nkc + case.nic * njv * case.nkc
        if self.dims == 3:
            n += case.nic * case.njc * nkv
        out = np.empty((n, 10))
        lib().orc_block_faces(self.h, b, out.ctypes.data_as(dp))
        return out

    def set_fs(self, b, rho, vx, vy, vz, p, a, u, T, massf=None, S=None):
        """Cold start of block b from flow-file cell values (no EOS call before the encode)."""
        arrs = [_d(x) for x in (rho, vx, vy, vz, p, a, u, T)]
        mf = _d(np.asarray(massf).reshape(len(rho), -1)) if massf is not None else None
        sv = _d(S) if S is not None else None
        rc = lib().orc_block_set_fs(self.h, b, *[x[1] for x in arrs], mf[1] if mf else None, sv[1] if sv else None)
        if rc != 0:
            raise RuntimeError(err())

    def face_S(self, b):
        """Shock indicator of every face of block b (face order ifi, ifj, ifk)."""
        out = np.empty(self.faces(b).shape[0])
        lib().orc_block_face_S(self.h, b, out.ctypes.data_as(dp))
        return out

    def face_flux(self, b, ncq):
        n = self.faces(b).shape[0]
        out = np.empty((n, ncq))
        lib().orc_block_face_flux(self.h, b, out.ctypes.data_as(dp))
        return out

    def run(self, max_steps, max_time=1.0e30):
        st = np.zeros(6)
        n = lib().orc_sim_run(self.h, max_steps, max_time, st.ctypes.data_as(dp))
        if n < 0:
            raise RuntimeError(err())
        return n, dict(time=st[0], dt=st[1], step=int(st[2]), cfl_max=st[3], dt_allow=st[4], attempts=int(st[5]))

    def step_fixed_dt(self, dt):
        st = np.zeros(6)
        rc = lib().orc_sim_step_fixed_dt(self.h, dt, st.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError(err())
        return dict(time=st[0], dt=st[1], step=int(st[2]))

    def stage(self, stage, dt):
        rc = lib().orc_sim_stage(self.h, stage, dt)
        if rc != 0:
            raise RuntimeError(err())

    def block_dt(self, b, dt_current, cfl, check_cfl):
        out = np.zeros(3)
        lib().orc_block_dt(self.h, b, dt_current, cfl, 1 if check_cfl else 0, out.ctypes.data_as(dp))
        return out

    def set_chemistry(self, chem):
        """Operator-split finite-rate chemistry after every step (an oracle Chem object)."""
        L = lib()
        L.orc_sim_set_chemistry.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_sim_set_chemistry(self.h, chem.h)
        self._chem = chem

    def set_dt(self, dt_global, step):
        lib().orc_sim_set_dt(self.h, dt_global, step)

    def set_time(self, t):
        lib().orc_sim_set_time(self.h, t)


def pj_detector(velL, aL, velR, aR, n, t1, t2, comp_tol=-0.30, shear_tol=0.20):
    """PJ shock detector (shockdetectors.d:21-113) on one face between two cells."""
    v = [_d(np.asarray(x, dtype=float)) for x in (velL, velR, n, t1, t2)]
    return lib().orc_pj_detector(v[0][1], aL, v[1][1], aR, v[2][1], v[3][1], v[4][1], comp_tol, shear_tol)


def flux(gas, kind, dims, L, R, massfL=None, massfR=None, lmr=False):
    """Unit flux: L, R = [rho, p, T, u, a, vx, vy, vz] in the face-local frame; lmr: Eilmer
    5's AUSMDV."""
    a, ap = _d(L)
    b, bp = _d(R)
    F = np.zeros(5 + max(gas.nsp, 1) + 1)
    mL = mR = None
    if massfL is not None:
        ml, mL = _d(massfL)
    if massfR is not None:
        mr, mR = _d(massfR)
    n = lib().orc_flux(gas.h, {"ausmdv": 0, "efm": 1, "hanel": 3}[kind], dims, ap, bp, mL, mR, F.ctypes.data_as(dp),
                       1 if lmr else 0)
    return F[:n].copy()


def exxef(sn):
    e1, e2 = C.c_double(), C.c_double()
    lib().orc_exxef(sn, C.byref(e1), C.byref(e2))
    return e1.value, e2.value


def l2r2_scalar(lens, q, apply_limiter=True, extrema_clipping=True, eps=1.0e-12, lmr=False):
    l, lp = _d(lens)
    qq, qp = _d(q)
    out = np.zeros(2)
    lib().orc_l2r2_scalar(lp, qp, 1 if apply_limiter else 0, 1 if extrema_clipping else 0, eps,
                          out.ctypes.data_as(dp), 1 if lmr else 0)
    return out[0], out[1]


# ---------------------------------------------------------------------------------
# Thermochemistry oracle (chem_oracle.cpp)
vp = C.c_void_p


def _chem_lib():
    L = lib()
    if not getattr(L, "_chem_ready", False):
        L.orc_tpg_sizes.argtypes = [ip]
        L.orc_tpg_create.restype = vp
        L.orc_tpg_create.argtypes = [vp]
        L.orc_cea_eval.argtypes = [vp, C.c_int, C.c_double, dp]
        L.orc_tpg_update.argtypes = [vp, C.c_int, dp, dp, ip]
        L.orc_chem_create.restype = vp
        L.orc_chem_create.argtypes = [vp, vp]
        L.orc_chem_destroy.argtypes = [vp]
        L.orc_chem_rate_constants.argtypes = [vp, C.c_double, dp, dp]
        L.orc_chem_split_rates.argtypes = [vp, dp, dp, dp]
        L.orc_chem_ode_step.argtypes = [vp, C.c_int, dp, C.c_double, dp, dp]
        L.orc_chem_update.argtypes = [vp, dp, dp, C.c_double, dp, dp]
        L.orc_chem_step_batch.argtypes = [vp, vp, C.c_int, dp, dp, dp, dp, dp, dp, C.c_double, ip, ip, ip, C.c_int]
        L.orc_gas_destroy_tpg.argtypes = [vp]
        from gdtk_amd.chem import Mechanism, TPGGas
        sizes = (C.c_int * 3)()
        L.orc_tpg_sizes(sizes)
        assert sizes[0] == C.sizeof(TPGGas) and sizes[1] == C.sizeof(Mechanism), "record layout mismatch"
        L._chem_ready = True
    return L


class TPG:
    """Thermally perfect gas of the oracle, from a gdtk_amd.chem.TPGGas record."""

    def __init__(self, rec):
        self.rec = rec
        self.nsp = rec.nsp
        self.h = _chem_lib().orc_tpg_create(C.byref(rec))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_gas_destroy_tpg(self.h)
            self.h = None

    def cea(self, isp, T):
        out = np.zeros(3)
        _chem_lib().orc_cea_eval(self.h, isp, T, out.ctypes.data_as(dp))
        return dict(Cp=out[0], h=out[1], s=out[2])

    def update(self, mode, massf, rho=0.0, p=0.0, T=0.0, u=0.0):
        st, stp = _d([rho, p, T, u, 0.0])
        mf, mfp = _d(massf)
        ne = C.c_int(0)
        rc = _chem_lib().orc_tpg_update(self.h, {"pT": 0, "rhou": 1, "rhoT": 2, "rhop": 3}[mode], stp, mfp,
                                        C.byref(ne))
        if rc != 0:
            raise RuntimeError(err())
        return dict(rho=st[0], p=st[1], T=st[2], u=st[3], a=st[4], newton_evals=ne.value)


class Chem:
    """ChemistryUpdate of the oracle from gas and mechanism records."""

    def __init__(self, gas_rec, mech_rec):
        self.gas_rec, self.mech_rec = gas_rec, mech_rec
        self.nsp = gas_rec.nsp
        self.nr = mech_rec.nr
        self.h = _chem_lib().orc_chem_create(C.byref(gas_rec), C.byref(mech_rec))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_chem_destroy(self.h)
            self.h = None

    def rate_constants(self, T):
        kf, kb = np.zeros(self.nr), np.zeros(self.nr)
        _chem_lib().orc_chem_rate_constants(self.h, T, kf.ctypes.data_as(dp), kb.ctypes.data_as(dp))
        return kf, kb

    def split_rates(self, conc):
        c, cp = _d(conc)
        q, L = np.zeros(self.nsp), np.zeros(self.nsp)
        _chem_lib().orc_chem_split_rates(self.h, cp, q.ctypes.data_as(dp), L.ctypes.data_as(dp))
        return q, L

    def ode_step(self, method, y0, h):
        y, ypp = _d(y0)
        out = np.zeros(self.nsp)
        hs = C.c_double(0.0)
        ok = _chem_lib().orc_chem_ode_step(self.h, {"alpha-qss": 0, "rkf": 1}[method], ypp, h,
                                           out.ctypes.data_as(dp), C.byref(hs))
        return bool(ok), out, hs.value

    def update(self, rho, p, T, u, massf, tInterval, dtSuggest):
        st, stp = _d([rho, p, T, u])
        mf, mfp = _d(massf)
        mf = mf.copy()
        mfp = mf.ctypes.data_as(dp)
        dts = C.c_double(dtSuggest)
        stats = np.zeros(4)
        rc = _chem_lib().orc_chem_update(self.h, stp, mfp, tInterval, C.byref(dts), stats.ctypes.data_as(dp))
        return rc, mf, dts.value, dict(subcycles=int(stats[0]), corrector_iters=int(stats[1]))


def chem_step_batch(gas_rec, mech_rec, rho, T, u, p, massf, dt_chem, dt, nthreads=1):
    """Batched thermochemical_increment; arrays are copied, results returned."""
    n = len(rho)
    T, u, p, massf, dt_chem = (np.array(a, dtype=np.float64, copy=True) for a in (T, u, p, massf, dt_chem))
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    status = np.zeros(n, dtype=np.int32)
    sub = np.zeros(n, dtype=np.int32)
    att = np.zeros(n, dtype=np.int32)
    failed = _chem_lib().orc_chem_step_batch(C.byref(gas_rec), C.byref(mech_rec), n, rho.ctypes.data_as(dp),
                                             T.ctypes.data_as(dp), u.ctypes.data_as(dp), p.ctypes.data_as(dp),
                                             massf.ctypes.data_as(dp), dt_chem.ctypes.data_as(dp), dt,
                                             status.ctypes.data_as(ip), sub.ctypes.data_as(ip),
                                             att.ctypes.data_as(ip), nthreads)
    return dict(T=T, u=u, p=p, massf=massf, dt_chem=dt_chem, status=status, substeps=sub, attempts=att,
                failed=failed)


# Two-temperature air (tt_oracle.cpp)
def _tt_lib():
    L = _chem_lib()
    if not getattr(L, "_tt_ready", False):
        L.orc_tt_create.restype = vp
        L.orc_tt_create.argtypes = [vp, vp, ip, dp]
        L.orc_tt_destroy.argtypes = [vp]
        L.orc_tt_update.argtypes = [vp, C.c_int, dp, dp, ip]
        L.orc_tt_species_energy.argtypes = [vp, C.c_int, C.c_double, dp]
        L.orc_tt_relaxation.argtypes = [vp, dp, dp, dp, dp]
        L.orc_tt_rate_constants.argtypes = [vp, C.c_double, C.c_double, dp, dp]
        L.orc_tt_step_batch.argtypes = [vp, vp, ip, dp, C.c_int] + [dp] * 8 + [C.c_double, ip, ip, ip, C.c_int]
        L._tt_ready = True
    return L


class TT:
    """TwoTemperatureAir + TwoTemperatureAirKinetics of the oracle.  molecular[nsp], A_MW[nsp]
    as in gdtk_amd.two_temp.tt_params."""

    def __init__(self, gas_rec, mech_rec, molecular, A_MW):
        self.gas_rec, self.mech_rec = gas_rec, mech_rec
        self.nsp, self.nr = gas_rec.nsp, mech_rec.nr
        self.mol = np.ascontiguousarray(molecular, dtype=np.int32)
        self.A = np.ascontiguousarray(A_MW, dtype=np.float64)
        self.h = _tt_lib().orc_tt_create(C.byref(gas_rec), C.byref(mech_rec), _i(self.mol)[1], _d(self.A)[1])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tt_destroy(self.h)
            self.h = None

    def update(self, mode, massf, rho=0.0, p=0.0, T=0.0, Tv=0.0, u=0.0, uv=0.0):
        """mode 'pT', 'rhou' or 'rhoT'; returns the state dict (and Newton iterations)."""
        st, stp = _d([rho, p, T, Tv, u, uv])
        mf, mfp = _d(massf)
        ni = C.c_int(0)
        rc = _tt_lib().orc_tt_update(self.h, {"pT": 0, "rhou": 1, "rhoT": 2}[mode], stp, mfp, C.byref(ni))
        if rc != 0:
            raise RuntimeError(err())
        return dict(rho=st[0], p=st[1], T=st[2], Tv=st[3], u=st[4], uv=st[5], newton_iters=ni.value)

    def species_energy(self, isp, T):
        out = np.zeros(2)
        _tt_lib().orc_tt_species_energy(self.h, isp, T, out.ctypes.data_as(dp))
        return dict(e_ve=out[0], Cv_ve=out[1])

    def relaxation(self, state, massf):
        """(tau per species (0 for atoms), energy-exchange rate of u_v) at a state dict."""
        st, stp = _d([state[k] for k in ("rho", "p", "T", "Tv", "u", "uv")])
        mf, mfp = _d(massf)
        tau = np.zeros(self.nsp)
        rate = C.c_double(0.0)
        _tt_lib().orc_tt_relaxation(self.h, stp, mfp, tau.ctypes.data_as(dp), C.byref(rate))
        return tau, rate.value

    def rate_constants(self, T, Tv):
        kf, kb = np.zeros(self.nr), np.zeros(self.nr)
        if _tt_lib().orc_tt_rate_constants(self.h, T, Tv, kf.ctypes.data_as(dp), kb.ctypes.data_as(dp)) != 0:
            raise RuntimeError(err())
        return kf, kb


def tt_step_batch(gas_rec, mech_rec, molecular, A_MW, rho, T, Tv, u, uv, p, massf, dt_chem, dt, nthreads=1):
    """Batched two-temperature thermochemical_increment; arrays are copied, results returned."""
    n = len(rho)
    T, Tv, u, uv, p, massf, dt_chem = (np.array(a, dtype=np.float64, copy=True)
                                       for a in (T, Tv, u, uv, p, massf, dt_chem))
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    mol = np.ascontiguousarray(molecular, dtype=np.int32)
    A = np.ascontiguousarray(A_MW, dtype=np.float64)
    status = np.zeros(n, dtype=np.int32)
    sub = np.zeros(n, dtype=np.int32)
    att = np.zeros(n, dtype=np.int32)
    failed = _tt_lib().orc_tt_step_batch(C.byref(gas_rec), C.byref(mech_rec), mol.ctypes.data_as(ip),
                                         A.ctypes.data_as(dp), n, rho.ctypes.data_as(dp), T.ctypes.data_as(dp),
                                         Tv.ctypes.data_as(dp), u.ctypes.data_as(dp), uv.ctypes.data_as(dp),
                                         p.ctypes.data_as(dp), massf.ctypes.data_as(dp),
                                         dt_chem.ctypes.data_as(dp), dt, status.ctypes.data_as(ip),
                                         sub.ctypes.data_as(ip), att.ctypes.data_as(ip), nthreads)
    return dict(T=T, Tv=Tv, u=u, uv=uv, p=p, massf=massf, dt_chem=dt_chem, status=status, substeps=sub,
                attempts=att, failed=failed)


def ffc_source(dims, face, other_face, orientation, a, b, n, nic, njc, nkc):
    """(i, j, k) of the full-face-copy source cell of ghost layer n at (a, b) of `face`."""
    out = (C.c_int * 3)()
    lib().orc_ffc_source(dims, face, other_face, orientation, a, b, n, nic, njc, nkc, C.cast(out, ip))
    return tuple(out)
