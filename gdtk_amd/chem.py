"""Thermochemistry records of the C ABI (gdtk_tpg_gas, gdtk_mechanism) built from the JSON
gas-model and mechanism data, and the ctypes host of the chemistry entry points.

The record layout is the one declared in include/gdtk_hip.h.  The data format follows the
reference's machine-readable gas and reaction files (src/gas/therm_perf_gas.d:40-120,
src/kinetics/reaction.d:460-540).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

MAX_SPECIES = 8
MAX_REACTIONS = 32
ODE_ALPHA_QSS = 0
ODE_RKF = 1


class TPGGas(C.Structure):
    _fields_ = [("nsp", C.c_int), ("M", C.c_double * MAX_SPECIES), ("nseg", C.c_int * MAX_SPECIES),
                ("T_breaks", (C.c_double * 4) * MAX_SPECIES), ("T_blends", (C.c_double * 2) * MAX_SPECIES),
                ("coeffs", ((C.c_double * 9) * 3) * MAX_SPECIES)]


class Reaction(C.Structure):
    _fields_ = [("collider", C.c_int), ("nreac", C.c_int), ("reac_sp", C.c_int * 3), ("reac_nu", C.c_int * 3),
                ("nprod", C.c_int), ("prod_sp", C.c_int * 3), ("prod_nu", C.c_int * 3),
                ("neff", C.c_int), ("eff_sp", C.c_int * MAX_SPECIES), ("eff", C.c_double * MAX_SPECIES),
                ("f_from_eq", C.c_int), ("fA", C.c_double), ("fn", C.c_double), ("fC", C.c_double),
                ("b_from_eq", C.c_int), ("bA", C.c_double), ("bn", C.c_double), ("bC", C.c_double),
                ("f_park", C.c_int), ("f_s", C.c_double), ("b_park", C.c_int), ("b_s", C.c_double)]


class Mechanism(C.Structure):
    _fields_ = [("nr", C.c_int), ("r", Reaction * MAX_REACTIONS), ("T_lower", C.c_double), ("T_upper", C.c_double),
                ("method", C.c_int), ("eps1", C.c_double), ("eps2", C.c_double), ("delta", C.c_double),
                ("maxIters", C.c_int), ("errTol", C.c_double), ("tightTempCoupling", C.c_int),
                ("maxSubcycles", C.c_int), ("maxAttempts", C.c_int), ("T_frozen", C.c_double)]


def gas_record(d: dict) -> TPGGas:
    """From a gas JSON: {species, M, thermo: [{T_breaks, T_blends, coeffs}]} (2 or 3 CEA segments)."""
    g = TPGGas()
    nsp = len(d["M"])
    if nsp > MAX_SPECIES:
        raise ValueError("too many species")
    g.nsp = nsp
    for i in range(nsp):
        g.M[i] = d["M"][i]
        th = d["thermo"][i]
        nseg = len(th["coeffs"])
        if not 2 <= nseg <= 3 or len(th["T_breaks"]) != nseg + 1 or len(th["T_blends"]) != nseg - 1:
            raise ValueError("CEA curves need 2 or 3 segments with matching breaks and blends")
        g.nseg[i] = nseg
        for k in range(nseg + 1):
            g.T_breaks[i][k] = th["T_breaks"][k]
        for k in range(nseg - 1):
            g.T_blends[i][k] = th["T_blends"][k]
        for s in range(nseg):
            for j in range(9):
                g.coeffs[i][s][j] = th["coeffs"][s][j]
    return g


def mech_record(d: dict, method=None, T_frozen=300.0, **overrides) -> Mechanism:
    """From a mechanism JSON {species, config, reactions}; method/overrides replace config."""
    m = Mechanism()
    rs = d["reactions"]
    if len(rs) > MAX_REACTIONS:
        raise ValueError("too many reactions")
    m.nr = len(rs)
    for k, r in enumerate(rs):
        rr = m.r[k]
        rr.collider = 1 if r["type"] == "anonymous_collider" else 0
        rr.nreac = len(r["reactants"])
        for q, (i, c) in enumerate(r["reactants"]):
            rr.reac_sp[q], rr.reac_nu[q] = int(i), int(c)
        rr.nprod = len(r["products"])
        for q, (i, c) in enumerate(r["products"]):
            rr.prod_sp[q], rr.prod_nu[q] = int(i), int(c)
        eff = r.get("efficiencies", [])
        rr.neff = len(eff)
        for q, (i, v) in enumerate(eff):
            rr.eff_sp[q], rr.eff[q] = int(i), float(v)
        for pre, key in (("f", "frc"), ("b", "brc")):
            rc = r[key]
            if rc["model"] == "fromEqConst":
                setattr(rr, pre + "_from_eq", 1)
            else:
                if rc["model"] not in ("Arrhenius", "Park"):
                    raise ValueError(f"rate constant model {rc['model']!r} is not supported")
                setattr(rr, pre + "_from_eq", 0)
                setattr(rr, pre + "A", float(rc["A"]))
                setattr(rr, pre + "n", float(rc["n"]))
                setattr(rr, pre + "C", float(rc["C"]))
                if rc["model"] == "Park":
                    # rate_constant.d:365-395: A Te^n exp(-C/Te), Te = T^s Tv^(1-s)
                    setattr(rr, pre + "_park", 1)
                    setattr(rr, pre + "_s", float(rc["s"]))
    cfg = d["config"]
    m.T_lower = float(cfg["tempLimits"]["lower"])
    m.T_upper = float(cfg["tempLimits"]["upper"])
    ode = cfg["odeStep"]
    meth = method if method is not None else ode["method"]
    m.method = ODE_RKF if meth == "rkf" else ODE_ALPHA_QSS
    eps1 = ode.get("eps1", 1.0e-3)
    m.eps1 = eps1
    m.eps2 = ode.get("eps2", eps1 / 2.0)
    m.delta = ode.get("delta", 1.0e-10)
    m.maxIters = int(ode.get("maxIters", 10))
    m.errTol = ode.get("errTol", 1.0e-3)
    m.tightTempCoupling = 1 if cfg.get("tightTempCoupling", False) else 0
    m.maxSubcycles = int(cfg.get("maxSubcycles", 10000))
    m.maxAttempts = int(cfg.get("maxAttempts", 4))
    m.T_frozen = T_frozen
    for k, v in overrides.items():
        setattr(m, k, v)
    return m


def element_matrix(species):
    """Atoms of N and O per species (rows N, O) for the 5-species air species names."""
    atoms = {"N2": (2, 0), "O2": (0, 2), "NO": (1, 1), "N": (1, 0), "O": (0, 1)}
    return np.array([[atoms[s][0] for s in species], [atoms[s][1] for s in species]], dtype=np.float64)


# --------------------------------------------------------------------------------------
# Host of the chemistry entry points (include/gdtk_hip.h), the batched counterpart of
# ThermochemicalReactor.opCall (src/kinetics/thermochemical_reactor.d:45-47) applied per
# cell as FVCell.thermochemical_increment does (src/eilmer/fvcell.d:858-972).

TUNE_PERSISTENT, TUNE_CHUNK, TUNE_REFILL, TUNE_BLOCKS_PER_CU, TUNE_HEAVY_RATIO = 1, 2, 3, 4, 5
TPG_MODES = {"pT": 0, "rhou": 1, "rhoT": 2}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bound = False


def _lib():
    global _bound
    from . import hip
    L = hip.lib()
    if not _bound:
        L.gdtk_hip_chem_create.argtypes = [C.POINTER(TPGGas), C.POINTER(Mechanism), _ip]
        L.gdtk_hip_chem_destroy.argtypes = [C.c_int]
        L.gdtk_hip_chem_step.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _ip, _dp]
        L.gdtk_hip_chem_step_device.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_double] + \
            [C.c_void_p] * 6 + [C.c_void_p]
        L.gdtk_hip_chem_stats.argtypes = [C.c_int, _dp, C.c_int]
        L.gdtk_hip_chem_substeps.argtypes = [C.c_int, C.c_int, _ip]
        L.gdtk_hip_chem_set_tuning.argtypes = [C.c_int, C.c_int, C.c_int]
        L.gdtk_hip_chem_autotune.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_double, _dp, C.c_int]
        L.gdtk_hip_tpg_update.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        _bound = True
    return L


def _check(rc):
    from . import hip
    if rc < 0:
        raise hip.HipError(rc, hip.lib().gdtk_hip_last_error().decode())
    return rc


def _arr(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Chemistry:
    """Thermally perfect gas + reaction mechanism on the GPU (gdtk_hip_chem_create)."""

    def __init__(self, gas: TPGGas, mech: Mechanism):
        self.gas, self.mech = gas, mech
        self.nsp, self.nr = gas.nsp, mech.nr
        h = C.c_int(-1)
        _check(_lib().gdtk_hip_chem_create(C.byref(gas), C.byref(mech), C.byref(h)))
        self.h = h.value

    def close(self):
        if getattr(self, "h", -1) >= 0:
            _lib().gdtk_hip_chem_destroy(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, rho, T, u, p, massf, dt_chem, dt):
        """One thermochemical increment of n cells on host arrays (copied, returned)."""
        n = len(rho)
        rho = _arr(rho, n)
        T, u, p, dtc = (_arr(a, n).copy() for a in (T, u, p, dt_chem))
        mf = _arr(massf, self.nsp * n).copy()
        status = np.zeros(n, dtype=np.int32)
        stats = np.zeros(4)
        failed = _check(_lib().gdtk_hip_chem_step(
            self.h, n, rho.ctypes.data_as(_dp), T.ctypes.data_as(_dp), u.ctypes.data_as(_dp),
            p.ctypes.data_as(_dp), mf.ctypes.data_as(_dp), dtc.ctypes.data_as(_dp), dt,
            status.ctypes.data_as(_ip), stats.ctypes.data_as(_dp)))
        return dict(T=T, u=u, p=p, massf=mf, dt_chem=dtc, status=status, failed=failed,
                    substeps_total=int(stats[0]), iters_total=int(stats[1]), substeps_max=int(stats[2]),
                    ms=float(stats[3]))

    def step_device(self, n, rho, T, u, p, massf, dt_chem, dt, T_out, u_out, p_out, massf_out, dt_chem_out,
                    status, stream=None):
        """Asynchronous step on device pointers (ints, e.g. torch tensor.data_ptr())."""
        _check(_lib().gdtk_hip_chem_step_device(self.h, n, rho, T, u, p, massf, dt_chem, dt, T_out, u_out, p_out,
                                                massf_out, dt_chem_out, status, stream))

    def stats(self):
        v = np.zeros(8)
        _check(_lib().gdtk_hip_chem_stats(self.h, v.ctypes.data_as(_dp), 8))
        return dict(substeps_total=int(v[0]), iters_total=int(v[1]), substeps_max=int(v[2]),
                    attempts_total=int(v[3]), ms_setup=v[4], ms_integrate=v[5], ms_finish=v[6], cells=int(v[7]))

    def substeps(self, n):
        out = np.zeros(n, dtype=np.int32)
        _check(_lib().gdtk_hip_chem_substeps(self.h, n, out.ctypes.data_as(_ip)))
        return out

    def set_tuning(self, persistent=None, chunk=None, refill=None, blocks_per_cu=None, heavy_ratio=None):
        for key, val in ((TUNE_PERSISTENT, persistent), (TUNE_CHUNK, chunk), (TUNE_REFILL, refill),
                         (TUNE_BLOCKS_PER_CU, blocks_per_cu), (TUNE_HEAVY_RATIO, heavy_ratio)):
            if val is not None:
                _check(_lib().gdtk_hip_chem_set_tuning(self.h, key, int(val)))

    def autotune(self, n, rho, T, u, p, massf, dt_chem, dt):
        out = np.zeros(5 * 64)
        k = _check(_lib().gdtk_hip_chem_autotune(self.h, n, rho, T, u, p, massf, dt_chem, dt,
                                                 out.ctypes.data_as(_dp), out.size))
        return [dict(persistent=int(r[0]), chunk=int(r[1]), refill=int(r[2]), blocks_per_cu=int(r[3]), ms=r[4])
                for r in out[:5 * k].reshape(k, 5)]

    def tpg_update(self, mode, massf, rho=None, p=None, T=None, u=None):
        """Gas-model update of n states: mode 'pT', 'rhou' or 'rhoT'; returns dict + failed count."""
        mf = _arr(massf)
        n = mf.size // self.nsp
        z = np.zeros(n)
        rho, p, T, u = (z.copy() if a is None else _arr(a, n).copy() for a in (rho, p, T, u))
        a = np.zeros(n)
        failed = _check(_lib().gdtk_hip_tpg_update(self.h, TPG_MODES[mode], n, rho.ctypes.data_as(_dp),
                                                   p.ctypes.data_as(_dp), T.ctypes.data_as(_dp),
                                                   u.ctypes.data_as(_dp), a.ctypes.data_as(_dp),
                                                   mf.ctypes.data_as(_dp)))
        return dict(rho=rho, p=p, T=T, u=u, a=a, failed=failed)
