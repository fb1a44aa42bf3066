"""GPU parity of the chemistry path: the HIP kernels (through the C ABI) against the CPU
oracle's restatement of FVCell.thermochemical_increment (src/eilmer/fvcell.d:858-972),
ChemistryUpdate (src/kinetics/chemistry_update.d) and the thermally perfect gas.

The state (T, u, p, mass fractions) is compared with the relative metric of
tests/parity.py at 1e-12 (FP64).  Per-cell status and substep counts must be identical: a
flipped step-size or convergence branch shows up there first.  The suggested dt_chem has
its own bar, 1e-10: it is h (1/sqrt(sigma) + 0.005) with sigma built from |yc - yp0|,
the difference of corrector and predictor, which is ~eps2 * y (chemistry_update.d:630-655),
so a relative change d of the concentrations moves dt_chem by up to ~d / eps2 = 2000 d.
The RKF step's suggestion is scale * h with scale = 0.9 err^-0.2 and err the norm of
the embedded error estimate, itself a difference of the 4th- and 5th-order solutions of
size ~errTol (chemistry_update.d:456-483): its conditioning is ~1/errTol, so with
errTol = 1e-6 the RKF test holds dt_chem to 1e-5 (the state still to 1e-12).
The H2-I2-HI case also meets the reference's own known answer (chemistry_update.d:739-750).
"""
import json
import os

import numpy as np
import pytest

from gdtk_amd import data as gdata
from gdtk_amd.chem import Chemistry, gas_record, mech_record
from gdtk_amd.chemcases import air_cells
from tests.parity import TOL, rel_err
from tests.test_oracle_chem import _nasa9_g_molar as _g_molar

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
R_U = 8.31451
DT = 1.0e-7
DTCHEM_TOL = 1.0e-10
DTCHEM_TOL_RKF = 1.0e-5


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _air(method=None, **kw):
    gas = gas_record(gdata.load("tpg-air-5sp"))
    mech = mech_record(gdata.load("park-air-5sp-1T"), method=method, **kw)
    return gas, mech


def _oracle_step(gas, mech, c, dtc, dt=DT):
    from oracle import oracle as O
    return O.chem_step_batch(gas, mech, c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, dt, nthreads=8)


def _compare(g, o, label, sub_g=None, tol=TOL, dtc_tol=DTCHEM_TOL):
    st_g, st_o = g["status"], o["status"]
    assert np.array_equal(st_g, st_o), f"{label}: status differs in {np.count_nonzero(st_g != st_o)} cells"
    if sub_g is not None:
        diff = np.nonzero(sub_g != o["substeps"])[0]
        assert diff.size == 0, f"{label}: substep counts differ in cells {diff[:10]}"
    errs = {k: rel_err(g[k], o[k]) for k in ("T", "u", "p", "massf", "dt_chem")}
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, "failed", int(np.count_nonzero(st_o)), flush=True)
    bad = {k: v for k, v in errs.items() if not v <= (dtc_tol if k == "dt_chem" else tol)}
    assert not bad, f"{label}: parity above {tol} (dt_chem {dtc_tol}): {bad}"


def _cell_errors(a, b, n):
    """Per-cell relative error of field arrays ([nfields*]n) as tests/parity.py defines it."""
    a, b = np.asarray(a).reshape(-1, n), np.asarray(b).reshape(-1, n)
    scale = np.maximum(np.abs(b), max(1e-8 * float(np.max(np.abs(b))), 1e-300))
    return np.max(np.abs(a - b) / scale, axis=0)


def _perturbed_steps(gas, mech, c, dtc, dt):
    """The oracle's step from inputs moved in the last bit: T * (1 +- 1e-15), u * (1 + 1e-15)."""
    return [_oracle_step(gas, mech, dict(c, T=c["T"] * (1.0 + 1.0e-15)), dtc, dt),
            _oracle_step(gas, mech, dict(c, T=c["T"] * (1.0 - 1.0e-15)), dtc, dt),
            _oracle_step(gas, mech, dict(c, u=c["u"] * (1.0 + 1.0e-15)), dtc, dt)]


def _compare_conditioned(g, o, o_perts, label, sub_g):
    """GPU against oracle where the oracle's own answer moves under a last-bit perturbation
    of its input (o_perts: _perturbed_steps): status and substeps identical; per field, no
    cell further from the oracle than 10 times the oracle's own largest move (or the 1e-12
    bar, if larger), and no more cells beyond 1e-12 than twice the oracle's own count plus
    1 %."""
    n = g["status"].size
    assert np.array_equal(g["status"], o["status"]), label
    assert np.array_equal(sub_g, o["substeps"]), label
    ok = o["status"] == 0
    for q in o_perts:
        ok &= q["status"] == 0
    rep = {}
    for k in ("T", "u", "p", "massf", "dt_chem"):
        e = _cell_errors(g[k], o[k], n)[ok]
        e_self = np.max([_cell_errors(q[k], o[k], n)[ok] for q in o_perts], axis=0)
        bar = max(TOL if k != "dt_chem" else DTCHEM_TOL, 10.0 * float(e_self.max()))
        far, far_self = int(np.count_nonzero(e > TOL)), int(np.count_nonzero(e_self > TOL))
        rep[k] = dict(max=float(e.max()), cells=far, oracle_self_max=float(e_self.max()), oracle_self_cells=far_self)
        assert e.max() <= bar, (label, k, rep[k])
        assert far <= 2 * far_self + 0.01 * n, (label, k, rep[k])
    print(label, rep, flush=True)


def test_tpg_update_modes():
    """pT, rhou (Newton from a perturbed start) and rhoT updates plus the sound speed,
    state by state against the oracle's ThermallyPerfectGas."""
    from oracle import oracle as O
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    og = O.TPG(gas)
    c = air_cells(400, seed=11, T_lo=200.0, T_hi=20000.0)
    n = 400
    Y = c["massf"].reshape(5, n)
    g_pT = ch.tpg_update("pT", c["massf"], p=c["p"], T=c["T"])
    T0 = c["T"] * (1.0 + 0.2 * np.sin(np.arange(n)))
    g_ru = ch.tpg_update("rhou", c["massf"], rho=c["rho"], T=T0, u=c["u"])
    g_rT = ch.tpg_update("rhoT", c["massf"], rho=c["rho"], T=c["T"])
    ref = {m: {k: np.zeros(n) for k in ("rho", "p", "T", "u", "a")} for m in ("pT", "rhou", "rhoT")}
    for i in range(n):
        for mode, kw in (("pT", dict(p=c["p"][i], T=c["T"][i])),
                         ("rhou", dict(rho=c["rho"][i], T=T0[i], u=c["u"][i])),
                         ("rhoT", dict(rho=c["rho"][i], T=c["T"][i]))):
            st = og.update(mode, Y[:, i], **kw)
            for k in ref[mode]:
                ref[mode][k][i] = st[k]
    for mode, g in (("pT", g_pT), ("rhou", g_ru), ("rhoT", g_rT)):
        assert g["failed"] == 0, mode
        errs = {k: rel_err(g[k], ref[mode][k]) for k in ("rho", "p", "T", "u", "a")}
        print(mode, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert max(errs.values()) <= TOL, (mode, errs)


def test_tpg_rhou_guess_offsets():
    """The rhou update's Newton iteration from guesses at set distances from the root:
    inside the bracket by a wide margin (the bracket signs then come from the first
    evaluation and a monotonicity bound), near its ends (499 and 501 K off, where the bound
    cannot decide and the end points are evaluated), beyond it (2000 K: the reference's
    second pass over [10, 1e5] K), and at 12 K (the bracket clamped to [10, 1010] K), against
    the oracle's ThermallyPerfectGas, which always evaluates the ends."""
    from oracle import oracle as O
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    og = O.TPG(gas)
    base = air_cells(64, seed=23, T_lo=300.0, T_hi=12000.0)
    offsets = np.array([0.0, 0.5, -0.5, 10.0, -10.0, 300.0, -300.0, 499.0, -499.0, 501.0, -501.0, 2000.0, -2000.0])
    n0 = base["T"].size
    Y0 = base["massf"].reshape(5, n0)
    rows = []
    for k in range(n0):
        for d in offsets:
            rows.append((k, max(base["T"][k] + d, 12.0)))
    n = len(rows)
    Y = np.zeros((5, n))
    rho, u, Tg = np.zeros(n), np.zeros(n), np.zeros(n)
    for i, (k, tg) in enumerate(rows):
        Y[:, i], rho[i], u[i], Tg[i] = Y0[:, k], base["rho"][k], base["u"][k], tg
    assert np.sum(Tg == 12.0) > 10  # guesses clamped near the floor: bracket [10, 1010] K
    g = ch.tpg_update("rhou", Y.reshape(-1), rho=rho, T=Tg, u=u)
    ref = {k: np.zeros(n) for k in ("p", "T", "u")}
    for i in range(n):
        st = og.update("rhou", Y[:, i], rho=rho[i], T=Tg[i], u=u[i])
        for k in ref:
            ref[k][i] = st[k]
    assert g["failed"] == 0
    errs = {k: rel_err(g[k], ref[k]) for k in ("p", "T")}
    errs["u"] = rel_err(g["u"], ref["u"], np.abs(ref["u"]) + g["p"] / rho)
    print("rhou from offset guesses", n, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("persistent", [0, 1])
def test_air_alpha_qss_cold_then_warm(persistent):
    """Two chemistry steps of 4096 uneven-stiffness air cells: the first from dt_chem = -1
    (step estimate), the second from the suggested dt_chem of the first."""
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    ch.set_tuning(persistent=persistent)
    n = 4096
    c = air_cells(n, seed=5)
    dtc = np.full(n, -1.0)
    g1 = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, DT)
    s1 = ch.substeps(n)
    o1 = _oracle_step(gas, mech, c, dtc)
    _compare(g1, o1, f"cold persistent={persistent}", s1)
    assert g1["failed"] == o1["failed"]
    assert g1["substeps_total"] == int(o1["substeps"].sum())
    assert g1["substeps_max"] == int(o1["substeps"].max())
    assert ch.stats()["attempts_total"] == int(o1["attempts"].sum())
    c2 = dict(rho=c["rho"], T=o1["T"], u=o1["u"], p=o1["p"], massf=o1["massf"])
    g2 = ch.step(c2["rho"], c2["T"], c2["u"], c2["p"], c2["massf"], o1["dt_chem"], DT)
    o2 = _oracle_step(gas, mech, c2, o1["dt_chem"])
    _compare(g2, o2, f"warm persistent={persistent}", ch.substeps(n))


@pytest.mark.parametrize("persistent", [1, 0])
def test_air_alpha_qss_tight_coupling(persistent):
    """tightTempCoupling (chemistry_update.d:265-269) on the device: after every subcycle the
    mass fractions, T(u) by Newton from the previous T and the rate constants at the new T.
    4096 uneven air cells over the flow step and over a 100 times longer interval, against
    the oracle (which tests/test_oracle_chem.py pins to mass action at the final temperature
    of an adiabatic reactor); the coupling changes the result.

    The coupled update is ill-conditioned in the reference's own arithmetic: each subcycle
    ends in a Newton T(u) started at the previous root, whose first correction is the
    subcycle's small change of T and often sits near the stopping test |dT| < 1e-6 K, and
    the iteration leaves u at the energy of its last evaluated temperature, which becomes the
    next target (therm_perf_gas_mix_eos.d:61-151).  A last-bit change decides some of those
    tests the other way and the carried u keeps the shift: the oracle moves by up to 2e-6
    in T when its input T moves by 1e-15.  The device's libm differs from the host's in the
    last bit, so the bar is the oracle's own sensitivity (_compare_conditioned).  Without the
    coupling the same cells agree at 1e-12 (test_air_alpha_qss_cold_then_warm)."""
    gas, mech = _air(tightTempCoupling=1)
    gas0, mech0 = _air()
    ch = Chemistry(gas, mech)
    ch0 = Chemistry(gas0, mech0)
    ch.set_tuning(persistent=persistent)
    n = 4096
    c = air_cells(n, seed=21)
    dtc = np.full(n, -1.0)
    for dt in (DT, 100.0 * DT):
        g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, dt)
        o = _oracle_step(gas, mech, c, dtc, dt)
        label = f"tight dt={dt:g} persistent={persistent}"
        _compare_conditioned(g, o, _perturbed_steps(gas, mech, c, dtc, dt), label, ch.substeps(n))
        g0 = ch0.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, dt)
        ok = (g["status"] == 0) & (g0["status"] == 0)
        moved = np.abs(g["T"] - g0["T"]) > 1e-6 * g0["T"]
        print(f"tight dt={dt:g}: T differs from the uncoupled step in {int(moved[ok].sum())} of {int(ok.sum())} cells",
              flush=True)
        assert moved[ok].sum() > 0.25 * ok.sum()


def test_tight_coupling_equilibrium_single_reaction():
    """The adiabatic O2 + M <=> O + O + M reactor of the oracle's mass-action pin, on the
    device: the oracle's final temperature and composition (within the oracle's own
    sensitivity, see the test above), the same substeps, and mass action at the device's own
    final temperature."""
    from oracle import oracle as O
    gd, md = gdata.load("tpg-air-5sp"), dict(gdata.load("park-air-5sp-1T"))
    md["reactions"] = [r for r in md["reactions"] if r["equation"].startswith("O2 + M")]
    md["config"] = dict(md["config"], tightTempCoupling=True)
    gas, mech = gas_record(gd), mech_record(md)
    M = np.array(gd["M"])
    T0 = np.array([4500.0, 4000.0, 6000.0])
    Y = np.zeros(15)
    Y[3:6] = 1.0  # species-major: O2 = 1 in every cell
    rho = np.array([2.0e3, 1.0e4, 5.0e2]) / (R_U / M[1] * T0)
    st = [O.TPG(gas).update("rhoT", [0, 1.0, 0, 0, 0], rho=r, T=t) for r, t in zip(rho, T0)]
    c = dict(rho=rho, T=T0, u=np.array([s["u"] for s in st]), p=np.array([s["p"] for s in st]), massf=Y)
    dtc = np.full(3, -1.0)
    ch = Chemistry(gas, mech)
    g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, 1.0)
    o = _oracle_step(gas, mech, c, dtc, 1.0)
    assert np.all(o["status"] == 0)
    _compare_conditioned(g, o, _perturbed_steps(gas, mech, c, dtc, 1.0), "tight equilibrium", ch.substeps(3))
    assert np.all(g["T"] < T0 - 500.0)
    # the device's own result obeys mass action at its final temperature
    import math
    Yg = g["massf"].reshape(5, 3)
    for i in range(3):
        conc = Yg[:, i] * rho[i] / M
        x_o, x_o2 = conc[4] / (conc[1] + conc[4]), conc[1] / (conc[1] + conc[4])
        Tf = g["T"][i]
        lhs = x_o ** 2 / x_o2 * (conc[1] + conc[4]) * R_U * Tf / 101.325e3
        kp = math.exp(-(2.0 * _g_molar(gd["thermo"][4], Tf) - _g_molar(gd["thermo"][1], Tf)) / (R_U * Tf))
        assert abs(lhs - kp) <= 2.0e-3 * max(lhs, kp), (i, Tf, lhs, kp)


def test_air_rkf_matches_oracle():
    gas, mech = _air(method="rkf", errTol=1.0e-6)
    ch = Chemistry(gas, mech)
    n = 1024
    c = air_cells(n, seed=8, T_lo=2000.0, T_hi=9000.0)
    dtc = np.full(n, -1.0)
    g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, DT)
    o = _oracle_step(gas, mech, c, dtc)
    _compare(g, o, "rkf", ch.substeps(n), dtc_tol=DTCHEM_TOL_RKF)


def test_h2_i2_known_answer_and_alpha_qss():
    """chemistry_update.d:739-750: H2 + I2 <=> 2 HI at 700 K, 60 000 s with the RKF step
    gives [HI] = 7.14201983840 (1e-9); the alpha-QSS update of the same cell matches the
    oracle."""
    from oracle import oracle as O
    gas = gas_record(_json("H2-I2-HI.json"))
    md = _json("H2-I2-inp.json")
    M = np.array(_json("H2-I2-HI.json")["M"])
    T, c0 = 700.0, 4.54
    p = 2.0 * c0 * R_U * T
    molef = np.array([0.5, 0.5, 0.0])
    massf = molef * M / np.sum(molef * M)
    og = O.TPG(gas)
    st = og.update("pT", massf, p=p, T=T)
    cell = dict(rho=np.array([st["rho"]]), T=np.array([T]), u=np.array([st["u"]]), p=np.array([p]),
                massf=massf.copy())
    for method in ("rkf", "alpha-qss"):
        mech = mech_record(md, method=method)
        ch = Chemistry(gas, mech)
        g = ch.step(cell["rho"], cell["T"], cell["u"], cell["p"], cell["massf"], np.array([200.0]), 60_000.0)
        o = O.chem_step_batch(gas, mech, cell["rho"], cell["T"], cell["u"], cell["p"], cell["massf"],
                              np.array([200.0]), 60_000.0)
        _compare(g, o, f"H2-I2 {method}", ch.substeps(1), dtc_tol=DTCHEM_TOL_RKF if method == "rkf" else DTCHEM_TOL)
        if method == "rkf":
            conc_hi = g["massf"][2] * st["rho"] / M[2]
            assert abs(conc_hi - 7.14201983840) <= 1.0e-9 * 7.14201983840


def test_frozen_invalid_and_empty_cells():
    """Cells at or below T_frozen are left unchanged with status 0; a non-positive or NaN
    density fails both attempts (status -1, dt_chem = -1, state unchanged); n = 0 is a no-op."""
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    n = 256
    c = air_cells(n, seed=2)
    c["T"][:32] = np.linspace(200.0, 300.0, 32)
    c["rho"][40:48] = -1.0
    c["rho"][48] = 0.0
    c["rho"][49] = np.nan
    dtc = np.full(n, 3.0e-9)
    g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, DT)
    o = _oracle_step(gas, mech, c, dtc)
    assert np.all(g["status"][:32] == 0) and np.all(g["T"][:32] == c["T"][:32])
    assert np.all(g["dt_chem"][:32] == 3.0e-9)
    assert np.all(g["status"][40:50] == -1) and np.all(g["dt_chem"][40:50] == -1.0)
    assert np.array_equal(g["massf"].reshape(5, n)[:, 40:50], c["massf"].reshape(5, n)[:, 40:50])
    # the oracle leaves failed cells' T/u/p/massf untouched as well
    for k in ("T", "u", "p"):
        o[k][40:50] = c[k][40:50]
    keep = np.ones(n, bool)
    keep[40:50] = False
    _compare({k: (v[keep] if k != "massf" else v.reshape(5, n)[:, keep]) for k, v in g.items()
              if k in ("T", "u", "p", "massf", "dt_chem", "status")},
             {k: (v[keep] if k != "massf" else v.reshape(5, n)[:, keep]) for k, v in o.items()
              if k in ("T", "u", "p", "massf", "dt_chem", "status")}, "frozen+invalid")
    e = ch.step(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), DT)
    assert e["failed"] == 0 and e["T"].size == 0


def test_queue_variants_identical_and_autotune():
    """Every launch shape of the integration kernel gives bit-identical cells; the
    autotuner times them on device arrays without changing its inputs."""
    import torch
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    n = 20000
    c = air_cells(n, seed=9)
    dtc = np.full(n, -1.0)
    ref = ref_sub = None
    # heavy_ratio 1 and 4 put many of these cells (no substep history, so the first substep
    # is the estimate) on the heavy list; 0 keeps the list empty
    for tu in (dict(persistent=0), dict(persistent=1, chunk=1, refill=1), dict(persistent=1, chunk=64, refill=8),
               dict(persistent=1, chunk=256, refill=1, blocks_per_cu=1), dict(persistent=1, heavy_ratio=0),
               dict(persistent=1, chunk=8, heavy_ratio=1), dict(persistent=-1)):
        ch.set_tuning(**{"blocks_per_cu": 0, "heavy_ratio": 4, **tu})
        g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, DT)
        sub = ch.substeps(n)
        if ref is None:
            ref, ref_sub = g, sub
        else:
            for k in ("T", "u", "p", "massf", "dt_chem", "status"):
                assert np.array_equal(g[k], ref[k]), (tu, k)
            assert np.array_equal(sub, ref_sub), (tu, "substeps")
    dev = torch.device("cuda:0")
    t = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in c.items()}
    td = torch.tensor(dtc, device=dev)
    snap = {k: v.clone() for k, v in t.items()}
    res = ch.autotune(n, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(),
                      t["massf"].data_ptr(), td.data_ptr(), DT)
    torch.cuda.synchronize()
    assert len(res) >= 2 and all(r["ms"] > 0 for r in res)
    for k in t:
        assert torch.equal(t[k], snap[k])
    g = ch.step(c["rho"], c["T"], c["u"], c["p"], c["massf"], dtc, DT)
    for k in ("T", "u", "p", "massf", "dt_chem", "status"):
        assert np.array_equal(g[k], ref[k]), ("tuned", k)


def test_device_entry_point_large_batch_sampled():
    """262 144 cells through gdtk_hip_chem_step_device on torch-held device arrays: each
    cell's result is independent of the batch, so a sample of cells checked against the
    oracle pins the whole launch at a size where the queue wraps many times."""
    import torch
    gas, mech = _air()
    ch = Chemistry(gas, mech)
    n = 1 << 18
    c = air_cells(n, seed=21)
    dev = torch.device("cuda:0")
    t = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in c.items()}
    dtc = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    out = {k: torch.empty_like(t[k]) for k in ("T", "u", "p", "massf")}
    dto = torch.empty_like(dtc)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ch.step_device(n, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(),
                   t["massf"].data_ptr(), dtc.data_ptr(), DT, out["T"].data_ptr(), out["u"].data_ptr(),
                   out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(), stream)
    torch.cuda.synchronize()
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(n, 3000, replace=False))
    Y = c["massf"].reshape(5, n)
    sub = {k: c[k][idx] for k in ("rho", "T", "u", "p")}
    sub["massf"] = np.ascontiguousarray(Y[:, idx]).reshape(-1)
    o = _oracle_step(gas, mech, sub, np.full(idx.size, -1.0))
    g = {k: out[k].cpu().numpy() for k in out}
    g = dict(T=g["T"][idx], u=g["u"][idx], p=g["p"][idx],
             massf=np.ascontiguousarray(g["massf"].reshape(5, n)[:, idx]).reshape(-1),
             dt_chem=dto.cpu().numpy()[idx], status=st.cpu().numpy()[idx])
    _compare(g, o, "device 2^18 sampled", ch.substeps(n)[idx])


def test_reactor_n2_reference_and_oracle():
    """The reference's N2 recombination reactor (examples/eilmer/2D/reactor-n2,
    reactor-test.rb:40-55): 200 operator-split increments of 1e-6 s of the batched reactor
    on 256 copies of the cell end at p = 1.494e5 Pa, T = 6381.7 K, Y = (0.8760, 0.1240)
    within 1e-3, and equal the oracle's increments at the 1e-12 bar after every step."""
    from oracle import oracle as O
    gas = gas_record(gdata.load("tpg-nitrogen-2sp"))
    mech = mech_record(gdata.load("nitrogen-2sp-2r"))
    ch = Chemistry(gas, mech)
    n = 256
    s = O.TPG(gas).update("pT", np.array([0.8, 0.2]), p=1.0e5, T=4000.0)
    rho = np.full(n, s["rho"])
    T, u, p = np.full(n, 4000.0), np.full(n, s["u"]), np.full(n, 1.0e5)
    Y = np.concatenate([np.full(n, 0.8), np.full(n, 0.2)])   # species-major
    dtc = np.full(n, -1.0)
    oT, ou, op, oY, odtc = T.copy(), u.copy(), p.copy(), Y.copy(), dtc.copy()
    worst = 0.0
    for _ in range(200):
        g = ch.step(rho, T, u, p, Y, dtc, 1.0e-6)
        o = O.chem_step_batch(gas, mech, rho, oT, ou, op, oY, odtc, 1.0e-6)
        assert np.all(g["status"] == 0) and o["failed"] == 0
        T, u, p, Y, dtc = g["T"], g["u"], g["p"], g["massf"], g["dt_chem"]
        oT, ou, op, oY, odtc = o["T"], o["u"], o["p"], o["massf"], o["dt_chem"]
        worst = max(worst, rel_err(T, oT), rel_err(p, op), rel_err(Y, oY, np.ones_like(Y)))
    print("reactor n2: worst per-step parity", worst, "final", p[0], T[0], Y[0], Y[n], flush=True)
    assert worst <= TOL
    refs = {"p": 1.494e5, "T": 6381.7, "Y_N2": 0.8760, "Y_N": 0.1240}
    got = {"p": p[0], "T": T[0], "Y_N2": Y[0], "Y_N": Y[n]}
    for k in refs:
        assert abs(got[k] - refs[k]) / refs[k] < 1.0e-3, (k, got[k], refs[k])
