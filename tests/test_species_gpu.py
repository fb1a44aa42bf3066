"""GPU parity of the multi-species thermally perfect flux path: the HIP library (through its
C ABI) against the CPU oracle per step, at the 1e-12 bar of tests/parity.py, for flow-state
fields, mass fractions and conserved quantities (species densities included)."""
import numpy as np
import pytest

from gdtk_amd import cases
from tests.parity import TOL, field_errors, rel_err

pytestmark = pytest.mark.gpu

FS_NAMES = ["rho", "vx", "vy", "vz", "u", "p", "T", "a"]


def _gpu(blocks, rec, flux="ausmdv", graph=True):
    from gdtk_amd import Simulation
    g = Simulation(2, flux=flux, tpg=rec)
    for b in blocks:
        g.add_block(b)
    g.set_graph(graph)
    return g


def _oracle(blocks, rec, flux="ausmdv", mutate=True):
    from oracle import oracle as O
    gas = O.TPG(rec)
    o = O.Sim(gas, 2, flux, mutate_cell_velocities=mutate)
    for b in blocks:
        o.add_block(b)
    o._gas = gas
    return o


def _check(g, o, nblocks, label, tol=TOL):
    """The internal energy of a mixture with formation enthalpies crosses zero (its datum is
    arbitrary), so u is compared relative to |u| + p/rho, the local energy scale R T."""
    worst = {}
    names = [n for n in FS_NAMES if n != "u"]
    for q in range(nblocks):
        errs = field_errors(lambda n: g.get(q, n), lambda n: o.get(q, n), names, g.get_U(q), o.get_U(q, g.ncq), 2)
        u_ref = o.get(q, "u")
        errs["u"] = rel_err(g.get(q, "u"), u_ref, np.abs(u_ref) + o.get(q, "p") / o.get(q, "rho"))
        for s in range(g.nsp):
            errs[f"Y{s}"] = rel_err(g.get(q, 300 + s), o.get(q, 300 + s), np.ones(g.cases[q].ncells))
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
    print(label, {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"{label}: parity above {tol}: {bad}"


@pytest.mark.parametrize("nsp,flux", [(5, "ausmdv"), (5, "efm"), (2, "ausmdv")])
def test_species_box_against_oracle(nsp, flux):
    case, rec = cases.species_box(48, 24, nsp=nsp)
    g, o = _gpu([case], rec, flux), _oracle([case], rec, flux)
    _, st_g = g.run(12)
    _, st_o = o.run(12)
    assert st_g["step"] == st_o["step"] == 12
    assert st_g["dt"] == pytest.approx(st_o["dt"], rel=1e-12)
    _check(g, o, 1, f"species box nsp={nsp} {flux}")


def test_cut_species_box_against_oracle():
    case, rec = cases.species_box(40, 20, nsp=5)
    blocks = cases.split_block(case, (0, 13, 40), (0, 9, 20))
    g, o = _gpu(blocks, rec), _oracle(blocks, rec)
    g.run(8)
    o.run(8)
    _check(g, o, len(blocks), "cut species box")


def test_closed_box_conserves_species():
    case, rec = cases.species_box(64, 32, nsp=5, closed=True)
    g = _gpu([case], rec)
    V = g.get(0, "volume")

    def totals():
        U = g.get_U(0)
        return np.array([np.sum(U[:, 4 + i] * V) for i in range(5)] + [np.sum(U[:, 0] * V), np.sum(U[:, 3] * V)])

    t0 = totals()
    g.run(30)
    t1 = totals()
    scale = np.maximum(np.abs(t0), np.max(np.abs(t0[:5])))
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale), (t1 - t0) / scale


def test_graph_and_direct_bitwise_species():
    case, rec = cases.species_box(40, 20, nsp=5)
    blocks = cases.split_block(case, (0, 21, 40), (0, 20))
    out = []
    for graph in (True, False):
        g = _gpu(blocks, rec, graph=graph)
        g.run(6)
        out.append([{n: g.get(q, n) for n in FS_NAMES + [300, 301, 302, 303, 304]} for q in range(len(blocks))])
        g.close()
    for q in range(len(blocks)):
        for n in out[0][q]:
            assert np.array_equal(out[0][q][n], out[1][q][n]), (q, n)


def test_species_step_retry():
    """A dt far above the CFL limit fails decoding: the attempt is repeated from the saved
    step-start state at dt * 0.2, and the result equals the oracle's step at the final dt."""
    case, rec = cases.species_box(32, 16, nsp=5)
    g = _gpu([case], rec)
    dt_ok = 0.4 * g.block_dt(0, 1.0, 0.5, False)[0]
    st = g.step_fixed_dt(60.0 * dt_ok)
    print("species retry:", st, flush=True)
    assert st["attempts"] >= 2
    o = _oracle([case], rec)
    o.step_fixed_dt(st["dt"])
    _check(g, o, 1, "species retry")


def test_reacting_species_box_against_oracle():
    """Flux steps with operator-split finite-rate chemistry (5-species air, Park 1-T
    mechanism, alpha-QSS): a hot, partly dissociated blob reacts while it is carried
    downstream; the HIP run equals the oracle's per step."""
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, mech_record
    from oracle import oracle as O
    case, rec = cases.species_box(40, 20, nsp=5)
    mech = mech_record(gdata.load("park-air-5sp-1T"))
    g = _gpu([case], rec)
    ch = Chemistry(rec, mech)
    g.set_chemistry(ch)
    o = _oracle([case], rec)
    oc = O.Chem(rec, mech)
    o.set_chemistry(oc)
    Y0 = g.get_massf(0)
    _, st_g = g.run(8)
    _, st_o = o.run(8)
    assert st_g["dt"] == pytest.approx(st_o["dt"], rel=1e-12)
    cs = g.chem_stats()
    print("reacting box:", st_g, cs, flush=True)
    assert cs["steps"] == 8
    # the chemistry changed the composition beyond what the flow alone would do: O atoms
    # formed in the blob
    assert np.max(g.get_massf(0)[:, 4] - Y0[:, 4]) > 1e-4
    _check(g, o, 1, "reacting species box")


def test_coupled_surrogate_with_zero_tolerance_is_the_integrator():
    """The coupled cylinder with the surrogate in the chemistry step: at a disagreement
    tolerance of zero every reacting cell goes to the integrator, so the run equals the
    run with the integrator alone bit for bit (the surrogate path's wiring: gather, queue,
    integrator, scatter, per-cell dt_chem)."""
    from gdtk_amd import Simulation
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, mech_record
    from gdtk_amd.surrogate import with_tolerance, train_air_surrogate
    blocks, rec, _ = cases.cylinder_blocks(ni=48, nj_block=12, njb=4)
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(rec, mech_record(gdata.load("park-air-5sp-1T")))
    sur, _ = train_air_surrogate(ch, gd, n_train=1 << 14, epochs=(1,), lrs=(1e-3,), batch=2048)
    sur0 = with_tolerance(sur, 0.0, x_abs_max=1.0e30)
    out = []
    for s in (None, sur0):
        g = Simulation(2, tpg=rec, dt_init=1e-10, dt_max=1e-6)
        for b in blocks:
            g.add_block(b)
        g.set_chemistry(ch, s)
        g.run(120)
        out.append([{n: g.get(q, n) for n in ("rho", "T", "p", 300, 304)} for q in range(len(blocks))])
        cs = g.chem_stats()
        print("coupled", "surrogate" if s else "integrator", cs, flush=True)
        if s is not None:
            assert cs["surrogate_queued"] > 0
        g.close()
    for q in range(len(blocks)):
        for n in out[0][q]:
            assert np.array_equal(out[0][q][n], out[1][q][n]), (q, n)
