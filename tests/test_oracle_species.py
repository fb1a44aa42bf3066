"""Oracle pins of the multi-species thermally perfect flux path, no GPU.

The reference has no unit test of this path short of whole simulations, so the oracle is
pinned by what the equations fix:
  - a mixture of species that share one molecular mass and one constant Cp is an ideal
    gas: the species bookkeeping (reconstruction of the species densities, mass-fraction
    scaling, species fluxes) and the mixture EOS (Newton T(u), p = rho R T, frozen sound
    speed) must reproduce the ideal-gas oracle's flow;
  - in a closed box every species' mass and the total energy are conserved;
  - a uniform composition stays uniform while the flow moves, up to the effect of the
    van Albada limiter's absolute epsilon (1e-12, onedinterp.d), which makes each species'
    limited slope depend slightly on that species' magnitude.
"""
import numpy as np
import pytest

from gdtk_amd import cases
from oracle import oracle as O

FIELDS = ("rho", "p", "T", "vx", "vy", "u", "a")


def _vols(sim, q):
    return sim.get(q, "volume")


def test_constant_cp_mixture_matches_ideal_gas():
    base = cases.sod_2d(40, 20)
    xc = cases.cell_centres(base)
    base.p = base.p * (1.0 + 0.2 * np.sin(8.0 * xc[:, 1]))
    base.massf = np.tile([0.3, 0.7], (base.ncells, 1))
    base.inflow_massf = np.tile([0.3, 0.7], (6, 1))
    tpg = O.TPG(cases.constant_cp_gas(2, M=0.02896, cp_on_R=3.5))
    ideal = O.IdealGas(0.02896, 1.4)
    out = {}
    for name, gas in (("tpg", tpg), ("ideal", ideal)):
        s = O.Sim(gas, 2, "ausmdv", mutate_cell_velocities=False)
        s.add_block(base)
        s.run(12)
        out[name] = {f: s.get(0, f) for f in FIELDS}
        s.close()
    for f in FIELDS:
        a, b = out["tpg"][f], out["ideal"][f]
        scale = np.max(np.abs(b)) if f in ("vx", "vy") else np.abs(b)
        err = np.max(np.abs(a - b) / np.maximum(scale, 1e-300))
        assert err < 1e-8, (f, err)


@pytest.mark.parametrize("nsp,flux", [(5, "ausmdv"), (2, "efm")])
def test_closed_box_conserves_species_and_energy(nsp, flux):
    case, rec = cases.species_box(24, 14, nsp=nsp, closed=True)
    gas = O.TPG(rec)
    s = O.Sim(gas, 2, flux)
    s.add_block(case)
    V = _vols(s, 0)

    def totals():
        U = s.get_U(0, 4 + nsp)
        return np.array([np.sum(U[:, 4 + i] * V) for i in range(nsp)] + [np.sum(U[:, 0] * V), np.sum(U[:, 3] * V)])

    t0 = totals()
    s.run(20)
    t1 = totals()
    scale = np.maximum(np.abs(t0), np.max(np.abs(t0[:nsp])))
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale), (t1 - t0) / scale
    # the mass fractions moved (the test is not vacuous)
    Y = np.stack([s.get(0, 300 + i) for i in range(nsp)], axis=1)
    assert np.max(np.abs(Y - case.massf)) > 1e-4
    s.close()


def test_uniform_composition_stays_uniform():
    case, rec = cases.species_box(30, 16, nsp=5)
    Y0 = np.array([0.6, 0.1, 0.05, 0.1, 0.15])
    case.massf = np.tile(Y0, (case.ncells, 1))
    case.inflow_massf = np.tile(Y0, (6, 1))
    gas = O.TPG(rec)
    s = O.Sim(gas, 2, "ausmdv")
    s.add_block(case)
    s.run(10)
    Y = np.stack([s.get(0, 300 + i) for i in range(5)], axis=1)
    assert np.max(np.abs(Y - Y0)) < 1e-6
    assert np.max(np.abs(s.get(0, "T") - case.T)) > 1.0   # the flow did move
    s.close()


@pytest.mark.parametrize("flux", ["ausmdv", "efm"])
def test_mirrored_flow_gives_mirrored_species(flux):
    """A composition blob carried by a flow along -x is the mirror image of the same blob
    carried along +x: the species fluxes must be upwinded from the side the mass comes from.
    EFM, as coded in the reference, treats its two sides differently where their gas
    constants differ (fluxcalc.d: the right state's RT is p/rho), so its case uses a
    mixture of species with one gas constant."""
    case, rec = cases.species_box(32, 12, nsp=5, closed=True)
    if flux == "efm":
        rec = cases.constant_cp_gas(5, M=0.02896, cp_on_R=3.5)
    case.vel[:, 0] = 300.0
    case.vel[:, 1] = 0.0
    nic, njc = case.nic, case.njc

    def mirror(a):
        return a.reshape(njc, nic, *a.shape[1:])[:, ::-1].reshape(a.shape)

    m = cases.BlockCase(2, nic, njc, 1, case.vertices, mirror(case.p), mirror(case.T), mirror(case.vel).copy(),
                        list(case.bc), case.inflow.copy())
    m.vel[:, 0] *= -1.0
    m.massf = mirror(case.massf)
    m.inflow_massf = case.inflow_massf
    Ys = []
    for c in (case, m):
        s = O.Sim(O.TPG(rec), 2, flux, mutate_cell_velocities=False)
        s.add_block(c)
        s.run(15)
        Ys.append(np.stack([s.get(0, 300 + i) for i in range(5)], axis=1))
        s.close()
    assert np.max(np.abs(Ys[0] - case.massf)) > 1e-3       # the blob moved
    assert np.max(np.abs(mirror(Ys[1]) - Ys[0])) < 1e-10
