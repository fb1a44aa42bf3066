"""GPU parity of the Eilmer 5 (src/lmr) numerics of the flux path against the oracle.

The library's lmr mode (gdtk_hip_sim_set_numerics) scales the van Albada epsilon by the
cell values and the cell size (lmr/onedinterp.d:147-151) and takes AUSMDV's common sound
speed as a smooth maximum (lmr/fluxcalc.d).  The rippled box (pressure ripples of 1e-9)
makes the reconstruction depend on the epsilon scaling: there the two numerics differ by
about 1e-8 relative in p after a few steps, so agreement with the oracle's lmr mode at the
1e-12 bar, and disagreement with its Eilmer 4 mode, shows which numerics ran.  The 3-D
Sod tube is the reference's lmr integration test (examples/lmr/3D/sod-shock-tube/sg,
test_sod_sg.py: 75 +/- 5 steps, probes within 1 %).
"""
import numpy as np
import pytest

from gdtk_amd import cases
from tests.parity import TOL, field_errors, rel_err

pytestmark = pytest.mark.gpu

NAMES = ("rho", "vx", "vy", "p", "T")


def _gpu(case, flux, numerics):
    from gdtk_amd import Simulation
    g = Simulation(case.dims, flux=flux, numerics=numerics)
    g.add_block(case)
    return g


def _oracle(case, flux, numerics):
    from oracle import oracle as O
    o = O.Sim(O.IdealGas(), case.dims, flux, numerics=numerics)
    o.add_block(case)
    return o


@pytest.mark.parametrize("flux", ["ausmdv", "efm", "adaptive"])
def test_rippled_box_matches_lmr_oracle(flux):
    case = cases.rippled_box()
    g = _gpu(case, flux, "lmr")
    o = _oracle(case, flux, "lmr")
    o4 = _oracle(case, flux, "eilmer4")
    _, sg = g.run(6)
    _, so = o.run(6)
    o4.run(6)
    assert sg["dt"] == pytest.approx(so["dt"], rel=1e-12)
    errs = field_errors(lambda k: g.get(0, k), lambda k: o.get(0, k), NAMES)
    other = rel_err(g.get(0, "p"), o4.get(0, "p"))
    print("lmr rippled box", flux, {k: f"{v:.2e}" for k, v in errs.items()}, f"vs eilmer4 p {other:.2e}", flush=True)
    assert all(v <= TOL for v in errs.values()), errs
    assert other > 1.0e3 * TOL, other


def test_eilmer4_mode_is_unchanged_on_rippled_box():
    case = cases.rippled_box()
    g = _gpu(case, "ausmdv", "eilmer4")
    o = _oracle(case, "ausmdv", "eilmer4")
    g.run(6)
    o.run(6)
    errs = field_errors(lambda k: g.get(0, k), lambda k: o.get(0, k), NAMES)
    assert all(v <= TOL for v in errs.values()), errs


def test_lmr_sod_tube_reference_run():
    case = cases.sod_3d_tube()
    g = _gpu(case, "ausmdv", "lmr")
    o = _oracle(case, "ausmdv", "lmr")
    _, sg = g.run(600, 0.6e-3)
    _, so = o.run(600, 0.6e-3)
    assert sg["step"] == so["step"] and abs(sg["step"] - 75) < 5
    assert sg["time"] == pytest.approx(so["time"], rel=1e-12)
    errs = field_errors(lambda k: g.get(0, k), lambda k: o.get(0, k), ("rho", "vx", "vy", "vz", "p", "T"))
    print("lmr sod 3-D", sg["step"], {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(v <= TOL for v in errs.values()), errs
    xc = cases.cell_centres(case)[:, 0]
    for x, ref in ((0.78, dict(rho=0.2647, p=30.2e3, T=398.0, vx=293.0)),
                   (0.6, dict(rho=0.4271, p=30.2e3, T=247.0, vx=293.0))):
        i = int(np.argmin(np.abs(xc - x)))
        for k, v in ref.items():
            assert abs(g.get(0, k)[i] - v) / (abs(v) + 1.0) < 0.01, (x, k)


def test_lmr_refuses_multi_species_gas():
    from gdtk_amd import Simulation, hip
    case, rec = cases.species_box(16, 8, nsp=5)
    with pytest.raises(hip.HipError, match="single-species"):
        Simulation(2, flux="ausmdv", tpg=rec, numerics="lmr")
    g = Simulation(2, flux="ausmdv")
    with pytest.raises(ValueError, match="unknown numerics"):
        g.set_numerics("eilmer3")
    with pytest.raises(hip.HipError, match="unknown numerics"):
        hip._check(hip.lib().gdtk_hip_sim_set_numerics(g.h, 7))
