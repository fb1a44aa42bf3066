"""GPU parity of the adaptive EFM/AUSMDV flux calculator and its shock detector: the HIP
library (through its C ABI) against the CPU oracle.

Besides the flow fields at the 1e-12 bar of tests/parity.py, every test compares the cells'
shock indicators exactly: they are 0/1 decisions, so one flipped cell would switch a face's
flux calculator.  The oracle runs with the in-place velocity rotation of interp_l2r2 off,
so that the detector sees the same cell velocities as the GPU (with it on they differ in
the last bits, which can move a decision sitting exactly on a tolerance).  The cases are
chosen so that rounding differences stay at rounding level over the run: the start-up of the
wedge amplifies them about a hundredfold every 20 steps with any flux calculator (the
oracle with and without the rotation differs by 1e-9 after 40 steps), so it is compared over
its first 12 steps.
"""
import numpy as np
import pytest

from gdtk_amd import cases
from tests.parity import TOL, field_errors
from tests.test_oracle_shock import converging_case

pytestmark = pytest.mark.gpu

FS_NAMES = ["rho", "vx", "vy", "vz", "u", "p", "T", "a"]


def _pair(blocks, dims, rec=None, graph=True, flux="adaptive", **kw):
    from gdtk_amd import Simulation
    from oracle import oracle as O
    g = Simulation(dims, flux=flux, tpg=rec, **kw)
    for b in blocks:
        g.add_block(b)
    g.set_graph(graph)
    gas = O.TPG(rec) if rec is not None else O.IdealGas()
    o = O.Sim(gas, dims, flux, mutate_cell_velocities=False, **kw)
    for b in blocks:
        o.add_block(b)
    o._gas = gas
    return g, o


def _check(g, o, nblocks, label, names=FS_NAMES, tol=TOL, shock=True):
    worst = {}
    marked = 0
    for q in range(nblocks):
        errs = field_errors(lambda n: g.get(q, n), lambda n: o.get(q, n), names, g.get_U(q), o.get_U(q, g.ncq),
                            g.dims)
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
        if not shock:
            continue
        Sg, So = g.get(q, "S"), o.get(q, "S")
        assert np.array_equal(Sg, So), f"{label}: block {q}: shock indicators differ in {np.sum(Sg != So)} cells"
        marked += int(So.sum())
    print(label, "marked", marked, {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"{label}: parity above {tol}: {bad}"
    return marked


def _sod_y(nic, njc, amp=0.0, seed=3):
    """Sod with the pressure varied along y, so that every cell carries a real velocity (in a
    resting gas on a distorted grid the velocities are rounding noise)."""
    c = cases.sod_2d(nic, njc)
    xc = cases.cell_centres(c)
    c.p = c.p * (1.0 + 0.3 * np.sin(6.0 * xc[:, 1]))
    return cases.perturb_vertices(c, amp=amp, seed=seed) if amp else c


def test_sod2d_adaptive():
    case = _sod_y(120, 40)
    g, o = _pair([case], 2)
    _, st_g = g.run(40)
    _, st_o = o.run(40)
    assert st_g["step"] == st_o["step"] == 40
    assert st_g["dt"] == pytest.approx(st_o["dt"], rel=1e-12)
    assert _check(g, o, 1, "sod2d adaptive") > 0


def test_wedge_adaptive():
    """The reference's Mach-10 6-degree wedge (adaptive flux), the first 12 steps."""
    case = cases.wedge_2d()
    g, o = _pair([case], 2)
    g.run(12)
    o.run(12)
    assert _check(g, o, 1, "wedge adaptive") > 0


@pytest.mark.parametrize("flux", ["hanel", "adaptive_hanel_ausmdv"])
def test_hanel_calculators(flux):
    """Hanel's splitting alone and as the shock-side calculator of the reference's default
    adaptive_hanel_ausmdv, on the Sod tube (2-D) and the Mach-6 box (3-D)."""
    case = _sod_y(120, 40)
    g, o = _pair([case], 2, flux=flux)
    g.run(30)
    o.run(30)
    _check(g, o, 1, f"sod2d {flux}", shock=flux != "hanel")
    case3 = cases.mach6_box(24, 16, 12, seed=11)
    g, o = _pair([case3], 3, flux=flux)
    g.run(5)
    o.run(5)
    _check(g, o, 1, f"mach6 3d {flux}", shock=flux != "hanel")


def test_mach6_box_3d_adaptive():
    case = cases.mach6_box(24, 16, 12, seed=11)
    g, o = _pair([case], 3)
    g.run(6)
    o.run(6)
    _check(g, o, 1, "mach6 3d adaptive")


def test_cut_sod_adaptive():
    """Sod split into 6 blocks: the shock crosses exchange boundaries, where a ghost cell's
    indicator is the copy taken before this step's detection."""
    case = _sod_y(60, 30, amp=0.2, seed=4)
    blocks = cases.split_block(case, (0, 17, 33, 60), (0, 13, 30))
    g, o = _pair(blocks, 2)
    g.run(60)
    o.run(60)
    assert _check(g, o, len(blocks), "cut sod adaptive") > 0


def test_turned_blocks_adaptive():
    from tests.exchange_cases import reorient
    case = cases.mach6_box(16, 8, 8, seed=5)
    blocks = cases.split_block(case, (0, 8, 16), (0, 8), (0, 8))
    blocks, _ = reorient(blocks, 1, (2, 0, 1), (False, False, False))
    assert any(o != 0 for o in blocks[0].other_orientation + blocks[1].other_orientation)
    g, o = _pair(blocks, 3)
    g.run(5)
    o.run(5)
    _check(g, o, len(blocks), "turned blocks adaptive")


@pytest.mark.parametrize("flux", ["adaptive", "adaptive_hanel_ausmdv", "hanel"])
def test_species_box_adaptive(flux):
    case, rec = cases.species_box(48, 24, nsp=5)
    g, o = _pair([case], 2, rec=rec, flux=flux)
    g.run(10)
    o.run(10)
    names = [n for n in FS_NAMES if n != "u"]
    _check(g, o, 1, f"species box {flux}", names=names, shock=flux != "hanel")


def test_converging_flow_every_face_efm():
    """Every face marked: the adaptive GPU run equals the GPU's own EFM run bit for bit."""
    from gdtk_amd import Simulation
    case = converging_case()
    opts = dict(compression_tolerance=-1.0e-300, shear_tolerance=1.0e300)
    a = Simulation(2, flux="adaptive", shock_detector=opts)
    a.add_block(case)
    b = Simulation(2, flux="efm", fused=False)
    b.add_block(case)
    a.run(8)
    b.run(8)
    assert np.all(a.get(0, "S") == 1.0)
    for n in FS_NAMES:
        assert np.array_equal(a.get(0, n), b.get(0, n)), n


@pytest.mark.parametrize("opts", [dict(damped_outflow=True),
                                  dict(frozen_shock_detector=True, shock_detector_freeze_step=12)])
def test_detector_options(opts):
    case = _sod_y(60, 12, amp=0.2, seed=8)
    case.bc[cases.EAST] = cases.BC_OUTFLOW_SIMPLE_EXTRAPOLATE
    g, o = _pair([case], 2, shock_detector=opts)
    g.run(30)
    o.run(30)
    _check(g, o, 1, f"detector options {opts}")


def test_graph_autotune_and_direct_bitwise():
    """Captured steps, kernel-by-kernel launches and a run interrupted by the autotuner give
    the same fields and indicators bit for bit."""
    from gdtk_amd import Simulation
    case = cases.wedge_2d()
    out = []
    for mode in ("graph", "direct", "autotune"):
        g = Simulation(2, flux="adaptive")
        g.add_block(case)
        g.set_graph(mode != "direct")
        g.run(40)
        if mode == "autotune":
            g.autotune(reps=2)
        g.run(40)
        out.append({n: g.get(0, n) for n in FS_NAMES + ["S"]})
    for other in out[1:]:
        for n in out[0]:
            assert np.array_equal(out[0][n], other[n]), n
