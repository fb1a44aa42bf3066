"""GPU parity of the coupled 2-D Mach-10 cylinder (5-species air, Park 1-T mechanism, flux +
operator-split alpha-QSS chemistry, simcore.d:1619-1638 and fvcell.d:858-972) against the
oracle, on a reduced grid of the same case (cases.cylinder_blocks: 4 blocks joined by
exchange boundaries, supersonic inflow, slip-wall body and symmetry line, simple outflow).

- From the free-stream start: the HIP library and the oracle agree at the 1e-12 bar of
  tests/parity.py after every chunk of steps while the bow shock forms.
- From a restart: the GPU runs long enough for a shock layer of hot, reacting cells, writes
  every block's flow file, and both cold-start from those files (simcore.d:299-336) and run
  on.  This compares the reacting state without the long free-stream history, over which
  the two libms' last-bit differences in exp/log/pow would be amplified by the shock.
"""
import numpy as np
import pytest

from gdtk_amd import cases
from oracle import flow_format as FF
from tests.parity import TOL, field_errors, rel_err

pytestmark = pytest.mark.gpu

AIR5 = ["N2", "O2", "NO", "N", "O"]
NAMES = ("rho", "vx", "vy", "p", "T")
GRID = dict(ni=40, nj_block=10, njb=4)


def _mech():
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import mech_record
    return mech_record(gdata.load("park-air-5sp-1T"))


def _gpu(blocks, rec, mech):
    from gdtk_amd import Simulation
    from gdtk_amd.chem import Chemistry
    g = Simulation(2, flux="ausmdv", tpg=rec, dt_init=1.0e-10, dt_max=1.0e-6)
    for b in blocks:
        g.add_block(b)
    g.set_chemistry(Chemistry(rec, mech))
    return g


def _oracle(blocks, rec, mech):
    from oracle import oracle as O
    gas = O.TPG(rec)
    o = O.Sim(gas, 2, "ausmdv", dt_init=1.0e-10, dt_max=1.0e-6, mutate_cell_velocities=False)
    for b in blocks:
        o.add_block(b)
    o._gas = gas
    o.set_chemistry(O.Chem(rec, mech))
    return o


def _errors(g, o, nblocks):
    worst = {}
    for q in range(nblocks):
        errs = field_errors(lambda k: g.get(q, k), lambda k: o.get(q, k), NAMES)
        u_ref = o.get(q, "u")
        errs["u"] = rel_err(g.get(q, "u"), u_ref, np.abs(u_ref) + o.get(q, "p") / o.get(q, "rho"))
        for s in range(5):
            errs[f"Y{s}"] = rel_err(g.get(q, 300 + s), o.get(q, 300 + s), np.ones(g.cases[q].ncells))
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
    return worst


def test_cylinder_free_stream_start_against_oracle():
    blocks, rec, _ = cases.cylinder_blocks(**GRID)
    mech = _mech()
    g, o = _gpu(blocks, rec, mech), _oracle(blocks, rec, mech)
    for chunk in range(4):
        _, sg = g.run(10)
        _, so = o.run(10)
        assert sg["step"] == so["step"] == 10 * (chunk + 1)
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-12)
        worst = _errors(g, o, len(blocks))
        print(f"cylinder free stream, step {sg['step']}:", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
        assert all(v <= TOL for v in worst.values()), worst
    cs = g.chem_stats()
    assert cs["steps"] == 40
    g.close()


def test_cylinder_restart_reacting_against_oracle(tmp_path):
    blocks, rec, _ = cases.cylinder_blocks(**GRID)
    mech = _mech()
    g0 = _gpu(blocks, rec, mech)
    _, st0 = g0.run(300)
    # a shock layer of hot cells that react: oxygen dissociated behind the bow shock
    T = np.concatenate([g0.get(q, "T") for q in range(len(blocks))])
    YO = np.concatenate([g0.get(q, 304) for q in range(len(blocks))])
    hot = T > 2000.0
    print(f"after 300 steps: t = {st0['time']:.3e} s, {int(hot.sum())} cells above 2000 K, "
          f"max T {T.max():.0f} K, max Y_O {YO.max():.3e}", flush=True)
    assert hot.sum() >= 50 and YO.max() > 1.0e-3
    fns = []
    for q in range(len(blocks)):
        fn = str(tmp_path / f"b{q:04d}.flow")
        g0.write_flow(q, fn, "rawbinary", species=AIR5)
        fns.append(fn)
    g0.close()

    from gdtk_amd import Simulation
    from gdtk_amd.chem import Chemistry
    from oracle import oracle as O
    g = Simulation(2, flux="ausmdv", tpg=rec, dt_init=1.0e-10, dt_max=1.0e-6)
    for b in blocks:
        g.add_block(b)
    gas = O.TPG(rec)
    o = O.Sim(gas, 2, "ausmdv", dt_init=1.0e-10, dt_max=1.0e-6, mutate_cell_velocities=False)
    for b in blocks:
        o.add_block(b)
    o._gas = gas
    names = FF.variable_list(AIR5)
    col = {nm: j for j, nm in enumerate(names)}
    t_file = None
    for q, (b, fn) in enumerate(zip(blocks, fns)):
        t = g.read_flow(q, fn, "rawbinary", species=AIR5)
        r, t2 = FF.read_rawbinary(fn, names, b.ncells)
        assert t2 == t
        t_file = t
        o.set_fs(q, r[:, col["rho"]], r[:, col["vel.x"]], r[:, col["vel.y"]], r[:, col["vel.z"]], r[:, col["p"]],
                 r[:, col["a"]], r[:, col["u"]], r[:, col["T"]], r[:, 15:20], r[:, col["S"]])
    g.set_time(t_file)
    o.set_time(t_file)
    g.set_chemistry(Chemistry(rec, mech))
    o.set_chemistry(O.Chem(rec, mech))
    for chunk in range(3):
        _, sg = g.run(4)
        _, so = o.run(4)
        assert sg["time"] == pytest.approx(so["time"], rel=1e-13)
        worst = _errors(g, o, len(blocks))
        print(f"cylinder restart, step {sg['step']}:", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
        assert all(v <= TOL for v in worst.values()), worst
    cs = g.chem_stats()
    print("chemistry after the restart:", cs, flush=True)
    assert cs["steps"] == 12
    g.close()
