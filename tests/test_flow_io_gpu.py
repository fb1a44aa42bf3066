"""GPU tests of the flow-solution files: what the library writes from device state, and the
restart from a file, against the oracle.

- A written file, parsed by the oracle's restatement of the format, holds exactly the
  device's fields (rho, vel, p, a, u, T, S, mass fractions, volume), the oracle's cell
  centroids and the caller's mu, k; its text is the restatement's text of those values.
- A restart (read + cold start, simcore.d:299-336) continues like the oracle restarted from
  the same file, at the 1e-12 bar of tests/parity.py; gziptext and rawbinary restarts are
  bit-identical; a restart stays within 1e-10 of the uninterrupted run (the cold start
  re-encodes U from the decoded state, a rounding-level change).
"""
import gzip

import numpy as np
import pytest

from gdtk_amd import cases
from oracle import flow_format as FF
from tests.parity import TOL, field_errors, rel_err

pytestmark = pytest.mark.gpu

AIR5 = ["N2", "O2", "NO", "N", "O"]
NAMES = ("rho", "vx", "vy", "vz", "p", "T")


def _sod_y(nic=48, njc=16):
    c = cases.sod_2d(nic, njc)
    yc = cases.cell_centres(c)[:, 1]
    c.p = c.p * (1.0 + 0.3 * np.sin(6.0 * yc))
    return cases.perturb_vertices(c, amp=0.15, seed=2)


def _gpu(case, flux="adaptive", rec=None):
    from gdtk_amd import Simulation
    g = Simulation(case.dims, flux=flux, tpg=rec)
    g.add_block(case)
    return g


@pytest.mark.parametrize("fmt", ["gziptext", "rawbinary"])
def test_written_file_holds_device_state(tmp_path, fmt):
    from oracle import oracle as O
    case = _sod_y()
    g = _gpu(case)
    n = case.ncells
    mu = np.linspace(1.0e-5, 2.0e-5, n)
    k = np.linspace(0.02, 0.03, n)
    g.set_transport(0, mu, k)
    _, st = g.run(15)
    fn = str(tmp_path / ("b0.flow.gz" if fmt == "gziptext" else "b0.flow"))
    g.write_flow(0, fn, fmt, label="block-0")
    names = FF.variable_list(["air"])
    o = O.Sim(O.IdealGas(), 2, "adaptive")
    o.add_block(case)
    pos = np.stack([o.get(0, "pos_x"), o.get(0, "pos_y"), o.get(0, "pos_z")], axis=1)
    expect = np.zeros((n, len(names)))
    expect[:, 0:3] = pos
    expect[:, 3] = g.get(0, "volume")
    for j, f in zip((4, 5, 6, 7, 8, 9), ("rho", "vx", "vy", "vz", "p", "a")):
        expect[:, j] = g.get(0, f)
    expect[:, 10], expect[:, 11] = mu, k
    expect[:, 14] = g.get(0, "S")
    expect[:, 15] = 1.0
    expect[:, 16], expect[:, 17] = g.get(0, "u"), g.get(0, "T")
    assert expect[:, 14].sum() > 0
    if fmt == "gziptext":
        with gzip.open(fn, "rt", newline="") as f:
            got = f.read()
        assert got == FF.text(expect, names, "block-0", st["time"], 2, case.nic, case.njc, 1)
    else:
        assert open(fn, "rb").read() == FF.rawbinary(expect, names, "block-0", st["time"], 2, case.nic, case.njc, 1)


def _restart_pair(case, fn, fmt, species, steps, rec=None, chem=None, flux="adaptive"):
    """GPU and oracle simulations both cold-started from the file."""
    from gdtk_amd import Simulation
    from oracle import oracle as O
    g = Simulation(case.dims, flux=flux, tpg=rec)
    g.add_block(case)
    t = g.read_flow(0, fn, fmt, species=species)
    g.set_time(t)
    gas = O.TPG(rec) if rec is not None else O.IdealGas()
    o = O.Sim(gas, case.dims, flux, mutate_cell_velocities=False)
    o.add_block(case)
    o._gas = gas
    names = FF.variable_list(species or ["air"])
    if fmt == "gziptext":
        r, t2, _ = FF.read_gziptext(fn, names, case.dims, case.nic, case.njc, case.nkc)
    else:
        r, t2 = _raw_records(fn, len(names), case.ncells)
    assert t2 == t
    nsp = len(species or ["air"])
    col = {nm: j for j, nm in enumerate(names)}
    mf = r[:, 15:15 + nsp] if nsp > 1 else None
    o.set_fs(0, r[:, col["rho"]], r[:, col["vel.x"]], r[:, col["vel.y"]], r[:, col["vel.z"]], r[:, col["p"]],
             r[:, col["a"]], r[:, col["u"]], r[:, col["T"]], mf, r[:, col["S"]])
    o.set_time(t2)
    if chem is not None:
        g.set_chemistry(chem[0])
        o.set_chemistry(chem[1])
    return g, o


def _raw_records(fn, nv, ncell):
    data = open(fn, "rb").read()
    p = len(b"structured_grid_flow 1.0")
    lab = int(np.frombuffer(data[p:p + 4], "<i4")[0])
    p += 4 + lab
    t = float(np.frombuffer(data[p:p + 8], "<f8")[0])
    p += 8
    nvar = int(np.frombuffer(data[p:p + 4], "<i4")[0])
    p += 4
    for _ in range(nvar):
        ln = int(np.frombuffer(data[p:p + 4], "<i4")[0])
        p += 4 + ln
    p += 16
    assert nvar == nv
    return np.frombuffer(data[p:], "<f8").reshape(ncell, nv), t


@pytest.mark.parametrize("fmt", ["gziptext", "rawbinary"])
def test_restart_matches_oracle_ideal(tmp_path, fmt):
    case = _sod_y()
    g0 = _gpu(case)
    g0.run(12)
    fn = str(tmp_path / "b0.flow")
    g0.write_flow(0, fn, fmt)
    g, o = _restart_pair(case, fn, fmt, None, 0)
    _, sg = g.run(10)
    _, so = o.run(10)
    assert sg["time"] == pytest.approx(so["time"], rel=1e-13)
    errs = field_errors(lambda k: g.get(0, k), lambda k: o.get(0, k), NAMES)
    print("restart ideal", fmt, errs, flush=True)
    assert all(v <= TOL for v in errs.values()), errs
    assert np.array_equal(g.get(0, "S"), o.get(0, "S"))
    # Against the run without the restart, over the same fixed time steps (a restart starts
    # its dt controller afresh from dt_init, simcore.d:720, so run() would take other steps).
    g1, _ = _restart_pair(case, fn, fmt, None, 0)
    dt = 0.5 * g0.state()["dt"]
    for _ in range(10):
        g0.step_fixed_dt(dt)
        g1.step_fixed_dt(dt)
    e2 = field_errors(lambda k: g1.get(0, k), lambda k: g0.get(0, k), NAMES)
    print("restart vs uninterrupted", fmt, e2, flush=True)
    assert all(v <= 1.0e-10 for v in e2.values()), e2


def test_restart_formats_bit_identical(tmp_path):
    case = _sod_y(40, 12)
    g0 = _gpu(case)
    g0.run(8)
    fa, fb = str(tmp_path / "a.flow.gz"), str(tmp_path / "b.flow")
    g0.write_flow(0, fa, "gziptext")
    g0.write_flow(0, fb, "rawbinary")
    out = []
    for fn, fmt in ((fa, "gziptext"), (fb, "rawbinary")):
        g = _gpu(case)
        g.read_flow(0, fn, fmt)
        g.run(6)
        out.append({k: g.get(0, k) for k in NAMES + ("S", "u", "a")})
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_restart_reacting_species(tmp_path):
    """5-species reacting air: the file carries the mass fractions and dt_chem; the restart
    resets dt_chem to -1 (simcore.d:335) in both and continues like the oracle."""
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, mech_record
    from oracle import oracle as O
    case, rec = cases.species_box(32, 16, nsp=5)
    mech = mech_record(gdata.load("park-air-5sp-1T"))
    g0 = _gpu(case, flux="ausmdv", rec=rec)
    g0.set_chemistry(Chemistry(rec, mech))
    g0.run(4)
    fn = str(tmp_path / "r.flow.gz")
    g0.write_flow(0, fn, "gziptext", species=AIR5)
    names = FF.variable_list(AIR5)
    r, _, _ = FF.read_gziptext(fn, names, 2, case.nic, case.njc, 1)
    assert np.array_equal(r[:, 15:20], g0.get_massf(0))
    assert np.all(r[:, 20] != -1.0)       # dt_chem after 4 reacting steps
    g, o = _restart_pair(case, fn, "gziptext", AIR5, 0, rec=rec,
                         chem=(Chemistry(rec, mech), O.Chem(rec, mech)), flux="ausmdv")
    g.run(3)
    o.run(3)
    errs = field_errors(lambda k: g.get(0, k), lambda k: o.get(0, k), ("rho", "vx", "vy", "p", "T"))
    for s in range(5):
        errs[f"Y{s}"] = rel_err(g.get(0, 300 + s), o.get(0, 300 + s), np.ones(case.ncells))
    print("restart reacting", errs, flush=True)
    assert all(v <= TOL for v in errs.values()), errs
