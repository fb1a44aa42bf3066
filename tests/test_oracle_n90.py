"""The oracle against the reference's own integration test of the reacting multi-species
path: examples/eilmer/2D/cylinder-dlr-n90/cpu-chem (n90.lua, n90-test.rb).

Nitrogen (N2/N thermally perfect gas, the 2-reaction dissociation mechanism
nitrogen-2sp-2r.lua, operator-split chemistry) at p = 500 Pa, T = 700 K, 5 km/s onto a
cylinder of radius 45 mm, with the reference's default flux calculator
adaptive_hanel_ausmdv, on the Billig-patch grid of 4 blocks joined by exchange boundaries,
from an impulsive start to 100 us.  n90-test.rb:35 expects 4428 +/- 30 steps, and at the
probe (-0.010, 0, 0) (:43-49) rho 1.9090e-2 (1 %), vel.x 579.7 (10 %), p 5.3735e4 (1 %),
massf[N] 1.749e-2 (15 %), T 9320.5 K (1 %).  About a minute of CPU time.
"""
import numpy as np

from gdtk_amd import cases

REFS = {"rho": 1.9090e-02, "vx": 5.797e+02, "p": 5.3735e+04, "YN": 1.749e-02, "T": 9.3205e+03}
TOLS = {"rho": 0.01, "vx": 0.1, "p": 0.01, "YN": 0.15, "T": 0.01}


def probe_cell(blocks, point):
    best = (np.inf, None, None)
    for q, b in enumerate(blocks):
        xc = cases.cell_centres(b)
        d = np.hypot(xc[:, 0] - point[0], xc[:, 1] - point[1])
        i = int(np.argmin(d))
        if d[i] < best[0]:
            best = (d[i], q, i)
    return best[1], best[2]


def test_n90_reference_integration(oracle_lib):
    from gdtk_amd import data
    from gdtk_amd.chem import mech_record
    O = oracle_lib
    blocks, rec, info = cases.n90_blocks()
    mech = mech_record(data.load("nitrogen-2sp-2r"))
    o = O.Sim(O.TPG(rec), 2, "adaptive_hanel_ausmdv", dt_init=info["dt_init"])
    for b in blocks:
        o.add_block(b)
    o.set_chemistry(O.Chem(rec, mech))
    _, st = o.run(40000, info["max_time"])
    assert abs(st["step"] - 4428) < 30, st["step"]
    q, i = probe_cell(blocks, info["probe"])
    got = {k: o.get(q, k)[i] for k in ("rho", "vx", "p", "T")}
    got["YN"] = o.get(q, 301)[i]
    for k in REFS:
        assert abs(got[k] - REFS[k]) / REFS[k] < TOLS[k], (k, got[k], REFS[k])
