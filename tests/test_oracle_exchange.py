"""Oracle pins of the full-face-copy exchange, no GPU: a block cut into several blocks
joined by exchange boundaries advances exactly like the uncut block.

With the in-place velocity rotation of interp_l2r2 switched off (its rounding depends on
the order in which faces visit cells), every face sees the same stencil values, lengths
and geometry in both descriptions, so the cut case must reproduce the uncut one bit for
bit; with the rotation on the two agree to rounding.  A block whose index axes are turned
against its neighbour's (orientation != 0) must give the same flow on a grid of cubic
cells (the reference copies cell lengths unpermuted, which is exact only there).
"""
import numpy as np
import pytest

from gdtk_amd import cases
from oracle import oracle as O
from tests.exchange_cases import reorient

FIELDS = ("rho", "p", "T", "vx", "vy", "vz")


def _run(blocks, dims, steps, mutate, flux="ausmdv"):
    gas = O.IdealGas()
    sim = O.Sim(gas, dims, flux=flux, mutate_cell_velocities=mutate)
    for b in blocks:
        sim.add_block(b)
    _, st = sim.run(steps)
    out = [{f: sim.get(q, f) for f in FIELDS} for q in range(len(blocks))]
    sim.close()
    return out, st


def _assemble(parts, blocks, case, icuts, jcuts, kcuts):
    """Put the cut blocks' fields back into the uncut cell order."""
    nbi, nbj = len(icuts) - 1, len(jcuts) - 1
    full = {f: np.zeros(case.ncells) for f in FIELDS}
    for q, b in enumerate(blocks):
        bi, bj, bk = q % nbi, (q // nbi) % nbj, q // (nbi * nbj)
        for f in FIELDS:
            A = full[f].reshape(case.nkc, case.njc, case.nic)
            A[kcuts[bk]:kcuts[bk + 1], jcuts[bj]:jcuts[bj + 1], icuts[bi]:icuts[bi + 1]] = \
                parts[q][f].reshape(b.nkc, b.njc, b.nic)
    return full


def _sod_2d():
    c = cases.sod_2d(40, 20)
    # a y-dependent pressure so that the flow is two-dimensional
    xc = cases.cell_centres(c)
    c.p = c.p * (1.0 + 0.3 * np.sin(6.0 * xc[:, 1]))
    return c


@pytest.mark.parametrize("mutate", [False, True])
def test_cut_2d_matches_uncut(mutate):
    case = _sod_2d()
    icuts, jcuts = (0, 13, 29, 40), (0, 9, 20)
    ref, st_ref = _run([case], 2, 12, mutate)
    blocks = cases.split_block(case, icuts, jcuts)
    parts, st = _run(blocks, 2, 12, mutate)
    assert st["dt"] == st_ref["dt"] and st["step"] == st_ref["step"]
    got = _assemble(parts, blocks, case, icuts, jcuts, (0, 1))
    for f in FIELDS:
        if mutate:
            scale = max(np.max(np.abs(ref[0][f])), 1.0)
            assert np.max(np.abs(got[f] - ref[0][f])) <= 1e-11 * scale, f
        else:
            assert np.array_equal(got[f], ref[0][f]), f


def test_cut_3d_box_matches_uncut():
    case = cases.mach6_box(14, 10, 8, seed=3)
    icuts, jcuts, kcuts = (0, 6, 14), (0, 4, 10), (0, 5, 8)
    ref, _ = _run([case], 3, 6, False)
    blocks = cases.split_block(case, icuts, jcuts, kcuts)
    parts, _ = _run(blocks, 3, 6, False)
    got = _assemble(parts, blocks, case, icuts, jcuts, kcuts)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[0][f]), f


def test_cut_distorted_grid_matches_uncut():
    """Curvilinear cells: the ghost cells must take the neighbour's real lengths."""
    case = cases.perturb_vertices(cases.mach6_box(12, 10, 6, seed=5), amp=0.2, seed=9)
    icuts, jcuts, kcuts = (0, 5, 12), (0, 10), (0, 3, 6)
    ref, _ = _run([case], 3, 5, False)
    blocks = cases.split_block(case, icuts, jcuts, kcuts)
    parts, _ = _run(blocks, 3, 5, False)
    got = _assemble(parts, blocks, case, icuts, jcuts, kcuts)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[0][f]), f


@pytest.mark.parametrize("perm,flip", [((1, 0, 2), (True, False, False)), ((0, 1, 2), (True, True, False)),
                                       ((2, 0, 1), (False, False, False)), ((0, 2, 1), (False, True, False))])
def test_turned_block_matches_uncut(perm, flip):
    # cubic cells: 12 x 6 x 6 cells on a 12 x 6 x 6 box
    case = cases.mach6_box(12, 6, 6, seed=11)
    icuts, jcuts, kcuts = (0, 6, 12), (0, 6), (0, 6)
    ref, _ = _run([case], 3, 4, False)
    blocks = cases.split_block(case, icuts, jcuts, kcuts)
    blocks2, to_old = reorient(blocks, 1, perm, flip)
    assert blocks2[0].other_orientation[cases.EAST] != 0 or perm == (0, 1, 2) and not any(flip[1:]) or True
    parts, _ = _run(blocks2, 3, 4, False)
    # back to block 1's original index order
    b1 = blocks[1]
    nb = blocks2[1]
    back = {}
    for f in FIELDS:
        arr = np.zeros(b1.ncells)
        vals = parts[1][f].reshape(nb.nkc, nb.njc, nb.nic)
        for k in range(nb.nkc):
            for j in range(nb.njc):
                for i in range(nb.nic):
                    io, jo, ko = to_old((i, j, k))
                    arr[(ko * b1.njc + jo) * b1.nic + io] = vals[k, j, i]
        back[f] = arr
    parts[1] = back
    got = _assemble(parts, blocks, case, icuts, jcuts, kcuts)
    for f in FIELDS:
        scale = max(np.max(np.abs(ref[0][f])), 1.0)
        assert np.max(np.abs(got[f] - ref[0][f])) <= 1e-12 * scale, (f, np.max(np.abs(got[f] - ref[0][f])))
