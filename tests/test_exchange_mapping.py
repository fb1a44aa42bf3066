"""Pins of the full-face-copy cell mapping (exchange BC), no GPU.

The reference's block-connection tables define each connection geometrically: which
corner vertices of the two blocks coincide (src/eilmer/blk_conn.lua:38-54 in 2-D and
:111-290 in 3-D, for every (face, other face, orientation)).  For a selection of those
rows this test builds block B in the position the vertex pairs fix, next to a unit-cube
block A, and checks that every ghost cell of A's face is mapped to the cell of B that
lies where the ghost cell lies.  This pins the orientation numbering of both the HIP
library's mapping (the code the exchange kernel runs, called on the host) and the
oracle's.
"""
import numpy as np
import pytest

from gdtk_amd import hip
from oracle import oracle as O

W, E, S, N, B, T = range(6)
FACE = dict(west=W, east=E, south=S, north=N, bottom=B, top=T)

# Corner vertices of a hexahedral block, unit-cube positions (i, j, k order).
C3 = {0: (0, 0, 0), 1: (1, 0, 0), 2: (1, 1, 0), 3: (0, 1, 0), 4: (0, 0, 1), 5: (1, 0, 1), 6: (1, 1, 1), 7: (0, 1, 1)}
C2 = {0: (0, 0, 0), 1: (1, 0, 0), 2: (1, 1, 0), 3: (0, 1, 0)}

# (A-vertex, B-vertex) pairs -> (A face, B face, orientation): blk_conn.lua:112-135,
# 161-165, 182-190, 207-210, 222-225, 232-235, 257-260, 272-275, 282-285.
ROWS_3D = [
    ([(3, 2), (7, 6), (6, 7), (2, 3)], "north", "north", 0),
    ([(3, 3), (7, 2), (6, 6), (2, 7)], "north", "north", 1),
    ([(3, 7), (7, 3), (6, 2), (2, 6)], "north", "north", 2),
    ([(3, 6), (7, 7), (6, 3), (2, 2)], "north", "north", 3),
    ([(3, 0), (7, 4), (6, 5), (2, 1)], "north", "south", 0),
    ([(3, 1), (7, 0), (6, 4), (2, 5)], "north", "south", 1),
    ([(3, 5), (7, 1), (6, 0), (2, 4)], "north", "south", 2),
    ([(3, 4), (7, 5), (6, 1), (2, 0)], "north", "south", 3),
    ([(3, 1), (7, 5), (6, 6), (2, 2)], "north", "east", 0),
    ([(3, 2), (7, 1), (6, 5), (2, 6)], "north", "east", 1),
    ([(3, 6), (7, 2), (6, 1), (2, 5)], "north", "east", 2),
    ([(3, 5), (7, 6), (6, 2), (2, 1)], "north", "east", 3),
    ([(3, 3), (7, 7), (6, 4), (2, 0)], "north", "west", 0),
    ([(3, 0), (7, 3), (6, 7), (2, 4)], "north", "west", 1),
    ([(3, 4), (7, 0), (6, 3), (2, 7)], "north", "west", 2),
    ([(3, 7), (7, 4), (6, 0), (2, 3)], "north", "west", 3),
    ([(3, 4), (7, 7), (6, 6), (2, 5)], "north", "top", 0),
    ([(3, 5), (7, 4), (6, 7), (2, 6)], "north", "top", 1),
    ([(3, 6), (7, 5), (6, 4), (2, 7)], "north", "top", 2),
    ([(3, 7), (7, 6), (6, 5), (2, 4)], "north", "top", 3),
    ([(3, 1), (7, 2), (6, 3), (2, 0)], "north", "bottom", 0),
    ([(3, 0), (7, 1), (6, 2), (2, 3)], "north", "bottom", 1),
    ([(3, 3), (7, 0), (6, 1), (2, 2)], "north", "bottom", 2),
    ([(3, 2), (7, 3), (6, 0), (2, 1)], "north", "bottom", 3),
    ([(1, 4), (5, 7), (4, 6), (0, 5)], "south", "top", 0),
    ([(1, 5), (5, 4), (4, 7), (0, 6)], "south", "top", 1),
    ([(1, 6), (5, 5), (4, 4), (0, 7)], "south", "top", 2),
    ([(1, 7), (5, 6), (4, 5), (0, 4)], "south", "top", 3),
    ([(2, 1), (6, 5), (5, 6), (1, 2)], "east", "east", 0),
    ([(2, 2), (6, 1), (5, 5), (1, 6)], "east", "east", 1),
    ([(2, 6), (6, 2), (5, 1), (1, 5)], "east", "east", 2),
    ([(2, 5), (6, 6), (5, 2), (1, 1)], "east", "east", 3),
    ([(2, 3), (6, 7), (5, 4), (1, 0)], "east", "west", 0),
    ([(2, 0), (6, 3), (5, 7), (1, 4)], "east", "west", 1),
    ([(2, 4), (6, 0), (5, 3), (1, 7)], "east", "west", 2),
    ([(2, 7), (6, 4), (5, 0), (1, 3)], "east", "west", 3),
    ([(0, 0), (4, 4), (7, 5), (3, 1)], "west", "south", 0),
    ([(0, 1), (4, 0), (7, 4), (3, 5)], "west", "south", 1),
    ([(0, 5), (4, 1), (7, 0), (3, 4)], "west", "south", 2),
    ([(0, 4), (4, 5), (7, 1), (3, 0)], "west", "south", 3),
    ([(0, 4), (4, 7), (7, 6), (3, 5)], "west", "top", 0),
    ([(0, 5), (4, 4), (7, 7), (3, 6)], "west", "top", 1),
    ([(0, 6), (4, 5), (7, 4), (3, 7)], "west", "top", 2),
    ([(0, 7), (4, 6), (7, 5), (3, 4)], "west", "top", 3),
    ([(5, 2), (6, 6), (7, 7), (4, 3)], "top", "north", 0),
    ([(5, 3), (6, 2), (7, 6), (4, 7)], "top", "north", 1),
    ([(5, 7), (6, 3), (7, 2), (4, 6)], "top", "north", 2),
    ([(5, 6), (6, 7), (7, 3), (4, 2)], "top", "north", 3),
    ([(5, 1), (6, 2), (7, 3), (4, 0)], "top", "bottom", 0),
    ([(5, 0), (6, 1), (7, 2), (4, 3)], "top", "bottom", 1),
    ([(5, 3), (6, 0), (7, 1), (4, 2)], "top", "bottom", 2),
    ([(5, 2), (6, 3), (7, 0), (4, 1)], "top", "bottom", 3),
    ([(0, 1), (3, 5), (2, 6), (1, 2)], "bottom", "east", 0),
    ([(0, 2), (3, 1), (2, 5), (1, 6)], "bottom", "east", 1),
    ([(0, 6), (3, 2), (2, 1), (1, 5)], "bottom", "east", 2),
    ([(0, 5), (3, 6), (2, 2), (1, 1)], "bottom", "east", 3),
    ([(0, 4), (3, 7), (2, 6), (1, 5)], "bottom", "top", 0),
    ([(0, 5), (3, 4), (2, 7), (1, 6)], "bottom", "top", 1),
    ([(0, 6), (3, 5), (2, 4), (1, 7)], "bottom", "top", 2),
    ([(0, 7), (3, 6), (2, 5), (1, 4)], "bottom", "top", 3),
]

# blk_conn.lua:39-54 (2-D, orientation always 0).
ROWS_2D = [
    ([(3, 0), (2, 1)], "north", "south"), ([(3, 3), (2, 0)], "north", "west"),
    ([(3, 2), (2, 3)], "north", "north"), ([(3, 1), (2, 2)], "north", "east"),
    ([(0, 0), (3, 1)], "west", "south"), ([(0, 3), (3, 0)], "west", "west"),
    ([(0, 2), (3, 3)], "west", "north"), ([(0, 1), (3, 2)], "west", "east"),
    ([(1, 0), (0, 1)], "south", "south"), ([(1, 3), (0, 0)], "south", "west"),
    ([(1, 2), (0, 3)], "south", "north"), ([(1, 1), (0, 2)], "south", "east"),
    ([(2, 0), (1, 1)], "east", "south"), ([(2, 3), (1, 0)], "east", "west"),
    ([(2, 2), (1, 3)], "east", "north"), ([(2, 1), (1, 2)], "east", "east"),
]

NORMAL_AXIS = {W: 0, E: 0, S: 1, N: 1, B: 2, T: 2}
OUT_SIGN = {W: -1, E: 1, S: -1, N: 1, B: -1, T: 1}
INFACE = {W: (1, 2), E: (1, 2), S: (0, 2), N: (0, 2), B: (0, 1), T: (0, 1)}


def _check_row(dims, pairs, fa, fb, orient, mapper):
    corners = C3 if dims == 3 else C2
    fa_name, fb_name = fa, fb
    fa, fb = FACE[fa], FACE[fb]
    nA = np.array([3, 4, 5] if dims == 3 else [3, 4, 1])
    nrm = np.zeros(3)
    nrm[NORMAL_AXIS[fa]] = OUT_SIGN[fa]
    hN = 1.0 / nA[NORMAL_AXIS[fa]]
    nBn = 2  # cells of B across its own normal direction
    posB = {}
    for a, b in pairs:
        posB[b] = np.array(corners[a], dtype=float)
    # B's other corners: across B's normal axis from the paired ones, away from A.
    axB = NORMAL_AXIS[fb]
    for b in list(posB):
        loc = list(corners[b])
        loc[axB] = 1 - loc[axB]
        opp = [c for c, v in corners.items() if tuple(v) == tuple(loc)][0]
        posB[opp] = posB[b] + nBn * hN * nrm
    assert len(posB) == len(corners)
    e = [posB[1] - posB[0], posB[3] - posB[0], (posB[4] - posB[0]) if dims == 3 else np.array([0, 0, 1.0])]
    nB = []
    for d in range(3):
        if dims == 2 and d == 2:
            nB.append(1)
            continue
        ax = int(np.argmax(np.abs(e[d])))
        assert np.isclose(np.abs(e[d]).sum(), np.abs(e[d][ax]))  # along one physical axis
        nB.append(nBn if ax == NORMAL_AXIS[fa] else int(nA[ax]))
    ia, ib = INFACE[fa]
    nfa = int(nA[ia]), (int(nA[ib]) if (dims == 3 or ib != 2) else 1)
    for a in range(nfa[0]):
        for b in range(nfa[1] if dims == 3 else 1):
            idx = [0, 0, 0]
            idx[ia] = a
            if dims == 3:
                idx[ib] = b
            idx[NORMAL_AXIS[fa]] = 0 if OUT_SIGN[fa] < 0 else nA[NORMAL_AXIS[fa]] - 1
            cA = (np.array(idx) + 0.5) / nA
            if dims == 2:
                cA[2] = 0.0
            for n in range(2):
                ghost = cA + (n + 1) * hN * nrm
                i, j, k = mapper(dims, fa, fb, orient, a, b, n, *nB)
                assert 0 <= i < nB[0] and 0 <= j < nB[1] and 0 <= k < nB[2], (fa_name, fb_name, orient, a, b, n)
                cB = posB[0] + (i + 0.5) / nB[0] * e[0] + (j + 0.5) / nB[1] * e[1]
                if dims == 3:
                    cB = cB + (k + 0.5) / nB[2] * e[2]
                assert np.allclose(cB, ghost, atol=1e-12), (fa_name, fb_name, orient, a, b, n, cB, ghost)


MAPPERS = [pytest.param(hip.exchange_source, id="hip"), pytest.param(O.ffc_source, id="oracle")]


@pytest.mark.parametrize("mapper", MAPPERS)
@pytest.mark.parametrize("row", ROWS_3D, ids=[f"{r[1]}-{r[2]}-{r[3]}" for r in ROWS_3D])
def test_3d_connection_geometry(row, mapper):
    pairs, fa, fb, orient = row
    _check_row(3, pairs, fa, fb, orient, mapper)


@pytest.mark.parametrize("mapper", MAPPERS)
@pytest.mark.parametrize("row", ROWS_2D, ids=[f"{r[1]}-{r[2]}" for r in ROWS_2D])
def test_2d_connection_geometry(row, mapper):
    pairs, fa, fb = row
    _check_row(2, pairs, fa, fb, 0, mapper)


def test_mappings_agree_everywhere():
    """Library and oracle agree on every face pair and orientation (uneven extents)."""
    for fa in range(6):
        for fb in range(6):
            for o in range(4):
                for a, b, n in [(0, 0, 0), (2, 1, 1), (4, 3, 0)]:
                    assert hip.exchange_source(3, fa, fb, o, a, b, n, 7, 8, 9) == \
                        O.ffc_source(3, fa, fb, o, a, b, n, 7, 8, 9)
