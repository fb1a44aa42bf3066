"""Synthetic workloads of the structured-block explicit update, with their sizes.

Each builder returns plain numpy arrays (vertex coordinates, per-cell initial p, T,
velocity, boundary-condition kinds, inflow states) that both the HIP block API
(gdtk_amd.block) and the CPU oracle accept.  Vertex ordering is the structured-grid
order of the reference, index (k*njv + j)*niv + i (src/eilmer/sfluidblock.d:315-328).

Gas: ideal air from src/gas/sample-data/ideal-air-gas-model.lua:5-7 (mMass 0.02896,
gamma 1.4, so R = 8.31451/0.02896 = 287.103 J/kg/K).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Boundary-condition kinds shared by the C ABI and the oracle.
BC_SLIP_WALL = 0
BC_INFLOW_SUPERSONIC = 1
BC_OUTFLOW_SIMPLE_FLUX = 2
BC_OUTFLOW_SIMPLE_EXTRAPOLATE = 3
BC_EXCHANGE = 4

# Face order of the reference (src/geom/elements/nomenclature.d:12-20).
WEST, EAST, SOUTH, NORTH, BOTTOM, TOP = range(6)

IDEAL_AIR_MMASS = 0.02896
IDEAL_AIR_GAMMA = 1.4
R_UNIVERSAL = 8.31451


@dataclass
class BlockCase:
    dims: int
    nic: int
    njc: int
    nkc: int
    vertices: np.ndarray            # (niv*njv*nkv, 3) float64
    p: np.ndarray                   # (ncells,) float64
    T: np.ndarray
    vel: np.ndarray                 # (ncells, 3)
    bc: list                        # 6 BC kinds
    inflow: np.ndarray = field(default_factory=lambda: np.zeros((6, 5)))  # [p,T,vx,vy,vz] per face
    other_block: list = field(default_factory=lambda: [-1] * 6)
    other_face: list = field(default_factory=lambda: [-1] * 6)
    other_orientation: list = field(default_factory=lambda: [0] * 6)
    # Multi-species gas: mass fractions per cell (ncells, nsp) and of the inflow faces (6, nsp).
    massf: np.ndarray = None
    inflow_massf: np.ndarray = None

    @property
    def ncells(self) -> int:
        return self.nic * self.njc * self.nkc


def box_vertices(nic, njc, nkc, lx, ly, lz, dims, x0=0.0, y0=0.0, z0=0.0):
    """Uniform Cartesian vertices, evaluated as x0 + lx*(i/nic) like a linear volume map."""
    niv, njv = nic + 1, njc + 1
    nkv = nkc + 1 if dims == 3 else 1
    r = np.arange(niv, dtype=np.float64) / nic
    s = np.arange(njv, dtype=np.float64) / njc
    t = np.arange(nkv, dtype=np.float64) / nkc if dims == 3 else np.zeros(1)
    K, J, I = np.meshgrid(t, s, r, indexing="ij")
    v = np.empty((nkv, njv, niv, 3))
    v[..., 0] = x0 + lx * I
    v[..., 1] = y0 + ly * J
    v[..., 2] = (z0 + lz * K) if dims == 3 else 0.0
    return v.reshape(-1, 3)


def cell_centres(case: BlockCase) -> np.ndarray:
    """Vertex-average cell centres (for initial-condition placement only)."""
    niv, njv = case.nic + 1, case.njc + 1
    nkv = case.nkc + 1 if case.dims == 3 else 1
    v = case.vertices.reshape(nkv, njv, niv, 3)
    if case.dims == 2:
        c = 0.25 * (v[0, :-1, :-1] + v[0, :-1, 1:] + v[0, 1:, :-1] + v[0, 1:, 1:])
        return c.reshape(-1, 3)
    c = 0.125 * (v[:-1, :-1, :-1] + v[:-1, :-1, 1:] + v[:-1, 1:, :-1] + v[:-1, 1:, 1:]
                 + v[1:, :-1, :-1] + v[1:, :-1, 1:] + v[1:, 1:, :-1] + v[1:, 1:, 1:])
    return c.reshape(-1, 3)


def sod_2d(nic=200, njc=100, T_right=278.7) -> BlockCase:
    """Plain method, small: 2-D Sod split at x = 0.5 on [0,1]x[0,0.5], slip walls all round.

    Left p=1e5 Pa, T=348.4 K; right p=1e4 Pa, T=278.7 K; u=0.
    """
    v = box_vertices(nic, njc, 1, 1.0, 0.5, 0.0, dims=2)
    case = BlockCase(2, nic, njc, 1, v, None, None, None, [BC_SLIP_WALL] * 6)
    xc = cell_centres(case)[:, 0]
    left = xc < 0.5
    case.p = np.where(left, 1.0e5, 1.0e4)
    case.T = np.where(left, 348.4, T_right)
    case.vel = np.zeros((case.ncells, 3))
    return case


def sod_3d_tube(nic=100) -> BlockCase:
    """The reference's 3-D Sod tube: 100x2x2 cells on 1.0 x 0.1 x 0.1 m, slip walls.

    examples/eilmer/3D/sod-shock-tube/sg/sod.lua:13-45 (p=1e5/1e4 Pa, T=348.4/278.8 K).
    """
    v = box_vertices(nic, 2, 2, 1.0, 0.1, 0.1, dims=3)
    case = BlockCase(3, nic, 2, 2, v, None, None, None, [BC_SLIP_WALL] * 6)
    xc = cell_centres(case)[:, 0]
    left = xc < 0.5
    case.p = np.where(left, 1.0e5, 1.0e4)
    case.T = np.where(left, 348.4, 278.8)
    case.vel = np.zeros((case.ncells, 3))
    return case


def rippled_box(nic=40, njc=12, ripple=1.0e-9, seed=5) -> BlockCase:
    """A 2-D box of air at 1e5 Pa, 300 K moving at about 50 m/s, with pressure ripples of
    relative size ripple, on a perturbed grid with slip walls.  Differences of u between
    neighbouring cells are then comparable to the van Albada smoothing of Eilmer 5
    (epsilon scaled by |u| and the cell size), so the reconstruction depends on it."""
    c = sod_2d(nic, njc)
    xc = cell_centres(c)
    c.p = np.full(c.ncells, 1.0e5) * (1.0 + ripple * np.sin(20.0 * xc[:, 0]) * np.cos(9.0 * xc[:, 1]))
    c.T = np.full(c.ncells, 300.0)
    c.vel = np.zeros((c.ncells, 3))
    c.vel[:, 0] = 50.0 + 5.0 * np.sin(7.0 * xc[:, 1])
    return perturb_vertices(c, amp=0.1, seed=seed)


def mach6_box(nic=256, njc=128, nkc=128, seed=1234, dims=3) -> BlockCase:
    """Flux-path bandwidth stress: Mach-6 supersonic box, unit cell spacing.

    Free stream p = 2e3 Pa, T = 220 K, u = 6a along +x.  Density and temperature are
    perturbed by +/-5 % with a sum of 8 seeded Fourier modes; pressure follows from the
    EOS (p = rho R T with the perturbed rho and T).  West face supersonic inflow, all
    other faces simple outflow (flux).
    """
    rng = np.random.default_rng(seed)
    nk = nkc if dims == 3 else 1
    v = box_vertices(nic, njc, nk, float(nic), float(njc), float(nk), dims=dims)
    bc = [BC_INFLOW_SUPERSONIC] + [BC_OUTFLOW_SIMPLE_FLUX] * 5
    case = BlockCase(dims, nic, njc, nk, v, None, None, None, bc)
    R = R_UNIVERSAL / IDEAL_AIR_MMASS
    p0, T0 = 2.0e3, 220.0
    a0 = np.sqrt(IDEAL_AIR_GAMMA * R * T0)
    u0 = 6.0 * a0
    xc = cell_centres(case)
    L = np.array([nic, njc, max(nk, 1)], dtype=np.float64)

    def field_():
        s = np.zeros(case.ncells)
        amps = rng.uniform(0.2, 1.0, 8)
        for m in range(8):
            kvec = 2.0 * np.pi * rng.integers(1, 6, 3) / L
            if dims == 2:
                kvec[2] = 0.0
            phi = rng.uniform(0.0, 2.0 * np.pi)
            s += amps[m] * np.sin(xc @ kvec + phi)
        return 0.05 * s / np.abs(amps).sum()

    drho = field_()
    dT = field_()
    rho0 = p0 / (R * T0)
    rho = rho0 * (1.0 + drho)
    case.T = T0 * (1.0 + dT)
    case.p = rho * R * case.T
    case.vel = np.zeros((case.ncells, 3))
    case.vel[:, 0] = u0
    case.inflow = np.zeros((6, 5))
    case.inflow[WEST] = [p0, T0, u0, 0.0, 0.0]
    return case


def perturb_vertices(case: BlockCase, amp=0.15, seed=7) -> BlockCase:
    """Smoothly distort the interior vertices (a curvilinear grid, boundary faces kept planar).

    The displacement is a seeded sum of sines scaled by amp times the local spacing, zero on
    the block boundary, so the grid stays valid but faces are no longer axis-aligned.
    """
    rng = np.random.default_rng(seed)
    niv, njv = case.nic + 1, case.njc + 1
    nkv = case.nkc + 1 if case.dims == 3 else 1
    v = case.vertices.reshape(nkv, njv, niv, 3).copy()
    r = np.linspace(0.0, 1.0, niv)[None, None, :]
    s = np.linspace(0.0, 1.0, njv)[None, :, None]
    t = np.linspace(0.0, 1.0, nkv)[:, None, None] if case.dims == 3 else np.zeros((1, 1, 1))
    bump = np.sin(np.pi * r) * np.sin(np.pi * s)
    if case.dims == 3:
        bump = bump * np.sin(np.pi * t)
    dx = (v[0, 0, -1, 0] - v[0, 0, 0, 0]) / case.nic
    dy = (v[0, -1, 0, 1] - v[0, 0, 0, 1]) / case.njc
    ph = rng.uniform(0.0, 2.0 * np.pi, 6)
    v[..., 0] += amp * dx * bump * np.sin(2.0 * np.pi * s + ph[0]) * np.cos(2.0 * np.pi * r + ph[1])
    v[..., 1] += amp * dy * bump * np.sin(2.0 * np.pi * r + ph[2]) * np.cos(2.0 * np.pi * s + ph[3])
    if case.dims == 3:
        dz = (v[-1, 0, 0, 2] - v[0, 0, 0, 2]) / case.nkc
        v[..., 2] += amp * dz * bump * np.sin(2.0 * np.pi * (r + s) + ph[4]) * np.cos(2.0 * np.pi * t + ph[5])
    out = BlockCase(case.dims, case.nic, case.njc, case.nkc, v.reshape(-1, 3), case.p, case.T, case.vel,
                    list(case.bc), case.inflow.copy(), list(case.other_block), list(case.other_face),
                    list(case.other_orientation))
    return out


def split_block(case: BlockCase, icuts, jcuts, kcuts=(0, 1)) -> list:
    """Cut a block into a grid of blocks joined by full-face-copy exchange boundaries.

    icuts/jcuts/kcuts are cell-index boundaries, e.g. (0, 5, 12) cuts 12 cells in i into
    5 + 7.  Blocks are numbered (bk*nbj + bj)*nbi + bi; outer faces keep the original
    boundary conditions and inflow states, inner faces exchange with their neighbours
    (orientation 0: index axes aligned).
    """
    icuts, jcuts, kcuts = list(icuts), list(jcuts), list(kcuts)
    if case.dims == 2:
        kcuts = [0, 1]
    assert icuts[0] == 0 and icuts[-1] == case.nic and jcuts[0] == 0 and jcuts[-1] == case.njc
    assert kcuts[0] == 0 and kcuts[-1] == case.nkc
    nbi, nbj, nbk = len(icuts) - 1, len(jcuts) - 1, len(kcuts) - 1
    niv, njv = case.nic + 1, case.njc + 1
    nkv = case.nkc + 1 if case.dims == 3 else 1
    V = case.vertices.reshape(nkv, njv, niv, 3)
    cell = lambda a: a.reshape(case.nkc, case.njc, case.nic, *a.shape[1:])
    P, Tt, Vel = cell(case.p), cell(case.T), cell(case.vel)
    Mf = cell(case.massf) if case.massf is not None else None
    bid = lambda bi, bj, bk: (bk * nbj + bj) * nbi + bi
    out = []
    for bk in range(nbk):
        for bj in range(nbj):
            for bi in range(nbi):
                i0, i1 = icuts[bi], icuts[bi + 1]
                j0, j1 = jcuts[bj], jcuts[bj + 1]
                k0, k1 = kcuts[bk], kcuts[bk + 1]
                kv = slice(k0, k1 + 1) if case.dims == 3 else slice(0, 1)
                v = V[kv, j0:j1 + 1, i0:i1 + 1].reshape(-1, 3).copy()
                sl = (slice(k0, k1), slice(j0, j1), slice(i0, i1))
                b = BlockCase(case.dims, i1 - i0, j1 - j0, k1 - k0, v, P[sl].reshape(-1).copy(),
                              Tt[sl].reshape(-1).copy(), Vel[sl].reshape(-1, 3).copy(), list(case.bc),
                              case.inflow.copy())
                if Mf is not None:
                    b.massf = Mf[sl].reshape(-1, Mf.shape[-1]).copy()
                if case.inflow_massf is not None:
                    b.inflow_massf = case.inflow_massf.copy()
                nbr = {WEST: (bi > 0, bid(bi - 1, bj, bk)), EAST: (bi < nbi - 1, bid(bi + 1, bj, bk)),
                       SOUTH: (bj > 0, bid(bi, bj - 1, bk)), NORTH: (bj < nbj - 1, bid(bi, bj + 1, bk)),
                       BOTTOM: (bk > 0, bid(bi, bj, bk - 1)), TOP: (bk < nbk - 1, bid(bi, bj, bk + 1))}
                for side, (inner, other) in nbr.items():
                    if inner:
                        b.bc[side] = BC_EXCHANGE
                        b.other_block[side] = other
                        b.other_face[side] = side ^ 1
                        b.other_orientation[side] = 0
                out.append(b)
    return out


def batched_uneven(nblk=(4, 4, 4), lo=16, hi=192, seed=2024, p0=2.0e3, T0=220.0, mach=(2.0, 10.0)) -> list:
    """Batched uneven blocks: a grid of nblk[0] x nblk[1] x nblk[2] structured blocks whose
    edge lengths per index direction are drawn uniformly from lo..hi cells (seeded; one draw
    per block column, row and layer, so that neighbouring faces match), on unit cells.

    Each block starts from its own ideal-air free stream, p0 and T0 with a Mach number drawn
    uniformly from `mach`, flowing along +x.  The west faces of the first block column are
    supersonic inflows of their block's free stream, the other outer faces simple outflows
    (flux), and every inner face a full-face-copy exchange with the neighbour block.
    """
    rng = np.random.default_rng(seed)
    ni = rng.integers(lo, hi + 1, nblk[0])
    nj = rng.integers(lo, hi + 1, nblk[1])
    nk = rng.integers(lo, hi + 1, nblk[2])
    x0 = np.concatenate([[0], np.cumsum(ni)])
    y0 = np.concatenate([[0], np.cumsum(nj)])
    z0 = np.concatenate([[0], np.cumsum(nk)])
    machs = rng.uniform(mach[0], mach[1], nblk[0] * nblk[1] * nblk[2])
    R = R_UNIVERSAL / IDEAL_AIR_MMASS
    a0 = np.sqrt(IDEAL_AIR_GAMMA * R * T0)
    nbi, nbj, nbk = nblk
    bid = lambda bi, bj, bk: (bk * nbj + bj) * nbi + bi
    out = []
    for bk in range(nbk):
        for bj in range(nbj):
            for bi in range(nbi):
                q = bid(bi, bj, bk)
                nic, njc, nkc = int(ni[bi]), int(nj[bj]), int(nk[bk])
                v = box_vertices(nic, njc, nkc, float(nic), float(njc), float(nkc), dims=3,
                                 x0=float(x0[bi]), y0=float(y0[bj]), z0=float(z0[bk]))
                n = nic * njc * nkc
                u = machs[q] * a0
                vel = np.zeros((n, 3))
                vel[:, 0] = u
                b = BlockCase(3, nic, njc, nkc, v, np.full(n, p0), np.full(n, T0), vel,
                              [BC_OUTFLOW_SIMPLE_FLUX] * 6)
                if bi == 0:
                    b.bc[WEST] = BC_INFLOW_SUPERSONIC
                    b.inflow[WEST] = [p0, T0, u, 0.0, 0.0]
                nbr = {WEST: (bi > 0, bid(bi - 1, bj, bk)), EAST: (bi < nbi - 1, bid(bi + 1, bj, bk)),
                       SOUTH: (bj > 0, bid(bi, bj - 1, bk)), NORTH: (bj < nbj - 1, bid(bi, bj + 1, bk)),
                       BOTTOM: (bk > 0, bid(bi, bj, bk - 1)), TOP: (bk < nbk - 1, bid(bi, bj, bk + 1))}
                for side, (inner, other) in nbr.items():
                    if inner:
                        b.bc[side] = BC_EXCHANGE
                        b.other_block[side] = other
                        b.other_face[side] = side ^ 1
                out.append(b)
    return out


def batched_ragged(nblocks=64, lo=16, hi=192, seed=2024, p0=2.0e3, T0=220.0, mach=(2.0, 10.0), only=None) -> list:
    """Batched uneven blocks with per-block edge lengths: nblocks 3-D structured blocks on
    unit cells, in pairs joined by a full-face-copy exchange.  Every block draws its own i
    length uniformly from lo..hi (seeded); the two blocks of a pair share j and k lengths
    drawn per pair (a full-face copy needs the same face), so no two pairs need have any
    edge length in common.

    Each block starts from its own ideal-air free stream, p0 and T0 with a Mach number drawn
    uniformly from `mach`, flowing along +x.  The first block of a pair has a supersonic
    inflow of its free stream on its west face and exchanges its east face with the west
    face of the second; every other face is a simple outflow (flux).  Pairs are laid out
    side by side in y and z (they do not touch).  only=q builds block q alone (same draws).
    """
    assert nblocks % 2 == 0
    rng = np.random.default_rng(seed)
    npair = nblocks // 2
    ni = rng.integers(lo, hi + 1, nblocks)
    nj = rng.integers(lo, hi + 1, npair)
    nk = rng.integers(lo, hi + 1, npair)
    machs = rng.uniform(mach[0], mach[1], nblocks)
    R = R_UNIVERSAL / IDEAL_AIR_MMASS
    a0 = np.sqrt(IDEAL_AIR_GAMMA * R * T0)
    side = int(np.ceil(np.sqrt(npair)))
    out = []
    for q in range(nblocks) if only is None else (only,):
        pr, second = q // 2, q % 2
        nic, njc, nkc = int(ni[q]), int(nj[pr]), int(nk[pr])
        x0 = float(ni[q - 1]) if second else 0.0
        v = box_vertices(nic, njc, nkc, float(nic), float(njc), float(nkc), dims=3,
                         x0=x0, y0=float((pr % side) * (hi + 1)), z0=float((pr // side) * (hi + 1)))
        n = nic * njc * nkc
        u = machs[q] * a0
        vel = np.zeros((n, 3))
        vel[:, 0] = u
        b = BlockCase(3, nic, njc, nkc, v, np.full(n, p0), np.full(n, T0), vel, [BC_OUTFLOW_SIMPLE_FLUX] * 6)
        if second:
            b.bc[WEST] = BC_EXCHANGE
            b.other_block[WEST], b.other_face[WEST] = q - 1, EAST
        else:
            b.bc[WEST] = BC_INFLOW_SUPERSONIC
            b.inflow[WEST] = [p0, T0, u, 0.0, 0.0]
            b.bc[EAST] = BC_EXCHANGE
            b.other_block[EAST], b.other_face[EAST] = q + 1, WEST
        out.append(b)
    return out


def k_slab(case: BlockCase, nk: int) -> BlockCase:
    """The first nk cell layers of a 3-D block as a block of its own: the block's boundary
    conditions, with its top face taking the block's top condition and every exchange side
    turned into a simple outflow (flux)."""
    sub = split_block(case, (0, case.nic), (0, case.njc), (0, nk, case.nkc) if nk < case.nkc else (0, case.nkc))[0]
    sub.bc[TOP] = case.bc[TOP]
    sub.inflow[TOP] = case.inflow[TOP]
    for f in range(6):
        if sub.bc[f] == BC_EXCHANGE:
            sub.bc[f] = BC_OUTFLOW_SIMPLE_FLUX
            sub.other_block[f], sub.other_face[f] = -1, -1
    return sub


# ---------------------------------------------------------------------------------------
# Multi-species thermally perfect gas cases


def air5_gas():
    """5-species air (N2, O2, NO, N, O) thermally perfect gas record with CEA curves."""
    from gdtk_amd import chem, data
    return chem.gas_record(data.load("tpg-air-5sp"))


def species_gas(names):
    """Thermally perfect gas of a subset of the 5-species air species, e.g. ("N2", "N")."""
    from gdtk_amd import chem, data
    d = data.load("tpg-air-5sp")
    idx = [d["species"].index(n) for n in names]
    sub = dict(species=[d["species"][i] for i in idx], M=[d["M"][i] for i in idx],
               thermo=[d["thermo"][i] for i in idx])
    return chem.gas_record(sub)


def constant_cp_gas(nsp, M=IDEAL_AIR_MMASS, cp_on_R=3.5):
    """A thermally perfect gas whose species all have the molecular mass M and the constant
    Cp = cp_on_R * R (NASA-9 curves with a2 = cp_on_R only): thermodynamically an ideal gas
    with gamma = cp_on_R / (cp_on_R - 1)."""
    from gdtk_amd import chem
    th = dict(T_breaks=[200.0, 1000.0, 6000.0, 20000.0], T_blends=[400.0, 1000.0],
              coeffs=[[0.0, 0.0, cp_on_R, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]] * 3)
    return chem.gas_record(dict(species=[f"X{i}" for i in range(nsp)], M=[M] * nsp, thermo=[th] * nsp))


def species_box(nic=48, njc=24, nsp=5, seed=3, mach=4.0, closed=False):
    """2-D box [0, 0.1] x [0, 0.05] m of a multi-species gas: a free stream of air at
    p = 2e3 Pa, T = 300 K (N2/O2 = 0.767/0.233, or N2 only for 2 species) flowing in at the
    given Mach number through the west face, with a seeded hot, partly dissociated blob
    (T up to 4000 K, mass fractions shifted towards the atoms) in the interior.  Simple
    outflow on the east face, slip walls south and north; closed=True puts slip walls on
    every face and starts from rest."""
    rng = np.random.default_rng(seed)
    v = box_vertices(nic, njc, 1, 0.1, 0.05, 0.0, dims=2)
    bc = [BC_SLIP_WALL] * 6 if closed else [BC_INFLOW_SUPERSONIC, BC_OUTFLOW_SIMPLE_FLUX, BC_SLIP_WALL, BC_SLIP_WALL,
                                             BC_SLIP_WALL, BC_SLIP_WALL]
    case = BlockCase(2, nic, njc, 1, v, None, None, None, bc)
    xc = cell_centres(case)
    cx, cy = rng.uniform(0.03, 0.07), rng.uniform(0.015, 0.035)
    r2 = ((xc[:, 0] - cx) ** 2 + (xc[:, 1] - cy) ** 2) / 0.012 ** 2
    blob = np.exp(-r2)
    if nsp == 5:
        Y0 = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
        Yh = np.array([0.60, 0.05, 0.05, 0.10, 0.20])
    else:
        Y0 = np.array([1.0, 0.0])
        Yh = np.array([0.7, 0.3])
    case.massf = (1.0 - blob)[:, None] * Y0[None, :] + blob[:, None] * Yh[None, :]
    case.massf /= case.massf.sum(axis=1, keepdims=True)
    case.T = 300.0 + 3700.0 * blob
    case.p = 2.0e3 * (1.0 + 2.0 * blob)
    gas = air5_gas() if nsp == 5 else species_gas(("N2", "N"))
    # free-stream sound speed of the mixture (Cp/Cv from the curves at 300 K)
    from gdtk_amd import data
    d = data.load("tpg-air-5sp")
    names = ["N2", "O2", "NO", "N", "O"] if nsp == 5 else ["N2", "N"]
    R = np.array([R_UNIVERSAL / d["M"][d["species"].index(n)] for n in names])
    Rmix = float(Y0 @ R)
    a0 = np.sqrt(1.4 * Rmix * 300.0)
    u0 = 0.0 if closed else mach * a0
    case.vel = np.zeros((case.ncells, 3))
    case.vel[:, 0] = u0 * (1.0 - blob)
    if closed:
        case.vel[:, 1] = 50.0 * np.sin(40.0 * xc[:, 0])
    case.inflow = np.zeros((6, 5))
    case.inflow[WEST] = [2.0e3, 300.0, u0, 0.0, 0.0]
    case.inflow_massf = np.tile(Y0, (6, 1))
    return case, gas


def coons_patch(south, north, west, east, r, s):
    """Transfinite interpolation between four boundary curves given as functions of their
    parameter in [0, 1] (south/north in r, west/east in s); returns (len(s), len(r), 2)
    vertex positions."""
    Sr, Nr = south(r)[None, :, :], north(r)[None, :, :]
    Ws, Es = west(s)[:, None, :], east(s)[:, None, :]
    R_ = r[None, :, None]
    S_ = s[:, None, None]
    p00, p10, p01, p11 = south(np.zeros(1))[0], south(np.ones(1))[0], north(np.zeros(1))[0], north(np.ones(1))[0]
    return ((1 - S_) * Sr + S_ * Nr + (1 - R_) * Ws + R_ * Es
            - ((1 - R_) * (1 - S_) * p00 + R_ * (1 - S_) * p10 + (1 - R_) * S_ * p01 + R_ * S_ * p11))


def _line(p0, p1):
    p0, p1 = np.asarray(p0, dtype=float), np.asarray(p1, dtype=float)
    return lambda t: p0[None, :] + np.asarray(t, dtype=float)[:, None] * (p1 - p0)[None, :]


def _polyline(points):
    """Straight segments through `points`, parameterised by arc length."""
    P = np.asarray(points, dtype=float)
    L = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(P, axis=0).T))])
    L /= L[-1]

    def f(t):
        t = np.asarray(t, dtype=float)
        return np.stack([np.interp(t, L, P[:, 0]), np.interp(t, L, P[:, 1])], axis=1)
    return f


def wedge_2d(nx=40, ny=40, mach=10.0, theta_deg=6.0, p_inf=1.0e5, T_inf=300.0):
    """The reference's Mach-10 flow of ideal air over a 6-degree wedge
    (examples/eilmer/2D/sharp-wedge-6-degrees/sg/wedge.lua:18-93): a 0.1 m run-up along the
    x axis to the wedge tip at the origin, then the 0.3 m wedge surface; 0.1 m high; a
    Coons patch of nx x ny cells.  West: supersonic inflow; south: slip wall; north and
    east: simple outflow (OutFlowBC_Simple is the flux version, bc.lua:1625).  Every cell
    starts at the inflow state.  The example runs the adaptive flux calculator."""
    tb = np.tan(np.radians(theta_deg))
    a, b, c = (-0.1, 0.0), (0.0, 0.0), (0.3, 0.3 * tb)
    d, e = (-0.1, 0.1), (0.3, 0.1)
    grid = coons_patch(_polyline([a, b, c]), _line(d, e), _line(a, d), _line(c, e),
                       np.linspace(0.0, 1.0, nx + 1), np.linspace(0.0, 1.0, ny + 1))
    v = np.zeros(((nx + 1) * (ny + 1), 3))
    v[:, 0] = grid[..., 0].reshape(-1)
    v[:, 1] = grid[..., 1].reshape(-1)
    R = R_UNIVERSAL / IDEAL_AIR_MMASS
    u_inf = mach * np.sqrt(IDEAL_AIR_GAMMA * R * T_inf)
    n = nx * ny
    vel = np.zeros((n, 3))
    vel[:, 0] = u_inf
    case = BlockCase(2, nx, ny, 1, v, np.full(n, p_inf), np.full(n, T_inf), vel,
                     [BC_INFLOW_SUPERSONIC, BC_OUTFLOW_SIMPLE_FLUX, BC_SLIP_WALL, BC_OUTFLOW_SIMPLE_FLUX,
                      BC_SLIP_WALL, BC_SLIP_WALL])
    case.inflow = np.zeros((6, 5))
    case.inflow[WEST] = [p_inf, T_inf, u_inf, 0.0, 0.0]
    return case


def billig_shock_x(y, R, mach):
    """Billig's shock-shape correlation for a circular cylinder in a gamma = 1.4 free stream
    (Billig, J. Spacecraft Rockets 4 (1967) 822): distance upstream of the cylinder centre
    of the shock at height y, from the stand-off Delta/R = 0.386 exp(4.67/M^2) and the
    shock's radius of curvature Rc/R = 1.386 exp(1.8/(M-1)^0.75) at the axis."""
    delta = R * 0.386 * np.exp(4.67 / mach ** 2)
    rc = R * 1.386 * np.exp(1.8 / (mach - 1.0) ** 0.75)
    theta = np.arcsin(1.0 / mach)
    cot2 = 1.0 / np.tan(theta) ** 2
    return R + delta - rc * cot2 * (np.sqrt(1.0 + (y * np.tan(theta) / rc) ** 2) - 1.0)


def _bezier_spline_path(points):
    """The reference's Spline through points (src/geom/gpath/polyline.d:52-108): natural
    spline weight points from d[i-1] + 4 d[i] + d[i+1] = 6 p[i] (solved exactly here; the
    reference iterates Gauss-Seidel to 1e-10), one cubic Bezier per interval with control
    points (2 d[i] + d[i+1])/3 and (d[i] + 2 d[i+1])/3, joined as a Polyline whose parameter
    breakpoints are the segments' cumulative lengths (polyline.d:142-163, 281-290), each
    length the sum of 100 chords (path.d:80-101)."""
    P = np.asarray(points, dtype=float)
    m = len(P) - 1
    A = np.zeros((m + 1, m + 1))
    rhs = np.zeros_like(P)
    A[0, 0] = A[m, m] = 1.0
    rhs[0], rhs[m] = P[0], P[m]
    for i in range(1, m):
        A[i, i - 1], A[i, i], A[i, i + 1] = 1.0, 4.0, 1.0
        rhs[i] = 6.0 * P[i]
    D = np.linalg.solve(A, rhs)
    segs = [(P[i], (2.0 * D[i] + D[i + 1]) / 3.0, (D[i] + 2.0 * D[i + 1]) / 3.0, P[i + 1]) for i in range(m)]

    def bez(seg, t):
        t = np.asarray(t, dtype=float)[:, None]
        b0, b1, b2, b3 = seg
        return (1 - t) ** 3 * b0 + 3 * (1 - t) ** 2 * t * b1 + 3 * (1 - t) * t ** 2 * b2 + t ** 3 * b3

    def chord_length(f):
        q = f(np.linspace(0.0, 1.0, 101))
        return float(np.sum(np.hypot(*np.diff(q, axis=0).T)))

    L = np.cumsum([chord_length(lambda t, sg=sg: bez(sg, t)) for sg in segs])
    tv = L / L[-1]

    def path(t):
        t = np.atleast_1d(np.asarray(t, dtype=float))
        out = np.empty((len(t), 2))
        for k, tk in enumerate(t):
            i = int(np.searchsorted(tv, tk, side="left"))
            i = min(i, m - 1)
            tl = tk / tv[0] if i == 0 else (tk - tv[i - 1]) / (tv[i] - tv[i - 1])
            out[k] = bez(segs[i], [tl])[0]
        return out
    return path


def _arc_length_parameterized(path, N=100):
    """ArcLengthParameterizedPath (src/geom/gpath/modifiedpath.d:55-148): arc lengths of N
    chords of the underlying path, then a reverse linear look-up of the parameter."""
    q = path(np.linspace(0.0, 1.0, N + 1))
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(q, axis=0).T))])

    def f(t):
        t = np.atleast_1d(np.asarray(t, dtype=float))
        u = np.empty_like(t)
        dt = 1.0 / N
        for k, tk in enumerate(t):
            if tk <= 0.0:
                u[k] = 0.0
                continue
            if tk >= 1.0:
                u[k] = 1.0
                continue
            Lt = tk * arc[-1]
            i = N
            while Lt < arc[i] and i > 0:
                i -= 1
            frac = (Lt - arc[i]) / (arc[i + 1] - arc[i])
            u[k] = (1.0 - frac) * dt * i + frac * dt * (i + 1)
        return path(u)
    return f


def roberts(t, end0, end1, beta):
    """RobertsFunction (src/geom/misc/univariatefunctions.d:110-182)."""
    t = np.asarray(t, dtype=float)
    if (not end0 and not end1) or beta <= 1.0:
        return t
    alpha = 0.5 if (end0 and end1) else 0.0
    reverse = end0 and not end1
    x = 1.0 - t if reverse else t
    lam = ((beta + 1.0) / (beta - 1.0)) ** ((x - alpha) / (1.0 - alpha))
    xb = ((beta + 2.0 * alpha) * lam - beta + 2.0 * alpha) / ((2.0 * alpha + 1.0) * (1.0 + lam))
    return 1.0 - xb if reverse else xb


def _nasa9_cp_over_R(th, T):
    """Cp/R of one NASA-9 segment set at T away from a blending range."""
    k = 0
    while k + 1 < len(th["coeffs"]) and T > th["T_breaks"][k + 1]:
        k += 1
    a = th["coeffs"][k]
    return a[0] / T ** 2 + a[1] / T + a[2] + a[3] * T + a[4] * T ** 2 + a[5] * T ** 3 + a[6] * T ** 4


def n90_blocks(niv=61, njv=41, njb=4):
    """The reference's cylinder in dissociating nitrogen
    (examples/eilmer/2D/cylinder-dlr-n90/cpu-chem/n90.lua): N2/N thermally perfect gas,
    inflow p = 500 Pa, T = 700 K, u = 5000 m/s of N2 onto a cylinder of radius 0.045 m
    centred at x = 0.045; the gas initially at p = 5 Pa, T = 300 K.  The grid is
    billig_patch.make_patch{Minf, R, xc = R, scale = 0.95} (src/lib/billig_patch.lua:11-120,
    billig.lua): a Coons patch between the scaled Billig shock shape (a spline through nine
    sample points, arc-length parameterised), the body arc and two lines, niv x njv vertices
    with Roberts clustering (end0, beta = 1.1) on the west and east edges
    (src/geom/grid/sgrid.d:596-617); split into njb blocks along j.  West: supersonic
    inflow; north of the last block: simple outflow (flux); east (the body) and the axis:
    slip walls.  Returns (blocks, gas record, info with Minf and the probe point)."""
    from gdtk_amd import chem, data
    d = data.load("tpg-nitrogen-2sp")
    rec = chem.gas_record(d)
    p_inf, T_inf, u_inf = 500.0, 700.0, 5000.0
    Rg = R_UNIVERSAL / d["M"][0]
    cp = Rg * _nasa9_cp_over_R(d["thermo"][0], T_inf)
    a_inf = np.sqrt(cp / (cp - Rg) * Rg * T_inf)
    Minf = u_inf / a_inf
    R, xc, scale = 0.045, 0.045, 0.95
    ys = np.array([0.0, 0.2, 0.4, 0.6, 1.0, 1.4, 1.6, 2.0, 2.37]) * R
    xs = billig_shock_x(ys, R, Minf)
    pts = np.stack([-scale * xs + xc, scale * ys], axis=1)
    shock = _arc_length_parameterized(_bezier_spline_path(pts))
    a_pt, b_pt = np.array([xc - R, 0.0]), np.array([xc, R])

    def body(t):
        th = 0.5 * np.pi * np.asarray(t, dtype=float)
        return np.stack([xc - R * np.cos(th), R * np.sin(th)], axis=1)
    r = np.linspace(0.0, 1.0, niv)
    sj = roberts(np.linspace(0.0, 1.0, njv), True, False, 1.1)
    grid = coons_patch(_line(pts[0], a_pt), _line(pts[-1], b_pt), shock, body, r, sj)
    nj_block = (njv - 1) // njb
    ni = niv - 1
    Y = np.array([1.0, 0.0])
    blocks = []
    for bj in range(njb):
        sub = grid[bj * nj_block:(bj + 1) * nj_block + 1]
        v = np.zeros((sub.shape[0] * sub.shape[1], 3))
        v[:, 0] = sub[..., 0].reshape(-1)
        v[:, 1] = sub[..., 1].reshape(-1)
        n = ni * nj_block
        b = BlockCase(2, ni, nj_block, 1, v, np.full(n, 5.0), np.full(n, 300.0), np.zeros((n, 3)),
                      [BC_INFLOW_SUPERSONIC, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL])
        b.massf = np.tile(Y, (n, 1))
        b.inflow = np.zeros((6, 5))
        b.inflow[WEST] = [p_inf, T_inf, u_inf, 0.0, 0.0]
        b.inflow_massf = np.tile(Y, (6, 1))
        if bj > 0:
            b.bc[SOUTH] = BC_EXCHANGE
            b.other_block[SOUTH] = bj - 1
            b.other_face[SOUTH] = NORTH
        if bj < njb - 1:
            b.bc[NORTH] = BC_EXCHANGE
            b.other_block[NORTH] = bj + 1
            b.other_face[NORTH] = SOUTH
        else:
            b.bc[NORTH] = BC_OUTFLOW_SIMPLE_FLUX
        blocks.append(b)
    return blocks, rec, dict(Minf=Minf, a_inf=a_inf, probe=(-0.010, 0.0), max_time=100.0e-6, dt_init=5.0e-9)


def cylinder_blocks(ni=256, nj_block=128, njb=8, R=0.01, mach=10.0, p_inf=500.0, T_inf=300.0, margin=1.15,
                    beta_i=1.0):
    """The coupled 2-D Mach-10 cylinder: 5-species air at p = 500 Pa, T = 300 K, Mach 10, on a
    cylinder of radius R = 1 cm; a grid that fits the expected shock (Billig's correlation
    scaled outwards by `margin`), ni cells from the outer boundary to the body and njb
    blocks of nj_block cells from the stagnation line round to the top of the cylinder.

    Block faces: west = supersonic inflow (the outer boundary), east = slip wall (the
    body), south of the first block = slip wall (the symmetry line), north of the last
    block = simple outflow (extrapolate), inner faces exchange.  The grid is a Coons patch
    (transfinite interpolation) between the outer curve, the body arc and two straight
    edges, with cells clustered towards the body in i when beta_i > 1.  Every cell starts
    in the free stream.
    """
    nj = nj_block * njb
    # Boundary curves, parameter s in [0, 1] from the axis (s = 0) to the top.
    s = np.linspace(0.0, 1.0, nj + 1)
    xc = 0.0
    ang = np.pi / 2 * s                    # body arc from the nose (-R, 0) to (0, R)
    body = np.stack([xc - R * np.cos(ang), R * np.sin(ang)], axis=1)
    # outer curve: the Billig shock (scaled) from y = 0 to y = 2.5 R, parameterised by
    # arc length so the cells are even along it
    yy = np.linspace(0.0, 2.5 * R, 4001)
    xs = xc - margin * billig_shock_x(yy, R, mach)
    seg = np.hypot(np.diff(xs), np.diff(yy))
    arc = np.concatenate([[0.0], np.cumsum(seg)])
    outer = np.stack([np.interp(s * arc[-1], arc, xs), np.interp(s * arc[-1], arc, yy)], axis=1)
    r = np.linspace(0.0, 1.0, ni + 1)
    if beta_i > 1.0:
        # Roberts-style stretching towards r = 1 (the body)
        bp = (beta_i + 1.0) / (beta_i - 1.0)
        eta = 1.0 - r
        r = 1.0 - ((beta_i + 1.0) - (beta_i - 1.0) * bp ** (1.0 - eta)) / (bp ** (1.0 - eta) + 1.0)
        r = (r - r[0]) / (r[-1] - r[0])
    # Coons patch: west = outer(s), east = body(s), south = line outer(0)->body(0),
    # north = line outer(1)->body(1)
    R_, S_ = np.meshgrid(r, s, indexing="xy")        # (nj+1, ni+1)
    W = outer[:, None, :]
    Ea = body[:, None, :]
    So = outer[0][None, None, :] + r[None, :, None] * (body[0] - outer[0])[None, None, :]
    No = outer[-1][None, None, :] + r[None, :, None] * (body[-1] - outer[-1])[None, None, :]
    P00, P10, P01, P11 = outer[0], body[0], outer[-1], body[-1]
    Rr = R_[..., None]
    Ss = S_[..., None]
    grid = ((1 - Rr) * W + Rr * Ea + (1 - Ss) * So + Ss * No
            - ((1 - Rr) * (1 - Ss) * P00 + Rr * (1 - Ss) * P10 + (1 - Rr) * Ss * P01 + Rr * Ss * P11))
    # gas and free stream
    rec = air5_gas()
    from gdtk_amd import data
    d = data.load("tpg-air-5sp")
    Y = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
    Rmix = float(Y @ (R_UNIVERSAL / np.array(d["M"])))
    a_inf = np.sqrt(1.4 * Rmix * T_inf)
    u_inf = mach * a_inf
    blocks = []
    for bj in range(njb):
        sub = grid[bj * nj_block:(bj + 1) * nj_block + 1]           # (nj_block+1, ni+1, 2)
        v = np.zeros((sub.shape[0] * sub.shape[1], 3))
        v[:, 0] = sub[..., 0].reshape(-1)
        v[:, 1] = sub[..., 1].reshape(-1)
        n = ni * nj_block
        vel = np.zeros((n, 3))
        vel[:, 0] = u_inf
        b = BlockCase(2, ni, nj_block, 1, v, np.full(n, p_inf), np.full(n, T_inf), vel,
                      [BC_INFLOW_SUPERSONIC, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL, BC_SLIP_WALL])
        b.massf = np.tile(Y, (n, 1))
        b.inflow = np.zeros((6, 5))
        b.inflow[WEST] = [p_inf, T_inf, u_inf, 0.0, 0.0]
        b.inflow_massf = np.tile(Y, (6, 1))
        if bj > 0:
            b.bc[SOUTH] = BC_EXCHANGE
            b.other_block[SOUTH] = bj - 1
            b.other_face[SOUTH] = NORTH
        if bj < njb - 1:
            b.bc[NORTH] = BC_EXCHANGE
            b.other_block[NORTH] = bj + 1
            b.other_face[NORTH] = SOUTH
        else:
            b.bc[NORTH] = BC_OUTFLOW_SIMPLE_EXTRAPOLATE
        blocks.append(b)
    return blocks, rec, dict(u_inf=u_inf, a_inf=a_inf, R=R, mach=mach)


def stagnation_line(sim, blocks, nose_x=-0.01):
    """Shock stand-off and stagnation temperature along the first block's j = 0 cell row:
    the shock is where the pressure first exceeds the mean of the free-stream and the
    largest pressure on the row (from the outer boundary inwards); the stagnation
    temperature is that of the cell next to the body."""
    b = blocks[0]
    p = sim.get(0, "p").reshape(b.njc, b.nic)[0]
    T = sim.get(0, "T").reshape(b.njc, b.nic)[0]
    v = b.vertices.reshape(b.njc + 1, b.nic + 1, 3)
    xcen = 0.25 * (v[0, :-1, 0] + v[0, 1:, 0] + v[1, :-1, 0] + v[1, 1:, 0])
    thr = 0.5 * (p[0] + p.max())
    k = int(np.argmax(p > thr))
    if k > 0:
        f = (thr - p[k - 1]) / (p[k] - p[k - 1])
        xs = xcen[k - 1] + f * (xcen[k] - xcen[k - 1])
    else:
        xs = xcen[0]
    return dict(standoff=nose_x - xs, T_stag=float(T[-1]), T_max=float(T.max()), p_max=float(p.max()))
