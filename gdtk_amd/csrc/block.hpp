// block.hpp -- device-resident structured block: HBM layout shared by host and kernels.
//
// Layout (all FP64, structure of arrays):
//   cells are stored on a padded grid with two ghost layers on every active side,
//   NI = nic+4, NJ = njc+4, NK = nkc+4 (3-D) or 1 (2-D); padded index
//   pc(i,j,k) = (i+2) + NI*((j+2) + NJ*(k+2)) for active-cell indices i,j,k (ghost layer n
//   on the west side sits at i = -1-n, on the east side at i = nic+n).  With ghosts at
//   those positions every face stencil (L1, L0, R0, R1) is the same index pattern.
//   Flow state: nfs field planes of ncell_pad doubles (two copies, ping-pong): rho, vx,
//   vy, vz, u, p, T, a, and for a multi-species gas massf[nsp] and rho_s[nsp].
//   Conserved quantities U: ncq planes (two copies, ping-pong as the reference swaps
//   U[0] and U[2]); dUdt0: ncq planes.
//   Cell lengths iLength/jLength/kLength: 3 planes (ghosts carry mirrored lengths,
//   src/eilmer/sfluidblock.d:884-1000); volume: 1 plane.
//   Shock indicator (adaptive flux calculator): one byte per padded cell.
//   Face geometry per index direction d: 10 planes (n, t1, t2, area) over
//   (nic+1)*njc*nkc (d=0), nic*(njc+1)*nkc (d=1), nic*njc*(nkc+1) (d=2), face (fi,fj,fk)
//   at fi + nfi*(fj + nfj*fk).
#pragma once
#include <cstdint>

namespace gdtk {

enum BCKind : int {
    BC_SLIP_WALL = 0,
    BC_INFLOW_SUPERSONIC = 1,
    BC_OUTFLOW_SIMPLE_FLUX = 2,
    BC_OUTFLOW_SIMPLE_EXTRAPOLATE = 3,
    BC_EXCHANGE = 4,
};

constexpr int NGHOST = 2;
// Flow-state planes: rho, vx, vy, vz, u, p, T, a, then for a multi-species gas the mass
// fractions and the species densities (nsp planes each).
constexpr int FS_BASE_PLANES = 8;
constexpr int FS_MAX_PLANES = FS_BASE_PLANES + 2 * 8;

struct BlockDev {
    int dims, nic, njc, nkc;
    int NI, NJ, NK;
    int64_t npad;                 // NI*NJ*NK
    int ncq;
    int nsp, nfs;                 // species (1: single-species gas) and flow-state planes
    const double* fs_in;          // nfs planes
    double* fs_out;
    const double* U0;             // ncq planes
    double* Unew;
    double* dUdt0;
    const double* len;            // 3 planes
    const double* vol;            // 1 plane
    const double* fgeom[3];       // 10 planes each
    double* flux[3];              // ncq planes each (face-flux kernels)
    int nf[3][3];                 // face-array extents per direction
    int64_t nface[3];
    int bc_kind[6];
    double inflow[6][FS_MAX_PLANES];  // an inflow face's flow state, plane order
    int* counters;                // [0] decode failures, [1] invalid cells
    // Shock indicator per padded cell (0 or 1; adaptive flux calculator only, else null):
    // interior cells hold the last detection, ghost cells the copy made by the ghost fill
    // of the stage that detected.
    unsigned char* shock;
    // The PJ detector's decision per face of each direction (face-array order; adaptive
    // flux calculator only): written and read by the detection launches alone.
    unsigned char* fshock[3];
};

__host__ __device__ inline int64_t pidx(const BlockDev& b, int i, int j, int k) {
    return (int64_t)(i + NGHOST) + (int64_t)b.NI * ((int64_t)(j + NGHOST) + (int64_t)b.NJ * (int64_t)(k + (b.dims == 3 ? NGHOST : 0)));
}

// In-face coordinates of a block side: west/east (j, k), south/north (i, k), bottom/top
// (i, j); extents A x B (B = 1 on the sides of a 2-D block).
__host__ __device__ inline void side_extents(int side, int nic, int njc, int nkc, int& A, int& B) {
    if (side <= 1) { A = njc; B = nkc; }
    else if (side <= 3) { A = nic; B = nkc; }
    else { A = nic; B = njc; }
}

// Source cell (active-cell indices of the source block) of ghost layer n at in-face
// position (a, b) of a full-face-copy exchange, following the case tables of
// set_up_cell_mapping_phase0 (src/eilmer/bc/ghost_cell_effect/full_face_copy.d:696-1480):
// orientation 0 keeps (a, b), orientation 2 reverses both, orientations 1 and 3 are the
// quarter turns (b, B'-1-a) and (A'-1-b, a) -- in that order when the destination side's
// in-face coordinates and its outward normal are right-handed (east, south, top), in the
// opposite order otherwise -- and the first source coordinate is reversed when both sides
// have the same handedness.  (A', B') are the source side's extents; 2-D sides use only
// the first coordinate.  Source layer n lies n cells inside the source side.
__host__ __device__ inline void ffc_source(int dims, int dside, int sside, int orient, int a, int b, int n, int snic,
                                           int snjc, int snkc, int& i, int& j, int& k) {
    const bool dright = (dside == 1) || (dside == 2) || (dside == 5);
    const bool sright = (sside == 1) || (sside == 2) || (sside == 5);
    int A, B;
    side_extents(sside, snic, snjc, snkc, A, B);
    int p, q;
    if (dims == 2 || orient == 0) { p = a; q = (dims == 2) ? 0 : b; }
    else if (orient == 2) { p = A - 1 - a; q = B - 1 - b; }
    else if ((orient == 1) == dright) { p = b; q = B - 1 - a; }
    else { p = A - 1 - b; q = a; }
    if (dright == sright) p = A - 1 - p;
    switch (sside) {
    case 0: i = n; j = p; k = q; break;
    case 1: i = snic - 1 - n; j = p; k = q; break;
    case 2: j = n; i = p; k = q; break;
    case 3: j = snjc - 1 - n; i = p; k = q; break;
    case 4: k = n; i = p; j = q; break;
    default: k = snkc - 1 - n; i = p; j = q; break;
    }
}

// One side of one block in the batched ghost-cell launch: its boundary kind, the
// exchange source (block, side, orientation) and the side's first ghost cell in the
// launch (ghost cells ordered (face position, layer) within a side).
struct SideDev {
    int blk, side, kind;
    int oblk, oside, orient;
    int na, nb;
    int64_t start;
};

// The faces of every direction of every block in one launch: per direction d, the first
// launch index of each block's faces (multiples of the 256-thread workgroup) and the
// direction's launch size; the directions follow each other in the launch.
struct FaceBatch {
    const int64_t* start[3];
    int64_t total[3];
};

// Direction of the workgroup starting at launch index w0, which becomes the index within
// that direction.
__host__ __device__ inline int face_batch_dir(const FaceBatch& fb, int64_t& w0) {
    int d = 0;
    if (w0 >= fb.total[0]) {
        w0 -= fb.total[0];
        d = 1;
        if (w0 >= fb.total[1]) {
            w0 -= fb.total[1];
            d = 2;
        }
    }
    return d;
}

// Batched stage launch over every block: per-block device records for this stage (their
// buffers already point at this stage's input and output copies), the prefix sums of the
// blocks' tile counts, and the per-XCD tile queues (8 counters, TILE_QSTRIDE apart).
constexpr int TILE_QSTRIDE = 32;
struct StageBatch {
    const BlockDev* tab;
    const int* tile_start;
    unsigned* queue;
    int nblocks, total_tiles, per_xcd, KC;
};

}  // namespace gdtk
