// block.hpp -- device-resident structured block: HBM layout shared by host and kernels.
//
// Layout (all FP64, structure of arrays):
//   cells are stored on a padded grid with two ghost layers on every active side,
//   NI = nic+4, NJ = njc+4, NK = nkc+4 (3-D) or 1 (2-D); padded index
//   pc(i,j,k) = (i+2) + NI*((j+2) + NJ*(k+2)) for active-cell indices i,j,k (ghost layer n
//   on the west side sits at i = -1-n, on the east side at i = nic+n).  With ghosts at
//   those positions every face stencil (L1, L0, R0, R1) is the same index pattern.
//   Flow state: FS_NBASE field planes of ncell_pad doubles (two copies, ping-pong).
//   Conserved quantities U: ncq planes (two copies, ping-pong as the reference swaps
//   U[0] and U[2]); dUdt0: ncq planes.
//   Cell lengths iLength/jLength/kLength: 3 planes (ghosts carry mirrored lengths,
//   src/eilmer/sfluidblock.d:884-1000); volume: 1 plane.
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

struct BlockDev {
    int dims, nic, njc, nkc;
    int NI, NJ, NK;
    int64_t npad;                 // NI*NJ*NK
    int ncq;
    const double* fs_in;          // FS_NBASE planes
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
    double inflow[6][8];          // rho, vx, vy, vz, u, p, T, a
    int* counters;                // [0] decode failures, [1] invalid cells
};

__host__ __device__ inline int64_t pidx(const BlockDev& b, int i, int j, int k) {
    return (int64_t)(i + NGHOST) + (int64_t)b.NI * ((int64_t)(j + NGHOST) + (int64_t)b.NJ * (int64_t)(k + (b.dims == 3 ? NGHOST : 0)));
}

}  // namespace gdtk
