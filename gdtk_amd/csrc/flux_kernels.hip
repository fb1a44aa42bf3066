// flux_kernels.hip -- HIP kernels of the structured-block explicit update (gfx950).
//
// Face-per-thread path: ghost fill, the face flux of every direction into face arrays
// (flux_face.hpp, instantiated in flux_face_*.hip), then a cell kernel that integrates the
// face fluxes, applies the predictor-corrector stage update and decodes the new flow state.
// The fused LDS-tiled stage kernel lives in fused_stage.hip and shares the device physics in
// gdtk_device.hpp.
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "gdtk_device.hpp"
#include "kernel_common.hpp"
#include "launchers.hpp"
#include "flux_face.hpp"

namespace gdtk {

// Padded index of ghost layer `layer` at in-face position (a, bb) of a block side.
__device__ __forceinline__ int64_t ghost_pidx(const BlockDev& b, int side, int a, int bb, int layer) {
    if (side == 0) return pidx(b, -1 - layer, a, bb);
    if (side == 1) return pidx(b, b.nic + layer, a, bb);
    if (side == 2) return pidx(b, a, -1 - layer, bb);
    if (side == 3) return pidx(b, a, b.njc + layer, bb);
    if (side == 4) return pidx(b, a, bb, -1 - layer);
    return pidx(b, a, bb, b.nkc + layer);
}

// Ghost-cell effect of one ghost cell (layer `layer` at in-face position (a, bb) of side
// `side`) for the boundary kinds that read the block's own cells:
// internal_copy_then_reflect.d:111-134, flow_state_copy.d:92-107,
// extrapolate_copy.d:124-140 (xOrder 0).
// With `shock`, the ghost cell also takes its source's shock indicator (0 for an inflow), as
// the reference's ghost cells copy the whole flow state, fs.S included.
__device__ __forceinline__ void ghost_one(const BlockDev& b, double* fs, int side, int a, int bb, int layer,
                                          bool shock) {
    int i0, j0, k0;  // face location in the side
    int d;           // index direction normal to the side
    int fi, fj, fk;  // face index
    if (side <= 1) {
        d = 0; j0 = a; k0 = bb; i0 = 0;
        fi = (side == 0) ? 0 : b.nic; fj = j0; fk = k0;
    } else if (side <= 3) {
        d = 1; i0 = a; k0 = bb; j0 = 0;
        fi = i0; fj = (side == 2) ? 0 : b.njc; fk = k0;
    } else {
        d = 2; i0 = a; j0 = bb; k0 = 0;
        fi = i0; fj = j0; fk = (side == 4) ? 0 : b.nkc;
    }
    const bool lo = (side % 2) == 0;
    const int nd = (d == 0) ? b.nic : (d == 1) ? b.njc : b.nkc;
    // Ghost position and interior source position along d.
    int g_at = lo ? (-1 - layer) : (nd + layer);
    int kind = b.bc_kind[side];
    int src_at;
    if (kind == BC_SLIP_WALL) src_at = lo ? layer : (nd - 1 - layer);
    else src_at = lo ? 0 : (nd - 1);
    auto at = [&](int s) -> int64_t {
        if (d == 0) return pidx(b, s, j0, k0);
        if (d == 1) return pidx(b, i0, s, k0);
        return pidx(b, i0, j0, s);
    };
    int64_t g = at(g_at);
    const int64_t np = b.npad;
    if (kind == BC_INFLOW_SUPERSONIC) {
        for (int f = 0; f < b.nfs; ++f) fs[g + f * np] = b.inflow[side][f];
        if (shock) b.shock[g] = 0;
        return;
    }
    if (kind == BC_EXCHANGE) return;
    int64_t s = at(src_at);
    for (int f = 0; f < b.nfs; ++f) fs[g + f * np] = fs[s + f * np];
    if (shock) b.shock[g] = b.shock[s];
    if (kind == BC_SLIP_WALL) {
        // ghost_cell.d:33-43
        FaceGeom fg = load_face(b.fgeom[d], b.nface[d], face_index(b, d, fi, fj, fk));
        Vec3d v{fs[s + FS_VX * np], fs[s + FS_VY * np], fs[s + FS_VZ * np]};
        Vec3d l = to_local(v, fg);
        l.x = -(l.x);
        Vec3d w = to_global(l, fg);
        fs[g + FS_VX * np] = w.x;
        fs[g + FS_VY * np] = w.y;
        fs[g + FS_VZ * np] = w.z;
    }
}

// Ghost cells of every side of every block in one launch: full-face-copy exchange from the
// source block's cells (simcore_exchange.d:96-135 -> full_face_copy.d:1888-1900, the
// same-process copy of the whole flow state) and the other ghost-cell effects.  Each block
// record's fs_in is the flow-state copy this stage reads, so the ghost cells are written
// there.  mode 1 copies the exchange sides' cell lengths instead (the geometry exchange
// of full_face_copy.d:1576-1625, done once); mode 2 is mode 0 with the shock indicators
// (the fill that precedes shock detection).
__global__ void k_ghost_batch(const BlockDev* __restrict__ tab, const SideDev* __restrict__ sides, int nsides,
                              int64_t total, int mode) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int lo = 0, hi = nsides - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sides[mid].start <= t) lo = mid;
        else hi = mid - 1;
    }
    const SideDev sd = sides[lo];
    const int64_t r = t - sd.start;
    const int layer = (int)(r % NGHOST);
    const int64_t face = r / NGHOST;
    const int a = (int)(face % sd.na), bb = (int)(face / sd.na);
    const BlockDev& b = tab[sd.blk];
    if (sd.kind == BC_EXCHANGE) {
        const BlockDev& o = tab[sd.oblk];
        int i, j, k;
        ffc_source(b.dims, sd.side, sd.oside, sd.orient, a, bb, layer, o.nic, o.njc, o.nkc, i, j, k);
        const int64_t src = pidx(o, i, j, k);
        const int64_t dst = ghost_pidx(b, sd.side, a, bb, layer);
        if (mode != 1) {
            double* fs = const_cast<double*>(b.fs_in);
            for (int f = 0; f < b.nfs; ++f) fs[dst + f * b.npad] = o.fs_in[src + f * o.npad];
            if (mode == 2) b.shock[dst] = o.shock[src];
        } else {
            double* len = const_cast<double*>(b.len);
            for (int f = 0; f < 3; ++f) len[dst + f * b.npad] = o.len[src + f * o.npad];
        }
    } else if (mode != 1) {
        ghost_one(b, const_cast<double*>(b.fs_in), sd.side, a, bb, layer, mode == 2);
    }
}

// PJ shock detector (shockdetectors.d:21-113) for the face between cells cl and cr with
// face frame g: compression (uR - uL)/a_min and shear max(|vL - vR|, |wL - wR|)/a_mean.
__device__ __forceinline__ bool pj_shock(const double* fs, int64_t np, int64_t cl, int64_t cr, const FaceGeom& g,
                                         double comp_tol, double shear_tol) {
    const Vec3d vl{fs[cl + FS_VX * np], fs[cl + FS_VY * np], fs[cl + FS_VZ * np]};
    const Vec3d vr{fs[cr + FS_VX * np], fs[cr + FS_VY * np], fs[cr + FS_VZ * np]};
    const double aL = fs[cl + FS_A * np], aR = fs[cr + FS_A * np];
    const double uL = vl.x * g.n.x + vl.y * g.n.y + vl.z * g.n.z;
    const double uR = vr.x * g.n.x + vr.y * g.n.y + vr.z * g.n.z;
    const double a_min = (aL < aR) ? aL : aR;
    const double comp = ((uR - uL) / a_min);
    const double vL = vl.x * g.t1.x + vl.y * g.t1.y + vl.z * g.t1.z;
    const double vR = vr.x * g.t1.x + vr.y * g.t1.y + vr.z * g.t1.z;
    const double wL = vl.x * g.t2.x + vl.y * g.t2.y + vl.z * g.t2.z;
    const double wR = vr.x * g.t2.x + vr.y * g.t2.y + vr.z * g.t2.z;
    const double sound_speed = 0.5 * (aL + aR);
    const double shear_y = fabs(vL - vR) / sound_speed;
    const double shear_z = fabs(wL - wR) / sound_speed;
    const double shear = fmax(shear_y, shear_z);
    return (shear < shear_tol) && (comp < comp_tol);
}

// Shock detection (detect_shock_points then shock_faces_to_cells, fluidblock.d:478-520,
// 568-580) in two launches.  Pass 1, one thread per face of every direction of every block
// (FaceBatch order, as the face-flux launch): the PJ detector's decision for the face, or 1
// for a face on an outflow_simple_extrapolate boundary when damped_outflow is set.  Each
// face's frame and its two cells are read once, where a cell-per-thread detection reads
// every interior face twice from workgroups that run on different XCDs.
template <int DIM>
__global__ __launch_bounds__(256) void k_shock_faces_batch(const BlockDev* __restrict__ tab, FaceBatch fb,
                                                           int nblocks, double comp_tol, double shear_tol,
                                                           int damped_outflow) {
    int64_t w0 = (int64_t)blockIdx.x * blockDim.x;
    const int d = __builtin_amdgcn_readfirstlane(face_batch_dir(fb, w0));
    const int64_t* __restrict__ start = fb.start[d];
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= w0) lo = mid;
        else hi = mid - 1;
    }
    const int q = __builtin_amdgcn_readfirstlane(lo);
    const BlockDev& b = tab[q];
    const int64_t f = w0 - start[q] + threadIdx.x;
    if (f >= b.nface[d]) return;
    const int fi = (int)(f % b.nf[d][0]);
    const int64_t rest = f / b.nf[d][0];
    const int fj = (int)(rest % b.nf[d][1]);
    const int fk = (int)(rest / b.nf[d][1]);
    const int pos = (d == 0) ? fi : (d == 1) ? fj : fk;
    const int nd = (d == 0) ? b.nic : (d == 1) ? b.njc : b.nkc;
    const int side = (pos == 0) ? 2 * d : (pos == nd) ? 2 * d + 1 : -1;
    bool shocked;
    if (damped_outflow && side >= 0 && b.bc_kind[side] == BC_OUTFLOW_SIMPLE_EXTRAPOLATE) {
        shocked = true;
    } else {
        const int64_t cl = pidx(b, fi - (d == 0), fj - (d == 1), fk - (d == 2));
        const int64_t cr = pidx(b, fi, fj, fk);
        shocked = pj_shock(b.fs_in, b.npad, cl, cr, load_face(b.fgeom[d], b.nface[d], f), comp_tol, shear_tol);
    }
    b.fshock[d][f] = shocked ? 1 : 0;
}

// Pass 2, one thread per interior cell: the cell is shocked if any of its faces is.  The
// strict rule (fluidblock.d:582-605) is applied where the flux kernels read the indicators:
// a face is shocked if either adjacent cell is (a face the detector marks has its interior
// cells marked too).
template <int DIM>
__global__ __launch_bounds__(256) void k_shock_cells_batch(const BlockDev* __restrict__ tab,
                                                           const int64_t* __restrict__ start, int nblocks) {
    const int64_t w0 = (int64_t)blockIdx.x * blockDim.x;
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= w0) lo = mid;
        else hi = mid - 1;
    }
    const int q = __builtin_amdgcn_readfirstlane(lo);
    const BlockDev& b = tab[q];
    const int64_t t = w0 - start[q] + threadIdx.x;
    if (t >= (int64_t)b.nic * b.njc * b.nkc) return;
    const int i = (int)(t % b.nic);
    const int64_t rest = t / b.nic;
    const int j = (int)(rest % b.njc);
    const int k = (int)(rest / b.njc);
    unsigned s = b.fshock[0][face_index(b, 0, i, j, k)] | b.fshock[0][face_index(b, 0, i + 1, j, k)] |
                 b.fshock[1][face_index(b, 1, i, j, k)] | b.fshock[1][face_index(b, 1, i, j + 1, k)];
    if (DIM == 3) s |= b.fshock[2][face_index(b, 2, i, j, k)] | b.fshock[2][face_index(b, 2, i, j, k + 1)];
    b.shock[pidx(b, i, j, k)] = s ? 1 : 0;
}

// Per interior cell: time derivatives (fvcell.d:824-854), the pc stage update
// (simcore_gasdynamic_step.d:1225-1321), decode (fvcell.d:586-821) and the
// invalid-cell check (fluidblock.d:607-676).
template <int DIM, int STAGE>
__device__ __forceinline__ void cell_update_one(const BlockDev& b, const IdealGasDev& gas, double dt, double low_T,
                                                double max_vel, double min_T, double max_T, int64_t t) {
    using C = CQ<DIM>;
    int i = (int)(t % b.nic);
    int64_t rest = t / b.nic;
    int j = (int)(rest % b.njc);
    int k = (int)(rest / b.njc);
    int64_t c = pidx(b, i, j, k);
    // Faces in the order W, E, S, N, B, T with outsigns -1, +1.
    const int nfaces = (DIM == 3) ? 6 : 4;
    int64_t fidx[6];
    int fdir[6];
    fidx[0] = face_index(b, 0, i, j, k); fdir[0] = 0;
    fidx[1] = face_index(b, 0, i + 1, j, k); fdir[1] = 0;
    fidx[2] = face_index(b, 1, i, j, k); fdir[2] = 1;
    fidx[3] = face_index(b, 1, i, j + 1, k); fdir[3] = 1;
    if (DIM == 3) {
        fidx[4] = face_index(b, 2, i, j, k); fdir[4] = 2;
        fidx[5] = face_index(b, 2, i, j, k + 1); fdir[5] = 2;
    }
    double area[6];
    for (int q = 0; q < nfaces; ++q) {
        double A = b.fgeom[fdir[q]][fidx[q] + 9 * b.nface[fdir[q]]];
        area[q] = ((q % 2 == 0) ? -1 : 1) * A;
    }
    double si[C::n];
    for (int jq = 0; jq < C::n; ++jq) {
        double s = 0.0;
        for (int q = 0; q < nfaces; ++q) s -= b.flux[fdir[q]][fidx[q] + jq * b.nface[fdir[q]]] * area[q];
        si[jq] = s;
    }
    update_decode_store<DIM, STAGE>(b, gas, c, si, dt, low_T, max_vel, min_T, max_T);
}

// Cells of every block in one launch (block ranges aligned as in k_face_flux_batch); dt
// is read from device memory.
template <int DIM, int STAGE>
__global__ __launch_bounds__(256) void k_cell_update_batch(const BlockDev* __restrict__ tab,
                                                           const int64_t* __restrict__ start, int nblocks,
                                                           const double* __restrict__ dtp, CellLimits lim,
                                                           IdealGasDev gas) {
    const int64_t w0 = (int64_t)blockIdx.x * blockDim.x;
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= w0) lo = mid;
        else hi = mid - 1;
    }
    const int q = __builtin_amdgcn_readfirstlane(lo);
    const BlockDev& b = tab[q];
    const int64_t t = w0 - start[q] + threadIdx.x;
    if (t >= (int64_t)b.nic * b.njc * b.nkc) return;
    cell_update_one<DIM, STAGE>(b, gas, *dtp, lim.low_T, lim.max_vel, lim.min_T, lim.max_T, t);
}

// The block time-step estimates of many blocks in two launches (fluidblock.d:987-1084 per
// block).  k_dt_batch: each lane takes one cell's signal frequency (fvcell.d:975-1058) and its
// dt_local and cfl_local; the minimum and the maximum are reduced across the wavefront with
// lane exchanges and across the workgroup's four wavefronts through LDS, and written as the
// workgroup's partial (part[2 g], part[2 g + 1]); the block's last cell stores its signal
// (fluidblock.d:1081).  k_dt_blocks: one workgroup per block reduces the block's partials the
// same way.  Minimum and maximum are exact, so the result does not depend on the order; like
// fmin/fmax a NaN never wins.  No atomics: a single large block has thousands of
// workgroups, whose atomics on one address would serialise.  A workgroup serves 256 cells of
// one block: start[q] is block q's first launch index (a multiple of 256; start[nblocks] the
// end), base the launch index of workgroup 0.
__device__ __forceinline__ void dt_wave_reduce(double& dtl, double& cfl) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        dtl = fmin(dtl, __shfl_xor(dtl, off));
        cfl = fmax(cfl, __shfl_xor(cfl, off));
    }
}

// Across the workgroup's wavefronts; the result is valid in thread 0.
__device__ __forceinline__ void dt_group_reduce(double& dtl, double& cfl) {
    __shared__ double wdt[4], wcfl[4];
    dt_wave_reduce(dtl, cfl);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wdt[wave] = dtl;
        wcfl[wave] = cfl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int w = 1; w < nw; ++w) {
            dtl = fmin(dtl, wdt[w]);
            cfl = fmax(cfl, wcfl[w]);
        }
    }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_dt_batch(const BlockDev* __restrict__ tab, const int64_t* __restrict__ start,
                                                  int nblocks, int64_t base, double dt_current, double cfl_value,
                                                  double* part, double* last_signal) {
    const int64_t w0 = base + (int64_t)blockIdx.x * blockDim.x;
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= w0) lo = mid;
        else hi = mid - 1;
    }
    const int q = __builtin_amdgcn_readfirstlane(lo);
    const BlockDev& b = tab[q];
    const int64_t t = w0 - start[q] + threadIdx.x;
    const int64_t ncell = (int64_t)b.nic * b.njc * b.nkc;
    double dtl = __builtin_huge_val(), cfl = -__builtin_huge_val();
    if (t < ncell) {
        const int i = (int)(t % b.nic);
        const int64_t rest = t / b.nic;
        const int j = (int)(rest % b.njc);
        const int k = (int)(rest / b.njc);
        const int64_t np = b.npad;
        const int64_t c = pidx(b, i, j, k);
        const double* fs = b.fs_in;
        const Vec3d v{fs[c + FS_VX * np], fs[c + FS_VY * np], fs[c + FS_VZ * np]};
        const double a = fs[c + FS_A * np];
        const FaceGeom gN = load_face(b.fgeom[1], b.nface[1], face_index(b, 1, i, j + 1, k));
        const FaceGeom gE = load_face(b.fgeom[0], b.nface[0], face_index(b, 0, i + 1, j, k));
        const double un_N = fabs(v.x * gN.n.x + v.y * gN.n.y + v.z * gN.n.z);
        const double un_E = fabs(v.x * gE.n.x + v.y * gE.n.y + v.z * gE.n.z);
        double signal = 0.0;
        const double signalN = (un_N + a) / b.len[c + 1 * np];
        signal = fmax(signal, signalN);
        const double signalE = (un_E + a) / b.len[c + 0 * np];
        signal = fmax(signal, signalE);
        if (DIM == 3) {
            const FaceGeom gT = load_face(b.fgeom[2], b.nface[2], face_index(b, 2, i, j, k + 1));
            const double un_T = fabs(v.x * gT.n.x + v.y * gT.n.y + v.z * gT.n.z);
            const double signalT = (un_T + a) / b.len[c + 2 * np];
            signal = fmax(signal, signalT);
        }
        cfl = dt_current * signal;
        dtl = cfl_value / signal;
        if (t == ncell - 1) last_signal[q] = signal;
    }
    dt_group_reduce(dtl, cfl);
    if (threadIdx.x == 0) {
        part[2 * (int64_t)blockIdx.x] = dtl;
        part[2 * (int64_t)blockIdx.x + 1] = cfl;
    }
}

// One workgroup per block: out[3q..3q+2] = {min dt_local, max cfl_local, last signal}.
__global__ __launch_bounds__(256) void k_dt_blocks(const int64_t* __restrict__ start, int64_t base,
                                                   const double* __restrict__ part,
                                                   const double* __restrict__ last_signal, double* out) {
    const int q = blockIdx.x;
    const int64_t g0 = (start[q] - base) / 256, g1 = (start[q + 1] - base) / 256;
    double dtl = __builtin_huge_val(), cfl = -__builtin_huge_val();
    for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
        dtl = fmin(dtl, part[2 * g]);
        cfl = fmax(cfl, part[2 * g + 1]);
    }
    dt_group_reduce(dtl, cfl);
    if (threadIdx.x == 0) {
        out[3 * q] = dtl;
        out[3 * q + 1] = cfl;
        out[3 * q + 2] = last_signal[q];
    }
}

// Encode U from the flow state (fvcell.d:511-583), single species: used at start-up.
template <int DIM>
__global__ void k_encode(BlockDev b, double* U) {
    using C = CQ<DIM>;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ncell = (int64_t)b.nic * b.njc * b.nkc;
    if (t >= ncell) return;
    int i = (int)(t % b.nic);
    int64_t rest = t / b.nic;
    int j = (int)(rest % b.njc);
    int k = (int)(rest / b.njc);
    const int64_t np = b.npad;
    int64_t c = pidx(b, i, j, k);
    const double* fs = b.fs_in;
    double rho = fs[c + FS_RHO * np];
    double vx = fs[c + FS_VX * np], vy = fs[c + FS_VY * np], vz = fs[c + FS_VZ * np];
    U[c + C::mass * np] = rho;
    U[c + C::xMom * np] = rho * vx;
    U[c + C::yMom * np] = rho * vy;
    if (DIM == 3) U[c + C::zMom * np] = rho * vz;
    double u = fs[c + FS_U * np];
    double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
    U[c + C::totEnergy * np] = rho * (u + ke);
}

// Decode U into the flow state without a stage update (start-up and step retries).
template <int DIM>
__global__ void k_decode(BlockDev b, IdealGasDev gas, double* U, double* fs, double low_T) {
    using C = CQ<DIM>;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ncell = (int64_t)b.nic * b.njc * b.nkc;
    if (t >= ncell) return;
    int i = (int)(t % b.nic);
    int64_t rest = t / b.nic;
    int j = (int)(rest % b.njc);
    int k = (int)(rest / b.njc);
    const int64_t np = b.npad;
    int64_t c = pidx(b, i, j, k);
    double Uc[C::n];
    for (int q = 0; q < C::n; ++q) Uc[q] = U[c + q * np];
    double rho, u, p, T, a;
    Vec3d vel;
    bool reenc;
    if (decode_ideal<DIM>(gas, Uc, low_T, rho, vel, u, p, T, a, reenc)) {
        atomicAdd(&b.counters[0], 1);
        return;
    }
    if (reenc)
        for (int q = 0; q < C::n; ++q) U[c + q * np] = Uc[q];
    fs[c + FS_RHO * np] = rho;
    fs[c + FS_VX * np] = vel.x;
    fs[c + FS_VY * np] = vel.y;
    fs[c + FS_VZ * np] = vel.z;
    fs[c + FS_U * np] = u;
    fs[c + FS_P * np] = p;
    fs[c + FS_T * np] = T;
    fs[c + FS_A * np] = a;
}

//---------------------------------------------------------------------------------
// Launchers

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_face_flux_batch(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                 int numerics, double eps, const IdealGasDev& gas, hipStream_t s) {
    const int64_t total = fb.total[0] + fb.total[1] + fb.total[2];
    if (total <= 0) return hipSuccess;
    const bool lmr = numerics == 1;
    if (dims == 3)
        return lmr ? launch_face_flux_3d_lmr(tab, fb, nblocks, flux_calc, eps, gas, total, s)
                   : launch_face_flux_3d_e4(tab, fb, nblocks, flux_calc, eps, gas, total, s);
    return lmr ? launch_face_flux_2d_lmr(tab, fb, nblocks, flux_calc, eps, gas, total, s)
               : launch_face_flux_2d_e4(tab, fb, nblocks, flux_calc, eps, gas, total, s);
}

hipError_t launch_shock_detect(const BlockDev* tab, const FaceBatch& fb, const int64_t* cstart, int nblocks,
                               int64_t ctotal, int dims, double comp_tol, double shear_tol, int damped_outflow,
                               hipStream_t s) {
    const int64_t ftotal = fb.total[0] + fb.total[1] + fb.total[2];
    if (ctotal <= 0 || ftotal <= 0) return hipSuccess;
    dim3 fgrid(nblk(ftotal, 256)), cgrid(nblk(ctotal, 256)), blk(256);
    if (dims == 3) {
        hipLaunchKernelGGL((k_shock_faces_batch<3>), fgrid, blk, 0, s, tab, fb, nblocks, comp_tol, shear_tol,
                           damped_outflow);
        hipLaunchKernelGGL((k_shock_cells_batch<3>), cgrid, blk, 0, s, tab, cstart, nblocks);
    } else {
        hipLaunchKernelGGL((k_shock_faces_batch<2>), fgrid, blk, 0, s, tab, fb, nblocks, comp_tol, shear_tol,
                           damped_outflow);
        hipLaunchKernelGGL((k_shock_cells_batch<2>), cgrid, blk, 0, s, tab, cstart, nblocks);
    }
    return hipGetLastError();
}

hipError_t launch_cell_update_batch(const BlockDev* tab, const int64_t* start, int nblocks, int64_t total, int dims,
                                   int stage, const double* dtp, const CellLimits& lim, const IdealGasDev& gas,
                                   hipStream_t s) {
    if (total <= 0) return hipSuccess;
    dim3 grid(nblk(total, 256)), blk(256);
    if (dims == 3) {
        if (stage == 1) hipLaunchKernelGGL((k_cell_update_batch<3, 1>), grid, blk, 0, s, tab, start, nblocks, dtp, lim, gas);
        else hipLaunchKernelGGL((k_cell_update_batch<3, 2>), grid, blk, 0, s, tab, start, nblocks, dtp, lim, gas);
    } else {
        if (stage == 1) hipLaunchKernelGGL((k_cell_update_batch<2, 1>), grid, blk, 0, s, tab, start, nblocks, dtp, lim, gas);
        else hipLaunchKernelGGL((k_cell_update_batch<2, 2>), grid, blk, 0, s, tab, start, nblocks, dtp, lim, gas);
    }
    return hipGetLastError();
}

hipError_t launch_ghost_batch(const BlockDev* tab, const SideDev* sides, int nsides, int64_t total, int mode,
                              hipStream_t s) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_batch, dim3(nblk(total, 256)), dim3(256), 0, s, tab, sides, nsides, total, mode);
    return hipGetLastError();
}

hipError_t launch_dt_batch(const BlockDev* tab, const int64_t* start, int nblocks, int64_t base, int64_t count,
                           int dims, double dt_current, double cfl, double* part, double* last_signal, double* out,
                           hipStream_t s) {
    if (count <= 0 || count % 256 != 0 || nblocks <= 0) return hipErrorInvalidValue;
    const unsigned nwg = (unsigned)(count / 256);
    if (dims == 3)
        hipLaunchKernelGGL((k_dt_batch<3>), dim3(nwg), dim3(256), 0, s, tab, start, nblocks, base, dt_current, cfl,
                           part, last_signal);
    else
        hipLaunchKernelGGL((k_dt_batch<2>), dim3(nwg), dim3(256), 0, s, tab, start, nblocks, base, dt_current, cfl,
                           part, last_signal);
    hipLaunchKernelGGL(k_dt_blocks, dim3((unsigned)nblocks), dim3(256), 0, s, start, base, part, last_signal, out);
    return hipGetLastError();
}

hipError_t launch_encode(const BlockDev& b, double* U, hipStream_t s) {
    int64_t n = (int64_t)b.nic * b.njc * b.nkc;
    if (b.dims == 3) hipLaunchKernelGGL((k_encode<3>), dim3(nblk(n, 256)), dim3(256), 0, s, b, U);
    else hipLaunchKernelGGL((k_encode<2>), dim3(nblk(n, 256)), dim3(256), 0, s, b, U);
    return hipGetLastError();
}

hipError_t launch_decode(const BlockDev& b, const IdealGasDev& gas, double* U, double* fs, double low_T, hipStream_t s) {
    int64_t n = (int64_t)b.nic * b.njc * b.nkc;
    if (b.dims == 3) hipLaunchKernelGGL((k_decode<3>), dim3(nblk(n, 256)), dim3(256), 0, s, b, gas, U, fs, low_T);
    else hipLaunchKernelGGL((k_decode<2>), dim3(nblk(n, 256)), dim3(256), 0, s, b, gas, U, fs, low_T);
    return hipGetLastError();
}

}  // namespace gdtk
