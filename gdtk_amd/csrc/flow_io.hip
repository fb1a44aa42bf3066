// flow_io.hip -- device side of the legacy flow-solution files: the flow-state planes of a
// block packed into the file's cell records (one record of nv doubles per active cell, in
// the reference's cell order) and unpacked back.  The record layout is that of
// cell_data_as_string / cell_data_to_raw_binary (src/eilmer/fluidblockio_old.d:1487-1691)
// for this path: pos.x pos.y pos.z volume rho vel.x vel.y vel.z p a mu k mu_t k_t S,
// massf[nsp] (1.0 for a single-species gas), dt_chem (multi-species gas), u, T.
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "gdtk_device.hpp"
#include "launchers.hpp"

namespace gdtk {

namespace {

// Record slot of the fields after the mass fractions.
__host__ __device__ inline int rec_u(int nsp) { return 15 + nsp + (nsp > 1 ? 1 : 0); }

// One thread per record value: value v of cell c is t = c*nv + v, so the record buffer is
// written with unit stride; the reads gather from the planes.
__global__ void k_pack_flow(BlockDev b, const double* __restrict__ pos, const double* __restrict__ mu,
                            const double* __restrict__ kc, const double* __restrict__ dtc, double* __restrict__ rec,
                            int nv) {
    const int64_t ncell = (int64_t)b.nic * b.njc * b.nkc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * nv) return;
    const int64_t c = t / nv;
    const int v = (int)(t - c * nv);
    const int i = (int)(c % b.nic);
    const int64_t rest = c / b.nic;
    const int64_t pc = pidx(b, i, (int)(rest % b.njc), (int)(rest / b.njc));
    const int64_t np = b.npad;
    const double* fs = b.fs_in;
    const int nsp = b.nsp;
    const int iu = rec_u(nsp);
    double x = 0.0;
    if (v < 3) x = pos[c + v * ncell];
    else if (v == 3) x = b.vol[pc];
    else if (v == 4) x = fs[pc + FS_RHO * np];
    else if (v <= 7) x = fs[pc + (FS_VX + v - 5) * np];
    else if (v == 8) x = fs[pc + FS_P * np];
    else if (v == 9) x = fs[pc + FS_A * np];
    else if (v == 10) x = mu ? mu[c] : 0.0;
    else if (v == 11) x = kc ? kc[c] : 0.0;
    else if (v <= 13) x = 0.0;  // mu_t, k_t: laminar
    else if (v == 14) x = b.shock ? (double)b.shock[pc] : 0.0;
    else if (v < 15 + nsp) x = (nsp > 1) ? fs[pc + (FS_BASE_PLANES + v - 15) * np] : 1.0;
    else if (v < iu) x = dtc ? dtc[c] : -1.0;
    else if (v == iu) x = fs[pc + FS_U * np];
    else x = fs[pc + FS_T * np];
    rec[t] = x;
}

// The cell state a flow file gives (scan_cell_data_from_fixed_order_string,
// fluidblockio_old.d:1693-1800): rho, vel, p, a, u, T and the mass fractions go to the
// flow-state planes (the species densities as rho * massf), S to the shock indicator, mu
// and k to the caller's planes.  Position, volume and dt_chem are not taken (the cold start
// of a run reads without overwriting the geometry and resets dt_chem, simcore.d:299-336).
__global__ void k_unpack_flow(BlockDev b, const double* __restrict__ rec, double* __restrict__ fs, double* mu,
                              double* kc, int nv) {
    const int64_t ncell = (int64_t)b.nic * b.njc * b.nkc;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int i = (int)(c % b.nic);
    const int64_t rest = c / b.nic;
    const int64_t pc = pidx(b, i, (int)(rest % b.njc), (int)(rest / b.njc));
    const int64_t np = b.npad;
    const double* r = rec + c * nv;
    const int nsp = b.nsp;
    const int iu = rec_u(nsp);
    fs[pc + FS_RHO * np] = r[4];
    fs[pc + FS_VX * np] = r[5];
    fs[pc + FS_VY * np] = r[6];
    fs[pc + FS_VZ * np] = r[7];
    fs[pc + FS_P * np] = r[8];
    fs[pc + FS_A * np] = r[9];
    fs[pc + FS_U * np] = r[iu];
    fs[pc + FS_T * np] = r[iu + 1];
    if (nsp > 1)
        for (int s = 0; s < nsp; ++s) {
            fs[pc + (FS_BASE_PLANES + s) * np] = r[15 + s];
            fs[pc + (FS_BASE_PLANES + nsp + s) * np] = r[4] * r[15 + s];
        }
    if (mu) mu[c] = r[10];
    if (kc) kc[c] = r[11];
    if (b.shock) b.shock[pc] = (r[14] > 0.0) ? 1 : 0;
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_pack_flow(const BlockDev& b, const double* pos, const double* mu, const double* kc, const double* dtc,
                            double* rec, int nv, hipStream_t s) {
    const int64_t n = (int64_t)b.nic * b.njc * b.nkc * nv;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_flow, dim3(nblk(n, 256)), dim3(256), 0, s, b, pos, mu, kc, dtc, rec, nv);
    return hipGetLastError();
}

hipError_t launch_unpack_flow(const BlockDev& b, const double* rec, double* fs, double* mu, double* kc, int nv,
                              hipStream_t s) {
    const int64_t n = (int64_t)b.nic * b.njc * b.nkc;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack_flow, dim3(nblk(n, 256)), dim3(256), 0, s, b, rec, fs, mu, kc, nv);
    return hipGetLastError();
}

}  // namespace gdtk
