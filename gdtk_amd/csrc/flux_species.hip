// flux_species.hip -- kernels of the explicit update for a multi-species thermally perfect
// gas: the face-per-thread flux and cell-update kernels over all blocks, and the
// encode/decode kernels of one block (start-up).  Instantiated for 2 species (N2, N)
// and 5 species (N2, O2, NO, N, O).
#include <hip/hip_runtime.h>

#include "flux_species.hpp"
#include "launchers.hpp"

namespace gdtk {

template <int DIM, int STAGE, int NSP>
__global__ __launch_bounds__(256) void k_cell_update_species(const BlockDev* __restrict__ tab,
                                                             const int64_t* __restrict__ start, int nblocks,
                                                             const double* __restrict__ dtp, CellLimits lim,
                                                             const ChemDev* __restrict__ cdp) {
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
    cell_update_species<DIM, STAGE, NSP>(b, *cdp, *dtp, lim.low_T, lim.max_vel, lim.min_T, lim.max_T, t);
}

// encode_conserved (fvcell.d:511-583) of every active cell of one block.
template <int DIM, int NSP>
__global__ void k_encode_species(BlockDev b, double* U) {
    using C = CQ<DIM, NSP>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)b.nic * b.njc * b.nkc) return;
    const int i = (int)(t % b.nic);
    const int64_t rest = t / b.nic;
    const int64_t c = pidx(b, i, (int)(rest % b.njc), (int)(rest / b.njc));
    const int64_t np = b.npad;
    const double* fs = b.fs_in;
    const double rho = fs[c + FS_RHO * np];
    const double vx = fs[c + FS_VX * np], vy = fs[c + FS_VY * np], vz = fs[c + FS_VZ * np];
    U[c + C::mass * np] = rho;
    U[c + C::xMom * np] = rho * vx;
    U[c + C::yMom * np] = rho * vy;
    if (DIM == 3) U[c + C::zMom * np] = rho * vz;
    const double u = fs[c + FS_U * np];
    const double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
    U[c + C::totEnergy * np] = rho * (u + ke);
#pragma unroll
    for (int s = 0; s < NSP; ++s) U[c + (C::species + s) * np] = rho * fs[c + (FS_BASE_PLANES + s) * np];
}

// decode_conserved of every active cell of one block into fs (start-up); the Newton
// iteration starts at the temperature held in fs.
template <int DIM, int NSP>
__global__ void k_decode_species(BlockDev b, const ChemDev* __restrict__ cdp, double* U, double* fs, double low_T) {
    using C = CQ<DIM, NSP>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)b.nic * b.njc * b.nkc) return;
    const int i = (int)(t % b.nic);
    const int64_t rest = t / b.nic;
    const int64_t c = pidx(b, i, (int)(rest % b.njc), (int)(rest / b.njc));
    const int64_t np = b.npad;
    double Uc[C::n];
#pragma unroll
    for (int q = 0; q < C::n; ++q) Uc[q] = U[c + q * np];
    double rho, u, p, a, massf[NSP], rho_s[NSP];
    double T = fs[c + FS_T * np];
    Vec3d vel;
    bool reenc;
    if (decode_species<DIM, NSP>(*cdp, Uc, low_T, rho, vel, u, p, T, a, massf, rho_s, reenc)) {
        atomicAdd(&b.counters[0], 1);
        return;
    }
#pragma unroll
    for (int q = 0; q < C::n; ++q) U[c + q * np] = Uc[q];
    store_species_state<NSP>(fs, np, c, rho, vel, u, p, T, a, massf, rho_s);
}

static inline unsigned nblk256(int64_t n) { return (unsigned)((n + 255) / 256); }

bool species_supported(int nsp) { return nsp == 2 || nsp == 5; }

template <int NSP>
static void cell_species(const BlockDev* tab, const int64_t* start, int nb, int64_t total, int dims, int stage,
                         const double* dtp, const CellLimits& lim, const ChemDev* cd, hipStream_t s) {
    dim3 g(nblk256(total)), blk(256);
    if (dims == 3) {
        if (stage == 1) hipLaunchKernelGGL((k_cell_update_species<3, 1, NSP>), g, blk, 0, s, tab, start, nb, dtp, lim, cd);
        else hipLaunchKernelGGL((k_cell_update_species<3, 2, NSP>), g, blk, 0, s, tab, start, nb, dtp, lim, cd);
    } else {
        if (stage == 1) hipLaunchKernelGGL((k_cell_update_species<2, 1, NSP>), g, blk, 0, s, tab, start, nb, dtp, lim, cd);
        else hipLaunchKernelGGL((k_cell_update_species<2, 2, NSP>), g, blk, 0, s, tab, start, nb, dtp, lim, cd);
    }
}

hipError_t launch_face_flux_species(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                   int numerics, double eps, int nsp, const ChemDev* cd, hipStream_t s) {
    const int64_t total = fb.total[0] + fb.total[1] + fb.total[2];
    if (total <= 0) return hipSuccess;
    if (numerics == 1) return launch_face_flux_species_lmr(tab, fb, nblocks, dims, flux_calc, eps, nsp, cd, total, s);
    if (nsp == 2) face_species<2, false>(tab, fb, nblocks, total, dims, flux_calc, eps, cd, s);
    else if (nsp == 5) face_species<5, false>(tab, fb, nblocks, total, dims, flux_calc, eps, cd, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_cell_update_species(const BlockDev* tab, const int64_t* start, int nblocks, int64_t total, int dims,
                                     int stage, const double* dtp, const CellLimits& lim, int nsp, const ChemDev* cd,
                                     hipStream_t s) {
    if (total <= 0) return hipSuccess;
    if (nsp == 2) cell_species<2>(tab, start, nblocks, total, dims, stage, dtp, lim, cd, s);
    else if (nsp == 5) cell_species<5>(tab, start, nblocks, total, dims, stage, dtp, lim, cd, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_encode_species(const BlockDev& b, double* U, hipStream_t s) {
    const int64_t n = (int64_t)b.nic * b.njc * b.nkc;
    dim3 g(nblk256(n)), blk(256);
#define GDTK_ENC(D, N) hipLaunchKernelGGL((k_encode_species<D, N>), g, blk, 0, s, b, U)
    if (b.nsp == 2) { if (b.dims == 3) GDTK_ENC(3, 2); else GDTK_ENC(2, 2); }
    else if (b.nsp == 5) { if (b.dims == 3) GDTK_ENC(3, 5); else GDTK_ENC(2, 5); }
    else return hipErrorInvalidValue;
#undef GDTK_ENC
    return hipGetLastError();
}

hipError_t launch_decode_species(const BlockDev& b, const ChemDev* cd, double* U, double* fs, double low_T,
                                 hipStream_t s) {
    const int64_t n = (int64_t)b.nic * b.njc * b.nkc;
    dim3 g(nblk256(n)), blk(256);
#define GDTK_DEC(D, N) hipLaunchKernelGGL((k_decode_species<D, N>), g, blk, 0, s, b, cd, U, fs, low_T)
    if (b.nsp == 2) { if (b.dims == 3) GDTK_DEC(3, 2); else GDTK_DEC(2, 2); }
    else if (b.nsp == 5) { if (b.dims == 3) GDTK_DEC(3, 5); else GDTK_DEC(2, 5); }
    else return hipErrorInvalidValue;
#undef GDTK_DEC
    return hipGetLastError();
}

}  // namespace gdtk

namespace gdtk {

// Operator-split chemistry of a simulation (simcore.d:1619-1638): the cells of every block
// gathered into dense arrays for the batched reactor, and the results written back as
// thermochemical_increment leaves them (fvcell.d:858-972): T, u, p and the mass fractions
// of the flow state, and U[0]'s species densities rho * massf.  The sound speed and the
// cell's species densities are left as they are, as in the reference.
template <bool SCATTER>
__global__ __launch_bounds__(256) void k_cells_chem(const BlockDev* __restrict__ tab, const int64_t* __restrict__ start,
                                                    const int64_t* __restrict__ coff, int nblocks, int nsp, int64_t n,
                                                    double* rho, double* T, double* u, double* p, double* massf,
                                                    const double* T_o, const double* u_o, const double* p_o,
                                                    const double* massf_o, const double* dtc_o, double* dtc,
                                                    const int* status, unsigned* nfail) {
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
    const int64_t c = pidx(b, i, (int)(rest % b.njc), (int)(rest / b.njc));
    const int64_t np = b.npad;
    const int64_t x = coff[q] + t;
    double* fs = const_cast<double*>(b.fs_in);
    if (!SCATTER) {
        rho[x] = fs[c + FS_RHO * np];
        T[x] = fs[c + FS_T * np];
        u[x] = fs[c + FS_U * np];
        p[x] = fs[c + FS_P * np];
        for (int s = 0; s < nsp; ++s) massf[(int64_t)s * n + x] = fs[c + (FS_BASE_PLANES + s) * np];
        return;
    }
    if (status[x] != 0) atomicAdd(nfail, 1u);
    fs[c + FS_T * np] = T_o[x];
    fs[c + FS_U * np] = u_o[x];
    fs[c + FS_P * np] = p_o[x];
    const double r = fs[c + FS_RHO * np];
    double* U = const_cast<double*>(b.U0);
    const int sp0 = (b.dims == 3) ? 5 : 4;
    for (int s = 0; s < nsp; ++s) {
        const double y = massf_o[(int64_t)s * n + x];
        fs[c + (FS_BASE_PLANES + s) * np] = y;
        U[c + (sp0 + s) * np] = r * y;
    }
    dtc[x] = dtc_o[x];
}

hipError_t launch_cells_gather(const BlockDev* tab, const int64_t* start, const int64_t* coff, int nblocks,
                               int64_t total, int nsp, int64_t n, double* rho, double* T, double* u, double* p,
                               double* massf, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cells_chem<false>, dim3(nblk256(total)), dim3(256), 0, s, tab, start, coff, nblocks, nsp, n,
                       rho, T, u, p, massf, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_cells_scatter(const BlockDev* tab, const int64_t* start, const int64_t* coff, int nblocks,
                                int64_t total, int nsp, int64_t n, const double* T_o, const double* u_o,
                                const double* p_o, const double* massf_o, const double* dtc_o, double* dtc,
                                const int* status, unsigned* nfail, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cells_chem<true>, dim3(nblk256(total)), dim3(256), 0, s, tab, start, coff, nblocks, nsp, n,
                       nullptr, nullptr, nullptr, nullptr, nullptr, T_o, u_o, p_o, massf_o, dtc_o, dtc, status, nfail);
    return hipGetLastError();
}

// Fills n doubles with v.
__global__ void k_fill_f64(double* p, int64_t n, double v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk256(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace gdtk
