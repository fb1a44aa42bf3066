// flux_face.hpp -- the ideal-gas face-flux kernel of the face-per-thread path, shared by
// the translation units that instantiate it (flux_face_<dims>_<numerics>.hip, one per
// dimension and numerics so that they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "gdtk_device.hpp"
#include "kernel_common.hpp"

namespace gdtk {

// One thread per face of index direction d: reconstruction + flux calculator, or the
// simple-outflow boundary flux effect on faces whose BC computes the flux itself.
template <int DIM, int FC, bool LMR>
__device__ __forceinline__ void face_flux_one(const BlockDev& b, const IdealGasDev& gas, int d, double eps,
                                              int64_t f) {
    using C = CQ<DIM>;
    int fi = (int)(f % b.nf[d][0]);
    int64_t rest = f / b.nf[d][0];
    int fj = (int)(rest % b.nf[d][1]);
    int fk = (int)(rest / b.nf[d][1]);
    const double* fs = b.fs_in;
    const int64_t np = b.npad;
    int pos = (d == 0) ? fi : (d == 1) ? fj : fk;
    int nd = (d == 0) ? b.nic : (d == 1) ? b.njc : b.nkc;
    auto at = [&](int s) -> int64_t {
        if (d == 0) return pidx(b, s, fj, fk);
        if (d == 1) return pidx(b, fi, s, fk);
        return pidx(b, fi, fj, s);
    };
    FaceGeom g = load_face(b.fgeom[d], b.nface[d], f);
    double F[C::n];
    int side = -1;
    if (pos == 0) side = 2 * d;
    else if (pos == nd) side = 2 * d + 1;
    double* out = b.flux[d];
    const int64_t nfc = b.nface[d];
    if (side >= 0 && b.bc_kind[side] == BC_OUTFLOW_SIMPLE_FLUX) {
        int outsign = (side % 2 == 0) ? -1 : 1;
        int64_t c = at(side % 2 == 0 ? 0 : nd - 1);
        Vec3d v{fs[c + FS_VX * np], fs[c + FS_VY * np], fs[c + FS_VZ * np]};
        outflow_flux<DIM>(g, outsign, fs[c + FS_RHO * np], fs[c + FS_U * np], fs[c + FS_P * np], v, F);
    } else {
        int64_t cL1 = at(pos - 2), cL0 = at(pos - 1), cR0 = at(pos), cR1 = at(pos + 1);
        const double* len = b.len + d * np;
        StencilCell L1 = load_sc(fs, np, cL1), L0 = load_sc(fs, np, cL0), R0 = load_sc(fs, np, cR0),
                    R1 = load_sc(fs, np, cR1);
        FallbackFS fb{fs, np, cL0, cR0};
        // Adaptive calculator: EFM where the face is shocked (strict rule), else AUSMDV.
        const bool shocked = flux_is_adaptive(FC) && ((b.shock[cL0] | b.shock[cR0]) != 0);
        face_flux_l2r2<DIM, FC, LMR>(gas, g, L1, L0, R0, R1, fs[cL0 + FS_A * np], fs[cR0 + FS_A * np], fb,
                                     len[cL1], len[cL0], len[cR0], len[cR1], eps, F, shocked);
    }
    for (int j = 0; j < C::n; ++j) out[f + j * nfc] = F[j];
}

// Faces of every direction of every block in one launch.  Each block's faces of a direction
// start at a multiple of the workgroup size (FaceBatch), so every workgroup serves one
// direction of one block and reads that block's record with scalar loads.  LMR selects
// Eilmer 5's limiter smoothing and AUSMDV sound speed.
template <int DIM, int FC, bool LMR>
__global__ __launch_bounds__(256, 4) void k_face_flux_batch(const BlockDev* __restrict__ tab, FaceBatch fb,
                                                         int nblocks, double eps, IdealGasDev gas) {
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
    face_flux_one<DIM, FC, LMR>(b, gas, d, eps, f);
}

// Launchers of k_face_flux_batch<DIM, FC, LMR> for each dimension and numerics (defined in
// flux_face_<dims>_<numerics>.hip); flux_calc selects FC.
hipError_t launch_face_flux_3d_e4(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc, double eps,
                                  const IdealGasDev& gas, int64_t total, hipStream_t s);
hipError_t launch_face_flux_2d_e4(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc, double eps,
                                  const IdealGasDev& gas, int64_t total, hipStream_t s);
hipError_t launch_face_flux_3d_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc, double eps,
                                   const IdealGasDev& gas, int64_t total, hipStream_t s);
hipError_t launch_face_flux_2d_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc, double eps,
                                   const IdealGasDev& gas, int64_t total, hipStream_t s);

// Instantiates and launches k_face_flux_batch<DIM, FC, LMR> for the flux calculator code.
template <int DIM, bool LMR>
hipError_t launch_face_flux_fc(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc, double eps,
                               const IdealGasDev& gas, int64_t total, hipStream_t s) {
    dim3 grid((unsigned)((total + 255) / 256)), blk(256);
#define GDTK_FACE(F) hipLaunchKernelGGL((k_face_flux_batch<DIM, F, LMR>), grid, blk, 0, s, tab, fb, nblocks, eps, gas)
    if (flux_calc == 4) GDTK_FACE(4);
    else if (flux_calc == 3) GDTK_FACE(3);
    else if (flux_calc == 2) GDTK_FACE(2);
    else if (flux_calc == 1) GDTK_FACE(1);
    else GDTK_FACE(0);
#undef GDTK_FACE
    return hipGetLastError();
}

}  // namespace gdtk
