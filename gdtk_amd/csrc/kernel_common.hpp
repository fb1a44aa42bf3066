// kernel_common.hpp -- device helpers shared by the face-per-thread and fused kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "gdtk_device.hpp"

namespace gdtk {

__device__ __forceinline__ FaceGeom load_face(const double* g, int64_t nface, int64_t f) {
    FaceGeom r;
    r.n.x = g[f];
    r.n.y = g[f + nface];
    r.n.z = g[f + 2 * nface];
    r.t1.x = g[f + 3 * nface];
    r.t1.y = g[f + 4 * nface];
    r.t1.z = g[f + 5 * nface];
    r.t2.x = g[f + 6 * nface];
    r.t2.y = g[f + 7 * nface];
    r.t2.z = g[f + 8 * nface];
    r.area = g[f + 9 * nface];
    return r;
}

__device__ __forceinline__ int64_t face_index(const BlockDev& b, int d, int fi, int fj, int fk) {
    return (int64_t)fi + (int64_t)b.nf[d][0] * ((int64_t)fj + (int64_t)b.nf[d][1] * (int64_t)fk);
}

__device__ __forceinline__ StencilCell load_sc(const double* fs, int64_t np, int64_t c) {
    StencilCell s;
    s.rho = fs[c + FS_RHO * np];
    s.u = fs[c + FS_U * np];
    s.vel.x = fs[c + FS_VX * np];
    s.vel.y = fs[c + FS_VY * np];
    s.vel.z = fs[c + FS_VZ * np];
    return s;
}

// Stage update and decode of one cell from its surface integrals si[] (fvcell.d:824-854,
// simcore_gasdynamic_step.d:1225-1321, fvcell.d:586-821, fluidblock.d:607-676).
template <int DIM, int STAGE>
__device__ __forceinline__ void update_decode_store(const BlockDev& b, const IdealGasDev& gas, int64_t c,
                                                    const double* si, double dt, double low_T, double max_vel,
                                                    double min_T, double max_T) {
    using C = CQ<DIM>;
    const int64_t np = b.npad;
    double vol_inv = 1.0 / b.vol[c];
    double U[C::n];
    for (int jq = 0; jq < C::n; ++jq) {
        double dUdt = vol_inv * si[jq] + 0.0;
        double u0 = b.U0[c + jq * np];
        if (STAGE == 1) {
            const double gamma_1 = 1.0;
            b.dUdt0[c + jq * np] = dUdt;
            U[jq] = u0 + dt * gamma_1 * dUdt;
        } else {
            const double gamma_1 = 0.5, gamma_2 = 0.5;
            U[jq] = u0 + dt * (gamma_1 * b.dUdt0[c + jq * np] + gamma_2 * dUdt);
        }
    }
    double rho, u, p, T, a;
    Vec3d vel;
    bool reenc;
    int bad = decode_ideal<DIM>(gas, U, low_T, rho, vel, u, p, T, a, reenc);
    if (bad) {
        atomicAdd(&b.counters[0], 1);
        return;
    }
    if (STAGE == 2)
        for (int jq = 0; jq < C::n; ++jq) b.Unew[c + jq * np] = U[jq];
    double* fo = b.fs_out;
    fo[c + FS_RHO * np] = rho;
    fo[c + FS_VX * np] = vel.x;
    fo[c + FS_VY * np] = vel.y;
    fo[c + FS_VZ * np] = vel.z;
    fo[c + FS_U * np] = u;
    fo[c + FS_P * np] = p;
    fo[c + FS_T * np] = T;
    fo[c + FS_A * np] = a;
    if (!check_data(rho, p, T, a, vel, max_vel, min_T, max_T)) atomicAdd(&b.counters[1], 1);
}

}  // namespace gdtk
