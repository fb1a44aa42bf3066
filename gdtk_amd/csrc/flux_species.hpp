// flux_species.hpp -- device code of the explicit update for a multi-species thermally
// perfect gas (NSP species, 2..8): face fluxes with species reconstruction, the stage
// update with species densities, decode/encode through the mixture EOS.
//
// Flow-state planes of a block: rho, vx, vy, vz, u, p, T, a (FS_*), then massf[NSP] at
// FS_BASE_PLANES + s and rho_s[NSP] at FS_BASE_PLANES + NSP + s.  Conserved quantities:
// mass, momentum, total energy, then the species densities (CQ<DIM, NSP>).  The gas is the
// gas part of a ChemDev table (species gas constants and CEA curves).  Every expression
// keeps the reference's operation order (cited per function); FP64, no FMA contraction.
#pragma once
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "chem_device.hpp"
#include "gdtk_device.hpp"
#include "kernel_common.hpp"

namespace gdtk {

template <int NSP>
struct SpeciesCell {
    double u;
    Vec3d vel;
    double rho_s[NSP];
};

template <int NSP>
__device__ __forceinline__ SpeciesCell<NSP> load_species_cell(const double* fs, int64_t np, int64_t c) {
    SpeciesCell<NSP> s;
    s.u = fs[c + FS_U * np];
    s.vel.x = fs[c + FS_VX * np];
    s.vel.y = fs[c + FS_VY * np];
    s.vel.z = fs[c + FS_VZ * np];
#pragma unroll
    for (int i = 0; i < NSP; ++i) s.rho_s[i] = fs[c + (FS_BASE_PLANES + NSP + i) * np];
    return s;
}

// The whole flow state of a cell as a face state (the reference's Lft.copy_values_from of
// the cell after a failed EOS call, onedinterp.d:45-55); velocity given separately.
template <int NSP>
__device__ __forceinline__ void copy_cell_state(const double* fs, int64_t np, int64_t c, Vec3d vel_local,
                                                FaceStateT<NSP>& s) {
    s.rho = fs[c + FS_RHO * np];
    s.u = fs[c + FS_U * np];
    s.p = fs[c + FS_P * np];
    s.T = fs[c + FS_T * np];
    s.a = fs[c + FS_A * np];
    s.vel = vel_local;
#pragma unroll
    for (int i = 0; i < NSP; ++i) s.massf[i] = fs[c + (FS_BASE_PLANES + i) * np];
}

// gas_model.d:373-407 with tolerance 0 and assert_error_tolerance 0.1 (the call of
// onedinterp.d:885-890): negative fractions clipped, then scaled to sum 1.  Returns false
// where the reference throws.
template <int NSP>
__device__ __forceinline__ bool scale_mass_fractions(double* massf) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        massf[i] = massf[i] >= 0.0 ? massf[i] : 0.0;
        s += massf[i];
    }
    if (fabs(s - 1.0) > 0.1) return false;
    if (fabs(s - 1.0) > 0.0) {
#pragma unroll
        for (int i = 0; i < NSP; ++i) massf[i] /= s;
    }
    return true;
}

// update_thermo_from_rhou of the thermally perfect mixture (therm_perf_gas.d:235-239):
// Newton T(u) from the state's T, then p = rho R_mix T (perf_gas_mix_eos.d:43-49).  The
// sound speed is not updated, as in the reference's face reconstruction.
template <int NSP>
__device__ __forceinline__ bool tpg_update_rhou(const ChemDev& cd, FaceStateT<NSP>& s) {
    if (!tpg_update_temperature<NSP>(cd, s.massf, s.T, s.u)) return false;
    s.p = s.rho * tpg_Rmix<NSP>(cd, s.massf) * s.T;
    return true;
}

// Flux of one face of direction d (face-per-thread): onedinterp.d:751-988 with species
// reconstruction (allow_reconstruction_for_species and scale_species_after_reconstruction
// true, thermo_interpolator rhou, local frame) and fluxcalc.d:54-184; or the simple-outflow
// flux of boundary_flux_effect.d:557-645.  Returns false where the reference throws
// (mass fractions of a reconstructed state far from summing to one).
template <int DIM, int FC, int NSP>
__device__ __forceinline__ bool face_flux_species(const BlockDev& b, const ChemDev& cd, int d, double eps,
                                                  int64_t f) {
    using C = CQ<DIM, NSP>;
    const int fi = (int)(f % b.nf[d][0]);
    const int64_t rest = f / b.nf[d][0];
    const int fj = (int)(rest % b.nf[d][1]);
    const int fk = (int)(rest / b.nf[d][1]);
    const double* fs = b.fs_in;
    const int64_t np = b.npad;
    const int pos = (d == 0) ? fi : (d == 1) ? fj : fk;
    const int nd = (d == 0) ? b.nic : (d == 1) ? b.njc : b.nkc;
    auto at = [&](int s) -> int64_t {
        if (d == 0) return pidx(b, s, fj, fk);
        if (d == 1) return pidx(b, fi, s, fk);
        return pidx(b, fi, fj, s);
    };
    const FaceGeom g = load_face(b.fgeom[d], b.nface[d], f);
    double F[C::n];
    int side = -1;
    if (pos == 0) side = 2 * d;
    else if (pos == nd) side = 2 * d + 1;
    bool ok = true;
    if (side >= 0 && b.bc_kind[side] == BC_OUTFLOW_SIMPLE_FLUX) {
        const int outsign = (side % 2 == 0) ? -1 : 1;
        const int64_t c = at(side % 2 == 0 ? 0 : nd - 1);
        const Vec3d v{fs[c + FS_VX * np], fs[c + FS_VY * np], fs[c + FS_VZ * np]};
        double mf[NSP];
#pragma unroll
        for (int i = 0; i < NSP; ++i) mf[i] = fs[c + (FS_BASE_PLANES + i) * np];
        outflow_flux<DIM, NSP>(g, outsign, fs[c + FS_RHO * np], fs[c + FS_U * np], fs[c + FS_P * np], v, F, mf);
    } else {
        const int64_t cL1 = at(pos - 2), cL0 = at(pos - 1), cR0 = at(pos), cR1 = at(pos + 1);
        const double* len = b.len + d * np;
        const SpeciesCell<NSP> L1 = load_species_cell<NSP>(fs, np, cL1), L0 = load_species_cell<NSP>(fs, np, cL0),
                               R0 = load_species_cell<NSP>(fs, np, cR0), R1 = load_species_cell<NSP>(fs, np, cR1);
        const Vec3d vL1 = to_local(L1.vel, g), vL0 = to_local(L0.vel, g), vR0 = to_local(R0.vel, g),
                    vR1 = to_local(R1.vel, g);
        L2R2 c;
        c.prepare(len[cL1], len[cL0], len[cR0], len[cR1]);
        FaceStateT<NSP> Ls, Rs;
        // Lft/Rght start as copies of the cells L0/R0 (onedinterp.d:127-128): T is the
        // Newton start of the EOS call and a is never updated.
        Ls.T = fs[cL0 + FS_T * np];
        Rs.T = fs[cR0 + FS_T * np];
        Ls.a = fs[cL0 + FS_A * np];
        Rs.a = fs[cR0 + FS_A * np];
        c.scalar(vL1.x, vL0.x, vR0.x, vR1.x, eps, Ls.vel.x, Rs.vel.x);
        c.scalar(vL1.y, vL0.y, vR0.y, vR1.y, eps, Ls.vel.y, Rs.vel.y);
        c.scalar(vL1.z, vL0.z, vR0.z, vR1.z, eps, Ls.vel.z, Rs.vel.z);
        double rsL[NSP], rsR[NSP];
#pragma unroll
        for (int i = 0; i < NSP; ++i) c.scalar(L1.rho_s[i], L0.rho_s[i], R0.rho_s[i], R1.rho_s[i], eps, rsL[i], rsR[i]);
        double rho_L = 0.0, rho_R = 0.0;
#pragma unroll
        for (int i = 0; i < NSP; ++i) {
            rho_L += rsL[i];
            rho_R += rsR[i];
        }
        Ls.rho = rho_L;
        Rs.rho = rho_R;
#pragma unroll
        for (int i = 0; i < NSP; ++i) {
            Ls.massf[i] = rsL[i] / Ls.rho;
            Rs.massf[i] = rsR[i] / Rs.rho;
        }
        ok = scale_mass_fractions<NSP>(Ls.massf) && scale_mass_fractions<NSP>(Rs.massf);
        c.scalar(L1.u, L0.u, R0.u, R1.u, eps, Ls.u, Rs.u);
        // The EOS calls of the two sides share one copy of the Newton code (a loop that is
        // not unrolled, with the side's state selected in and out).
#pragma clang loop unroll(disable)
        for (int side = 0; side < 2; ++side) {
            FaceStateT<NSP> st = side ? Rs : Ls;
            if (!tpg_update_rhou<NSP>(cd, st))
                copy_cell_state<NSP>(fs, np, side ? cR0 : cL0, side ? vR0 : vL0, st);
            if (side) Rs = st;
            else Ls = st;
        }
        Ls.vel = to_local(to_global(Ls.vel, g), g);
        Rs.vel = to_local(to_global(Rs.vel, g), g);
#pragma unroll
        for (int j = 0; j < C::n; ++j) F[j] = 0.0;
        // FC 2: adaptive_efm_ausmdv (fluxcalc.d:1315-1332), EFM where the face is shocked.
        const bool shocked = flux_is_adaptive(FC) && ((b.shock[cL0] | b.shock[cR0]) != 0);
        if (FC == 1 || (FC == 2 && shocked))
            efmflx<DIM, NSP>(Ls, Rs, tpg_Cv<NSP>(cd, Ls.massf, Ls.T), tpg_Cv<NSP>(cd, Rs.massf, Rs.T), F);
        else if (FC == 3 || (FC == 4 && shocked))
            hanel<DIM, NSP>(Ls, Rs, F);
        else
            ausmdv<DIM, NSP>(Ls, Rs, F);
        // The grid-velocity terms of fluxcalc.d:159-171 add signed zeros (see face_flux_l2r2).
        const double fz = (DIM == 3) ? F[C::zMom] : 0.0;
        const Vec3d m = to_global(Vec3d{F[C::xMom], F[C::yMom], fz}, g);
        F[C::xMom] = m.x;
        F[C::yMom] = m.y;
        if (DIM == 3) F[C::zMom] = m.z;
    }
    double* out = b.flux[d];
    const int64_t nfc = b.nface[d];
#pragma unroll
    for (int j = 0; j < C::n; ++j) out[f + j * nfc] = F[j];
    return ok;
}

// decode_conserved for the multi-species gas (fvcell.d:586-821, cqi.mass present):
// species densities clipped at zero and scaled to the total density, mass fractions,
// the EOS from rho, u with Newton started at T, and on EOS failure the low-T reset with
// re-encoding (ignore_low_T_thermo_update_failure).  U is updated in place as in the
// reference.  Returns 1 where the reference throws.
template <int DIM, int NSP>
__device__ __forceinline__ int decode_species(const ChemDev& cd, double* U, double low_T, double& rho, Vec3d& vel,
                                              double& u, double& p, double& T, double& a, double* massf,
                                              double* rho_s, bool& reencoded) {
    using C = CQ<DIM, NSP>;
    reencoded = false;
    rho = U[C::mass];
    if (!(rho > 0.0)) return 1;
    const double dinv = 1.0 / rho;
    const double zMom = (DIM == 3) ? U[C::zMom] : 0.0;
    vel.x = U[C::xMom] * dinv;
    vel.y = U[C::yMom] * dinv;
    vel.z = zMom * dinv;
    const double rE = U[C::totEnergy];
    double uu = rE * dinv;
    const double ke = 0.5 * (vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
    uu -= ke;
    u = uu;
    // fvcell.d:716-738
    double rhos_sum = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        if (U[C::species + i] < 0.0) U[C::species + i] = 0.0;
        rhos_sum += U[C::species + i];
    }
    if (fabs(rhos_sum - rho) > 0.1) return 1;
    if (fabs(rhos_sum - rho) > 0.0) {
        const double scale_factor = rho / rhos_sum;
#pragma unroll
        for (int i = 0; i < NSP; ++i) U[C::species + i] *= scale_factor;
    }
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        massf[i] = U[C::species + i] * dinv;
        rho_s[i] = U[C::species + i];
    }
    if (tpg_update_temperature<NSP>(cd, massf, T, u)) {
        p = rho * tpg_Rmix<NSP>(cd, massf) * T;
    } else {
        // update_thermo_from_rhoT at the suggested low temperature, then encode_conserved
        T = low_T;
        u = tpg_energy<NSP>(cd, massf, T);
        p = rho * tpg_Rmix<NSP>(cd, massf) * T;
        U[C::mass] = rho;
        U[C::xMom] = rho * vel.x;
        U[C::yMom] = rho * vel.y;
        if (DIM == 3) U[C::zMom] = rho * vel.z;
        const double ke2 = 0.5 * (vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
        U[C::totEnergy] = rho * (u + ke2);
#pragma unroll
        for (int i = 0; i < NSP; ++i) U[C::species + i] = rho * massf[i];
        reencoded = true;
    }
    if (T <= 0.0) return 1;
    a = tpg_sound_speed<NSP>(cd, massf, T);
    return 0;
}

// Writes a decoded cell's flow state to the planes of fs.
template <int NSP>
__device__ __forceinline__ void store_species_state(double* fs, int64_t np, int64_t c, double rho, Vec3d vel, double u,
                                                    double p, double T, double a, const double* massf,
                                                    const double* rho_s) {
    fs[c + FS_RHO * np] = rho;
    fs[c + FS_VX * np] = vel.x;
    fs[c + FS_VY * np] = vel.y;
    fs[c + FS_VZ * np] = vel.z;
    fs[c + FS_U * np] = u;
    fs[c + FS_P * np] = p;
    fs[c + FS_T * np] = T;
    fs[c + FS_A * np] = a;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        fs[c + (FS_BASE_PLANES + i) * np] = massf[i];
        fs[c + (FS_BASE_PLANES + NSP + i) * np] = rho_s[i];
    }
}

// time_derivatives (fvcell.d:824-854), the pc stage update
// (simcore_gasdynamic_step.d:1225-1321), decode and check_data (flowstate.d:363-390 with
// the mass-fraction sum test of gas_state.d:226-268) of one cell.
template <int DIM, int STAGE, int NSP>
__device__ __forceinline__ void cell_update_species(const BlockDev& b, const ChemDev& cd, double dt, double low_T,
                                                    double max_vel, double min_T, double max_T, int64_t t) {
    using C = CQ<DIM, NSP>;
    const int i = (int)(t % b.nic);
    const int64_t rest = t / b.nic;
    const int j = (int)(rest % b.njc);
    const int k = (int)(rest / b.njc);
    const int64_t c = pidx(b, i, j, k);
    const int64_t np = b.npad;
    constexpr int nfaces = (DIM == 3) ? 6 : 4;
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
#pragma unroll
    for (int q = 0; q < nfaces; ++q) {
        const double A = b.fgeom[fdir[q]][fidx[q] + 9 * b.nface[fdir[q]]];
        area[q] = ((q % 2 == 0) ? -1 : 1) * A;
    }
    const double vol_inv = 1.0 / b.vol[c];
    double U[C::n];
#pragma unroll
    for (int jq = 0; jq < C::n; ++jq) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < nfaces; ++q) s -= b.flux[fdir[q]][fidx[q] + jq * b.nface[fdir[q]]] * area[q];
        const double dUdt = vol_inv * s + 0.0;
        const double u0 = b.U0[c + jq * np];
        if (STAGE == 1) {
            const double gamma_1 = 1.0;
            b.dUdt0[c + jq * np] = dUdt;
            U[jq] = u0 + dt * gamma_1 * dUdt;
        } else {
            const double gamma_1 = 0.5, gamma_2 = 0.5;
            U[jq] = u0 + dt * (gamma_1 * b.dUdt0[c + jq * np] + gamma_2 * dUdt);
        }
    }
    double rho, u, p, a, massf[NSP], rho_s[NSP];
    double T = b.fs_in[c + FS_T * np];  // the cell's temperature starts the Newton iteration
    Vec3d vel;
    bool reenc;
    if (decode_species<DIM, NSP>(cd, U, low_T, rho, vel, u, p, T, a, massf, rho_s, reenc)) {
        atomicAdd(&b.counters[0], 1);
        return;
    }
    if (STAGE == 2)
#pragma unroll
        for (int jq = 0; jq < C::n; ++jq) b.Unew[c + jq * np] = U[jq];
    store_species_state<NSP>(b.fs_out, np, c, rho, vel, u, p, T, a, massf, rho_s);
    bool ok = check_data(rho, p, T, a, vel, max_vel, min_T, max_T);
    double f_sum = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; ++q) f_sum += massf[q];
    if (f_sum < 0.99 || f_sum > 1.01 || !isfinite(f_sum)) ok = false;
    if (!ok) atomicAdd(&b.counters[1], 1);
}

}  // namespace gdtk
