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
// sound speed is not updated, as in the reference's face reconstruction.  The species
// evaluations run one after another (SERIAL): 162 instead of 202 VGPRs, 3 waves per SIMD
// instead of 2, and face launches 26 % shorter on the Mach-10 cylinder.  (Holding the face
// frame and velocity outside registers during the iteration reaches 4 waves per SIMD with
// a small spill, 1 % faster: not kept.)
template <int NSP>
__device__ __forceinline__ bool tpg_update_rhou(const ChemDev& cd, FaceStateT<NSP>& s) {
    if (!tpg_update_temperature<NSP, true>(cd, s.massf, s.T, s.u)) return false;
    s.p = s.rho * tpg_Rmix<NSP>(cd, s.massf) * s.T;
    return true;
}

// A double from the other lane of an adjacent lane pair (lanes 2k and 2k + 1).
__device__ __forceinline__ double pair_swap(double v) { return __shfl_xor(v, 1); }

// Flux of one face of direction d, two lanes per face: onedinterp.d:751-988 with species
// reconstruction (allow_reconstruction_for_species and scale_species_after_reconstruction
// true, thermo_interpolator rhou, local frame) and fluxcalc.d:54-184; or the simple-outflow
// flux of boundary_flux_effect.d:557-645.  The even lane of a pair builds the left state
// and the odd lane the right one (one side of the l2r2 reconstruction, mass-fraction
// scaling, the Newton T(u) of update_thermo_from_rhou with its fallback to the cell state),
// the pair swaps the states, and both lanes evaluate the flux, each storing every other
// component.  Each side's arithmetic is the reference's for that side, so a face's two
// Newton iterations run side by side and a lane holds one side's state.  Both lanes of a
// pair take the same branches up to the swap.  Returns false where the reference throws
// (either side's mass fractions far from summing to one).
//
// LMR: Eilmer 5's reconstruction (lmr/onedinterp.d:137-169, 268-286, 314-315): the mass
// fractions themselves are reconstructed and scaled, then the density, each with the
// dimensionalised limiter epsilon; AUSMDV takes the smooth-maximum sound speed.
template <int DIM, int FC, int NSP, bool LMR = false>
__device__ __forceinline__ bool face_flux_species_pair(const BlockDev& b, const ChemDev& cd, int d, double eps,
                                                       int64_t f, int right) {
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
    double* out = b.flux[d];
    const int64_t nfc = b.nface[d];
    if (side >= 0 && b.bc_kind[side] == BC_OUTFLOW_SIMPLE_FLUX) {
        if (right) return true;
        const int outsign = (side % 2 == 0) ? -1 : 1;
        const int64_t c = at(side % 2 == 0 ? 0 : nd - 1);
        const Vec3d v{fs[c + FS_VX * np], fs[c + FS_VY * np], fs[c + FS_VZ * np]};
        double mf[NSP];
#pragma unroll
        for (int i = 0; i < NSP; ++i) mf[i] = fs[c + (FS_BASE_PLANES + i) * np];
        outflow_flux<DIM, NSP>(g, outsign, fs[c + FS_RHO * np], fs[c + FS_U * np], fs[c + FS_P * np], v, F, mf);
#pragma unroll
        for (int j = 0; j < C::n; ++j) out[f + j * nfc] = F[j];
        return true;
    }
    // This lane's three stencil cells: L1, L0, R0 (left) or L0, R0, R1 (right); c1 is the
    // lane's own cell, the one the face state starts as a copy of (onedinterp.d:127-128).
    const int64_t cL0 = at(pos - 1), cR0 = at(pos);
    const int64_t c0 = right ? cL0 : at(pos - 2);
    const int64_t c1 = right ? cR0 : cL0;
    const int64_t c2 = right ? at(pos + 1) : cR0;
    const double* len = b.len + d * np;
    const SpeciesCell<NSP> S0 = load_species_cell<NSP>(fs, np, c0), S1 = load_species_cell<NSP>(fs, np, c1),
                           S2 = load_species_cell<NSP>(fs, np, c2);
    const Vec3d v0 = to_local(S0.vel, g), v1 = to_local(S1.vel, g), v2 = to_local(S2.vel, g);
    const double l0 = len[c0], l1 = len[c1], l2 = len[c2];
    FaceStateT<NSP> st;
    st.T = fs[c1 + FS_T * np];
    st.a = fs[c1 + FS_A * np];
    bool ok;
    auto build = [&](const auto& c) {
        st.vel.x = c.template value<LMR>(v0.x, v1.x, v2.x, eps);
        st.vel.y = c.template value<LMR>(v0.y, v1.y, v2.y, eps);
        st.vel.z = c.template value<LMR>(v0.z, v1.z, v2.z, eps);
        if (LMR) {
#pragma unroll
            for (int i = 0; i < NSP; ++i) {
                const int64_t pl = (FS_BASE_PLANES + i) * np;
                st.massf[i] = c.template value<true>(fs[c0 + pl], fs[c1 + pl], fs[c2 + pl], eps);
            }
            ok = scale_mass_fractions<NSP>(st.massf);
            st.rho = c.template value<true>(fs[c0 + FS_RHO * np], fs[c1 + FS_RHO * np], fs[c2 + FS_RHO * np], eps);
        } else {
            double rs[NSP];
#pragma unroll
            for (int i = 0; i < NSP; ++i) rs[i] = c.value(S0.rho_s[i], S1.rho_s[i], S2.rho_s[i], eps);
            double rho = 0.0;
#pragma unroll
            for (int i = 0; i < NSP; ++i) rho += rs[i];
            st.rho = rho;
#pragma unroll
            for (int i = 0; i < NSP; ++i) st.massf[i] = rs[i] / st.rho;
            ok = scale_mass_fractions<NSP>(st.massf);
        }
        st.u = c.template value<LMR>(S0.u, S1.u, S2.u, eps);
    };
    if (right) {
        L2R2Side<true> c;
        c.prepare(0.0, l0, l1, l2);
        build(c);
    } else {
        L2R2Side<false> c;
        c.prepare(l0, l1, l2, 0.0);
        build(c);
    }
    if (!tpg_update_rhou<NSP>(cd, st)) copy_cell_state<NSP>(fs, np, c1, v1, st);
    st.vel = to_local(to_global(st.vel, g), g);
    // EFM's specific heats, each at its own side.
    const bool shocked = flux_is_adaptive(FC) && ((b.shock[cL0] | b.shock[cR0]) != 0);
    const bool efm = FC == 1 || (FC == 2 && shocked);
    const double Cv = efm ? tpg_Cv<NSP>(cd, st.massf, st.T) : 0.0;
    // swap the two sides' states
    FaceStateT<NSP> o;
    o.rho = pair_swap(st.rho);
    o.p = pair_swap(st.p);
    o.T = pair_swap(st.T);
    o.u = pair_swap(st.u);
    o.a = pair_swap(st.a);
    o.vel.x = pair_swap(st.vel.x);
    o.vel.y = pair_swap(st.vel.y);
    o.vel.z = pair_swap(st.vel.z);
#pragma unroll
    for (int i = 0; i < NSP; ++i) o.massf[i] = pair_swap(st.massf[i]);
    const double Cv_o = pair_swap(Cv);
    const bool ok_o = pair_swap(ok ? 1.0 : 0.0) != 0.0;
    const FaceStateT<NSP>& Ls = right ? o : st;
    const FaceStateT<NSP>& Rs = right ? st : o;
#pragma unroll
    for (int j = 0; j < C::n; ++j) F[j] = 0.0;
    if (efm) efmflx<DIM, NSP>(Ls, Rs, right ? Cv_o : Cv, right ? Cv : Cv_o, F);
    else if (FC == 3 || (FC == 4 && shocked)) hanel<DIM, NSP>(Ls, Rs, F);
    else ausmdv<DIM, NSP, LMR>(Ls, Rs, F);
    const double fz = (DIM == 3) ? F[C::zMom] : 0.0;
    const Vec3d m = to_global(Vec3d{F[C::xMom], F[C::yMom], fz}, g);
    F[C::xMom] = m.x;
    F[C::yMom] = m.y;
    if (DIM == 3) F[C::zMom] = m.z;
#pragma unroll
    for (int j = 0; j < C::n; ++j)
        if ((j & 1) == right) out[f + j * nfc] = F[j];
    return right ? true : (ok && ok_o);
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
    if (tpg_update_temperature<NSP, true>(cd, massf, T, u)) {
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

// Two lanes per face (face_flux_species_pair): a workgroup of 256 lanes serves 128 faces of
// one direction of one block (the FaceBatch ranges start at multiples of 256 faces).
// LMR: Eilmer 5 reconstruction and AUSMDV (face_flux_species_pair).
// 3 waves per SIMD (168 VGPRs): the compiler fits the kernel there without spilling.
template <int DIM, int FC, int NSP, bool LMR>
__global__ __launch_bounds__(256, 3) void k_face_flux_species_pair(const BlockDev* __restrict__ tab, FaceBatch fb,
                                                                int nblocks, double eps,
                                                                const ChemDev* __restrict__ cdp) {
    int64_t w0 = (int64_t)blockIdx.x * (blockDim.x / 2);
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
    const int64_t f = w0 - start[q] + (threadIdx.x >> 1);
    if (f >= b.nface[d]) return;
    if (!face_flux_species_pair<DIM, FC, NSP, LMR>(b, *cdp, d, eps, f, (int)(threadIdx.x & 1)))
        atomicAdd(&b.counters[0], 1);
}

// Launches the face kernel for the flux calculator code (total = FaceBatch launch size).
template <int NSP, bool LMR>
inline void face_species(const BlockDev* tab, const FaceBatch& fb, int nb, int64_t total, int dims, int fc,
                         double eps, const ChemDev* cd, hipStream_t s) {
    dim3 g((unsigned)((2 * total + 255) / 256)), blk(256);
#define GDTK_FS(D, F) hipLaunchKernelGGL((k_face_flux_species_pair<D, F, NSP, LMR>), g, blk, 0, s, tab, fb, nb, eps, cd)
    if (dims == 3) {
        if (fc == 4) GDTK_FS(3, 4);
        else if (fc == 3) GDTK_FS(3, 3);
        else if (fc == 2) GDTK_FS(3, 2);
        else if (fc == 1) GDTK_FS(3, 1);
        else GDTK_FS(3, 0);
    } else {
        if (fc == 4) GDTK_FS(2, 4);
        else if (fc == 3) GDTK_FS(2, 3);
        else if (fc == 2) GDTK_FS(2, 2);
        else if (fc == 1) GDTK_FS(2, 1);
        else GDTK_FS(2, 0);
    }
#undef GDTK_FS
}

}  // namespace gdtk
