// fused_stage.hip -- one kernel per predictor-corrector stage: reconstruction, fluxes,
// time derivatives, stage update, decode and validity check of every cell of a block.
//
// Mapping (gfx950, 64-wide waves): a workgroup of 64 x 4 threads owns a tile of 64 (i) x
// 4 (j) cells and marches up a chunk of KC planes in k (2.5-D blocking).  For each plane
// the reconstruction inputs of the tile plus a two-cell halo in i and j are staged in LDS
// (coalesced along i); the k-direction stencil of a thread's own column is read through
// the caches (each plane arrives from HBM once and is reused from L1/L2 by the next
// iterations).  Each thread computes the west (i), south (j) and top (k) face of its
// cell, all with one copy of the face code in a non-unrolled job loop to keep register
// pressure low; the tile's east and north boundary faces are computed by lanes of waves 0
// and 1.  F*A of every i/j face goes to LDS as soon as it is computed and each cell sums
// its six faces in the reference order W, E, S, N, B, T (fvcell.d:835-845), so results
// match the face-per-thread path bit for bit.  The top-face flux of plane k is the
// bottom-face flux of plane k+1 (registers).
// Workgroups are numbered so that each XCD receives a contiguous range of tiles, which
// keeps the halo re-reads of neighbouring tiles in the same L2.
#include <hip/hip_runtime.h>

#include "kernel_common.hpp"
#include "launchers.hpp"

namespace gdtk {

constexpr int FT_X = 64, FT_Y = 4, FH_X = FT_X + 4, FH_Y = FT_Y + 4;
enum : int { SL_RHO = 0, SL_VX, SL_VY, SL_VZ, SL_U, SL_A, SL_N };

#ifndef GDTK_FUSED_WAVES_PER_EU
#define GDTK_FUSED_WAVES_PER_EU 2
#endif

// F*A of one face (reconstruction + flux calculator, or the simple-outflow flux of the
// interior cell where the face's boundary computes its own flux).  outflow: 0 none,
// -1/+1 = outsign, the interior cell being R0 (-1) or L0 (+1).
template <int DIM, int FC>
__device__ __forceinline__ void face_FA(const BlockDev& b, const IdealGasDev& gas, const FaceGeom& g,
                                        const StencilCell* c, double aL0, double aR0, const double* len,
                                        int64_t cL0, int64_t cR0, int outflow, double eps, double* FA) {
    using C = CQ<DIM>;
    double F[C::n];
    if (outflow != 0) {
        const StencilCell& in = (outflow < 0) ? c[2] : c[1];
        int64_t ci = (outflow < 0) ? cR0 : cL0;
        outflow_flux<DIM>(g, outflow, in.rho, in.u, b.fs_in[ci + FS_P * b.npad], in.vel, F);
    } else {
        FallbackFS fb{b.fs_in, b.npad, cL0, cR0};
        face_flux_l2r2<DIM, FC>(gas, g, c[0], c[1], c[2], c[3], aL0, aR0, fb, len[0], len[1], len[2], len[3], eps,
                                F);
    }
    for (int q = 0; q < C::n; ++q) FA[q] = F[q] * g.area;
}

template <int DIM, int FC, int STAGE>
__global__ __launch_bounds__(FT_X * FT_Y, GDTK_FUSED_WAVES_PER_EU) void k_fused_stage(
    BlockDev b, IdealGasDev gas, double dt, double eps, CellLimits lim, int KC, int gx, int gy, int gz, int per_xcd) {
    using C = CQ<DIM>;
    constexpr int NQ = C::n;
    __shared__ double S[SL_N][FH_Y][FH_X];
    // F*A exchange: i-faces [FT_Y][FT_X+1][NQ] and j-faces [FT_Y+1][FT_X][NQ], written by the
    // face jobs as they finish so that no flux is held in registers across jobs.
    __shared__ double FAi[FT_Y * (FT_X + 1) * NQ];
    __shared__ double FAj[(FT_Y + 1) * FT_X * NQ];

    // XCD-aware numbering: hardware workgroup w runs on XCD (w % 8); give XCD x the
    // contiguous logical range [x*per_xcd, (x+1)*per_xcd).
    const int hw = blockIdx.x;
    const int logical = (hw % 8) * per_xcd + hw / 8;
    const int total = gx * gy * gz;
    if (logical >= total) return;
    const int bx = logical % gx;
    const int by = (logical / gx) % gy;
    const int bz = logical / (gx * gy);

    const int ti = threadIdx.x, tj = threadIdx.y;
    const int tid = tj * FT_X + ti;
    const int i0 = bx * FT_X, j0 = by * FT_Y;
    const int k0 = (DIM == 3) ? bz * KC : 0;
    const int ncol = min(FT_X, b.nic - i0), nrow = min(FT_Y, b.njc - j0);
    const int kend = (DIM == 3) ? min(k0 + KC, b.nkc) : 1;
    const int i = i0 + ti, j = j0 + tj;
    const bool act = (ti < ncol) && (tj < nrow);
    const bool doE = (tj == 0) && (ti < nrow);
    const bool doN = (tj == 1) && (ti < ncol);
    const int64_t np = b.npad;
    const double* fs = b.fs_in;

    double FAb[NQ];
    for (int q = 0; q < NQ; ++q) FAb[q] = 0.0;

    for (int k = (DIM == 3 ? k0 - 1 : k0); k < kend; ++k) {
        const bool pre = (DIM == 3) && (k < k0);  // k0-1: only the bottom face of plane k0
        if (!pre) {
            __syncthreads();
            // Stage plane k of the tile and its halo (no corner cells) into LDS.
            for (int idx = tid; idx < FH_X * FH_Y; idx += FT_X * FT_Y) {
                const int hx = idx % FH_X, hy = idx / FH_X;
                const bool xin = (hx >= 2) && (hx < 2 + ncol), yin = (hy >= 2) && (hy < 2 + nrow);
                const bool need = (xin && hy < nrow + 4) || (yin && hx < ncol + 4);
                if (need) {
                    const int64_t c = pidx(b, i0 - 2 + hx, j0 - 2 + hy, k);
                    S[SL_RHO][hy][hx] = fs[c + FS_RHO * np];
                    S[SL_VX][hy][hx] = fs[c + FS_VX * np];
                    S[SL_VY][hy][hx] = fs[c + FS_VY * np];
                    S[SL_VZ][hy][hx] = fs[c + FS_VZ * np];
                    S[SL_U][hy][hx] = fs[c + FS_U * np];
                    S[SL_A][hy][hx] = fs[c + FS_A * np];
                }
            }
            __syncthreads();
        }
        // Face jobs: 0 west face (i-face at i), 1 south face (j-face at j), 2 the tile's
        // east (wave 0) or north (wave 1) boundary faces, all from LDS; 3 the top face
        // (k-face at k+1) from the cached flow-state planes.  One copy of the face code.
        double FAt[NQ];
        const int job_lo = pre ? 3 : 0;
#pragma clang loop unroll(disable)
        for (int job = job_lo; job < (DIM == 3 ? 4 : 3); ++job) {
            bool doit;
            int d;
            int fi, fj, fk;     // face index
            int hx0 = 0, hy0 = 0, sx = 0, sy = 0;  // LDS stencil start and step
            int of = 0;
            if (job == 0) {
                doit = act; d = 0; fi = i; fj = j; fk = k;
                hx0 = ti; hy0 = tj + 2; sx = 1;
                of = (i == 0 && b.bc_kind[0] == BC_OUTFLOW_SIMPLE_FLUX) ? -1 : 0;
            } else if (job == 1) {
                doit = act; d = 1; fi = i; fj = j; fk = k;
                hx0 = ti + 2; hy0 = tj; sy = 1;
                of = (j == 0 && b.bc_kind[2] == BC_OUTFLOW_SIMPLE_FLUX) ? -1 : 0;
            } else if (job == 2) {
                if (tj == 0) {
                    doit = doE; d = 0; fi = i0 + ncol; fj = j0 + ti; fk = k;
                    hx0 = ncol; hy0 = ti + 2; sx = 1;
                    of = (fi == b.nic && b.bc_kind[1] == BC_OUTFLOW_SIMPLE_FLUX) ? 1 : 0;
                } else {
                    doit = doN; d = 1; fi = i0 + ti; fj = j0 + nrow; fk = k;
                    hx0 = ti + 2; hy0 = nrow; sy = 1;
                    of = (fj == b.njc && b.bc_kind[3] == BC_OUTFLOW_SIMPLE_FLUX) ? 1 : 0;
                }
            } else {
                doit = act; d = 2; fi = i; fj = j; fk = k + 1;
                if (fk == 0 && b.bc_kind[4] == BC_OUTFLOW_SIMPLE_FLUX) of = -1;
                if (fk == b.nkc && b.bc_kind[5] == BC_OUTFLOW_SIMPLE_FLUX) of = 1;
            }
            if (!doit) continue;
            StencilCell cs[4];
            double len[4];
            double aL0, aR0;
            // Padded cell index of stencil cell m: the face sits between m = 1 and m = 2.
            auto cell_at = [&](int m) -> int64_t {
                if (d == 0) return pidx(b, fi - 2 + m, fj, fk);
                if (d == 1) return pidx(b, fi, fj - 2 + m, fk);
                return pidx(b, fi, fj, fk - 2 + m);
            };
            if (job < 3) {
                for (int m = 0; m < 4; ++m) {
                    const int hx = hx0 + m * sx, hy = hy0 + m * sy;
                    cs[m].rho = S[SL_RHO][hy][hx];
                    cs[m].u = S[SL_U][hy][hx];
                    cs[m].vel.x = S[SL_VX][hy][hx];
                    cs[m].vel.y = S[SL_VY][hy][hx];
                    cs[m].vel.z = S[SL_VZ][hy][hx];
                }
                aL0 = S[SL_A][hy0 + sy][hx0 + sx];
                aR0 = S[SL_A][hy0 + 2 * sy][hx0 + 2 * sx];
            } else {
                for (int m = 0; m < 4; ++m) cs[m] = load_sc(fs, np, cell_at(m));
                aL0 = fs[cell_at(1) + FS_A * np];
                aR0 = fs[cell_at(2) + FS_A * np];
            }
            for (int m = 0; m < 4; ++m) len[m] = b.len[cell_at(m) + d * np];
            FaceGeom g = load_face(b.fgeom[d], b.nface[d], face_index(b, d, fi, fj, fk));
            double FA[NQ];
            face_FA<DIM, FC>(b, gas, g, cs, aL0, aR0, len, cell_at(1), cell_at(2), of, eps, FA);
            if (job == 3) {
                for (int q = 0; q < NQ; ++q) FAt[q] = FA[q];
            } else {
                // i-faces at tile column (fi - i0), j-faces at tile row (fj - j0).
                double* dst = (d == 0) ? &FAi[((fj - j0) * (FT_X + 1) + (fi - i0)) * NQ]
                                       : &FAj[((fj - j0) * FT_X + (fi - i0)) * NQ];
                for (int q = 0; q < NQ; ++q) dst[q] = FA[q];
            }
        }
        if (pre) {
            // The top face of plane k0-1 is the bottom face of plane k0.
            if (act)
                for (int q = 0; q < NQ; ++q) FAb[q] = FAt[q];
            continue;
        }
        __syncthreads();  // every face of the plane is in the exchange arrays
        if (act) {
            // time_derivatives order W, E, S, N, B, T (fvcell.d:835-845), F*(outsign*A).
            double si[NQ];
            for (int q = 0; q < NQ; ++q) {
                double s = 0.0;
                s -= -FAi[(tj * (FT_X + 1) + ti) * NQ + q];
                s -= FAi[(tj * (FT_X + 1) + ti + 1) * NQ + q];
                s -= -FAj[(tj * FT_X + ti) * NQ + q];
                s -= FAj[((tj + 1) * FT_X + ti) * NQ + q];
                if (DIM == 3) {
                    s -= -FAb[q];
                    s -= FAt[q];
                }
                si[q] = s;
            }
            update_decode_store<DIM, STAGE>(b, gas, pidx(b, i, j, k), si, dt, lim.low_T, lim.max_vel, lim.min_T,
                                            lim.max_T);
            if (DIM == 3)
                for (int q = 0; q < NQ; ++q) FAb[q] = FAt[q];
        }
    }
}

static int g_kchunk = 16;

hipError_t launch_fused_stage(const BlockDev& b, const IdealGasDev& gas, int stage, int flux_calc, double dt,
                              double eps, const CellLimits& lim, hipStream_t s) {
    const int KC = (b.dims == 3) ? g_kchunk : 1;
    const int gx = (b.nic + FT_X - 1) / FT_X, gy = (b.njc + FT_Y - 1) / FT_Y;
    const int gz = (b.dims == 3) ? (b.nkc + KC - 1) / KC : 1;
    const int total = gx * gy * gz;
    const int per_xcd = (total + 7) / 8;
    dim3 grid(8 * per_xcd), blk(FT_X, FT_Y);
#define GDTK_FUSED(D, F, ST) \
    hipLaunchKernelGGL((k_fused_stage<D, F, ST>), grid, blk, 0, s, b, gas, dt, eps, lim, KC, gx, gy, gz, per_xcd)
    if (b.dims == 3) {
        if (flux_calc == 1) { if (stage == 1) GDTK_FUSED(3, 1, 1); else GDTK_FUSED(3, 1, 2); }
        else { if (stage == 1) GDTK_FUSED(3, 0, 1); else GDTK_FUSED(3, 0, 2); }
    } else {
        if (flux_calc == 1) { if (stage == 1) GDTK_FUSED(2, 1, 1); else GDTK_FUSED(2, 1, 2); }
        else { if (stage == 1) GDTK_FUSED(2, 0, 1); else GDTK_FUSED(2, 0, 2); }
    }
#undef GDTK_FUSED
    return hipGetLastError();
}

void set_fused_kchunk(int kc) {
    if (kc >= 1) g_kchunk = kc;
}

}  // namespace gdtk
