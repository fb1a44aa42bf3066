// fused_stage.hip -- one kernel per predictor-corrector stage: reconstruction, fluxes,
// time derivatives, stage update, decode and validity check of every cell of a block.
//
// Mapping (gfx950, 64-wide waves): a workgroup of 256 threads owns a tile of TX (i) x TY
// (j) cells, TX * TY = 256 (64 x 4, 32 x 8 or 16 x 16, chosen by the autotuner: narrow
// tiles waste fewer lanes on blocks whose i extent is not a multiple of 64 and compute
// fewer tile-boundary faces), and marches up a chunk of KC planes in k (2.5-D blocking).  For each plane
// the reconstruction inputs of the tile plus a two-cell halo in i and j are staged in LDS
// (coalesced along i); the k-direction stencil of a thread's own column is read through
// the caches (each plane arrives from HBM once and is reused from L1/L2 by the next
// iterations).  Each thread computes the west (i), south (j) and top (k) face of its
// cell, all with one copy of the face code in a non-unrolled job loop to keep register
// pressure low; the tile's east and north boundary faces are computed by the first
// TY + TX threads.  F*A of every i/j face goes to LDS as soon as it is computed and each cell sums
// its six faces in the reference order W, E, S, N, B, T (fvcell.d:835-845), so results
// match the face-per-thread path bit for bit.  The top-face flux of plane k is the
// bottom-face flux of plane k+1 (registers).
// One persistent launch per stage covers every block of the simulation: the tiles of all
// blocks form one list (block by block, i fastest), split into eight contiguous ranges, one
// per XCD (workgroup id mod 8), so that neighbouring tiles, whose halos overlap, are
// processed by the same L2.  A workgroup takes the next tile of its XCD's range from a
// queue counter and, when that range is empty, takes tiles from the other ranges, so
// blocks and tiles of uneven size keep every CU busy to the end.
#include <hip/hip_runtime.h>

#include "kernel_common.hpp"
#include "launchers.hpp"

namespace gdtk {

// Tile shapes: TX x TY cells with TX * TY = FT_THREADS, one thread per cell; the
// autotuner picks TX (16, 32 or 64).
constexpr int FT_THREADS = 256;
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

// LDS of one tile: the reconstruction inputs of a plane with its halo, and the F*A
// exchange arrays -- i-faces [FT_Y][FT_X+1][NQ] and j-faces [FT_Y+1][FT_X][NQ], written by
// the face jobs as they finish so that no flux is held in registers across jobs.
template <int NQ, int TX>
struct TileLds {
    static constexpr int TY = FT_THREADS / TX, HX = TX + 4, HY = TY + 4;
    double S[SL_N][HY][HX];
    double FAi[TY * (TX + 1) * NQ];
    double FAj[(TY + 1) * TX * NQ];
};

// One tile (TX x TY cells, k planes [bz*KC, bz*KC + KC)) of one block.
template <int DIM, int FC, int STAGE, int TX>
__device__ __forceinline__ void fused_tile(const BlockDev& b, const IdealGasDev& gas, double dt, double eps,
                                           const CellLimits& lim, int KC, int bx, int by, int bz,
                                           TileLds<CQ<DIM>::n, TX>& L) {
    using C = CQ<DIM>;
    constexpr int NQ = C::n;
    constexpr int FT_X = TX, FT_Y = FT_THREADS / TX, FH_X = FT_X + 4, FH_Y = FT_Y + 4;
    auto& S = L.S;
    double* FAi = L.FAi;
    double* FAj = L.FAj;

    const int tid = threadIdx.x;
    const int ti = tid % FT_X, tj = tid / FT_X;
    const int i0 = bx * FT_X, j0 = by * FT_Y;
    const int k0 = (DIM == 3) ? bz * KC : 0;
    const int ncol = min(FT_X, b.nic - i0), nrow = min(FT_Y, b.njc - j0);
    const int kend = (DIM == 3) ? min(k0 + KC, b.nkc) : 1;
    const int i = i0 + ti, j = j0 + tj;
    const bool act = (ti < ncol) && (tj < nrow);
    // The tile's east faces (one per row) go to threads 0..FT_Y-1, its north faces (one
    // per column) to threads FT_Y..FT_Y+FT_X-1.
    const bool eastJob = tid < FT_Y;
    const bool doE = eastJob && (tid < nrow);
    const bool doN = !eastJob && (tid - FT_Y < ncol) && (tid < FT_Y + FT_X);
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
        // east or north boundary faces, all from LDS; 3 the top face
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
                if (eastJob) {
                    doit = doE; d = 0; fi = i0 + ncol; fj = j0 + tid; fk = k;
                    hx0 = ncol; hy0 = tid + 2; sx = 1;
                    of = (fi == b.nic && b.bc_kind[1] == BC_OUTFLOW_SIMPLE_FLUX) ? 1 : 0;
                } else {
                    const int c = tid - FT_Y;
                    doit = doN; d = 1; fi = i0 + c; fj = j0 + nrow; fk = k;
                    hx0 = c + 2; hy0 = nrow; sy = 1;
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

// Persistent stage kernel over every block (see the header comment).
template <int DIM, int FC, int STAGE, int TX>
__global__ __launch_bounds__(FT_THREADS, GDTK_FUSED_WAVES_PER_EU) void k_fused_batch(
    StageBatch sb, IdealGasDev gas, const double* __restrict__ dtp, double eps, CellLimits lim) {
    constexpr int TY = FT_THREADS / TX;
    __shared__ TileLds<CQ<DIM>::n, TX> L;
    __shared__ int s_tile;
    const int xcd = blockIdx.x & 7;
    const double dt = *dtp;
    const bool leader = threadIdx.x == 0;
    int q = 0;  // queues tried so far (leader only)
    while (true) {
        if (leader) {
            int t = -1;
            for (; q < 8; ++q) {
                const int x = (xcd + q) & 7;
                const int lo = x * sb.per_xcd;
                const int hi = min(lo + sb.per_xcd, sb.total_tiles);
                if (lo >= hi) continue;
                const unsigned got = atomicAdd(&sb.queue[x * TILE_QSTRIDE], 1u);
                if ((long long)lo + got < (long long)hi) {
                    t = lo + (int)got;
                    break;
                }
            }
            s_tile = t;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t < 0) break;
        // Block of tile t: the last block whose first tile is <= t.
        int lo = 0, hi = sb.nblocks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sb.tile_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int blk = __builtin_amdgcn_readfirstlane(lo);
        const BlockDev& b = sb.tab[blk];
        const int lt = t - sb.tile_start[blk];
        const int gx = (b.nic + TX - 1) / TX, gy = (b.njc + TY - 1) / TY;
        fused_tile<DIM, FC, STAGE, TX>(b, gas, dt, eps, lim, (DIM == 3) ? sb.KC : 1, lt % gx, (lt / gx) % gy,
                                       lt / (gx * gy), L);
        __syncthreads();  // the tile's LDS and s_tile are free again
    }
}

int fused_tiles(const BlockDev& b, int kc, int tx) {
    const int ty = FT_THREADS / tx;
    const int gx = (b.nic + tx - 1) / tx, gy = (b.njc + ty - 1) / ty;
    const int gz = (b.dims == 3) ? (b.nkc + kc - 1) / kc : 1;
    return gx * gy * gz;
}

bool fused_tile_x_ok(int tx) { return tx == 16 || tx == 32 || tx == 64; }

template <int DIM, int FC, int TX>
static int occupancy_of() {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fused_batch<DIM, FC, 2, TX>, FT_THREADS, 0) !=
            hipSuccess ||
        occ < 1)
        occ = 1;
    return occ;
}

int fused_batch_grid(int dims, int flux_calc, int tx) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    int occ = 1;
#define GDTK_OCC(TXV)                                                                            \
    occ = (dims == 3) ? (flux_calc == 1 ? occupancy_of<3, 1, TXV>() : occupancy_of<3, 0, TXV>()) \
                      : (flux_calc == 1 ? occupancy_of<2, 1, TXV>() : occupancy_of<2, 0, TXV>())
    if (tx == 16) GDTK_OCC(16);
    else if (tx == 32) GDTK_OCC(32);
    else GDTK_OCC(64);
#undef GDTK_OCC
    return ncu * occ;
}

template <int TX>
static void launch_fused_tx(const StageBatch& sb, int dims, int stage, int flux_calc, const double* dtp, double eps,
                            const CellLimits& lim, const IdealGasDev& gas, dim3 g, hipStream_t s) {
    dim3 blk(FT_THREADS);
#define GDTK_FUSED(D, F, ST) \
    hipLaunchKernelGGL((k_fused_batch<D, F, ST, TX>), g, blk, 0, s, sb, gas, dtp, eps, lim)
    if (dims == 3) {
        if (flux_calc == 1) { if (stage == 1) GDTK_FUSED(3, 1, 1); else GDTK_FUSED(3, 1, 2); }
        else { if (stage == 1) GDTK_FUSED(3, 0, 1); else GDTK_FUSED(3, 0, 2); }
    } else {
        if (flux_calc == 1) { if (stage == 1) GDTK_FUSED(2, 1, 1); else GDTK_FUSED(2, 1, 2); }
        else { if (stage == 1) GDTK_FUSED(2, 0, 1); else GDTK_FUSED(2, 0, 2); }
    }
#undef GDTK_FUSED
}

hipError_t launch_fused_batch(const StageBatch& sb, int dims, int stage, int flux_calc, const double* dtp,
                              double eps, const CellLimits& lim, const IdealGasDev& gas, int grid, int tx,
                              hipStream_t s) {
    if (sb.total_tiles <= 0) return hipSuccess;
    if (!fused_tile_x_ok(tx) || (flux_calc != 0 && flux_calc != 1)) return hipErrorInvalidValue;
    grid = (grid + 7) / 8 * 8;
    dim3 g(grid);
    if (tx == 16) launch_fused_tx<16>(sb, dims, stage, flux_calc, dtp, eps, lim, gas, g, s);
    else if (tx == 32) launch_fused_tx<32>(sb, dims, stage, flux_calc, dtp, eps, lim, gas, g, s);
    else launch_fused_tx<64>(sb, dims, stage, flux_calc, dtp, eps, lim, gas, g, s);
    return hipGetLastError();
}

// Clears n unsigned counters (queue and cell counters) with a kernel, so that the
// clearing can be part of a captured step.
__global__ void k_clear_u32(unsigned* p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

hipError_t launch_clear_u32(unsigned* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Copies of several device buffers in one launch (the step-start flow state of every
// block), so that the copies are one node of a captured step.  Segment q covers chunks
// [segs[q].start, segs[q + 1].start) of 1024 doubles; each lane moves 4 doubles of one
// chunk (two 16-byte loads and stores where the chunk is whole).
__global__ __launch_bounds__(256) void k_copy_segments(const CopySeg* __restrict__ segs, int nseg) {
    const int64_t chunk = blockIdx.x;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].start <= chunk) lo = mid;
        else hi = mid - 1;
    }
    const CopySeg sg = segs[__builtin_amdgcn_readfirstlane(lo)];
    const int64_t i0 = (chunk - sg.start) * 1024 + 4 * (int64_t)threadIdx.x;
    if (i0 + 4 <= sg.n) {
        const double2* src = reinterpret_cast<const double2*>(sg.src + i0);
        double2* dst = reinterpret_cast<double2*>(sg.dst + i0);
        const double2 a = src[0], b = src[1];
        dst[0] = a;
        dst[1] = b;
    } else {
        for (int64_t i = i0; i < sg.n && i < i0 + 4; ++i) sg.dst[i] = sg.src[i];
    }
}

hipError_t launch_copy_segments(const CopySeg* segs, int nseg, int64_t nchunks, hipStream_t s) {
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_segments, dim3((unsigned)nchunks), dim3(256), 0, s, segs, nseg);
    return hipGetLastError();
}

// Writes one double on the device (the step's dt), ordered on the stream.
__global__ void k_set_f64(double* p, double v) {
    if (threadIdx.x == 0) *p = v;
}

hipError_t launch_set_f64(double* p, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

}  // namespace gdtk
