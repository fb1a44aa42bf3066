// mlp_kernels.hip -- the chemistry surrogate network on gfx950: FP32 training kernels and a
// BF16 MFMA inference kernel (no counterpart in the reference).
//
// Network: x[n_in] -> (W_l x + b_l, GELU) for the hidden layers of width H -> W_out h + b_out,
// with GELU in its tanh form (Hendrycks & Gimpel 2016):
//   GELU(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) = z / (1 + exp(-2 u)),
//   u = sqrt(2/pi) (z + 0.044715 z^3).
// Activations are cell-major [batch][width] in training.
//
// Inference keeps a wave's 16 cells in registers through every layer.  Each layer is
// computed transposed, Y^T = W X^T, with v_mfma_f32_16x16x32_bf16:
//   A operand (16 x 32) = a 16-row block of W,  lane l holds A[l%16][slot (l/16, j)], j < 8
//   B operand (32 x 16) = 32 inputs x 16 cells,  lane l holds B[slot (l/16, j)][l%16]
//   D (16 x 16) = 16 outputs x 16 cells,          lane l holds D[4 (l/16) + i][l%16], i < 4
// The sum over K pairs slot (g, j) of A with slot (g, j) of B, so any assignment of input
// indices to slots works if A and B agree.  The B fragment of k-step s of the next layer is
// taken straight from this layer's D registers of row blocks 2s and 2s+1 (after bias, GELU
// and rounding to BF16): slot (g, j) holds input 32 s + 4 g + j for j < 4 and
// 32 s + 16 + 4 g + (j - 4) for j >= 4.  The packed weights of that layer use the same slot
// map, so activations never leave registers.  The first layer's slots are k = 8 g + j over
// the zero-padded inputs.  Packed fragments are 64 lanes x 8 BF16 = 1 KiB, one 16-byte load
// per lane.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp_launch.hpp"

namespace gdtk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr float GELU_K = 0.7978845608028654f;   // sqrt(2/pi)
constexpr float GELU_C = 0.044715f;
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + tanhf(GELU_K * (z + GELU_C * z * z * z))); }
__device__ __forceinline__ float gelu_grad(float z) {
    // d/dz = (1 + t)/2 + z/2 (1 - t^2) sqrt(2/pi) (1 + 3 c z^2), t = tanh(u)
    const float t = tanhf(GELU_K * (z + GELU_C * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * GELU_K * (1.0f + 3.0f * GELU_C * z * z);
}
// The same function for inference: z / (1 + 2^(z (A + B z^2))) with the hardware exp2 and
// reciprocal (A = -2 log2(e) sqrt(2/pi), B = 0.044715 A); the result is rounded to BF16.
__device__ __forceinline__ float gelu_fast(float z) {
    constexpr float A = -2.0f * 1.4426950408889634f * 0.7978845608028654f;
    constexpr float B = 0.044715f * A;
    const float w = z * (A + B * (z * z));
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(w));
}

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Slot map of layer inputs: k-step s, lane group g, element j -> input index.
__device__ __host__ inline int slot_first(int s, int g, int j) { return 32 * s + 8 * g + j; }
__device__ __host__ inline int slot_chain(int s, int g, int j) { return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }

//---------------------------------------------------------------------------------
// Lane-layout probe: D = A B for one 16x32 A and 32x16 B (row-major FP32 inputs, exact in
// BF16), through the MFMA instruction and the slot map above with slots k = 8 g + j.
__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    const int l = threadIdx.x;
    const int r = l & 15, g = l >> 4;
    u16x8 a, b;
    for (int j = 0; j < 8; ++j) {
        a[j] = f2bf(A[r * 32 + slot_first(0, g, j)]);
        b[j] = f2bf(B[slot_first(0, g, j) * 16 + r]);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc, 0,
                                                  0, 0);
    for (int i = 0; i < 4; ++i) D[(4 * g + i) * 16 + r] = acc[i];
}

hipError_t launch_mfma_probe(const float* A, const float* B, float* D, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

//---------------------------------------------------------------------------------
// Packing of FP32 weights W[out][in] (row-major) into BF16 fragments:
// frag[((rb * nks) + s) * 64 + lane][j] = W[16 rb + l%16][slot(s, l/16, j)], zero outside.
__global__ void k_mlp_pack(const float* __restrict__ W, int n_out, int n_in, int nrb, int nks, int chain,
                           unsigned short* __restrict__ frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = nrb * nks * 64 * 8;
    if (t >= total) return;
    const int j = t & 7, lane = (t >> 3) & 63, fs = t >> 9;
    const int s = fs % nks, rb = fs / nks;
    const int row = 16 * rb + (lane & 15), g = lane >> 4;
    const int k = chain ? slot_chain(s, g, j) : slot_first(s, g, j);
    const float v = (row < n_out && k < n_in) ? W[(size_t)row * n_in + k] : 0.0f;
    frag[t] = f2bf(v);
}

hipError_t launch_mlp_pack(const float* W, int n_out, int n_in, int chain, unsigned short* frag, hipStream_t s) {
    const int nrb = (n_out + 15) / 16, nks = (n_in + 31) / 32;
    const int total = nrb * nks * 64 * 8;
    hipLaunchKernelGGL(k_mlp_pack, dim3((total + 255) / 256), dim3(256), 0, s, W, n_out, n_in, nrb, nks, chain, frag);
    return hipGetLastError();
}

//---------------------------------------------------------------------------------
// Inference: x SoA [n_in][n] FP32 -> y SoA [n_out][n] FP32.  H hidden width (multiple of
// 32), NH hidden layers, n_in <= 32, n_out <= 16.  frag: packed layers in order; bias:
// FP32 biases in order.
//
// A workgroup of 4 waves serves 128 cells, each wave two 16-cell column blocks (CB = 2), so
// every weight fragment read feeds two MFMAs.  The H x H layers stream their fragments
// through LDS one k-step at a time (NRB fragments = 16 KiB for H = 256), double-buffered:
// the global loads of k-step s+1 are in flight while k-step s is multiplied.  The small
// first and last layers read their fragments straight from global memory (L2).
constexpr int INF_WAVES = 4, INF_CB = 2;
constexpr int INF_CELLS = INF_WAVES * INF_CB * 16;

// count (may be null): the number of leading cells to evaluate, read from device memory (at
// most n, the arrays' stride); workgroups past it exit at once.
template <int H, int NH>
__global__ __launch_bounds__(256, 1) void k_mlp_infer(const float* __restrict__ x, int n_in, int n,
                                                      const unsigned* __restrict__ count,
                                                      const unsigned short* __restrict__ frag,
                                                      const float* __restrict__ bias, int n_out, float* __restrict__ y) {
    const int stride = n;
    if (count) {
        const unsigned c = *count;
        n = (c < (unsigned)n) ? (int)c : n;
        if ((int)(blockIdx.x * INF_CELLS) >= n) return;
    }
    constexpr int NRB = H / 16, NKS = H / 32;
    constexpr int FRAG_PER_STEP = NRB * 64;                 // 16-byte fragments per k-step
    constexpr int LOADS = FRAG_PER_STEP / 256;               // per thread
    static_assert(FRAG_PER_STEP % 256 == 0, "k-step fragments must split over the workgroup");
    __shared__ bf16x8 lds[2][FRAG_PER_STEP];
    __shared__ float bias_s[NH * H + 16];
    for (int k = threadIdx.x; k < NH * H + n_out; k += 256) bias_s[k] = bias[k];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int cell0 = blockIdx.x * INF_CELLS + wave * (INF_CB * 16);
    const bf16x8* F = reinterpret_cast<const bf16x8*>(frag);

    bf16x8 act[INF_CB][NKS];
    f32x4 acc[INF_CB][NRB];
    // first layer: one k-step over the zero-padded inputs (slots k = 8 g + j)
#pragma unroll
    for (int cb = 0; cb < INF_CB; ++cb) {
        const int cell = cell0 + cb * 16 + r;
        u16x8 xb;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            xb[j] = (cell < n && k < n_in) ? f2bf(x[(size_t)k * stride + cell]) : (unsigned short)0;
        }
        act[cb][0] = __builtin_bit_cast(bf16x8, xb);
    }
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
        const bf16x8 a = F[rb * 64 + lane];
#pragma unroll
        for (int cb = 0; cb < INF_CB; ++cb)
            acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, act[cb][0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    const bf16x8* Fl = F + NRB * 64;
    __syncthreads();  // biases in LDS
    const float* bl = bias_s;
    for (int layer = 0; layer < NH; ++layer) {
        // bias + GELU + BF16 -> B fragments of the next layer
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            const float4 b0 = *reinterpret_cast<const float4*>(bl + 32 * s + 4 * g);
            const float4 b1 = *reinterpret_cast<const float4*>(bl + 32 * s + 16 + 4 * g);
#pragma unroll
            for (int cb = 0; cb < INF_CB; ++cb) {
                bf16x8 v;
                v[0] = (__bf16)gelu_fast(acc[cb][2 * s][0] + b0.x);
                v[1] = (__bf16)gelu_fast(acc[cb][2 * s][1] + b0.y);
                v[2] = (__bf16)gelu_fast(acc[cb][2 * s][2] + b0.z);
                v[3] = (__bf16)gelu_fast(acc[cb][2 * s][3] + b0.w);
                v[4] = (__bf16)gelu_fast(acc[cb][2 * s + 1][0] + b1.x);
                v[5] = (__bf16)gelu_fast(acc[cb][2 * s + 1][1] + b1.y);
                v[6] = (__bf16)gelu_fast(acc[cb][2 * s + 1][2] + b1.z);
                v[7] = (__bf16)gelu_fast(acc[cb][2 * s + 1][3] + b1.w);
                act[cb][s] = v;
            }
        }
        bl += H;
        if (layer == NH - 1) break;
#pragma unroll
        for (int cb = 0; cb < INF_CB; ++cb)
#pragma unroll
            for (int rb = 0; rb < NRB; ++rb) acc[cb][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
        // k-step s of this layer: fragments Fl[(rb * NKS + s) * 64 + lane] for rb < NRB
        bf16x8 pre[LOADS];
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const int f = threadIdx.x + 256 * q;  // f = rb * 64 + lane'
            pre[q] = Fl[((f >> 6) * NKS + 0) * 64 + (f & 63)];
        }
#pragma unroll
        for (int q = 0; q < LOADS; ++q) lds[0][threadIdx.x + 256 * q] = pre[q];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            const int buf = s & 1;
            if (s + 1 < NKS) {
#pragma unroll
                for (int q = 0; q < LOADS; ++q) {
                    const int f = threadIdx.x + 256 * q;
                    pre[q] = Fl[((f >> 6) * NKS + (s + 1)) * 64 + (f & 63)];
                }
            }
#pragma unroll
            for (int rb = 0; rb < NRB; ++rb) {
                const bf16x8 a = lds[buf][rb * 64 + lane];
#pragma unroll
                for (int cb = 0; cb < INF_CB; ++cb)
                    acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, act[cb][s], acc[cb][rb], 0, 0, 0);
            }
            if (s + 1 < NKS) {
#pragma unroll
                for (int q = 0; q < LOADS; ++q) lds[buf ^ 1][threadIdx.x + 256 * q] = pre[q];
            }
            __syncthreads();
        }
        Fl += NRB * NKS * 64;
    }
    // output layer: one 16-row block, fragments from global
#pragma unroll
    for (int cb = 0; cb < INF_CB; ++cb) {
        f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < NKS; ++s) o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Fl[s * 64 + lane], act[cb][s], o, 0, 0, 0);
        const int cell = cell0 + cb * 16 + r;
        if (cell < n) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 4 * g + i;
                if (row < n_out) y[(size_t)row * stride + cell] = o[i] + bl[row];
            }
        }
    }
}

hipError_t launch_mlp_infer(int H, int NH, const float* x, int n_in, int n, const unsigned short* frag,
                            const float* bias, int n_out, float* y, hipStream_t s, const unsigned* count) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + INF_CELLS - 1) / INF_CELLS);
    if (H == 256 && NH == 3) {
        hipLaunchKernelGGL((k_mlp_infer<256, 3>), dim3(grid), dim3(256), 0, s, x, n_in, n, count, frag, bias, n_out,
                           y);
    } else if (H == 64 && NH == 2) {
        hipLaunchKernelGGL((k_mlp_infer<64, 2>), dim3(grid), dim3(256), 0, s, x, n_in, n, count, frag, bias, n_out,
                           y);
    } else {
        return hipErrorNotSupported;
    }
    return hipGetLastError();
}

//---------------------------------------------------------------------------------
// FP32 GEMM for training: C[M][N] (+)= sum_k A(m,k) B(k,n) with
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C row-major with leading dim ldc.
// 64x64 tiles, 256 threads with 4x4 outputs each, k tiles of 16 through LDS.
constexpr int GT = 64, GK = 16;

// Split K: workgroup z covers k in [z kc, (z+1) kc) and writes its partial product to
// C + z * M * ldc (summed afterwards in z order by k_sum_parts, so results do not depend on
// scheduling).
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, const float* __restrict__ A, int sam, int sak,
                                                  const float* __restrict__ B, int sbk, int sbn, float* __restrict__ C,
                                                  int ldc, int accumulate, int kc) {
    __shared__ float As[GK][GT + 1];
    __shared__ float Bs[GK][GT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int kbeg = blockIdx.z * kc;
    const int kend = (kbeg + kc < K) ? kbeg + kc : K;
    C += (size_t)blockIdx.z * M * ldc;
    float c[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) c[a][b] = 0.0f;
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        for (int e = threadIdx.x; e < GK * GT; e += 256) {
            const int kk = e / GT, mm = e % GT;
            const int m = m0 + mm, k = k0 + kk;
            As[kk][mm] = (m < M && k < kend) ? A[(size_t)m * sam + (size_t)k * sak] : 0.0f;
            const int nn = mm, nidx = n0 + nn;
            Bs[kk][nn] = (nidx < N && k < kend) ? B[(size_t)k * sbk + (size_t)nidx * sbn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk][ty * 4 + q];
                b[q] = Bs[kk][tx * 4 + q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) c[p][q] += a[p] * b[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int m = m0 + ty * 4 + p;
        if (m >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int nidx = n0 + tx * 4 + q;
            if (nidx >= N) continue;
            float* dst = C + (size_t)m * ldc + nidx;
            *dst = accumulate ? *dst + c[p][q] : c[p][q];
        }
    }
}

// C[m][n] = sum_z P[z][m][n] (row-major, leading dimension ldc), z ascending.
__global__ void k_sum_parts(int M, int N, int ldc, int nz, const float* __restrict__ P, float* __restrict__ C) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * N) return;
    const size_t m = t / N, nn = t % N;
    float acc = 0.0f;
    for (int z = 0; z < nz; ++z) acc += P[(size_t)z * M * ldc + m * ldc + nn];
    C[m * ldc + nn] = acc;
}

hipError_t launch_gemm_f32(int M, int N, int K, const float* A, int sam, int sak, const float* B, int sbk, int sbn,
                           float* C, int ldc, int accumulate, hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    dim3 grid((unsigned)((N + GT - 1) / GT), (unsigned)((M + GT - 1) / GT), 1);
    hipLaunchKernelGGL(k_gemm_f32, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate, K);
    return hipGetLastError();
}

int gemm_split(int M, int N, int K) {
    const int tiles = ((N + GT - 1) / GT) * ((M + GT - 1) / GT);
    int z = 256 / (tiles > 0 ? tiles : 1);
    z = z < 1 ? 1 : (z > 64 ? 64 : z);
    while (z > 1 && K / z < 4 * GK) z >>= 1;
    return z;
}

hipError_t launch_gemm_f32_split(int M, int N, int K, const float* A, int sam, int sak, const float* B, int sbk,
                                 int sbn, float* C, int ldc, float* parts, int nz, hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (nz <= 1) return launch_gemm_f32(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, 0, s);
    const int kc = ((K + nz - 1) / nz + GK - 1) / GK * GK;
    dim3 grid((unsigned)((N + GT - 1) / GT), (unsigned)((M + GT - 1) / GT), (unsigned)nz);
    hipLaunchKernelGGL(k_gemm_f32, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, parts, ldc, 0, kc);
    const size_t mn = (size_t)M * N;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, M, N, ldc, nz, parts, C);
    return hipGetLastError();
}

__device__ __forceinline__ float round_bf16(float f) { return __uint_as_float((unsigned)f2bf(f) << 16); }

// Z[b][j] += bias[j]; A = GELU(Z) (act != 0) or A = Z.  With bf16 != 0 the GELU output is
// rounded to BF16, as the inference kernel rounds it (BF16-aware training; the gradient
// passes the rounding unchanged).
__global__ void k_bias_act(int rows, int cols, const float* __restrict__ bias, float* __restrict__ Z,
                           float* __restrict__ Aout, int act, int bf16) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)rows * cols) return;
    const float z = Z[t] + bias[t % cols];
    Z[t] = z;
    const float a = act ? gelu(z) : z;
    Aout[t] = (act && bf16) ? round_bf16(a) : a;
}

hipError_t launch_bias_act(int rows, int cols, const float* bias, float* Z, float* A, int act, int bf16,
                           hipStream_t s) {
    const size_t n = (size_t)rows * cols;
    hipLaunchKernelGGL(k_bias_act, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, cols, bias, Z, A, act,
                       bf16);
    return hipGetLastError();
}

// out[i] = BF16(in[i]) (as FP32 values)
__global__ void k_round_bf16(size_t n, const float* __restrict__ in, float* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = round_bf16(in[t]);
}

hipError_t launch_round_bf16(size_t n, const float* in, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_round_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, in, out);
    return hipGetLastError();
}

// dZ = dA * GELU'(Z)
__global__ void k_gelu_back(size_t n, const float* __restrict__ Z, const float* __restrict__ dA, float* __restrict__ dZ) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dZ[t] = dA[t] * gelu_grad(Z[t]);
}

hipError_t launch_gelu_back(size_t n, const float* Z, const float* dA, float* dZ, hipStream_t s) {
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, Z, dA, dZ);
    return hipGetLastError();
}

// Mean-squared error over rows x cols with per-row weights 1: dY = 2 (Yp - Yt) / (rows cols);
// the loss sum is accumulated per workgroup into part[blockIdx].
__global__ void k_mse(int rows, int cols, const float* __restrict__ Yp, const float* __restrict__ Yt,
                      float* __restrict__ dY, float* __restrict__ part) {
    __shared__ float red[256];
    const size_t n = (size_t)rows * cols;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (t < n) {
        const float d = Yp[t] - Yt[t];
        dY[t] = 2.0f * d / (float)n;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int mse_parts(int rows, int cols) { return (int)(((size_t)rows * cols + 255) / 256); }

hipError_t launch_mse(int rows, int cols, const float* Yp, const float* Yt, float* dY, float* part, hipStream_t s) {
    hipLaunchKernelGGL(k_mse, dim3((unsigned)mse_parts(rows, cols)), dim3(256), 0, s, rows, cols, Yp, Yt, dY, part);
    return hipGetLastError();
}

// db[j] = sum_b dZ[b][j]: chunks of COLSUM_ROWS rows summed per workgroup into part[chunk][j],
// then the chunks summed in order.
constexpr int COLSUM_ROWS = 128;
__global__ void k_colsum_part(int rows, int cols, const float* __restrict__ dZ, float* __restrict__ part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int b0 = blockIdx.y * COLSUM_ROWS;
    const int b1 = (b0 + COLSUM_ROWS < rows) ? b0 + COLSUM_ROWS : rows;
    float acc = 0.0f;
    for (int b = b0; b < b1; ++b) acc += dZ[(size_t)b * cols + j];
    part[(size_t)blockIdx.y * cols + j] = acc;
}

int colsum_parts(int rows) { return (rows + COLSUM_ROWS - 1) / COLSUM_ROWS; }

hipError_t launch_colsum(int rows, int cols, const float* dZ, float* db, float* part, hipStream_t s) {
    const int nch = colsum_parts(rows);
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((cols + 63) / 64), (unsigned)nch), dim3(64), 0, s, rows, cols, dZ,
                       part);
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, 1, cols, cols, nch, part, db);
    return hipGetLastError();
}

// Adam (Kingma & Ba 2015) with bias correction; t is the 1-based step.
__global__ void k_adam(size_t n, float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t launch_adam(size_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       int t, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1,
                       c2);
    return hipGetLastError();
}

// Gather rows of a SoA [width][n] array into a cell-major batch [B][width]: out[b][f] =
// src[f*n + idx[b]].
__global__ void k_gather_batch(int B, int width, int n, const int* __restrict__ idx, const float* __restrict__ src,
                               float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * width) return;
    const int b = t / width, f = t % width;
    out[t] = src[(size_t)f * n + idx[b]];
}

hipError_t launch_gather_batch(int B, int width, int n, const int* idx, const float* src, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((B * width + 255) / 256)), dim3(256), 0, s, B, width, n, idx, src,
                       out);
    return hipGetLastError();
}

}  // namespace gdtk
