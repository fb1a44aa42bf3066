// mlp_launch.hpp -- launchers of the surrogate-network kernels (mlp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace gdtk {

hipError_t launch_mfma_probe(const float* A, const float* B, float* D, hipStream_t s);
hipError_t launch_mlp_pack(const float* W, int n_out, int n_in, int chain, unsigned short* frag, hipStream_t s);
hipError_t launch_mlp_infer(int H, int NH, const float* x, int n_in, int n, const unsigned short* frag,
                            const float* bias, int n_out, float* y, hipStream_t s, const unsigned* count = nullptr);
hipError_t launch_gemm_f32(int M, int N, int K, const float* A, int sam, int sak, const float* B, int sbk, int sbn,
                           float* C, int ldc, int accumulate, hipStream_t s);
hipError_t launch_bias_act(int rows, int cols, const float* bias, float* Z, float* A, int act, int bf16,
                           hipStream_t s);
hipError_t launch_round_bf16(size_t n, const float* in, float* out, hipStream_t s);
hipError_t launch_gelu_back(size_t n, const float* Z, const float* dA, float* dZ, hipStream_t s);
int mse_parts(int rows, int cols);
hipError_t launch_mse(int rows, int cols, const float* Yp, const float* Yt, float* dY, float* part, hipStream_t s);
int gemm_split(int M, int N, int K);
hipError_t launch_gemm_f32_split(int M, int N, int K, const float* A, int sam, int sak, const float* B, int sbk,
                                 int sbn, float* C, int ldc, float* parts, int nz, hipStream_t s);
int colsum_parts(int rows);
hipError_t launch_colsum(int rows, int cols, const float* dZ, float* db, float* part, hipStream_t s);
hipError_t launch_adam(size_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       int t, hipStream_t s);
hipError_t launch_gather_batch(int B, int width, int n, const int* idx, const float* src, float* out, hipStream_t s);

}  // namespace gdtk
