// mlp_host.hpp -- the network handle of the C ABI (defined in mlp.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "host_common.hpp"
#include "mlp_launch.hpp"

namespace gdtk {

struct Mlp {
    int n_in = 0, H = 0, NH = 0, n_out = 0;
    int nlayer = 0;  // NH + 1
    std::vector<int> in_w, out_w;
    std::vector<size_t> woff, boff;
    size_t nparam = 0;
    DevArr<float> params, grads, am, av;
    int adam_t = 0;
    // BF16-aware training: the forward pass uses BF16-rounded inputs, weights (wq, refreshed
    // after every update) and hidden activations, as the inference kernel does; gradients
    // pass the roundings unchanged and update the FP32 masters.
    int bf16_train = 0;
    DevArr<float> wq;
    bool wq_fresh = false;
    // BF16 inference fragments and the concatenated biases
    DevArr<unsigned short> frag;
    DevArr<float> bias_cat;
    bool packed = false;
    // training batch buffers
    int bcap = 0;
    std::vector<DevArr<float>> Z, A;
    DevArr<float> xb, yb, dA, dZ, part, xh, yh, gparts, cparts;
    DevArr<int> idx;
    hipStream_t stream = nullptr;

    ~Mlp() {
        if (stream) (void)hipStreamDestroy(stream);
    }

    void layout() {
        nlayer = NH + 1;
        in_w.assign(nlayer, H);
        out_w.assign(nlayer, H);
        in_w[0] = n_in;
        out_w[nlayer - 1] = n_out;
        woff.resize(nlayer);
        boff.resize(nlayer);
        size_t o = 0;
        for (int l = 0; l < nlayer; ++l) {
            woff[l] = o;
            o += (size_t)out_w[l] * in_w[l];
            boff[l] = o;
            o += out_w[l];
        }
        nparam = o;
    }

    void ensure_batch(int B) {
        if (B <= bcap) return;
        bcap = B;
        Z.clear();
        A.clear();
        Z.resize(nlayer);
        A.resize(nlayer);
        for (int l = 0; l < nlayer; ++l) {
            Z[l].alloc((size_t)B * out_w[l]);
            A[l].alloc((size_t)B * out_w[l]);
        }
        const size_t wmax = (size_t)std::max(std::max(H, n_in), n_out);
        xb.alloc((size_t)B * n_in);
        yb.alloc((size_t)B * n_out);
        dA.alloc((size_t)B * wmax);
        dZ.alloc((size_t)B * wmax);
        part.alloc((size_t)mse_parts(B, n_out));
        size_t gp = 0;
        for (int l = 0; l < nlayer; ++l)
            gp = std::max(gp, (size_t)gemm_split(out_w[l], in_w[l], B) * out_w[l] * in_w[l]);
        gparts.alloc(gp);
        cparts.alloc((size_t)colsum_parts(B) * wmax);
    }

    const float* weights(hipStream_t s) {
        if (!bf16_train) return params.p;
        if (!wq_fresh) {
            wq.reserve(nparam);
            HIPCHK(launch_round_bf16(nparam, params.p, wq.p, s));
            wq_fresh = true;
        }
        return wq.p;
    }

    // Forward pass of a cell-major batch X [B][n_in]; output A[last] = Z[last].
    void forward(int B, const float* X, hipStream_t s) {
        const float* in = X;
        if (bf16_train) {
            HIPCHK(launch_round_bf16((size_t)B * n_in, X, xq_buf(B), s));
            in = xq.p;
        }
        const float* Wt = weights(s);
        for (int l = 0; l < nlayer; ++l) {
            const int ni = in_w[l], no = out_w[l];
            // Z[b][o] = sum_k in[b][k] W[o][k]; the biases stay FP32 (as in inference)
            HIPCHK(launch_gemm_f32(B, no, ni, in, ni, 1, Wt + woff[l], 1, ni, Z[l].p, no, 0, s));
            HIPCHK(launch_bias_act(B, no, params.p + boff[l], Z[l].p, A[l].p, l < nlayer - 1 ? 1 : 0, bf16_train, s));
            in = A[l].p;
        }
    }
    DevArr<float> xq;
    float* xq_buf(int B) {
        xq.reserve((size_t)B * n_in);
        return xq.p;
    }

    // Backward pass after forward(B, X) and the loss gradient in dZ (output layer).
    void backward(int B, const float* X, hipStream_t s) {
        const float* Wt = weights(s);
        for (int l = nlayer - 1; l >= 0; --l) {
            const int ni = in_w[l], no = out_w[l];
            const float* in = (l == 0) ? (bf16_train ? xq.p : X) : A[l - 1].p;
            // dW[o][k] = sum_b dZ[b][o] in[b][k]  (split over the batch)
            HIPCHK(launch_gemm_f32_split(no, ni, B, dZ.p, 1, no, in, ni, 1, grads.p + woff[l], ni, gparts.p,
                                         gemm_split(no, ni, B), s));
            HIPCHK(launch_colsum(B, no, dZ.p, grads.p + boff[l], cparts.p, s));
            if (l == 0) break;
            // dA[b][k] = sum_o dZ[b][o] W[o][k]
            HIPCHK(launch_gemm_f32(B, ni, no, dZ.p, no, 1, Wt + woff[l], ni, 1, dA.p, ni, 0, s));
            HIPCHK(launch_gelu_back((size_t)B * ni, Z[l - 1].p, dA.p, dZ.p, s));
        }
    }

    // Mean-squared-error loss of the batch (sum of part[] / (B n_out)) and its gradient.
    void loss_grad(int B, const float* X, const float* Y, hipStream_t s) {
        forward(B, X, s);
        HIPCHK(launch_mse(B, n_out, A[nlayer - 1].p, Y, dZ.p, part.p, s));
        backward(B, X, s);
    }

    double read_loss(int B, hipStream_t s) {
        const int np = mse_parts(B, n_out);
        std::vector<float> h((size_t)np);
        HIPCHK(hipMemcpyAsync(h.data(), part.p, np * sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double acc = 0.0;
        for (float v : h) acc += v;
        return acc / ((double)B * n_out);
    }

    void adam(float lr, hipStream_t s) {
        ++adam_t;
        HIPCHK(launch_adam(nparam, params.p, grads.p, am.p, av.p, lr, 0.9f, 0.999f, 1.0e-8f, adam_t, s));
        packed = false;
        wq_fresh = false;
    }

    void pack(hipStream_t s) {
        if (packed) return;
        size_t nfrag = 0;
        std::vector<size_t> foff(nlayer);
        for (int l = 0; l < nlayer; ++l) {
            foff[l] = nfrag;
            nfrag += (size_t)((out_w[l] + 15) / 16) * ((in_w[l] + 31) / 32) * 64 * 8;
        }
        frag.reserve(nfrag);
        size_t nb = 0;
        for (int l = 0; l < nlayer; ++l) nb += out_w[l];
        bias_cat.reserve(nb);
        size_t bo = 0;
        for (int l = 0; l < nlayer; ++l) {
            HIPCHK(launch_mlp_pack(params.p + woff[l], out_w[l], in_w[l], l > 0 ? 1 : 0, frag.p + foff[l], s));
            HIPCHK(hipMemcpyAsync(bias_cat.p + bo, params.p + boff[l], out_w[l] * sizeof(float),
                                  hipMemcpyDeviceToDevice, s));
            bo += out_w[l];
        }
        packed = true;
    }
};

Mlp& get_mlp(int h);
// BF16 inference of n cells: x SoA [n_in][n] -> y SoA [n_out][n] (device arrays).
// Inference over x, y [n_in/n_out][n]; with count (device memory, may be null) only the
// first *count cells are evaluated (the arrays keep the stride n).
void mlp_infer(Mlp& m, int n, const float* x, float* y, hipStream_t s, const unsigned* count = nullptr);

}  // namespace gdtk
