// mlp.cpp -- host side of the chemistry surrogate network (C ABI in include/gdtk_hip.h).
//
// The network and its training have no counterpart in the reference: they learn the map
// (state, dt) -> mass-fraction change from samples of this library's own alpha-QSS
// integrator.  Parameters are FP32 masters on the device, laid out layer by layer as
// W_l [out_l][in_l] (row-major) followed by b_l [out_l].  Training runs the FP32 kernels
// (forward, MSE loss, backward, Adam); inference packs the weights into BF16 MFMA
// fragments and runs one fused kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/gdtk_hip.h"
#include "host_common.hpp"
#include "mlp_host.hpp"
#include "mlp_launch.hpp"

namespace gdtk {

static std::mutex g_mlp_mu;
static std::vector<std::unique_ptr<Mlp>> g_mlps;

Mlp& get_mlp(int h) {
    std::lock_guard<std::mutex> lk(g_mlp_mu);
    if (h < 0 || h >= (int)g_mlps.size() || !g_mlps[h]) throw Error(GDTK_ERR_ARG, "bad network handle");
    return *g_mlps[h];
}

// BF16 inference on device SoA arrays (used by the surrogate step as well).
void mlp_infer(Mlp& m, int n, const float* x, float* y, hipStream_t s, const unsigned* count) {
    m.pack(s);
    HIPCHK(launch_mlp_infer(m.H, m.NH, x, m.n_in, n, m.frag.p, m.bias_cat.p, m.n_out, y, s, count));
}

}  // namespace gdtk

using namespace gdtk;

extern "C" {

int gdtk_hip_mlp_create(const gdtk_mlp_desc* d, unsigned long long seed, int* h) {
    return guarded([&] {
        if (!d || !h) throw Error(GDTK_ERR_ARG, "null argument");
        if (d->n_in < 1 || d->n_in > 32 || d->n_out < 1 || d->n_out > 16)
            throw Error(GDTK_ERR_UNSUPPORTED, "inputs must be 1..32 and outputs 1..16");
        if (!((d->hidden == 256 && d->n_hidden_layers == 3) || (d->hidden == 64 && d->n_hidden_layers == 2)))
            throw Error(GDTK_ERR_UNSUPPORTED, "hidden layers: 3 x 256 or 2 x 64");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw Error(GDTK_ERR_HIP, "no HIP device");
        auto m = std::make_unique<Mlp>();
        m->n_in = d->n_in;
        m->H = d->hidden;
        m->NH = d->n_hidden_layers;
        m->n_out = d->n_out;
        m->layout();
        // He-normal weights for the GELU layers (std sqrt(2/in)), zero biases.
        std::vector<float> p(m->nparam, 0.0f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < m->nlayer; ++l) {
            std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / m->in_w[l]));
            for (size_t i = 0; i < (size_t)m->out_w[l] * m->in_w[l]; ++i) p[m->woff[l] + i] = (float)nd(rng);
        }
        m->params.alloc(m->nparam);
        m->grads.alloc(m->nparam);
        m->am.alloc(m->nparam);
        m->av.alloc(m->nparam);
        HIPCHK(hipMemcpy(m->params.p, p.data(), m->nparam * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        std::lock_guard<std::mutex> lk(g_mlp_mu);
        g_mlps.push_back(std::move(m));
        *h = (int)g_mlps.size() - 1;
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_destroy(int h) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_mlp_mu);
        if (h < 0 || h >= (int)g_mlps.size() || !g_mlps[h]) throw Error(GDTK_ERR_ARG, "bad network handle");
        HIPCHK(hipStreamSynchronize(g_mlps[h]->stream));
        g_mlps[h].reset();
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_nparams(int h) {
    return guarded([&] { return (int)get_mlp(h).nparam; });
}

int gdtk_hip_mlp_get_params(int h, float* out, int n) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (!out || n != (int)m.nparam) throw Error(GDTK_ERR_ARG, "n must equal the parameter count");
        HIPCHK(hipStreamSynchronize(m.stream));
        HIPCHK(hipMemcpy(out, m.params.p, m.nparam * sizeof(float), hipMemcpyDeviceToHost));
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_set_params(int h, const float* in, int n) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (!in || n != (int)m.nparam) throw Error(GDTK_ERR_ARG, "n must equal the parameter count");
        HIPCHK(hipStreamSynchronize(m.stream));
        HIPCHK(hipMemcpy(m.params.p, in, m.nparam * sizeof(float), hipMemcpyHostToDevice));
        m.packed = false;
        m.wq_fresh = false;
        m.adam_t = 0;
        HIPCHK(hipMemset(m.am.p, 0, m.nparam * sizeof(float)));
        HIPCHK(hipMemset(m.av.p, 0, m.nparam * sizeof(float)));
        HIPCHK(hipDeviceSynchronize());
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_set_bf16_training(int h, int on) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        m.bf16_train = on ? 1 : 0;
        m.wq_fresh = false;
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_forward(int h, int batch, const float* x, float* y) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (batch < 1 || !x || !y) throw Error(GDTK_ERR_ARG, "bad arguments");
        m.ensure_batch(batch);
        hipStream_t s = m.stream;
        HIPCHK(hipMemcpyAsync(m.xb.p, x, (size_t)batch * m.n_in * sizeof(float), hipMemcpyHostToDevice, s));
        m.forward(batch, m.xb.p, s);
        HIPCHK(hipMemcpyAsync(y, m.A[m.nlayer - 1].p, (size_t)batch * m.n_out * sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_loss_grad(int h, int batch, const float* x, const float* y, float* loss, float* grad) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (batch < 1 || !x || !y) throw Error(GDTK_ERR_ARG, "bad arguments");
        m.ensure_batch(batch);
        hipStream_t s = m.stream;
        HIPCHK(hipMemcpyAsync(m.xb.p, x, (size_t)batch * m.n_in * sizeof(float), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(m.yb.p, y, (size_t)batch * m.n_out * sizeof(float), hipMemcpyHostToDevice, s));
        m.loss_grad(batch, m.xb.p, m.yb.p, s);
        const double L = m.read_loss(batch, s);
        if (loss) *loss = (float)L;
        if (grad) {
            HIPCHK(hipMemcpyAsync(grad, m.grads.p, m.nparam * sizeof(float), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        return GDTK_OK;
    });
}

int gdtk_hip_mlp_train(int h, int n, const float* x, const float* y, int batch, int epochs, float lr,
                       unsigned long long seed, float* loss_hist, int nhist) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (n < 1 || !x || !y || batch < 1 || epochs < 1 || !(lr > 0.0f)) throw Error(GDTK_ERR_ARG, "bad arguments");
        batch = std::min(batch, n);
        m.ensure_batch(batch);
        hipStream_t s = m.stream;
        m.idx.reserve((size_t)n);
        std::vector<int> perm((size_t)n);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 rng(seed);
        const int nb = n / batch;
        std::vector<double> ep_loss;
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(perm.begin(), perm.end(), rng);
            HIPCHK(hipMemcpyAsync(m.idx.p, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
            double acc = 0.0;
            int nacc = 0;
            for (int b = 0; b < nb; ++b) {
                const int* ib = m.idx.p + (size_t)b * batch;
                HIPCHK(launch_gather_batch(batch, m.n_in, n, ib, x, m.xb.p, s));
                HIPCHK(launch_gather_batch(batch, m.n_out, n, ib, y, m.yb.p, s));
                m.loss_grad(batch, m.xb.p, m.yb.p, s);
                m.adam(lr, s);
                // the loss of every 16th batch is read back (a host sync), the rest run ahead
                if ((b & 15) == 15 || b == nb - 1) {
                    acc += m.read_loss(batch, s);
                    ++nacc;
                }
            }
            const double le = nacc ? acc / nacc : 0.0;
            if (loss_hist && e < nhist) loss_hist[e] = (float)le;
            ep_loss.push_back(le);
        }
        HIPCHK(hipStreamSynchronize(s));
        return epochs;
    });
}

int gdtk_hip_mlp_infer(int h, int n, const float* x, float* y, void* stream) {
    return guarded([&] {
        Mlp& m = get_mlp(h);
        if (n < 0 || (n > 0 && (!x || !y))) throw Error(GDTK_ERR_ARG, "bad arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // the packing runs on the network's stream; order it before the caller's stream
        m.pack(m.stream);
        HIPCHK(hipStreamSynchronize(m.stream));
        mlp_infer(m, n, x, y, s);
        return GDTK_OK;
    });
}

int gdtk_hip_mfma_probe(const float* A, const float* B, float* D) {
    return guarded([&] {
        if (!A || !B || !D) throw Error(GDTK_ERR_ARG, "null argument");
        DevArr<float> a, b, d;
        a.alloc(16 * 32);
        b.alloc(32 * 16);
        d.alloc(16 * 16);
        HIPCHK(hipMemcpy(a.p, A, 16 * 32 * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.p, B, 32 * 16 * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(launch_mfma_probe(a.p, b.p, d.p, nullptr));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(D, d.p, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
        return GDTK_OK;
    });
}

}  // extern "C"
