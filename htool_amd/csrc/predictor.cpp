// Learned ACA rank predictor: features, device state, training loop and the
// C ABI entry points declared in include/htool_amd.h.
#include "predictor.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

namespace htg {

std::string &last_error() {
    static thread_local std::string msg;
    return msg;
}

void block_features(const ClusterNode &t, const ClusterNode &s, double eta, double epsilon, float *f) {
    const double dx = t.ctr[0] - s.ctr[0], dy = t.ctr[1] - s.ctr[1], dz = t.ctr[2] - s.ctr[2];
    const double d  = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double rr = t.rad + s.rad;
    const double tiny = 1e-300;
    f[0] = (float)std::log((double)t.size);
    f[1] = (float)std::log((double)s.size);
    f[2] = (float)std::log(std::max(d, tiny));
    f[3] = (float)std::log(std::max(2 * t.rad, tiny));
    f[4] = (float)std::log(std::max(2 * s.rad, tiny));
    f[5] = (float)eta;
    f[6] = (float)std::log(epsilon);
    f[7] = (float)std::log(std::max(std::max(d - rr, 1e-6 * rr), tiny));
}

int aca_terms_needed(int m, int n, int rank) {
    if (rank > 0)
        return rank;
    return (int)(((long long)m * n) / (m + n));
}

void predictor_predict(const htg_predictor &p, const float *feat, int n, float *out_terms) {
    if (n <= 0)
        return;
    DevArray<float> d_feat, d_out;
    d_feat.upload(feat, (size_t)n * MLP_IN);
    d_out.alloc(n);
    launch_mlp_predict(p.params.p, p.mean.p, p.stdv.p, d_feat.p, n, d_out.p, nullptr);
    HTG_HIP(hipGetLastError());
    HTG_HIP(hipMemcpy(out_terms, d_out.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        out_terms[i] = std::expm1(std::max(out_terms[i], 0.0f));
}

namespace {

// splitmix64 (Steele, Lea, Flood 2014): seeds the initial weights and the shuffles.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

void upload_norm(htg_predictor &p) {
    p.mean.upload(p.h_mean, MLP_IN);
    p.stdv.upload(p.h_std, MLP_IN);
}

struct Scratch {
    DevArray<float> part;
    DevArray<double> loss;
    void alloc(int B) {
        part.alloc((size_t)mlp_tiles(B) * MLP_NPARAMS);
        loss.alloc(B);
    }
};

} // namespace
} // namespace htg

using namespace htg;

extern "C" int htg_predictor_create(unsigned long long seed, htg_predictor **out) {
    if (!out)
        return HTG_ERR_INVALID;
    *out = nullptr;
    return guard([&] {
        auto *p = new htg_predictor();
        try {
            std::vector<float> w(MLP_NPARAMS);
            Rng r{seed};
            auto fill = [&](int off, int count, int fan_in) {
                const double b = 1.0 / std::sqrt((double)fan_in);
                for (int i = 0; i < count; i++)
                    w[off + i] = (float)((2 * r.uniform() - 1) * b);
            };
            fill(MLP_W1, MLP_H * MLP_IN, MLP_IN);
            fill(MLP_B1, MLP_H, MLP_IN);
            fill(MLP_W2, MLP_H * MLP_H, MLP_H);
            fill(MLP_B2, MLP_H, MLP_H);
            fill(MLP_W3, MLP_H, MLP_H);
            fill(MLP_B3, 1, MLP_H);
            p->params.upload(w);
            p->m.alloc(MLP_NPARAMS);
            p->v.alloc(MLP_NPARAMS);
            p->m.zero();
            p->v.zero();
            upload_norm(*p);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

extern "C" int htg_predictor_destroy(htg_predictor *p) {
    delete p;
    return HTG_OK;
}

extern "C" int htg_predictor_set_params(htg_predictor *p, const float *params, const float *mean, const float *stdv) {
    if (!p)
        return HTG_ERR_INVALID;
    return guard([&] {
        if (params)
            HTG_HIP(hipMemcpy(p->params.p, params, MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice));
        if (mean)
            std::memcpy(p->h_mean, mean, sizeof(p->h_mean));
        if (stdv) {
            for (int i = 0; i < MLP_IN; i++)
                if (!(stdv[i] > 0))
                    throw Error(HTG_ERR_INVALID, "feature scales must be positive");
            std::memcpy(p->h_std, stdv, sizeof(p->h_std));
        }
        upload_norm(*p);
        p->m.zero();
        p->v.zero();
        p->step = 0;
    });
}

extern "C" int htg_predictor_get_params(const htg_predictor *p, float *params, float *mean, float *stdv) {
    if (!p)
        return HTG_ERR_INVALID;
    return guard([&] {
        if (params)
            HTG_HIP(hipMemcpy(params, p->params.p, MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
        if (mean)
            std::memcpy(mean, p->h_mean, sizeof(p->h_mean));
        if (stdv)
            std::memcpy(stdv, p->h_std, sizeof(p->h_std));
    });
}

extern "C" int htg_predictor_loss_grad(const htg_predictor *p, const float *feat, const float *target, int n,
                                       double *loss, float *grad) {
    if (!p || !feat || !target || n <= 0)
        return HTG_ERR_INVALID;
    return guard([&] {
        DevArray<float> d_feat, d_target, d_grad;
        d_feat.upload(feat, (size_t)n * MLP_IN);
        d_target.upload(target, n);
        d_grad.alloc(MLP_NPARAMS);
        Scratch w;
        w.alloc(n);
        launch_mlp_grad(p->params.p, p->mean.p, p->stdv.p, d_feat.p, d_target.p, nullptr, n, w.part.p, w.loss.p,
                        d_grad.p, nullptr);
        HTG_HIP(hipGetLastError());
        std::vector<double> l(n);
        HTG_HIP(hipMemcpy(l.data(), w.loss.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        if (loss) {
            double s = 0;
            for (double x : l)
                s += x;
            *loss = s / n;
        }
        if (grad)
            HTG_HIP(hipMemcpy(grad, d_grad.p, MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
    });
}

extern "C" int htg_predictor_train(htg_predictor *p, const float *feat, const float *target, int n, int epochs,
                                   int batch, float lr, unsigned long long seed, int fit_norm, double *epoch_loss) {
    if (!p || !feat || !target || n <= 0 || epochs < 0 || batch <= 0 || !(lr > 0))
        return HTG_ERR_INVALID;
    return guard([&] {
        if (fit_norm) {
            for (int i = 0; i < MLP_IN; i++) {
                double s = 0, s2 = 0;
                for (int b = 0; b < n; b++) {
                    const double x = feat[(size_t)b * MLP_IN + i];
                    s += x;
                    s2 += x * x;
                }
                const double mu = s / n, var = std::max(s2 / n - mu * mu, 0.0);
                p->h_mean[i]    = (float)mu;
                p->h_std[i]     = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
            }
            upload_norm(*p);
        }
        const int B = std::min(batch, n);
        DevArray<float> d_feat, d_target, d_grad;
        DevArray<int> d_idx;
        d_feat.upload(feat, (size_t)n * MLP_IN);
        d_target.upload(target, n);
        d_grad.alloc(MLP_NPARAMS);
        Scratch w;
        w.alloc(B);
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        Rng r{seed};
        d_idx.alloc(n);
        std::vector<double> l(B);
        for (int e = 0; e < epochs; e++) {
            for (int i = n - 1; i > 0; i--) // Fisher-Yates
                std::swap(order[i], order[(size_t)(r.next() % (uint64_t)(i + 1))]);
            d_idx.upload(order);
            double sum = 0;
            for (int b0 = 0; b0 + B <= n; b0 += B) {
                launch_mlp_grad(p->params.p, p->mean.p, p->stdv.p, d_feat.p, d_target.p, d_idx.p + b0, B, w.part.p,
                                w.loss.p, d_grad.p, nullptr);
                launch_adam(p->params.p, p->m.p, p->v.p, d_grad.p, lr, ++p->step, nullptr);
                if (epoch_loss) {
                    HTG_HIP(hipMemcpy(l.data(), w.loss.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost));
                    for (double x : l)
                        sum += x;
                }
            }
            HTG_HIP(hipGetLastError());
            if (epoch_loss)
                epoch_loss[e] = sum / ((n / B) * (double)B);
        }
        HTG_HIP(hipDeviceSynchronize());
    });
}

extern "C" int htg_predictor_predict(const htg_predictor *p, const float *feat, int n, float *out) {
    if (!p || !feat || !out || n < 0)
        return HTG_ERR_INVALID;
    return guard([&] {
        if (n == 0)
            return;
        DevArray<float> d_feat, d_out;
        d_feat.upload(feat, (size_t)n * MLP_IN);
        d_out.alloc(n);
        launch_mlp_predict(p->params.p, p->mean.p, p->stdv.p, d_feat.p, n, d_out.p, nullptr);
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipMemcpy(out, d_out.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    });
}
