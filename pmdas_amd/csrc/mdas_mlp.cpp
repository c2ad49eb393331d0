// C ABI of the learned proposal: the MLP handle (initialisation, parameter
// transfer, feature normalisation, inference, gradient, Adam training) and
// the per-grid learned mode (row collection, MLP scores in the error pass).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "mdas.h"
#include "mdas_handles.hpp"
#include "mlp_device.hpp"

struct mdas_mlp {
    float* p = nullptr;             // [kMlpFloats] FP32 parameters + normalisation
    unsigned short* w1p = nullptr;  // [64*32]
    unsigned short* w2p = nullptr;  // [64*64]
    float* w2r = nullptr;           // [64*64]
    float* w3r = nullptr;           // [64]
    float* m = nullptr;             // Adam moments [kMlpParams]
    float* v = nullptr;
    float* gpart = nullptr;         // [kMlpGradBlocks*4][kMlpParams] wave partials
    float* lpart = nullptr;         // [kMlpGradBlocks*4]
    float* gsum = nullptr;          // [kMlpParams + 1] summed gradient and loss
    double* spart = nullptr;        // feature statistics partials
    long long t = 0;                // Adam steps taken
    mdas::MlpDev dev() const { return mdas::MlpDev{p, w1p, w2p, w2r, w3r}; }
};

namespace {

using mdas::set_error;
const int kStatBlocks = 1024;

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(MDAS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void free_mlp(mdas_mlp* m) {
    void* ptrs[] = {m->p, m->w1p, m->w2p, m->w2r, m->w3r, m->m, m->v, m->gpart, m->lpart, m->gsum,
                    m->spart};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
}

// Default initialisation: weights and biases U(-1/sqrt(fan_in), 1/sqrt(fan_in))
// from a seeded LCG (the common default for linear layers), mu = 0,
// 1/sigma = 1.
void init_params(uint32_t seed, std::vector<float>& h) {
    h.assign(mdas::kMlpFloats, 0.0f);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
    auto u01 = [&s]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int n, double bound) {
        for (int k = 0; k < n; ++k) h[off + k] = (float)((2.0 * u01() - 1.0) * bound);
    };
    fill(mdas::kOffW1, 64 * 16, 1.0 / std::sqrt(16.0));
    fill(mdas::kOffB1, 64, 1.0 / std::sqrt(16.0));
    fill(mdas::kOffW2, 64 * 64, 1.0 / 8.0);
    fill(mdas::kOffB2, 64, 1.0 / 8.0);
    fill(mdas::kOffW3, 64, 1.0 / 8.0);
    fill(mdas::kOffB3, 1, 1.0 / 8.0);
    for (int k = 0; k < 16; ++k) h[mdas::kOffIsd + k] = 1.0f;
}

int upload(mdas_mlp* m, const float* host) {
    HIP_OK(hipMemcpy(m->p, host, sizeof(float) * mdas::kMlpFloats, hipMemcpyHostToDevice));
    HIP_OK(mdas::launch_mlp_pack(m->p, m->w1p, m->w2p, m->w2r, m->w3r, 0));
    HIP_OK(hipMemset(m->m, 0, sizeof(float) * mdas::kMlpParams));
    HIP_OK(hipMemset(m->v, 0, sizeof(float) * mdas::kMlpParams));
    HIP_OK(hipDeviceSynchronize());
    m->t = 0;
    return MDAS_OK;
}

// Multipliers for the per-epoch row permutation q -> (a q + c) mod n.
const unsigned long long kPrimes[] = {1000003ull,  2000029ull,  3000017ull,  4000037ull,
                                      5000011ull,  6000047ull,  7000003ull,  8000009ull,
                                      9000011ull,  10000019ull, 11000027ull, 12000017ull,
                                      13000027ull, 14000029ull, 15000017ull, 16000057ull};

}  // namespace

extern "C" {

int mdas_mlp_create(uint32_t seed, mdas_mlp_handle* out) {
    if (!out) return set_error(MDAS_ERR_INVALID, "null output handle");
    *out = nullptr;
    mdas_mlp* m = new mdas_mlp;
    const size_t waves = (size_t)mdas::kMlpGradBlocks * 4;
    hipError_t e = hipMalloc((void**)&m->p, sizeof(float) * mdas::kMlpFloats);
    if (e == hipSuccess) e = hipMalloc((void**)&m->w1p, sizeof(unsigned short) * 64 * 32);
    if (e == hipSuccess) e = hipMalloc((void**)&m->w2p, sizeof(unsigned short) * 64 * 64);
    if (e == hipSuccess) e = hipMalloc((void**)&m->w2r, sizeof(float) * 64 * 64);
    if (e == hipSuccess) e = hipMalloc((void**)&m->w3r, sizeof(float) * 64);
    if (e == hipSuccess) e = hipMalloc((void**)&m->m, sizeof(float) * mdas::kMlpParams);
    if (e == hipSuccess) e = hipMalloc((void**)&m->v, sizeof(float) * mdas::kMlpParams);
    if (e == hipSuccess) e = hipMalloc((void**)&m->gpart, sizeof(float) * waves * mdas::kMlpParams);
    if (e == hipSuccess) e = hipMalloc((void**)&m->lpart, sizeof(float) * waves);
    if (e == hipSuccess) e = hipMalloc((void**)&m->gsum, sizeof(float) * (mdas::kMlpParams + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&m->spart, sizeof(double) * kStatBlocks * 32);
    if (e != hipSuccess) {
        free_mlp(m);
        delete m;
        return set_error(MDAS_ERR_HIP, std::string("device allocation failed: ") + hipGetErrorString(e));
    }
    std::vector<float> h;
    init_params(seed, h);
    int rc = upload(m, h.data());
    if (rc) {
        free_mlp(m);
        delete m;
        return rc;
    }
    *out = m;
    return MDAS_OK;
}

int mdas_mlp_destroy(mdas_mlp_handle m) {
    if (!m) return MDAS_OK;
    (void)hipDeviceSynchronize();
    free_mlp(m);
    delete m;
    return MDAS_OK;
}

int mdas_mlp_set_params(mdas_mlp_handle m, const float* host, int n) {
    if (!m || !host) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n != mdas::kMlpFloats) return set_error(MDAS_ERR_INVALID, "parameter count must be MDAS_MLP_FLOATS");
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(host[k])) return set_error(MDAS_ERR_INVALID, "parameters must be finite");
    return upload(m, host);
}

int mdas_mlp_get_params(mdas_mlp_handle m, float* host, int n) {
    if (!m || !host) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n != mdas::kMlpFloats) return set_error(MDAS_ERR_INVALID, "parameter count must be MDAS_MLP_FLOATS");
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(host, m->p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MDAS_OK;
}

int mdas_mlp_fit_norm(mdas_mlp_handle m, const float* rows, long long n, void* stream) {
    if (!m || !rows) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n < 2) return set_error(MDAS_ERR_INVALID, "need at least 2 rows");
    hipStream_t st = (hipStream_t)stream;
    HIP_OK(mdas::launch_feat_stats(rows, n, m->spart, kStatBlocks, st));
    std::vector<double> part((size_t)kStatBlocks * 32);
    HIP_OK(hipMemcpyAsync(part.data(), m->spart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    std::vector<float> h(mdas::kMlpFloats);
    HIP_OK(hipMemcpy(h.data(), m->p, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    for (int j = 0; j < 16; ++j) {
        double s = 0.0, s2 = 0.0;
        for (int b = 0; b < kStatBlocks; ++b) {
            s += part[(size_t)b * 32 + j];
            s2 += part[(size_t)b * 32 + 16 + j];
        }
        const double mu = s / (double)n;
        const double var = std::max(0.0, s2 / (double)n - mu * mu);
        const double sd = std::sqrt(var);
        h[mdas::kOffMu + j] = (float)mu;
        h[mdas::kOffIsd + j] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;
    }
    // keep the Adam state: only the input normalisation changes
    HIP_OK(hipMemcpy(m->p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    return MDAS_OK;
}

int mdas_mlp_forward(mdas_mlp_handle m, const float* rows, long long n, float* out, void* stream) {
    if (!m || !rows || !out) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n < 0) return set_error(MDAS_ERR_INVALID, "negative row count");
    if (n == 0) return MDAS_OK;
    HIP_OK(mdas::launch_mlp_forward(m->dev(), rows, n, out, (hipStream_t)stream));
    return MDAS_OK;
}

int mdas_mlp_grad(mdas_mlp_handle m, const float* rows, const float* targets, long long n,
                  float* grad_host, float* loss_host, void* stream) {
    if (!m || !rows || !targets || !grad_host) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n < 1 || n > (1LL << 30)) return set_error(MDAS_ERR_INVALID, "row count out of range");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    const int parts = mdas::launch_mlp_grad(m->dev(), rows, targets, n, 0, (int)n, 1, 0, m->gpart,
                                            m->lpart, st, &e);
    HIP_OK(e);
    HIP_OK(mdas::launch_mlp_sum(m->gpart, m->lpart, parts, m->gsum, m->gsum + mdas::kMlpParams, st));
    std::vector<float> g(mdas::kMlpParams + 1);
    HIP_OK(hipMemcpyAsync(g.data(), m->gsum, sizeof(float) * g.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    memcpy(grad_host, g.data(), sizeof(float) * mdas::kMlpParams);
    if (loss_host) *loss_host = g[mdas::kMlpParams] / (float)n;
    return MDAS_OK;
}

int mdas_mlp_train(mdas_mlp_handle m, const float* rows, const float* targets, long long n,
                   int epochs, int batch, float lr, uint32_t seed, float* epoch_loss_host,
                   void* stream) {
    if (!m || !rows || !targets) return set_error(MDAS_ERR_INVALID, "null argument");
    if (n < 1 || n > (1LL << 40)) return set_error(MDAS_ERR_INVALID, "row count out of range");
    if (epochs < 1 || batch < 1 || batch > (1 << 26) || !(lr > 0.0f))
        return set_error(MDAS_ERR_INVALID, "epochs, batch and lr must be positive");
    hipStream_t st = (hipStream_t)stream;
    float* loss_acc = nullptr;
    HIP_OK(hipMalloc((void**)&loss_acc, sizeof(float) * epochs));
    int rc = MDAS_OK;
    hipError_t e = hipMemsetAsync(loss_acc, 0, sizeof(float) * epochs, st);
    const unsigned long long un = (unsigned long long)n;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        // a multiplier coprime to n, and an offset, per epoch
        unsigned long long pa = 1, pc = 0;
        if (n > 1) {
            const unsigned h = (seed * 2654435761u) ^ (unsigned)(ep * 40503u);
            for (int k = 0; k < 16; ++k) {
                const unsigned long long cand = kPrimes[(h + k) & 15];
                if (un % cand != 0ull) {
                    pa = cand % un;
                    if (pa == 0ull) pa = 1;
                    break;
                }
            }
            pc = ((unsigned long long)h * 2246822519ull) % un;
        }
        for (long long s0 = 0; s0 < n && e == hipSuccess; s0 += batch) {
            const int cnt = (int)std::min<long long>(batch, n - s0);
            const int parts = mdas::launch_mlp_grad(m->dev(), rows, targets, n, s0, cnt, pa, pc,
                                                    m->gpart, m->lpart, st, &e);
            if (e != hipSuccess) break;
            ++m->t;
            const float bc1 = (float)(1.0 - std::pow(0.9, (double)m->t));
            const float bc2 = (float)(1.0 - std::pow(0.999, (double)m->t));
            e = mdas::launch_mlp_adam(m->p, m->m, m->v, m->gpart, m->lpart, parts, lr, bc1, bc2,
                                      loss_acc + ep, m->w1p, m->w2p, m->w2r, m->w3r, st);
        }
    }
    if (e == hipSuccess && epoch_loss_host) {
        std::vector<float> l(epochs);
        e = hipMemcpyAsync(l.data(), loss_acc, sizeof(float) * epochs, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        for (int k = 0; k < epochs; ++k) epoch_loss_host[k] = l[k] / (float)n;
    } else if (e == hipSuccess) {
        e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) rc = set_error(MDAS_ERR_HIP, std::string("training: ") + hipGetErrorString(e));
    (void)hipFree(loss_acc);
    return rc;
}

int mdas_set_learned(mdas_handle h, int mode, mdas_mlp_handle m, float* rows, float* labels,
                     long long cap_rows) {
    if (!h) return set_error(MDAS_ERR_INVALID, "null handle");
    if (mode & ~(MDAS_LEARN_COLLECT | MDAS_LEARN_SCORE | MDAS_LEARN_IMAGE_LABEL | MDAS_LEARN_KEEP_FEATURES))
        return set_error(MDAS_ERR_INVALID, "unknown learned mode bits");
    if ((mode & MDAS_LEARN_SCORE) && !m)
        return set_error(MDAS_ERR_INVALID, "scoring needs a network");
    if ((mode & MDAS_LEARN_COLLECT) && (!rows || !labels || cap_rows < 0))
        return set_error(MDAS_ERR_INVALID, "collection needs row and label buffers");
    mdas::Grid& g = h->g;
    HIP_OK(hipDeviceSynchronize());
    if (mode != 0 && g.learn.feat == nullptr) {
        HIP_OK(hipMalloc((void**)&g.learn.feat, sizeof(float) * mdas::kFeat * (size_t)g.cap_l));
        HIP_OK(hipMalloc((void**)&g.learn.perr, sizeof(float) * (size_t)g.cap_l));
        HIP_OK(hipMalloc((void**)&g.learn.rows, sizeof(long long)));
    }
    if (mode != 0 && g.dev.sext == nullptr)
        HIP_OK(hipMalloc((void**)&g.dev.sext, sizeof(int4) * 2 * ((size_t)g.cap_l / 2 + 1)));
    g.learn_mode = mode;
    g.learn.X = (mode & MDAS_LEARN_COLLECT) ? rows : nullptr;
    g.learn.Y = (mode & MDAS_LEARN_COLLECT) ? labels : nullptr;
    g.learn.cap_rows = (mode & MDAS_LEARN_COLLECT) ? cap_rows : 0;
    g.learn.mlp = (mode & MDAS_LEARN_SCORE) ? m->dev() : mdas::MlpDev{};
    g.learn.label_mode = (mode & MDAS_LEARN_IMAGE_LABEL) ? 1 : 0;
    g.learn.keep = (mode & (MDAS_LEARN_COLLECT | MDAS_LEARN_KEEP_FEATURES)) ? 1 : 0;
    g.learn.kappa_e = g.learn.label_mode ? 0.0f : 1.0f;
    g.learn.kappa_y = 1.0f;
    g.learn.ylo = g.learn.label_mode ? -120.0f : -10.0f;
    g.learn.yhi = g.learn.label_mode ? 120.0f : 0.0f;
    if (g.learn.rows) HIP_OK(hipMemset(g.learn.rows, 0, sizeof(long long)));
    return MDAS_OK;
}

int mdas_set_learned_exponent(mdas_handle h, float kappa) {
    if (!h) return set_error(MDAS_ERR_INVALID, "null handle");
    if (!(kappa > 0.0f) || kappa > 16.0f) return set_error(MDAS_ERR_INVALID, "kappa must be in (0, 16]");
    h->g.learn.kappa_e = h->g.learn.label_mode ? 0.0f : kappa;
    h->g.learn.kappa_y = kappa;
    return MDAS_OK;
}

int mdas_set_learned_exponents(mdas_handle h, float kappa_e, float kappa_y) {
    if (!h) return set_error(MDAS_ERR_INVALID, "null handle");
    if (!(kappa_e >= 0.0f) || kappa_e > 16.0f || !(kappa_y >= 0.0f) || kappa_y > 16.0f ||
        kappa_e + kappa_y == 0.0f)
        return set_error(MDAS_ERR_INVALID, "exponents must be in [0, 16], not both 0");
    h->g.learn.kappa_e = kappa_e;
    h->g.learn.kappa_y = kappa_y;
    return MDAS_OK;
}

int mdas_get_learned_rows(mdas_handle h, long long* offered) {
    if (!h || !offered) return set_error(MDAS_ERR_INVALID, "null argument");
    *offered = 0;
    if (!h->g.learn.rows) return MDAS_OK;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(offered, h->g.learn.rows, sizeof(long long), hipMemcpyDeviceToHost));
    return MDAS_OK;
}

int mdas_get_leaf_features(mdas_handle h, float* feat_host, float* perr_host, long long n) {
    if (!h) return set_error(MDAS_ERR_INVALID, "null handle");
    const mdas::Grid& g = h->g;
    if (!g.learn.feat) return set_error(MDAS_ERR_STATE, "learned mode was never enabled");
    if (!g.learn.keep)
        return set_error(MDAS_ERR_STATE, "features are kept with MDAS_LEARN_COLLECT or MDAS_LEARN_KEEP_FEATURES");
    if (n < 0 || n > g.cap_l) return set_error(MDAS_ERR_INVALID, "leaf count out of range");
    HIP_OK(hipDeviceSynchronize());
    if (feat_host)
        HIP_OK(hipMemcpy(feat_host, g.learn.feat, sizeof(float) * mdas::kFeat * n, hipMemcpyDeviceToHost));
    if (perr_host) HIP_OK(hipMemcpy(perr_host, g.learn.perr, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MDAS_OK;
}

}  // extern "C"
