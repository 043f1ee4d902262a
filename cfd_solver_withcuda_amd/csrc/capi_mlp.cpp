// Coordinate-MLP surrogate (include/cfd_hip.h, cfd_mlp_*): (x, y, Mach, thickness) -> (rho, u,
// v, p), depth hidden SiLU layers of `width`, BF16 weights and activations on MFMA with FP32
// accumulation, trained with Adam on FP32 master weights. Its prediction initialises the FP64
// pseudo-time march of the plain solver.
//
// Device buffers per layer l (inputs K_l, outputs M_l, both padded to multiples of 32; rows of
// a minibatch padded to a multiple of 512):
//   X_l  bf16 [rows x K_l] and its transpose Xt_l [K_l x rows]  (X_{l+1} = silu(Z_l))
//   Z_l  fp32 [rows x M_l] pre-activations, dZ_l / dZt_l bf16 gradients
//   W_l  fp32 master [M_l x K_l] (unpadded), bf16 copies Wb_l [M_l x K_l] and Wt_l [K_l x M_l]
#include "../../include/cfd_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mlp_kernels.h"

using namespace cfd;

namespace cfd {
int capi_fail(int code, const std::string &msg);
int capi_hip_fail(hipError_t e, const char *where);
}  // namespace cfd

namespace {

#define MHIP(expr)                                                   \
    do {                                                             \
        hipError_t _e = (expr);                                      \
        if (_e != hipSuccess) return cfd::capi_hip_fail(_e, #expr);  \
    } while (0)

int pad32(int n) { return (n + 31) / 32 * 32; }
constexpr int kRowAlign = 512;  // minibatch rows: multiple of 32 * split-K factor
constexpr int kSplit = 16;      // split-K factor of the weight-gradient GEMMs (rows a multiple of 512)
constexpr int kSplitWide = 32;  // ... where the rows are a multiple of 1024 (more workgroups in flight)

}  // namespace

struct cfd_mlp_s {
    int n_in = 0, width = 0, depth = 0, n_out = 0, max_rows = 0;
    int L = 0;                    // weight layers = depth + 1
    std::vector<int> Kv, Mv, Kp, Mp;  // valid / padded inputs and outputs per layer
    std::vector<size_t> woff, boff;   // offsets into the concatenated parameter vectors
    size_t nw = 0, nb = 0;
    float *W = nullptr, *b = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
    float *gW = nullptr, *gb = nullptr;
    std::vector<uint16_t *> Wb, Wt, X, Xt, dZ, dZt;
    std::vector<uint16_t *> Zh;     // bf16 pre-activations of the hidden layers
    float *Zo = nullptr;            // fp32 outputs of the last layer
    float *part = nullptr;          // split-K partials of every layer's weight gradient
    cfd::MlpLayers tab{};           // shapes and buffers of all layers, for the all-layer launches
    float *norm = nullptr;  // in_mean[n_in], in_scale[n_in], out_mean[n_out], out_scale[n_out]
    std::vector<float> norm_h;
    float *feat = nullptr, *targ = nullptr, *lrow = nullptr;
    int *idx = nullptr;
    size_t data_cap = 0;
    int adam_t = 0;
    int graph_mode = 0;             // training steps: 0 launch by launch (faster here), 1 one captured graph per step
    float *sched = nullptr;         // [3 * steps] learning rate and Adam bias corrections per step
    size_t sched_cap = 0;
    int *d_step = nullptr;          // step of the captured graph's next replay
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    ~cfd_mlp_s() {
        if (stream) (void)hipStreamDestroy(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (feat) (void)hipFree(feat);
        if (targ) (void)hipFree(targ);
        if (sched) (void)hipFree(sched);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t n) {
        *p = nullptr;
        hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) {
            allocs.push_back(*p);
            e = hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T));
        }
        return e;
    }
};

namespace {

int pack_all(cfd_mlp_s *m, int *d_step = nullptr) {
    MHIP(mlp_pack_layers(m->W, m->tab, d_step, m->stream));
    return 0;
}

// Forward over `rows` (padded) minibatch rows already encoded in X[0]; leaves the output layer's
// fp32 result in Zo ([rows x Mp[L-1]]) and the hidden layers' bf16 pre-activations in Zh.
int forward(cfd_mlp_s *m, int rows) {
    for (int l = 0; l < m->L; l++) {
        const bool last = (l == m->L - 1);
        MHIP(mlp_gemm(last ? MLP_EPI_BIAS : MLP_EPI_BIAS_SILU, m->X[l], m->Kp[l], m->Wb[l], m->Kp[l], rows,
                      m->Mp[l], m->Kp[l], 1, m->b + m->boff[l], nullptr,
                      last ? static_cast<void *>(m->Zo) : static_cast<void *>(m->Zh[l]),
                      last ? nullptr : m->X[l + 1], last ? nullptr : m->Xt[l + 1], m->Mp[l], rows,
                      m->Mv[l], m->stream));
    }
    return 0;
}

// Backward from dZ[L-1]/dZt[L-1] (set by the loss kernel); weight and bias gradients into gW, gb.
int backward(cfd_mlp_s *m, int rows) {
    const int split = (rows % (32 * kSplitWide) == 0) ? kSplitWide : kSplit;
    for (int l = m->L - 1; l >= 0; l--) {
        // dW_l = dZ_l^T X_l: A = dZt_l [Mp x rows], B = Xt_l [Kp x rows], split over rows
        MHIP(mlp_gemm(MLP_EPI_PARTIAL, m->dZt[l], rows, m->Xt[l], rows, m->Mp[l], m->Kp[l], rows, split,
                      nullptr, nullptr, m->part + m->tab.poff[l], nullptr, nullptr, 0, 0, m->Mp[l], m->stream));
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}): A = dZ_l [rows x Mp], B = Wt_l [Kp x Mp]
            MHIP(mlp_gemm(MLP_EPI_DSILU, m->dZ[l], m->Mp[l], m->Wt[l], m->Mp[l], rows, m->Kp[l], m->Mp[l], 1,
                          nullptr, m->Zh[l - 1], nullptr, m->dZ[l - 1], m->dZt[l - 1], m->Kp[l], rows,
                          m->Mv[l - 1], m->stream));
        }
    }
    // every layer's split-K partials summed into gW, and every layer's bias gradient, in one
    // launch each
    MHIP(mlp_reduce_layers(m->part, split, m->tab, m->gW, m->stream));
    MHIP(mlp_row_sum_layers(m->tab, rows, m->gb, m->stream));
    return 0;
}

int upload_data(cfd_mlp_s *m, const float *feat, const float *targ, int n) {
    const size_t need = (size_t)n;
    if (need > m->data_cap) {
        if (m->feat) MHIP(hipFree(m->feat));
        if (m->targ) MHIP(hipFree(m->targ));
        m->feat = m->targ = nullptr;
        MHIP(hipMalloc((void **)&m->feat, need * m->n_in * sizeof(float)));
        MHIP(hipMalloc((void **)&m->targ, need * m->n_out * sizeof(float)));
        m->data_cap = need;
    }
    MHIP(hipMemcpy(m->feat, feat, need * m->n_in * sizeof(float), hipMemcpyHostToDevice));
    if (targ) MHIP(hipMemcpy(m->targ, targ, need * m->n_out * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

const float *nrm(const cfd_mlp_s *m, int which) {
    const int o[4] = {0, m->n_in, 2 * m->n_in, 2 * m->n_in + m->n_out};
    return m->norm + o[which];
}

}  // namespace

extern "C" {

int cfd_mlp_create(int n_in, int width, int depth, int n_out, int max_rows, unsigned long long seed,
                   cfd_mlp_t **out) {
    if (!out || n_in <= 0 || n_in > 32 || width <= 0 || depth <= 0 || depth >= cfd::kMlpMaxLayers || n_out <= 0 ||
        n_out > 32 || max_rows <= 0)
        return capi_fail(CFD_E_ARG, "cfd_mlp_create: bad arguments");
    *out = nullptr;
    cfd_mlp_s *m = new cfd_mlp_s();
    m->n_in = n_in;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    m->max_rows = (max_rows + kRowAlign - 1) / kRowAlign * kRowAlign;
    m->L = depth + 1;
    for (int l = 0; l < m->L; l++) {
        const int K = (l == 0) ? n_in : width, M = (l == m->L - 1) ? n_out : width;
        m->Kv.push_back(K);
        m->Mv.push_back(M);
        m->Kp.push_back(pad32(K));
        m->Mp.push_back(pad32(M));
        m->woff.push_back(m->nw);
        m->boff.push_back(m->nb);
        m->nw += (size_t)M * K;
        m->nb += M;
    }
    hipError_t e = hipSuccess;
    const int R = m->max_rows;
    int maxMK = 0;
    size_t sumMK = 0;
    for (int l = 0; l < m->L; l++) {
        maxMK = std::max(maxMK, m->Mp[l] * m->Kp[l]);
        m->tab.poff[l] = (long long)(kSplitWide * sumMK);
        sumMK += (size_t)m->Mp[l] * m->Kp[l];
    }
    bool ok = (e = m->alloc(&m->W, m->nw)) == hipSuccess && (e = m->alloc(&m->b, m->nb)) == hipSuccess &&
              (e = m->alloc(&m->mW, m->nw)) == hipSuccess && (e = m->alloc(&m->vW, m->nw)) == hipSuccess &&
              (e = m->alloc(&m->mb, m->nb)) == hipSuccess && (e = m->alloc(&m->vb, m->nb)) == hipSuccess &&
              (e = m->alloc(&m->gW, m->nw)) == hipSuccess && (e = m->alloc(&m->gb, m->nb)) == hipSuccess &&
              (e = m->alloc(&m->part, (size_t)kSplitWide * sumMK)) == hipSuccess &&
              (e = m->alloc(&m->norm, 2 * (size_t)(n_in + n_out))) == hipSuccess &&
              (e = m->alloc(&m->lrow, (size_t)R)) == hipSuccess && (e = m->alloc(&m->idx, (size_t)R)) == hipSuccess &&
              (e = m->alloc(&m->d_step, 1)) == hipSuccess;
    m->Wb.resize(m->L);
    m->Wt.resize(m->L);
    m->X.resize(m->L);
    m->Xt.resize(m->L);
    m->Zh.resize(m->L, nullptr);
    m->dZ.resize(m->L);
    m->dZt.resize(m->L);
    for (int l = 0; ok && l < m->L; l++) {
        const size_t mk = (size_t)m->Mp[l] * m->Kp[l];
        ok = (e = m->alloc(&m->Wb[l], mk)) == hipSuccess && (e = m->alloc(&m->Wt[l], mk)) == hipSuccess &&
             (e = m->alloc(&m->X[l], (size_t)R * m->Kp[l])) == hipSuccess &&
             (e = m->alloc(&m->Xt[l], (size_t)R * m->Kp[l])) == hipSuccess &&
             (e = (l < m->L - 1 ? m->alloc(&m->Zh[l], (size_t)R * m->Mp[l])
                                : m->alloc(&m->Zo, (size_t)R * m->Mp[l]))) == hipSuccess &&
             (e = m->alloc(&m->dZ[l], (size_t)R * m->Mp[l])) == hipSuccess &&
             (e = m->alloc(&m->dZt[l], (size_t)R * m->Mp[l])) == hipSuccess;
    }
    m->tab.L = m->L;
    for (int l = 0; ok && l < m->L; l++) {
        m->tab.M[l] = m->Mv[l];
        m->tab.K[l] = m->Kv[l];
        m->tab.Mp[l] = m->Mp[l];
        m->tab.Kp[l] = m->Kp[l];
        m->tab.woff[l] = (long long)m->woff[l];
        m->tab.boff[l] = (long long)m->boff[l];
        m->tab.Wb[l] = m->Wb[l];
        m->tab.Wt[l] = m->Wt[l];
        m->tab.dZt[l] = m->dZt[l];
    }
    if (ok) e = hipStreamCreate(&m->stream);
    if (!ok || e != hipSuccess) {
        delete m;
        return capi_hip_fail(e, "cfd_mlp_create alloc");
    }
    // He-uniform initial weights for SiLU layers (bound sqrt(6 / fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> Wh(m->nw);
    for (int l = 0; l < m->L; l++) {
        const double bound = std::sqrt(6.0 / m->Kv[l]) * ((l == m->L - 1) ? 0.5 : 1.0);
        std::uniform_real_distribution<double> U(-bound, bound);
        for (size_t i = 0; i < (size_t)m->Mv[l] * m->Kv[l]; i++) Wh[m->woff[l] + i] = (float)U(rng);
    }
    m->norm_h.assign(2 * (n_in + n_out), 0.f);
    for (int k = 0; k < n_in; k++) m->norm_h[n_in + k] = 1.f;
    for (int k = 0; k < n_out; k++) m->norm_h[2 * n_in + n_out + k] = 1.f;
    if ((e = hipMemcpy(m->W, Wh.data(), Wh.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(m->norm, m->norm_h.data(), m->norm_h.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
        delete m;
        return capi_hip_fail(e, "cfd_mlp_create init");
    }
    int rc = pack_all(m);
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

void cfd_mlp_destroy(cfd_mlp_t *m) { delete m; }

int cfd_mlp_num_params(const cfd_mlp_t *m, long long *nw, long long *nb) {
    if (!m) return capi_fail(CFD_E_ARG, "cfd_mlp_num_params: null mlp");
    if (nw) *nw = (long long)m->nw;
    if (nb) *nb = (long long)m->nb;
    return 0;
}

int cfd_mlp_set_params(cfd_mlp_t *m, const float *W, const float *b) {
    if (!m || !W || !b) return capi_fail(CFD_E_ARG, "cfd_mlp_set_params: null argument");
    MHIP(hipMemcpy(m->W, W, m->nw * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemcpy(m->b, b, m->nb * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemset(m->mW, 0, m->nw * sizeof(float)));
    MHIP(hipMemset(m->vW, 0, m->nw * sizeof(float)));
    MHIP(hipMemset(m->mb, 0, m->nb * sizeof(float)));
    MHIP(hipMemset(m->vb, 0, m->nb * sizeof(float)));
    m->adam_t = 0;
    int rc = pack_all(m);
    if (rc) return rc;
    MHIP(hipStreamSynchronize(m->stream));
    return 0;
}

int cfd_mlp_get_params(cfd_mlp_t *m, float *W, float *b) {
    if (!m) return capi_fail(CFD_E_ARG, "cfd_mlp_get_params: null mlp");
    MHIP(hipStreamSynchronize(m->stream));
    if (W) MHIP(hipMemcpy(W, m->W, m->nw * sizeof(float), hipMemcpyDeviceToHost));
    if (b) MHIP(hipMemcpy(b, m->b, m->nb * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int cfd_mlp_set_normalization(cfd_mlp_t *m, const float *in_mean, const float *in_scale,
                              const float *out_mean, const float *out_scale) {
    if (!m || !in_mean || !in_scale || !out_mean || !out_scale)
        return capi_fail(CFD_E_ARG, "cfd_mlp_set_normalization: null argument");
    float *h = m->norm_h.data();
    std::memcpy(h, in_mean, m->n_in * sizeof(float));
    std::memcpy(h + m->n_in, in_scale, m->n_in * sizeof(float));
    std::memcpy(h + 2 * m->n_in, out_mean, m->n_out * sizeof(float));
    std::memcpy(h + 2 * m->n_in + m->n_out, out_scale, m->n_out * sizeof(float));
    MHIP(hipMemcpy(m->norm, h, m->norm_h.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int cfd_mlp_gradients(cfd_mlp_t *m, const float *feat, const float *targ, int n, float *gW,
                      float *gb, double *loss) {
    if (!m || !feat || !targ || n <= 0 || n % kRowAlign || n > m->max_rows)
        return capi_fail(CFD_E_ARG, "cfd_mlp_gradients: n must be a positive multiple of 512 within max_rows");
    int rc = upload_data(m, feat, targ, n);
    if (rc) return rc;
    MHIP(mlp_encode(m->feat, m->n_in, nullptr, n, nrm(m, 0), nrm(m, 1), m->Kp[0], m->X[0], m->Xt[0], n, m->stream));
    rc = forward(m, n);
    if (rc) return rc;
    const int Lm = m->L - 1;
    MHIP(mlp_loss(m->Zo, m->Mp[Lm], m->targ, m->n_out, nullptr, n, nrm(m, 2), nrm(m, 3), m->dZ[Lm], m->dZt[Lm], n,
                  m->lrow, m->stream));
    rc = backward(m, n);
    if (rc) return rc;
    MHIP(hipStreamSynchronize(m->stream));
    if (gW) MHIP(hipMemcpy(gW, m->gW, m->nw * sizeof(float), hipMemcpyDeviceToHost));
    if (gb) MHIP(hipMemcpy(gb, m->gb, m->nb * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) {
        std::vector<float> lr(n);
        MHIP(hipMemcpy(lr.data(), m->lrow, n * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float v : lr) s += v;
        *loss = s / n;
    }
    return 0;
}

int cfd_mlp_train(cfd_mlp_t *m, const float *feat, const float *targ, int n, int steps, int batch,
                  float lr, float lr_final, unsigned long long seed, float *loss_hist, int log_every) {
    if (!m || !feat || !targ || n <= 0 || steps < 0 || batch <= 0 || batch % kRowAlign || batch > m->max_rows ||
        log_every <= 0)
        return capi_fail(CFD_E_ARG, "cfd_mlp_train: bad arguments (batch must be a multiple of 512 within max_rows)");
    int rc = upload_data(m, feat, targ, n);
    if (rc) return rc;
    const int Lm = m->L - 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<float> lrow(batch);
    // cosine decay from lr to lr_final; Adam's bias corrections of its global step count
    std::vector<float> sch(3 * (size_t)std::max(steps, 1));
    for (int s = 0; s < steps; s++) {
        const int t = m->adam_t + s + 1;
        const double f = steps > 1 ? 0.5 * (1.0 + std::cos(M_PI * s / (double)(steps - 1))) : 1.0;
        sch[3 * (size_t)s] = (float)(lr_final + (lr - lr_final) * f);
        sch[3 * (size_t)s + 1] = 1.0f / (1.0f - std::pow(b1, (float)t));
        sch[3 * (size_t)s + 2] = 1.0f / (1.0f - std::pow(b2, (float)t));
    }
    auto log_loss = [&](int s) -> int {
        if (loss_hist && s % log_every == 0) {
            MHIP(hipMemcpyAsync(lrow.data(), m->lrow, batch * sizeof(float), hipMemcpyDeviceToHost, m->stream));
            MHIP(hipStreamSynchronize(m->stream));
            double acc = 0.0;
            for (float v : lrow) acc += v;
            loss_hist[s / log_every] = (float)(acc / batch);
        }
        return 0;
    };
    if (m->graph_mode && steps > 0) {
        // one training step captured once and replayed: the step index, learning rate and bias
        // corrections come from device memory, so every replay is the same launch sequence
        if ((size_t)steps > m->sched_cap) {
            if (m->sched) MHIP(hipFree(m->sched));
            m->sched = nullptr;
            MHIP(hipMalloc((void **)&m->sched, 3 * (size_t)steps * sizeof(float)));
            m->sched_cap = steps;
        }
        MHIP(hipMemcpy(m->sched, sch.data(), 3 * (size_t)steps * sizeof(float), hipMemcpyHostToDevice));
        const int zero = 0;
        MHIP(hipMemcpy(m->d_step, &zero, sizeof(int), hipMemcpyHostToDevice));
        auto launches = [&]() -> int {
            MHIP(mlp_sample_dev(seed, m->d_step, batch, n, m->idx, m->stream));
            MHIP(mlp_encode(m->feat, m->n_in, m->idx, batch, nrm(m, 0), nrm(m, 1), m->Kp[0], m->X[0], m->Xt[0],
                            batch, m->stream));
            if (int r = forward(m, batch)) return r;
            MHIP(mlp_loss(m->Zo, m->Mp[Lm], m->targ, m->n_out, m->idx, batch, nrm(m, 2), nrm(m, 3), m->dZ[Lm],
                          m->dZt[Lm], batch, m->lrow, m->stream));
            if (int r = backward(m, batch)) return r;
            MHIP(mlp_adam2_dev(m->W, m->mW, m->vW, m->gW, (int)m->nw, m->b, m->mb, m->vb, m->gb, (int)m->nb,
                               m->sched, m->d_step, b1, b2, eps, m->stream));
            // the repacking launch also advances the step index read by the next replay
            return pack_all(m, m->d_step);
        };
        MHIP(hipStreamSynchronize(m->stream));
        MHIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        rc = launches();
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(m->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return capi_hip_fail(ec, "cfd_mlp_train capture");
        hipGraphExec_t ge = nullptr;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return capi_hip_fail(ei, "cfd_mlp_train instantiate");
        for (int s = 0; s < steps && !rc; s++) {
            hipError_t el = hipGraphLaunch(ge, m->stream);
            if (el != hipSuccess) rc = capi_hip_fail(el, "cfd_mlp_train graph launch");
            else rc = log_loss(s);
        }
        (void)hipStreamSynchronize(m->stream);
        (void)hipGraphExecDestroy(ge);
        if (rc) return rc;
        m->adam_t += steps;
        return 0;
    }
    for (int s = 0; s < steps; s++) {
        MHIP(mlp_sample(seed, s, batch, n, m->idx, m->stream));
        MHIP(mlp_encode(m->feat, m->n_in, m->idx, batch, nrm(m, 0), nrm(m, 1), m->Kp[0], m->X[0], m->Xt[0], batch,
                        m->stream));
        rc = forward(m, batch);
        if (rc) return rc;
        MHIP(mlp_loss(m->Zo, m->Mp[Lm], m->targ, m->n_out, m->idx, batch, nrm(m, 2), nrm(m, 3), m->dZ[Lm],
                      m->dZt[Lm], batch, m->lrow, m->stream));
        rc = backward(m, batch);
        if (rc) return rc;
        m->adam_t++;
        const float lrs = sch[3 * (size_t)s], c1 = sch[3 * (size_t)s + 1], c2 = sch[3 * (size_t)s + 2];
        MHIP(mlp_adam2(m->W, m->mW, m->vW, m->gW, (int)m->nw, m->b, m->mb, m->vb, m->gb, (int)m->nb, lrs, b1, b2,
                       eps, c1, c2, m->stream));
        rc = pack_all(m);
        if (rc) return rc;
        if ((rc = log_loss(s))) return rc;
    }
    MHIP(hipStreamSynchronize(m->stream));
    return 0;
}

int cfd_mlp_set_launch(cfd_mlp_t *m, int mode) {
    if (!m || (mode != 0 && mode != 1)) return capi_fail(CFD_E_ARG, "cfd_mlp_set_launch: mode must be 0 or 1");
    m->graph_mode = mode;
    return 0;
}

int cfd_mlp_predict(cfd_mlp_t *m, const float *feat, int n, float *out) {
    if (!m || !feat || !out || n <= 0) return capi_fail(CFD_E_ARG, "cfd_mlp_predict: bad arguments");
    const int Lm = m->L - 1, C = m->Mp[Lm];
    std::vector<float> z((size_t)m->max_rows * C);
    const float *om = m->norm_h.data() + 2 * m->n_in, *os = om + m->n_out;
    for (int r0 = 0; r0 < n; r0 += m->max_rows) {
        const int cnt = std::min(m->max_rows, n - r0);
        const int rows = (cnt + kRowAlign - 1) / kRowAlign * kRowAlign;
        std::vector<float> f((size_t)rows * m->n_in, 0.f);
        std::memcpy(f.data(), feat + (size_t)r0 * m->n_in, (size_t)cnt * m->n_in * sizeof(float));
        int rc = upload_data(m, f.data(), nullptr, rows);
        if (rc) return rc;
        MHIP(mlp_encode(m->feat, m->n_in, nullptr, rows, nrm(m, 0), nrm(m, 1), m->Kp[0], m->X[0], m->Xt[0], rows,
                        m->stream));
        rc = forward(m, rows);
        if (rc) return rc;
        MHIP(hipMemcpyAsync(z.data(), m->Zo, (size_t)rows * C * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP(hipStreamSynchronize(m->stream));
        for (int r = 0; r < cnt; r++)
            for (int k = 0; k < m->n_out; k++)
                out[(size_t)(r0 + r) * m->n_out + k] = z[(size_t)r * C + k] / os[k] + om[k];
    }
    return 0;
}

int cfd_mlp_time_step(cfd_mlp_t *m, int batch, int reps, float *ms) {
    if (!m || !ms || batch <= 0 || batch % kRowAlign || batch > m->max_rows || reps <= 0 || m->data_cap == 0)
        return capi_fail(CFD_E_ARG, "cfd_mlp_time_step: bad arguments (train once first)");
    const int Lm = m->L - 1;
    hipEvent_t e0, e1;
    MHIP(hipEventCreate(&e0));
    MHIP(hipEventCreate(&e1));
    MHIP(hipEventRecord(e0, m->stream));
    for (int s = 0; s < reps; s++) {
        MHIP(mlp_sample(7, s, batch, (int)m->data_cap, m->idx, m->stream));
        MHIP(mlp_encode(m->feat, m->n_in, m->idx, batch, nrm(m, 0), nrm(m, 1), m->Kp[0], m->X[0], m->Xt[0], batch,
                        m->stream));
        int rc = forward(m, batch);
        if (rc) return rc;
        MHIP(mlp_loss(m->Zo, m->Mp[Lm], m->targ, m->n_out, m->idx, batch, nrm(m, 2), nrm(m, 3), m->dZ[Lm],
                      m->dZt[Lm], batch, m->lrow, m->stream));
        rc = backward(m, batch);
        if (rc) return rc;
    }
    MHIP(hipEventRecord(e1, m->stream));
    MHIP(hipEventSynchronize(e1));
    MHIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // extern "C"
