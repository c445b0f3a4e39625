// md_train_api.hip -- host side of the surrogate trainer (C ABI in
// include/moldyn_hip.h, md_trainer_*): a data set of frames resident in HBM,
// full-data-set loss and gradient, and Adam epochs over minibatches of whole
// frames, all on the device with the kernels of md_mlp_train.hip.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "md_train.h"
#include "moldyn_hip.h"

using namespace md;

struct md_trainer {
    hipStream_t st = nullptr;
    TrainState t{};
    std::vector<void*> allocs;
    int64_t batch_rows = 0;  // capacity of the minibatch arrays (atoms)
    int step = 0;            // Adam steps taken
    double* loss_part = nullptr;
    int32_t* frames_seq = nullptr;  // 0, 1, ..., n_frames - 1
    int32_t* frames_ord = nullptr;  // the epoch's frame order
    double* batch_loss = nullptr;   // per minibatch: summed squared energy and force errors
    std::string err;
};

static thread_local std::string g_train_error;

static int tfail(md_trainer* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_train_error = msg;
    return code;
}

#define TCHK(c, expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return tfail((c), MD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int talloc(md_trainer* c, void** p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        *p = nullptr;
        return tfail(c, MD_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + ")");
    }
    c->allocs.push_back(*p);
    return MD_OK;
}

template <typename T>
static int talloc_t(md_trainer* c, T** p, size_t n) {
    void* v;
    int rc = talloc(c, &v, sizeof(T) * n);
    *p = (T*)v;
    return rc;
}

// Minibatch work arrays for up to `rows` atoms.
static int ensure_batch(md_trainer* c, int64_t rows) {
    if (rows <= c->batch_rows) return MD_OK;
    TrainState& t = c->t;
    const size_t r = (size_t)rows;
    int rc;
    float** m32[] = {&t.X, &t.Gam, &t.U};
    for (float** p : m32)
        if ((rc = talloc_t(c, p, r * kMlpK))) return rc;
    float** m128[] = {&t.H1, &t.S1, &t.S1p, &t.H2, &t.S2, &t.S2p, &t.D2, &t.T, &t.D1, &t.V1, &t.PP, &t.V2, &t.R1, &t.W3T};
    for (float** p : m128)
        if ((rc = talloc_t(c, p, r * kMlpH))) return rc;
    if ((rc = talloc_t(c, &t.E, r))) return rc;
    if ((rc = talloc_t(c, &t.dy, r))) return rc;
    if ((rc = talloc_t(c, &t.F, r * 3))) return rc;
    if ((rc = talloc_t(c, &t.phi, r * 3))) return rc;
    const int64_t blocks = (int64_t)((c->t.n_atoms + 255) / 256) * (rows / c->t.n_atoms + 1);
    if ((rc = talloc_t(c, &c->loss_part, (size_t)(2 * blocks + 2)))) return rc;
    c->batch_rows = rows;
    return MD_OK;
}

// Forward, forces and (optionally) gradient of the nf frames frames[0..nf)
// (device array); the summed squared energy and force errors go to
// out[0..1] (device).  Nothing waits for the device.
static int run_batch(md_trainer* c, const int32_t* frames, int nf, bool grad, double* out) {
    TrainState& t = c->t;
    const int64_t M = (int64_t)nf * t.n_atoms;
    int rc;
    if ((rc = ensure_batch(c, M))) return rc;
    t.bf = frames;
    TCHK(c, tr_forward_backward(t, nf, c->st));
    TCHK(c, tr_forces(t, nf, c->loss_part, c->st));
    TCHK(c, tr_loss_sum(t, nf, c->loss_part, out, c->st));
    if (grad) TCHK(c, tr_gradient(t, nf, c->st));
    return MD_OK;
}

// per-minibatch loss sums of the last run, copied to the host
static int read_batch_loss(md_trainer* c, int nb, std::vector<double>& h) {
    h.assign((size_t)2 * nb, 0.0);
    TCHK(c, hipMemcpyAsync(h.data(), c->batch_loss, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
    TCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

extern "C" {

const char* md_trainer_last_error(const md_trainer* c) { return c ? c->err.c_str() : g_train_error.c_str(); }

void md_trainer_destroy(md_trainer* c) {
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int md_trainer_create(md_trainer** out, int n_frames, int n_atoms, const double* pos, const double* boxes,
                      const double* e_ref, const double* f_ref, double sigma) {
    if (!out || n_frames < 1 || n_atoms < 2 || !pos || !boxes || !e_ref || !f_ref || !(sigma > 0))
        return tfail(nullptr, MD_ERR_ARG, "bad argument");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return tfail(nullptr, MD_ERR_NO_DEVICE, "no HIP device");
    md_trainer* c = new md_trainer();
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return tfail(nullptr, MD_ERR_HIP, "hipStreamCreate failed");
    }
    TrainState& t = c->t;
    t.n_atoms = n_atoms;
    t.n_frames = n_frames;
    const size_t na = (size_t)n_frames * n_atoms;
    const double w = 0.1 * sigma;
    for (int k = 0; k < kMlpK; ++k) {
        t.mu[k] = (float)(sigma * (0.8 + 1.7 * k / (kMlpK - 1)));
        t.g_mean[k] = 0.0f;
        t.g_inv_std[k] = 1.0f;
    }
    t.rc = (float)(2.5 * sigma);
    t.inv_2w2 = (float)(1.0 / (2.0 * w * w));
    t.inv_w2 = (float)(1.0 / (w * w));
    t.e_scale = 1.0f;
    t.splits = 128;
    t.colsum_blocks = 256;
    int rc;
    double *dpos, *dbox;
    float *de, *df;
    std::vector<float> he(na), hf(na * 3);
    for (size_t i = 0; i < na; ++i) he[i] = (float)e_ref[i];
    for (size_t i = 0; i < na * 3; ++i) hf[i] = (float)f_ref[i];
    if ((rc = talloc_t(c, &dpos, na * 3)) || (rc = talloc_t(c, &dbox, (size_t)n_frames * 3)) ||
        (rc = talloc_t(c, &de, na)) || (rc = talloc_t(c, &df, na * 3)) || (rc = talloc_t(c, &t.G, na * kMlpK)) ||
        (rc = talloc_t(c, &t.theta, kMlpNP)) || (rc = talloc_t(c, &t.grad, kMlpNP)) ||
        (rc = talloc_t(c, &t.adam_m, kMlpNP)) || (rc = talloc_t(c, &t.adam_v, kMlpNP)) ||
        (rc = talloc_t(c, &t.gpart, (size_t)std::max(t.splits, t.colsum_blocks) * kMlpH * kMlpH)) ||
        (rc = talloc_t(c, &t.ncnt, na)) || (rc = talloc_t(c, &c->frames_seq, (size_t)n_frames)) ||
        (rc = talloc_t(c, &c->frames_ord, (size_t)n_frames)) ||
        (rc = talloc_t(c, &c->batch_loss, (size_t)2 * n_frames + 2))) {
        g_train_error = c->err;
        md_trainer_destroy(c);
        return rc;
    }
    t.pos = dpos;
    t.box = dbox;
    t.Eref = de;
    t.Fref = df;
    hipError_t e = hipMemcpy(dpos, pos, sizeof(double) * na * 3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dbox, boxes, sizeof(double) * n_frames * 3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(de, he.data(), sizeof(float) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, hf.data(), sizeof(float) * na * 3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(t.adam_m, 0, sizeof(float) * kMlpNP);
    if (e == hipSuccess) e = hipMemset(t.adam_v, 0, sizeof(float) * kMlpNP);
    std::vector<int32_t> seq((size_t)n_frames);
    for (int f = 0; f < n_frames; ++f) seq[(size_t)f] = f;
    if (e == hipSuccess) e = hipMemcpy(c->frames_seq, seq.data(), sizeof(int32_t) * n_frames, hipMemcpyHostToDevice);
    // pair lists of every frame: counts and their maximum, then the indices
    int32_t* dmax = nullptr;
    int32_t cap = 0;
    if ((rc = talloc_t(c, &dmax, 1))) {
        g_train_error = c->err;
        md_trainer_destroy(c);
        return rc;
    }
    if (e == hipSuccess) e = hipMemset(dmax, 0, sizeof(int32_t));
    for (int f0 = 0; f0 < n_frames && e == hipSuccess; f0 += 64)
        e = tr_count_pairs(t, f0, std::min(64, n_frames - f0), dmax, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(&cap, dmax, sizeof(int32_t), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) {
        t.cap = std::max(cap, 1);
        if ((rc = talloc_t(c, &t.nbr, na * (size_t)t.cap))) {
            g_train_error = c->err;
            md_trainer_destroy(c);
            return rc;
        }
    }
    for (int f0 = 0; f0 < n_frames && e == hipSuccess; f0 += 64)
        e = tr_fill_pairs(t, f0, std::min(64, n_frames - f0), c->st);
    // features of every frame (they do not depend on the parameters)
    for (int f0 = 0; f0 < n_frames && e == hipSuccess; f0 += 64) e = tr_features(t, f0, std::min(64, n_frames - f0), c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        g_train_error = std::string("trainer set-up: ") + hipGetErrorString(e);
        md_trainer_destroy(c);
        return MD_ERR_HIP;
    }
    *out = c;
    return MD_OK;
}

int md_trainer_feature_stats(md_trainer* c, double out[128]) {
    if (!c || !out) return tfail(c, MD_ERR_ARG, "bad argument");
    const int blocks = 256;
    double* part;
    int rc;
    if ((rc = talloc_t(c, &part, (size_t)blocks * 4 * kMlpK))) return rc;
    const int64_t rows = (int64_t)c->t.n_frames * c->t.n_atoms;
    TCHK(c, tr_feature_stats(c->t, rows, part, blocks, c->st));
    std::vector<double> h((size_t)blocks * 4 * kMlpK);
    TCHK(c, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
    TCHK(c, hipStreamSynchronize(c->st));
    for (int k = 0; k < kMlpK; ++k) {
        double s = 0, s2 = 0, mn = INFINITY, mx = -INFINITY;
        for (int b = 0; b < blocks; ++b) {
            const double* p = &h[(size_t)b * 4 * kMlpK];
            s += p[k];
            s2 += p[kMlpK + k];
            mn = std::min(mn, p[2 * kMlpK + k]);
            mx = std::max(mx, p[3 * kMlpK + k]);
        }
        const double mean = s / rows;
        out[k] = mean;
        out[kMlpK + k] = sqrt(std::max(0.0, s2 / rows - mean * mean));
        out[2 * kMlpK + k] = mn;
        out[3 * kMlpK + k] = mx;
    }
    return MD_OK;
}

int md_trainer_set_params(md_trainer* c, const float* theta, const float* g_mean, const float* g_inv_std,
                          double e_scale, double e_shift) {
    if (!c || !theta || !g_mean || !g_inv_std) return tfail(c, MD_ERR_ARG, "bad argument");
    for (int k = 0; k < kMlpK; ++k) {
        c->t.g_mean[k] = g_mean[k];
        c->t.g_inv_std[k] = g_inv_std[k];
    }
    c->t.e_scale = (float)e_scale;
    c->t.e_shift = (float)e_shift;
    TCHK(c, hipMemcpyAsync(c->t.theta, theta, sizeof(float) * kMlpNP, hipMemcpyHostToDevice, c->st));
    TCHK(c, hipMemsetAsync(c->t.adam_m, 0, sizeof(float) * kMlpNP, c->st));
    TCHK(c, hipMemsetAsync(c->t.adam_v, 0, sizeof(float) * kMlpNP, c->st));
    TCHK(c, hipStreamSynchronize(c->st));
    c->step = 0;
    return MD_OK;
}

int md_trainer_get_params(md_trainer* c, float* theta) {
    if (!c || !theta) return tfail(c, MD_ERR_ARG, "bad argument");
    TCHK(c, hipMemcpyAsync(theta, c->t.theta, sizeof(float) * kMlpNP, hipMemcpyDeviceToHost, c->st));
    TCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

// Loss over the whole data set, out = {loss, energy MSE, force MSE per
// component}; with grad != NULL also its gradient (MD_MLP_NUM_PARAMS floats).
int md_trainer_loss(md_trainer* c, double w_e, double w_f, int batch_frames, double out[3], float* grad) {
    if (!c || !out || batch_frames < 1) return tfail(c, MD_ERR_ARG, "bad argument");
    TrainState& t = c->t;
    t.w_e = (float)w_e;
    t.w_f = (float)w_f;
    t.n_total = (int64_t)t.n_frames * t.n_atoms;
    if (grad) TCHK(c, hipMemsetAsync(t.grad, 0, sizeof(float) * kMlpNP, c->st));
    int nb = 0;
    for (int f0 = 0; f0 < t.n_frames; f0 += batch_frames, ++nb) {
        const int nf = std::min(batch_frames, t.n_frames - f0);
        int rc = run_batch(c, c->frames_seq + f0, nf, grad != nullptr, c->batch_loss + 2 * nb);
        if (rc) return rc;
    }
    std::vector<double> bl;
    int rc = read_batch_loss(c, nb, bl);
    if (rc) return rc;
    double se = 0, sf = 0;
    for (int q = 0; q < nb; ++q) {
        se += bl[(size_t)2 * q];
        sf += bl[(size_t)2 * q + 1];
    }
    const double n = (double)t.n_total;
    out[1] = se / n;
    out[2] = sf / (3.0 * n);
    out[0] = w_e * out[1] + w_f * out[2];
    if (grad) {
        TCHK(c, hipMemcpyAsync(grad, t.grad, sizeof(float) * kMlpNP, hipMemcpyDeviceToHost, c->st));
        TCHK(c, hipStreamSynchronize(c->st));
    }
    return MD_OK;
}

// One pass over the data set in minibatches of whole frames (frame order
// given by `order`, or sequential), one Adam step per minibatch on its mean
// loss; out = mean minibatch {loss, energy MSE, force MSE}.
int md_trainer_epoch(md_trainer* c, int batch_frames, double lr, double w_e, double w_f, const int32_t* order,
                     double out[3]) {
    if (!c || !out || batch_frames < 1 || !(lr > 0)) return tfail(c, MD_ERR_ARG, "bad argument");
    TrainState& t = c->t;
    t.w_e = (float)w_e;
    t.w_f = (float)w_f;
    // minibatch k holds the frames order[k batch_frames ...] (or sequential)
    if (order) {
        for (int q = 0; q < t.n_frames; ++q)
            if (order[q] < 0 || order[q] >= t.n_frames) return tfail(c, MD_ERR_ARG, "order holds a bad frame index");
        TCHK(c, hipMemcpyAsync(c->frames_ord, order, sizeof(int32_t) * t.n_frames, hipMemcpyHostToDevice, c->st));
    }
    const int32_t* frames = order ? c->frames_ord : c->frames_seq;
    std::vector<int64_t> rows;
    int nb = 0;
    for (int b0 = 0; b0 < t.n_frames; b0 += batch_frames, ++nb) {
        const int nf = std::min(batch_frames, t.n_frames - b0);
        t.n_total = (int64_t)nf * t.n_atoms;
        rows.push_back(t.n_total);
        TCHK(c, hipMemsetAsync(t.grad, 0, sizeof(float) * kMlpNP, c->st));
        int rc = run_batch(c, frames + b0, nf, true, c->batch_loss + 2 * nb);
        if (rc) return rc;
        TCHK(c, tr_adam(t, (float)lr, ++c->step, c->st));
    }
    std::vector<double> bl;
    int rc = read_batch_loss(c, nb, bl);
    if (rc) return rc;
    double acc[3] = {0, 0, 0};
    for (int q = 0; q < nb; ++q) {
        const double n = (double)rows[(size_t)q];
        const double se = bl[(size_t)2 * q], sf = bl[(size_t)2 * q + 1];
        acc[1] += se / n;
        acc[2] += sf / (3.0 * n);
        acc[0] += w_e * se / n + w_f * sf / (3.0 * n);
    }
    for (int k = 0; k < 3; ++k) out[k] = acc[k] / std::max(nb, 1);
    return MD_OK;
}

// Per-atom energies and forces of one frame under the current parameters
// (FP32 training kernels).
int md_trainer_predict(md_trainer* c, int frame, double* e_out, double* f_out) {
    if (!c || frame < 0 || frame >= c->t.n_frames) return tfail(c, MD_ERR_ARG, "bad argument");
    TrainState& t = c->t;
    t.n_total = t.n_atoms;
    int rc = run_batch(c, c->frames_seq + frame, 1, false, c->batch_loss);
    if (rc) return rc;
    const int n = t.n_atoms;
    std::vector<float> he((size_t)n), hf((size_t)n * 3);
    TCHK(c, hipMemcpyAsync(he.data(), t.E, sizeof(float) * n, hipMemcpyDeviceToHost, c->st));
    TCHK(c, hipMemcpyAsync(hf.data(), t.F, sizeof(float) * n * 3, hipMemcpyDeviceToHost, c->st));
    TCHK(c, hipStreamSynchronize(c->st));
    for (int i = 0; i < n; ++i) {
        if (e_out) e_out[i] = he[(size_t)i];
        if (f_out)
            for (int k = 0; k < 3; ++k) f_out[3 * i + k] = hf[(size_t)(3 * i + k)];
    }
    return MD_OK;
}

}  // extern "C"
