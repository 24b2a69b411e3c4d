// C-ABI of the learned indicator's trainer: device buffers, epochs and batches, Adam steps.
// The kernels (forward + back-propagation, weight gradients, update, packing) are in
// indicator.hip; every reduction over samples is done in a fixed order, so a training run is
// reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rkdg.h"
#include "indicator.hpp"
#include "launch.hpp"

using rkdg::mlp::NPARAM;

static_assert(sizeof(rkdg_mlp_weights) == (NPARAM + 1) * sizeof(float), "rkdg_mlp_weights must be the flat parameters");

struct rkdg_mlp_trainer {
    hipStream_t stream = nullptr;
    std::string err;
    int cap = 0, n = 0;
    float threshold = 0.0f;
    std::vector<void*> allocs;
    float *X = nullptr, *y = nullptr;          // training set [36][cap], [cap]
    int* idx = nullptr;                        // [cap] sample order of the current epoch
    float *Xs = nullptr, *ys = nullptr;        // staging for forward / gradient [36][cap], [cap]
    int* ident = nullptr;                      // [cap] 0, 1, 2, ...
    float *theta = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    uint16_t* wpack = nullptr;
    // per-batch buffers for up to bcap padded samples
    int bcap = 0;
    uint16_t *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *dz1 = nullptr, *dz2 = nullptr, *dz3 = nullptr;
    float *h3 = nullptr, *dz = nullptr, *loss = nullptr, *logit = nullptr, *part = nullptr, *loss_part = nullptr;
    float* step_loss = nullptr;  // [steps of an epoch]
    float* scal = nullptr;       // [4] scalar results
    int step_loss_cap = 0;
    long long t = 0;             // Adam steps taken

    ~rkdg_mlp_trainer() {
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    bool fail(const std::string& what, hipError_t e) {
        err = what + ": " + hipGetErrorString(e);
        return false;
    }
    template <typename T>
    bool alloc(T*& p, size_t count) {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, count * sizeof(T) + 256);
        if (e != hipSuccess) return fail("hipMalloc", e);
        allocs.push_back(v);
        p = (T*)v;
        return true;
    }
    bool release(void* p) {
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) (void)hipFree(*it), allocs.erase(it);
        return true;
    }
    static int pad(int b) { return (b + 31) / 32 * 32; }
    bool ensure_batch(int bpad) {
        if (bpad <= bcap) return true;
        for (void* p : {(void*)xb, (void*)h1, (void*)h2, (void*)dz1, (void*)dz2, (void*)dz3, (void*)h3, (void*)dz,
                        (void*)loss, (void*)logit, (void*)part, (void*)loss_part})
            if (p) release(p);
        const size_t b = (size_t)bpad, ns = (size_t)rkdg::mlp::wgrad_slices_cap(bpad);
        if (!alloc(xb, 48 * b) || !alloc(h1, 64 * b) || !alloc(h2, 64 * b) || !alloc(dz1, 64 * b) ||
            !alloc(dz2, 64 * b) || !alloc(dz3, 64 * b) || !alloc(h3, 64 * b) || !alloc(dz, b) || !alloc(loss, b) ||
            !alloc(logit, b) || !alloc(part, ns * NPARAM) || !alloc(loss_part, ns))
            return false;
        bcap = bpad;
        return true;
    }
    rkdg::TrainArgs args(const float* Xm, long long ldx, const float* ym, const int* ix, int B, float pw) const {
        rkdg::TrainArgs T;
        std::memset(&T, 0, sizeof T);
        T.B = B, T.Bpad = pad(B), T.X = Xm, T.ldx = ldx, T.y = ym, T.idx = ix, T.pos_weight = pw;
        T.wpack = wpack, T.theta = theta;
        T.xb = xb, T.h1 = h1, T.h2 = h2, T.dz1 = dz1, T.dz2 = dz2, T.dz3 = dz3;
        T.h3 = h3, T.dz = dz, T.loss = loss, T.logit = logit;
        return T;
    }
    static bool finite(const float* a, size_t n) {
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(a[i])) return false;
        return true;
    }
    // x [36][m] (host) -> staging [36][cap]
    bool stage_in(const float* x, const float* yh, int m) {
        hipError_t e = hipMemcpy2DAsync(Xs, (size_t)cap * sizeof(float), x, (size_t)m * sizeof(float),
                                        (size_t)m * sizeof(float), rkdg::mlp::NIN, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess && yh) e = hipMemcpyAsync(ys, yh, (size_t)m * sizeof(float), hipMemcpyHostToDevice, stream);
        return e == hipSuccess || fail("copy in", e);
    }
};

static rkdg_status terr(rkdg_mlp_trainer* t, const std::string& msg, rkdg_status s = RKDG_ERR_ARG) {
    if (t) t->err = msg;
    return s;
}

extern "C" {

rkdg_status rkdg_mlp_trainer_create(const rkdg_mlp_weights* init, int32_t capacity, rkdg_mlp_trainer** out) {
    if (!init || !out || capacity < 1) return RKDG_ERR_ARG;
    *out = nullptr;
    auto* t = new (std::nothrow) rkdg_mlp_trainer;
    if (!t) return RKDG_ERR_HIP;
    t->cap = capacity;
    t->threshold = init->threshold;
    const size_t c = (size_t)capacity;
    bool ok = hipStreamCreate(&t->stream) == hipSuccess;
    ok = ok && t->alloc(t->X, rkdg::mlp::NIN * c) && t->alloc(t->y, c) && t->alloc(t->idx, c) &&
         t->alloc(t->Xs, rkdg::mlp::NIN * c) && t->alloc(t->ys, c) && t->alloc(t->ident, c) &&
         t->alloc(t->theta, NPARAM) && t->alloc(t->m1, NPARAM) && t->alloc(t->m2, NPARAM) &&
         t->alloc(t->grad, NPARAM) && t->alloc(t->scal, 4) && t->alloc(t->wpack, (size_t)rkdg::mlp::NPACKED * rkdg::mlp::PACK);
    if (ok) {
        std::vector<int> id(c);
        std::iota(id.begin(), id.end(), 0);
        ok = hipMemcpy(t->ident, id.data(), c * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(t->theta, init, NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemset(t->m1, 0, NPARAM * sizeof(float)) == hipSuccess &&
             hipMemset(t->m2, 0, NPARAM * sizeof(float)) == hipSuccess &&
             rkdg::launch_mlp_pack(t->theta, t->wpack, t->stream) == hipSuccess &&
             hipStreamSynchronize(t->stream) == hipSuccess;
    }
    if (!ok) {
        delete t;
        return RKDG_ERR_HIP;
    }
    *out = t;
    return RKDG_OK;
}

rkdg_status rkdg_mlp_set_data(rkdg_mlp_trainer* t, const float* x, const float* y, int32_t n) {
    if (!t || !x || !y || n < 1) return terr(t, "bad arguments");
    if (n > t->cap) return terr(t, "more samples than the trainer's capacity");
    if (!rkdg_mlp_trainer::finite(x, (size_t)rkdg::mlp::NIN * n) || !rkdg_mlp_trainer::finite(y, n))
        return terr(t, "non-finite training data");
    hipError_t e = hipMemcpy2D(t->X, (size_t)t->cap * sizeof(float), x, (size_t)n * sizeof(float),
                               (size_t)n * sizeof(float), rkdg::mlp::NIN, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t->y, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return t->fail("set_data", e), RKDG_ERR_HIP;
    t->n = n;
    return RKDG_OK;
}

rkdg_status rkdg_mlp_train(rkdg_mlp_trainer* t, const rkdg_train_params* prm, float* epoch_loss) {
    if (!t || !prm) return RKDG_ERR_ARG;
    if (t->n < 1) return terr(t, "no training data");
    if (prm->batch < 1 || prm->epochs < 0 || !(prm->lr > 0.0f) || !(prm->beta1 >= 0.0f && prm->beta1 < 1.0f) ||
        !(prm->beta2 >= 0.0f && prm->beta2 < 1.0f) || !(prm->eps > 0.0f) || !(prm->pos_weight > 0.0f))
        return terr(t, "bad training parameters");
    const int n = t->n, batch = std::min(prm->batch, n);
    const int steps = (n + batch - 1) / batch;
    if (!t->ensure_batch(rkdg_mlp_trainer::pad(batch))) return RKDG_ERR_HIP;
    if (steps > t->step_loss_cap) {
        if (t->step_loss) t->release(t->step_loss);
        if (!t->alloc(t->step_loss, (size_t)steps)) return RKDG_ERR_HIP;
        t->step_loss_cap = steps;
    }
    std::mt19937_64 rng(prm->seed);
    // Sample orders of consecutive epochs alternate between two host buffers: a host thread
    // shuffles the next epoch's while this one launches and the GPU trains on the current one (the
    // copy from the other buffer completed before the previous epoch's synchronisation; only that
    // thread draws from the generator meanwhile, so the sequence of orders is unchanged).
    std::vector<int> perm[2] = {std::vector<int>(n), std::vector<int>(n)};
    auto shuffle = [&](std::vector<int>& pm) {
        std::iota(pm.begin(), pm.end(), 0);
        for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with a fixed generator
            const int j = (int)(rng() % (uint64_t)(i + 1));
            std::swap(pm[i], pm[j]);
        }
    };
    std::vector<float> sl(steps);
    if (prm->epochs > 0) shuffle(perm[0]);
    for (int ep = 0; ep < prm->epochs; ++ep) {
        const std::vector<int>& pm = perm[ep & 1];
        hipError_t e = hipMemcpyAsync(t->idx, pm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, t->stream);
        std::thread next;
        bool shuffled = ep + 1 >= prm->epochs;
        if (!shuffled) try {
                next = std::thread(shuffle, std::ref(perm[(ep + 1) & 1]));
                shuffled = true;
            } catch (const std::exception&) {  // no thread to be had: shuffle after the launches
            }
        for (int s = 0; s < steps && e == hipSuccess; ++s) {
            const int b0 = s * batch, B = std::min(batch, n - b0);
            const rkdg::TrainArgs T = t->args(t->X, t->cap, t->y, t->idx + b0, B, prm->pos_weight);
            const int nslices = rkdg::mlp::wgrad_slices(T.Bpad);
            ++t->t;
            const float c1 = (float)(1.0 / (1.0 - std::pow((double)prm->beta1, (double)t->t)));
            const float c2 = (float)(1.0 / (1.0 - std::pow((double)prm->beta2, (double)t->t)));
            e = rkdg::launch_mlp_fwdbwd(T, true, t->stream);
            if (e == hipSuccess) e = rkdg::launch_mlp_wgrad(T, t->part, t->loss_part, t->stream);
            if (e == hipSuccess)
                e = rkdg::launch_mlp_update(nslices, t->part, t->loss_part, nullptr, t->theta, t->m1, t->m2, prm->lr,
                                            prm->beta1, prm->beta2, prm->eps, c1, c2, true, t->step_loss + s, t->stream);
            if (e == hipSuccess) e = rkdg::launch_mlp_pack(t->theta, t->wpack, t->stream);
        }
        if (next.joinable()) next.join();
        if (!shuffled) shuffle(perm[(ep + 1) & 1]);
        if (e == hipSuccess) e = hipMemcpyAsync(sl.data(), t->step_loss, steps * sizeof(float), hipMemcpyDeviceToHost, t->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
        if (e != hipSuccess) return t->fail("training step", e), RKDG_ERR_HIP;
        if (epoch_loss) {
            double acc = 0.0;
            for (int s = 0; s < steps; ++s) acc += sl[s] / (double)std::min(batch, n - s * batch);
            epoch_loss[ep] = (float)(acc / steps);
        }
    }
    return RKDG_OK;
}

rkdg_status rkdg_mlp_get_weights(rkdg_mlp_trainer* t, rkdg_mlp_weights* out) {
    if (!t || !out) return RKDG_ERR_ARG;
    hipError_t e = hipStreamSynchronize(t->stream);
    if (e == hipSuccess) e = hipMemcpy(out, t->theta, NPARAM * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return t->fail("get_weights", e), RKDG_ERR_HIP;
    out->threshold = t->threshold;
    return RKDG_OK;
}

rkdg_status rkdg_mlp_forward(rkdg_mlp_trainer* t, const float* x, int32_t n, float* logit) {
    if (!t || !x || !logit || n < 1) return terr(t, "bad arguments");
    if (n > t->cap) return terr(t, "more samples than the trainer's capacity");
    if (!rkdg_mlp_trainer::finite(x, (size_t)rkdg::mlp::NIN * n)) return terr(t, "non-finite features");
    if (!t->ensure_batch(rkdg_mlp_trainer::pad(n)) || !t->stage_in(x, nullptr, n)) return RKDG_ERR_HIP;
    const rkdg::TrainArgs T = t->args(t->Xs, t->cap, t->ys, t->ident, n, 1.0f);
    hipError_t e = rkdg::launch_mlp_fwdbwd(T, false, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(logit, t->logit, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) return t->fail("forward", e), RKDG_ERR_HIP;
    return RKDG_OK;
}

rkdg_status rkdg_mlp_gradient(rkdg_mlp_trainer* t, const float* x, const float* y, int32_t n, float pos_weight,
                              float* loss, rkdg_mlp_weights* grad) {
    if (!t || !x || !y || n < 1 || !(pos_weight > 0.0f)) return terr(t, "bad arguments");
    if (n > t->cap) return terr(t, "more samples than the trainer's capacity");
    if (!rkdg_mlp_trainer::finite(x, (size_t)rkdg::mlp::NIN * n) || !rkdg_mlp_trainer::finite(y, n))
        return terr(t, "non-finite samples");
    if (!t->ensure_batch(rkdg_mlp_trainer::pad(n)) || !t->stage_in(x, y, n)) return RKDG_ERR_HIP;
    const rkdg::TrainArgs T = t->args(t->Xs, t->cap, t->ys, t->ident, n, pos_weight);
    const int nslices = rkdg::mlp::wgrad_slices(T.Bpad);
    float lsum = 0.0f;
    hipError_t e = rkdg::launch_mlp_fwdbwd(T, true, t->stream);
    if (e == hipSuccess) e = rkdg::launch_mlp_wgrad(T, t->part, t->loss_part, t->stream);
    if (e == hipSuccess)
        e = rkdg::launch_mlp_update(nslices, t->part, t->loss_part, t->grad, t->theta, t->m1, t->m2, 0.0f, 0.0f, 0.0f,
                                    0.0f, 0.0f, 0.0f, false, t->scal, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&lsum, t->scal, sizeof(float), hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, t->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) return t->fail("gradient", e), RKDG_ERR_HIP;
    if (grad) grad->threshold = 0.0f;
    if (loss) *loss = lsum / (float)n;
    return RKDG_OK;
}

rkdg_status rkdg_mlp_forward_noalloc(const rkdg_mlp_weights* w, const float* x, int32_t n, float* logit) {
    if (!w || !x || !logit || n < 1 || n > RKDG_MLP_FORWARD_NOALLOC_MAX) return RKDG_ERR_ARG;
    if (!rkdg_mlp_trainer::finite(x, (size_t)rkdg::mlp::NIN * n)) return RKDG_ERR_ARG;
    return rkdg::mlp_forward_noalloc(reinterpret_cast<const float*>(w), x, n, logit) == hipSuccess ? RKDG_OK
                                                                                                  : RKDG_ERR_HIP;
}

const char* rkdg_mlp_trainer_error(const rkdg_mlp_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

void rkdg_mlp_trainer_destroy(rkdg_mlp_trainer* t) { delete t; }

}  // extern "C"
