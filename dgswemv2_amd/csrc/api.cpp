// C-ABI of the RKDG stage engine: context lifecycle, HBM buffers, stage orchestration.
//
// The host drives the reference's step/stage loop (Problem::step_serial / stage_serial,
// kernels_processor/rkdg_swe_proc_serial_step.hpp:9-67) and the ESSPRKStepper bookkeeping;
// every per-element and per-edge computation runs in the HIP kernels of kernels.hip.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <limits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/rkdg.h"
#include "launch.hpp"
#include "model.hpp"

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace rkdg {
// Sets the message rkdg_create_error() returns (also used by the mesh reader).
void set_create_error(const std::string& msg) { g_create_error = msg; }
}  // namespace rkdg

namespace {

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual rkdg_status init() = 0;
    virtual rkdg_status stage() = 0;
    virtual rkdg_status set_state(const double* q, const uint8_t* wet) = 0;
    virtual rkdg_status get_state(double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled) = 0;
    virtual rkdg_status nan_key(unsigned long long* key) = 0;
    virtual void collect(rkdg_perf_report* r) = 0;  // adds timed kernel milliseconds, recycles events
    std::string err;
    bool timing = false;
    int64_t stages_timed = 0;
};

const unsigned long long kNoNan = ~0ull;

template <typename Real>
struct Engine final : EngineBase {
    rkdg::Model& M;
    hipStream_t stream;
    int block = 128;  // work-items per block (measured best for the stage kernels on 1M elements)
    std::vector<void*> allocs;
    std::vector<Real*> q;      // reference state slots 0..nstages
    std::vector<Real*> sol;    // per stage, only where a later stage reads it
    Real *bath = nullptr, *geom = nullptr, *egeom = nullptr, *vxy = nullptr, *gmn2 = nullptr;
    Real *famp = nullptr, *fphase = nullptr, *levee_gp = nullptr;
    int *nb = nullptr, *edge = nullptr, *eforce = nullptr, *neighbor = nullptr;
    uint32_t* meta = nullptr;
    uint8_t* wet[2] = {nullptr, nullptr};
    int wcur = 0;
    double* bath_vrtx = nullptr;
    unsigned long long* d_nan = nullptr;
    // limiter
    double *median = nullptr, *alpha = nullptr, *rsq = nullptr, *len = nullptr, *radius = nullptr,
           *bath_bc = nullptr, *baryctr = nullptr;
    uint8_t *set_a = nullptr, *set_b = nullptr, *troubled = nullptr, *limited = nullptr;
    // timing
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> stage_ev, lim_ev;

    bool fast;  // fast form (FP32 bulk, FP64 fast) instead of the reference-order form

    hipError_t launch_plain(const rkdg::StageArgs<Real>& A) {
        if constexpr (sizeof(Real) == 8)
            return rkdg::launch_stage<Real, false>(M.p, M.prm.wetting_drying != 0, M.closed_forms, A, stream, block);
        else
            return hipErrorInvalidValue;  // FP32 exists only in the fast form
    }
    hipError_t launch_fast(const rkdg::StageArgs<Real>& A) {
        return rkdg::launch_stage<Real, true>(M.p, M.prm.wetting_drying != 0, M.closed_forms, A, stream, block);
    }

    Engine(rkdg::Model& m, hipStream_t s, bool fast_form) : M(m), stream(s), fast(fast_form) {
        if (const char* b = std::getenv("RKDG_BLOCK")) block = std::atoi(b);  // tuning experiments
        if (block < 64 || block > 256 || block % 64) block = 128;
    }
    ~Engine() override {
        for (auto& pr : stage_ev) hipEventDestroy(pr.first), hipEventDestroy(pr.second);
        for (auto& pr : lim_ev) hipEventDestroy(pr.first), hipEventDestroy(pr.second);
        for (hipEvent_t ev : ev_pool) hipEventDestroy(ev);
        for (void* p : allocs) hipFree(p);
    }

    bool fail(const char* what, hipError_t e) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return false;
    }
    template <typename T>
    bool alloc(T*& p, size_t n) {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, n * sizeof(T) + 16);
        if (e != hipSuccess) return fail("hipMalloc", e);
        allocs.push_back(v);
        p = (T*)v;
        return true;
    }
    template <typename T, typename S>
    bool upload(T*& p, const std::vector<S>& h) {
        if (!alloc(p, h.size() ? h.size() : 1)) return false;
        if (h.empty()) return true;
        std::vector<T> tmp(h.begin(), h.end());
        hipError_t e = hipMemcpy(p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice);
        return e == hipSuccess ? true : fail("hipMemcpy", e);
    }

    rkdg_status init() override {
        const size_t ne = M.ne, S = 3 * (size_t)M.ndof * ne;
        const int ns = M.stepper.nstages;
        q.assign(ns + 1, nullptr);
        sol.assign(ns, nullptr);
        for (int s = 0; s <= ns; ++s)
            if (!alloc(q[s], S)) return RKDG_ERR_HIP;
        for (int s = 0; s < ns; ++s) {
            bool needed = false;  // a later stage combines this stage's M^-1 rhs
            for (int i = s + 1; i < ns; ++i) needed |= (M.stepper.brk[i][s] != 0.0);
            if (needed && !alloc(sol[s], S)) return RKDG_ERR_HIP;
        }
        if (!upload(bath, M.bath) || !upload(geom, M.geom) || !upload(egeom, M.egeom) || !upload(vxy, M.vxy) ||
            !upload(gmn2, M.gmn2) || !upload(famp, M.famp) || !upload(fphase, M.fphase) ||
            !upload(levee_gp, M.levee_gp) || !upload(nb, M.nb) ||
            !upload(edge, M.edge) || !upload(eforce, M.eforce) || !upload(meta, M.meta) ||
            !upload(neighbor, M.neighbor) || !upload(bath_vrtx, M.bath_vrtx))
            return RKDG_ERR_HIP;
        if (!upload(wet[0], M.wet0) || !upload(wet[1], M.wet0)) return RKDG_ERR_HIP;
        if (M.prm.slope_limiting) {
            if (!upload(median, M.sl_median) || !upload(alpha, M.sl_alpha) || !upload(rsq, M.sl_rsq) ||
                !upload(len, M.sl_len) || !upload(radius, M.sl_radius) || !upload(bath_bc, M.sl_bath_bc) ||
                !alloc(baryctr, 3 * ne) || !alloc(set_a, ne) || !alloc(set_b, ne) || !alloc(troubled, ne) ||
                !alloc(limited, 3 * ne))
                return RKDG_ERR_HIP;
            if (hipMemset(baryctr, 0, 3 * ne * sizeof(double)) != hipSuccess ||
                hipMemset(troubled, 0, ne) != hipSuccess || hipMemset(limited, 0, 3 * ne) != hipSuccess)
                return fail("hipMemset", hipGetLastError()), RKDG_ERR_HIP;
        }
        if (!alloc(d_nan, 1)) return RKDG_ERR_HIP;
        if (hipMemcpy(d_nan, &kNoNan, sizeof kNoNan, hipMemcpyHostToDevice) != hipSuccess)
            return fail("hipMemcpy", hipGetLastError()), RKDG_ERR_HIP;
        std::vector<Real> q0(M.q0.begin(), M.q0.end());
        if (hipMemcpy(q[0], q0.data(), q0.size() * sizeof(Real), hipMemcpyHostToDevice) != hipSuccess)
            return fail("hipMemcpy state", hipGetLastError()), RKDG_ERR_HIP;
        return RKDG_OK;
    }

    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }

    rkdg_status stage() override {
        rkdg::Stepper& st = M.stepper;
        const int s = st.stage, ns = st.nstages;
        rkdg::StageArgs<Real> A;
        std::memset(&A, 0, sizeof A);
        A.ne = M.ne, A.nedge = M.nedge;
        A.q_s = q[s];
        A.q_out = q[s + 1];
        A.sol_out = sol[s];
        for (int k = 0; k < s; ++k) {
            const double a = st.ark[s][k], c = st.dt * st.brk[s][k];
            if (a == 0.0 && c == 0.0) continue;
            A.term_q[A.nterm] = q[k];
            A.term_sol[A.nterm] = c != 0.0 ? sol[k] : nullptr;
            A.term_a[A.nterm] = (Real)a;
            A.term_c[A.nterm] = (Real)c;
            ++A.nterm;
        }
        A.cur_a = (Real)st.ark[s][s];
        A.cur_c = (Real)(st.dt * st.brk[s][s]);
        A.bath = bath, A.geom = geom, A.nb = nb, A.edge = edge, A.meta = meta, A.egeom = egeom;
        A.wet_in = wet[wcur], A.wet_out = wet[wcur ^ 1];
        A.bath_vrtx = bath_vrtx, A.gmn2 = gmn2, A.vxy = vxy, A.eforce = eforce, A.famp = famp, A.fphase = fphase;
        A.levee_gp = levee_gp;
        A.ncon = M.ncon;
        for (int c = 0; c < M.ncon; ++c)
            A.freq[c] = (Real)M.freq[c], A.ffact[c] = (Real)M.ffact[c], A.eqarg[c] = (Real)M.eqarg[c];
        A.t = st.stage_time();
        A.ramp = st.ramp;
        A.g = (Real)M.prm.g;
        A.cf = (Real)M.prm.friction_cf;
        A.friction = M.prm.friction;
        A.fsrc = M.prm.function_source;
        A.problem = M.prm.problem;
        for (int i = 0; i < 4; ++i) A.pparam[i] = M.prm.problem_param[i];
        A.h_o = M.h_o, A.h_thr = M.h_thr;
        A.nan_check = M.prm.slope_limiting ? 0 : 1;
        A.nan_key = d_nan;
        A.stage_key = (unsigned long long)st.step * ns + s;

        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing) e0 = take_event(), hipEventRecord(e0, stream);
        hipError_t e = (fast ? launch_fast(A) : launch_plain(A));
        if (e != hipSuccess) return fail("stage kernel launch", e), RKDG_ERR_HIP;
        if (timing) e1 = take_event(), hipEventRecord(e1, stream), stage_ev.push_back({e0, e1});
        ++stages_timed;

        if (s + 1 == ns) std::swap(q[0], q[ns]);  // UpdateState's final swap
        const unsigned long long key = A.stage_key;
        st.advance();
        if (M.prm.wetting_drying) wcur ^= 1;

        if (M.prm.slope_limiting) {
            rkdg::LimiterArgs<Real> L;
            std::memset(&L, 0, sizeof L);
            L.ne = M.ne, L.nedge = M.nedge;
            L.q = q[st.stage];
            L.bath = bath, L.nb = nb, L.edge = edge, L.meta = meta, L.egeom = egeom;
            L.wet = wet[wcur];
            L.bath_vrtx = bath_vrtx, L.neighbor = neighbor;
            L.median = median, L.alpha = alpha, L.rsq = rsq, L.len = len, L.radius = radius, L.bath_bc = bath_bc;
            L.baryctr = baryctr, L.set_a = set_a, L.set_b = set_b, L.troubled = troubled, L.limited = limited;
            L.g = M.prm.g, L.M = M.sl_M, L.nu = M.sl_nu, L.p = M.p;
            L.nan_check = 1, L.nan_key = d_nan, L.stage_key = key;
            hipEvent_t l0 = nullptr, l1 = nullptr;
            if (timing) l0 = take_event(), hipEventRecord(l0, stream);
            e = rkdg::launch_limiter<Real>(M.p, L, stream, block);
            if (e != hipSuccess) return fail("limiter kernel launch", e), RKDG_ERR_HIP;
            if (timing) l1 = take_event(), hipEventRecord(l1, stream), lim_ev.push_back({l0, l1});
        }
        return RKDG_OK;
    }

    rkdg_status set_state(const double* qh, const uint8_t* wh) override {
        const size_t S = 3 * (size_t)M.ndof * M.ne;
        if (qh) {
            std::vector<Real> tmp(qh, qh + S);
            hipError_t e = hipMemcpyAsync(q[M.stepper.stage], tmp.data(), S * sizeof(Real), hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return fail("set_state", e), RKDG_ERR_HIP;
        }
        if (wh) {
            hipError_t e = hipMemcpyAsync(wet[wcur], wh, M.ne, hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return fail("set_state wet", e), RKDG_ERR_HIP;
        }
        return RKDG_OK;
    }

    rkdg_status get_state(double* qh, uint8_t* wh, uint8_t* lh, uint8_t* th) override {
        const size_t ne = M.ne, S = 3 * (size_t)M.ndof * ne;
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return fail("synchronize", e), RKDG_ERR_HIP;
        if (qh) {
            std::vector<Real> tmp(S);
            e = hipMemcpy(tmp.data(), q[M.stepper.stage], S * sizeof(Real), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return fail("get_state", e), RKDG_ERR_HIP;
            for (size_t i = 0; i < S; ++i) qh[i] = (double)tmp[i];
        }
        if (wh) {
            e = hipMemcpy(wh, wet[wcur], ne, hipMemcpyDeviceToHost);
            if (e != hipSuccess) return fail("get_state wet", e), RKDG_ERR_HIP;
        }
        if (lh) {
            if (limited)
                e = hipMemcpy(lh, limited, 3 * ne, hipMemcpyDeviceToHost);
            else
                std::memset(lh, 0, 3 * ne);
            if (e != hipSuccess) return fail("get_state limited", e), RKDG_ERR_HIP;
        }
        if (th) {
            if (troubled)
                e = hipMemcpy(th, troubled, ne, hipMemcpyDeviceToHost);
            else
                std::memset(th, 0, ne);
            if (e != hipSuccess) return fail("get_state troubled", e), RKDG_ERR_HIP;
        }
        return RKDG_OK;
    }

    rkdg_status nan_key(unsigned long long* key) override {
        hipError_t e = hipMemcpyAsync(key, d_nan, sizeof *key, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return fail("nan readback", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }

    void collect(rkdg_perf_report* r) override {
        hipStreamSynchronize(stream);
        double sms = 0, lms = 0;
        for (auto& pr : stage_ev) {
            float ms = 0;
            hipEventElapsedTime(&ms, pr.first, pr.second);
            sms += ms;
            ev_pool.push_back(pr.first), ev_pool.push_back(pr.second);
        }
        for (auto& pr : lim_ev) {
            float ms = 0;
            hipEventElapsedTime(&ms, pr.first, pr.second);
            lms += ms;
            ev_pool.push_back(pr.first), ev_pool.push_back(pr.second);
        }
        stage_ev.clear(), lim_ev.clear();
        r->stage_kernel_ms += sms;
        r->limiter_kernel_ms += lms;
    }
};

}  // namespace

struct rkdg_ctx {
    rkdg::Model model;
    hipStream_t stream = nullptr;
    std::unique_ptr<EngineBase> eng;
    rkdg_precision precision = RKDG_FP64_PLAIN;
    std::string err;
    rkdg_perf_report acc{};
    bool timing = false;
    double total_ms = 0;
};

static rkdg_status set_err(rkdg_ctx* c, rkdg_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

extern "C" {

const char* rkdg_create_error(void) { return g_create_error.c_str(); }

rkdg_status rkdg_create(const rkdg_mesh_desc* mesh, const rkdg_params* params, rkdg_precision precision,
                        rkdg_ctx** out) {
    g_create_error.clear();
    if (!mesh || !params || !out) {
        g_create_error = "null argument";
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    if (precision != RKDG_FP64_PLAIN && precision != RKDG_FP32_BULK && precision != RKDG_FP64_FAST) {
        g_create_error = "unknown precision";
        return RKDG_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device visible";
        return RKDG_ERR_NO_DEVICE;
    }
    std::unique_ptr<rkdg_ctx> c(new (std::nothrow) rkdg_ctx());
    if (!c) return RKDG_ERR_ARG;
    c->precision = precision;
    std::string msg;
    try {
        msg = rkdg::build_model(*mesh, *params, c->model, precision == RKDG_FP64_PLAIN);
    } catch (const std::exception& ex) {
        msg = ex.what();
    }
    if (!msg.empty()) {
        g_create_error = msg;
        return RKDG_ERR_ARG;
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        return RKDG_ERR_HIP;
    }
    if (precision == RKDG_FP32_BULK)
        c->eng.reset(new Engine<float>(c->model, c->stream, true));
    else
        c->eng.reset(new Engine<double>(c->model, c->stream, precision == RKDG_FP64_FAST));
    rkdg_status s = c->eng->init();
    if (s != RKDG_OK) {
        g_create_error = c->eng->err;
        c->eng.reset();
        hipStreamDestroy(c->stream);
        return s;
    }
    *out = c.release();
    return RKDG_OK;
}

void rkdg_destroy(rkdg_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    ctx->eng.reset();
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rkdg_last_error(const rkdg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rkdg_status rkdg_set_state(rkdg_ctx* ctx, const double* q, const uint8_t* wet) {
    if (!ctx) return RKDG_ERR_ARG;
    rkdg_status s = ctx->eng->set_state(q, wet);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_get_state(rkdg_ctx* ctx, double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled) {
    if (!ctx) return RKDG_ERR_ARG;
    rkdg_status s = ctx->eng->get_state(q, wet, limited, troubled);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

static rkdg_status check_nan(rkdg_ctx* ctx) {
    unsigned long long key = kNoNan;
    rkdg_status s = ctx->eng->nan_key(&key);
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    if (key != kNoNan) return set_err(ctx, RKDG_ERR_NAN, "NaN found in the state");
    return RKDG_OK;
}

rkdg_status rkdg_stage(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    rkdg_status s = ctx->eng->stage();
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    return check_nan(ctx);
}

rkdg_status rkdg_step(rkdg_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return RKDG_ERR_ARG;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (ctx->timing) {
        hipEventCreate(&t0), hipEventCreate(&t1);
        hipEventRecord(t0, ctx->stream);
    }
    const int ns = ctx->model.stepper.nstages;
    for (int i = 0; i < nsteps; ++i)
        for (int k = 0; k < ns; ++k) {
            rkdg_status s = ctx->eng->stage();
            if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
        }
    if (ctx->timing) {
        hipEventRecord(t1, ctx->stream);
        hipEventSynchronize(t1);
        float ms = 0;
        hipEventElapsedTime(&ms, t0, t1);
        ctx->total_ms += ms;
        hipEventDestroy(t0), hipEventDestroy(t1);
    }
    return check_nan(ctx);
}

rkdg_status rkdg_sync(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RKDG_OK : set_err(ctx, RKDG_ERR_HIP, hipGetErrorString(e));
}

rkdg_status rkdg_get_time(rkdg_ctx* ctx, double* t, int32_t* step, int32_t* stage) {
    if (!ctx) return RKDG_ERR_ARG;
    if (t) *t = ctx->model.stepper.t;
    if (step) *step = ctx->model.stepper.step;
    if (stage) *stage = ctx->model.stepper.stage;
    return RKDG_OK;
}

rkdg_status rkdg_set_time(rkdg_ctx* ctx, double t, int32_t step) {
    if (!ctx || step < 0) return RKDG_ERR_ARG;
    rkdg::Stepper& st = ctx->model.stepper;
    if (st.stage != 0) return set_err(ctx, RKDG_ERR_ARG, "rkdg_set_time needs a step boundary (stage 0)");
    st.t = t;
    st.step = step;
    if (!(std::abs(st.ramp_duration) < std::numeric_limits<double>::epsilon() * 100))
        st.ramp = std::tanh(2 * (st.stage_time() / 86400) / st.ramp_duration);
    return RKDG_OK;
}

rkdg_status rkdg_nan_info(rkdg_ctx* ctx, int32_t* element, int32_t* step, int32_t* stage) {
    if (!ctx) return RKDG_ERR_ARG;
    unsigned long long key = kNoNan;
    rkdg_status s = ctx->eng->nan_key(&key);
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    const int ns = ctx->model.stepper.nstages;
    if (key == kNoNan) {
        if (element) *element = -1;
        if (step) *step = -1;
        if (stage) *stage = -1;
        return RKDG_OK;
    }
    const unsigned long long sk = key >> 32;
    if (element) *element = (int32_t)(key & 0xffffffffu);
    if (step) *step = (int32_t)(sk / ns);
    if (stage) *stage = (int32_t)(sk % ns);
    return RKDG_OK;
}

rkdg_status rkdg_l2_residual(rkdg_ctx* ctx, double* l2) {
    if (!ctx || !l2) return RKDG_ERR_ARG;
    const rkdg::Model& M = ctx->model;
    std::vector<double> q(3 * (size_t)M.ndof * M.ne);
    rkdg_status s = ctx->eng->get_state(q.data(), nullptr, nullptr, nullptr);
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    *l2 = rkdg::l2_residual(M, q, M.stepper.stage_time());
    return RKDG_OK;
}

rkdg_status rkdg_reset_timers(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    hipStreamSynchronize(ctx->stream);
    rkdg_perf_report scratch{};
    ctx->eng->collect(&scratch);
    ctx->eng->timing = true;
    ctx->eng->stages_timed = 0;
    ctx->acc = rkdg_perf_report{};
    ctx->timing = true;
    ctx->total_ms = 0;
    return RKDG_OK;
}

rkdg_status rkdg_report(rkdg_ctx* ctx, rkdg_perf_report* r) {
    if (!ctx || !r) return RKDG_ERR_ARG;
    const rkdg::Model& M = ctx->model;
    ctx->eng->collect(&ctx->acc);
    const int64_t staged = ctx->eng->stages_timed;
    *r = ctx->acc;
    r->n_elements = M.ne;
    r->p = M.p, r->ndof = M.ndof, r->nstages = M.stepper.nstages;
    r->precision = ctx->precision;
    r->stages_timed = staged;
    r->total_ms = ctx->total_ms;
    // Algorithmic HBM bytes of one stage launch averaged over the scheme's stages: every
    // array element the kernel must read or write, each counted once.
    const double b = ctx->precision == RKDG_FP32_BULK ? 4.0 : 8.0;
    const double S = 3.0 * M.ndof * b;
    double reads_q = 0;
    const int ns = M.stepper.nstages;
    for (int s = 0; s < ns; ++s) {
        reads_q += 1;  // q_s
        for (int k = 0; k < s; ++k) {
            if (M.stepper.ark[s][k] != 0.0 || M.stepper.brk[s][k] != 0.0) reads_q += 1;
            if (M.stepper.brk[s][k] != 0.0) reads_q += 1;
        }
    }
    reads_q /= ns;
    double per_elt = (reads_q + 1.0) * S + M.ndof * b + 5 * b + 28.0;
    if (M.prm.wetting_drying) per_elt += 2.0 + 24.0;
    r->stage_bytes_alg = per_elt * M.ne + 3.0 * b * M.nedge;
    return RKDG_OK;
}

rkdg_status rkdg_master_tables(int32_t p, int32_t* ndof, int32_t* ngp, double* phi, double* dz1, double* dz2,
                               double* w) {
    if (p < 0 || p > 10) return RKDG_ERR_ARG;
    rkdg::Master m = rkdg::make_master(p);
    if (ndof) *ndof = m.ndof;
    if (ngp) *ngp = m.ngp;
    if (phi) std::memcpy(phi, m.phi.data(), m.phi.size() * sizeof(double));
    if (dz1) std::memcpy(dz1, m.dphi[0].data(), m.dphi[0].size() * sizeof(double));
    if (dz2) std::memcpy(dz2, m.dphi[1].data(), m.dphi[1].size() * sizeof(double));
    if (w) std::memcpy(w, m.rule.w.data(), m.rule.w.size() * sizeof(double));
    return RKDG_OK;
}

rkdg_status rkdg_ssprk_coeffs(int32_t nstages, int32_t order, double* ark, double* brk, double* drk) {
    rkdg::Stepper st;
    if (!st.init(nstages, order, 1.0, 0.0)) return RKDG_ERR_ARG;
    for (int i = 0; i < nstages; ++i)
        for (int k = 0; k < nstages; ++k) {
            if (ark) ark[i * nstages + k] = st.ark[i][k];
            if (brk) brk[i * nstages + k] = st.brk[i][k];
        }
    if (drk)
        for (int i = 0; i < nstages; ++i) drk[i] = st.drk[i];
    return RKDG_OK;
}

}  // extern "C"
