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
#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/rkdg.h"
#include "launch.hpp"
#include "model.hpp"
#include "writers.hpp"

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace rkdg {
// Sets the message rkdg_create_error() returns (also used by the mesh reader).
void set_create_error(const std::string& msg) { g_create_error = msg; }

// Polynomial orders the kernels are built for; the reference takes any p >= 0.
static bool order_built(int p) { return p >= 1 && p <= 3; }
}  // namespace rkdg

namespace {

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual rkdg_status init() = 0;
    virtual rkdg_status stage() = 0;
    // the stage in two parts (a partitioned run exchanges its halo between them): the stage
    // kernel (with the stepper advance), then the limiter (and an ensemble's stepper advance)
    virtual rkdg_status stage_begin() = 0;
    virtual rkdg_status stage_finish() = 0;
    bool mid_stage = false;      // between stage_begin and stage_finish

    virtual rkdg_status set_state(const double* q, const uint8_t* wet) = 0;
    virtual rkdg_status get_state(double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled) = 0;
    virtual rkdg_status nan_key(unsigned long long* key) = 0;
    virtual void collect(rkdg_perf_report* r) = 0;  // adds timed kernel milliseconds, recycles events
    // ensembles (merged models)
    virtual rkdg_status inst_get(int i, double* q, uint8_t* wet, uint8_t* lim, uint8_t* trb) = 0;
    virtual rkdg_status inst_set(int i, const double* q, const uint8_t* wet) = 0;
    virtual rkdg_status inst_time(int i, double* t, int* step, int* stage) = 0;
    virtual rkdg_status snapshot(bool restore) = 0;  // save / restore the step-boundary state
    // ensembles whose instances ramp their forcing: the per-stage ramps of the next nsteps steps
    // (nsteps < kRampRows), computed on the host and queued for the device stepper
    static constexpr int kRampRows = 128;
    virtual rkdg_status ramp_prepare(int nsteps) = 0;
    virtual bool ramps() const = 0;
    // learned troubled-cell indicator
    virtual rkdg_status set_indicator(const rkdg_mlp_weights* w, int judge) = 0;
    virtual rkdg_status indicator_stats(rkdg_indicator_stats* st, bool reset) = 0;
    virtual rkdg_status set_capture(bool on) = 0;
    virtual rkdg_status get_capture(float* feat, uint8_t* label, uint8_t* ok, float* logit, float* change) = 0;
    // halo exchange of a partitioned mesh
    virtual rkdg_status halo_define(int slot, const int32_t* idx, int n) = 0;
    virtual rkdg_status halo_buffer(int slot, void** buf, int64_t* bytes) = 0;
    virtual rkdg_status halo_move(int slot, bool pack, const void* src, int64_t src_bytes) = 0;
    // work-groups of this engine's persistent stage kernel resident on one CU at this block size
    virtual int persistent_resident(int block) const = 0;
    int block = 128;            // work-items per block (measured best for the stage kernels on 1M elements)
    int persistent_grid = 0;    // > 0: persistent stage kernel on this many work-groups (ensembles)
    bool capturing = false;     // inside a graph capture: no host-side event records
    std::string err;
    bool timing = false;
    int64_t stages_timed = 0;
};

const unsigned long long kNoNan = ~0ull;

template <typename Real>
struct Engine final : EngineBase {
    rkdg::Model& M;
    hipStream_t stream;
    std::vector<void*> allocs;
    std::vector<Real*> q;      // reference state slots 0..nstages
    std::vector<Real*> sol;    // per stage, only where a later stage reads it
    Real *bath = nullptr, *geom = nullptr, *egeom = nullptr, *vxy = nullptr, *gmn2 = nullptr;
    Real *sp_vol = nullptr, *sp_edge = nullptr;  // spherical correction (null: sp = 1)
    Real *famp = nullptr, *fphase = nullptr, *levee_gp = nullptr;
    Real* con = nullptr;  // [ncon][3] frequency, forcing factor, equilibrium argument
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
    // ensemble tables (M.ens.n > 0)
    int *tile_inst_d = nullptr, *inst_end_d = nullptr, *inst_step_d = nullptr, *inst_stage_d = nullptr;
    double *inst_tr_d = nullptr, *inst_t_d = nullptr, *inst_dt_d = nullptr, *inst_rd_d = nullptr, *drk_d = nullptr;
    Real* coef_d = nullptr;
    unsigned long long* step_ctr_d = nullptr;
    // ramped ensembles: ramp of each instance at each stage of kRampRows steps from ramp_base,
    // [row][stage][instance], filled on the host (pinned staging copy) by ramp_prepare
    bool ramp_on = false;
    double* ramp_tab_d = nullptr;
    long long* ramp_base_d = nullptr;
    double* ramp_tab_h = nullptr;  // kRampRows * ns * n values, then the base step
    std::vector<void*> snap;  // snapshot buffers (autotuning)
    // learned indicator: packed bf16 weights, vectors, judge counters, capture buffers
    bool ind_on = false, ind_judge = false, cap_on = false, cap_valid = false, cap_logits = false;
    uint16_t* ind_pack = nullptr;
    float* ind_vec = nullptr;
    unsigned long long* ind_cnt = nullptr;
    int *ind_list = nullptr, *ind_count = nullptr;
    int64_t ind_stages = 0;
    float *cap_feat = nullptr, *cap_logit = nullptr, *cap_change = nullptr;
    uint8_t *cap_label = nullptr, *cap_ok = nullptr;
    // timing
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> stage_ev, lim_ev;

    bool fast;  // fast form (FP32 bulk, FP64 fast) instead of the reference-order form

    hipError_t launch_plain(const rkdg::StageArgs<Real>& A) {
        if constexpr (sizeof(Real) == 8)
            return rkdg::launch_stage<Real, false>(M.p, M.prm.wetting_drying != 0, M.closed_forms, A, stream, block,
                                                   persistent_grid);
        else
            return hipErrorInvalidValue;  // FP32 exists only in the fast form
    }
    int persistent_resident(int b) const override {
        const bool wd = M.prm.wetting_drying != 0;
        if constexpr (sizeof(Real) == 8)
            return fast ? rkdg::stage_persistent_resident<Real, true>(M.p, wd, b)
                        : rkdg::stage_persistent_resident<Real, false>(M.p, wd, b);
        else
            return rkdg::stage_persistent_resident<Real, true>(M.p, wd, b);
    }
    hipError_t launch_fast(const rkdg::StageArgs<Real>& A) {
        return rkdg::launch_stage<Real, true>(M.p, M.prm.wetting_drying != 0, M.closed_forms, A, stream, block,
                                              persistent_grid);
    }

    Engine(rkdg::Model& m, hipStream_t s, bool fast_form) : M(m), stream(s), fast(fast_form) {
        if (const char* b = std::getenv("RKDG_BLOCK")) block = std::atoi(b);  // tuning experiments
        if (block < 64 || block > 256 || block % 64) block = 128;
    }
    // halo slots: listed local elements and a [3 ndof + 1][n] buffer each
    struct Halo {
        int n = 0;
        int* idx = nullptr;
        Real* buf = nullptr;
    };
    std::array<Halo, 64> halo{};

    rkdg_status halo_define(int slot, const int32_t* idx, int n) override {
        if (slot < 0 || slot >= (int)halo.size() || n < 0 || (n > 0 && !idx))
            return fail("bad halo slot", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (M.ens.n > 0) return fail("no halo exchange for ensembles", hipErrorInvalidValue), RKDG_ERR_ARG;
        for (int i = 0; i < n; ++i)
            if (idx[i] < 0 || idx[i] >= M.ne) return fail("halo element out of range", hipErrorInvalidValue), RKDG_ERR_ARG;
        Halo& h = halo[slot];
        if (h.idx) return fail("halo slot already defined", hipErrorInvalidValue), RKDG_ERR_ARG;
        const std::vector<int> hi(idx, idx + n);
        if (!upload(h.idx, hi) || !alloc(h.buf, (size_t)(3 * M.ndof + 1) * (n ? n : 1))) return RKDG_ERR_HIP;
        h.n = n;
        return RKDG_OK;
    }
    rkdg_status halo_buffer(int slot, void** buf, int64_t* bytes) override {
        if (slot < 0 || slot >= (int)halo.size() || !halo[slot].idx)
            return fail("halo slot not defined", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (buf) *buf = halo[slot].buf;
        if (bytes) *bytes = (int64_t)(3 * M.ndof + 1) * halo[slot].n * (int64_t)sizeof(Real);
        return RKDG_OK;
    }
    rkdg_status halo_move(int slot, bool pack, const void* src, int64_t src_bytes) override {
        if (slot < 0 || slot >= (int)halo.size() || !halo[slot].idx)
            return fail("halo slot not defined", hipErrorInvalidValue), RKDG_ERR_ARG;
        const Halo& h = halo[slot];
        if (!pack && !src) return fail("null halo buffer", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (!pack && src_bytes < (int64_t)(3 * M.ndof + 1) * h.n * (int64_t)sizeof(Real))
            return fail("halo buffer smaller than the slot's data", hipErrorInvalidValue), RKDG_ERR_ARG;
        Real* cur = q[M.stepper.stage];
        Real* buf = pack ? h.buf : const_cast<Real*>(static_cast<const Real*>(src));
        hipError_t e = rkdg::launch_halo<Real>(pack, M.ne, 3 * M.ndof, cur, wet[wcur], h.idx, h.n, buf, stream);
        return e == hipSuccess ? RKDG_OK : (fail("halo kernel launch", e), RKDG_ERR_HIP);
    }

    ~Engine() override {
        for (auto& pr : stage_ev) hipEventDestroy(pr.first), hipEventDestroy(pr.second);
        for (auto& pr : lim_ev) hipEventDestroy(pr.first), hipEventDestroy(pr.second);
        for (hipEvent_t ev : ev_pool) hipEventDestroy(ev);
        for (void* p : allocs) hipFree(p);
        if (ramp_tab_h) hipHostFree(ramp_tab_h);
        for (int i = 0; i < 2; ++i) {
            if (pin[i]) hipHostFree(pin[i]);
            if (pin_ev[i]) hipEventDestroy(pin_ev[i]);
        }
    }

    // State transfers between the caller's FP64 host arrays and a device state slot go through two
    // pinned staging buffers (allocated on first use): the device copy of one chunk runs while the
    // previous chunk is converted into (or out of) the caller's array, instead of a pageable copy
    // of the whole state into a freshly allocated host vector.
    static constexpr size_t kPinBytes = size_t(8) << 20;
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    hipError_t ensure_pinned() {
        for (int i = 0; i < 2; ++i) {
            if (!pin[i]) {
                const hipError_t e = hipHostMalloc(&pin[i], kPinBytes, hipHostMallocDefault);
                if (e != hipSuccess) return pin[i] = nullptr, e;
            }
            if (!pin_ev[i]) {
                const hipError_t e = hipEventCreateWithFlags(&pin_ev[i], hipEventDisableTiming);
                if (e != hipSuccess) return pin_ev[i] = nullptr, e;
            }
        }
        return hipSuccess;
    }
    // device -> host: chunk c is copied into buffer c % 2 while chunk c - 1 is converted; buffer
    // c % 2 was last read when chunk c - 2 was converted, before chunk c's copy is issued
    hipError_t read_state(const Real* dev, double* out, size_t n) {
        hipError_t e = ensure_pinned();
        if (e != hipSuccess) return e;
        const size_t per = kPinBytes / sizeof(Real), nch = (n + per - 1) / per;
        for (size_t c = 0; c <= nch && e == hipSuccess; ++c) {
            if (c < nch) {
                const size_t off = c * per, cnt = std::min(per, n - off);
                e = hipMemcpyAsync(pin[c & 1], dev + off, cnt * sizeof(Real), hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipEventRecord(pin_ev[c & 1], stream);
            }
            if (e == hipSuccess && c > 0) {
                const size_t pc = c - 1, off = pc * per, cnt = std::min(per, n - off);
                e = hipEventSynchronize(pin_ev[pc & 1]);
                const Real* src = static_cast<const Real*>(pin[pc & 1]);
                if (e == hipSuccess)
                    for (size_t i = 0; i < cnt; ++i) out[off + i] = (double)src[i];
            }
        }
        return e;
    }
    // host -> device: chunk c is converted into buffer c % 2 once that buffer's previous copy
    // (chunk c - 2, or a copy of an earlier call still in flight) has completed, then copied
    // while chunk c + 1 is converted; the copies are asynchronous on the context's stream
    hipError_t write_state(Real* dev, const double* in, size_t n) {
        hipError_t e = ensure_pinned();
        if (e != hipSuccess) return e;
        const size_t per = kPinBytes / sizeof(Real), nch = (n + per - 1) / per;
        for (size_t c = 0; c < nch && e == hipSuccess; ++c) {
            const size_t off = c * per, cnt = std::min(per, n - off);
            e = hipEventSynchronize(pin_ev[c & 1]);  // returns at once for an event never recorded
            Real* dst = static_cast<Real*>(pin[c & 1]);
            if (e == hipSuccess) {
                for (size_t i = 0; i < cnt; ++i) dst[i] = (Real)in[off + i];
                e = hipMemcpyAsync(dev + off, dst, cnt * sizeof(Real), hipMemcpyHostToDevice, stream);
            }
            if (e == hipSuccess) e = hipEventRecord(pin_ev[c & 1], stream);
        }
        return e;
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
        // every slot starts zeroed: ensemble padding elements are never written, and the limiter's
        // NaN scan reads them
        for (int s = 0; s <= ns; ++s)
            if (!alloc(q[s], S) || hipMemset(q[s], 0, S * sizeof(Real)) != hipSuccess) return RKDG_ERR_HIP;
        for (int s = 0; s < ns; ++s) {
            bool needed = false;  // a later stage combines this stage's M^-1 rhs
            for (int i = s + 1; i < ns; ++i) needed |= (M.stepper.brk[i][s] != 0.0);
            if (needed && (!alloc(sol[s], S) || hipMemset(sol[s], 0, S * sizeof(Real)) != hipSuccess))
                return RKDG_ERR_HIP;
        }
        if (!M.sp_vol.empty() && (!upload(sp_vol, M.sp_vol) || !upload(sp_edge, M.sp_edge))) return RKDG_ERR_HIP;
        if (!upload(bath, M.bath) || !upload(geom, M.geom) || !upload(egeom, M.egeom) || !upload(vxy, M.vxy) ||
            !upload(gmn2, M.gmn2) || !upload(famp, M.famp) || !upload(fphase, M.fphase) ||
            !upload(levee_gp, M.levee_gp) || !upload(nb, M.nb) ||
            !upload(edge, M.edge) || !upload(eforce, M.eforce) || !upload(meta, M.meta) ||
            !upload(neighbor, M.neighbor) || !upload(bath_vrtx, M.bath_vrtx))
            return RKDG_ERR_HIP;
        if (!upload(wet[0], M.wet0) || !upload(wet[1], M.wet0)) return RKDG_ERR_HIP;
        std::vector<double> cons;
        for (int c = 0; c < M.ncon; ++c) cons.push_back(M.freq[c]), cons.push_back(M.ffact[c]), cons.push_back(M.eqarg[c]);
        if (!upload(con, cons)) return RKDG_ERR_HIP;
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
        if (M.ens.n > 0 && !init_ensemble()) return RKDG_ERR_HIP;
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

    bool init_ensemble() {
        const auto& E = M.ens;
        const int n = E.n, ns = M.stepper.nstages;
        std::vector<int> end(n), zero_i(n, 0);
        std::vector<double> zero_d(n, 0.0), tr(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            end[i] = E.base[i] + E.ne[i];
            const bool no_ramp = std::abs(E.ramp_duration[i]) < std::numeric_limits<double>::epsilon() * 100;
            tr[2 * i] = 0.0 + E.dt[i] * M.stepper.drk[0];
            tr[2 * i + 1] = no_ramp ? 1.0 : 0.0;
        }
        // per stage and instance: cur_c, then term_c of the terms stage() lists, all dt_i * brk
        std::vector<double> coef((size_t)ns * n * 9, 0.0);
        for (int st = 0; st < ns; ++st)
            for (int i = 0; i < n; ++i) {
                double* c = &coef[((size_t)st * n + i) * 9];
                c[0] = E.dt[i] * M.stepper.brk[st][st];
                int j = 0;
                for (int k = 0; k < st; ++k) {
                    if (M.stepper.ark[st][k] == 0.0 && M.stepper.brk[st][k] == 0.0) continue;
                    c[1 + j++] = E.dt[i] * M.stepper.brk[st][k];
                }
            }
        if (!upload(tile_inst_d, E.tile_inst) || !upload(inst_end_d, end) || !upload(inst_dt_d, E.dt) ||
            !upload(inst_rd_d, E.ramp_duration) || !upload(drk_d, M.stepper.drk) || !upload(inst_t_d, zero_d) ||
            !upload(inst_step_d, zero_i) || !upload(inst_stage_d, zero_i) || !upload(inst_tr_d, tr) ||
            !upload(coef_d, coef) || !alloc(step_ctr_d, 1))
            return false;
        for (int i = 0; i < n; ++i)
            ramp_on |= !(std::abs(E.ramp_duration[i]) < std::numeric_limits<double>::epsilon() * 100);
        if (ramp_on) {
            const size_t nv = (size_t)kRampRows * ns * n;
            if (!alloc(ramp_tab_d, nv) || !alloc(ramp_base_d, 1)) return false;
            hipError_t e = hipHostMalloc((void**)&ramp_tab_h, (nv + 1) * sizeof(double), hipHostMallocDefault);
            if (e != hipSuccess) return fail("hipHostMalloc", e);
        }
        return hipMemset(step_ctr_d, 0, sizeof(unsigned long long)) == hipSuccess ||
               fail("hipMemset", hipGetLastError());
    }

    rkdg::GateArgs<Real> gate_args(const rkdg::LimiterArgs<Real>& L) const {
        rkdg::GateArgs<Real> G;
        std::memset(&G, 0, sizeof G);
        G.ne = M.ne, G.q = L.q, G.wet = L.wet, G.neighbor = neighbor, G.bath_bc = bath_bc, G.g = M.prm.g;
        if (cap_on) G.feat = cap_feat, G.ok = cap_ok;
        return G;
    }
    rkdg::JudgeArgs judge_args() const {
        rkdg::JudgeArgs J;
        std::memset(&J, 0, sizeof J);
        J.h_floor = rkdg::mlp::H_FLOOR, J.tol = sizeof(Real) == 4 ? rkdg::mlp::SIG_TOL_FP32 : rkdg::mlp::SIG_TOL;
        if (cap_on) J.label = cap_label, J.change = cap_change;
        return J;
    }
    // reference limiter: trouble check, dilation, limit (with capture: features and labels of the
    // state before the limit pass)
    hipError_t plain_limiter(const rkdg::LimiterArgs<Real>& L) {
        hipError_t e = rkdg::launch_limiter_prepare<Real>(M.p, L, stream, block);
        if (e == hipSuccess && cap_on) {
            e = rkdg::launch_gate<Real>(M.p, gate_args(L), false, stream);
            if (e == hipSuccess) e = rkdg::launch_limiter_judge<Real>(M.p, L, judge_args(), stream, block);
            cap_valid = true, cap_logits = false;
        }
        if (e == hipSuccess) e = rkdg::launch_limiter_limit<Real>(M.p, L, stream, block);
        return e;
    }
    // learned form: the gate writes the barycentres, the troubled (gate) flags and the list of
    // open elements; the limit pass runs over that list
    hipError_t learned_limiter(const rkdg::LimiterArgs<Real>& L) {
        rkdg::GateArgs<Real> G = gate_args(L);
        G.wpack = ind_pack, G.vec = ind_vec, G.baryctr = baryctr, G.flag = troubled;
        G.limited = limited, G.list = ind_list, G.list_count = ind_count, G.list_cap = rkdg::gate_grid(M.ne).cap;
        if (cap_on) G.logit = cap_logit;
        hipError_t e = rkdg::launch_gate<Real>(M.p, G, true, stream);
        if (e == hipSuccess && (ind_judge || cap_on)) {
            rkdg::JudgeArgs J = judge_args();
            J.flag = troubled;
            if (ind_judge) J.cnt = ind_cnt;
            e = rkdg::launch_limiter_judge<Real>(M.p, L, J, stream, block);
        }
        if (cap_on) cap_valid = true, cap_logits = true;
        if (e == hipSuccess) e = rkdg::launch_limiter_list<Real>(M.p, L, ind_list, ind_count, stream);
        ++ind_stages;
        return e;
    }

    rkdg_status set_indicator(const rkdg_mlp_weights* w, int judge) override {
        if (!w) {
            ind_on = false;
            return RKDG_OK;
        }
        if (!M.prm.slope_limiting) return fail("the indicator needs the slope limiter on", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (M.ens.n > 0) return fail("no indicator for ensembles", hipErrorInvalidValue), RKDG_ERR_ARG;
        namespace ml = rkdg::mlp;
        const float* theta = reinterpret_cast<const float*>(w);
        std::vector<uint16_t> pack(3 * (size_t)ml::PACK);
        for (int m = 0; m < 3; ++m)
            for (int i = 0; i < ml::PACK; ++i) {
                const int src = ml::pack_source(m, i);
                pack[(size_t)m * ml::PACK + i] = src < 0 ? (uint16_t)0 : ml::bf16_bits(theta[src]);
            }
        std::vector<float> vec(ml::NVEC);
        std::copy(w->b1, w->b1 + 64, vec.begin());
        std::copy(w->b2, w->b2 + 64, vec.begin() + 64);
        std::copy(w->b3, w->b3 + 64, vec.begin() + 128);
        std::copy(w->w4, w->w4 + 64, vec.begin() + 192);
        vec[256] = w->b4, vec[257] = w->threshold;
        if (!ind_pack && (!alloc(ind_pack, pack.size()) || !alloc(ind_vec, vec.size()) || !alloc(ind_cnt, 8) ||
                          !alloc(ind_list, (size_t)rkdg::gate_grid(M.ne).waves * rkdg::gate_grid(M.ne).cap) ||
                          !alloc(ind_count, (size_t)rkdg::gate_grid(M.ne).waves)))
            return RKDG_ERR_HIP;
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipMemcpy(ind_pack, pack.data(), pack.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ind_vec, vec.data(), vec.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(ind_cnt, 0, 8 * sizeof(unsigned long long));
        if (e != hipSuccess) return fail("set_indicator", e), RKDG_ERR_HIP;
        ind_on = true, ind_judge = judge != 0, ind_stages = 0;
        return RKDG_OK;
    }
    rkdg_status indicator_stats(rkdg_indicator_stats* st, bool reset) override {
        std::memset(st, 0, sizeof *st);
        if (!ind_cnt) return RKDG_OK;
        unsigned long long c[8];
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipMemcpy(c, ind_cnt, sizeof c, hipMemcpyDeviceToHost);
        if (e == hipSuccess && reset) e = hipMemset(ind_cnt, 0, sizeof c);
        if (e != hipSuccess) return fail("indicator stats", e), RKDG_ERR_HIP;
        st->stages = ind_stages;
        st->candidates = (int64_t)c[0], st->flagged = (int64_t)c[1], st->write_backs = (int64_t)c[2];
        st->significant = (int64_t)c[3], st->missed = (int64_t)c[4], st->missed_significant = (int64_t)c[5];
        std::memcpy(&st->max_missed_change, &c[6], sizeof(double));
        if (reset) ind_stages = 0;
        return RKDG_OK;
    }
    rkdg_status set_capture(bool on) override {
        if (on && !M.prm.slope_limiting) return fail("capture needs the slope limiter on", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (on && M.ens.n > 0) return fail("no capture for ensembles", hipErrorInvalidValue), RKDG_ERR_ARG;
        const size_t ne = M.ne;
        if (on && !cap_feat &&
            (!alloc(cap_feat, rkdg::mlp::NIN * ne) || !alloc(cap_logit, ne) || !alloc(cap_label, ne) || !alloc(cap_ok, ne) ||
             !alloc(cap_change, ne)))
            return RKDG_ERR_HIP;
        cap_on = on;
        return RKDG_OK;
    }
    rkdg_status get_capture(float* feat, uint8_t* label, uint8_t* ok, float* logit, float* change) override {
        if (!cap_valid) return fail("nothing captured yet", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (logit && !cap_logits) return fail("the last captured stage ran without an indicator", hipErrorInvalidValue), RKDG_ERR_ARG;
        const size_t ne = M.ne;
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess && feat) e = hipMemcpy(feat, cap_feat, rkdg::mlp::NIN * ne * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess && label) e = hipMemcpy(label, cap_label, ne, hipMemcpyDeviceToHost);
        if (e == hipSuccess && ok) e = hipMemcpy(ok, cap_ok, ne, hipMemcpyDeviceToHost);
        if (e == hipSuccess && logit) e = hipMemcpy(logit, cap_logit, ne * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess && change) e = hipMemcpy(change, cap_change, ne * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("get_capture", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }

    rkdg_status stage() override {
        rkdg_status r = stage_begin();
        return r == RKDG_OK ? stage_finish() : r;
    }
    unsigned long long mid_key = 0;  // stage_key of the stage begun
    bool mid_tm = false;

    rkdg_status stage_begin() override {
        if (mid_stage) return fail("stage_begin twice without stage_finish", hipErrorInvalidValue), RKDG_ERR_ARG;
        rkdg::Stepper& st = M.stepper;
        const int s = st.stage, ns = st.nstages;
        rkdg::StageArgs<Real> A;
        std::memset(&A, 0, sizeof A);
        const bool ens = M.ens.n > 0;
        A.ne = M.ne, A.nedge = M.nedge;
        A.q_s = q[s];
        // an ensemble writes its last stage in place into q0 (each element reads only its own q0),
        // so a captured step replays with fixed pointers
        A.q_out = (ens && s + 1 == ns) ? q[0] : q[s + 1];
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
        A.sp_vol = sp_vol, A.sp_edge = sp_edge;
        A.ncon = M.ncon;
        A.con = con;
        A.t = st.stage_time();
        A.ramp = st.ramp;
        A.g = (Real)M.prm.g;
        A.cf = (Real)M.prm.friction_cf;
        A.friction = M.prm.friction;
        A.fsrc = M.prm.function_source;
        A.problem = M.prm.problem;
        for (int i = 0; i < 4; ++i) A.pparam[i] = M.prm.problem_param[i];
        A.h_o = M.h_o, A.h_thr = M.h_thr;
        A.nan_check = (M.prm.slope_limiting && !ind_on) ? 0 : 1;  // the gated limiter scans only what it limits
        A.nan_key = d_nan;
        A.stage_key = (unsigned long long)st.step * ns + s;
        if (ens) {
            A.tile_inst = tile_inst_d, A.inst_end = inst_end_d, A.inst_tr = inst_tr_d;
            A.inst_coef = coef_d + (size_t)s * M.ens.n * 9;
            A.step_ctr = step_ctr_d, A.key_ns = ns, A.stage_key = (unsigned long long)s;
        }

        const bool tm = timing && !capturing;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (tm) e0 = take_event(), hipEventRecord(e0, stream);
        hipError_t e = (fast ? launch_fast(A) : launch_plain(A));
        if (e != hipSuccess) return fail("stage kernel launch", e), RKDG_ERR_HIP;
        if (tm) e1 = take_event(), hipEventRecord(e1, stream), stage_ev.push_back({e0, e1});
        ++stages_timed;

        if (!ens && s + 1 == ns) std::swap(q[0], q[ns]);  // UpdateState's final swap
        const unsigned long long key = A.stage_key;
        st.advance();
        if (M.prm.wetting_drying) {
            if (ens && s + 1 == ns && (ns & 1)) {
                // keep the step-start buffer fixed for graph replay: copy the new flags back
                e = rkdg::launch_copy_bytes((size_t)M.ne, wet[wcur ^ 1], wet[wcur], stream);
                if (e != hipSuccess) return fail("wet copy launch", e), RKDG_ERR_HIP;
            } else {
                wcur ^= 1;
            }
        }

        mid_key = key, mid_tm = tm, mid_stage = true;
        return RKDG_OK;
    }

    rkdg_status stage_finish() override {
        if (!mid_stage) return fail("stage_finish without stage_begin", hipErrorInvalidValue), RKDG_ERR_ARG;
        mid_stage = false;
        rkdg::Stepper& st = M.stepper;
        const int ns = st.nstages;
        const bool ens = M.ens.n > 0;
        const unsigned long long key = mid_key;
        const bool tm = mid_tm;
        hipError_t e = hipSuccess;
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
            if (ens) L.step_ctr = step_ctr_d, L.key_ns = ns;
            hipEvent_t l0 = nullptr, l1 = nullptr;
            if (tm) l0 = take_event(), hipEventRecord(l0, stream);
            e = ind_on ? learned_limiter(L) : plain_limiter(L);
            if (e != hipSuccess) return fail("limiter kernel launch", e), RKDG_ERR_HIP;
            if (tm) l1 = take_event(), hipEventRecord(l1, stream), lim_ev.push_back({l0, l1});
        }
        if (ens) {  // per-instance stepper advance, after the NaN keys of this stage are taken
            e = rkdg::launch_batch_advance(M.ens.n, ns, inst_t_d, inst_dt_d, inst_rd_d, inst_step_d, inst_stage_d, drk_d,
                                           inst_tr_d, step_ctr_d, ramp_tab_d, ramp_base_d, kRampRows, stream);
            if (e != hipSuccess) return fail("advance launch", e), RKDG_ERR_HIP;
        }
        return RKDG_OK;
    }

    // ---- ensemble instance access (at a step boundary: the state is in q0)
    rkdg_status inst_copy(int i, double* qh, const double* qin, uint8_t* wh, const uint8_t* win) {
        if (M.ens.n == 0 || i < 0 || i >= M.ens.n) return fail("bad instance", hipErrorInvalidValue), RKDG_ERR_ARG;
        if (M.stepper.stage != 0) return fail("instance access needs a step boundary", hipErrorInvalidValue), RKDG_ERR_ARG;
        const size_t base = M.ens.base[i], ne = M.ens.ne[i], rows = 3 * (size_t)M.ndof;
        hipError_t e = hipStreamSynchronize(stream);
        // each mode row of the instance is a contiguous run of ne values at `base` in a row of the
        // merged range: moved through the pinned staging buffers row by row
        for (size_t r = 0; r < rows && e == hipSuccess && (qh || qin); ++r) {
            Real* dev = q[0] + r * (size_t)M.ne + base;
            e = qin ? write_state(dev, qin + r * ne, ne) : read_state(dev, qh + r * ne, ne);
        }
        if (e == hipSuccess && qin) e = hipStreamSynchronize(stream);
        if (e == hipSuccess && win) e = hipMemcpy(wet[wcur] + base, win, ne, hipMemcpyHostToDevice);
        if (e == hipSuccess && wh) e = hipMemcpy(wh, wet[wcur] + base, ne, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("instance copy", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }
    rkdg_status inst_get(int i, double* qh, uint8_t* wh, uint8_t* lh, uint8_t* th) override {
        rkdg_status st = inst_copy(i, qh, nullptr, wh, nullptr);
        if (st != RKDG_OK) return st;
        const size_t base = M.ens.base[i], ne = M.ens.ne[i];
        hipError_t e = hipSuccess;
        if (lh) {
            if (limited)
                e = hipMemcpy2D(lh, ne, limited + base, M.ne, ne, 3, hipMemcpyDeviceToHost);
            else
                std::memset(lh, 0, 3 * ne);
        }
        if (th && e == hipSuccess) {
            if (troubled)
                e = hipMemcpy(th, troubled + base, ne, hipMemcpyDeviceToHost);
            else
                std::memset(th, 0, ne);
        }
        if (e != hipSuccess) return fail("instance flags", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }
    rkdg_status inst_set(int i, const double* qh, const uint8_t* wh) override {
        return inst_copy(i, nullptr, qh, nullptr, wh);
    }
    bool ramps() const override { return ramp_on; }
    // The ramp of every instance at every stage of steps step..step+nsteps, the arithmetic of the
    // host stepper (ESSPRKStepper::operator++ / GetRamp, explicit_ssp_rk_stepper.hpp:46-65: t +=
    // dt per step, stage time t + dt drk[s], tanh(2 (t_s / 86400) / ramp_duration)), so every
    // instance ramps bit for bit as its single-mesh run; the device stepper looks the values up.
    // Waits for the stream (the staging copy is reused and the device time is read).
    rkdg_status ramp_prepare(int nsteps) override {
        if (!ramp_on) return RKDG_OK;
        if (nsteps < 0 || nsteps >= kRampRows || M.stepper.stage != 0) return RKDG_ERR_ARG;
        const int n = M.ens.n, ns = M.stepper.nstages;
        std::vector<double> t(n);
        std::vector<int> step(n);
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipMemcpy(t.data(), inst_t_d, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(step.data(), inst_step_d, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("ramp table", e), RKDG_ERR_HIP;
        const size_t nv = (size_t)kRampRows * ns * n;
        for (int i = 0; i < n; ++i) {
            const double dt = M.ens.dt[i], rd = M.ens.ramp_duration[i];
            const bool no_ramp = std::abs(rd) < std::numeric_limits<double>::epsilon() * 100;
            double ti = t[i];
            for (int j = 0; j <= nsteps; ++j) {
                for (int k = 0; k < ns; ++k) {
                    const double ts = ti + dt * M.stepper.drk[k];
                    ramp_tab_h[((size_t)j * ns + k) * n + i] = no_ramp ? 1.0 : std::tanh(2 * (ts / 86400) / rd);
                }
                ti += dt;
            }
        }
        long long base = step[0];
        std::memcpy(ramp_tab_h + nv, &base, sizeof base);
        e = hipMemcpyAsync(ramp_tab_d, ramp_tab_h, (size_t)(nsteps + 1) * ns * n * sizeof(double), hipMemcpyHostToDevice,
                           stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ramp_base_d, ramp_tab_h + nv, sizeof base, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return fail("ramp table upload", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }
    rkdg_status inst_time(int i, double* t, int* step, int* stage) override {
        if (M.ens.n == 0 || i < 0 || i >= M.ens.n) return RKDG_ERR_ARG;
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess && t) e = hipMemcpy(t, inst_t_d + i, sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && step) e = hipMemcpy(step, inst_step_d + i, sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess && stage) e = hipMemcpy(stage, inst_stage_d + i, sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("instance time", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }
    rkdg_status snapshot(bool restore) override {
        if (M.ens.n == 0 || M.stepper.stage != 0) return RKDG_ERR_ARG;
        const size_t n = M.ens.n, S = 3 * (size_t)M.ndof * M.ne;
        struct Buf { void* dev; size_t bytes; };
        const Buf bufs[] = {{q[0], S * sizeof(Real)}, {wet[0], (size_t)M.ne}, {wet[1], (size_t)M.ne},
                            {inst_t_d, n * sizeof(double)}, {inst_tr_d, 2 * n * sizeof(double)},
                            {inst_step_d, n * sizeof(int)}, {inst_stage_d, n * sizeof(int)},
                            {step_ctr_d, sizeof(unsigned long long)}};
        const size_t nb = sizeof bufs / sizeof bufs[0];
        hipError_t e = hipStreamSynchronize(stream);
        if (!restore && snap.empty())
            for (size_t k = 0; k < nb && e == hipSuccess; ++k) {
                void* p = nullptr;
                e = hipMalloc(&p, bufs[k].bytes);
                if (e == hipSuccess) snap.push_back(p), allocs.push_back(p);
            }
        if (snap.size() != nb) return fail("snapshot", e == hipSuccess ? hipErrorInvalidValue : e), RKDG_ERR_HIP;
        for (size_t k = 0; k < nb && e == hipSuccess; ++k)
            e = restore ? hipMemcpy(bufs[k].dev, snap[k], bufs[k].bytes, hipMemcpyDeviceToDevice)
                        : hipMemcpy(snap[k], bufs[k].dev, bufs[k].bytes, hipMemcpyDeviceToDevice);
        if (e != hipSuccess) return fail("snapshot copy", e), RKDG_ERR_HIP;
        return RKDG_OK;
    }

    rkdg_status set_state(const double* qh, const uint8_t* wh) override {
        const size_t S = 3 * (size_t)M.ndof * M.ne;
        if (qh) {
            hipError_t e = write_state(q[M.stepper.stage], qh, S);
            // the new state replaces the one a recorded NaN was found in: clear the record
            if (e == hipSuccess) e = hipMemcpyAsync(d_nan, &kNoNan, sizeof kNoNan, hipMemcpyHostToDevice, stream);
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
            e = read_state(q[M.stepper.stage], qh, S);
            if (e != hipSuccess) return fail("get_state", e), RKDG_ERR_HIP;
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
    // ensemble: one merged model and engine per polynomial order
    std::vector<std::unique_ptr<rkdg::Model>> gm;
    std::vector<std::unique_ptr<EngineBase>> ge;
    std::vector<std::array<int, 2>> loc;  // instance -> (group, index in group)
    rkdg_schedule sched{128, 0, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool batch() const { return !ge.empty(); }
    int nstages() const { return batch() ? gm[0]->stepper.nstages : model.stepper.nstages; }
    void drop_graph() {
        if (gexec) hipGraphExecDestroy(gexec), gexec = nullptr;
        if (graph) hipGraphDestroy(graph), graph = nullptr;
    }
};

// Algorithmic HBM bytes of one stage launch averaged over the scheme's stages: every array element
// the kernel must read or write, each counted once; ne_real excludes ensemble padding.
static double stage_bytes(const rkdg::Model& M, rkdg_precision precision, double ne_real) {
    const double b = precision == RKDG_FP32_BULK ? 4.0 : 8.0;
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
    // modal bathymetry: every mode in the plain form, the three linear ones in the fast forms
    const int nbath = precision == RKDG_FP64_PLAIN ? M.ndof : std::min(M.ndof, 3);
    double per_elt = (reads_q + 1.0) * S + nbath * b + 5 * b + 28.0;
    if (M.prm.wetting_drying) per_elt += 2.0 + 24.0;
    return per_elt * ne_real + 3.0 * b * M.nedge;
}

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
        return params->p >= 0 && !rkdg::order_built(params->p) ? RKDG_ERR_UNSUPPORTED : RKDG_ERR_ARG;
    }
    // the mesh and parameters are checked first, so argument errors do not depend on a device
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device visible";
        return RKDG_ERR_NO_DEVICE;
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
    ctx->drop_graph();
    ctx->ge.clear();
    ctx->eng.reset();
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rkdg_last_error(const rkdg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rkdg_status rkdg_set_state(rkdg_ctx* ctx, const double* q, const uint8_t* wet) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    rkdg_status s = ctx->eng->set_state(q, wet);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_get_state(rkdg_ctx* ctx, double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    rkdg_status s = ctx->eng->get_state(q, wet, limited, troubled);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

static rkdg_status check_nan(rkdg_ctx* ctx) {
    std::vector<EngineBase*> engines;
    if (ctx->batch())
        for (auto& g : ctx->ge) engines.push_back(g.get());
    else
        engines.push_back(ctx->eng.get());
    for (EngineBase* g : engines) {
        unsigned long long key = kNoNan;
        rkdg_status s = g->nan_key(&key);
        if (s != RKDG_OK) return set_err(ctx, s, g->err);
        if (key != kNoNan) return set_err(ctx, RKDG_ERR_NAN, "NaN found in the state");
    }
    return RKDG_OK;
}

// One ensemble step: every stage of every order group, eagerly or by replaying the captured graph.
static rkdg_status batch_capture(rkdg_ctx* ctx) {
    ctx->drop_graph();
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return set_err(ctx, RKDG_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    rkdg_status st = RKDG_OK;
    const int ns = ctx->nstages();
    for (int k = 0; k < ns && st == RKDG_OK; ++k)
        for (auto& g : ctx->ge) {
            g->capturing = true;
            st = g->stage();
            g->capturing = false;
            if (st != RKDG_OK) {
                set_err(ctx, st, g->err);
                break;
            }
        }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != RKDG_OK) {
        if (graph) hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return set_err(ctx, RKDG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    ctx->graph = graph;
    e = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        ctx->drop_graph();
        return set_err(ctx, RKDG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return RKDG_OK;
}

static rkdg_status batch_steps_chunk(rkdg_ctx* ctx, int nsteps);

// Ensemble steps.  Instances that ramp their forcing take the ramp of each stage from a table the
// host fills for up to kRampRows - 1 steps at a time (EngineBase::ramp_prepare).
static rkdg_status batch_steps(rkdg_ctx* ctx, int nsteps) {
    bool ramps = false;
    for (auto& g : ctx->ge) ramps |= g->ramps();
    if (!ramps) return batch_steps_chunk(ctx, nsteps);
    while (nsteps > 0) {
        const int n = std::min(nsteps, EngineBase::kRampRows - 1);
        for (auto& g : ctx->ge) {
            rkdg_status st = g->ramp_prepare(n);
            if (st != RKDG_OK) return set_err(ctx, st, g->err);
        }
        rkdg_status st = batch_steps_chunk(ctx, n);
        if (st != RKDG_OK) return st;
        nsteps -= n;
    }
    return RKDG_OK;
}

static rkdg_status batch_steps_chunk(rkdg_ctx* ctx, int nsteps) {
    if (ctx->sched.graph) {
        if (!ctx->gexec) {
            rkdg_status st = batch_capture(ctx);
            if (st != RKDG_OK) return st;
        }
        for (int i = 0; i < nsteps; ++i) {
            hipError_t e = hipGraphLaunch(ctx->gexec, ctx->stream);
            if (e != hipSuccess) return set_err(ctx, RKDG_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        }
        return RKDG_OK;
    }
    const int ns = ctx->nstages();
    for (int i = 0; i < nsteps; ++i)
        for (int k = 0; k < ns; ++k)
            for (auto& g : ctx->ge) {
                rkdg_status st = g->stage();
                if (st != RKDG_OK) return set_err(ctx, st, g->err);
            }
    return RKDG_OK;
}

rkdg_status rkdg_stage(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    rkdg_status s = ctx->eng->stage();
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    return check_nan(ctx);
}

rkdg_status rkdg_stage_begin(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    rkdg_status s = ctx->eng->stage_begin();
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_stage_finish(rkdg_ctx* ctx) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    rkdg_status s = ctx->eng->stage_finish();
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    return check_nan(ctx);
}

rkdg_status rkdg_halo_define(rkdg_ctx* ctx, int32_t slot, const int32_t* local_elements, int32_t n) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no halo exchange");
    rkdg_status s = ctx->eng->halo_define(slot, local_elements, n);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_halo_buffer(rkdg_ctx* ctx, int32_t slot, void** device_buffer, int64_t* bytes) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no halo exchange");
    rkdg_status s = ctx->eng->halo_buffer(slot, device_buffer, bytes);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_halo_pack(rkdg_ctx* ctx, int32_t slot) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no halo exchange");
    rkdg_status s = ctx->eng->halo_move(slot, true, nullptr, 0);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_halo_unpack(rkdg_ctx* ctx, int32_t slot, const void* device_buffer, int64_t bytes) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no halo exchange");
    rkdg_status s = ctx->eng->halo_move(slot, false, device_buffer, bytes);
    return s == RKDG_OK ? s : set_err(ctx, s, ctx->eng->err);
}

rkdg_status rkdg_step(rkdg_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return RKDG_ERR_ARG;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (ctx->timing) {
        hipEventCreate(&t0), hipEventCreate(&t1);
        hipEventRecord(t0, ctx->stream);
    }
    if (ctx->batch()) {
        rkdg_status s = batch_steps(ctx, nsteps);
        if (s != RKDG_OK) return s;
    } else {
        const int ns = ctx->model.stepper.nstages;
        for (int i = 0; i < nsteps; ++i)
            for (int k = 0; k < ns; ++k) {
                rkdg_status s = ctx->eng->stage();
                if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
            }
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
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    if (t) *t = ctx->model.stepper.t;
    if (step) *step = ctx->model.stepper.step;
    if (stage) *stage = ctx->model.stepper.stage;
    return RKDG_OK;
}

rkdg_status rkdg_set_time(rkdg_ctx* ctx, double t, int32_t step) {
    if (!ctx || step < 0) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
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
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
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
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: use the rkdg_batch_* calls");
    const rkdg::Model& M = ctx->model;
    std::vector<double> q(3 * (size_t)M.ndof * M.ne);
    rkdg_status s = ctx->eng->get_state(q.data(), nullptr, nullptr, nullptr);
    if (s != RKDG_OK) return set_err(ctx, s, ctx->eng->err);
    *l2 = rkdg::l2_residual(M, q, M.stepper.stage_time());
    return RKDG_OK;
}

rkdg_status rkdg_reset_timers(rkdg_ctx* ctx) { return rkdg_set_timing(ctx, 1); }

rkdg_status rkdg_set_timing(rkdg_ctx* ctx, int32_t per_kernel) {
    if (!ctx) return RKDG_ERR_ARG;
    hipStreamSynchronize(ctx->stream);
    rkdg_perf_report scratch{};
    std::vector<EngineBase*> engines;
    if (ctx->batch())
        for (auto& g : ctx->ge) engines.push_back(g.get());
    else
        engines.push_back(ctx->eng.get());
    for (EngineBase* g : engines) {
        g->collect(&scratch);
        g->timing = per_kernel != 0;
        g->stages_timed = 0;
    }
    ctx->acc = rkdg_perf_report{};
    ctx->timing = true;
    ctx->total_ms = 0;
    return RKDG_OK;
}

rkdg_status rkdg_report(rkdg_ctx* ctx, rkdg_perf_report* r) {
    if (!ctx || !r) return RKDG_ERR_ARG;
    if (ctx->batch()) {
        int64_t staged = 0, ne = 0;
        double bytes = 0;
        for (size_t k = 0; k < ctx->ge.size(); ++k) {
            ctx->ge[k]->collect(&ctx->acc);
            staged = std::max(staged, ctx->ge[k]->stages_timed);
            const rkdg::Model& M = *ctx->gm[k];
            int64_t real = 0;
            for (int n : M.ens.ne) real += n;
            ne += real;
            bytes += stage_bytes(M, ctx->precision, (double)real);
        }
        *r = ctx->acc;
        r->n_elements = ne;
        r->p = ctx->gm.size() == 1 ? ctx->gm[0]->p : 0;
        r->ndof = ctx->gm.size() == 1 ? ctx->gm[0]->ndof : 0;
        r->nstages = ctx->nstages();
        r->precision = ctx->precision;
        r->stages_timed = staged;
        r->total_ms = ctx->total_ms;
        r->stage_bytes_alg = bytes;  // one stage of every group
        return RKDG_OK;
    }
    const rkdg::Model& M = ctx->model;
    ctx->eng->collect(&ctx->acc);
    const int64_t staged = ctx->eng->stages_timed;
    *r = ctx->acc;
    r->n_elements = M.ne;
    r->p = M.p, r->ndof = M.ndof, r->nstages = M.stepper.nstages;
    r->precision = ctx->precision;
    r->stages_timed = staged;
    r->total_ms = ctx->total_ms;
    r->stage_bytes_alg = stage_bytes(M, ctx->precision, (double)M.ne);
    return RKDG_OK;
}

// ---- ensembles
rkdg_status rkdg_create_batch(int32_t n, const rkdg_mesh_desc* meshes, const rkdg_params* params,
                              rkdg_precision precision, rkdg_ctx** out) {
    g_create_error.clear();
    if (n < 1 || !meshes || !params || !out) {
        g_create_error = "null argument or empty batch";
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    if (precision != RKDG_FP64_PLAIN && precision != RKDG_FP32_BULK && precision != RKDG_FP64_FAST) {
        g_create_error = "unknown precision";
        return RKDG_ERR_ARG;
    }
    std::unique_ptr<rkdg_ctx> c(new (std::nothrow) rkdg_ctx());
    if (!c) return RKDG_ERR_ARG;
    c->precision = precision;
    // build each instance, then merge per polynomial order
    std::vector<std::unique_ptr<rkdg::Model>> parts(n);
    std::map<int, std::vector<int>> by_p;
    for (int i = 0; i < n; ++i) {
        parts[i].reset(new rkdg::Model());
        std::string msg;
        try {
            msg = rkdg::build_model(meshes[i], params[i], *parts[i], precision == RKDG_FP64_PLAIN);
        } catch (const std::exception& ex) {
            msg = ex.what();
        }
        if (msg.empty() && !(params[i].dt > 0)) msg = "dt must be positive";
        if (!msg.empty()) {
            g_create_error = "instance " + std::to_string(i) + ": " + msg;
            return params[i].p >= 0 && !rkdg::order_built(params[i].p) ? RKDG_ERR_UNSUPPORTED : RKDG_ERR_ARG;
        }
        by_p[params[i].p].push_back(i);
    }
    c->loc.assign(n, {0, 0});
    for (auto& kv : by_p) {
        std::vector<const rkdg::Model*> ms;
        for (int i : kv.second) {
            c->loc[i] = {(int)c->gm.size(), (int)ms.size()};
            ms.push_back(parts[i].get());
        }
        std::unique_ptr<rkdg::Model> g(new rkdg::Model());
        std::string msg = rkdg::merge_models(ms, *g);
        if (!msg.empty()) {
            g_create_error = msg;
            return RKDG_ERR_ARG;
        }
        for (int i : kv.second) parts[i].reset();  // free the instance copies as groups complete
        c->gm.push_back(std::move(g));
    }
    for (size_t k = 1; k < c->gm.size(); ++k)
        if (c->gm[k]->stepper.nstages != c->gm[0]->stepper.nstages) {
            g_create_error = "all instances need the same number of Runge-Kutta stages";
            return RKDG_ERR_ARG;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device visible";
        return RKDG_ERR_NO_DEVICE;
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        return RKDG_ERR_HIP;
    }
    for (auto& g : c->gm) {
        std::unique_ptr<EngineBase> eng;
        if (precision == RKDG_FP32_BULK)
            eng.reset(new Engine<float>(*g, c->stream, true));
        else
            eng.reset(new Engine<double>(*g, c->stream, precision == RKDG_FP64_FAST));
        rkdg_status s = eng->init();
        if (s != RKDG_OK) {
            g_create_error = eng->err;
            c->ge.clear();
            hipStreamDestroy(c->stream);
            return s;
        }
        eng->block = c->sched.block;
        c->ge.push_back(std::move(eng));
    }
    *out = c.release();
    return RKDG_OK;
}

int32_t rkdg_batch_size(const rkdg_ctx* ctx) { return ctx ? (int32_t)ctx->loc.size() : 0; }

static EngineBase* batch_engine(rkdg_ctx* ctx, int32_t i, int* local) {
    if (!ctx || !ctx->batch() || i < 0 || i >= (int32_t)ctx->loc.size()) return nullptr;
    *local = ctx->loc[i][1];
    return ctx->ge[ctx->loc[i][0]].get();
}

rkdg_status rkdg_batch_get_state(rkdg_ctx* ctx, int32_t instance, double* q, uint8_t* wet, uint8_t* limited,
                                 uint8_t* troubled) {
    int li = 0;
    EngineBase* g = batch_engine(ctx, instance, &li);
    if (!g) return ctx ? set_err(ctx, RKDG_ERR_ARG, "not a batch context or bad instance") : RKDG_ERR_ARG;
    rkdg_status s = g->inst_get(li, q, wet, limited, troubled);
    return s == RKDG_OK ? s : set_err(ctx, s, g->err);
}

rkdg_status rkdg_batch_set_state(rkdg_ctx* ctx, int32_t instance, const double* q, const uint8_t* wet) {
    int li = 0;
    EngineBase* g = batch_engine(ctx, instance, &li);
    if (!g) return ctx ? set_err(ctx, RKDG_ERR_ARG, "not a batch context or bad instance") : RKDG_ERR_ARG;
    rkdg_status s = g->inst_set(li, q, wet);
    return s == RKDG_OK ? s : set_err(ctx, s, g->err);
}

rkdg_status rkdg_batch_get_time(rkdg_ctx* ctx, int32_t instance, double* t, int32_t* step, int32_t* stage) {
    int li = 0;
    EngineBase* g = batch_engine(ctx, instance, &li);
    if (!g) return ctx ? set_err(ctx, RKDG_ERR_ARG, "not a batch context or bad instance") : RKDG_ERR_ARG;
    int st = 0, sg = 0;
    rkdg_status s = g->inst_time(li, t, &st, &sg);
    if (step) *step = st;
    if (stage) *stage = sg;
    return s == RKDG_OK ? s : set_err(ctx, s, g->err);
}

rkdg_status rkdg_batch_nan_info(rkdg_ctx* ctx, int32_t* instance, int32_t* element, int32_t* step, int32_t* stage) {
    if (!ctx || !ctx->batch()) return RKDG_ERR_ARG;
    if (instance) *instance = -1;
    if (element) *element = -1;
    if (step) *step = -1;
    if (stage) *stage = -1;
    const int ns = ctx->nstages();
    unsigned long long best = kNoNan;
    int best_g = -1;
    for (size_t k = 0; k < ctx->ge.size(); ++k) {
        unsigned long long key = kNoNan;
        rkdg_status s = ctx->ge[k]->nan_key(&key);
        if (s != RKDG_OK) return set_err(ctx, s, ctx->ge[k]->err);
        if (key < best) best = key, best_g = (int)k;
    }
    if (best_g < 0) return RKDG_OK;
    const rkdg::Model& M = *ctx->gm[best_g];
    const int e = (int)(best & 0xffffffffu);
    for (int li = 0; li < M.ens.n; ++li)
        if (e >= M.ens.base[li] && e < M.ens.base[li] + M.ens.ne[li]) {
            for (size_t i = 0; i < ctx->loc.size(); ++i)
                if (ctx->loc[i][0] == best_g && ctx->loc[i][1] == li && instance) *instance = (int32_t)i;
            if (element) *element = e - M.ens.base[li];
        }
    if (step) *step = (int32_t)((best >> 32) / ns);
    if (stage) *stage = (int32_t)((best >> 32) % ns);
    return RKDG_OK;
}

static bool schedule_valid(const rkdg_schedule& s) {
    return (s.block == 64 || s.block == 128 || s.block == 256) && s.groups_per_xcd >= -1 &&
           s.groups_per_xcd <= 4096 && (s.graph == 0 || s.graph == 1);
}

rkdg_status rkdg_batch_set_schedule(rkdg_ctx* ctx, const rkdg_schedule* sc) {
    if (!ctx || !sc) return RKDG_ERR_ARG;
    if (!ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "not a batch context");
    if (!schedule_valid(*sc)) return set_err(ctx, RKDG_ERR_ARG, "invalid schedule");
    for (auto& g : ctx->gm)
        if (sc->groups_per_xcd != 0 && g->closed_forms)
            return set_err(ctx, RKDG_ERR_ARG, "the persistent kernel is not built for closed-form problems");
    // -1: each order's persistent kernel on as many work-groups as are resident at once on every CU
    std::vector<int> grid(ctx->ge.size());
    for (size_t k = 0; k < ctx->ge.size(); ++k) {
        if (sc->groups_per_xcd >= 0) {
            grid[k] = 8 * sc->groups_per_xcd;
            continue;
        }
        int dev = 0, cu = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
        const int r = ctx->ge[k]->persistent_resident(sc->block);
        if (r < 1 || cu < 8) return set_err(ctx, RKDG_ERR_HIP, "occupancy of the persistent stage kernel unknown");
        grid[k] = 8 * (cu / 8) * r;
    }
    hipStreamSynchronize(ctx->stream);
    ctx->sched = *sc;
    for (size_t k = 0; k < ctx->ge.size(); ++k) ctx->ge[k]->block = sc->block, ctx->ge[k]->persistent_grid = grid[k];
    ctx->drop_graph();
    return RKDG_OK;
}

rkdg_status rkdg_batch_get_schedule(const rkdg_ctx* ctx, rkdg_schedule* sc) {
    if (!ctx || !sc || !ctx->batch()) return RKDG_ERR_ARG;
    *sc = ctx->sched;
    return RKDG_OK;
}

rkdg_status rkdg_batch_autotune(rkdg_ctx* ctx, int32_t steps_per_trial, rkdg_schedule* best, double* best_ms) {
    if (!ctx || steps_per_trial < 1) return RKDG_ERR_ARG;
    if (!ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "not a batch context");
    for (auto& g : ctx->ge) {
        rkdg_status s = g->snapshot(false);
        if (s != RKDG_OK) return set_err(ctx, s, "autotune needs a step boundary");
    }
    bool cf = false;
    for (auto& g : ctx->gm) cf |= g->closed_forms;
    // one work-group per tile, or the persistent kernel on the resident work-groups (-1)
    std::vector<rkdg_schedule> cand;
    for (int b : {64, 128, 256})
        for (int pers : {0, -1})
            for (int gr : {0, 1}) {
                if (pers && cf) continue;
                cand.push_back(rkdg_schedule{b, pers, gr});
            }
    const bool was_timing = ctx->timing;
    rkdg_schedule win = ctx->sched;
    double win_ms = 1e300;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipEventCreate(&t0), hipEventCreate(&t1);
    rkdg_status st = RKDG_OK;
    for (const rkdg_schedule& sc : cand) {
        if ((st = rkdg_batch_set_schedule(ctx, &sc)) != RKDG_OK) break;
        for (auto& g : ctx->ge)
            if ((st = g->snapshot(true)) != RKDG_OK) break;
        if (st != RKDG_OK) break;
        ctx->timing = false;
        if ((st = batch_steps(ctx, 1)) != RKDG_OK) break;  // warm-up (and graph capture)
        hipEventRecord(t0, ctx->stream);
        if ((st = batch_steps(ctx, steps_per_trial)) != RKDG_OK) break;
        hipEventRecord(t1, ctx->stream);
        hipEventSynchronize(t1);
        float ms = 0;
        hipEventElapsedTime(&ms, t0, t1);
        const double per = ms / steps_per_trial;
        if (per < win_ms) win_ms = per, win = sc;
    }
    hipEventDestroy(t0), hipEventDestroy(t1);
    ctx->timing = was_timing;
    if (st == RKDG_OK) st = check_nan(ctx);
    for (auto& g : ctx->ge) g->snapshot(true);
    rkdg_status s2 = rkdg_batch_set_schedule(ctx, &win);
    if (st != RKDG_OK) return st;
    if (s2 != RKDG_OK) return s2;
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
    return RKDG_OK;
}

rkdg_status rkdg_set_indicator(rkdg_ctx* ctx, const rkdg_mlp_weights* w, int32_t judge) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no learned indicator");
    rkdg_status st = ctx->eng->set_indicator(w, judge);
    if (st != RKDG_OK) ctx->err = ctx->eng->err;
    return st;
}

rkdg_status rkdg_get_indicator_stats(rkdg_ctx* ctx, rkdg_indicator_stats* stats, int32_t reset) {
    if (!ctx || !stats) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no learned indicator");
    rkdg_status st = ctx->eng->indicator_stats(stats, reset != 0);
    if (st != RKDG_OK) ctx->err = ctx->eng->err;
    return st;
}

rkdg_status rkdg_capture(rkdg_ctx* ctx, int32_t on) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no capture");
    rkdg_status st = ctx->eng->set_capture(on != 0);
    if (st != RKDG_OK) ctx->err = ctx->eng->err;
    return st;
}

rkdg_status rkdg_get_capture(rkdg_ctx* ctx, float* features, uint8_t* label, uint8_t* may_act, float* logit,
                             float* change) {
    if (!ctx) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: no capture");
    rkdg_status st = ctx->eng->get_capture(features, label, may_act, logit, change);
    if (st != RKDG_OK) ctx->err = ctx->eng->err;
    return st;
}

// Writes the current state at a step boundary (the reference writes state[0] from
// Writer::WriteOutput, simulation/writer.hpp:121-196).
enum class OutputKind { modal, vtu, vtk };

static rkdg_status write_output(rkdg_ctx* ctx, const char* prefix, OutputKind kind) {
    if (!ctx || !prefix) return RKDG_ERR_ARG;
    if (ctx->batch()) return set_err(ctx, RKDG_ERR_ARG, "batch context: write each instance from its own context");
    const rkdg::Model& M = ctx->model;
    if (M.stepper.stage != 0) return set_err(ctx, RKDG_ERR_ARG, "output is written at a step boundary");
    const size_t ne = M.ne;
    std::vector<double> q(3 * (size_t)M.ndof * ne);
    std::vector<uint8_t> wet(ne);
    rkdg_status st = ctx->eng->get_state(q.data(), wet.data(), nullptr, nullptr);
    if (st != RKDG_OK) return set_err(ctx, st, ctx->eng->err);
    std::string err;
    const int step = (int)M.stepper.step;
    const bool ok = kind == OutputKind::vtu   ? rkdg::write_vtu(M, q.data(), wet.data(), step, prefix, err)
                    : kind == OutputKind::vtk ? rkdg::write_vtk(M, q.data(), wet.data(), step, prefix, err)
                                              : rkdg::write_modal(M, q.data(), M.stepper.stage_time(), step, prefix, err);
    return ok ? RKDG_OK : set_err(ctx, RKDG_ERR_ARG, err);
}

rkdg_status rkdg_write_modal(rkdg_ctx* ctx, const char* prefix) { return write_output(ctx, prefix, OutputKind::modal); }

rkdg_status rkdg_write_vtu(rkdg_ctx* ctx, const char* prefix) { return write_output(ctx, prefix, OutputKind::vtu); }

rkdg_status rkdg_write_vtk(rkdg_ctx* ctx, const char* prefix) { return write_output(ctx, prefix, OutputKind::vtk); }

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
