// C ABI of the LMM engine: context management, model tables, product lowering,
// kernel orchestration and error mapping. See include/finmath_lmm.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fm_internal.hpp"
#include "hip_check.hpp"

namespace fm {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// Set while an entry point that must not allocate runs on this thread: a context buffer
// that would have to grow is an error instead of a hipMalloc.
static thread_local bool g_no_alloc = false;

struct NoAllocScope {
    NoAllocScope() { g_no_alloc = true; }
    ~NoAllocScope() { g_no_alloc = false; }
};

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void* get(size_t need) {
        if (need == 0) need = 8;
        if (need > bytes) {
            if (g_no_alloc)
                throw FmError(FM_ERR_STATE, "the context would have to allocate device memory (" + std::to_string(need) +
                                            " bytes): call the allocating entry point once with this model first");
            if (p) { HIP_CHECK(hipFree(p)); p = nullptr; bytes = 0; }
            HIP_CHECK(hipMalloc(&p, need));
            bytes = need;
        }
        return p;
    }
    template <class T> T* as() const { return (T*)p; }
};

template <class T>
static T* upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    T* d = (T*)b.get(v.size() * sizeof(T));
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

// Device memory handed in by the caller for the per-call buffers of one pass (wave
// partials, product tables, reduction output), taken in 256-byte aligned pieces.
struct ExternalScratch {
    char* base = nullptr;
    size_t bytes = 0, used = 0;
    static size_t piece(size_t n) { return (std::max<size_t>(n, 8) + 255) & ~(size_t)255; }
    void* take(size_t n) {
        const size_t need = piece(n);
        if (used + need > bytes) throw FmError(FM_ERR_INTERNAL, "scratch carve-up exceeds its checked size");
        void* p = base + used;
        used += need;
        return p;
    }
    template <class T> T* upload(const std::vector<T>& v, hipStream_t s) {
        T* d = (T*)take(v.size() * sizeof(T));
        if (!v.empty()) HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        return d;
    }
};

// TimeDiscretizationFromArray tick rounding (time/TimeDiscretizationFromArray.java:387-389).
static double round_tick(double t) {
    const double tick = 1.0 / (365.0 * 24.0);
    return std::rint(t / tick) * tick;
}
// java.util.Arrays.binarySearch semantics (index, or -(insertion point) - 1).
static int jsearch(const std::vector<double>& a, double key) {
    int lo = 0, hi = (int)a.size() - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (a[(size_t)mid] < key) lo = mid + 1;
        else if (a[(size_t)mid] > key) hi = mid - 1;
        else return mid;
    }
    return -(lo + 1);
}

} // namespace fm

using namespace fm;

struct fm_lmm {
    std::mutex mu;
    fm_lmm_config cfg{};
    std::vector<double> times, tenor;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    double timings[5] = {0, 0, 0, 0, 0};

    // model inputs
    std::vector<double> L0, lam, var, pd;
    bool has_L0 = false, has_cov = false, has_pd = false;
    std::vector<double> vf_sig, vf_fac, vf_c;       // volatility factorization for the adjoint
    bool has_vf = false;
    std::vector<double> lvb;                        // local-volatility offsets b_ij [nT][n]
    double lv_alpha = 0.0;
    bool has_lv = false;
    bool has_sv = false;                            // stochastic volatility scaling in d_svs [nT][P]
    bool steps_ready = false;                       // d_times, d_sqrt_dt uploaded
    bool tables_dirty = true;

    // derived host tables
    std::vector<int32_t> first_alive, tenor_ti, row_of;
    int num_rows = 0;
    std::vector<double> init_value;     // time-0 process value per component
    double n0_terminal = 1.0;           // N(T_0) under the terminal measure (deterministic)

    // generator
    StreamLayout lay;
    bool states_ready = false;
    bool dw_ready = false;
    bool simulated = false;
    std::vector<double> zb;             // numeraire-adjustment averages of the last simulate
    bool uniform_dt = false;            // all simulation steps have the same length
    std::vector<std::vector<double>> berm_fixing;   // fixing dates of the Bermudans of the last pass

    DevBuf d_times, d_first, d_tenor_ti, d_rec, d_delta, d_L0, d_Y0, d_Y0cap, d_sqrt_dt, d_lvb, d_svs;
    DevBuf d_states, d_poly, d_dw, d_row_of, d_paths, d_num;
    DevBuf d_products, d_periods, d_wave_sum, d_wave_m2, d_contrib, d_out, d_swapmeta, d_scale;
    DevBuf d_ckpt, d_ckpt_first, d_seg, d_adj_num, d_kcur, d_fsig, d_sig, d_cdiag, d_wave_grad, d_grad_out;
    DevBuf d_berm_vals, d_berm_under, d_berm_ex, d_berm_part, d_berm_coef, d_berm_col, d_berm_scale;
    // the Bermudan induction's adjoint: triggers and choices recorded by the pricing pass
    // ([exercise dates of all Bermudans][P]), the per-path adjoint state, the period weights
    DevBuf d_berm_trig, d_berm_diff, d_berm_vb, d_berm_ub, d_berm_weight, d_berm_bcoef, d_berm_wcol;
    DevBuf d_berm_wsum, d_berm_wm2, d_berm_stat, d_berm_pinv;
    bool berm_record = false;                 // the next pricing pass records triggers and choices
    std::vector<int> berm_trig_off;           // first recorded date of each Bermudan
    int dirac_method = FM_DIRAC_DISCRETE_DELTA;
    double dirac_width = 0.05;

    int n() const { return cfg.num_libors; }
    int nT() const { return cfg.num_time_steps; }
    int F() const { return cfg.num_factors; }
    int P() const { return cfg.num_paths; }
    bool fp32() const { return cfg.precision == FM_PRECISION_FP32_MIXED; }
    bool terminal() const { return cfg.measure == FM_MEASURE_TERMINAL; }
    bool pc() const { return cfg.scheme == FM_SCHEME_PREDICTOR_CORRECTOR || cfg.scheme == FM_SCHEME_PREDICTOR_CORRECTOR_FUNCTIONAL; }
    // the state is carried between steps (EULER, PREDICTOR_CORRECTOR); the functional
    // schemes re-derive it from the value (EulerSchemeFromProcessModel.java:234-237)
    bool carry() const { return cfg.scheme == FM_SCHEME_EULER || cfg.scheme == FM_SCHEME_PREDICTOR_CORRECTOR; }
    // carrying the state is the functional scheme when the inverse transform is the
    // identity and no cap applies
    bool carry_matters() const {
        return carry() && (cfg.state_space == FM_STATE_SPACE_LOGNORMAL || !std::isinf(cfg.libor_cap));
    }
    // spot measure, functional (or equivalent) Euler, loadings without local volatility:
    // the kernel the adjoint differentiates
    bool plain_euler() const { return !terminal() && !pc() && !carry_matters() && !has_lv && !has_sv; }
    size_t real_size() const { return fp32() ? sizeof(float) : sizeof(double); }
    int process_time_index(double t) const { return jsearch(times, round_tick(t)); }
    int libor_index(double t) const { return jsearch(tenor, round_tick(t)); }
};

namespace {

int fail(const std::exception& e) {
    if (auto* fe = dynamic_cast<const FmError*>(&e)) { set_last_error(fe->what()); return fe->code; }
    if (dynamic_cast<const std::bad_alloc*>(&e)) { set_last_error("host out of memory"); return FM_ERR_OUT_OF_MEMORY; }
    set_last_error(e.what());
    return FM_ERR_INTERNAL;
}

#define FM_TRY try {
#define FM_CATCH } catch (const std::exception& e) { return fail(e); } catch (...) { set_last_error("unknown error"); return FM_ERR_INTERNAL; }

void require(bool cond, int code, const std::string& msg) { if (!cond) throw FmError(code, msg); }

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) { HIP_CHECK(hipGetDevice(&prev)); if (prev != dev) HIP_CHECK(hipSetDevice(dev)); }
    ~DeviceGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void validate_config(const fm_lmm_config* c) {
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "config is NULL");
    require(c->num_time_steps >= 1, FM_ERR_INVALID_ARGUMENT, "number of time steps must be >= 1");
    require(c->num_libors >= 1, FM_ERR_INVALID_ARGUMENT, "number of forward rates must be >= 1");
    require(c->num_factors >= 1, FM_ERR_INVALID_ARGUMENT,
            "Number of factors must be greater or equal 1 (given " + std::to_string(c->num_factors) + ").");
    require(c->num_paths >= 1, FM_ERR_INVALID_ARGUMENT,
            "Number of paths must be greater or equal 1 (given " + std::to_string(c->num_paths) + ").");
    require(c->times != nullptr && c->tenor != nullptr, FM_ERR_INVALID_ARGUMENT, "times and tenor must be given");
    for (int i = 0; i < c->num_time_steps; i++)
        require(c->times[i + 1] > c->times[i], FM_ERR_INVALID_ARGUMENT, "simulation times must increase");
    for (int i = 0; i < c->num_libors; i++)
        require(c->tenor[i + 1] > c->tenor[i], FM_ERR_INVALID_ARGUMENT, "tenor dates must increase");
    require(c->times[0] == 0.0 && c->tenor[0] == 0.0, FM_ERR_UNSUPPORTED, "simulation and tenor must start at time 0");
    require(c->measure == FM_MEASURE_SPOT || c->measure == FM_MEASURE_TERMINAL, FM_ERR_INVALID_ARGUMENT, "unknown measure");
    require(c->state_space == FM_STATE_SPACE_NORMAL || c->state_space == FM_STATE_SPACE_LOGNORMAL,
            FM_ERR_INVALID_ARGUMENT, "unknown state space");
    require(c->scheme >= FM_SCHEME_EULER && c->scheme <= FM_SCHEME_PREDICTOR_CORRECTOR_FUNCTIONAL,
            FM_ERR_INVALID_ARGUMENT, "unknown scheme");
    require(c->precision == FM_PRECISION_FP64 || c->precision == FM_PRECISION_FP32_MIXED ||
            c->precision == FM_PRECISION_FP64_LOGSTATE, FM_ERR_INVALID_ARGUMENT, "unknown precision");
    require(!std::isnan(c->libor_cap), FM_ERR_INVALID_ARGUMENT, "liborCap must not be NaN");
    require(c->path_offset >= 0, FM_ERR_INVALID_ARGUMENT, "path offset must not be negative");
    // the generator jumps at most 2^41 words (the cached jump polynomials x^(2^e), e <= 40)
    const double words = 2.0 * ((double)c->path_offset + (double)c->num_paths) * c->num_time_steps * c->num_factors;
    require(words < 2199023255552.0, FM_ERR_UNSUPPORTED, "path range beyond 2^41 generator words");
    require(c->num_libors <= 80 && c->num_factors <= 10, FM_ERR_UNSUPPORTED,
            "the fused kernel holds at most 80 forward rates and 10 factors");
}

// Rebuild derived host tables and upload the model (on the context stream).
void prepare_model(fm_lmm* c) {
    require(c->has_L0, FM_ERR_STATE, "initial forwards not set");
    require(c->has_cov, FM_ERR_STATE, "covariance tables not set");
    if (!c->tables_dirty) return;
    const int n = c->n(), nT = c->nT(), F = c->F();
    hipStream_t s = c->stream;
    // first alive component at each step (LMM.getDrift:1126-1130)
    c->first_alive.assign((size_t)nT, 0);
    for (int i = 0; i < nT; i++) {
        int f = jsearch(c->tenor, round_tick(c->times[(size_t)i])) + 1;
        if (f < 0) f = -f - 1 + 1;
        c->first_alive[(size_t)i] = std::min(f, n);
    }
    // process time index of each tenor date, rounded up (LMM:1055-1058)
    c->tenor_ti.assign((size_t)n + 1, nT + 1);
    for (int k = 0; k <= n; k++) {
        int ti = c->process_time_index(c->tenor[(size_t)k]);
        if (ti < 0) ti = -ti - 1;
        c->tenor_ti[(size_t)k] = ti;
    }
    // time-0 process value: transform(initial state) (Euler:187-192, LMM:1081-1093,1200-1212)
    c->init_value.assign((size_t)n, 0.0);
    for (int j = 0; j < n; j++) {
        double v;
        if (c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL) v = std::exp(std::log(std::max(c->L0[(size_t)j], 0.0)));
        else v = c->L0[(size_t)j];
        if (!std::isinf(c->cfg.libor_cap)) v = std::min(v, c->cfg.libor_cap);
        c->init_value[(size_t)j] = v;
    }
    // materialised-process row map
    c->row_of.assign((size_t)(nT + 1) * n, -1);
    int r = 0;
    for (int ti = 1; ti <= nT; ti++)
        for (int j = 0; j < n; j++) {
            if (j >= c->first_alive[(size_t)ti - 1]) c->row_of[(size_t)ti * n + j] = r++;
            else c->row_of[(size_t)ti * n + j] = c->row_of[(size_t)(ti - 1) * n + j];
        }
    c->num_rows = r;
    std::vector<double> delta((size_t)n), sq((size_t)nT);
    for (int j = 0; j < n; j++) delta[(size_t)j] = c->tenor[(size_t)j + 1] - c->tenor[(size_t)j];
    for (int i = 0; i < nT; i++) sq[(size_t)i] = std::sqrt(c->times[(size_t)i + 1] - c->times[(size_t)i]);
    c->uniform_dt = std::all_of(sq.begin(), sq.end(), [&](double v) { return v == sq[0]; });
    const bool lognormal = c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL;
    // records per (step i, component j): lambda_j.(t_i) zero-padded to the factor capacity,
    // variance*(-0.5) (0 in the normal state space), d/d and 1/d
    const int fcap = fcap_of(F), rs = rec_size_of(fcap);
    std::vector<double> rec((size_t)nT * n * rs, 0.0);
    for (int i = 0; i < nT; i++)
        for (int j = 0; j < n; j++) {
            double* r = rec.data() + ((size_t)i * n + j) * rs;
            for (int k = 0; k < F; k++) r[k] = c->lam[((size_t)i * n + j) * F + k];
            r[fcap] = lognormal ? c->var[(size_t)i * n + j] * (-0.5) : 0.0;
            // Scalar.of(+-d).discount(L, d) = 1/(L (d/(+-d)) + 1/(+-d)) (LMM:1145,1167)
            const double d = delta[(size_t)j], m = c->terminal() ? -d : d;
            r[fcap + 1] = d / m;
            r[fcap + 2] = 1.0 / m;
        }
    upload(c->d_times, c->times, s);
    upload(c->d_first, c->first_alive, s);
    upload(c->d_tenor_ti, c->tenor_ti, s);
    upload(c->d_delta, delta, s);
    upload(c->d_sqrt_dt, sq, s);
    upload(c->d_row_of, c->row_of, s);
    // host staging buffers stay alive until the stream synchronisation below
    std::vector<float> rf, lvf;
    std::vector<double> l0d;
    if (c->has_lv) {
        if (c->fp32()) { lvf.assign(c->lvb.begin(), c->lvb.end()); upload(c->d_lvb, lvf, s); }
        else upload(c->d_lvb, c->lvb, s);
    }
    if (c->fp32()) {
        rf.assign(rec.begin(), rec.end());
        upload(c->d_rec, rf, s);
        l0d.resize((size_t)n);                              // float-rounded initial state, passed as double
        for (int j = 0; j < n; j++) l0d[(size_t)j] = (double)(float)c->init_value[(size_t)j];
        upload(c->d_L0, l0d, s);
    } else {
        upload(c->d_rec, rec, s);
        upload(c->d_L0, c->init_value, s);
    }
    // initial state (LMM.getInitialState:1081-1093): log(max(L_j(0), 0)) or L_j(0), uncapped
    std::vector<double> y0((size_t)n);
    for (int j = 0; j < n; j++)
        y0[(size_t)j] = lognormal ? std::log(std::max(c->L0[(size_t)j], 0.0)) : c->L0[(size_t)j];
    upload(c->d_Y0, y0, s);
    // the log-forward kernel caps its log-states (Y <= log(cap))
    std::vector<double> y0c(y0);
    if (!std::isinf(c->cfg.libor_cap))
        for (double& y : y0c) y = std::min(y, std::log(c->cfg.libor_cap));
    upload(c->d_Y0cap, y0c, s);
    // terminal measure N(T_0): the Scalar chain of the time-0 values (LMM:1028-1042,
    // stochastic/Scalar.java:320-328)
    c->n0_terminal = 1.0;
    for (int j = 0; j < n; j++) {
        const double v = c->n0_terminal;
        c->n0_terminal = 1.0 / (c->init_value[(size_t)j] * (delta[(size_t)j] / v) + 1.0 / v);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    c->tables_dirty = false;
}

void ensure_states(fm_lmm* c) {
    if (c->states_ready) return;
    const uint64_t total = 2ull * (uint64_t)c->P() * (uint64_t)c->nT() * (uint64_t)c->F();
    c->lay = choose_layout(total);
    uint32_t seedw[624];
    mt_seed_window((int64_t)c->cfg.seed, seedw);
    // path p of the seed's stream starts at word 2 p nT F: jump to the first path of this context
    if (c->cfg.path_offset > 0)
        mt_jump_window_by(seedw, 2ull * (uint64_t)c->cfg.path_offset * (uint64_t)c->nT() * (uint64_t)c->F());
    uint32_t* st = (uint32_t*)c->d_states.get((size_t)c->lay.num_sub * 624 * sizeof(uint32_t));
    uint64_t* poly = (uint64_t*)c->d_poly.get((size_t)64 * 312 * sizeof(uint64_t));
    launch_jump_tree(c->lay, seedw, st, poly, c->stream);
    c->states_ready = true;
}

// simulation times and sqrt(dt) on the device (also for a context used only as a Brownian driver)
void ensure_time_steps(fm_lmm* c) {
    if (c->steps_ready) return;
    const int nT = c->nT();
    std::vector<double> sq((size_t)nT);
    for (int i = 0; i < nT; i++) sq[(size_t)i] = std::sqrt(c->times[(size_t)i + 1] - c->times[(size_t)i]);
    c->uniform_dt = std::all_of(sq.begin(), sq.end(), [&](double v) { return v == sq[0]; });
    upload(c->d_times, c->times, c->stream);
    upload(c->d_sqrt_dt, sq, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->steps_ready = true;
}

void ensure_brownian(fm_lmm* c) {
    ensure_time_steps(c);
    ensure_states(c);
    if (c->dw_ready) return;
    const size_t count = (size_t)c->P() * c->nT() * c->F();
    void* out = c->d_dw.get(count * c->real_size());
    launch_brownian(c->lay, c->d_states.as<uint32_t>(), c->nT(), c->F(), c->d_sqrt_dt.as<double>(),
                    c->uniform_dt ? 1 : 0, out, c->fp32() ? 1 : 0, c->stream);
    c->dw_ready = true;
}

// Backward induction plan of one Bermudan swaption.
struct BermPlan {
    int slot = -1;
    int callable = 1;
    int basis = FM_BASIS_REFERENCE;
    int num_basis = 6;
    std::vector<int> per_col, per_k;        // per period: payoff column and its numeraire index
    std::vector<int> ex_period, ex_col, ex_k;   // per exercise date (ascending): period, first basis column, N index
    std::vector<double> fixing;                 // fixing dates of the periods
};

// Products lowered for the kernel.
struct Lowered {
    std::vector<BermPlan> berm;
    std::vector<KProduct> kp;
    std::vector<KPeriod> per;
    int num_slots = 0;
    int num_contrib = 0;
    std::vector<int> slot_of;                 // product q -> reduction slot
    std::vector<int> slot_k;                  // slot -> numeraire tenor index (single-date)
    std::vector<int32_t> swap_col0, swap_ncol, swap_slot;
    std::vector<int> contrib_k;               // contribution column -> numeraire tenor index
};

// kcur of the kernel after the numeraire advance at time index ti
int kcur_at(const fm_lmm* c, int ti) {
    int k = 0;
    while (k < c->n() && c->tenor_ti[(size_t)k] <= ti) k++;
    return k;
}

// Process time index used by getForwardRate for `time` (LMM:1243-1252, ROUND_NEAREST).
int forward_time_index(const fm_lmm* c, double time) {
    int ti = c->process_time_index(time);
    if (ti < 0) {
        ti = -ti - 2;
        if (ti + 1 <= c->nT() && time - c->times[(size_t)ti] > c->times[(size_t)ti + 1] - time) ti++;
    }
    return ti;
}

// Swaptions evaluated at one time index over the same periods (first forward, count, period
// lengths, discounting taus and adjustments) and divided by the same numeraire differ only in
// strike and notional: the path kernel values up to kSwaptionGroup of them in one walk over the
// periods. Each group is made consecutive (the order of products within a time index changes no
// result: each has its own reduction slot) and its first member carries the group's size - 1.
// The members' periods are copied back to back (member s at the first's offset + s * periods),
// each period with its swaption's notional.
static bool same_swaption_periods(const KProduct& a, const KProduct& b, const std::vector<KPeriod>& per) {
    if (a.type != FM_PRODUCT_SWAPTION || b.type != FM_PRODUCT_SWAPTION) return false;
    if (a.eval_index != b.eval_index || a.first_j != b.first_j || a.num_periods != b.num_periods ||
        a.num_k != b.num_k)
        return false;
    for (int i = 0; i < a.num_periods; i++) {
        const KPeriod& p = per[(size_t)(a.param_off + i)];
        const KPeriod& q = per[(size_t)(b.param_off + i)];
        if (p.period_length != q.period_length || p.tau != q.tau || p.adj != q.adj) return false;
    }
    return true;
}

static void group_swaptions(Lowered& L) {
    std::vector<KProduct>& kp = L.kp;
    for (size_t i = 0; i < kp.size(); i++) kp[i].pad2 = 0;
    for (size_t i = 0; i < kp.size();) {
        size_t e = i + 1;
        if (kp[i].type == FM_PRODUCT_SWAPTION) {
            for (size_t f = i + 1; f < kp.size() && kp[f].eval_index == kp[i].eval_index &&
                                   e - i < (size_t)kSwaptionGroup; f++) {
                if (!same_swaption_periods(kp[i], kp[f], L.per)) continue;
                std::rotate(kp.begin() + (long)e, kp.begin() + (long)f, kp.begin() + (long)f + 1);
                e++;
            }
            kp[i].pad2 = (int32_t)(e - i - 1);
            if (e - i > 1) {
                const int np = kp[i].num_periods;
                const int base = (int)L.per.size();
                for (size_t f = i; f < e; f++) {
                    for (int q = 0; q < np; q++) {
                        KPeriod pe = L.per[(size_t)(kp[f].param_off + q)];
                        pe.notional = kp[f].notional;
                        L.per.push_back(pe);
                    }
                    kp[f].param_off = base + (int)(f - i) * np;
                }
            }
        }
        i = e;
    }
}

Lowered lower_products(fm_lmm* c, const fm_product* prods, int count) {
    Lowered L;
    require(count >= 0 && (count == 0 || prods != nullptr), FM_ERR_INVALID_ARGUMENT, "products array is NULL");
    const int n = c->n();
    auto numeraire_index = [&](double T) {
        int k = c->libor_index(T);
        require(k >= 0, FM_ERR_UNSUPPORTED, "numeraire requested off the tenor grid (interpolation not implemented)");
        return k;
    };
    auto check_available = [&](int ti, int k) {
        require(ti >= 0 && ti <= c->nT(), FM_ERR_INVALID_ARGUMENT, "product date outside the simulation horizon");
        if (c->terminal()) {
            // the kernel holds N(T_k) from the time index of T_k on (and the one before it)
            int kn = 0;
            while (kn <= n && c->tenor_ti[(size_t)kn] <= ti) kn++;
            require(k == kn - 1, FM_ERR_UNSUPPORTED, "numeraire date not on the simulation time grid");
            return;
        }
        const int kc = kcur_at(c, ti);
        require(k == kc || k == kc - 1, FM_ERR_UNSUPPORTED, "numeraire date not adjacent to the evaluation date");
    };
    // Terminal measure: N(T_pay) is known only at the time index of T_pay, so a payoff of
    // a forward fixed earlier is evaluated there; the forward is frozen in between (its
    // process value aliases the fixing value, EulerSchemeFromProcessModel.java:279-290).
    auto defer_to_payment = [&](KProduct& k, int j) {
        if (!c->terminal()) return;
        const int te = c->tenor_ti[(size_t)k.num_k];
        if (te <= k.eval_index || te > c->nT()) return;
        for (int ti = k.eval_index; ti < te; ti++)
            require(c->first_alive[(size_t)ti] > j, FM_ERR_UNSUPPORTED, "forward not fixed before its payment date");
        k.eval_index = te;
    };
    L.slot_of.assign((size_t)count, -1);
    for (int q = 0; q < count; q++) {
        const fm_product& p = prods[q];
        require(p.num_periods >= 1, FM_ERR_INVALID_ARGUMENT, "product must have at least one period");
        if (p.type == FM_PRODUCT_SWAPTION) {
            require(p.fixing_dates && p.payment_dates && p.swaprates, FM_ERR_INVALID_ARGUMENT, "swaption arrays missing");
            KProduct k{};
            k.type = FM_PRODUCT_SWAPTION;
            const double Te = p.exercise_date;
            for (int i = 0; i < p.num_periods; i++)
                require(p.fixing_dates[i] >= Te, FM_ERR_UNSUPPORTED, "swaption periods fixing before exercise");
            k.eval_index = forward_time_index(c, Te);
            const int j0 = c->libor_index(p.fixing_dates[0]);
            require(j0 >= 0, FM_ERR_UNSUPPORTED, "swaption fixing off the tenor grid");
            for (int i = 0; i < p.num_periods; i++) {
                const int ji = c->libor_index(p.fixing_dates[i]);
                const int jp = c->libor_index(p.payment_dates[i]);
                require(ji == j0 + i && jp == ji + 1 && jp <= n, FM_ERR_UNSUPPORTED,
                        "swaption periods must be consecutive single tenor periods");
                KPeriod pe{};
                pe.strike = p.swaprates[i];
                pe.period_length = p.period_lengths ? p.period_lengths[i] : p.payment_dates[i] - p.fixing_dates[i];
                pe.tau = p.payment_dates[i] - std::max(p.fixing_dates[i], Te);
                pe.adj = p.discounting_adjustment ? p.discounting_adjustment[i] : 1.0;
                pe.inv_adj = 1.0 / pe.adj;
                pe.notional = p.notional;
                pe.j = ji; pe.pay_k = jp;
                L.per.push_back(pe);
            }
            k.first_j = j0;
            k.num_periods = p.num_periods;
            k.notional = p.notional;
            k.exercise = Te;
            k.param_off = (int32_t)L.per.size() - p.num_periods;
            k.num_k = numeraire_index(Te);
            check_available(k.eval_index, k.num_k);
            k.out_slot = L.num_slots++;
            L.slot_of[(size_t)q] = k.out_slot;
            L.slot_k.push_back(k.num_k);
            L.kp.push_back(k);
        } else if (p.type == FM_PRODUCT_CAPLET) {
            require(p.period_lengths && p.swaprates, FM_ERR_INVALID_ARGUMENT, "caplet arrays missing");
            const double Tm = p.exercise_date, dl = p.period_lengths[0];
            const int j = c->libor_index(Tm), jp = c->libor_index(Tm + dl);
            require(j >= 0 && jp == j + 1, FM_ERR_UNSUPPORTED, "caplet period must be a tenor period");
            KProduct k{};
            k.type = FM_PRODUCT_CAPLET;
            k.eval_index = forward_time_index(c, Tm);
            k.first_j = j; k.num_periods = 1;
            k.is_floorlet = p.is_floorlet;
            KPeriod pe{};
            pe.strike = p.swaprates[0]; pe.period_length = dl; pe.tau = 0; pe.adj = 1.0;
            pe.notional = p.notional;   // daycount fraction
            pe.j = j; pe.pay_k = jp;
            L.per.push_back(pe);
            k.param_off = (int32_t)L.per.size() - 1;
            k.num_k = numeraire_index(Tm + dl);
            defer_to_payment(k, j);
            check_available(k.eval_index, k.num_k);
            k.out_slot = L.num_slots++;
            L.slot_of[(size_t)q] = k.out_slot;
            L.slot_k.push_back(k.num_k);
            L.kp.push_back(k);
        } else if (p.type == FM_PRODUCT_SWAP) {
            require(p.fixing_dates && p.payment_dates && p.swaprates, FM_ERR_INVALID_ARGUMENT, "swap arrays missing");
            const int slot = L.num_slots++;
            L.slot_of[(size_t)q] = slot;
            L.slot_k.push_back(-1);
            L.swap_col0.push_back(L.num_contrib);
            L.swap_ncol.push_back(p.num_periods);
            L.swap_slot.push_back(slot);
            for (int i = 0; i < p.num_periods; i++) {
                const double fd = p.fixing_dates[i], pd = p.payment_dates[i];
                const int j = c->libor_index(fd), jp = c->libor_index(pd);
                require(j >= 0 && jp == j + 1, FM_ERR_UNSUPPORTED, "swap periods must be tenor periods");
                KProduct k{};
                k.type = FM_PRODUCT_SWAP;
                k.eval_index = forward_time_index(c, fd);
                k.first_j = j; k.num_periods = 1;
                k.flags = p.is_pay_fix ? 1 : 0;
                KPeriod pe{};
                pe.strike = p.swaprates[i]; pe.period_length = pd - fd; pe.tau = 0; pe.adj = 1.0;
                pe.notional = p.notionals ? p.notionals[i] : 1.0;
                pe.j = j; pe.pay_k = jp;
                L.per.push_back(pe);
                k.param_off = (int32_t)L.per.size() - 1;
                k.num_k = numeraire_index(pd);
                defer_to_payment(k, j);
                check_available(k.eval_index, k.num_k);
                k.out_slot = L.num_contrib++;
                L.contrib_k.push_back(k.num_k);
                L.kp.push_back(k);
            }
        } else if (p.type == FM_PRODUCT_BERMUDAN_SWAPTION) {
            // BermudanSwaption.getValues (BermudanSwaption.java:91-157): periods become payoff
            // columns (payoff / N(T_pay)), exercise dates become raw basis columns.
            require(p.fixing_dates && p.payment_dates && p.period_lengths && p.swaprates && p.exercise_flags,
                    FM_ERR_INVALID_ARGUMENT, "Bermudan swaption arrays missing");
            require(p.basis == FM_BASIS_REFERENCE || p.basis == FM_BASIS_SWAP_RATE, FM_ERR_INVALID_ARGUMENT,
                    "unknown regression basis " + std::to_string(p.basis));
            const int m = p.num_periods;
            BermPlan b;
            b.slot = L.num_slots++;
            L.slot_of[(size_t)q] = b.slot;
            L.slot_k.push_back(-1);
            b.callable = p.is_cancelable ? 0 : 1;
            b.basis = p.basis;
            b.num_basis = p.basis == FM_BASIS_REFERENCE ? 6 : 4;
            b.fixing.assign(p.fixing_dates, p.fixing_dates + p.num_periods);
            const int j0 = c->libor_index(p.fixing_dates[0]);
            for (int i = 0; i < m; i++) {
                const double fd = p.fixing_dates[i], pd = p.payment_dates[i], pl = p.period_lengths[i];
                const int j = c->libor_index(fd), jp = c->libor_index(pd);
                require(j >= 0 && j == j0 + i && jp == j + 1 && c->libor_index(fd + pl) == j + 1, FM_ERR_UNSUPPORTED,
                        "Bermudan periods must be consecutive single tenor periods");
                KProduct k{};
                k.type = FM_PRODUCT_SWAP;
                k.eval_index = forward_time_index(c, fd);
                k.first_j = j; k.num_periods = 1;
                k.flags = 1;
                KPeriod pe{};
                pe.strike = p.swaprates[i]; pe.period_length = pl; pe.tau = 0; pe.adj = 1.0;
                pe.notional = p.notionals ? p.notionals[i] : 1.0;
                pe.j = j; pe.pay_k = jp;
                L.per.push_back(pe);
                k.param_off = (int32_t)L.per.size() - 1;
                k.num_k = numeraire_index(pd);
                defer_to_payment(k, j);
                check_available(k.eval_index, k.num_k);
                k.out_slot = L.num_contrib++;
                L.contrib_k.push_back(k.num_k);
                L.kp.push_back(k);
                b.per_col.push_back(k.out_slot);
                b.per_k.push_back(k.num_k);
            }
            for (int i = 0; i < m; i++) {
                if (!p.exercise_flags[i]) continue;
                const double fd = p.fixing_dates[i];
                KProduct k{};
                k.type = K_BERM_BASIS;
                k.eval_index = forward_time_index(c, fd);
                k.first_j = j0 + i;
                k.num_periods = m - i;
                k.flags = p.basis;
                k.param_off = (int32_t)L.per.size();
                if (p.basis == FM_BASIS_REFERENCE) {
                    // the exercise date is the fixing date of period i (getBasisFunctions :227-233)
                    KPeriod pe{};
                    pe.tau = p.payment_dates[i] - fd;
                    pe.period_length = p.payment_dates[m - 1] - p.fixing_dates[i];
                    pe.j = c->libor_index(p.payment_dates[i]);
                    pe.pay_k = c->libor_index(p.payment_dates[m - 1]);
                    pe.adj = 1.0;
                    L.per.push_back(pe);
                } else {
                    for (int r = i; r < m; r++) {
                        KPeriod pe{};
                        pe.period_length = p.period_lengths[r];
                        pe.j = j0 + r; pe.pay_k = j0 + r + 1; pe.adj = 1.0;
                        L.per.push_back(pe);
                    }
                }
                k.num_k = numeraire_index(fd);
                check_available(k.eval_index, k.num_k);
                k.out_slot = L.num_contrib;
                const int ncols = p.basis == FM_BASIS_REFERENCE ? 3 : 2;
                for (int r = 0; r < ncols; r++) L.contrib_k.push_back(-1);
                L.num_contrib += ncols;
                L.kp.push_back(k);
                b.ex_period.push_back(i);
                b.ex_col.push_back(k.out_slot);
                b.ex_k.push_back(k.num_k);
            }
            L.berm.push_back(std::move(b));
        } else {
            throw FmError(FM_ERR_INVALID_ARGUMENT, "unknown product type " + std::to_string(p.type));
        }
    }
    std::stable_sort(L.kp.begin(), L.kp.end(), [](const KProduct& a, const KProduct& b) { return a.eval_index < b.eval_index; });
    group_swaptions(L);
    return L;
}

// Adjusted numeraire at time 0 by which the products multiply (Swaption.java:192-196):
// N(0) avg(N(0)/N(0)) / P^d(0), a Scalar (Scalar.div: (1/adj) (N zb), stochastic/Scalar.java:296-298).
double numeraire0(const fm_lmm* c, const std::vector<double>& adj) {
    const double N0 = c->terminal() ? c->n0_terminal : 1.0;
    return c->has_pd ? (1.0 / adj[0]) * (N0 * c->zb[0]) : N0;
}

// Scale turning mean(payoff / N(T_k)) into the product value: the numeraire adjustment
// N(T) avg(N(0)/N(T)) / P^d(T) of T_k (LMM:869-874) and, under the terminal measure, the
// time-0 numeraire (1 under the spot measure without discount curve).
double value_scale(const fm_lmm* c, int k, const std::vector<double>& adj) {
    double sc = (k >= 0 && c->has_pd) ? adj[(size_t)k] / c->zb[(size_t)k] : 1.0;
    if (c->terminal()) sc *= numeraire0(c, adj);
    return sc;
}

std::vector<double> adjustment_chain(const fm_lmm* c) {
    const int n = c->n();
    std::vector<double> adj((size_t)n + 1, 1.0);
    if (!c->has_pd) return adj;
    double a0 = c->pd[0];
    adj[0] = a0;
    for (int i = 0; i < n; i++) {
        const double ts = c->tenor[(size_t)i + 1] - c->tenor[(size_t)i];
        const double fr = (c->pd[(size_t)i] / c->pd[(size_t)i + 1] - 1.0) / ts;
        a0 = 1.0 / (fr * (ts / a0) + 1.0 / a0);         // Scalar.discount (stochastic/Scalar.java:320-328)
        adj[(size_t)i + 1] = a0;
    }
    return adj;
}

struct Timer {
    fm_lmm* c;
    explicit Timer(fm_lmm* cc) : c(cc) { for (int i = 0; i < 5; i++) c->timings[i] = 0.0; }
    void mark(int e) { HIP_CHECK(hipEventRecord(c->ev[e], c->stream)); }
    float span(int a, int b) { float ms = 0; HIP_CHECK(hipEventElapsedTime(&ms, c->ev[a], c->ev[b])); return ms; }
};

// Regression basis reference of exercise date e of a Bermudan plan (numeraire adjustment of
// getNumeraire, LIBORMarketModelFromCovarianceModel.java:869-874).
BermBasisRef basis_ref(const fm_lmm* c, const BermPlan& b, int e, const std::vector<double>& adj) {
    BermBasisRef r{};
    r.kind = b.basis;
    r.col0 = b.ex_col[(size_t)e];
    const int k = b.ex_k[(size_t)e];
    const bool scalar_num = c->terminal() ? (k == 0 || k == c->n()) : k <= 1;
    r.nmode = !c->has_pd ? 0 : (scalar_num ? 1 : 2);
    r.zb = c->zb[(size_t)k];
    r.adj = adj[(size_t)k];
    return r;
}

// Backward induction of one Bermudan over the columns of the last path pass
// (BermudanSwaption.java:100-146): one step launch per exercise date (latest first)
// followed by its regression solve, then the final step into the product's slot.
void run_bermudan(fm_lmm* c, const BermPlan& b, int bi, const std::vector<double>& adj, const double* contrib,
                  double* wsum, double* wm2, int stride) {
    hipStream_t s = c->stream;
    const int P = c->P(), m = (int)b.per_col.size(), nex = (int)b.ex_period.size();
    std::vector<int32_t> cols((size_t)m);
    std::vector<double> scales((size_t)m);
    for (int r = 0; r < m; r++) {            // processing order: descending period index
        const int i = m - 1 - r, k = b.per_k[(size_t)i];
        cols[(size_t)r] = b.per_col[(size_t)i];
        scales[(size_t)r] = c->has_pd ? adj[(size_t)k] / c->zb[(size_t)k] : 1.0;
    }
    const int32_t* dcol = upload(c->d_berm_col, cols, s);
    const double* dscale = upload(c->d_berm_scale, scales, s);
    double* vals = (double*)c->d_berm_vals.get((size_t)P * sizeof(double));
    double* under = (double*)c->d_berm_under.get((size_t)P * sizeof(double));
    // exercise period per path of Bermudan bi (kept for the adjoint); the buffer is sized
    // for all Bermudans of the call by run_pass
    int32_t* ex = c->d_berm_ex.as<int32_t>() + (size_t)bi * P;
    HIP_CHECK(hipMemsetAsync(ex, 0xff, (size_t)P * sizeof(int32_t), s));
    const int nb = berm_num_blocks(P);
    double* part = (double*)c->d_berm_part.get((size_t)nb * kBermParts * sizeof(double));
    double* coef = (double*)c->d_berm_coef.get((size_t)std::max(nex, 1) * kBermMaxBasis * sizeof(double));
    BermStep st{};
    st.P = P;
    st.callable = b.callable;
    st.w = 1.0 / (double)P;
    st.contrib = contrib;
    st.values = vals; st.under = under; st.ex_index = ex;
    st.partials = part;
    st.wave_sum = wsum; st.wave_m2 = wm2; st.stride = stride; st.slot = b.slot;
    st.n0 = numeraire0(c, adj);
    int done = m;       // periods >= done have been added
    int prev = -1;      // exercise date whose decision is still to be applied
    for (int e = nex - 1; e >= -1; e--) {
        const int lo = e >= 0 ? b.ex_period[(size_t)e] : 0;
        BermStep t = st;
        t.first = done == m ? 1 : 0;
        t.apply = prev >= 0 ? 1 : 0;
        if (prev >= 0) {
            t.apply_basis = basis_ref(c, b, prev, adj);
            t.apply_coef = coef + (size_t)prev * kBermMaxBasis;
            t.apply_period = b.ex_period[(size_t)prev];
            if (c->berm_record) {
                const size_t row = (size_t)(c->berm_trig_off[(size_t)bi] + prev) * P;
                t.rec_trig = c->d_berm_trig.as<double>() + row;
                t.rec_diff = c->d_berm_diff.as<double>() + row;
            }
        }
        t.per_col = dcol + (m - done);
        t.per_scale = dscale + (m - done);
        t.num_per = done - lo;
        t.regress = e >= 0 ? 1 : 0;
        if (e >= 0) t.reg_basis = basis_ref(c, b, e, adj);
        t.final_ = e < 0 ? 1 : 0;
        launch_berm_step(t, b.basis, s);
        if (e >= 0) {
            // recording for the adjoint: keep the date's pseudo-inverse, which projects the
            // trigger's adjoint on the same basis
            double* factors = c->berm_record
                ? c->d_berm_pinv.as<double>() + (size_t)(c->berm_trig_off[(size_t)bi] + e) * kBermPinv : nullptr;
            launch_berm_solve(part, nb, b.num_basis, P, coef + (size_t)e * kBermMaxBasis, s, factors);
            prev = e;
        }
        done = lo;
    }
    HIP_CHECK(hipStreamSynchronize(s));      // cols and scales are host sources of the uploads above
}

// The adjoint of Bermudan bi's induction (lmm_bermudan.hip): exercise dates ascending, from the
// triggers and choices its pricing pass recorded; leaves dV/d(payoff) of every period per path in
// weight[period column][P]. The DISCRETE_DELTA width of each date is width x sd(trigger).
void run_bermudan_adjoint(fm_lmm* c, const BermPlan& b, int bi, const std::vector<double>& adj, const double* contrib,
                          double* weight) {
    hipStream_t s = c->stream;
    const int P = c->P(), m = (int)b.per_col.size(), nex = (int)b.ex_period.size();
    const size_t row0 = (size_t)c->berm_trig_off[(size_t)bi] * P;
    const double* trig = c->d_berm_trig.as<double>() + row0;
    const double* diff = c->d_berm_diff.as<double>() + row0;
    const double* pinv = c->d_berm_pinv.as<double>() + (size_t)c->berm_trig_off[(size_t)bi] * kBermPinv;
    // the centred second moment of every date's trigger, for the DISCRETE_DELTA width
    double* stat = (double*)c->d_berm_stat.get((size_t)2 * std::max(nex, 1) * sizeof(double));
    if (c->dirac_method == FM_DIRAC_DISCRETE_DELTA && nex > 0) {
        const int num_waves = (P + 63) / 64;
        double* wsum = (double*)c->d_berm_wsum.get((size_t)num_waves * nex * sizeof(double));
        double* wm2 = (double*)c->d_berm_wm2.get((size_t)num_waves * nex * sizeof(double));
        launch_berm_trigger_stats(trig, nex, P, wsum, wm2, nex, s);
        launch_reduce_partials(wsum, wm2, num_waves, nex, nex, 0, P, stat, stat + nex, s);
    }
    const int32_t* wcol = upload(c->d_berm_wcol, b.per_col, s);       // ascending periods
    double* vb = (double*)c->d_berm_vb.get((size_t)P * sizeof(double));
    double* ub = (double*)c->d_berm_ub.get((size_t)P * sizeof(double));
    const int nb = berm_num_blocks(P);
    double* part = (double*)c->d_berm_part.get((size_t)nb * kBermParts * sizeof(double));
    double* bcoef = (double*)c->d_berm_bcoef.get((size_t)std::max(nex, 1) * kBermMaxBasis * sizeof(double));
    BermAdjStep a{};
    a.P = P;
    a.callable = b.callable;
    a.method = c->dirac_method;
    a.contrib = contrib;
    a.weight = weight;
    a.vb = vb; a.ub = ub;
    a.partials = part;
    int lo = 0;         // first period whose weight is still to be written
    for (int e = 0; e <= nex; e++) {
        BermAdjStep t = a;
        t.first = e == 0 ? 1 : 0;
        t.finish = (e > 0 && c->dirac_method != FM_DIRAC_ZERO) ? 1 : 0;
        if (t.finish) {
            t.fin_basis = basis_ref(c, b, e - 1, adj);
            t.fin_coef = bcoef + (size_t)(e - 1) * kBermMaxBasis;
        }
        const int hi = e < nex ? b.ex_period[(size_t)e] : m;
        t.w_col = wcol + lo;
        t.num_w = hi - lo;
        t.decide = e < nex ? 1 : 0;
        if (t.decide) {
            t.dec_basis = basis_ref(c, b, e, adj);
            t.trig = trig + (size_t)e * P;
            t.diff = diff + (size_t)e * P;
            t.trig_m2 = stat + nex + e;
            t.width = c->dirac_width;
        }
        launch_berm_adjoint_step(t, b.basis, s);
        if (t.decide && c->dirac_method != FM_DIRAC_ZERO)
            launch_berm_project(part, nb, b.num_basis, P, pinv + (size_t)e * kBermPinv,
                                bcoef + (size_t)e * kBermMaxBasis, s);
        lo = hi;
    }
    // no synchronisation: the upload's host source (the plan's per_col) lives until the caller's
    // final synchronisation
}

// Scratch bytes of a pass over single-date products with the per-call buffers in caller memory
// (run_pass with an ExternalScratch).
size_t external_scratch_bytes(const fm_lmm* c, const Lowered& L) {
    const size_t num_waves = ((size_t)c->P() + 63) / 64, stride = (size_t)L.num_slots + c->n() + 1;
    return 2 * ExternalScratch::piece(num_waves * stride * sizeof(double)) +
           ExternalScratch::piece(L.kp.size() * sizeof(KProduct)) + ExternalScratch::piece(L.per.size() * sizeof(KPeriod)) +
           ExternalScratch::piece(stride * 2 * sizeof(double));
}

// Run generator + fused step (+ reductions). Returns per-column sums and M2 (host).
// With xs, the wave partials, product tables and reduction output live in that caller memory
// and are cleared by a fill kernel; such a pass takes single-date products only.
void run_pass(fm_lmm* c, const Lowered& L, bool store, std::vector<double>& sums, std::vector<double>& m2s,
              void* ckpt = nullptr, int ckpt_stride = 1, double* num_out = nullptr, ExternalScratch* xs = nullptr,
              const int32_t* ckpt_first = nullptr) {
    prepare_model(c);
    hipStream_t s = c->stream;
    const int n = c->n(), P = c->P();
    if (xs) require(L.num_contrib == 0 && L.berm.empty() && !store && !ckpt, FM_ERR_INTERNAL,
                    "a pass in caller scratch takes single-date products only");
    Timer tm(c);
    tm.mark(0);
    if (c->cfg.regenerate_increments) c->dw_ready = false;
    const bool regen = !c->dw_ready;
    ensure_brownian(c);
    tm.mark(1);
    const int64_t num_waves = ((int64_t)P + 63) / 64;
    const int stride = L.num_slots + n + 1;
    const size_t wbytes = (size_t)num_waves * stride * sizeof(double);
    double* wsum = (double*)(xs ? xs->take(wbytes) : c->d_wave_sum.get(wbytes));
    double* wm2 = (double*)(xs ? xs->take(wbytes) : c->d_wave_m2.get(wbytes));
    if (xs) {
        launch_fill_zero(wsum, (size_t)num_waves * stride, s);
        launch_fill_zero(wm2, (size_t)num_waves * stride, s);
    } else {
        HIP_CHECK(hipMemsetAsync(wsum, 0, wbytes, s));
        HIP_CHECK(hipMemsetAsync(wm2, 0, wbytes, s));
    }
    StepArgs a{};
    a.nT = c->nT(); a.n = n; a.F = c->F(); a.P = P;
    a.precision = c->cfg.precision;
    a.state_space = c->cfg.state_space;
    a.store_paths = store ? 1 : 0;
    a.times = c->d_times.as<double>();
    a.first_alive = c->d_first.as<int32_t>();
    a.tenor_time_index = c->d_tenor_ti.as<int32_t>();
    a.rec = c->d_rec.p;
    a.lognormal = c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL ? 1 : 0;
    a.capped = std::isinf(c->cfg.libor_cap) ? 0 : 1;
    a.delta = c->d_delta.as<double>();
    a.L0 = c->d_L0.as<double>();
    a.Y0 = c->d_Y0.as<double>();
    a.logcap = std::isinf(c->cfg.libor_cap) ? HUGE_VAL : std::log(c->cfg.libor_cap);
    a.terminal = c->terminal() ? 1 : 0;
    a.carry = c->carry() ? 1 : 0;
    a.pc = c->pc() ? 1 : 0;
    a.dw = c->d_dw.p;
    a.cap = c->cfg.libor_cap;
    a.scheme = c->cfg.scheme;
    a.num_products = (int)L.kp.size();
    a.products = xs ? xs->upload(L.kp, s) : upload(c->d_products, L.kp, s);
    a.periods = xs ? xs->upload(L.per, s) : upload(c->d_periods, L.per, s);
    a.num_slots = L.num_slots;
    a.num_contrib = L.num_contrib;
    a.contrib = L.num_contrib ? (double*)c->d_contrib.get((size_t)L.num_contrib * P * sizeof(double)) : nullptr;
    a.wave_sum = wsum; a.wave_m2 = wm2;
    a.paths = store ? c->d_paths.get((size_t)std::max(c->num_rows, 1) * P * c->real_size()) : nullptr;
    a.row_of = c->d_row_of.as<int32_t>();
    a.numeraire = store ? (double*)c->d_num.get((size_t)(n + 1) * P * sizeof(double)) : num_out;
    a.ckpt = ckpt;
    a.ckpt_stride = ckpt ? ckpt_stride : 1;
    a.ckpt_first = ckpt_first;
    tm.mark(2);
    // kernel variant: the spot-measure Euler kernels, else the general kernel (terminal
    // measure, carried state, predictor-corrector)
    // the log-forward kernel caps the log-state, as EULER_FUNCTIONAL does up to rounding;
    // EULER with a finite cap caps only the value and runs on the general kernel
    const bool logstate_kernel = c->cfg.precision == FM_PRECISION_FP64_LOGSTATE &&
        c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL && !c->terminal() && !c->pc() &&
        (!c->carry() || std::isinf(c->cfg.libor_cap)) && !c->has_lv && !c->has_sv;
    if (logstate_kernel) {
        a.Y0 = c->d_Y0cap.as<double>();
        launch_lmm_step_fp64_logstate(a, s);
    } else if (c->has_lv || c->has_sv) {
        a.lv_alpha = c->lv_alpha;
        a.lvb = c->has_lv ? c->d_lvb.p : nullptr;
        a.svs = c->has_sv ? c->d_svs.p : nullptr;
        if (c->fp32()) launch_lmm_step_lv_fp32(a, s);
        else launch_lmm_step_lv_fp64(a, s);
    } else if (c->plain_euler()) {
        // the spot kernel's initial states: FP64 log-normal the log-forwards capped at log(cap),
        // otherwise the capped initial forwards
        a.Y0 = (!c->fp32() && c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL) ? c->d_Y0cap.as<double>()
                                                                               : c->d_L0.as<double>();
        if (c->fp32()) launch_lmm_step_fp32(a, s);
        else launch_lmm_step_fp64(a, s);
    } else {
        // the general kernel: FP64 log-normal functional schemes carry the log-forwards, capped
        // at log(cap), like the spot kernel
        if (!c->fp32() && c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL && !c->carry())
            a.Y0 = c->d_Y0cap.as<double>();
        if (c->pc()) {
            if (c->fp32()) launch_lmm_step_pc_fp32(a, s);
            else launch_lmm_step_pc_fp64(a, s);
        } else {
            if (c->fp32()) launch_lmm_step_general_fp32(a, s);
            else launch_lmm_step_general_fp64(a, s);
        }
    }
    tm.mark(3);
    // numeraire-adjustment averages first (the swap contributions need them)
    const int cols = L.num_slots + n + 1;
    double* dout = (double*)(xs ? xs->take((size_t)cols * 2 * sizeof(double)) : c->d_out.get((size_t)cols * 2 * sizeof(double)));
    launch_reduce_partials(wsum, wm2, (int)num_waves, stride, L.num_slots, n + 1, P, dout, dout + cols, s);
    sums.assign((size_t)cols, 0.0); m2s.assign((size_t)cols, 0.0);
    HIP_CHECK(hipMemcpyAsync(sums.data(), dout, cols * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(m2s.data(), dout + cols, cols * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    // zb_k = average of N(0)/N(T_k) (LMM:872)
    c->zb.assign((size_t)n + 1, 1.0);
    if (c->terminal()) {
        // N(T_0) and N(T_n) = 1 are Scalars: (1/N0) N0 and (1/1) N0
        const double N0 = c->n0_terminal;
        for (int k = 1; k < n; k++) c->zb[(size_t)k] = sums[(size_t)L.num_slots + k] / (double)P;
        c->zb[0] = (1.0 / N0) * N0;
        c->zb[(size_t)n] = N0;
    } else {
        // N(T_0) = 1 and N(T_1) is deterministic
        for (int k = 1; k <= n; k++) c->zb[(size_t)k] = sums[(size_t)L.num_slots + k] / (double)P;
        if (n >= 1) {
            const double N1 = c->init_value[0] * ((c->tenor[1] - c->tenor[0]) * 1.0) + 1.0;
            c->zb[1] = (1.0 / N1) * 1.0;
        }
    }
    if (!L.swap_slot.empty() || !L.berm.empty()) {
        const std::vector<double> adj = adjustment_chain(c);
        if (!L.swap_slot.empty()) {
            std::vector<double> scale((size_t)L.num_contrib, 1.0);
            for (int col = 0; col < L.num_contrib; col++) {
                const int k = L.contrib_k[(size_t)col];
                if (k >= 0) scale[(size_t)col] = value_scale(c, k, adj);
            }
            const double* dscale = upload(c->d_scale, scale, s);
            std::vector<int32_t> meta;
            meta.insert(meta.end(), L.swap_col0.begin(), L.swap_col0.end());
            meta.insert(meta.end(), L.swap_ncol.begin(), L.swap_ncol.end());
            meta.insert(meta.end(), L.swap_slot.begin(), L.swap_slot.end());
            const int32_t* dmeta = upload(c->d_swapmeta, meta, s);
            const int ns = (int)L.swap_slot.size();
            launch_swap_finalize(a.contrib, P, ns, dmeta, dmeta + ns, dmeta + 2 * ns, dscale, wsum, wm2, stride, s);
        }
        if (!L.berm.empty()) c->d_berm_ex.get(L.berm.size() * (size_t)P * sizeof(int32_t));
        c->berm_fixing.clear();
        for (const BermPlan& b : L.berm) c->berm_fixing.push_back(b.fixing);
        for (size_t bi = 0; bi < L.berm.size(); bi++)
            run_bermudan(c, L.berm[bi], (int)bi, adj, a.contrib, wsum, wm2, stride);
        launch_reduce_partials(wsum, wm2, (int)num_waves, stride, L.num_slots, 0, P, dout, dout + cols, s);
        std::vector<double> s2((size_t)cols), m2b((size_t)cols);
        HIP_CHECK(hipMemcpyAsync(s2.data(), dout, L.num_slots * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(m2b.data(), dout + cols, L.num_slots * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (int ss : L.swap_slot) { sums[(size_t)ss] = s2[(size_t)ss]; m2s[(size_t)ss] = m2b[(size_t)ss]; }
        for (const BermPlan& b : L.berm) { sums[(size_t)b.slot] = s2[(size_t)b.slot]; m2s[(size_t)b.slot] = m2b[(size_t)b.slot]; }
    }
    tm.mark(4);
    HIP_CHECK(hipEventSynchronize(c->ev[4]));
    c->timings[0] = regen ? tm.span(0, 1) : 0.0;
    c->timings[1] = tm.span(2, 3);
    c->timings[2] = tm.span(3, 4);
    c->timings[4] = tm.span(0, 4);
}

} // namespace

extern "C" {

const char* fm_last_error(void) { return g_last_error.c_str(); }

int fm_lmm_create(const fm_lmm_config* config, fm_lmm** out) {
    FM_TRY
    require(out != nullptr, FM_ERR_INVALID_ARGUMENT, "output handle pointer is NULL");
    *out = nullptr;
    validate_config(config);
    std::unique_ptr<fm_lmm> c(new fm_lmm());
    c->cfg = *config;
    c->times.assign(config->times, config->times + config->num_time_steps + 1);
    c->tenor.assign(config->tenor, config->tenor + config->num_libors + 1);
    c->cfg.times = nullptr; c->cfg.tenor = nullptr;
    if (c->cfg.checkpoint_stride <= 0) c->cfg.checkpoint_stride = 4;
    HIP_CHECK(hipGetDevice(&c->device));
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_CHECK(hipEventCreate(&e));
    *out = c.release();
    return FM_OK;
    FM_CATCH
}

void fm_lmm_destroy(fm_lmm* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int fm_lmm_set_initial_forwards(fm_lmm* c, const double* forwards) {
    FM_TRY
    require(c && forwards, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    c->L0.assign(forwards, forwards + c->n());
    for (double v : c->L0) require(std::isfinite(v), FM_ERR_INVALID_ARGUMENT, "initial forwards must be finite");
    c->has_L0 = true; c->tables_dirty = true; c->simulated = false;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_set_covariance(fm_lmm* c, const double* factor_loading, const double* variance) {
    FM_TRY
    require(c && factor_loading && variance, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t nl = (size_t)c->nT() * c->n() * c->F(), nv = (size_t)c->nT() * c->n();
    c->lam.assign(factor_loading, factor_loading + nl);
    c->var.assign(variance, variance + nv);
    for (double v : c->lam) require(std::isfinite(v), FM_ERR_INVALID_ARGUMENT, "factor loadings must be finite");
    for (double v : c->var) require(std::isfinite(v), FM_ERR_INVALID_ARGUMENT, "variances must be finite");
    c->has_cov = true; c->tables_dirty = true; c->simulated = false;
    c->has_vf = false;      // a factorization set before belongs to the previous tables
    return FM_OK;
    FM_CATCH
}

int fm_lmm_set_local_volatility(fm_lmm* c, double alpha, const double* offsets) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> lk(c->mu);
    if (offsets) {
        require(std::isfinite(alpha), FM_ERR_INVALID_ARGUMENT, "the local-volatility blend must be finite");
        std::vector<double> b(offsets, offsets + (size_t)c->nT() * c->n());
        for (double v : b) require(std::isfinite(v), FM_ERR_INVALID_ARGUMENT, "local-volatility offsets must be finite");
        c->lvb = std::move(b);
        c->lv_alpha = alpha;
        c->has_lv = true;
    } else {
        c->lvb.clear();
        c->lv_alpha = 0.0;
        c->has_lv = false;
    }
    c->tables_dirty = true; c->simulated = false;
    return FM_OK;
    FM_CATCH
}

extern "C++" {
namespace {
// Checks of a scaling's driver context, under both contexts' locks; generates its increments.
void prepare_driver(fm_lmm* c, fm_lmm* driver, std::initializer_list<int32_t> factors) {
    require(driver->device == c->device, FM_ERR_INVALID_ARGUMENT, "the driver lives on another device");
    require(driver->P() == c->P() && driver->cfg.path_offset == c->cfg.path_offset, FM_ERR_INVALID_ARGUMENT,
            "the driver must simulate the same paths");
    require(driver->times == c->times, FM_ERR_INVALID_ARGUMENT, "the driver must have the same time discretization");
    require(!driver->fp32(), FM_ERR_UNSUPPORTED, "the driver's increments must be FP64");
    for (int32_t f : factors)
        require(f >= 0 && f < driver->F(), FM_ERR_INVALID_ARGUMENT, "driver factor out of range");
    ensure_brownian(driver);
    HIP_CHECK(hipStreamSynchronize(driver->stream));
    ensure_time_steps(c);
}
} // namespace
} // extern "C++"

int fm_lmm_set_stochastic_volatility(fm_lmm* c, double nu, double rho, fm_lmm* driver, int32_t factor_a,
                                     int32_t factor_b) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    if (!driver) {
        std::lock_guard<std::mutex> lk(c->mu);
        c->has_sv = false;
        c->tables_dirty = true; c->simulated = false;
        return FM_OK;
    }
    std::unique_lock<std::mutex> lk_c(c->mu, std::defer_lock), lk_d(driver->mu, std::defer_lock);
    if (driver == c) lk_c.lock(); else std::lock(lk_c, lk_d);
    require(std::isfinite(nu) && std::isfinite(rho) && rho >= -1.0 && rho <= 1.0, FM_ERR_INVALID_ARGUMENT,
            "nu must be finite and rho in [-1, 1]");
    DeviceGuard g(c->device);
    prepare_driver(c, driver, {factor_a, factor_b});
    // drift nu.squared().mult(-0.5), loadings rho.mult(nu) and rho.squared().sub(1).mult(-1).sqrt().mult(nu)
    // (LIBORCovarianceModelStochasticVolatility.java:206-213)
    const double drift = (nu * nu) * -0.5;
    const double l0 = rho * nu;
    const double l1 = std::sqrt((rho * rho - 1.0) * -1.0) * nu;
    void* out = c->d_svs.get((size_t)c->nT() * c->P() * c->real_size());
    launch_stochastic_volatility(driver->d_dw.as<double>(), c->nT(), driver->F(), factor_a, factor_b, c->P(),
                                 c->d_times.as<double>(), drift, l0, l1, out, c->fp32() ? 1 : 0, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->has_sv = true;
    c->tables_dirty = true; c->simulated = false;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_set_heston_volatility(fm_lmm* c, double kappa, double theta, double xi, fm_lmm* driver, int32_t factor) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    if (!driver) return fm_lmm_set_stochastic_volatility(c, 0.0, 0.0, nullptr, 0, 0);
    std::unique_lock<std::mutex> lk_c(c->mu, std::defer_lock), lk_d(driver->mu, std::defer_lock);
    if (driver == c) lk_c.lock(); else std::lock(lk_c, lk_d);
    require(std::isfinite(kappa) && std::isfinite(theta) && std::isfinite(xi), FM_ERR_INVALID_ARGUMENT,
            "kappa, theta and xi must be finite");
    DeviceGuard g(c->device);
    prepare_driver(c, driver, {factor});
    void* out = c->d_svs.get((size_t)c->nT() * c->P() * c->real_size());
    launch_heston_volatility(driver->d_dw.as<double>(), c->nT(), driver->F(), factor, c->P(), c->d_times.as<double>(),
                             kappa, theta, xi, out, c->fp32() ? 1 : 0, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->has_sv = true;
    c->tables_dirty = true; c->simulated = false;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_set_discount_curve(fm_lmm* c, const double* df) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> lk(c->mu);
    if (df) {
        c->pd.assign(df, df + c->n() + 1);
        for (double v : c->pd) require(std::isfinite(v) && v > 0.0, FM_ERR_INVALID_ARGUMENT, "discount factors must be positive");
        c->has_pd = true;
    } else { c->pd.clear(); c->has_pd = false; }
    return FM_OK;
    FM_CATCH
}

int fm_lmm_simulate(fm_lmm* c) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> lk(c->mu);
    require(c->cfg.store_paths != 0, FM_ERR_STATE, "context created without store_paths");
    DeviceGuard g(c->device);
    Lowered none;
    std::vector<double> sums, m2s;
    run_pass(c, none, true, sums, m2s);
    c->simulated = true;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_get_process_value(fm_lmm* c, int32_t t, int32_t j, double* out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(t >= 0 && t <= c->nT() && j >= 0 && j < c->n(), FM_ERR_INVALID_ARGUMENT, "process index out of range");
    require(c->simulated, FM_ERR_STATE, "process not simulated (call fm_lmm_simulate)");
    DeviceGuard g(c->device);
    const int P = c->P();
    const int r = c->row_of[(size_t)t * c->n() + j];
    if (r < 0) { std::fill(out, out + P, c->init_value[(size_t)j]); return FM_OK; }
    if (c->fp32()) {
        std::vector<float> tmp((size_t)P);
        HIP_CHECK(hipMemcpy(tmp.data(), (const float*)c->d_paths.p + (size_t)r * P, P * sizeof(float), hipMemcpyDeviceToHost));
        for (int p = 0; p < P; p++) out[p] = tmp[(size_t)p];
    } else {
        HIP_CHECK(hipMemcpy(out, (const double*)c->d_paths.p + (size_t)r * P, P * sizeof(double), hipMemcpyDeviceToHost));
    }
    return FM_OK;
    FM_CATCH
}

int fm_lmm_process_value_device_ptr(fm_lmm* c, int32_t t, int32_t j, const void** out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(t >= 0 && t <= c->nT() && j >= 0 && j < c->n(), FM_ERR_INVALID_ARGUMENT, "process index out of range");
    require(c->simulated, FM_ERR_STATE, "process not simulated (call fm_lmm_simulate)");
    const int r = c->row_of[(size_t)t * c->n() + j];
    require(r >= 0, FM_ERR_STATE, "time-0 values are deterministic (no device realisations)");
    *out = (const char*)c->d_paths.p + (size_t)r * c->P() * c->real_size();
    return FM_OK;
    FM_CATCH
}

int fm_lmm_get_numeraire(fm_lmm* c, double time, double* out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(c->simulated, FM_ERR_STATE, "process not simulated (call fm_lmm_simulate)");
    DeviceGuard g(c->device);
    const int k = c->libor_index(time);
    require(k >= 0 && k <= c->n(), FM_ERR_UNSUPPORTED, "numeraire requested off the tenor grid");
    // spot: N(T_k) is known at the time index of T_{k-1}; terminal: at that of T_k
    const int known = c->terminal() ? k : std::max(k - 1, 0);
    require(c->tenor_ti[(size_t)known] <= c->nT(), FM_ERR_INVALID_ARGUMENT, "numeraire date beyond the simulation");
    const int P = c->P();
    // deterministic numeraires are Scalars: N(T_0) = 1 (spot), N(T_0) and N(T_n) = 1 (terminal)
    const bool scalar_num = c->terminal() ? (k == 0 || k == c->n()) : k == 0;
    if (scalar_num) std::fill(out, out + P, c->terminal() && k == 0 ? c->n0_terminal : 1.0);
    else HIP_CHECK(hipMemcpy(out, c->d_num.as<double>() + (size_t)k * P, P * sizeof(double), hipMemcpyDeviceToHost));
    if (c->has_pd) {
        // N(T) * avg(N(0)/N(T)) / P^d(T) (LMM:869-874); Scalar numeraires divide as
        // (1/adj) (N zb) (stochastic/Scalar.java:296-298)
        const std::vector<double> adj = adjustment_chain(c);
        const double zbk = c->zb[(size_t)k], ak = adj[(size_t)k];
        const bool scalar_adj = c->terminal() ? scalar_num : k <= 1;
        if (scalar_adj) { for (int p = 0; p < P; p++) out[p] = (1.0 / ak) * (out[p] * zbk); }
        else { for (int p = 0; p < P; p++) out[p] = (out[p] * zbk) / ak; }
    }
    return FM_OK;
    FM_CATCH
}

int fm_lmm_process_value_rv(fm_lmm* c, int32_t t, int32_t j, fm_rv** out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    require(t >= 0 && t <= c->nT() && j >= 0 && j < c->n(), FM_ERR_INVALID_ARGUMENT, "process index out of range");
    require(c->simulated, FM_ERR_STATE, "process not simulated (call fm_lmm_simulate)");
    DeviceGuard g(c->device);
    const int r = c->row_of[(size_t)t * c->n() + j];
    if (r < 0) { *out = rv_constant(c->times[(size_t)t], c->init_value[(size_t)j]); return FM_OK; }
    const char* src = (const char*)c->d_paths.p + (size_t)r * c->P() * c->real_size();
    *out = rv_from_device(c->device, src, c->fp32(), c->P(), c->times[(size_t)t]);
    return FM_OK;
    FM_CATCH
}

int fm_lmm_numeraire_rv(fm_lmm* c, double time, fm_rv** out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    require(c->simulated, FM_ERR_STATE, "process not simulated (call fm_lmm_simulate)");
    DeviceGuard g(c->device);
    const int k = c->libor_index(time);
    require(k >= 0 && k <= c->n(), FM_ERR_UNSUPPORTED, "numeraire requested off the tenor grid");
    const int known = c->terminal() ? k : std::max(k - 1, 0);
    require(c->tenor_ti[(size_t)known] <= c->nT(), FM_ERR_INVALID_ARGUMENT, "numeraire date beyond the simulation");
    const bool scalar_num = c->terminal() ? (k == 0 || k == c->n()) : k == 0;
    const std::vector<double> adj = c->has_pd ? adjustment_chain(c) : std::vector<double>();
    if (scalar_num) {
        double v = c->terminal() && k == 0 ? c->n0_terminal : 1.0;
        if (c->has_pd) v = (1.0 / adj[(size_t)k]) * (v * c->zb[(size_t)k]);
        *out = rv_constant(time, v);
        return FM_OK;
    }
    fm_rv* raw = rv_from_device(c->device, c->d_num.as<double>() + (size_t)k * c->P(), false, c->P(), time);
    if (!c->has_pd) { *out = raw; return FM_OK; }
    // N(T) * avg(N(0)/N(T)) / P^d(T) (LMM:869-874); spot N(T_1) is a Scalar: (1/adj) (N zb)
    const bool scalar_adj = !c->terminal() && k <= 1;
    try {
        *out = rv_scale_numeraire(raw, c->zb[(size_t)k], adj[(size_t)k], scalar_adj);
    } catch (...) {
        fm_rv_release(raw);
        throw;
    }
    fm_rv_release(raw);
    return FM_OK;
    FM_CATCH
}

int fm_lmm_weights_rv(fm_lmm* c, fm_rv** out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = rv_constant(0.0, 1.0 / (double)c->P());
    return FM_OK;
    FM_CATCH
}

int fm_lmm_get_brownian_increment(fm_lmm* c, int32_t t, int32_t k, double* out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(t >= 0 && t < c->nT() && k >= 0 && k < c->F(), FM_ERR_INVALID_ARGUMENT, "increment index out of range");
    DeviceGuard g(c->device);
    prepare_model(c);
    ensure_brownian(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const int P = c->P();
    const size_t es = c->real_size();
    const size_t pitch = (size_t)c->nT() * c->F() * es;
    std::vector<unsigned char> tmp((size_t)P * es);
    HIP_CHECK(hipMemcpy2D(tmp.data(), es, (const char*)c->d_dw.p + ((size_t)t * c->F() + k) * es, pitch, es, P,
                          hipMemcpyDeviceToHost));
    if (c->fp32()) { const float* f = (const float*)tmp.data(); for (int p = 0; p < P; p++) out[p] = f[p]; }
    else std::memcpy(out, tmp.data(), (size_t)P * sizeof(double));
    return FM_OK;
    FM_CATCH
}

int fm_lmm_value_products(fm_lmm* c, const fm_product* products, int32_t count, double* values, double* errors) {
    return fm::value_products_in(c, products, count, values, errors, nullptr, 0);
}

} // extern "C"

namespace fm {

int context_device(const fm_lmm* c) { return c->device; }

int value_products_in(fm_lmm* c, const fm_product* products, int32_t count, double* values, double* errors,
                      void* scratch, size_t scratch_bytes) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    require(count >= 0, FM_ERR_INVALID_ARGUMENT, "negative product count");
    require(count == 0 || values != nullptr, FM_ERR_INVALID_ARGUMENT, "values is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard g(c->device);
    std::unique_ptr<NoAllocScope> no_alloc;
    if (scratch) no_alloc.reset(new NoAllocScope());
    prepare_model(c);
    Lowered L = lower_products(c, products, count);
    ExternalScratch xs;
    if (scratch) {
        require(L.num_contrib == 0 && L.berm.empty(), FM_ERR_UNSUPPORTED,
                "without allocation only single-date products (swaptions, caplets) are valued: swaps and "
                "Bermudans need per-path columns");
        const size_t need = external_scratch_bytes(c, L);
        require(need <= scratch_bytes, FM_ERR_INVALID_ARGUMENT,
                "the pass needs " + std::to_string(need) + " bytes of scratch, more than the " +
                std::to_string(scratch_bytes) + " bytes of the no-allocation entry point (fewer paths or products)");
        xs.base = (char*)scratch;
        xs.bytes = scratch_bytes;
    }
    std::vector<double> sums, m2s;
    run_pass(c, L, false, sums, m2s, nullptr, 1, nullptr, scratch ? &xs : nullptr);
    const std::vector<double> adj = adjustment_chain(c);
    const double P = (double)c->P();
    for (int q = 0; q < count; q++) {
        const int slot = L.slot_of[(size_t)q];
        const int k = L.slot_k[(size_t)slot];
        const double scale = k >= 0 ? value_scale(c, k, adj) : 1.0;
        const double mean = sums[(size_t)slot] / P;
        const double var = m2s[(size_t)slot] / P;
        values[q] = scale * mean;
        if (errors) errors[q] = std::fabs(scale) * std::sqrt(var) / std::sqrt(P);
        require(std::isfinite(values[q]), FM_ERR_NUMERICAL, "product value is not finite (NaN/Inf on some path)");
    }
    return FM_OK;
    FM_CATCH
}

} // namespace fm

extern "C" {

int fm_lmm_set_volatility_factorization(fm_lmm* c, const double* volatility, const double* factors,
                                        const double* variance_factor) {
    FM_TRY
    require(c && volatility && factors, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    c->vf_sig.assign(volatility, volatility + (size_t)c->nT() * c->n());
    c->vf_fac.assign(factors, factors + (size_t)c->n() * c->F());
    if (variance_factor) c->vf_c.assign(variance_factor, variance_factor + c->n());
    else c->vf_c.assign((size_t)c->n(), 1.0);
    c->has_vf = true;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_set_dirac_delta(fm_lmm* c, int32_t method, double width) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    require(method == FM_DIRAC_DISCRETE_DELTA || method == FM_DIRAC_ONE || method == FM_DIRAC_ZERO ||
                method == FM_DIRAC_REGRESSION_ON_DENSITY || method == FM_DIRAC_REGRESSION_ON_DISTRIBUTION,
            FM_ERR_INVALID_ARGUMENT, "unknown Dirac delta approximation method");
    require(method != FM_DIRAC_REGRESSION_ON_DENSITY && method != FM_DIRAC_REGRESSION_ON_DISTRIBUTION,
            FM_ERR_UNSUPPORTED, "the regression Dirac delta approximations are not implemented");
    require(width > 0.0, FM_ERR_INVALID_ARGUMENT, "the Dirac delta width must be positive");
    std::lock_guard<std::mutex> lk(c->mu);
    c->dirac_method = method;
    c->dirac_width = width;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_adjoint(fm_lmm* c, const fm_product* products, int32_t count, const double* weights, double* value,
                   double* d_initial_forwards, double* d_volatility) {
    FM_TRY
    require(c != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL context");
    require(count >= 1 && products != nullptr, FM_ERR_INVALID_ARGUMENT, "no products");
    std::lock_guard<std::mutex> lk(c->mu);
    require(c->has_vf, FM_ERR_STATE, "volatility factorization not set (fm_lmm_set_volatility_factorization)");
    require(c->cfg.precision != FM_PRECISION_FP64_LOGSTATE, FM_ERR_UNSUPPORTED,
            "the adjoint runs on the FP64 and FP32_MIXED path kernels");
    require(c->plain_euler(), FM_ERR_UNSUPPORTED,
            "the adjoint differentiates the spot-measure Euler scheme (EULER_FUNCTIONAL)");
    {
        // the factorization must be the one of the tables that drive the path:
        // lambda_jk(t_i) = sigma_j(t_i) f_jk and variance_j(t_i) = sigma_j(t_i)^2 c_j
        const int n = c->n(), nT = c->nT(), F = c->F();
        double lscale = 0.0, vscale = 0.0, lworst = 0.0, vworst = 0.0;
        for (double v : c->lam) lscale = std::max(lscale, std::fabs(v));
        for (double v : c->var) vscale = std::max(vscale, std::fabs(v));
        for (int i = 0; i < nT; i++)
            for (int j = 0; j < n; j++) {
                const double sg = c->vf_sig[(size_t)i * n + j];
                for (int k = 0; k < F; k++)
                    lworst = std::max(lworst, std::fabs(c->lam[((size_t)i * n + j) * F + k] - sg * c->vf_fac[(size_t)j * F + k]));
                vworst = std::max(vworst, std::fabs(c->var[(size_t)i * n + j] - sg * sg * c->vf_c[(size_t)j]));
            }
        require(lworst <= 1e-12 * lscale && vworst <= 1e-12 * vscale, FM_ERR_STATE,
                "the volatility factorization does not reproduce the covariance tables "
                "(lambda = sigma f, variance = sigma^2 c)");
    }
    DeviceGuard g(c->device);
    prepare_model(c);
    const int n = c->n(), nT = c->nT(), P = c->P();
    hipStream_t s = c->stream;
    Lowered L = lower_products(c, products, count);
    const int stride = std::max(1, c->cfg.checkpoint_stride);
    const int ncp = nT / stride + 1;
    void* ckpt = c->d_ckpt.get((size_t)ncp * n * P * c->real_size());
    double* num = (double*)c->d_adj_num.get((size_t)(n + 1) * P * sizeof(double));
    hipEvent_t e0 = c->ev[5];
    HIP_CHECK(hipEventRecord(e0, s));
    std::vector<double> sums, m2s;
    // Bermudans under ONE or DISCRETE_DELTA: the pricing pass records each decision's trigger and
    // choice for the induction's adjoint
    const bool berm_adjoint = !L.berm.empty() && c->dirac_method != FM_DIRAC_ZERO;
    if (berm_adjoint) {
        c->berm_trig_off.assign(L.berm.size(), 0);
        int total = 0;
        for (size_t bi = 0; bi < L.berm.size(); bi++) {
            c->berm_trig_off[bi] = total;
            total += (int)L.berm[bi].ex_period.size();
        }
        c->d_berm_trig.get((size_t)std::max(total, 1) * P * sizeof(double));
        c->d_berm_diff.get((size_t)std::max(total, 1) * P * sizeof(double));
        c->d_berm_pinv.get((size_t)std::max(total, 1) * kBermPinv * sizeof(double));
        c->berm_record = true;
    }
    // the first component the reverse sweep reads from each checkpoint's segment [c, c + stride) on:
    // the alive components of its steps (first_alive[c - 1] for the state the step into c left),
    // the forwards of the products evaluated from c on, and the numeraire's forwards from
    // kcur_at(c - 1) on; rows below are neither written nor read
    std::vector<int32_t> cfirst((size_t)ncp, 0);
    for (int cc = 0; cc < ncp; cc++) {
        const int c0 = cc * stride;
        int f = n;
        if (c0 < nT) f = std::min(f, c->first_alive[(size_t)c0]);
        if (c0 >= 1) f = std::min({f, c->first_alive[(size_t)c0 - 1], kcur_at(c, c0 - 1)});
        for (const KProduct& k : L.kp)
            if (k.eval_index >= c0) f = std::min(f, (int)k.first_j);
        cfirst[(size_t)cc] = c0 == 0 ? 0 : f;
    }
    const int32_t* d_cfirst = upload(c->d_ckpt_first, cfirst, s);
    try {
        run_pass(c, L, false, sums, m2s, ckpt, stride, num, nullptr, d_cfirst);
    } catch (...) {
        c->berm_record = false;
        throw;
    }
    c->berm_record = false;
    const double fwd_ms = c->timings[4];
    // objective weights per lowered product: w_q * (adjustment scale) / P
    const std::vector<double> adj = adjustment_chain(c);
    std::vector<double> wq((size_t)count, 1.0);
    if (weights) for (int q = 0; q < count; q++) wq[(size_t)q] = weights[q];
    // map each lowered entry back to its product: swaptions/caplets by slot, swap periods by column
    std::vector<int> slot_to_q((size_t)L.num_slots, -1), col_to_q((size_t)L.num_contrib, -1);
    for (int q = 0; q < count; q++) slot_to_q[(size_t)L.slot_of[(size_t)q]] = q;
    for (size_t i = 0; i < L.swap_slot.size(); i++) {
        const int q = slot_to_q[(size_t)L.swap_slot[i]];
        for (int cc = 0; cc < L.swap_ncol[i]; cc++) col_to_q[(size_t)(L.swap_col0[i] + cc)] = q;
    }
    // Bermudan periods: swap periods gated per path by the exercise period of the pass above
    std::vector<int> col_berm((size_t)L.num_contrib, -1), col_period((size_t)L.num_contrib, -1);
    for (size_t bi = 0; bi < L.berm.size(); bi++) {
        const BermPlan& b = L.berm[bi];
        for (size_t i = 0; i < b.per_col.size(); i++) {
            col_to_q[(size_t)b.per_col[i]] = slot_to_q[(size_t)b.slot];
            col_berm[(size_t)b.per_col[i]] = (int)bi;
            col_period[(size_t)b.per_col[i]] = (int)i;
        }
    }
    std::vector<KProduct> kp = L.kp;
    for (auto& k : kp) {
        if (k.type == K_BERM_BASIS) {
            k.weight = 0.0;
        } else if (k.type == FM_PRODUCT_SWAP) {
            const int q = col_to_q[(size_t)k.out_slot];
            const double sc = c->has_pd ? adj[(size_t)k.num_k] / c->zb[(size_t)k.num_k] : 1.0;
            k.weight = wq[(size_t)q] * sc / (double)P;
            const int bi = col_berm[(size_t)k.out_slot];
            if (bi >= 0) {
                k.type = K_BERM_PERIOD;
                k.is_floorlet = col_period[(size_t)k.out_slot];
                k.pad2 = bi;
                k.flags = 1 | (L.berm[(size_t)bi].callable ? 2 : 0);
            }
        } else {
            const int q = slot_to_q[(size_t)k.out_slot];
            const double sc = c->has_pd ? adj[(size_t)k.num_k] / c->zb[(size_t)k.num_k] : 1.0;
            k.weight = wq[(size_t)q] * sc / (double)P;
        }
    }
    std::vector<double> kweights(kp.size());
    for (size_t i = 0; i < kp.size(); i++) kweights[i] = kp[i].weight;
    std::vector<int32_t> kcur((size_t)nT + 1);
    for (int t = 0; t <= nT; t++) kcur[(size_t)t] = kcur_at(c, t);
    StepArgs a{};
    a.nT = nT; a.n = n; a.F = c->F(); a.P = P;
    a.precision = c->cfg.precision; a.state_space = c->cfg.state_space;
    a.times = c->d_times.as<double>(); a.first_alive = c->d_first.as<int32_t>();
    a.tenor_time_index = c->d_tenor_ti.as<int32_t>();
    a.rec = c->d_rec.p;
    a.lognormal = c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL ? 1 : 0;
    a.capped = std::isinf(c->cfg.libor_cap) ? 0 : 1;
    a.delta = c->d_delta.as<double>();
    a.L0 = c->d_L0.as<double>(); a.dw = c->d_dw.p; a.cap = c->cfg.libor_cap; a.scheme = c->cfg.scheme;
    a.num_products = (int)kp.size();
    a.products = upload(c->d_products, kp, s);
    a.periods = upload(c->d_periods, L.per, s);
    a.numeraire = num; a.ckpt = ckpt; a.ckpt_stride = stride; a.ckpt_first = d_cfirst;
    AdjArgs ga{};
    ga.stride = stride;
    ga.kcur_at = upload(c->d_kcur, kcur, s);
    ga.weight = upload(c->d_scale, kweights, s);
    ga.berm_ex = L.berm.empty() ? nullptr : c->d_berm_ex.as<int32_t>();
    ga.berm_weight = nullptr;
    if (berm_adjoint) {
        // dV/d(payoff) of every Bermudan period per path, [contribution column][P]
        double* w = (double*)c->d_berm_weight.get((size_t)L.num_contrib * P * sizeof(double));
        for (size_t bi = 0; bi < L.berm.size(); bi++)
            run_bermudan_adjoint(c, L.berm[bi], (int)bi, adj, c->d_contrib.as<double>(), w);
        ga.berm_weight = w;
    }
    ga.seg = c->d_seg.get((size_t)std::max(stride - 2, 1) * n * P * c->real_size());
    // factor matrix padded to the factor capacity and -sigma_j(t_i) c_j, in the path precision
    std::vector<double> fpad((size_t)n * fcap_of(c->F()), 0.0), vterm((size_t)nT * n, 0.0);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < c->F(); k++) fpad[(size_t)j * fcap_of(c->F()) + k] = c->vf_fac[(size_t)j * c->F() + k];
    if (c->cfg.state_space == FM_STATE_SPACE_LOGNORMAL)
        for (int i = 0; i < nT; i++)
            for (int j = 0; j < n; j++) vterm[(size_t)i * n + j] = -c->vf_sig[(size_t)i * n + j] * c->vf_c[(size_t)j];
    // host sources stay alive until the stream synchronisation at the end of this call
    const std::vector<float> fpad_f(fpad.begin(), fpad.end()), vterm_f(vterm.begin(), vterm.end());
    if (c->fp32()) {
        ga.fsig = upload(c->d_fsig, fpad_f, s);
        ga.vterm = upload(c->d_sig, vterm_f, s);
    } else {
        ga.fsig = upload(c->d_fsig, fpad, s);
        ga.vterm = upload(c->d_sig, vterm, s);
    }
    const int64_t num_waves = ((int64_t)P + 63) / 64;
    const int gstride = nT * n + n;
    // every wavefront's row (nT x n dV/dsigma entries, then n dV/dL(0)) is written in full by the sweep
    ga.wave_grad = (double*)c->d_wave_grad.get((size_t)num_waves * gstride * sizeof(double));
    HIP_CHECK(hipEventRecord(c->ev[0], s));
    if (c->fp32()) launch_lmm_adjoint_fp32(a, ga, s); else launch_lmm_adjoint_fp64(a, ga, s);
    HIP_CHECK(hipEventRecord(c->ev[1], s));
    double* gout = (double*)c->d_grad_out.get((size_t)(gstride + reduce_columns_scratch(gstride)) * sizeof(double));
    launch_reduce_columns(ga.wave_grad, (int)num_waves, gstride, gstride, gout + gstride, gout, s);
    std::vector<double> grad((size_t)gstride);
    HIP_CHECK(hipMemcpyAsync(grad.data(), gout, gstride * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipEventRecord(c->ev[2], s));
    HIP_CHECK(hipStreamSynchronize(s));
    float ms_adj = 0, ms_all = 0;
    HIP_CHECK(hipEventElapsedTime(&ms_adj, c->ev[0], c->ev[1]));
    HIP_CHECK(hipEventElapsedTime(&ms_all, e0, c->ev[2]));
    c->timings[3] = ms_adj;
    c->timings[4] = ms_all;
    (void)fwd_ms;
    if (value) {
        double v = 0.0;
        for (int q = 0; q < count; q++) {
            const int slot = L.slot_of[(size_t)q];
            const int k = L.slot_k[(size_t)slot];
            const double sc = (k >= 0 && c->has_pd) ? adj[(size_t)k] / c->zb[(size_t)k] : 1.0;
            v += wq[(size_t)q] * sc * sums[(size_t)slot] / (double)P;
        }
        *value = v;
    }
    if (d_volatility) for (size_t i = 0; i < (size_t)nT * n; i++) d_volatility[i] = grad[i];
    if (d_initial_forwards) for (int j = 0; j < n; j++) d_initial_forwards[j] = grad[(size_t)nT * n + j];
    for (double x : grad) require(std::isfinite(x), FM_ERR_NUMERICAL, "adjoint is not finite");
    return FM_OK;
    FM_CATCH
}

// ---- serialisation: "FMLMMSER", version, configuration, grids, tables (native little-endian)
extern "C++" {
namespace {
constexpr char kSerMagic[8] = {'F', 'M', 'L', 'M', 'M', 'S', 'E', 'R'};
constexpr int32_t kSerVersion = 2;      // 2: the adjoint's Dirac delta method and width

struct Writer {
    std::vector<unsigned char> b;
    template <class T> void put(const T& v) {
        const unsigned char* p = (const unsigned char*)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    void put_doubles(const std::vector<double>& v) {
        put<int64_t>((int64_t)v.size());
        const unsigned char* p = (const unsigned char*)v.data();
        b.insert(b.end(), p, p + v.size() * sizeof(double));
    }
};

struct Reader {
    const unsigned char* p;
    int64_t left;
    template <class T> T get() {
        require(left >= (int64_t)sizeof(T), FM_ERR_INVALID_ARGUMENT, "serialised context is truncated");
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        left -= (int64_t)sizeof(T);
        return v;
    }
    std::vector<double> get_doubles(int64_t expect) {
        const int64_t n = get<int64_t>();
        require(n == expect, FM_ERR_INVALID_ARGUMENT, "serialised table has an unexpected length");
        require(left >= n * (int64_t)sizeof(double), FM_ERR_INVALID_ARGUMENT, "serialised context is truncated");
        std::vector<double> v((size_t)n);
        std::memcpy(v.data(), p, (size_t)n * sizeof(double));
        p += n * (int64_t)sizeof(double);
        left -= n * (int64_t)sizeof(double);
        return v;
    }
};
} // namespace
} // extern "C++"

int fm_lmm_serialize(fm_lmm* c, void* buffer, int64_t* size) {
    FM_TRY
    require(c && size, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(!c->has_sv, FM_ERR_UNSUPPORTED,
            "a stochastic volatility scaling depends on another context's Brownian motion and is not serialised");
    Writer w;
    w.b.insert(w.b.end(), kSerMagic, kSerMagic + 8);
    w.put<int32_t>(kSerVersion);
    const fm_lmm_config& g = c->cfg;
    for (int32_t v : {g.num_time_steps, g.num_libors, g.num_factors, g.num_paths, g.seed, g.measure, g.state_space,
                      g.scheme, g.precision, g.store_paths, g.checkpoint_stride, g.regenerate_increments})
        w.put<int32_t>(v);
    w.put<int64_t>(g.path_offset);
    w.put<double>(g.libor_cap);
    w.put_doubles(c->times);
    w.put_doubles(c->tenor);
    w.put<int32_t>(c->has_L0); w.put<int32_t>(c->has_cov); w.put<int32_t>(c->has_pd); w.put<int32_t>(c->has_vf);
    w.put<int32_t>(c->has_lv);
    if (c->has_L0) w.put_doubles(c->L0);
    if (c->has_cov) { w.put_doubles(c->lam); w.put_doubles(c->var); }
    if (c->has_pd) w.put_doubles(c->pd);
    if (c->has_vf) { w.put_doubles(c->vf_sig); w.put_doubles(c->vf_fac); w.put_doubles(c->vf_c); }
    if (c->has_lv) { w.put<double>(c->lv_alpha); w.put_doubles(c->lvb); }
    w.put<int32_t>(c->dirac_method);
    w.put<double>(c->dirac_width);
    const int64_t need = (int64_t)w.b.size();
    if (buffer) {
        require(*size >= need, FM_ERR_INVALID_ARGUMENT, "buffer too small for the serialised context");
        std::memcpy(buffer, w.b.data(), (size_t)need);
    }
    *size = need;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_deserialize(const void* buffer, int64_t size, fm_lmm** out) {
    FM_TRY
    require(buffer && out && size > 0, FM_ERR_INVALID_ARGUMENT, "invalid arguments");
    *out = nullptr;
    Reader r{(const unsigned char*)buffer, size};
    char magic[8];
    for (char& m : magic) m = r.get<char>();
    require(std::memcmp(magic, kSerMagic, 8) == 0, FM_ERR_INVALID_ARGUMENT, "not a serialised context");
    const int32_t version = r.get<int32_t>();
    require(version == 1 || version == kSerVersion, FM_ERR_INVALID_ARGUMENT, "unsupported serialisation version");
    fm_lmm_config g{};
    g.num_time_steps = r.get<int32_t>(); g.num_libors = r.get<int32_t>(); g.num_factors = r.get<int32_t>();
    g.num_paths = r.get<int32_t>(); g.seed = r.get<int32_t>(); g.measure = r.get<int32_t>();
    g.state_space = r.get<int32_t>(); g.scheme = r.get<int32_t>(); g.precision = r.get<int32_t>();
    g.store_paths = r.get<int32_t>(); g.checkpoint_stride = r.get<int32_t>(); g.regenerate_increments = r.get<int32_t>();
    g.path_offset = r.get<int64_t>();
    g.libor_cap = r.get<double>();
    require(g.num_time_steps >= 1 && g.num_libors >= 1 && g.num_factors >= 1 && g.num_time_steps < (1 << 24) &&
            g.num_libors <= 80 && g.num_factors <= 10, FM_ERR_INVALID_ARGUMENT, "serialised shape out of range");
    const int nT = g.num_time_steps, n = g.num_libors, F = g.num_factors;
    const std::vector<double> times = r.get_doubles((int64_t)nT + 1), tenor = r.get_doubles((int64_t)n + 1);
    g.times = times.data();
    g.tenor = tenor.data();
    const int32_t hL0 = r.get<int32_t>(), hcov = r.get<int32_t>(), hpd = r.get<int32_t>(), hvf = r.get<int32_t>();
    const int32_t hlv = r.get<int32_t>();
    std::vector<double> L0, lam, var, pd, sig, fac, cv, lvb;
    double lv_alpha = 0.0;
    if (hL0) L0 = r.get_doubles(n);
    if (hcov) { lam = r.get_doubles((int64_t)nT * n * F); var = r.get_doubles((int64_t)nT * n); }
    if (hpd) pd = r.get_doubles((int64_t)n + 1);
    if (hvf) { sig = r.get_doubles((int64_t)nT * n); fac = r.get_doubles((int64_t)n * F); cv = r.get_doubles(n); }
    if (hlv) { lv_alpha = r.get<double>(); lvb = r.get_doubles((int64_t)nT * n); }
    int32_t dirac_method = FM_DIRAC_DISCRETE_DELTA;      // version 1: the defaults
    double dirac_width = 0.05;
    if (version >= 2) { dirac_method = r.get<int32_t>(); dirac_width = r.get<double>(); }
    require(r.left == 0, FM_ERR_INVALID_ARGUMENT, "trailing bytes after the serialised context");
    fm_lmm* c = nullptr;
    int rc = fm_lmm_create(&g, &c);
    if (rc != FM_OK) return rc;
    std::unique_ptr<fm_lmm, void (*)(fm_lmm*)> guard(c, fm_lmm_destroy);
    if (hL0 && (rc = fm_lmm_set_initial_forwards(c, L0.data())) != FM_OK) return rc;
    if (hcov && (rc = fm_lmm_set_covariance(c, lam.data(), var.data())) != FM_OK) return rc;
    if (hpd && (rc = fm_lmm_set_discount_curve(c, pd.data())) != FM_OK) return rc;
    if (hvf && (rc = fm_lmm_set_volatility_factorization(c, sig.data(), fac.data(), cv.data())) != FM_OK) return rc;
    if (hlv && (rc = fm_lmm_set_local_volatility(c, lv_alpha, lvb.data())) != FM_OK) return rc;
    if ((rc = fm_lmm_set_dirac_delta(c, dirac_method, dirac_width)) != FM_OK) return rc;
    *out = guard.release();
    return FM_OK;
    FM_CATCH
}

int fm_lmm_bermudan_exercise_times(fm_lmm* c, int32_t bermudan, double* out) {
    FM_TRY
    require(c && out, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    require(bermudan >= 0 && bermudan < (int)c->berm_fixing.size(), FM_ERR_STATE,
            "no Bermudan swaption with this index in the last valuation");
    DeviceGuard g(c->device);
    const int P = c->P();
    std::vector<int32_t> ex((size_t)P);
    HIP_CHECK(hipMemcpy(ex.data(), c->d_berm_ex.as<int32_t>() + (size_t)bermudan * P, (size_t)P * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    const std::vector<double>& fix = c->berm_fixing[(size_t)bermudan];
    for (int p = 0; p < P; p++) out[p] = ex[(size_t)p] >= 0 ? fix[(size_t)ex[(size_t)p]] : HUGE_VAL;
    return FM_OK;
    FM_CATCH
}

int fm_lmm_last_timings(fm_lmm* c, double* ms, int32_t count) {
    FM_TRY
    require(c && ms, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int i = 0; i < count && i < 5; i++) ms[i] = c->timings[i];
    return FM_OK;
    FM_CATCH
}

static int mt_generate(int64_t seed, uint64_t offset, uint64_t count, std::vector<uint32_t>& words) {
    StreamLayout lay = choose_layout(offset + count + (offset + count) % 2);
    uint32_t seedw[624];
    mt_seed_window(seed, seedw);
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf st, poly, out;
    try {
        uint32_t* dst = (uint32_t*)st.get((size_t)lay.num_sub * 624 * sizeof(uint32_t));
        launch_jump_tree(lay, seedw, dst, (uint64_t*)poly.get((size_t)64 * 312 * sizeof(uint64_t)), s);
        uint32_t* dout = (uint32_t*)out.get((size_t)(offset + count) * sizeof(uint32_t));
        launch_words(lay, dst, offset + count, dout, s);
        words.resize((size_t)count);
        HIP_CHECK(hipMemcpyAsync(words.data(), dout + offset, count * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) { (void)hipStreamDestroy(s); throw; }
    HIP_CHECK(hipStreamDestroy(s));
    return FM_OK;
}

int fm_mt19937_words(int64_t seed, uint64_t offset, uint64_t count, uint32_t* out) {
    FM_TRY
    require(out != nullptr || count == 0, FM_ERR_INVALID_ARGUMENT, "NULL output");
    if (count == 0) return FM_OK;
    std::vector<uint32_t> w;
    mt_generate(seed, offset, count, w);
    std::memcpy(out, w.data(), count * sizeof(uint32_t));
    return FM_OK;
    FM_CATCH
}

int fm_mt19937_doubles(int64_t seed, uint64_t offset, uint64_t count, double* out) {
    FM_TRY
    require(out != nullptr || count == 0, FM_ERR_INVALID_ARGUMENT, "NULL output");
    if (count == 0) return FM_OK;
    std::vector<uint32_t> w;
    mt_generate(seed, 2 * offset, 2 * count, w);
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t bits = ((uint64_t)(w[2 * i] >> 6) << 26) | (uint64_t)(w[2 * i + 1] >> 6);
        out[i] = (double)bits * 0x1.0p-52;
    }
    return FM_OK;
    FM_CATCH
}

int fm_device_info(char* buffer, int32_t length) {
    FM_TRY
    require(buffer && length > 0, FM_ERR_INVALID_ARGUMENT, "NULL buffer");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    std::string s = std::string("finmath-lmm 0.1; device ") + std::to_string(dev) + " " + prop.name + " " +
                    prop.gcnArchName + " CUs=" + std::to_string(prop.multiProcessorCount);
    std::snprintf(buffer, (size_t)length, "%s", s.c_str());
    return FM_OK;
    FM_CATCH
}

} // extern "C"
