// parametric_engine.cpp — the C++ facade (see the header) over the C ABI.
#include "parametric_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

#include "../../include/ucuda.h"
#include "data_export.hpp"

namespace ucuda {
namespace {

// Device output buffer. On an engine without a device (compile-only) the
// buffer is a placeholder address that the C ABI never dereferences: it
// validates the request and reports UC_ERR_NO_DEVICE before any device work.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = false;
    bool alloc(size_t count, bool device) {
        n = count ? count : 1;
        if (!device) {
            p = reinterpret_cast<T*>(static_cast<uintptr_t>(256));
            return true;
        }
        owned = hipMalloc((void**)&p, n * sizeof(T)) == hipSuccess;
        return owned;
    }
    bool download(std::vector<T>& out, size_t count) const {
        out.resize(count);
        return count == 0 || hipMemcpy(out.data(), p, count * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess;
    }
    // `planes` planes of `count` values (plane stride `count`) into
    // host[k * total + begin ...]: one piece of a plane-major grid output
    bool download_piece(T* host, size_t total, size_t begin, size_t count, size_t planes) const {
        return hipMemcpy2D(host + begin, total * sizeof(T), p, count * sizeof(T), count * sizeof(T), planes,
                           hipMemcpyDeviceToHost) == hipSuccess;
    }
    ~DevBuf() { if (owned) (void)hipFree(p); }
};

std::string key_of(const System& s) {
    std::string k = s.latex ? "L|" : "P|";
    for (auto& v : s.vars) k += v + ",";
    k += "|";
    for (auto& v : s.params) k += v + ",";
    k += "|";
    for (auto& v : s.rhs) k += v + ";";
    return k + "|" + s.time_symbol;
}

uc_axis make_axis(bool over_h, bool over_var, int var_index, int param_index, int n, double lo, double hi, bool lg) {
    uc_axis a{};
    a.kind = over_h ? UC_AXIS_STEP : (over_var ? UC_AXIS_VAR : UC_AXIS_PARAM);
    a.index = over_var ? var_index : param_index;
    a.n_pts = n;
    a.lo = lo;
    a.hi = hi;
    a.log_scale = lg ? 1 : 0;
    return a;
}

bool cancelled(const std::shared_ptr<std::atomic<bool>>& c) { return c && c->load(std::memory_order_relaxed); }
void set_progress(const std::shared_ptr<std::atomic<float>>& p, float v) {
    if (p) p->store(v, std::memory_order_relaxed);
}
template <class R>
R cancelled_result() {
    R r;
    r.cancelled = true;
    r.error = "Cancelled by user";
    return r;
}
const char* kNoCpu = "CPU sweeps are not part of this engine (use_cpu); the GPU path runs the same algorithm";

// Pieces of a grid run: at least 2^18 cells (a full GPU at the map
// kernels' occupancy), at most 16.
constexpr long long kMinPiece = 1LL << 18;
long long piece_size(long long total) { return std::max(kMinPiece, (total + 15) / 16); }

// Runs `total` cells piece by piece: `run(begin, count)` launches one piece
// and downloads its outputs (false on error). The cancel token is checked
// and progress advanced between pieces. 0 done, 1 cancelled, -1 error.
template <class F>
int run_pieces(long long total, const std::shared_ptr<std::atomic<bool>>& cancel,
               const std::shared_ptr<std::atomic<float>>& progress, F run) {
    const long long piece = piece_size(total);
    for (long long b = 0; b < total; b += piece) {
        if (b > 0 && cancelled(cancel)) return 1;
        const long long count = std::min(piece, total - b);
        if (!run(b, count)) return -1;
        set_progress(progress, (float)((double)(b + count) / (double)total));
    }
    return 0;
}

int par_or_var(bool v1, bool v2) { return (v1 == v2) ? (v1 ? 0 : 1) : 2; }

}  // namespace

struct ParametricEngine::Impl {
    uc_engine* eng = nullptr;
    int device = -1;
    std::mutex mu;
    std::map<std::string, uc_system*> systems;
    std::map<std::string, std::string> rk4_text;  // system key -> codegen's RK4 step body

    ~Impl() {
        for (auto& kv : systems) uc_system_destroy(kv.second);
        if (eng) uc_engine_destroy(eng);
    }
    std::string err() const { return eng ? uc_last_error(eng) : "no engine"; }

    uc_system* rhs_system(const System& s, std::string& error) {
        const std::string k = key_of(s);
        auto it = systems.find(k);
        if (it != systems.end()) return it->second;
        std::vector<const char*> v, p, r;
        for (auto& x : s.vars) v.push_back(x.c_str());
        for (auto& x : s.params) p.push_back(x.c_str());
        for (auto& x : s.rhs) r.push_back(x.c_str());
        uc_system_desc d{v.data(), (int)v.size(), p.data(), (int)p.size(), r.data(), s.latex ? 1 : 0,
                         s.time_symbol.empty() ? nullptr : s.time_symbol.c_str()};
        uc_system* out = nullptr;
        if (uc_system_create(eng, &d, &out) != UC_OK) { error = err(); return nullptr; }
        systems[k] = out;
        return out;
    }
    uc_system* krs_system(const std::string& body, int n_vars, int n_values, std::string& error) {
        const std::string k = "KRS|" + std::to_string(n_vars) + "|" + std::to_string(n_values) + "|" + body;
        auto it = systems.find(k);
        if (it != systems.end()) return it->second;
        uc_system* out = nullptr;
        if (uc_system_create_krs(eng, body.c_str(), n_vars, n_values, &out) != UC_OK) { error = err(); return nullptr; }
        systems[k] = out;
        return out;
    }
    const std::string& rk4_of(const System& s) {
        const std::string k = key_of(s);
        auto it = rk4_text.find(k);
        if (it != rk4_text.end()) return it->second;
        std::string text;
        try {
            text = codegen_scheme(s, Scheme::RK4);
        } catch (...) {
        }
        return rk4_text[k] = text;
    }
    // The system a request runs: its right-hand side when `system` is set and
    // the step body is empty or codegen's RK4 text of it (or the analysis
    // needs the right-hand side), else the step body itself (`stepped`).
    uc_system* resolve(const System& sys, const std::string& krs, int n_vars, int n_values, bool need_rhs,
                       std::string& error, bool* stepped = nullptr) {
        if (stepped) *stepped = false;
        if (!sys.vars.empty()) {
            if (need_rhs || krs.empty() || krs == rk4_of(sys)) return rhs_system(sys, error);
        } else if (need_rhs) {
            error = "this analysis integrates the right-hand side (adaptive Dormand-Prince): set `system`";
            return nullptr;
        }
        if (krs.empty()) {
            error = "neither a step body (krs_body) nor a system (vars / params / rhs) was given";
            return nullptr;
        }
        if (stepped) *stepped = true;
        return krs_system(krs, n_vars, n_values, error);
    }
    void common(uc_common& c, const std::vector<double>& ic, const std::vector<double>& vals, double h,
                double transient, double t_max, double max_value, int precision) {
        c.initial_conditions = ic.data();
        c.base_values = vals.data();
        c.n_values = (int)vals.size();
        c.h = h;
        c.transient_time = transient;
        c.t_max = t_max;
        c.max_value = max_value;
        c.precision = precision;
        c.strict = 0;
        c.ic_seed = 0;
        c.cell_begin = 0;
        c.cell_count = 0;
    }
};

ParametricEngine::ParametricEngine(int device) : impl_(new Impl) {
    impl_->device = device;
    if (uc_engine_create(device, &impl_->eng) != UC_OK) impl_->eng = nullptr;
}
ParametricEngine::~ParametricEngine() = default;

namespace {
template <class Req>
int vars_of(const Req& req) {
    return req.amountOfX > 0 ? req.amountOfX : (int)req.system.vars.size();
}
}  // namespace

Bifurcation1DResult ParametricEngine::run_bifurcation_1d(const Bifurcation1DRequest& req) {
    Bifurcation1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (req.use_cpu) { res.error = kNoCpu; return res; }
    if (cancelled(req.cancel)) return cancelled_result<Bifurcation1DResult>();
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error);
    if (!s) return res;
    uc_bif_request r{};
    impl_->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                  req.precision);
    r.c.n_axes = 1;
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    r.writable_var = req.writable_var;
    r.pre_scaller = req.pre_scaller;
    r.max_peaks = req.max_peaks;
    const double hw = req.sweep_over_h ? std::min(req.param_lo, req.param_hi) : req.h;
    // continuation keeps every peak of the block, as the reference does
    if (req.continuation) r.max_peaks = (int)std::ceil(req.t_max / hw / req.pre_scaller) / 2 + 2;
    const size_t n = (size_t)std::max(0, req.n_pts), mp = (size_t)std::max(1, r.max_peaks);
    std::vector<int> f(n);
    std::vector<double> pk(mp * n), iv(mp * n);
    if (req.continuation) {
        DevBuf<int> flags;
        DevBuf<double> peaks, ivals;
        if (!flags.alloc(n, dev) || !peaks.alloc(mp * n, dev) || !ivals.alloc(mp * n, dev)) { res.error = "hipMalloc failed"; return res; }
        uc_bif_output o{flags.p, peaks.p, ivals.p, nullptr, 0, nullptr};
        if (uc_run_bif_cont(impl_->eng, s, &r, req.continuation_reverse ? 1 : 0, &o, nullptr) != UC_OK) {
            res.error = impl_->err();
            return res;
        }
        if (hipDeviceSynchronize() != hipSuccess || !flags.download(f, n) || !peaks.download(pk, mp * n) ||
            !ivals.download(iv, mp * n)) { res.error = "device error"; return res; }
    } else {
        const size_t piece = (size_t)std::min<long long>(piece_size((long long)n), (long long)n);
        DevBuf<int> flags;
        DevBuf<double> peaks, ivals;
        if (!flags.alloc(piece, dev) || !peaks.alloc(mp * piece, dev) || !ivals.alloc(mp * piece, dev)) {
            res.error = "hipMalloc failed";
            return res;
        }
        const int st = run_pieces((long long)n, req.cancel, req.progress, [&](long long b, long long cnt) {
            r.c.cell_begin = b;
            r.c.cell_count = cnt;
            uc_bif_output o{flags.p, peaks.p, ivals.p, nullptr, 0, nullptr};
            if (uc_run_bif(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return false; }
            if (hipDeviceSynchronize() != hipSuccess || !flags.download_piece(f.data(), n, b, cnt, 1) ||
                !peaks.download_piece(pk.data(), n, b, cnt, mp) || !ivals.download_piece(iv.data(), n, b, cnt, mp)) {
                res.error = "device error";
                return false;
            }
            return true;
        });
        if (st == 1) return cancelled_result<Bifurcation1DResult>();
        if (st < 0) return res;
    }
    res.param_lo = req.param_lo;
    res.param_hi = req.param_hi;
    res.n_pts = req.n_pts;
    res.record_steps = (int)std::ceil(req.t_max / hw / req.pre_scaller);
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    res.flags = f;
    res.bifurcation_points.assign(n, {});
    res.peak_times.assign(n, {});
    for (size_t c = 0; c < n; ++c) {
        // continuation over h: the chain reports peak times in units of the
        // base step; the reference rescales them to the point's own step
        // (run_bif1d_continuation, U-CUDA/parametric_engine.cpp:4718-4727)
        double scale = 1.0;
        if (req.continuation && req.sweep_over_h) {
            const double t = (double)c / (double)(req.n_pts > 1 ? req.n_pts - 1 : 1);
            double hp;
            if (req.log_scale) {
                const double l0 = std::log10(req.param_lo), l1 = std::log10(req.param_hi);
                hp = std::pow(10.0, req.continuation_reverse ? (l1 - (l1 - l0) * t) : (l0 + (l1 - l0) * t));
            } else {
                hp = req.continuation_reverse ? (req.param_hi - (req.param_hi - req.param_lo) * t)
                                              : (req.param_lo + (req.param_hi - req.param_lo) * t);
            }
            scale = hp / req.h;
        }
        for (int k = 0; k < f[c]; ++k) {
            res.bifurcation_points[c].push_back(pk[(size_t)k * n + c]);
            double v = iv[(size_t)k * n + c];
            if (scale != 1.0) v *= scale;
            res.peak_times[c].push_back(v);
        }
    }
    auto& sn = res.snapshot;
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.tMax = req.t_max;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.preScaller = req.pre_scaller;
    sn.writableVar = req.writable_var;
    sn.indexOfMutVar = req.sweep_over_var ? req.var_sweep_index : req.param_index;
    sn.range_lo = req.param_lo;
    sn.range_hi = req.param_hi;
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::bif1d_config(cfg, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                          req.pre_scaller, req.writable_var, sn.indexOfMutVar, req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t c = 0; c < n; ++c)
            csv::bif1d_rows(out, csv::param_value_at(c, req.n_pts, req.param_lo, req.param_hi), f[c],
                            res.bifurcation_points[c].data(), res.peak_times[c].data());
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

namespace {
template <class I, class Req>
void lyap_common(I* impl, const Req& req, uc_lyap_request& r, int n_axes, bool stepped) {
    impl->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                 req.precision);
    r.c.n_axes = n_axes;
    r.NT = req.NT;
    r.eps = req.eps;
    // a stepped body has no Jacobian: the reference's finite-difference method
    r.method = stepped ? UC_METHOD_FINITE_DIFF : req.method;
    r.probe_seed = 1234567891ULL;
}

template <class Req, class Snap>
void lle1d_snapshot(const Req& req, Snap& sn) {
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.tMax = req.t_max;
    sn.NT = req.NT;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.eps = req.eps;
    sn.indexOfMutVar = req.sweep_over_var ? req.var_sweep_index : req.param_index;
    sn.range_lo = req.param_lo;
    sn.range_hi = req.param_hi;
}

// Runs an LLE / LS map (1-D or 2-D, `n` cells, K exponents) into host
// vectors: piece by piece (cancel + progress), or as one chain.
template <class Req>
int run_lyap_map(uc_engine* eng, bool dev, uc_system* s, uc_lyap_request& r, const Req& req, bool spectrum,
                 bool chain, bool reverse, size_t n, size_t K, std::vector<double>& vals, std::vector<int>& flags,
                 std::string& error) {
    vals.assign(n * K, 0.0);
    flags.assign(n, 0);
    if (chain) {
        DevBuf<double> v;
        DevBuf<int> f;
        if (!v.alloc(n * K, dev) || !f.alloc(n, dev)) { error = "hipMalloc failed"; return -1; }
        uc_lyap_output o{v.p, f.p, nullptr};
        const int rc = spectrum ? uc_run_ls_cont(eng, s, &r, reverse ? 1 : 0, &o, nullptr)
                                : uc_run_lle_cont(eng, s, &r, reverse ? 1 : 0, &o, nullptr);
        if (rc != UC_OK) { error = uc_last_error(eng); return -1; }
        if (hipDeviceSynchronize() != hipSuccess || !v.download(vals, n * K) || !f.download(flags, n)) {
            error = "device error";
            return -1;
        }
        return 0;
    }
    const size_t piece = (size_t)std::min<long long>(piece_size((long long)n), (long long)n);
    DevBuf<double> v;
    DevBuf<int> f;
    if (!v.alloc(piece * K, dev) || !f.alloc(piece, dev)) { error = "hipMalloc failed"; return -1; }
    return run_pieces((long long)n, req.cancel, req.progress, [&](long long b, long long cnt) {
        r.c.cell_begin = b;
        r.c.cell_count = cnt;
        uc_lyap_output o{v.p, f.p, nullptr};
        const int rc = spectrum ? uc_run_ls(eng, s, &r, &o, nullptr) : uc_run_lle(eng, s, &r, &o, nullptr);
        if (rc != UC_OK) { error = uc_last_error(eng); return false; }
        if (hipDeviceSynchronize() != hipSuccess || !v.download_piece(vals.data(), n, b, cnt, K) ||
            !f.download_piece(flags.data(), n, b, cnt, 1)) {
            error = "device error";
            return false;
        }
        return true;
    });
}
}  // namespace

LLE1DResult ParametricEngine::run_lle_1d(const LLE1DRequest& req) {
    LLE1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (req.use_cpu) { res.error = kNoCpu; return res; }
    if (cancelled(req.cancel)) return cancelled_result<LLE1DResult>();
    bool stepped = false;
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error, &stepped);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 1, stepped);
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    const size_t n = (size_t)std::max(0, req.n_pts);
    std::vector<double> v;
    const int st = run_lyap_map(impl_->eng, dev, s, r, req, false, req.continuation, req.continuation_reverse, n, 1, v, res.flags,
                                 res.error);
    if (st == 1) return cancelled_result<LLE1DResult>();
    if (st < 0) return res;
    res.n_pts = req.n_pts;
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    res.param_lo = req.param_lo;
    res.param_hi = req.param_hi;
    res.lyapunov.resize(n);
    for (size_t i = 0; i < n; ++i)
        res.lyapunov[i] = (res.flags[i] == 1) ? v[i] : std::numeric_limits<double>::quiet_NaN();
    lle1d_snapshot(req, res.snapshot);
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::curve1d_config(cfg, "1D LLE", req.base_values, req.initial_conditions, req.t_max, req.NT,
                            req.transient_time, req.h, req.eps, req.param_index, req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t i = 0; i < n; ++i)
            csv::lle1d_row(out, csv::param_value_at(i, req.n_pts, req.param_lo, req.param_hi), res.lyapunov[i]);
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

LS1DResult ParametricEngine::run_ls_1d(const LS1DRequest& req) {
    LS1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (req.use_cpu) { res.error = kNoCpu; return res; }
    if (cancelled(req.cancel)) return cancelled_result<LS1DResult>();
    bool stepped = false;
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error, &stepped);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 1, stepped);
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    const size_t n = (size_t)std::max(0, req.n_pts), N = (size_t)uc_system_n_vars(s);
    std::vector<double> v;
    const int st = run_lyap_map(impl_->eng, dev, s, r, req, true, req.continuation, req.continuation_reverse, n, N, v, res.flags,
                                 res.error);
    if (st == 1) return cancelled_result<LS1DResult>();
    if (st < 0) return res;
    res.n_pts = req.n_pts;
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    res.n_exponents = (int)N;
    res.param_lo = req.param_lo;
    res.param_hi = req.param_hi;
    res.spectrum.assign(n, std::vector<double>(N));
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < N; ++k) res.spectrum[i][k] = v[k * n + i];
    lle1d_snapshot(req, res.snapshot);
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::curve1d_config(cfg, "1D LS", req.base_values, req.initial_conditions, req.t_max, req.NT,
                            req.transient_time, req.h, req.eps, req.param_index, req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t i = 0; i < n; ++i)
            csv::ls1d_row(out, csv::param_value_at(i, req.n_pts, req.param_lo, req.param_hi), res.spectrum[i].data(),
                          (int)N);
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

namespace {
template <class Req>
void axes_2d(const Req& req, uc_common& c) {
    c.n_axes = 2;
    c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                          req.param_lo, req.param_hi, req.log_scale);
    c.axis[1] = make_axis(req.sweep_over_h_2, req.sweep_over_var_2, req.var_sweep_index_2, req.param_index_2,
                          req.n_pts, req.param_lo_2, req.param_hi_2, req.log_scale_2);
}
void min_max(const double* v, const int* f, size_t n, double& lo, double& hi) {
    bool any = false;
    lo = hi = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (f[i] != 1 || !std::isfinite(v[i]) || v[i] == 999.0 || v[i] == -999.0) continue;
        if (!any) { lo = hi = v[i]; any = true; }
        lo = std::min(lo, v[i]);
        hi = std::max(hi, v[i]);
    }
}
template <class Req, class Snap>
void map2d_snapshot(const Req& req, Snap& sn) {
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.par_or_var = par_or_var(req.sweep_over_var, req.sweep_over_var_2);
    sn.tMax = req.t_max;
    sn.NT = req.NT;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.eps = req.eps;
    sn.indexOfMutVar = req.sweep_over_var ? req.var_sweep_index : req.param_index;
    sn.indexOfMutVar2 = req.sweep_over_var_2 ? req.var_sweep_index_2 : req.param_index_2;
    sn.range1_lo = req.param_lo;
    sn.range1_hi = req.param_hi;
    sn.range2_lo = req.param_lo_2;
    sn.range2_hi = req.param_hi_2;
    sn.n_pts = req.n_pts;
}
}  // namespace

Dft1DResult ParametricEngine::run_dft_1d(const Dft1DRequest& req) {
    Dft1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (req.use_cpu) { res.error = kNoCpu; return res; }
    if (cancelled(req.cancel)) return cancelled_result<Dft1DResult>();
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error);
    if (!s) return res;
    uc_dft_request r{};
    impl_->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                  req.precision);
    r.c.n_axes = 1;
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    r.writable_var = req.writable_var;
    r.pre_scaller = req.pre_scaller;
    r.n_freq = req.n_freq;
    r.freq_lo = req.freq_lo;
    r.freq_hi = req.freq_hi;
    r.freq_log_scale = req.freq_log_scale ? 1 : 0;
    r.window_type = req.window_type;
    const size_t n = (size_t)std::max(0, req.n_pts), nf = (size_t)std::max(1, req.n_freq);
    DevBuf<double> ak, bk;
    DevBuf<int> flags;
    if (!ak.alloc(n * nf, dev) || !bk.alloc(n * nf, dev) || !flags.alloc(n, dev)) { res.error = "hipMalloc failed"; return res; }
    uc_dft_output o{ak.p, bk.p, flags.p};
    const int rc = req.continuation
                       ? uc_run_dft_cont(impl_->eng, s, &r, req.continuation_reverse ? 1 : 0, &o, nullptr)
                       : uc_run_dft(impl_->eng, s, &r, &o, nullptr);
    if (rc != UC_OK) { res.error = impl_->err(); return res; }
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    if (!ak.download(res.ak_cos, n * nf) || !bk.download(res.bk_sin, n * nf) || !flags.download(res.flags, n)) {
        res.error = "device error";
        return res;
    }
    res.n_pts = req.n_pts;
    res.n_freq = req.n_freq;
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.freq_lo = req.freq_lo; res.freq_hi = req.freq_hi;
    auto& sn = res.snapshot;
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.tMax = req.t_max;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.preScaller = req.pre_scaller;
    sn.writableVar = req.writable_var;
    sn.indexOfMutVar = req.sweep_over_h ? -1 : req.sweep_over_var ? req.var_sweep_index : req.param_index;
    sn.range_lo = req.param_lo;
    sn.range_hi = req.param_hi;
    sn.n_freq = req.n_freq;
    sn.freq_lo = req.freq_lo;
    sn.freq_hi = req.freq_hi;
    sn.window_type = req.window_type;
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::dft1d_config(cfg, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                          req.pre_scaller, req.writable_var, sn.indexOfMutVar, req.param_lo, req.param_hi, req.n_freq,
                          req.freq_lo, req.freq_hi, req.window_type);
        std::ofstream a(req.csv_output_path + "_AkCOS.csv"), b(req.csv_output_path + "_BkSIN.csv");
        csv::dft1d_header(a, req.param_lo, req.param_hi, req.freq_lo, req.freq_hi);
        csv::dft1d_header(b, req.param_lo, req.param_hi, req.freq_lo, req.freq_hi);
        csv::dft1d_matrix(a, res.ak_cos.data(), req.n_pts, req.n_freq);
        csv::dft1d_matrix(b, res.bk_sin.data(), req.n_pts, req.n_freq);
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

FastSyncResult ParametricEngine::run_fastsync(const FastSyncRequest& req) {
    FastSyncResult res;
    res.mode = req.mode;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (cancelled(req.cancel)) { res = cancelled_result<FastSyncResult>(); res.mode = req.mode; return res; }
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.values.size(), false, res.error);
    if (!s) return res;
    const size_t n = (size_t)uc_system_n_vars(s);
    if (req.k_forward.size() != n || req.k_backward.size() != n || req.ic_master.size() != n || req.ic_slave.size() != n) {
        res.error = "k_forward / k_backward / ic_master / ic_slave must hold amountOfX values";
        return res;
    }
    uc_fastsync_request r{};
    r.values = req.values.data();
    r.n_values = (int)req.values.size();
    r.mode = req.mode;
    r.h = req.h;
    r.iter_of_synchr = req.iter_of_synchr;
    r.pre_scaller = req.pre_scaller;
    r.k_forward = req.k_forward.data();
    r.k_backward = req.k_backward.data();
    r.ic_master = req.ic_master.data();
    r.ic_slave = req.ic_slave.data();
    r.t_max = req.t_max;
    r.transient_time = req.transient_time;
    r.window = req.window;
    r.axis_x_var = req.axis_x_var;
    r.axis_y_var = req.axis_y_var;
    r.axis_x_lo = req.axis_x_lo; r.axis_x_hi = req.axis_x_hi;
    r.axis_y_lo = req.axis_y_lo; r.axis_y_hi = req.axis_y_hi;
    r.n_pts = req.n_pts;
    r.grid_swap_master_slave = req.grid_swap_master_slave ? 1 : 0;
    r.type_of_synch = req.type_of_synch;
    r.error_estim = req.error_estim;
    r.fs_error_trs = req.fs_error_trs;
    const long long count = uc_fastsync_points(&r);
    if (count < 0) {
        // the C ABI names the first violated rule
        (void)uc_run_fastsync(impl_->eng, s, &r, nullptr, nullptr);
        res.error = impl_->err();
        if (res.error.empty()) res.error = "invalid fast synchro request";
        return res;
    }
    DevBuf<double> err, traj;
    if (!err.alloc((size_t)count, dev) || (req.mode == 0 && !traj.alloc((size_t)count * n, dev))) {
        res.error = "hipMalloc failed";
        return res;
    }
    uc_fastsync_output o{err.p, req.mode == 0 ? traj.p : nullptr};
    if (uc_run_fastsync(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<double> values;
    if (!err.download(values, (size_t)count)) { res.error = "device error"; return res; }
    double vmin = std::numeric_limits<double>::infinity(), vmax = -vmin;
    for (double v : values) {
        if (!std::isfinite(v)) continue;
        vmin = std::min(vmin, v);
        vmax = std::max(vmax, v);
    }
    res.min_val = std::isfinite(vmin) ? vmin : 0.0;
    res.max_val = std::isfinite(vmax) ? vmax : 0.0;
    if (req.mode == 0) {
        if (!traj.download(res.traj_full, (size_t)count * n)) { res.error = "device error"; return res; }
        res.sync_error = std::move(values);
        res.n_pts_traj = (int)count;
        res.amountOfX_traj = (int)n;
    } else {
        res.heatmap = std::move(values);
        res.n_pts_grid = req.n_pts;
        res.axis_x_lo = req.axis_x_lo; res.axis_x_hi = req.axis_x_hi;
        res.axis_y_lo = req.axis_y_lo; res.axis_y_hi = req.axis_y_hi;
        res.axis_x_var = req.axis_x_var; res.axis_y_var = req.axis_y_var;
    }
    auto& sn = res.snapshot;
    sn.mode = req.mode;
    sn.values = req.values;
    sn.ic_master = req.ic_master;
    sn.ic_slave = req.ic_slave;
    sn.k_forward = req.k_forward;
    sn.k_backward = req.k_backward;
    sn.var_names = req.var_names;
    sn.h = req.h;
    sn.iter_of_synchr = req.iter_of_synchr;
    sn.preScaller = req.pre_scaller;
    sn.window = req.window;
    sn.type_of_synch = req.type_of_synch;
    sn.error_estim = req.error_estim;
    sn.fs_error_trs = req.fs_error_trs;
    sn.tMax = req.t_max;
    sn.transientTime = req.transient_time;
    sn.axis_x_var = req.axis_x_var; sn.axis_y_var = req.axis_y_var;
    sn.axis_x_lo = req.axis_x_lo; sn.axis_x_hi = req.axis_x_hi;
    sn.axis_y_lo = req.axis_y_lo; sn.axis_y_hi = req.axis_y_hi;
    sn.n_pts = req.n_pts;
    sn.grid_swap_master_slave = req.grid_swap_master_slave;
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::fastsync_config(cfg, req.mode, req.values, req.ic_master, req.ic_slave, req.k_forward, req.k_backward,
                             req.h, req.iter_of_synchr, req.pre_scaller, req.window, req.type_of_synch,
                             req.error_estim, req.fs_error_trs, req.t_max, req.transient_time, req.axis_x_var,
                             req.axis_y_var, req.axis_x_lo, req.axis_x_hi, req.axis_y_lo, req.axis_y_hi, req.n_pts,
                             req.grid_swap_master_slave);
        std::ofstream data(req.csv_output_path);
        if (req.mode == 0)
            csv::fastsync_attractor(data, res.traj_full.data(), res.sync_error.data(), res.n_pts_traj, (int)n,
                                    req.var_names);
        else
            csv::fastsync_grid(data, res.heatmap.data(), req.n_pts, req.axis_x_lo, req.axis_x_hi, req.axis_y_lo,
                               req.axis_y_hi);
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

Bifurcation2DResult ParametricEngine::run_bifurcation_2d(const Bifurcation2DRequest& req) {
    Bifurcation2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (cancelled(req.cancel)) return cancelled_result<Bifurcation2DResult>();
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error);
    if (!s) return res;
    uc_bif_request r{};
    impl_->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                  req.precision);
    axes_2d(req, r.c);
    r.writable_var = req.writable_var;
    r.pre_scaller = req.pre_scaller;
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    DevBuf<int> periods;
    if (!periods.alloc(n, dev)) { res.error = "hipMalloc failed"; return res; }
    if (uc_run_bif_periods(impl_->eng, s, &r, req.eps_dbscan, periods.p, nullptr) != UC_OK) {
        res.error = impl_->err();
        return res;
    }
    std::vector<int> per;
    if (!periods.download(per, n)) { res.error = "device error"; return res; }
    res.values.resize(n);
    res.flags.resize(n);
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    for (size_t i = 0; i < n; ++i) {
        res.values[i] = per[i] < 0 ? -1.0 : (double)per[i];
        res.flags[i] = per[i] < 0 ? -1 : 1;
        // colour range over flagged-ok, non-negative cells (run_bif2d, parametric_engine.cpp:5897-5906)
        if (res.flags[i] >= 0 && res.values[i] >= 0.0) { lo = std::min(lo, res.values[i]); hi = std::max(hi, res.values[i]); }
    }
    res.min_val = std::isfinite(lo) ? lo : 0.0;
    res.max_val = std::isfinite(hi) ? hi : 0.0;
    res.n_pts = req.n_pts;
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.param_lo_2 = req.param_lo_2; res.param_hi_2 = req.param_hi_2;
    auto& sn = res.snapshot;
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.par_or_var = par_or_var(req.sweep_over_var, req.sweep_over_var_2);
    sn.tMax = req.t_max;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.preScaller = req.pre_scaller;
    sn.eps_dbscan = req.eps_dbscan;
    sn.writableVar = req.writable_var;
    sn.indexOfMutVar = req.sweep_over_var ? req.var_sweep_index : req.param_index;
    sn.indexOfMutVar2 = req.sweep_over_var_2 ? req.var_sweep_index_2 : req.param_index_2;
    sn.range1_lo = req.param_lo; sn.range1_hi = req.param_hi;
    sn.range2_lo = req.param_lo_2; sn.range2_hi = req.param_hi_2;
    sn.n_pts = req.n_pts;
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::bif2d_config(cfg, sn.par_or_var, req.base_values, req.initial_conditions, req.t_max, req.transient_time,
                          req.h, req.pre_scaller, req.eps_dbscan, req.writable_var, sn.indexOfMutVar,
                          sn.indexOfMutVar2, req.param_lo, req.param_hi, req.param_lo_2, req.param_hi_2, req.n_pts);
        std::ofstream out(req.csv_output_path);
        csv::grid(out, req.n_pts, res.values.data());
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

LLE2DResult ParametricEngine::run_lle_2d(const LLE2DRequest& req) {
    LLE2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (cancelled(req.cancel)) return cancelled_result<LLE2DResult>();
    bool stepped = false;
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error, &stepped);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 2, stepped);
    axes_2d(req, r.c);
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    const int st = run_lyap_map(impl_->eng, dev, s, r, req, false, false, false, n, 1, res.values, res.flags, res.error);
    if (st == 1) return cancelled_result<LLE2DResult>();
    if (st < 0) return res;
    for (size_t i = 0; i < n; ++i)
        if (res.flags[i] != 1) res.values[i] = std::numeric_limits<double>::quiet_NaN();
    min_max(res.values.data(), res.flags.data(), n, res.min_val, res.max_val);
    res.n_pts = req.n_pts;
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.param_lo_2 = req.param_lo_2; res.param_hi_2 = req.param_hi_2;
    map2d_snapshot(req, res.snapshot);
    if (!req.csv_output_path.empty()) {
        const auto& sn = res.snapshot;
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::map2d_config(cfg, "2D LLE", sn.par_or_var, req.base_values, req.initial_conditions, req.t_max, req.NT,
                          req.transient_time, req.h, req.eps, sn.indexOfMutVar, sn.indexOfMutVar2, req.param_lo,
                          req.param_hi, req.param_lo_2, req.param_hi_2, req.n_pts, 0);
        std::ofstream out(req.csv_output_path);
        csv::grid(out, req.n_pts, res.values.data());
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

LS2DResult ParametricEngine::run_ls_2d(const LS2DRequest& req) {
    LS2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (cancelled(req.cancel)) return cancelled_result<LS2DResult>();
    bool stepped = false;
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), false,
                                  res.error, &stepped);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 2, stepped);
    axes_2d(req, r.c);
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts, N = (size_t)uc_system_n_vars(s);
    const int st = run_lyap_map(impl_->eng, dev, s, r, req, true, false, false, n, N, res.values, res.flags, res.error);
    if (st == 1) return cancelled_result<LS2DResult>();
    if (st < 0) return res;
    res.n_pts = req.n_pts;
    res.n_exponents = (int)N;
    res.min_val.resize(N);
    res.max_val.resize(N);
    for (size_t k = 0; k < N; ++k) min_max(res.values.data() + k * n, res.flags.data(), n, res.min_val[k], res.max_val[k]);
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.param_lo_2 = req.param_lo_2; res.param_hi_2 = req.param_hi_2;
    map2d_snapshot(req, res.snapshot);
    res.snapshot.n_exponents = (int)N;
    if (!req.csv_output_path.empty()) {
        const auto& sn = res.snapshot;
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::map2d_config(cfg, "2D LS", sn.par_or_var, req.base_values, req.initial_conditions, req.t_max, req.NT,
                          req.transient_time, req.h, req.eps, sn.indexOfMutVar, sn.indexOfMutVar2, req.param_lo,
                          req.param_hi, req.param_lo_2, req.param_hi_2, req.n_pts, (int)N);
        std::ofstream out(req.csv_output_path);
        csv::ls2d_cells(out, req.n_pts, (int)N, res.values.data());
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

namespace {
void feature_range(const std::vector<double>& f, const std::vector<int>& ok, double& lo, double& hi) {
    lo = hi = 0.0;
    bool any = false;
    for (size_t i = 0; i < f.size(); ++i) {
        if (!ok[i] || !std::isfinite(f[i]) || f[i] == 999.0) continue;
        if (!any) { lo = hi = f[i]; any = true; }
        lo = std::min(lo, f[i]);
        hi = std::max(hi, f[i]);
    }
}

// Clusters cells by their two features: the features are quantised with
// step eps, grid boxes that touch (including diagonally) are merged, and
// clusters are numbered 1..K by their smallest box (feature 1, then 2).
// Cells with ok == 0 or non-finite features get 0.
int cluster_features(const std::vector<double>& f1, const std::vector<double>& f2, const std::vector<int>& ok,
                     double eps, std::vector<int>& label) {
    const size_t n = f1.size();
    label.assign(n, 0);
    std::map<std::pair<long long, long long>, int> box;  // box -> node
    std::vector<std::pair<long long, long long>> key(n);
    std::vector<int> node_of(n, -1);
    std::vector<std::pair<long long, long long>> boxes;
    for (size_t i = 0; i < n; ++i) {
        if (!ok[i] || !std::isfinite(f1[i]) || !std::isfinite(f2[i]) || f1[i] == 999.0) continue;
        key[i] = {(long long)std::floor(f1[i] / eps + 0.5), (long long)std::floor(f2[i] / eps + 0.5)};
        auto it = box.find(key[i]);
        if (it == box.end()) {
            it = box.emplace(key[i], (int)boxes.size()).first;
            boxes.push_back(key[i]);
        }
        node_of[i] = it->second;
    }
    std::vector<int> parent(boxes.size());
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int(int)> find = [&](int a) { return parent[a] == a ? a : parent[a] = find(parent[a]); };
    for (size_t b = 0; b < boxes.size(); ++b)
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) {
                auto it = box.find({boxes[b].first + dx, boxes[b].second + dy});
                if (it != box.end()) {
                    const int ra = find((int)b), rb = find(it->second);
                    if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
                }
            }
    // canonical order: a cluster's smallest box
    std::map<int, std::pair<long long, long long>> least;
    for (size_t b = 0; b < boxes.size(); ++b) {
        const int r = find((int)b);
        auto it = least.find(r);
        if (it == least.end() || boxes[b] < it->second) least[r] = boxes[b];
    }
    std::vector<std::pair<std::pair<long long, long long>, int>> order;
    for (auto& kv : least) order.push_back({kv.second, kv.first});
    std::sort(order.begin(), order.end());
    std::map<int, int> id;
    for (size_t k = 0; k < order.size(); ++k) id[order[k].second] = (int)k + 1;
    for (size_t i = 0; i < n; ++i)
        if (node_of[i] >= 0) label[i] = id[find(node_of[i])];
    return (int)order.size();
}
}  // namespace

BasinsResult ParametricEngine::run_basins(const BasinsRequest& req) {
    BasinsResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    const bool dev = impl_->device >= 0;
    if (cancelled(req.cancel)) return cancelled_result<BasinsResult>();
    uc_system* s = impl_->resolve(req.system, req.krs_body, vars_of(req), (int)req.base_values.size(), true,
                                  res.error);
    if (!s) return res;
    if (!(req.section_period > 0.0)) { res.error = "section_period must be > 0"; return res; }
    if (!(req.eps_dbscan > 0.0)) { res.error = "eps_dbscan must be > 0"; return res; }
    int n_sections = req.n_sections;
    if (n_sections <= 0)
        n_sections = std::max(17, (int)std::ceil((req.transient_time + req.t_max) / req.section_period));
    if (req.progress_phase) req.progress_phase->store(1, std::memory_order_relaxed);
    set_progress(req.progress, 0.0f);
    uc_basins_request r{};
    impl_->common(r.c, req.initial_conditions, req.base_values, req.h, 0.0, 1.0, req.max_value, req.precision);
    r.c.n_axes = 2;
    r.c.axis[0] = make_axis(false, true, req.axis_x_var, 0, req.n_pts, req.axis_x_lo, req.axis_x_hi, false);
    r.c.axis[1] = make_axis(false, true, req.axis_y_var, 0, req.n_pts, req.axis_y_lo, req.axis_y_hi, false);
    r.rtol = req.rtol;
    r.atol = req.atol;
    r.h_min = 1e-12;
    r.h_max = req.section_period / 4.0;
    r.section_period = req.section_period;
    r.n_sections = n_sections;
    r.n_record = 16;
    r.max_period = 8;
    r.label_vars = 0;
    r.period_tol = 1e-4;
    r.quantum = req.eps_dbscan;
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    const int lbl = std::min(2, uc_system_n_vars(s));
    DevBuf<int> labels, periods;
    DevBuf<double> points;
    if (!labels.alloc(n, dev) || !periods.alloc(n, dev) || !points.alloc(n * (size_t)lbl, dev)) { res.error = "hipMalloc failed"; return res; }
    int n_att = 0;
    uc_basins_output o{labels.p, periods.p, points.p, nullptr, 0, &n_att, nullptr, nullptr};
    if (uc_run_basins(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<double> pts;
    if (!labels.download(res.basin_idx, n) || !periods.download(res.periods, n) || !points.download(pts, n * (size_t)lbl)) {
        res.error = "device error";
        return res;
    }
    if (req.progress_phase) req.progress_phase->store(2, std::memory_order_relaxed);
    res.helpful_array.resize(n);
    res.avg_peaks.assign(n, 999.0);
    res.avg_intervals.assign(n, 999.0);
    res.cycle_x.assign(n, std::nan(""));
    for (size_t i = 0; i < n; ++i) {
        const bool settled = res.periods[i] >= 0;
        res.helpful_array[i] = settled ? 1 : 0;
        if (!settled) continue;
        res.avg_peaks[i] = pts[i];
        res.avg_intervals[i] = lbl > 1 ? pts[n + i] : 0.0;
        if (res.periods[i] > 0) res.cycle_x[i] = pts[i];
    }
    feature_range(res.avg_peaks, res.helpful_array, res.avg_peaks_min, res.avg_peaks_max);
    feature_range(res.avg_intervals, res.helpful_array, res.avg_intervals_min, res.avg_intervals_max);
    res.n_clusters = n_att;
    res.min_cluster_idx = 0;
    res.n_pts = req.n_pts;
    res.axis_x_lo = req.axis_x_lo; res.axis_x_hi = req.axis_x_hi;
    res.axis_y_lo = req.axis_y_lo; res.axis_y_hi = req.axis_y_hi;
    res.axis_x_var = req.axis_x_var; res.axis_y_var = req.axis_y_var;
    auto& sn = res.snapshot;
    sn.values = req.base_values;
    sn.initial_conditions = req.initial_conditions;
    sn.tMax = req.t_max;
    sn.transientTime = req.transient_time;
    sn.h = req.h;
    sn.preScaller = req.pre_scaller;
    sn.eps_dbscan = req.eps_dbscan;
    sn.writableVar = req.writable_var;
    sn.axis_x_var = req.axis_x_var; sn.axis_y_var = req.axis_y_var;
    sn.axis_x_lo = req.axis_x_lo; sn.axis_x_hi = req.axis_x_hi;
    sn.axis_y_lo = req.axis_y_lo; sn.axis_y_hi = req.axis_y_hi;
    sn.n_pts = req.n_pts;
    sn.feature1 = req.feature1; sn.feature2 = req.feature2;
    sn.mult1 = req.mult1; sn.mult2 = req.mult2;
    if (!req.csv_output_path.empty()) {
        // the reference's five files (U-CUDA/data_export.cpp:472-544): the
        // label grid, _1 / _2 the two features (here the canonical cycle
        // point, 999 where unbound), _3 the settled / unbound array
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::basins_config(cfg, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                           req.pre_scaller, req.eps_dbscan, req.writable_var, req.axis_x_var, req.axis_y_var,
                           req.axis_x_lo, req.axis_x_hi, req.axis_y_lo, req.axis_y_hi);
        auto with_ranges = [&](const std::string& path) {
            auto f = std::make_unique<std::ofstream>(path);
            csv::basins_ranges(*f, req.axis_x_lo, req.axis_x_hi, req.axis_y_lo, req.axis_y_hi);
            return f;
        };
        csv::grid_int(*with_ranges(req.csv_output_path), req.n_pts, res.basin_idx.data());
        csv::grid_finite(*with_ranges(req.csv_output_path + "_1.csv"), req.n_pts, res.avg_peaks.data());
        csv::grid_finite(*with_ranges(req.csv_output_path + "_2.csv"), req.n_pts, res.avg_intervals.data());
        csv::grid_int(*with_ranges(req.csv_output_path + "_3.csv"), req.n_pts, res.helpful_array.data());
    }
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

BasinsReclusterResult ParametricEngine::run_basins_recluster(const BasinsReclusterRequest& req) {
    BasinsReclusterResult res;
    if (cancelled(req.cancel)) return cancelled_result<BasinsReclusterResult>();
    const size_t n = (size_t)std::max(0, req.n_pts) * (size_t)std::max(0, req.n_pts);
    if (req.avg_peaks.size() != n || req.avg_intervals.size() != n || req.helpful_array.size() != n) {
        res.error = "avg_peaks, avg_intervals and helpful_array must hold n_pts * n_pts values";
        return res;
    }
    if (!(req.eps_dbscan > 0.0)) { res.error = "eps_dbscan must be > 0"; return res; }
    std::vector<int> ok(n);
    for (size_t i = 0; i < n; ++i) ok[i] = req.helpful_array[i] == 1;
    res.n_clusters = cluster_features(req.avg_peaks, req.avg_intervals, ok, req.eps_dbscan, res.basin_idx);
    res.min_cluster_idx = 0;
    res.n_pts = req.n_pts;
    set_progress(req.progress, 1.0f);
    res.ok = true;
    return res;
}

}  // namespace ucuda
