// parametric_engine.cpp — the C++ facade (see the header) over the C ABI.
#include "parametric_engine.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <map>
#include <mutex>

#include "../../include/ucuda.h"
#include "data_export.hpp"

#include <fstream>

namespace ucuda {
namespace {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool alloc(size_t count) {
        n = count ? count : 1;
        return hipMalloc((void**)&p, n * sizeof(T)) == hipSuccess;
    }
    bool download(std::vector<T>& out, size_t count) const {
        out.resize(count);
        return count == 0 || hipMemcpy(out.data(), p, count * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
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
void done(const std::shared_ptr<std::atomic<float>>& p) { if (p) p->store(1.0f, std::memory_order_relaxed); }

}  // namespace

struct ParametricEngine::Impl {
    uc_engine* eng = nullptr;
    std::mutex mu;
    std::map<std::string, uc_system*> systems;

    ~Impl() {
        for (auto& kv : systems) uc_system_destroy(kv.second);
        if (eng) uc_engine_destroy(eng);
    }
    std::string err() const { return eng ? uc_last_error(eng) : "no engine"; }
    uc_system* system(const System& s, std::string& error) {
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
    }
};

ParametricEngine::ParametricEngine(int device) : impl_(new Impl) {
    if (uc_engine_create(device, &impl_->eng) != UC_OK) impl_->eng = nullptr;
}
ParametricEngine::~ParametricEngine() = default;

Bifurcation1DResult ParametricEngine::run_bifurcation_1d(const Bifurcation1DRequest& req) {
    Bifurcation1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (req.use_cpu) { res.error = "CPU sweeps are not part of this engine (use_cpu); the GPU path runs the same algorithm"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
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
    const size_t n = (size_t)req.n_pts, mp = (size_t)std::max(1, r.max_peaks);
    DevBuf<int> flags;
    DevBuf<double> peaks, ivals;
    if (!flags.alloc(n) || !peaks.alloc(mp * n) || !ivals.alloc(mp * n)) { res.error = "hipMalloc failed"; return res; }
    uc_bif_output o{flags.p, peaks.p, ivals.p, nullptr, 0, nullptr};
    const int rc = req.continuation
                       ? uc_run_bif_cont(impl_->eng, s, &r, req.continuation_reverse ? 1 : 0, &o, nullptr)
                       : uc_run_bif(impl_->eng, s, &r, &o, nullptr);
    if (rc != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<int> f;
    std::vector<double> pk, iv;
    if (hipDeviceSynchronize() != hipSuccess || !flags.download(f, n) || !peaks.download(pk, mp * n) ||
        !ivals.download(iv, mp * n)) { res.error = "device error"; return res; }
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
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::bif1d_config(cfg, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                          req.pre_scaller, req.writable_var, req.sweep_over_var ? req.var_sweep_index : req.param_index,
                          req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t c = 0; c < n; ++c)
            csv::bif1d_rows(out, csv::param_value_at(c, req.n_pts, req.param_lo, req.param_hi), f[c],
                            res.bifurcation_points[c].data(), res.peak_times[c].data());
    }
    done(req.progress);
    res.ok = true;
    return res;
}

namespace {
template <class I, class Req>
int lyap_common(I* impl, const Req& req, uc_lyap_request& r, int n_axes) {
    impl->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                 req.precision);
    r.c.n_axes = n_axes;
    r.NT = req.NT;
    r.eps = req.eps;
    r.method = req.method;
    r.probe_seed = 1234567891ULL;
    return 0;
}
}  // namespace

LLE1DResult ParametricEngine::run_lle_1d(const LLE1DRequest& req) {
    LLE1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (req.use_cpu) { res.error = "CPU sweeps are not part of this engine (use_cpu); the GPU path runs the same algorithm"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 1);
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    const size_t n = (size_t)req.n_pts;
    DevBuf<double> vals;
    DevBuf<int> flags;
    if (!vals.alloc(n) || !flags.alloc(n)) { res.error = "hipMalloc failed"; return res; }
    uc_lyap_output o{vals.p, flags.p, nullptr};
    const int rc = req.continuation
                       ? uc_run_lle_cont(impl_->eng, s, &r, req.continuation_reverse ? 1 : 0, &o, nullptr)
                       : uc_run_lle(impl_->eng, s, &r, &o, nullptr);
    if (rc != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<double> v;
    if (hipDeviceSynchronize() != hipSuccess || !vals.download(v, n) || !flags.download(res.flags, n)) {
        res.error = "device error";
        return res;
    }
    res.n_pts = req.n_pts;
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    res.param_lo = req.param_lo;
    res.param_hi = req.param_hi;
    res.lyapunov.resize(n);
    for (size_t i = 0; i < n; ++i)
        res.lyapunov[i] = (res.flags[i] == 1) ? v[i] : std::numeric_limits<double>::quiet_NaN();
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::curve1d_config(cfg, "1D LLE", req.base_values, req.initial_conditions, req.t_max, req.NT,
                            req.transient_time, req.h, req.eps, req.param_index, req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t i = 0; i < n; ++i)
            csv::lle1d_row(out, csv::param_value_at(i, req.n_pts, req.param_lo, req.param_hi), res.lyapunov[i]);
    }
    done(req.progress);
    res.ok = true;
    return res;
}

LS1DResult ParametricEngine::run_ls_1d(const LS1DRequest& req) {
    LS1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (req.use_cpu) { res.error = "CPU sweeps are not part of this engine (use_cpu); the GPU path runs the same algorithm"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 1);
    r.c.axis[0] = make_axis(req.sweep_over_h, req.sweep_over_var, req.var_sweep_index, req.param_index, req.n_pts,
                            req.param_lo, req.param_hi, req.log_scale);
    const size_t n = (size_t)req.n_pts, N = req.system.vars.size();
    DevBuf<double> vals;
    DevBuf<int> flags;
    if (!vals.alloc(n * N) || !flags.alloc(n)) { res.error = "hipMalloc failed"; return res; }
    uc_lyap_output o{vals.p, flags.p, nullptr};
    const int rc = req.continuation
                       ? uc_run_ls_cont(impl_->eng, s, &r, req.continuation_reverse ? 1 : 0, &o, nullptr)
                       : uc_run_ls(impl_->eng, s, &r, &o, nullptr);
    if (rc != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<double> v;
    if (hipDeviceSynchronize() != hipSuccess || !vals.download(v, n * N) || !flags.download(res.flags, n)) {
        res.error = "device error";
        return res;
    }
    res.n_pts = req.n_pts;
    res.continuation_reverse = req.continuation && req.continuation_reverse;
    res.n_exponents = (int)N;
    res.param_lo = req.param_lo;
    res.param_hi = req.param_hi;
    res.spectrum.assign(n, std::vector<double>(N));
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < N; ++k) res.spectrum[i][k] = v[k * n + i];
    if (!req.csv_output_path.empty() && !req.continuation) {  // the reference writes no file for continuation
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::curve1d_config(cfg, "1D LS", req.base_values, req.initial_conditions, req.t_max, req.NT,
                            req.transient_time, req.h, req.eps, req.param_index, req.param_lo, req.param_hi);
        std::ofstream out(req.csv_output_path);
        for (size_t i = 0; i < n; ++i)
            csv::ls1d_row(out, csv::param_value_at(i, req.n_pts, req.param_lo, req.param_hi), res.spectrum[i].data(),
                          (int)N);
    }
    done(req.progress);
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
}  // namespace

Dft1DResult ParametricEngine::run_dft_1d(const Dft1DRequest& req) {
    Dft1DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (req.use_cpu) { res.error = "CPU sweeps are not part of this engine (use_cpu); the GPU path runs the same algorithm"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
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
    const size_t n = (size_t)req.n_pts, nf = (size_t)std::max(1, req.n_freq);
    DevBuf<double> ak, bk;
    DevBuf<int> flags;
    if (!ak.alloc(n * nf) || !bk.alloc(n * nf) || !flags.alloc(n)) { res.error = "hipMalloc failed"; return res; }
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
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::dft1d_config(cfg, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                          req.pre_scaller, req.writable_var,
                          req.sweep_over_h ? -1 : req.sweep_over_var ? req.var_sweep_index : req.param_index,
                          req.param_lo, req.param_hi, req.n_freq, req.freq_lo, req.freq_hi, req.window_type);
        std::ofstream a(req.csv_output_path + "_AkCOS.csv"), b(req.csv_output_path + "_BkSIN.csv");
        csv::dft1d_header(a, req.param_lo, req.param_hi, req.freq_lo, req.freq_hi);
        csv::dft1d_header(b, req.param_lo, req.param_hi, req.freq_lo, req.freq_hi);
        csv::dft1d_matrix(a, res.ak_cos.data(), req.n_pts, req.n_freq);
        csv::dft1d_matrix(b, res.bk_sin.data(), req.n_pts, req.n_freq);
    }
    done(req.progress);
    res.ok = true;
    return res;
}

FastSyncResult ParametricEngine::run_fastsync(const FastSyncRequest& req) {
    FastSyncResult res;
    res.mode = req.mode;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    const size_t n = req.system.vars.size();
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
    if (!err.alloc((size_t)count) || (req.mode == 0 && !traj.alloc((size_t)count * n))) {
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
    done(req.progress);
    res.ok = true;
    return res;
}

Bifurcation2DResult ParametricEngine::run_bifurcation_2d(const Bifurcation2DRequest& req) {
    Bifurcation2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    uc_bif_request r{};
    impl_->common(r.c, req.initial_conditions, req.base_values, req.h, req.transient_time, req.t_max, req.max_value,
                  req.precision);
    axes_2d(req, r.c);
    r.writable_var = req.writable_var;
    r.pre_scaller = req.pre_scaller;
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    DevBuf<int> periods;
    if (!periods.alloc(n)) { res.error = "hipMalloc failed"; return res; }
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
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        const int pov = (req.sweep_over_var == req.sweep_over_var_2) ? (req.sweep_over_var ? 0 : 1) : 2;
        csv::bif2d_config(cfg, pov, req.base_values, req.initial_conditions, req.t_max, req.transient_time, req.h,
                          req.pre_scaller, req.eps_dbscan, req.writable_var,
                          req.sweep_over_var ? req.var_sweep_index : req.param_index,
                          req.sweep_over_var_2 ? req.var_sweep_index_2 : req.param_index_2, req.param_lo, req.param_hi,
                          req.param_lo_2, req.param_hi_2, req.n_pts);
        std::ofstream out(req.csv_output_path);
        csv::grid(out, req.n_pts, res.values.data());
    }
    done(req.progress);
    res.ok = true;
    return res;
}

LLE2DResult ParametricEngine::run_lle_2d(const LLE2DRequest& req) {
    LLE2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 2);
    axes_2d(req, r.c);
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    DevBuf<double> vals;
    DevBuf<int> flags;
    if (!vals.alloc(n) || !flags.alloc(n)) { res.error = "hipMalloc failed"; return res; }
    uc_lyap_output o{vals.p, flags.p, nullptr};
    if (uc_run_lle(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return res; }
    if (hipDeviceSynchronize() != hipSuccess || !vals.download(res.values, n) || !flags.download(res.flags, n)) {
        res.error = "device error";
        return res;
    }
    for (size_t i = 0; i < n; ++i)
        if (res.flags[i] != 1) res.values[i] = std::numeric_limits<double>::quiet_NaN();
    min_max(res.values.data(), res.flags.data(), n, res.min_val, res.max_val);
    res.n_pts = req.n_pts;
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.param_lo_2 = req.param_lo_2; res.param_hi_2 = req.param_hi_2;
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        const int pov = (req.sweep_over_var == req.sweep_over_var_2) ? (req.sweep_over_var ? 0 : 1) : 2;
        csv::map2d_config(cfg, "2D LLE", pov, req.base_values, req.initial_conditions, req.t_max, req.NT,
                          req.transient_time, req.h, req.eps, req.sweep_over_var ? req.var_sweep_index : req.param_index,
                          req.sweep_over_var_2 ? req.var_sweep_index_2 : req.param_index_2, req.param_lo, req.param_hi,
                          req.param_lo_2, req.param_hi_2, req.n_pts, 0);
        std::ofstream out(req.csv_output_path);
        csv::grid(out, req.n_pts, res.values.data());
    }
    done(req.progress);
    res.ok = true;
    return res;
}

LS2DResult ParametricEngine::run_ls_2d(const LS2DRequest& req) {
    LS2DResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
    uc_lyap_request r{};
    lyap_common(impl_.get(), req, r, 2);
    axes_2d(req, r.c);
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts, N = req.system.vars.size();
    DevBuf<double> vals;
    DevBuf<int> flags;
    if (!vals.alloc(n * N) || !flags.alloc(n)) { res.error = "hipMalloc failed"; return res; }
    uc_lyap_output o{vals.p, flags.p, nullptr};
    if (uc_run_ls(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return res; }
    if (hipDeviceSynchronize() != hipSuccess || !vals.download(res.values, n * N) || !flags.download(res.flags, n)) {
        res.error = "device error";
        return res;
    }
    res.n_pts = req.n_pts;
    res.n_exponents = (int)N;
    res.min_val.resize(N);
    res.max_val.resize(N);
    for (size_t k = 0; k < N; ++k) min_max(res.values.data() + k * n, res.flags.data(), n, res.min_val[k], res.max_val[k]);
    res.param_lo = req.param_lo; res.param_hi = req.param_hi;
    res.param_lo_2 = req.param_lo_2; res.param_hi_2 = req.param_hi_2;
    if (!req.csv_output_path.empty()) {
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        const int pov = (req.sweep_over_var == req.sweep_over_var_2) ? (req.sweep_over_var ? 0 : 1) : 2;
        csv::map2d_config(cfg, "2D LS", pov, req.base_values, req.initial_conditions, req.t_max, req.NT,
                          req.transient_time, req.h, req.eps, req.sweep_over_var ? req.var_sweep_index : req.param_index,
                          req.sweep_over_var_2 ? req.var_sweep_index_2 : req.param_index_2, req.param_lo, req.param_hi,
                          req.param_lo_2, req.param_hi_2, req.n_pts, (int)N);
        std::ofstream out(req.csv_output_path);
        csv::ls2d_cells(out, req.n_pts, (int)N, res.values.data());
    }
    done(req.progress);
    res.ok = true;
    return res;
}

BasinsResult ParametricEngine::run_basins(const BasinsRequest& req) {
    BasinsResult res;
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (!impl_->eng) { res.error = "engine not initialised"; return res; }
    if (cancelled(req.cancel)) { res.cancelled = true; res.error = "Cancelled by user"; return res; }
    uc_system* s = impl_->system(req.system, res.error);
    if (!s) return res;
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
    r.n_sections = req.n_sections;
    r.n_record = 16;
    r.max_period = 8;
    r.label_vars = 0;
    r.period_tol = 1e-4;
    r.quantum = 1e-3;
    const size_t n = (size_t)req.n_pts * (size_t)req.n_pts;
    const int lbl = req.system.vars.size() < 2 ? (int)req.system.vars.size() : 2;
    DevBuf<int> labels, periods;
    DevBuf<double> points;
    if (!labels.alloc(n) || !periods.alloc(n) || !points.alloc(n * (size_t)lbl)) { res.error = "hipMalloc failed"; return res; }
    int n_att = 0;
    uc_basins_output o{labels.p, periods.p, points.p, nullptr, 0, &n_att, nullptr, nullptr};
    if (uc_run_basins(impl_->eng, s, &r, &o, nullptr) != UC_OK) { res.error = impl_->err(); return res; }
    std::vector<double> pts;
    if (!labels.download(res.basin_idx, n) || !periods.download(res.periods, n) || !points.download(pts, n * (size_t)lbl)) {
        res.error = "device error";
        return res;
    }
    res.cycle_x.assign(pts.begin(), pts.begin() + (std::ptrdiff_t)n);
    for (size_t i = 0; i < n; ++i)
        if (res.periods[i] <= 0) res.cycle_x[i] = std::nan("");
    res.helpful_array.resize(n);
    for (size_t i = 0; i < n; ++i) res.helpful_array[i] = res.periods[i] >= 0 ? 1 : 0;
    res.n_clusters = n_att;
    res.min_cluster_idx = 0;
    res.n_pts = req.n_pts;
    res.axis_x_lo = req.axis_x_lo; res.axis_x_hi = req.axis_x_hi;
    res.axis_y_lo = req.axis_y_lo; res.axis_y_hi = req.axis_y_hi;
    res.axis_x_var = req.axis_x_var; res.axis_y_var = req.axis_y_var;
    if (!req.csv_output_path.empty()) {
        // the reference's five files (U-CUDA/data_export.cpp:472-544); its
        // DBSCAN features (_1 mean peak, _2 mean interval) have no equivalent
        // here: _1 holds the canonical cycle point's first coordinate and _2
        // the Poincare period, non-finite / no cycle written as 999
        std::ofstream cfg(req.csv_output_path + "_config.csv");
        csv::basins_config(cfg, req.base_values, req.initial_conditions, req.section_period * req.n_sections, 0.0,
                           req.h, 1, 1e-3, 0, req.axis_x_var, req.axis_y_var, req.axis_x_lo, req.axis_x_hi,
                           req.axis_y_lo, req.axis_y_hi);
        auto with_ranges = [&](const std::string& path) {
            auto f = std::make_unique<std::ofstream>(path);
            csv::basins_ranges(*f, req.axis_x_lo, req.axis_x_hi, req.axis_y_lo, req.axis_y_hi);
            return f;
        };
        csv::grid_int(*with_ranges(req.csv_output_path), req.n_pts, res.basin_idx.data());
        csv::grid_finite(*with_ranges(req.csv_output_path + "_1.csv"), req.n_pts, res.cycle_x.data());
        std::vector<double> per(res.periods.begin(), res.periods.end());
        for (auto& v : per) v = v > 0 ? v : std::nan("");
        csv::grid_finite(*with_ranges(req.csv_output_path + "_2.csv"), req.n_pts, per.data());
        csv::grid_int(*with_ranges(req.csv_output_path + "_3.csv"), req.n_pts, res.helpful_array.data());
    }
    done(req.progress);
    res.ok = true;
    return res;
}

}  // namespace ucuda
