// parametric_engine.hpp — drop-in C++ facade over the C ABI.
//
// Same class, method and field names as the reference facade
// (U-CUDA/parametric_engine.h:27-126, 211-482, 581-889), so a session layer
// written against it keeps compiling: Request in, Result by value, no
// exceptions (errors in Result::error), results carry their `snapshot`.
//
// Which integrator runs (krs_body and `system`):
//  * `krs_body` is the session's step body (codegen_scheme output for the
//    chosen scheme, or a custom scheme). When `system` is empty, or the body
//    is not codegen's RK4 text of `system`, the body itself is compiled and
//    stepped (uc_system_create_krs): bifurcation, DFT, fast synchronisation
//    and the reference's finite-difference Lyapunov method, exactly the
//    reference's algorithms on that step.
//  * When `system` (vars / params / right-hand sides) is set and the body is
//    empty or its RK4 text, the engine integrates the right-hand side with
//    its RK4, which adds the variational Lyapunov method (`method`) and is
//    the only input basins accept (adaptive Dormand-Prince needs the
//    right-hand side; a basins request with a body only is an error).
//
// Cancellation and progress: maps of more than 2^18 cells (bifurcation,
// LLE, LS) run in up to 16 pieces of the grid; the cancel token is checked
// and `progress` advanced between pieces (the reference's chunk loop,
// U-CUDA/parametric_engine.cpp:1803-1886). Other analyses check the token
// before starting; basins report progress_phase 1 (integration) then 2
// (labels).
//
// Additions: `system`, `precision` (UC_PREC_F64 / UC_PREC_F32C), `method`,
// `max_peaks`, and for basins the adaptive integrator's fields. Results use
// the reference's sentinels: LLE 999 -> NaN with flag -1; LS keeps 999 with
// flag -1; bifurcation flags are the peak count, -1 (fixed point / bad step)
// or 0 (diverged); 2-D maps are row-major values[iy * n + ix], LS planes
// values[k * n * n + iy * n + ix].
#pragma once
#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "codegen.hpp"
#include "data_export.hpp"

namespace ucuda {

struct Bifurcation1DRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> initial_conditions;
    std::vector<double> base_values;
    int param_index = 0;
    bool sweep_over_var = false;
    int var_sweep_index = 0;
    bool sweep_over_h = false;
    bool log_scale = false;
    bool continuation = false;
    bool continuation_reverse = false;
    bool use_cpu = false;
    double param_lo = 0.0, param_hi = 1.0;
    int n_pts = 1000;
    int writable_var = 0;
    double h = 0.01, transient_time = 0.0, t_max = 0.0;
    int pre_scaller = 1;
    double max_value = 1.0e6;
    int max_peaks = 2500;
    int precision = 0;  // UC_PREC_F64 / UC_PREC_F32C (F32C records its window in double)
    std::string csv_output_path;  // non-empty: write <path> and <path>_config.csv in the reference format
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct Bifurcation1DResult {
    bool ok = false, cancelled = false;
    std::string error;
    bool continuation_reverse = false;  // rows in visiting order: hi first when reversed
    double param_lo = 0.0, param_hi = 1.0;
    int n_pts = 0, record_steps = 0;
    std::vector<std::vector<double>> bifurcation_points;
    std::vector<std::vector<double>> peak_times;
    std::vector<int> flags;
    data_export::Bif1DSnapshot snapshot;
};

// 2-D bifurcation map: period = DBSCAN cluster count of each cell's peaks
// (U-CUDA/parametric_engine.h:499-569).
struct Bifurcation2DRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> initial_conditions, base_values;
    bool sweep_over_var = false, sweep_over_h = false, log_scale = false;
    int var_sweep_index = 0, param_index = 0;
    bool sweep_over_var_2 = false, sweep_over_h_2 = false, log_scale_2 = false;
    int var_sweep_index_2 = 0, param_index_2 = 0;
    double param_lo = 0.0, param_hi = 1.0, param_lo_2 = 0.0, param_hi_2 = 1.0;
    int n_pts = 200;
    int writable_var = 0;
    double h = 0.01, transient_time = 0.0, t_max = 100.0;
    int pre_scaller = 1;
    double max_value = 1.0e6;
    double eps_dbscan = 0.1;
    int precision = 0;
    std::string csv_output_path;  // non-empty: write <path> and <path>_config.csv in the reference format
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct Bifurcation2DResult {
    bool ok = false, cancelled = false;
    std::string error;
    int n_pts = 0;
    double param_lo = 0.0, param_hi = 1.0, param_lo_2 = 0.0, param_hi_2 = 1.0;
    std::vector<double> values;  // [iy * n + ix] period; -1 fixed point / bad step, 0 no peaks / diverged
    std::vector<int> flags;      // 1 ok, -1 when the period is negative
    double min_val = 0.0, max_val = 0.0;
    data_export::Bif2DSnapshot snapshot;
};

// 1-D DFT sweep (U-CUDA/parametric_engine.h:140-226).
struct Dft1DRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> initial_conditions, base_values;
    int param_index = 0;
    bool sweep_over_var = false;
    int var_sweep_index = 0;
    bool sweep_over_h = false, log_scale = false;
    bool continuation = false, continuation_reverse = false, use_cpu = false;
    double param_lo = 0.0, param_hi = 1.0;
    int n_pts = 1000;
    int writable_var = 0;
    double h = 0.01, transient_time = 0.0, t_max = 0.0;
    int pre_scaller = 1;
    double max_value = 1.0e6;
    int n_freq = 200;
    double freq_lo = 0.0, freq_hi = 10.0;
    bool freq_log_scale = false;
    int window_type = 1;  // 0 none, 1 Hanning, 2 Hamming
    int precision = 0;
    std::string csv_output_path;  // non-empty: <path>_config.csv, <path>_AkCOS.csv, <path>_BkSIN.csv
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct Dft1DResult {
    bool ok = false, cancelled = false;
    std::string error;
    bool continuation_reverse = false;
    double param_lo = 0.0, param_hi = 1.0;
    int n_pts = 0, n_freq = 0;
    double freq_lo = 0.0, freq_hi = 0.0;
    std::vector<double> ak_cos, bk_sin;  // [pt * n_freq + f]
    std::vector<int> flags;              // 1 ok, 0 diverged, -1 fixed point
    data_export::Dft1DSnapshot snapshot;
};

struct LLE1DRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> initial_conditions, base_values;
    bool sweep_over_var = false;
    int var_sweep_index = 0;
    bool sweep_over_h = false, log_scale = false;
    int param_index = 0;
    bool continuation = false, continuation_reverse = false, use_cpu = false;
    double param_lo = 0.0, param_hi = 1.0;
    int n_pts = 500;
    double h = 0.01, transient_time = 0.0, t_max = 100.0, NT = 1.0, eps = 1.0e-4, max_value = 1.0e6;
    // method: UC_METHOD_VARIATIONAL (needs `system`) or UC_METHOD_FINITE_DIFF
    // (the reference's; always used when the step body is stepped)
    int precision = 0, method = 0;
    std::string csv_output_path;  // non-empty: write <path> and <path>_config.csv in the reference format
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct LLE1DResult {
    bool ok = false, cancelled = false;
    std::string error;
    bool continuation_reverse = false;
    int n_pts = 0;
    double param_lo = 0.0, param_hi = 1.0;
    std::vector<double> lyapunov;  // NaN where the point diverged
    std::vector<int> flags;        // 1 ok, -1 diverged
    data_export::LLE1DSnapshot snapshot;
};

using LS1DRequest = LLE1DRequest;
struct LS1DResult {
    bool ok = false, cancelled = false;
    std::string error;
    bool continuation_reverse = false;
    int n_pts = 0, n_exponents = 0;
    double param_lo = 0.0, param_hi = 1.0;
    std::vector<std::vector<double>> spectrum;  // 999 where the point diverged (NaN in continuation)
    std::vector<int> flags;
    data_export::LS1DSnapshot snapshot;
};

struct LLE2DRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> initial_conditions, base_values;
    bool sweep_over_var = false, sweep_over_var_2 = false;
    bool sweep_over_h = false, sweep_over_h_2 = false;
    bool log_scale = false, log_scale_2 = false;
    int param_index = 0, var_sweep_index = 0, param_index_2 = 0, var_sweep_index_2 = 0;
    double param_lo = 0.0, param_hi = 1.0, param_lo_2 = 0.0, param_hi_2 = 1.0;
    int n_pts = 200;
    double h = 0.01, transient_time = 0.0, t_max = 100.0, NT = 1.0, eps = 1.0e-4, max_value = 1.0e6;
    int precision = 0, method = 0;
    std::string csv_output_path;  // non-empty: write <path> and <path>_config.csv in the reference format
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct LLE2DResult {
    bool ok = false, cancelled = false;
    std::string error;
    int n_pts = 0;
    double param_lo = 0.0, param_hi = 1.0, param_lo_2 = 0.0, param_hi_2 = 1.0;
    std::vector<double> values;  // [iy * n + ix], NaN where diverged
    std::vector<int> flags;
    double min_val = 0.0, max_val = 0.0;
    data_export::LLE2DSnapshot snapshot;
};

using LS2DRequest = LLE2DRequest;
struct LS2DResult {
    bool ok = false, cancelled = false;
    std::string error;
    int n_pts = 0, n_exponents = 0;
    double param_lo = 0.0, param_hi = 1.0, param_lo_2 = 0.0, param_hi_2 = 1.0;
    std::vector<double> values;  // [k * n * n + iy * n + ix]
    std::vector<int> flags;
    std::vector<double> min_val, max_val;
    data_export::LS2DSnapshot snapshot;
};

// Basins of attraction (U-CUDA/parametric_engine.h:659-723). This engine
// labels attractors by adaptive Dormand-Prince 5(4) and Poincare sections
// every section_period, not by DBSCAN over averaged peaks. The reference's
// fields map as follows:
//   transient_time + t_max -> the integrated time: n_sections =
//     ceil((transient_time + t_max) / section_period) when n_sections is 0;
//   eps_dbscan -> the clustering radius: attractor signatures are quantised
//     with this step and adjacent ones of equal period merged;
//   h -> the integrator's first step;
//   writable_var, pre_scaller, feature1 / feature2, mult1 / mult2 -> kept
//     in the snapshot (the features are fixed: the canonical cycle point).
struct BasinsRequest {
    System system;          // required: the adaptive integrator needs the right-hand side
    std::string krs_body;   // the session's step body (not used by the adaptive integrator)
    int amountOfX = 0;
    std::vector<double> initial_conditions, base_values;
    int axis_x_var = 0, axis_y_var = 1;
    double axis_x_lo = -10.0, axis_x_hi = 10.0, axis_y_lo = -10.0, axis_y_hi = 10.0;
    int n_pts = 200;
    int writable_var = 0;
    double h = 0.01;
    double t_max = 1000.0;
    double transient_time = 10000.0;
    int pre_scaller = 1;
    double max_value = 1.0e6;
    double eps_dbscan = 0.5;
    int feature1 = 0, feature2 = 1;
    double mult1 = 1.0, mult2 = 1.0;
    // the adaptive integrator
    double rtol = 1e-9, atol = 1e-9;
    double section_period = 6.283185307179586;
    int n_sections = 0;     // 0: from transient_time + t_max
    int precision = 0;
    std::string csv_output_path;  // non-empty: write <path> and <path>_config.csv in the reference format
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
    std::shared_ptr<std::atomic<int>> progress_phase;  // 1 integration, 2 labels
};

struct BasinsResult {
    bool ok = false, cancelled = false;
    std::string error;
    int n_pts = 0;
    double axis_x_lo = 0.0, axis_x_hi = 0.0, axis_y_lo = 0.0, axis_y_hi = 0.0;
    int axis_x_var = 0, axis_y_var = 1;
    std::vector<int> basin_idx;        // >= 1 attractor (canonical order), 0 unbound / unresolved
    // per-cell features (the canonical cycle point's two coordinates; the
    // mean section point when no cycle closes); 999 for unbound cells
    std::vector<double> avg_peaks;
    std::vector<double> avg_intervals;
    std::vector<int> helpful_array;    // 1 settled on an attractor, 0 unbound / unresolved
    int n_clusters = 0;
    int min_cluster_idx = 0;
    double avg_peaks_min = 0.0, avg_peaks_max = 0.0;
    double avg_intervals_min = 0.0, avg_intervals_max = 0.0;
    std::vector<int> periods;          // Poincare period per cell (0 none, -1 unbound, -2 unresolved)
    std::vector<double> cycle_x;       // avg_peaks with NaN where no cycle closed
    data_export::BasinsSnapshot snapshot;
};

// Re-labelling from a previous run_basins's features without integrating
// again (U-CUDA/parametric_engine.h:733-752): cells with helpful_array 1 are
// grouped by their features quantised with eps_dbscan, adjacent groups
// merged, clusters numbered 1..K in canonical order (smallest feature
// pair first); the others get 0.
struct BasinsReclusterRequest {
    std::string krs_body;
    int amountOfX = 0;
    int n_pts = 0;
    double eps_dbscan = 0.5;
    std::vector<double> avg_peaks, avg_intervals;
    std::vector<int> helpful_array;
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct BasinsReclusterResult {
    bool ok = false, cancelled = false;
    std::string error;
    int n_pts = 0;
    std::vector<int> basin_idx;
    int n_clusters = 0;
    int min_cluster_idx = 0;
};

// Fast synchronisation (U-CUDA/parametric_engine.h:767-842).
struct FastSyncRequest {
    System system;
    std::string krs_body;
    int amountOfX = 0;
    std::vector<double> values;  // a[]
    int mode = 0;                // 0 on attractor, 1 on grid
    double h = 0.01;
    int iter_of_synchr = 100;
    int pre_scaller = 1;
    double max_value = 1e6;      // kept for the request's shape; the reference's loops do not test it
    std::vector<double> k_forward, k_backward, ic_master, ic_slave;
    double t_max = 100.0, transient_time = 0.0;
    double window = 50.0;
    int axis_x_var = 0, axis_y_var = 1;
    double axis_x_lo = -10.0, axis_x_hi = 10.0, axis_y_lo = -10.0, axis_y_hi = 10.0;
    int n_pts = 200;
    bool grid_swap_master_slave = false;
    int type_of_synch = 0;
    int error_estim = 2;
    double fs_error_trs = 1e-12;
    std::string csv_output_path;  // non-empty: <path> and <path>_config.csv
    std::vector<std::string> var_names;
    std::shared_ptr<std::atomic<bool>> cancel;
    std::shared_ptr<std::atomic<float>> progress;
};

struct FastSyncResult {
    bool ok = false, cancelled = false;
    std::string error;
    int mode = 0;
    int n_pts_traj = 0, amountOfX_traj = 0;
    std::vector<double> traj_full;   // mode 0: [point * amountOfX + var]
    std::vector<double> sync_error;  // mode 0: per point
    int n_pts_grid = 0;
    double axis_x_lo = 0.0, axis_x_hi = 0.0, axis_y_lo = 0.0, axis_y_hi = 0.0;
    int axis_x_var = 0, axis_y_var = 1;
    std::vector<double> heatmap;     // mode 1: [iy * n + ix]
    double min_val = 0.0, max_val = 0.0;
    data_export::FastSyncSnapshot snapshot;
};

class ParametricEngine {
public:
    explicit ParametricEngine(int device = 0);
    ~ParametricEngine();
    ParametricEngine(const ParametricEngine&) = delete;
    ParametricEngine& operator=(const ParametricEngine&) = delete;

    Bifurcation1DResult run_bifurcation_1d(const Bifurcation1DRequest& req);
    Dft1DResult run_dft_1d(const Dft1DRequest& req);
    Bifurcation2DResult run_bifurcation_2d(const Bifurcation2DRequest& req);
    LLE1DResult run_lle_1d(const LLE1DRequest& req);
    LS1DResult run_ls_1d(const LS1DRequest& req);
    LLE2DResult run_lle_2d(const LLE2DRequest& req);
    LS2DResult run_ls_2d(const LS2DRequest& req);
    BasinsResult run_basins(const BasinsRequest& req);
    BasinsReclusterResult run_basins_recluster(const BasinsReclusterRequest& req);
    FastSyncResult run_fastsync(const FastSyncRequest& req);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ucuda
