// data_export.hpp — CSV files in the reference's on-disk format, so results
// can be diffed against U-CUDA's exports (U-CUDA/data_export.cpp:16-58,
// 176-294, 300-466, 472-544). Numbers are printed with 15 significant
// digits (set_precision, U-CUDA/configCUDA.h:137); separators ", ".
#pragma once
#include <fstream>
#include <string>
#include <vector>

namespace ucuda {

// Snapshots of the request fields a result's CSV header needs, carried in
// every *Result (the reference's data_export snapshot structs,
// U-CUDA/data_export.h:44-210, same member names) so a session or GUI can
// write the files after the run.
namespace data_export {
struct Bif1DSnapshot {
    std::vector<double> values, initial_conditions;
    double tMax = 0.0, transientTime = 0.0, h = 0.0;
    int preScaller = 0, writableVar = 0, indexOfMutVar = 0;
    double range_lo = 0.0, range_hi = 0.0;
};
struct Dft1DSnapshot {
    std::vector<double> values, initial_conditions;
    double tMax = 0.0, transientTime = 0.0, h = 0.0;
    int preScaller = 0, writableVar = 0, indexOfMutVar = 0;
    double range_lo = 0.0, range_hi = 0.0;
    int n_freq = 0;
    double freq_lo = 0.0, freq_hi = 0.0;
    int window_type = 1;
};
struct LLE1DSnapshot {
    std::vector<double> values, initial_conditions;
    double tMax = 0.0, NT = 0.0, transientTime = 0.0, h = 0.0, eps = 0.0;
    int indexOfMutVar = 0;
    double range_lo = 0.0, range_hi = 0.0;
};
using LS1DSnapshot = LLE1DSnapshot;
struct Bif2DSnapshot {
    std::vector<double> values, initial_conditions;
    int par_or_var = 1;
    double tMax = 0.0, transientTime = 0.0, h = 0.0;
    int preScaller = 0;
    double eps_dbscan = 0.0;
    int writableVar = 0, indexOfMutVar = 0, indexOfMutVar2 = 0;
    double range1_lo = 0.0, range1_hi = 0.0, range2_lo = 0.0, range2_hi = 0.0;
    int n_pts = 0;
};
struct LLE2DSnapshot {
    std::vector<double> values, initial_conditions;
    int par_or_var = 1;
    double tMax = 0.0, NT = 0.0, transientTime = 0.0, h = 0.0, eps = 0.0;
    int indexOfMutVar = 0, indexOfMutVar2 = 0;
    double range1_lo = 0.0, range1_hi = 0.0, range2_lo = 0.0, range2_hi = 0.0;
    int n_pts = 0;
};
struct LS2DSnapshot : LLE2DSnapshot {
    int n_exponents = 0;
};
struct BasinsSnapshot {
    std::vector<double> values, initial_conditions;
    double tMax = 0.0, transientTime = 0.0, h = 0.0;
    int preScaller = 0;
    double eps_dbscan = 0.0;
    int writableVar = 0, axis_x_var = 0, axis_y_var = 0;
    double axis_x_lo = 0.0, axis_x_hi = 0.0, axis_y_lo = 0.0, axis_y_hi = 0.0;
    int n_pts = 0, feature1 = 0, feature2 = 0;
    double mult1 = 0.0, mult2 = 0.0;
};
struct FastSyncSnapshot {
    int mode = 0;
    std::vector<double> values, ic_master, ic_slave, k_forward, k_backward;
    std::vector<std::string> var_names;
    double h = 0.0;
    int iter_of_synchr = 0, preScaller = 0;
    double window = 0.0;
    int type_of_synch = 0, error_estim = 0;
    double fs_error_trs = 0.0, tMax = 0.0, transientTime = 0.0;
    int axis_x_var = 0, axis_y_var = 0;
    double axis_x_lo = 0.0, axis_x_hi = 0.0, axis_y_lo = 0.0, axis_y_hi = 0.0;
    int n_pts = 0;
    bool grid_swap_master_slave = false;
};
}  // namespace data_export

namespace csv {

// Header block "a[n] = { ... }" / "X0[n] = { ... }".
void values_and_ic(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic);

// Sweep value written next to 1-D results: lo + (hi - lo) * i / (n - 1), n <= 1 -> lo.
double param_value_at(size_t i, int n_pts, double lo, double hi);

void bif1d_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                  double transient, double h, int pre_scaller, int writable_var, int index_of_mut_var, double lo,
                  double hi);
// One row per peak "param, peak, interval"; "param, 0, 0" without peaks,
// "param, 0, -1" for a fixed point.
void bif1d_rows(std::ostream& out, double param, int npeaks, const double* peaks, const double* times);

// DFT: config, then per file the two range lines "lo, hi" and one row of
// n_freq coefficients per sweep point.
void dft1d_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                  double transient, double h, int pre_scaller, int writable_var, int index_of_mut_var, double lo,
                  double hi, int n_freq, double freq_lo, double freq_hi, int window_type);
void dft1d_header(std::ostream& out, double lo, double hi, double freq_lo, double freq_hi);
void dft1d_matrix(std::ostream& out, const double* matrix, int n_rows, int n_freq);

void curve1d_config(std::ostream& out, const char* title, const std::vector<double>& values,
                    const std::vector<double>& ic, double t_max, double NT, double transient, double h, double eps,
                    int index_of_mut_var, double lo, double hi);
void lle1d_row(std::ostream& out, double param, double lyapunov);
void ls1d_row(std::ostream& out, double param, const double* exponents, int n_exponents);

void map2d_config(std::ostream& out, const char* title, int par_or_var, const std::vector<double>& values,
                  const std::vector<double>& ic, double t_max, double NT, double transient, double h, double eps,
                  int index1, int index2, double lo1, double hi1, double lo2, double hi2, int n_pts, int n_exponents);
void bif2d_config(std::ostream& out, int par_or_var, const std::vector<double>& values, const std::vector<double>& ic,
                  double t_max, double transient, double h, int pre_scaller, double eps_dbscan, int writable_var,
                  int index1, int index2, double lo1, double hi1, double lo2, double hi2, int n_pts);
// n_pts rows of n_pts values (values[iy * n + ix]).
void grid(std::ostream& out, int n_pts, const double* values);
// one row per cell (iy * n + ix) with its n_exponents values (plane-major input).
void ls2d_cells(std::ostream& out, int n_pts, int n_exponents, const double* values);
// basins: "lo hi" range lines then the integer grid.
void basins_ranges(std::ostream& out, double xlo, double xhi, double ylo, double yhi);
void grid_int(std::ostream& out, int n_pts, const int* values);
// n_pts rows of doubles with non-finite values written as 999.
void grid_finite(std::ostream& out, int n_pts, const double* values);
void basins_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                   double transient, double h, int pre_scaller, double eps, int writable_var, int axis_x_var,
                   int axis_y_var, double xlo, double xhi, double ylo, double yhi);

// Fast synchronisation (write_fastsync_config / _attractor / _grid,
// U-CUDA/data_export.cpp:553-657).
void fastsync_config(std::ostream& out, int mode, const std::vector<double>& values, const std::vector<double>& ic_master,
                     const std::vector<double>& ic_slave, const std::vector<double>& k_forward,
                     const std::vector<double>& k_backward, double h, int iter_of_synchr, int pre_scaller,
                     double window, int type_of_synch, int error_estim, double fs_error_trs, double t_max,
                     double transient, int axis_x_var, int axis_y_var, double x_lo, double x_hi, double y_lo,
                     double y_hi, int n_pts, bool swap);
void fastsync_attractor(std::ostream& out, const double* traj, const double* err, int n_pts, int n_vars,
                        const std::vector<std::string>& var_names);
void fastsync_grid(std::ostream& out, const double* heatmap, int n, double x_lo, double x_hi, double y_lo,
                   double y_hi);

}  // namespace csv
}  // namespace ucuda
