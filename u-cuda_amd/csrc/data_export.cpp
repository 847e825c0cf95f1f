// data_export.cpp — CSV writers (see the header for the format source).
#include "data_export.hpp"

#include <cmath>
#include <iomanip>

namespace ucuda {
namespace csv {

namespace {
constexpr int kPrecision = 15;

void list(std::ostream& out, const char* name, const std::vector<double>& v) {
    out << name << "[" << v.size() << "] = { ";
    for (size_t k = 0; k < v.size(); ++k) {
        out << v[k];
        out << (k + 1 != v.size() ? ", " : " }\n");
    }
}
}  // namespace

void values_and_ic(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic) {
    list(out, "a", values);
    list(out, "X0", ic);
}

double param_value_at(size_t i, int n_pts, double lo, double hi) {
    if (n_pts <= 1) return lo;
    return lo + (hi - lo) * (double)i / (double)(n_pts - 1);
}

void bif1d_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                  double transient, double h, int pre_scaller, int writable_var, int index_of_mut_var, double lo,
                  double hi) {
    out << std::setprecision(kPrecision);
    out << "1D classical bifurcation\nParameter estimation\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nTT = " << transient << "\nh = " << h << "\ndecimator = " << pre_scaller
        << "\nindexVar for peakfinder = " << writable_var << "\nindexPar for estimation = " << index_of_mut_var
        << "\nstart value = " << lo << ", stop value = " << hi << "\n";
}

void bif1d_rows(std::ostream& out, double param, int npeaks, const double* peaks, const double* times) {
    out << std::setprecision(kPrecision);
    if (npeaks == 0) out << param << ", " << 0 << ", " << 0 << '\n';
    else if (npeaks == -1) out << param << ", " << 0 << ", " << -1 << '\n';
    else
        for (int j = 0; j < npeaks; ++j) out << param << ", " << peaks[j] << ", " << times[j] << '\n';
}

void dft1d_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                  double transient, double h, int pre_scaller, int writable_var, int index_of_mut_var, double lo,
                  double hi, int n_freq, double freq_lo, double freq_hi, int window_type) {
    out << std::setprecision(kPrecision);
    out << "1D bifurcation DFT\nParameter estimation\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nTT = " << transient << "\nh = " << h << "\ndecimator = " << pre_scaller
        << "\nindexVar for DFT = " << writable_var << "\nindexPar for estimation = " << index_of_mut_var
        << "\nstart value = " << lo << ", stop value = " << hi << "\n";
    out << "n_freq = " << n_freq << "\nfreq_lo = " << freq_lo << ", freq_hi = " << freq_hi << "\n";
    out << "window = " << (window_type == 0 ? "None" : window_type == 2 ? "Hamming" : "Hanning") << "\n";
}

void dft1d_header(std::ostream& out, double lo, double hi, double freq_lo, double freq_hi) {
    out << std::setprecision(kPrecision) << lo << ", " << hi << "\n" << freq_lo << ", " << freq_hi << "\n";
}

void dft1d_matrix(std::ostream& out, const double* matrix, int n_rows, int n_freq) {
    out << std::setprecision(kPrecision);
    for (int i = 0; i < n_rows; ++i) {
        for (int j = 0; j < n_freq; ++j) {
            out << matrix[(size_t)i * n_freq + j];
            out << (j != n_freq - 1 ? ", " : "\n");
        }
    }
}

void curve1d_config(std::ostream& out, const char* title, const std::vector<double>& values,
                    const std::vector<double>& ic, double t_max, double NT, double transient, double h, double eps,
                    int index_of_mut_var, double lo, double hi) {
    out << std::setprecision(kPrecision);
    out << title << "\nParameter estimation\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nNT = " << NT << "\nTT = " << transient << "\n";
    out << "h = " << h << "\neps = " << eps << "\n";
    out << "indexPar = " << index_of_mut_var << "\n";
    out << "start value = " << lo << ", stop value = " << hi << "\n";
}

void lle1d_row(std::ostream& out, double param, double lyapunov) {
    out << std::setprecision(kPrecision) << param << ", " << lyapunov << '\n';
}

void ls1d_row(std::ostream& out, double param, const double* exponents, int n_exponents) {
    out << std::setprecision(kPrecision) << param;
    for (int j = 0; j < n_exponents; ++j) out << ", " << exponents[j];
    out << '\n';
}

void map2d_config(std::ostream& out, const char* title, int par_or_var, const std::vector<double>& values,
                  const std::vector<double>& ic, double t_max, double NT, double transient, double h, double eps,
                  int index1, int index2, double lo1, double hi1, double lo2, double hi2, int n_pts, int n_exponents) {
    out << std::setprecision(kPrecision) << title << "\n";
    if (par_or_var == 1) out << "Parameter estimation\n";
    else if (par_or_var == 0) out << "Initial conditions estimation\n";
    else if (par_or_var == 2) out << "Mixed: x=IC, y=parameter\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nNT = " << NT << "\nTT = " << transient << "\n";
    out << "h = " << h << "\neps = " << eps << "\n";
    out << "indices = " << index1 << ", " << index2 << "\n";
    out << "axis1: " << lo1 << " .. " << hi1 << "\n";
    out << "axis2: " << lo2 << " .. " << hi2 << "\n";
    out << "n_pts = " << n_pts << "x" << n_pts << "\n";
    if (n_exponents > 0) out << "exponents = " << n_exponents << "\n";
}

void bif2d_config(std::ostream& out, int par_or_var, const std::vector<double>& values, const std::vector<double>& ic,
                  double t_max, double transient, double h, int pre_scaller, double eps_dbscan, int writable_var,
                  int index1, int index2, double lo1, double hi1, double lo2, double hi2, int n_pts) {
    out << std::setprecision(kPrecision) << "2D bifurcation (DBSCAN)\n";
    if (par_or_var == 1) out << "Parameter estimation\n";
    else if (par_or_var == 0) out << "Initial conditions estimation\n";
    else if (par_or_var == 2) out << "Mixed: x=IC, y=parameter\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nTT = " << transient << "\nh = " << h << "\ndecimator = " << pre_scaller << "\n";
    out << "eps_DBSCAN = " << eps_dbscan << "\n";
    out << "indexVar for peakfinder = " << writable_var << "\n";
    out << "indices = " << index1 << ", " << index2 << "\n";
    out << "axis1: " << lo1 << " .. " << hi1 << "\n";
    out << "axis2: " << lo2 << " .. " << hi2 << "\n";
    out << "n_pts = " << n_pts << "x" << n_pts << "\n";
}

void grid(std::ostream& out, int n_pts, const double* values) {
    out << std::setprecision(kPrecision);
    for (int iy = 0; iy < n_pts; ++iy) {
        for (int ix = 0; ix < n_pts; ++ix) {
            out << values[(size_t)iy * n_pts + ix];
            if (ix + 1 < n_pts) out << ", ";
        }
        out << '\n';
    }
}

void ls2d_cells(std::ostream& out, int n_pts, int n_exponents, const double* values) {
    out << std::setprecision(kPrecision);
    const size_t total = (size_t)n_pts * n_pts;
    for (size_t cell = 0; cell < total; ++cell) {
        for (int k = 0; k < n_exponents; ++k) {
            out << values[(size_t)k * total + cell];
            if (k + 1 < n_exponents) out << ", ";
        }
        out << '\n';
    }
}

void basins_ranges(std::ostream& out, double xlo, double xhi, double ylo, double yhi) {
    out << std::setprecision(kPrecision) << xlo << " " << xhi << "\n" << ylo << " " << yhi << "\n";
}

void grid_int(std::ostream& out, int n_pts, const int* values) {
    for (int iy = 0; iy < n_pts; ++iy) {
        for (int ix = 0; ix < n_pts; ++ix) {
            out << values[(size_t)iy * n_pts + ix];
            if (ix + 1 < n_pts) out << ", ";
        }
        out << '\n';
    }
}

void grid_finite(std::ostream& out, int n_pts, const double* values) {
    out << std::setprecision(kPrecision);
    for (int iy = 0; iy < n_pts; ++iy) {
        for (int ix = 0; ix < n_pts; ++ix) {
            const double v = values[(size_t)iy * n_pts + ix];
            out << (std::isfinite(v) ? v : 999.0);
            if (ix + 1 < n_pts) out << ", ";
        }
        out << '\n';
    }
}

void basins_config(std::ostream& out, const std::vector<double>& values, const std::vector<double>& ic, double t_max,
                   double transient, double h, int pre_scaller, double eps, int writable_var, int axis_x_var,
                   int axis_y_var, double xlo, double xhi, double ylo, double yhi) {
    out << std::setprecision(kPrecision) << "Basins of attraction\n";
    values_and_ic(out, values, ic);
    out << "CT = " << t_max << "\nTT = " << transient << "\nh = " << h << "\n";
    out << "decimator = " << pre_scaller << "\neps_DBSCAN = " << eps << "\n";
    out << "indexVar for peakfinder = " << writable_var << "\n";
    out << "indexVar for estimation = " << axis_x_var << ", " << axis_y_var << "\n";
    out << "start value_1 = " << xlo << ", stop value_1 = " << xhi << "\n";
    out << "start value_2 = " << ylo << ", stop value_2 = " << yhi << "\n";
}

static void fs_vec(std::ostream& out, const char* label, const std::vector<double>& v) {
    out << label << "[" << v.size() << "] = { ";
    for (size_t k = 0; k < v.size(); ++k) out << v[k] << (k + 1 != v.size() ? ", " : " }\n");
}

void fastsync_config(std::ostream& out, int mode, const std::vector<double>& values, const std::vector<double>& ic_master,
                     const std::vector<double>& ic_slave, const std::vector<double>& k_forward,
                     const std::vector<double>& k_backward, double h, int iter_of_synchr, int pre_scaller,
                     double window, int type_of_synch, int error_estim, double fs_error_trs, double t_max,
                     double transient, int axis_x_var, int axis_y_var, double x_lo, double x_hi, double y_lo,
                     double y_hi, int n_pts, bool swap) {
    out << std::setprecision(kPrecision);
    out << (mode == 0 ? "Fast Synchro (On Attractor)\n" : "Fast Synchro (On Grid)\n");
    out << "a[" << values.size() << "] = { ";
    for (size_t k = 0; k < values.size(); ++k) out << values[k] << (k + 1 != values.size() ? ", " : " }\n");
    fs_vec(out, "IC_master", ic_master);
    fs_vec(out, "IC_slave", ic_slave);
    fs_vec(out, "k_forward", k_forward);
    fs_vec(out, "k_backward", k_backward);
    out << "h = " << h << "\niter_of_synchr = " << iter_of_synchr << "\ndecimator = " << pre_scaller
        << "\nwindow = " << window << "\n";
    const char* estim = error_estim == 1 ? "# iters to reach FS_error_trs"
                        : error_estim == 2 ? "RMS at last point" : "RMS on last iter";
    out << "type_of_synch = " << (type_of_synch == 1 ? "Bidirectional" : "Unidirectional")
        << ", error_estim = " << estim << ", fs_error_trs = " << fs_error_trs << "\n";
    if (mode == 0) {
        out << "CT = " << t_max << "\nTT = " << transient << "\n";
    } else {
        out << "TT = " << transient << "\n";
        out << "axis_x_var = " << axis_x_var << ", axis_y_var = " << axis_y_var << "\n";
        out << "axis_x: " << x_lo << " .. " << x_hi << "\n";
        out << "axis_y: " << y_lo << " .. " << y_hi << "\n";
        out << "n_pts = " << n_pts << "x" << n_pts << "\n";
        out << "grid_swap_master_slave = " << (swap ? 1 : 0) << "\n";
    }
}

void fastsync_attractor(std::ostream& out, const double* traj, const double* err, int n_pts, int n_vars,
                        const std::vector<std::string>& var_names) {
    out << std::setprecision(kPrecision);
    for (int j = 0; j < n_vars; ++j) {
        if (j < (int)var_names.size()) out << var_names[(size_t)j];
        else out << "x" << j;
        out << ",";
    }
    out << "sync_error\n";
    for (int i = 0; i < n_pts; ++i) {
        for (int j = 0; j < n_vars; ++j) out << traj[(size_t)i * n_vars + j] << ",";
        out << err[i] << "\n";
    }
}

void fastsync_grid(std::ostream& out, const double* heatmap, int n, double x_lo, double x_hi, double y_lo,
                   double y_hi) {
    out << std::setprecision(kPrecision);
    out << x_lo << " " << x_hi << "\n" << y_lo << " " << y_hi << "\n";
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) out << heatmap[(size_t)iy * n + ix] << (ix + 1 < n ? "," : "\n");
}

}  // namespace csv
}  // namespace ucuda
