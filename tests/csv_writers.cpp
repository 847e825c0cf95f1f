// Writes every CSV writer of u-cuda_amd/csrc/data_export.cpp for fixed
// inputs, one file per writer call, into the directory argv[1]. The inputs
// exercise the number formatting (thirds, tiny, huge, negative zero,
// non-finite values where a writer maps them) and every branch of the
// header lines (par_or_var, window names, attractor and grid modes).
// tests/test_csv_formats.py compares the files with the reference layout.
#include <cmath>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "data_export.hpp"

using namespace ucuda::csv;

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    auto file = [&](const char* name) { return std::ofstream(dir + "/" + name); };
    const std::vector<double> values = {10.0, 28.0, 8.0 / 3.0, -0.5, 1e-20};
    const std::vector<double> ic = {0.1, 1.0 / 3.0, 2.5e10};
    const double t_max = 100.0, transient = 50.5, h = 0.01, NT = 0.1, eps = 1e-4;

    { auto o = file("bif1d_config"); bif1d_config(o, values, ic, t_max, transient, h, 2, 1, 2, 0.0, 250.0); }
    {
        auto o = file("bif1d_rows");
        const double pk[2] = {12.345678901234567, -0.0};
        const double tm[2] = {0.7, 1e-7};
        bif1d_rows(o, param_value_at(0, 5, 0.0, 250.0), 0, pk, tm);
        bif1d_rows(o, param_value_at(3, 5, 0.0, 250.0), -1, pk, tm);
        bif1d_rows(o, param_value_at(4, 5, 0.0, 250.0), 2, pk, tm);
        bif1d_rows(o, param_value_at(0, 1, 7.5, 250.0), 0, pk, tm);
    }
    for (int w = 0; w < 3; ++w) {
        auto o = file(("dft1d_config_" + std::to_string(w)).c_str());
        dft1d_config(o, values, ic, t_max, transient, h, 1, 0, 3, -1.0, 1.0, 3, 0.5, 123456789012345678.0, w);
    }
    { auto o = file("dft1d_header"); dft1d_header(o, -1.0, 1.0, 0.5, 123456789012345678.0); }
    {
        auto o = file("dft1d_matrix");
        const double m[6] = {1.0 / 7.0, 2e-300, 3.0, -4.5e15, 0.0, 1e21};
        dft1d_matrix(o, m, 2, 3);
    }
    { auto o = file("lle1d_config"); curve1d_config(o, "1D LLE", values, ic, t_max, NT, transient, h, eps, 1, 0.05, 2.0); }
    { auto o = file("ls1d_config"); curve1d_config(o, "1D LS", values, ic, t_max, NT, transient, h, eps, 1, 0.05, 2.0); }
    {
        auto o = file("lle1d_rows");
        lle1d_row(o, param_value_at(1, 3, 0.05, 2.0), 0.0713);
        lle1d_row(o, param_value_at(2, 3, 0.05, 2.0), -1.0 / 3.0);
    }
    {
        auto o = file("ls1d_rows");
        const double ex[3] = {0.0713, 1e-12, -5.394};
        ls1d_row(o, 0.05, ex, 3);
        ls1d_row(o, 2.0, ex, 1);
    }
    for (int pv = 0; pv < 3; ++pv) {
        auto o = file(("bif2d_config_" + std::to_string(pv)).c_str());
        bif2d_config(o, pv, values, ic, t_max, transient, h, 3, 0.1, 2, 1, 2, -2.5, 2.5, 0.05, 20.0, 3);
    }
    { auto o = file("lle2d_config"); map2d_config(o, "2D LLE", 1, values, ic, t_max, NT, transient, h, eps, 2, 3, 0.05, 2.0, 1.0, 20.0, 512, 0); }
    { auto o = file("ls2d_config"); map2d_config(o, "2D LS", 2, values, ic, t_max, NT, transient, h, eps, 0, 3, -2.5, 2.5, 1.0, 20.0, 4, 3); }
    {
        auto o = file("grid");
        const double g[9] = {0.1, -1.0 / 3.0, 1e300, 7.0, 0.0, -0.0, 999.0, 2.5e-8, 1234567.0};
        grid(o, 3, g);
    }
    {
        auto o = file("ls2d_cells");
        // plane-major: values[k * 4 + cell], 3 exponents over 2 x 2 cells
        const double v[12] = {0.1, 0.2, 0.3, 0.4, -0.1, -0.2, -0.3, -0.4, 1.0 / 3.0, 2.0 / 3.0, 1e-9, 5.0};
        ls2d_cells(o, 2, 3, v);
    }
    {
        auto o = file("basins_config");
        basins_config(o, values, ic, t_max, transient, h, 1, 0.25, 0, 0, 1, -2.5, 2.5, -2.0, 2.0);
    }
    {
        auto o = file("basins_labels");
        basins_ranges(o, -2.5, 2.5, -2.0, 2.0);
        const int lab[4] = {0, 1, -1, 12};
        grid_int(o, 2, lab);
    }
    {
        auto o = file("basins_finite");
        basins_ranges(o, -2.5, 2.5, -2.0, 2.0);
        const double f[4] = {std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::infinity(),
                             -std::numeric_limits<double>::infinity(), 2.5};
        grid_finite(o, 2, f);
    }
    const std::vector<double> icm = {1.0, 2.0, 3.0}, ics = {1.5, -2.0, 1.0 / 3.0}, kf = {0.5, 0.0, 0.0},
                              kb = {0.0, 0.25, 1e-3};
    for (int mode = 0; mode < 2; ++mode) {
        auto o = file(("fastsync_config_" + std::to_string(mode)).c_str());
        fastsync_config(o, mode, values, icm, ics, kf, kb, h, 40, 2, 5.5, mode, mode == 0 ? 1 : 2, 1e-6, t_max,
                        transient, 0, 2, -20.0, 20.0, 0.0, 50.0, 128, mode == 1);
    }
    {
        auto o = file("fastsync_config_estim0");
        fastsync_config(o, 0, values, icm, ics, kf, kb, h, 40, 2, 5.5, 0, 0, 1e-6, t_max, transient, 0, 0, 0.0, 0.0,
                        0.0, 0.0, 0, false);
    }
    {
        auto o = file("fastsync_attractor");
        const double traj[6] = {1.0, 2.0, 1.0 / 3.0, -4.0, 5e-5, 6.0};
        const double err[2] = {0.125, 1e-17};
        fastsync_attractor(o, traj, err, 2, 3, {"x", "y"});
    }
    {
        auto o = file("fastsync_grid");
        const double hm[4] = {1.0, 2.0 / 3.0, 40.0, -1.0};
        fastsync_grid(o, hm, 2, -20.0, 20.0, 0.0, 50.0);
    }
    return 0;
}
