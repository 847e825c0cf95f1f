// facade_demo.cpp — drives the C++ facade the way the reference's session
// layer drives ParametricEngine (build a Request, run, read the Result),
// and prints a JSON summary. Used by the GPU tests to check that the facade
// and the C ABI give the same numbers. With a directory argument every run
// also writes the reference's CSV files there (csv_output_path).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "parametric_engine.hpp"

using namespace ucuda;

namespace {
// A 1024 x 1024 Rossler LLE map runs in four pieces; a watcher sets the
// cancel token once the first piece has reported progress. The run must
// come back cancelled, with progress short of 1.
int cancel_demo() {
    ParametricEngine eng(0);
    System rossler{{"x", "y", "z"}, {"a", "b", "c"}, {"-y-z", "x+a*y", "b+z*(x-c)"}, false, ""};
    LLE2DRequest l;
    l.system = rossler;
    l.amountOfX = 3;
    l.initial_conditions = {1.0, 1.0, 1.0};
    l.base_values = {0.5, 0.2, 0.2, 5.7};
    l.param_index = 2;
    l.param_lo = 0.05;
    l.param_hi = 2.0;
    l.param_index_2 = 3;
    l.param_lo_2 = 1.0;
    l.param_hi_2 = 20.0;
    l.n_pts = 1024;
    l.h = 0.01;
    l.transient_time = 5.0;
    l.t_max = 300.0;
    l.NT = 0.1;
    l.cancel = std::make_shared<std::atomic<bool>>(false);
    l.progress = std::make_shared<std::atomic<float>>(0.0f);
    float seen = 0.0f;
    std::thread watcher([&] {
        const auto t0 = std::chrono::steady_clock::now();
        while (std::chrono::steady_clock::now() - t0 < std::chrono::seconds(60)) {
            seen = l.progress->load();
            if (seen > 0.0f) { l.cancel->store(true); return; }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    });
    LLE2DResult r = eng.run_lle_2d(l);
    watcher.join();
    std::printf("{\"cancelled\": %d, \"ok\": %d, \"progress\": %.6f, \"seen\": %.6f, \"error\": \"%s\"}\n",
                r.cancelled ? 1 : 0, r.ok ? 1 : 0, l.progress->load(), seen, r.error.c_str());
    return 0;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "--cancel") == 0) return cancel_demo();
    const std::string dir = argc > 1 ? std::string(argv[1]) + "/" : std::string();
    auto csv = [&](const char* name) { return dir.empty() ? std::string() : dir + name; };
    ParametricEngine eng(0);
    System lorenz{{"x", "y", "z"}, {"sigma", "rho", "beta"}, {"sigma*(y-x)", "x*(rho-z)-y", "x*y-beta*z"}, false, ""};
    Bifurcation1DRequest b;
    b.system = lorenz;
    b.amountOfX = 3;
    b.initial_conditions = {0.1, 0.1, 0.1};
    b.base_values = {0.5, 10.0, 28.0, 8.0 / 3.0};
    b.param_index = 2;
    b.param_lo = 0.0;
    b.param_hi = 250.0;
    b.n_pts = 16;
    b.h = 0.01;
    b.transient_time = 20.0;
    b.t_max = 20.0;
    b.max_peaks = 64;
    b.csv_output_path = csv("bif1d.csv");
    Bifurcation1DResult br = eng.run_bifurcation_1d(b);
    if (!br.ok) { std::printf("{\"error\": \"%s\"}\n", br.error.c_str()); return 1; }

    System rossler{{"x", "y", "z"}, {"a", "b", "c"}, {"-y-z", "x+a*y", "b+z*(x-c)"}, false, ""};
    LLE2DRequest l;
    l.system = rossler;
    l.amountOfX = 3;
    l.initial_conditions = {1.0, 1.0, 1.0};
    l.base_values = {0.5, 0.2, 0.2, 5.7};
    l.param_index = 2;
    l.param_lo = 0.05;
    l.param_hi = 2.0;
    l.param_index_2 = 3;
    l.param_lo_2 = 1.0;
    l.param_hi_2 = 20.0;
    l.n_pts = 8;
    l.h = 0.01;
    l.transient_time = 5.0;
    l.t_max = 50.0;
    l.NT = 0.1;
    l.csv_output_path = csv("lle2d.csv");
    LLE2DResult lr = eng.run_lle_2d(l);
    if (!lr.ok) { std::printf("{\"error\": \"%s\"}\n", lr.error.c_str()); return 1; }

    LS2DRequest ls = l;
    ls.n_pts = 4;
    ls.csv_output_path = csv("ls2d.csv");
    LS2DResult lsr = eng.run_ls_2d(ls);
    if (!lsr.ok) { std::printf("{\"error\": \"%s\"}\n", lsr.error.c_str()); return 1; }

    Bifurcation2DRequest b2;
    b2.system = lorenz;
    b2.amountOfX = 3;
    b2.initial_conditions = {0.1, 0.1, 0.1};
    b2.base_values = {0.5, 10.0, 28.0, 8.0 / 3.0};
    b2.param_index = 2;
    b2.param_lo = 20.0;
    b2.param_hi = 250.0;
    b2.param_index_2 = 1;
    b2.param_lo_2 = 5.0;
    b2.param_hi_2 = 16.0;
    b2.n_pts = 12;
    b2.h = 0.01;
    b2.transient_time = 30.0;
    b2.t_max = 30.0;
    b2.eps_dbscan = 0.1;
    b2.csv_output_path = csv("bif2d.csv");
    Bifurcation2DResult b2r = eng.run_bifurcation_2d(b2);
    if (!b2r.ok) { std::printf("{\"error\": \"%s\"}\n", b2r.error.c_str()); return 1; }

    Dft1DRequest dq;
    dq.system = lorenz;
    dq.amountOfX = 3;
    dq.initial_conditions = {0.1, 0.1, 0.1};
    dq.base_values = {0.5, 10.0, 28.0, 8.0 / 3.0};
    dq.param_index = 2;
    dq.param_lo = 0.5;
    dq.param_hi = 250.0;
    dq.n_pts = 16;
    dq.h = 0.01;
    dq.transient_time = 10.0;
    dq.t_max = 20.0;
    dq.n_freq = 32;
    dq.freq_lo = 0.0;
    dq.freq_hi = 5.0;
    dq.csv_output_path = csv("dft1d");
    Dft1DResult dr = eng.run_dft_1d(dq);
    if (!dr.ok) { std::printf("{\"error\": \"%s\"}\n", dr.error.c_str()); return 1; }

    System duffing{{"x", "v"}, {"delta", "gamma", "omega"}, {"v", "-delta*v+x-x^3+gamma*cos(omega*t)"}, false, "t"};
    BasinsRequest bq;
    bq.system = duffing;
    bq.amountOfX = 2;
    bq.initial_conditions = {0.0, 0.0};
    bq.base_values = {0.5, 0.15, 0.3, 1.0};
    bq.n_pts = 16;
    bq.axis_x_lo = bq.axis_y_lo = -2.5;
    bq.axis_x_hi = bq.axis_y_hi = 2.5;
    bq.n_sections = 40;
    bq.csv_output_path = csv("basins.csv");
    BasinsResult bres = eng.run_basins(bq);
    if (!bres.ok) { std::printf("{\"error\": \"%s\"}\n", bres.error.c_str()); return 1; }

    FastSyncRequest fq;
    fq.system = lorenz;
    fq.amountOfX = 3;
    fq.values = {0.5, 10.0, 28.0, 8.0 / 3.0};
    fq.mode = 1;
    fq.h = 0.01;
    fq.iter_of_synchr = 3;
    fq.k_forward = {1.0, 0.0, 0.0};
    fq.k_backward = {15.0, 15.0, 15.0};
    fq.ic_master = {0.1, 0.1, 0.1};
    fq.ic_slave = {3.0, -2.0, 20.0};
    fq.transient_time = 1.0;
    fq.window = 1.2;
    fq.axis_x_var = 0;
    fq.axis_y_var = 2;
    fq.axis_x_lo = -15.0; fq.axis_x_hi = 15.0;
    fq.axis_y_lo = 5.0; fq.axis_y_hi = 40.0;
    fq.n_pts = 10;
    fq.csv_output_path = csv("fastsync_grid.csv");
    FastSyncResult fr = eng.run_fastsync(fq);
    if (!fr.ok) { std::printf("{\"error\": \"%s\"}\n", fr.error.c_str()); return 1; }
    fq.mode = 0;
    fq.t_max = 2.0;
    fq.pre_scaller = 4;
    fq.window = 1.0;
    fq.var_names = {"x", "y", "z"};
    fq.csv_output_path = csv("fastsync_attr.csv");
    FastSyncResult fa = eng.run_fastsync(fq);
    if (!fa.ok) { std::printf("{\"error\": \"%s\"}\n", fa.error.c_str()); return 1; }

    System fold{{"x"}, {"r"}, {"r+x-x^3"}, false, ""};
    LLE1DRequest cq;
    cq.system = fold;
    cq.amountOfX = 1;
    cq.initial_conditions = {0.0};
    cq.base_values = {0.5, 0.0};
    cq.param_index = 1;
    cq.param_lo = -1.0;
    cq.param_hi = 1.0;
    cq.n_pts = 9;
    cq.h = 0.01;
    cq.transient_time = 20.0;
    cq.t_max = 20.0;
    cq.NT = 0.1;
    cq.continuation = true;
    LLE1DResult cf = eng.run_lle_1d(cq);
    cq.continuation_reverse = true;
    LLE1DResult cr = eng.run_lle_1d(cq);
    if (!cf.ok || !cr.ok) { std::printf("{\"error\": \"%s\"}\n", (cf.ok ? cr : cf).error.c_str()); return 1; }

    // the session's step body alone (DOPRI78 text, no right-hand side): the
    // body itself is stepped, Lyapunov by the reference's finite differences
    Bifurcation1DRequest kb;
    kb.krs_body = codegen_scheme(lorenz, Scheme::DOPRI78);
    kb.amountOfX = 3;
    kb.initial_conditions = {0.1, 0.1, 0.1};
    kb.base_values = {0.5, 10.0, 28.0, 8.0 / 3.0};
    kb.param_index = 2;
    kb.param_lo = 5.0;
    kb.param_hi = 20.0;
    kb.n_pts = 16;
    kb.h = 0.01;
    kb.transient_time = 2.0;
    kb.t_max = 10.0;
    kb.max_peaks = 32;
    Bifurcation1DResult kbr = eng.run_bifurcation_1d(kb);
    LLE1DRequest kl;
    kl.krs_body = kb.krs_body;
    kl.amountOfX = 3;
    kl.initial_conditions = {0.1, 0.1, 0.1};
    kl.base_values = kb.base_values;
    kl.param_index = 2;
    kl.param_lo = 5.0;
    kl.param_hi = 20.0;
    kl.n_pts = 16;
    kl.h = 0.01;
    kl.transient_time = 2.0;
    kl.t_max = 10.0;
    kl.NT = 0.1;
    LLE1DResult klr = eng.run_lle_1d(kl);
    if (!kbr.ok || !klr.ok) { std::printf("{\"error\": \"%s\"}\n", (kbr.ok ? klr.error : kbr.error).c_str()); return 1; }

    std::printf("{\"krs_bif_flags\": [");
    for (size_t i = 0; i < kbr.flags.size(); ++i) std::printf("%s%d", i ? ", " : "", kbr.flags[i]);
    std::printf("], \"krs_bif_peaks\": [");
    for (size_t i = 0; i < kbr.flags.size(); ++i) {
        std::printf("%s[", i ? ", " : "");
        for (size_t k = 0; k < kbr.bifurcation_points[i].size(); ++k)
            std::printf("%s%.17g", k ? ", " : "", kbr.bifurcation_points[i][k]);
        std::printf("]");
    }
    std::printf("], \"krs_lle\": [");
    for (size_t i = 0; i < klr.lyapunov.size(); ++i) std::printf("%s%.17g", i ? ", " : "", klr.lyapunov[i]);
    std::printf("], \"krs_snapshot_index\": %d, ", kbr.snapshot.indexOfMutVar);
    std::printf("\"bif_flags\": [");
    for (size_t i = 0; i < br.flags.size(); ++i) std::printf("%s%d", i ? ", " : "", br.flags[i]);
    std::printf("], \"bif_first_peaks\": [");
    for (size_t i = 0; i < br.flags.size(); ++i)
        std::printf("%s%.17g", i ? ", " : "", br.bifurcation_points[i].empty() ? 0.0 : br.bifurcation_points[i][0]);
    std::printf("], \"lle2d\": [");
    for (size_t i = 0; i < lr.values.size(); ++i) std::printf("%s%.17g", i ? ", " : "", std::isnan(lr.values[i]) ? 999.0 : lr.values[i]);
    std::printf("], \"ls2d_sum\": [");
    const size_t n = lsr.flags.size();
    for (size_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < lsr.n_exponents; ++k) s += lsr.values[(size_t)k * n + i];
        std::printf("%s%.17g", i ? ", " : "", s);
    }
    std::printf("], \"bif2d\": [");
    for (size_t i = 0; i < b2r.values.size(); ++i) std::printf("%s%.17g", i ? ", " : "", b2r.values[i]);
    std::printf("], \"dft_flags\": [");
    for (size_t i = 0; i < dr.flags.size(); ++i) std::printf("%s%d", i ? ", " : "", dr.flags[i]);
    std::printf("], \"basins\": [");
    for (size_t i = 0; i < bres.basin_idx.size(); ++i) std::printf("%s%d", i ? ", " : "", bres.basin_idx[i]);
    std::printf("], \"fs_grid\": [");
    for (size_t i = 0; i < fr.heatmap.size(); ++i) std::printf("%s%.17g", i ? ", " : "", fr.heatmap[i]);
    std::printf("], \"fs_attr\": [");
    for (size_t i = 0; i < fa.sync_error.size(); ++i) std::printf("%s%.17g", i ? ", " : "", fa.sync_error[i]);
    std::printf("], \"cont_fwd\": [");
    for (size_t i = 0; i < cf.lyapunov.size(); ++i) std::printf("%s%.17g", i ? ", " : "", cf.lyapunov[i]);
    std::printf("], \"cont_rev\": [");
    for (size_t i = 0; i < cr.lyapunov.size(); ++i) std::printf("%s%.17g", i ? ", " : "", cr.lyapunov[i]);
    std::printf("], \"cont_rev_flag\": %d", cr.continuation_reverse ? 1 : 0);
    std::printf(", \"lle_min\": %.17g, \"lle_max\": %.17g}\n", lr.min_val, lr.max_val);
    return 0;
}
