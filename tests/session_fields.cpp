// session_fields.cpp — a session layer's use of the facade, written the way
// the reference's analysis sessions build their requests
// (U-CUDA/analysis_session.cpp:477-520 bifurcation, :777-810 LLE 1-D,
// :829-870 LLE 2-D, :1298-1345 basins, :1440-1490 recluster): the same
// fields, the step body from codegen_scheme for the session's scheme, and
// the result snapshot read back. Built and run by tests/test_abi.py.
//
// With a device-less engine (argv[1] == "cpu") every run must get through
// request validation to the device stage, and the host-only recluster must
// label synthetic features; prints one JSON line.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "parametric_engine.hpp"

using namespace ucuda;

namespace {
System lorenz() { return System{{"x", "y", "z"}, {"sigma", "rho", "beta"}, {"sigma*(y-x)", "x*(rho-z)-y", "x*y-beta*z"}, false, ""}; }

template <class Req>
void common_fields(Req& req, const System& sys, Scheme scheme) {
    req.krs_body = codegen_scheme(sys, scheme);
    req.amountOfX = (int)sys.vars.size();
    req.initial_conditions.assign((size_t)req.amountOfX, 0.1);
    req.base_values.assign(sys.params.size() + 1, 0.0);
    req.base_values[0] = 0.5;  // a[0]: the CD scheme's symmetry coefficient
    req.base_values[1] = 10.0;
    req.base_values[2] = 28.0;
    req.base_values[3] = 8.0 / 3.0;
}

std::string quote(const std::string& s) {
    std::string o = "\"";
    for (char c : s) o += (c == '"' || c == '\\') ? std::string("\\") + c : (c == '\n' ? std::string(" ") : std::string(1, c));
    return o + "\"";
}
}  // namespace

int main(int argc, char** argv) {
    const bool cpu = argc > 1 && std::strcmp(argv[1], "cpu") == 0;
    ParametricEngine engine(cpu ? -1 : 0);
    const System sys = lorenz();
    auto cancel = std::make_shared<std::atomic<bool>>(false);
    auto progress = std::make_shared<std::atomic<float>>(0.0f);
    auto phase = std::make_shared<std::atomic<int>>(0);

    // bifurcation 1-D (build_bif1d_request)
    Bifurcation1DRequest b;
    common_fields(b, sys, Scheme::DOPRI78);
    b.param_index = 2;
    b.sweep_over_var = false;
    b.sweep_over_h = false;
    b.log_scale = false;
    b.var_sweep_index = 0;
    b.continuation = false;
    b.continuation_reverse = false;
    b.use_cpu = false;
    b.param_lo = 20.0;
    b.param_hi = 30.0;
    b.n_pts = 8;
    b.writable_var = 0;
    b.h = 0.01;
    b.t_max = 10.0;
    b.transient_time = 10.0;
    b.pre_scaller = 1;
    b.max_value = 1.0e6;
    b.csv_output_path = std::string{};
    b.cancel = cancel;
    b.progress = progress;
    Bifurcation1DResult br = engine.run_bifurcation_1d(b);

    // LLE 1-D (build_lle1d_request)
    LLE1DRequest l;
    l.continuation = false;
    l.continuation_reverse = false;
    l.use_cpu = false;
    common_fields(l, sys, Scheme::DOPRI78);
    l.param_index = 2;
    l.sweep_over_var = false;
    l.sweep_over_h = false;
    l.log_scale = false;
    l.var_sweep_index = 0;
    l.param_lo = 20.0;
    l.param_hi = 30.0;
    l.n_pts = 8;
    l.h = 0.01;
    l.t_max = 10.0;
    l.transient_time = 5.0;
    l.max_value = 1.0e6;
    l.NT = 0.1;
    l.eps = 1.0e-4;
    l.csv_output_path = std::string{};
    l.cancel = cancel;
    l.progress = progress;
    LLE1DResult lr = engine.run_lle_1d(l);

    // LLE 2-D (build_lle2d_request)
    LLE2DRequest l2;
    common_fields(l2, sys, Scheme::RK4);
    l2.sweep_over_var = false;
    l2.sweep_over_var_2 = false;
    l2.sweep_over_h = false;
    l2.sweep_over_h_2 = false;
    l2.log_scale = false;
    l2.log_scale_2 = false;
    l2.param_index = 2;
    l2.var_sweep_index = 0;
    l2.param_index_2 = 1;
    l2.var_sweep_index_2 = 0;
    l2.param_lo = 20.0;
    l2.param_hi = 30.0;
    l2.param_lo_2 = 8.0;
    l2.param_hi_2 = 12.0;
    l2.n_pts = 4;
    l2.h = 0.01;
    l2.t_max = 10.0;
    l2.transient_time = 5.0;
    l2.max_value = 1.0e6;
    l2.NT = 0.1;
    l2.eps = 1.0e-4;
    l2.csv_output_path = std::string{};
    l2.cancel = cancel;
    l2.progress = progress;
    LLE2DResult l2r = engine.run_lle_2d(l2);

    // basins (build_basins_request): a step body alone cannot run the
    // adaptive integrator, so the session's system goes along
    const System duffing{{"x", "v"}, {"delta", "gamma", "omega"}, {"v", "-delta*v+x-x^3+gamma*cos(omega*t)"}, false, "t"};
    BasinsRequest q;
    q.krs_body = codegen_scheme(duffing, Scheme::RK4);
    q.system = duffing;
    q.amountOfX = 2;
    q.initial_conditions = {0.0, 0.0};
    q.base_values = {0.5, 0.15, 0.3, 1.0};
    q.axis_x_var = 0;
    q.axis_y_var = 1;
    q.axis_x_lo = -2.5;
    q.axis_x_hi = 2.5;
    q.axis_y_lo = -2.5;
    q.axis_y_hi = 2.5;
    q.n_pts = 8;
    q.writable_var = 0;
    q.h = 0.01;
    q.t_max = 60.0;
    q.transient_time = 60.0;
    q.pre_scaller = 1;
    q.max_value = 1.0e6;
    q.eps_dbscan = 0.05;
    q.csv_output_path = std::string{};
    q.feature1 = 0;
    q.feature2 = 1;
    q.mult1 = 1.0;
    q.mult2 = 1.0;
    q.cancel = cancel;
    q.progress = progress;
    q.progress_phase = phase;
    BasinsResult qr = engine.run_basins(q);
    BasinsRequest q_body_only = q;
    q_body_only.system = System{};
    BasinsResult qr_body_only = engine.run_basins(q_body_only);

    // recluster from a previous result's features (the session's "Clustering"
    // button): features of three synthetic attractors, one of them split in
    // two boxes that touch, and unbound cells
    BasinsReclusterRequest rq;
    rq.krs_body = q.krs_body;
    rq.amountOfX = q.amountOfX;
    rq.n_pts = 3;
    rq.eps_dbscan = 0.1;
    rq.avg_peaks = {1.00, 1.02, 1.13, -1.0, -1.01, 999.0, 0.5, 0.5, 999.0};
    rq.avg_intervals = {0.0, 0.01, 0.0, 0.2, 0.21, 999.0, -0.3, -0.3, 999.0};
    rq.helpful_array = {1, 1, 1, 1, 1, 0, 1, 1, 0};
    rq.cancel = cancel;
    rq.progress = progress;
    BasinsReclusterResult rc = engine.run_basins_recluster(rq);
    BasinsResult merged = qr;  // what the session does with the recluster result
    merged.snapshot.eps_dbscan = rq.eps_dbscan;
    if (rc.ok) {
        merged.basin_idx = rc.basin_idx;
        merged.n_clusters = rc.n_clusters;
        merged.min_cluster_idx = rc.min_cluster_idx;
    }

    std::printf("{\"bif_error\": %s, \"lle1d_error\": %s, \"lle2d_error\": %s, \"basins_error\": %s, "
                "\"basins_body_only_error\": %s, \"recluster_ok\": %d, \"recluster_labels\": [",
                quote(br.error).c_str(), quote(lr.error).c_str(), quote(l2r.error).c_str(), quote(qr.error).c_str(),
                quote(qr_body_only.error).c_str(), rc.ok ? 1 : 0);
    for (size_t i = 0; i < rc.basin_idx.size(); ++i) std::printf("%s%d", i ? ", " : "", rc.basin_idx[i]);
    std::printf("], \"recluster_n\": %d, \"snapshot_eps\": %.17g, \"ok\": [%d, %d, %d, %d]}\n", rc.n_clusters,
                merged.snapshot.eps_dbscan, br.ok ? 1 : 0, lr.ok ? 1 : 0, l2r.ok ? 1 : 0, qr.ok ? 1 : 0);
    return 0;
}
