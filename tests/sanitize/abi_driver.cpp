// ASan/UBSan driver for the host library on a device-less engine (run by
// tests/sanitize/run.sh): systems, codegen of every scheme, step bodies,
// request validation of the entry points, axis values.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "ucuda.h"

static int fails = 0;
#define EXPECT(cond) do { if (!(cond)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); ++fails; } } while (0)

struct Sys { std::vector<const char*> v, p, r; const char* t; int latex; };

int main() {
    uc_engine* e = nullptr;
    EXPECT(uc_engine_create(-1, &e) == UC_OK);
    std::printf("version %s\n", uc_version());
    std::vector<Sys> systems = {
        {{"x", "y", "z"}, {"sigma", "rho", "beta"}, {"sigma*(y-x)", "x*(rho-z)-y", "x*y-beta*z"}, nullptr, 0},
        {{"x", "y", "z"}, {"a", "b", "c"}, {"-y-z", "x+a*y", "b+z*(x-c)"}, nullptr, 0},
        {{"x", "v"}, {"delta", "gamma", "omega"}, {"v", "-delta*v+x-x^3+gamma*cos(omega*t)"}, "t", 0},
        {{"x", "y", "z", "w"}, {"a", "d"}, {"-y-z", "x+a*y+w", "3+x*z", "-0.5*z+d*w"}, nullptr, 0},
        {{"x", "y"}, {"a"}, {"\\frac{y}{2}", "-a \\cdot \\sin{x}"}, nullptr, 1},
        // harmonic forcing terms of every linear form, and terms that stay calls
        {{"x", "v"}, {"d", "g", "w", "f"},
         {"v", "-d*v - x^3 + g*cos(w*t) + g*sin(2*w*t + f) - cos(t/w) + sin(x*t) + cos(w)*t + sin(-(w*t) - f)"
               " + cos(f - w*t)"}, "t", 0},
    };
    std::vector<uc_system*> made;
    for (auto& S : systems) {
        uc_system_desc d{S.v.data(), (int)S.v.size(), S.p.data(), (int)S.p.size(), S.r.data(), S.latex, S.t};
        uc_system* s = nullptr;
        int rc = uc_system_create(e, &d, &s);
        EXPECT(rc == UC_OK);
        if (rc) { std::printf("  %s\n", uc_last_error(e)); continue; }
        made.push_back(s);
        for (int scheme = 0; scheme <= 5; ++scheme) {
            size_t len = 0;
            rc = uc_codegen_scheme(e, &d, scheme, nullptr, 0, &len);
            if (rc != UC_OK) { std::printf("  scheme %d: %s\n", scheme, uc_last_error(e)); continue; }
            std::vector<char> buf(len + 1);
            EXPECT(uc_codegen_scheme(e, &d, scheme, buf.data(), buf.size(), &len) == UC_OK);
            // the step body as a custom system
            uc_system* k = nullptr;
            rc = uc_system_create_krs(e, buf.data(), (int)S.v.size(), (int)S.p.size() + 1, &k);
            if (rc == UC_OK) uc_system_destroy(k);
        }
        for (int dialect = 0; dialect <= 1; ++dialect) {
            size_t len = 0;
            EXPECT(uc_system_source(s, dialect, nullptr, 0, &len) == UC_OK);
            std::vector<char> buf(len + 1);
            EXPECT(uc_system_source(s, dialect, buf.data(), buf.size(), &len) == UC_OK);
        }
        int rf = 0, rcalls = 0, nz = 0, jf = 0;
        EXPECT(uc_system_op_counts(s, &rf, &rcalls, &nz, &jf) == UC_OK);
        std::vector<double> X(S.v.size(), 0.3), a(S.p.size() + 1, 0.7), dX(S.v.size());
        EXPECT(uc_system_eval_rhs(s, X.data(), a.data(), 0.1, dX.data()) == UC_OK);
    }
    // malformed systems and step bodies
    {
        const char* v[] = {"x"}; const char* p[] = {"a"}; const char* r[] = {"a*(x"};
        uc_system_desc d{v, 1, p, 1, r, 0, nullptr};
        uc_system* s = nullptr;
        EXPECT(uc_system_create(e, &d, &s) != UC_OK);
        std::printf("  parse error: %s\n", uc_last_error(e));
        uc_system* k = nullptr;
        EXPECT(uc_system_create_krs(e, "X[0] = X[5] + a[9];", 1, 2, &k) != UC_OK);
        std::printf("  krs error: %s\n", uc_last_error(e));
    }
    // requests on a device-less engine: valid ones reach the device stage,
    // invalid ones are rejected first
    uc_system* s = made.empty() ? nullptr : made[0];
    if (s) {
        double ic[3] = {0.1, 0.1, 0.1}, vals[4] = {0.5, 10.0, 28.0, 8.0 / 3.0};
        uc_common c{};
        c.initial_conditions = ic; c.base_values = vals; c.n_values = 4; c.n_axes = 2;
        c.axis[0] = uc_axis{UC_AXIS_PARAM, 2, 16, 20.0, 30.0, 0};
        c.axis[1] = uc_axis{UC_AXIS_PARAM, 1, 8, 5.0, 15.0, 1};
        c.h = 0.01; c.transient_time = 1.0; c.t_max = 2.0; c.max_value = 1e6;
        std::vector<double> ax(16);
        EXPECT(uc_axis_values(&c.axis[0], ax.data()) == UC_OK);
        EXPECT(uc_grid_cells(&c) == 128);
        double dummy[4096]; int iflag[4096]; long long steps[512];
        uc_bif_request br{}; br.c = c; br.max_peaks = 8; br.pre_scaller = 1;
        uc_bif_output bo{iflag, dummy, dummy, nullptr, 0, nullptr};
        int rc = uc_run_bif(e, s, &br, &bo, nullptr);
        std::printf("  bif rc %d: %s\n", rc, uc_last_error(e));
        uc_lyap_request lr{}; lr.c = c; lr.NT = 0.1; lr.eps = 1e-4; lr.probe_seed = 1234567891ULL;
        uc_lyap_output lo{dummy, iflag, nullptr};
        rc = uc_run_lle(e, s, &lr, &lo, nullptr); std::printf("  lle rc %d: %s\n", rc, uc_last_error(e));
        rc = uc_run_ls(e, s, &lr, &lo, nullptr); std::printf("  ls rc %d: %s\n", rc, uc_last_error(e));
        rc = uc_run_lle_noalloc(e, s, &lr, &lo, nullptr); std::printf("  noalloc rc %d: %s\n", rc, uc_last_error(e));
        lr.c.cell_begin = 100; lr.c.cell_count = 50;
        rc = uc_run_lle(e, s, &lr, &lo, nullptr); EXPECT(rc != UC_OK);
        std::printf("  range rc %d: %s\n", rc, uc_last_error(e));
        lr.c.cell_begin = 0; lr.c.cell_count = 0;
        uc_basins_request q{}; q.c = c;
        q.c.axis[0] = uc_axis{UC_AXIS_VAR, 0, 8, -2.0, 2.0, 0}; q.c.axis[1] = uc_axis{UC_AXIS_VAR, 1, 8, -2.0, 2.0, 0};
        q.rtol = q.atol = 1e-9; q.h_min = 1e-12; q.h_max = 1.0; q.section_period = 6.283185307179586;
        q.n_sections = 20; q.n_record = 16; q.max_period = 8; q.period_tol = 1e-4; q.quantum = 1e-3;
        int na = 0, aper[8]; double apt[16];
        uc_basins_output qo{iflag, iflag + 1024, nullptr, steps, 8, &na, aper, apt};
        rc = uc_run_basins(e, s, &q, &qo, nullptr); std::printf("  basins rc %d: %s\n", rc, uc_last_error(e));
        q.max_period = 20; rc = uc_run_basins(e, s, &q, &qo, nullptr); EXPECT(rc == UC_ERR_ARGUMENT);
        std::printf("  basins bad rc %d: %s\n", rc, uc_last_error(e));
        // invalid common fields
        uc_lyap_request bad = lr; bad.c.n_axes = 3; EXPECT(uc_run_lle(e, s, &bad, &lo, nullptr) == UC_ERR_ARGUMENT);
        bad = lr; bad.c.h = -1.0; rc = uc_run_lle(e, s, &bad, &lo, nullptr); std::printf("  h<0 rc %d: %s\n", rc, uc_last_error(e));
        bad = lr; bad.c.axis[0].index = 99; EXPECT(uc_run_lle(e, s, &bad, &lo, nullptr) == UC_ERR_ARGUMENT);
        bad = lr; bad.c.n_values = 1; EXPECT(uc_run_lle(e, s, &bad, &lo, nullptr) == UC_ERR_ARGUMENT);
        uc_bif_request bb = br; bb.max_peaks = -1; EXPECT(uc_run_bif(e, s, &bb, &bo, nullptr) == UC_ERR_ARGUMENT);
        // batch / server need a device
        uc_batch* b = nullptr; rc = uc_batch_create(e, &b); std::printf("  batch rc %d\n", rc);
        if (rc == UC_OK) uc_batch_destroy(b);
        uc_server* srv = nullptr; rc = uc_server_create(e, nullptr, &srv); std::printf("  server rc %d\n", rc);
        if (rc == UC_OK) uc_server_destroy(srv);
    }
    for (auto* m : made) uc_system_destroy(m);
    uc_engine_destroy(e);
    std::printf("%s (%d failures)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
