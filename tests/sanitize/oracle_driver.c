/* ASan/UBSan driver for the oracle (run by tests/sanitize/run.sh): every
 * analysis on small inputs. The basins entry points are the ones oracle.py
 * binds (declared here, not in the header). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ucuda_oracle.h"

typedef struct oc_basins_job {
    oc_job base;
    double rtol, atol, h_min, h_max, section_period, period_tol;
    int n_sections, n_record, max_period, label_vars;
    long long max_attempts;
} oc_basins_job;
void oc_basins_cells(oc_rhs_fn rhs, const oc_basins_job* bj, const long long* cells, long long n_sel, int* period,
                     double* point, long long* steps, int threads);
int oc_section_signature(const double* buf, int L, int D, int max_period, double tol, double* point);
double oc_dp5_step(oc_rhs_fn rhs, const double* x, const double* a, double t, double h, int n, double rtol,
                   double atol, double* x5, double* err);

static void lor_rhs(const double* X, const double* a, double t, double* d) {
    (void)t;
    d[0] = a[1] * (X[1] - X[0]); d[1] = X[0] * (a[2] - X[2]) - X[1]; d[2] = X[0] * X[1] - a[3] * X[2];
}
static void lor_tan(const double* X, const double* a, double t, const double* V, double* JV, int K) {
    (void)t;
    for (int k = 0; k < K; ++k) {
        const double* v = V + 3 * k; double* o = JV + 3 * k;
        o[0] = -a[1] * v[0] + a[1] * v[1];
        o[1] = (a[2] - X[2]) * v[0] - v[1] - X[0] * v[2];
        o[2] = X[1] * v[0] + X[0] * v[1] - a[3] * v[2];
    }
}
static void lor_step(double* X, const double* a, double h) { oc_rk4(lor_rhs, X, a, h, 3); }
static void duf_rhs(const double* X, const double* a, double t, double* d) {
    d[0] = X[1]; d[1] = -a[1] * X[1] + X[0] - X[0] * X[0] * X[0] + a[2] * cos(a[3] * t);
}

int main(void) {
    double ic[3] = {0.1, 0.1, 0.1}, vals[4] = {0.5, 10.0, 28.0, 8.0 / 3.0};
    oc_job j;
    memset(&j, 0, sizeof j);
    j.n_vars = 3; j.n_values = 4; j.ic = ic; j.params = vals; j.n_axes = 2;
    j.axis_kind[0] = OC_AXIS_PARAM; j.axis_index[0] = 2; j.axis_n[0] = 12; j.axis_lo[0] = 10; j.axis_hi[0] = 200;
    j.axis_kind[1] = OC_AXIS_PARAM; j.axis_index[1] = 1; j.axis_n[1] = 3; j.axis_lo[1] = 6; j.axis_hi[1] = 14;
    j.h = 0.01; j.transient_time = 2.0; j.t_max = 5.0; j.max_value = 1e6;
    j.writable_var = 0; j.pre_scaller = 1; j.max_peaks = 64; j.NT = 0.1; j.eps = 1e-4; j.probe_seed = 1234567891ULL;
    const long long n = 36;
    long long cells[36];
    for (int i = 0; i < n; ++i) cells[i] = i;
    int* flags = calloc(n, sizeof(int));
    double* peaks = calloc((size_t)n * 64, sizeof(double));
    double* intervals = calloc((size_t)n * 64, sizeof(double));
    double* traj = calloc((size_t)n * 32, sizeof(double));
    oc_bif_cells(lor_step, &j, cells, n, flags, peaks, intervals, traj, 32, 4);
    oc_bif_cells_noisy(lor_step, &j, cells, n, flags, peaks, intervals, NULL, 0, 2, 1e-7, 7, 100);
    double* out = calloc((size_t)n * 3, sizeof(double));
    oc_lle_fd_cells(lor_step, &j, cells, n, out, flags, 4);
    oc_ls_fd_cells(lor_step, &j, cells, n, out, flags, 4);
    oc_lyap_var_cells(lor_rhs, lor_tan, 1, &j, cells, n, out, flags, 4);
    oc_lyap_var_cells(lor_rhs, lor_tan, 3, &j, cells, n, out, flags, 4);
    oc_lyap_var_cells_noisy(lor_rhs, lor_tan, 3, &j, cells, n, out, flags, 2, 1e-7, 11);
    /* DFT over the cells */
    double* ak = calloc((size_t)n * 40, sizeof(double));
    double* bk = calloc((size_t)n * 40, sizeof(double));
    oc_dft_cells(lor_step, &j, cells, n, 40, 0.0, 10.0, 0, 1, ak, bk, flags, 4);
    /* peak finder + DBSCAN on a stored series */
    double series[500], pk[500], tm[500];
    for (int i = 0; i < 500; ++i) series[i] = sin(0.1 * i) + 0.3 * sin(0.37 * i);
    int np = oc_peak_finder(series, 500, pk, tm, 0.01);
    int* lab = calloc(500, sizeof(int));
    int* slot = calloc(500, sizeof(int));
    int per = oc_dbscan_count(pk, np, 0.1, lab, slot);
    /* continuation chains along axis 0 (3 chains) */
    double* cv = calloc((size_t)n * 3, sizeof(double));
    int* cf = calloc(n, sizeof(int));
    oc_cont_lyap(lor_step, &j, 0, 1, cv, cf, 2);
    oc_cont_lyap(lor_step, &j, 1, 3, cv, cf, 2);
    int* blens = calloc(n, sizeof(int));
    oc_cont_bif(lor_step, &j, 0, 0, cf, peaks, intervals, NULL, blens, 2);
    /* fast synchronisation on a grid */
    double kf[3] = {1, 0, 0}, kb[3] = {15, 15, 15}, m0[3] = {0.1, 0.1, 0.1}, s0[3] = {3, -2, 20};
    oc_fs f;
    memset(&f, 0, sizeof f);
    f.n_vars = 3; f.a = vals; f.h = 0.01; f.iter_of_synchr = 3; f.pre_scaller = 1; f.type_of_synch = 0;
    f.error_estim = 2; f.fs_error_trs = 1e-3; f.k_forward = kf; f.k_backward = kb;
    double fso[16];
    long long fc[16];
    for (int i = 0; i < 16; ++i) fc[i] = i;
    oc_fs_grid(lor_step, &f, m0, s0, 0.5, 1.0, 4, 0, 2, -15, 15, 5, 40, 0, fc, 16, fso, 2);
    /* basins: forced Duffing 6 x 6 ICs, 30 sections */
    double dic[2] = {0, 0}, dv[4] = {0, 0.15, 0.3, 1.0};
    oc_basins_job bj;
    memset(&bj, 0, sizeof bj);
    bj.base.n_vars = 2; bj.base.n_values = 4; bj.base.ic = dic; bj.base.params = dv; bj.base.n_axes = 2;
    bj.base.axis_kind[0] = OC_AXIS_VAR; bj.base.axis_index[0] = 0; bj.base.axis_n[0] = 6;
    bj.base.axis_lo[0] = -2.5; bj.base.axis_hi[0] = 2.5;
    bj.base.axis_kind[1] = OC_AXIS_VAR; bj.base.axis_index[1] = 1; bj.base.axis_n[1] = 6;
    bj.base.axis_lo[1] = -2.5; bj.base.axis_hi[1] = 2.5;
    bj.base.h = 0.01; bj.base.max_value = 1e6;
    bj.rtol = bj.atol = 1e-9; bj.h_min = 1e-12; bj.h_max = 1.0; bj.section_period = 2 * M_PI;
    bj.period_tol = 1e-4; bj.n_sections = 30; bj.n_record = 16; bj.max_period = 8; bj.label_vars = 2;
    bj.max_attempts = 100000000LL;
    int bper[36];
    double bpt[72];
    long long bst[72];
    oc_basins_cells(duf_rhs, &bj, cells, 36, bper, bpt, bst, 4);
    double sigbuf[16 * 4], spt[4];
    for (int i = 0; i < 64; ++i) sigbuf[i] = (i / 4) % 3 == 0 ? 1.0 : 2.0;
    int sp = oc_section_signature(sigbuf, 16, 2, 8, 1e-4, spt);
    double x5[2], err[2], x0[2] = {0.5, 0.1};
    double en = oc_dp5_step(duf_rhs, x0, dv, 0.0, 0.05, 2, 1e-9, 1e-9, x5, err);
    printf("bif flag0 %d, lle %g, dft %g, peaks %d period %d, cont %g, fs %g, basins p0 %d, sig %d, dp5 %g\n",
           flags[0], out[0], ak[0], np, per, cv[0], fso[0], bper[0], sp, en);
    free(flags); free(peaks); free(intervals); free(traj); free(out); free(ak); free(bk); free(lab); free(slot);
    free(cv); free(cf); free(blens);
    printf("ok\n");
    return 0;
}
