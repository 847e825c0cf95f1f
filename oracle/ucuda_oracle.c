/* ucuda_oracle.c — FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see the header).
 *
 * Every function cites the reference code it restates. Where the
 * reference's CPU twin and GPU kernel disagree, the GPU kernel's semantics
 * are used (the 2-D maps have no CPU twin in the reference), and the choice
 * is stated at the function.
 */
#include "ucuda_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OC_CHECK_INTERVAL 100        /* configCUDA.h:12 */
#define OC_EPS_FIXED_POINT 1e-8      /* configCUDA.h:84 */
#define OC_EPS_PEAK_DELTA 1e-14      /* configCUDA.h:85 */
#define OC_PEAK_THRESHOLD (-1e25)    /* configCUDA.h:86 */
#define OC_MAX_PEAKS 2500            /* configCUDA.h:88 */
#define OC_PI 3.1415926535897932384626433832795
#define OC_E 2.7182818284590452353602874713527
#define OC_MAXN 32
#define OC_MAXV 64

typedef unsigned long long u64;

/* ---------------------------------------------------------------- RNG ----
 * cpu_curand_init / cpu_curand_uniform, U-CUDA/parametric_engine.cpp:196-209
 * (same constants as the kernel stub, U-CUDA/kernels/lle1d.template.cu:39-57). */
void oc_rng_init(u64 seed, u64 seq, u64 off, u64* state) {
    u64 z = seed + seq * 0x9E3779B97F4A7C15ULL + off;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    *state = z ^ (z >> 31);
}

float oc_rng_uniform(u64* state) {
    *state = *state * 6364136223846793005ULL + 1442695040888963407ULL;
    return (float)(((*state >> 40) & 0xFFFFFFULL) + 1ULL) / 16777216.0f;
}

/* ---------------------------------------------------------------- grid ---
 * getValueByIdx / getValueByIdx_log, U-CUDA/cudaLibrary.cu:1344-1396 (the
 * kernel formula; the CPU twin's lo + (hi-lo)*t differs in rounding and in
 * the n == 1 case, parametric_engine.cpp:177-186). */
void oc_axis_values(int log_scale, int n, double lo, double hi, double* out) {
    for (int i = 0; i < n; ++i) {
        if (log_scale) {
            if (n == 1) { out[i] = pow(10.0, log10(lo)); continue; }
            const double l0 = log10(lo), l1 = log10(hi);
            out[i] = pow(10.0, l0 + (double)i * ((l1 - l0) / (double)(n - 1)));
        } else {
            if (n == 1) { out[i] = hi; continue; }
            const double t = (double)i / (double)(n - 1);
            out[i] = (1.0 - t) * lo + t * hi;
        }
    }
}

static double axis_at(const oc_job* j, int ax, long long idx) {
    double v;
    /* one value only: cheap to recompute, identical formula */
    const int n = j->axis_n[ax];
    if (j->axis_log[ax]) {
        if (n == 1) return pow(10.0, log10(j->axis_lo[ax]));
        const double l0 = log10(j->axis_lo[ax]), l1 = log10(j->axis_hi[ax]);
        v = pow(10.0, l0 + (double)idx * ((l1 - l0) / (double)(n - 1)));
    } else {
        if (n == 1) return j->axis_hi[ax];
        const double t = (double)idx / (double)(n - 1);
        v = (1.0 - t) * j->axis_lo[ax] + t * j->axis_hi[ax];
    }
    return v;
}

/* Cell -> parameters / initial state / step, as calculateDiscreteModelCUDA
 * does (U-CUDA/cudaLibrary.cu:990-1058): axis 0 index = cell % n0, axis 1
 * index = (cell / n0) % n1. Optional per-cell random initial state keyed by
 * (ic_seed, cell). Returns 0 when the step is not positive. */
int oc_cell_setup(const oc_job* j, long long cell, double* a, double* x0, double* h) {
    for (int i = 0; i < j->n_values; ++i) a[i] = j->params[i];
    if (j->ic_seed) {
        u64 st;
        oc_rng_init(j->ic_seed, (u64)cell, 0ULL, &st);
        for (int i = 0; i < j->n_vars; ++i) x0[i] = j->ic_lo + (j->ic_hi - j->ic_lo) * (double)oc_rng_uniform(&st);
    } else {
        for (int i = 0; i < j->n_vars; ++i) x0[i] = j->ic[i];
    }
    *h = j->h;
    for (int ax = 0; ax < j->n_axes; ++ax) {
        const long long idx = ax == 0 ? cell % j->axis_n[0] : (cell / j->axis_n[0]) % j->axis_n[1];
        const double v = axis_at(j, ax, idx);
        if (j->axis_kind[ax] == OC_AXIS_PARAM) a[j->axis_index[ax]] = v;
        else if (j->axis_kind[ax] == OC_AXIS_VAR) x0[j->axis_index[ax]] = v;
        else if (j->axis_kind[ax] == OC_AXIS_STEP) *h = v;
    }
    return *h > 0.0;
}

static int sweeps_step(const oc_job* j) {
    for (int ax = 0; ax < j->n_axes; ++ax)
        if (j->axis_kind[ax] == OC_AXIS_STEP) return 1;
    return 0;
}

static double worst_step(const oc_job* j) {
    for (int ax = 0; ax < j->n_axes; ++ax)
        if (j->axis_kind[ax] == OC_AXIS_STEP) return j->axis_lo[ax] < j->axis_hi[ax] ? j->axis_lo[ax] : j->axis_hi[ax];
    return j->h;
}

/* ----------------------------------------------------------- step loop ---
 * cpu_loop_model, U-CUDA/parametric_engine.cpp:59-101 (≡ the kernel's
 * loopCalculateDiscreteModel_int, cudaLibrary.cu:782-895). A generic
 * "advance" callback lets the same loop drive a compiled KRS step, the
 * engine-order RK4 on a right-hand side, or state + tangent vectors. */
typedef void (*oc_adv_fn)(void* ctx, double* x, const double* a, double h);

static int loop_generic(oc_adv_fn adv, void* ctx, double* x, const double* a, double h, long long iters, int n,
                        int pre, int wvar, double max_value, double* data) {
    for (long long i = 0; i < iters; ++i) {
        if (data) {
            if (wvar < 0) {
                if (n >= 3) data[i] = x[0] + OC_PI * x[1] + OC_E * x[2];
                else if (n == 2) data[i] = x[0] + OC_PI * x[1];
                else data[i] = x[0];
            } else {
                data[i] = x[wvar];
            }
            for (int p = 0; p < pre; ++p) adv(ctx, x, a, h);
        } else {
            adv(ctx, x, a, h);
        }
        if (i % OC_CHECK_INTERVAL == 0) {
            double chk = 0.0;
            for (int q = 0; q < n; ++q) chk += fabs(x[q]);
            if (isnan(chk) || isinf(chk)) return 0;
            if (max_value != 0.0 && fabs(chk) > max_value) return 0;
        }
    }
    double prev[OC_MAXN];
    memcpy(prev, x, sizeof(double) * (size_t)n);
    for (int p = 0; p < pre; ++p) adv(ctx, x, a, h);
    double d = 0.0;
    for (int q = 0; q < n; ++q) d += fabs(x[q] - prev[q]);
    if (fabs(d) < OC_EPS_FIXED_POINT) return -1;
    return 1;
}

static void adv_krs(void* ctx, double* x, const double* a, double h) { ((oc_step_fn)ctx)(x, a, h); }

int oc_loop_model(oc_step_fn step, double* x, const double* a, double h, long long iters, int n, int pre,
                  int wvar, double max_value, double* data) {
    return loop_generic(adv_krs, (void*)step, x, a, h, iters, n, pre, wvar, max_value, data);
}

/* --------------------------------------------------------- peak finder ---
 * cpu_peak_finder, U-CUDA/parametric_engine.cpp:111-166 (≡ peakFinder,
 * cudaLibrary.cu:1503-1633): maxima with a plateau walk, parabolic
 * refinement, first peak dropped, times turned into intervals * h. */
int oc_peak_finder(const double* d, long long n, double* peaks, double* times, double h) {
    if (n < 5) return 0;
    int count = 0;
    for (long long i = 2; i + 2 < n; ++i) {
        if (d[i] - d[i - 1] > OC_EPS_PEAK_DELTA && d[i] > OC_PEAK_THRESHOLD && d[i] >= d[i + 1]) {
            for (long long j = i; j + 2 < n; ++j) {
                if (d[j] < d[j + 1]) { i = j + 1; break; }
                if (d[j] - d[j + 1] > OC_EPS_PEAK_DELTA) {
                    const double denom = d[j - 1] - 2.0 * d[j] + d[j + 1];
                    double delta = 0.0;
                    if (fabs(denom) > 1e-12) delta = 0.5 * (d[j - 1] - d[j + 1]) / denom;
                    peaks[count] = d[j] - 0.25 * (d[j - 1] - d[j + 1]) * delta;
                    times[count] = (double)(j - 1) + delta;
                    ++count;
                    i = j + 1;
                    break;
                }
            }
        }
    }
    if (count > 1) {
        for (int i = 0; i < count - 1; ++i) {
            peaks[i] = peaks[i + 1];
            times[i] = (times[i + 1] - times[i]) * h;
        }
        count -= 1;
    } else {
        count = 0;
    }
    if (count >= OC_MAX_PEAKS) count = OC_MAX_PEAKS;
    return count;
}

/* ------------------------------------------------------- Gram-Schmidt ---
 * cpu_gram_schmidt / cpu_projection_operator, U-CUDA/parametric_engine.cpp:858-885
 * (classical: projections of the ORIGINAL a_i on the unnormalised b_j). */
void oc_cgs(const double* a, double* b, int n, double* den) {
    for (int i = 0; i < n; ++i) {
        for (int q = 0; q < n; ++q) b[q + i * n] = a[q + i * n];
        for (int jj = 0; jj < i; ++jj) {
            double num = 0.0, dd = 0.0;
            for (int q = 0; q < n; ++q) { num += a[i * n + q] * b[jj * n + q]; dd += b[jj * n + q] * b[jj * n + q]; }
            const double f = (dd == 0.0) ? 0.0 : num / dd;
            for (int q = 0; q < n; ++q) b[i * n + q] -= f * b[jj * n + q];
        }
    }
    for (int i = 0; i < n; ++i) {
        double nrm = 0.0;
        for (int q = 0; q < n; ++q) nrm += b[i * n + q] * b[i * n + q];
        nrm = sqrt(nrm);
        for (int q = 0; q < n; ++q) b[i * n + q] = (nrm == 0.0) ? 0.0 : b[i * n + q] / nrm;
        if (den) den[i] = nrm;
    }
}

/* Modified Gram-Schmidt used by the variational method (K row vectors). */
void oc_mgs(double* v, int K, int n, double* d) {
    for (int k = 0; k < K; ++k) {
        for (int q = 0; q < k; ++q) {
            double r = 0.0;
            for (int i = 0; i < n; ++i) r = r + v[k * n + i] * v[q * n + i];
            for (int i = 0; i < n; ++i) v[k * n + i] = v[k * n + i] - r * v[q * n + i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm = nrm + v[k * n + i] * v[k * n + i];
        nrm = sqrt(nrm);
        d[k] = nrm;
        for (int i = 0; i < n; ++i) v[k * n + i] = (nrm == 0.0) ? 0.0 : v[k * n + i] / nrm;
    }
}

/* ----------------------------------------------------------------- RK4 ---
 * Classic RK4 in the stage order of the reference KRS body
 * (U-CUDA/codegen.cpp:597-606): X1 = X + 0.5*h*k (stages 0, 1), X1 = X + h*k
 * (stage 2), X += h*(k0 + 2*k1 + 2*k2 + k3)/6, summed left to right. */
static void rk4_at(oc_rhs_fn rhs, double* x, const double* a, double h, double t, int n) {
    double x1[OC_MAXN], k[OC_MAXN], acc[OC_MAXN];
    const double hh = 0.5 * h;
    rhs(x, a, t, k);
    for (int i = 0; i < n; ++i) { acc[i] = k[i]; x1[i] = x[i] + hh * k[i]; }
    rhs(x1, a, t + hh, k);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = x[i] + hh * k[i]; }
    rhs(x1, a, t + hh, k);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = x[i] + h * k[i]; }
    rhs(x1, a, t + h, k);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + k[i]; x[i] = x[i] + h * acc[i] / 6.0; }
}

void oc_rk4(oc_rhs_fn rhs, double* x, const double* a, double h, int n) { rk4_at(rhs, x, a, h, 0.0, n); }

/* RK4 applied to the variational system V' = J(X) V with the same stages
 * (the exact derivative of the discrete RK4 map). */
static void rk4_tan_at(oc_rhs_fn rhs, oc_tan_fn tan, double* x, double* V, int K, const double* a, double h, double t,
                       int n) {
    double x1[OC_MAXN], k[OC_MAXN], acc[OC_MAXN];
    double v1[OC_MAXN * OC_MAXN], l[OC_MAXN * OC_MAXN], lacc[OC_MAXN * OC_MAXN];
    const double hh = 0.5 * h;
    const int m = K * n;
    rhs(x, a, t, k);
    tan(x, a, t, V, l, K);
    for (int i = 0; i < n; ++i) { acc[i] = k[i]; x1[i] = x[i] + hh * k[i]; }
    for (int i = 0; i < m; ++i) { lacc[i] = l[i]; v1[i] = V[i] + hh * l[i]; }
    rhs(x1, a, t + hh, k);
    tan(x1, a, t + hh, v1, l, K);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = x[i] + hh * k[i]; }
    for (int i = 0; i < m; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + hh * l[i]; }
    rhs(x1, a, t + hh, k);
    tan(x1, a, t + hh, v1, l, K);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = x[i] + h * k[i]; }
    for (int i = 0; i < m; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + h * l[i]; }
    rhs(x1, a, t + h, k);
    tan(x1, a, t + h, v1, l, K);
    for (int i = 0; i < n; ++i) { acc[i] = acc[i] + k[i]; x[i] = x[i] + h * acc[i] / 6.0; }
    for (int i = 0; i < m; ++i) { lacc[i] = lacc[i] + l[i]; V[i] = V[i] + h * lacc[i] / 6.0; }
}

void oc_rk4_tan(oc_rhs_fn rhs, oc_tan_fn tan, double* x, double* V, int K, const double* a, double h, int n) {
    rk4_tan_at(rhs, tan, x, V, K, a, h, 0.0, n);
}

/* right-hand-side stepping with the system time t advanced by h per step
 * (t is read only by systems with a time symbol) */
typedef struct { oc_rhs_fn rhs; oc_tan_fn tan; double* V; int K; int n; double t; } tan_ctx;
static void adv_rhs(void* ctx, double* x, const double* a, double h) {
    tan_ctx* c = (tan_ctx*)ctx;
    rk4_at(c->rhs, x, a, h, c->t, c->n);
    c->t += h;
}
static void adv_tan(void* ctx, double* x, const double* a, double h) {
    tan_ctx* c = (tan_ctx*)ctx;
    rk4_tan_at(c->rhs, c->tan, x, c->V, c->K, a, h, c->t, c->n);
    c->t += h;
}

/* Random unit probes (LLEKernelCUDA :2603-2626, LSKernelCUDA :2980-2997):
 * K vectors of n draws u - 0.5, each normalised; subsequence = global cell. */
static void probe_basis(u64 seed, long long cell, int K, int n, double* z) {
    u64 st;
    oc_rng_init(seed, (u64)cell, 0ULL, &st);
    for (int k = 0; k < K; ++k) {
        double p = 0.0;
        for (int i = 0; i < n; ++i) { z[k * n + i] = (double)(oc_rng_uniform(&st) - 0.5); p += z[k * n + i] * z[k * n + i]; }
        p = sqrt(p);
        for (int i = 0; i < n; ++i) z[k * n + i] /= p;
    }
}

/* -------------------------------------------------------- bifurcation ---
 * One cell of calculateDiscreteModelCUDA + peakFinderCUDA
 * (cudaLibrary.cu:944-1077, 1641-1674; host counts parametric_engine.cpp:1796-1797).
 * Transient as the kernel runs it: one step per iteration, pre_scaller
 * trailing steps (the CPU twin passes pre_scaller = 1 there, :472-475).
 * Recorded run only for transient flag 1 / -1. Under an h axis the counts
 * are recomputed per cell with ceil (cudaLibrary.cu:1028-1030). */
/* KRS step with a relative state kick after every `every`-th step
 * (noise > 0); see noise_kick below. */
typedef struct noisy_step {
    oc_step_fn step;
    double noise;
    unsigned long long rng;
    long long count, every;
    int n;
} noisy_step;
static void noise_kick(double* x, int n, double noise, unsigned long long* s);
static void adv_krs_noisy(void* ctx, double* x, const double* a, double h) {
    noisy_step* c = (noisy_step*)ctx;
    c->step(x, a, h);
    if (++c->count % c->every == 0) noise_kick(x, c->n, c->noise, &c->rng);
}

/* jr (optional): the recorded run uses jr's parameters and its own noise
 * size noise_rec (a transient at one precision, the record at another). */
static void bif_cell(oc_step_fn step, const oc_job* j, long long cell, long long iters_max, double* data,
                     double* pk, double* tm, int* flag_out, double* peaks, double* intervals, long long stride,
                     double* traj, int traj_len, double noise, unsigned long long noise_seed, int every,
                     const oc_job* jr, double noise_rec) {
    double a[OC_MAXV], x[OC_MAXN], h;
    if (!oc_cell_setup(j, cell, a, x, &h)) { *flag_out = -1; return; }
    long long skip = j->h > 0.0 ? (long long)(j->transient_time / j->h) : 0, iters = iters_max;
    if (sweeps_step(j)) {
        skip = (long long)ceil(j->transient_time / h);
        long long it = (long long)ceil(j->t_max / h / (double)j->pre_scaller);
        if (it < iters) iters = it;
    }
    noisy_step ns = {step, noise, noise_seed ^ ((unsigned long long)cell * 0xD1B54A32D192ED03ull), 0,
                     every > 0 ? every : 100, j->n_vars};
    const oc_adv_fn adv = noise > 0.0 ? adv_krs_noisy : adv_krs;
    void* ctx = noise > 0.0 ? (void*)&ns : (void*)step;
    int flag = loop_generic(adv, ctx, x, a, h, skip, j->n_vars, j->pre_scaller, 0, j->max_value, NULL);
    if (!(flag == 1 || flag == -1)) return (void)(*flag_out = flag);
    if (jr) {
        double xr[OC_MAXN], hr;
        if (!oc_cell_setup(jr, cell, a, xr, &hr)) { *flag_out = -1; return; }
        noisy_step nr = {step, noise_rec, ns.rng ^ 0x9E3779B97F4A7C15ull, 0, every > 0 ? every : 100, j->n_vars};
        const oc_adv_fn advr = noise_rec > 0.0 ? adv_krs_noisy : adv_krs;
        void* ctxr = noise_rec > 0.0 ? (void*)&nr : (void*)step;
        flag = loop_generic(advr, ctxr, x, a, hr, iters, j->n_vars, j->pre_scaller, j->writable_var, j->max_value, data);
    } else {
        flag = loop_generic(adv, ctx, x, a, h, iters, j->n_vars, j->pre_scaller, j->writable_var, j->max_value, data);
    }
    for (int i = 0; i < traj_len && i < iters; ++i) traj[(long long)i * stride] = data[i];
    if (flag == -1 || flag == 0) { *flag_out = flag; return; }
    int cnt = oc_peak_finder(data, iters, pk, tm, h * (double)j->pre_scaller);
    if (cnt > j->max_peaks) cnt = j->max_peaks;
    *flag_out = cnt;
    for (int k = 0; k < cnt; ++k) {
        peaks[(long long)k * stride] = pk[k];
        intervals[(long long)k * stride] = tm[k];
    }
}

/* Transient of job j with noise, record with job jr's parameters and
 * noise_rec (jr NULL: the record continues with j and noise). */
void oc_bif_cells_phased(oc_step_fn step, const oc_job* j, const oc_job* jr, const long long* cells, long long n_sel,
                         int* flags, double* peaks, double* intervals, double* traj, int traj_len, int threads,
                         double noise, double noise_rec, unsigned long long noise_seed, int noise_every) {
    const long long iters_max = (long long)ceil(j->t_max / worst_step(j) / j->pre_scaller);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)(iters_max + 8));
        double* pk = (double*)malloc(sizeof(double) * (size_t)(iters_max + 8));
        double* tm = (double*)malloc(sizeof(double) * (size_t)(iters_max + 8));
#pragma omp for schedule(dynamic, 1)
        for (long long s = 0; s < n_sel; ++s)
            bif_cell(step, j, cells[s], iters_max, data, pk, tm, &flags[s], peaks + s, intervals + s, n_sel,
                     traj ? traj + s : NULL, traj ? traj_len : 0, noise, noise_seed, noise_every, jr, noise_rec);
        free(data);
        free(pk);
        free(tm);
    }
}

void oc_bif_cells_noisy(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int* flags,
                        double* peaks, double* intervals, double* traj, int traj_len, int threads, double noise,
                        unsigned long long noise_seed, int noise_every) {
    oc_bif_cells_phased(step, j, NULL, cells, n_sel, flags, peaks, intervals, traj, traj_len, threads, noise, noise,
                        noise_seed, noise_every);
}

void oc_bif_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int* flags,
                  double* peaks, double* intervals, double* traj, int traj_len, int threads) {
    oc_bif_cells_noisy(step, j, cells, n_sel, flags, peaks, intervals, traj, traj_len, threads, 0.0, 0, 100);
}

/* ------------------------------------------------------------ DBSCAN ---
 * The reference's per-cell peak clustering for 2-D bifurcation maps
 * (dbscan, U-CUDA/cudaLibrary.cu:2088-2166, called by dbscanCUDA
 * :2303-2327 as compiled into kernels/bifurcation2d.template.cu), restated
 * step for step because its result is the map value: labels[0..n) and a
 * neighbour-slot array[0..n) start at 0; a peak without a label opens a new
 * cluster and labels every unlabelled peak k < n - 1 (the last peak is never
 * scanned as a neighbour) within eps, writing slot[k] = k and counting it;
 * while that count is positive the loop variable is replaced by
 * slot[count - 1] (cleared afterwards) and the scan repeats from there; the
 * loop variable then advances by one. Distance = sqrt(dp^2 + di^2) with the
 * interval term multiplied by 0 (configCUDA.h:91-92: mult_peak 1,
 * mult_interval 0). Returns -1 / 0 / 1 for a count of -1 / 0 / 1 peaks,
 * else the number of clusters minus one. `labels` and `slots` hold n ints. */
int oc_dbscan_count(const double* peaks, int n, double eps, int* labels, int* slots) {
    if (n == -1) return -1;
    if (n == 0) return 0;
    if (n == 1) return 1;
    int cluster = 0, neighbours = 0;
    for (int q = 0; q < n; ++q) { labels[q] = 0; slots[q] = 0; }
    for (long long i = 0; i < n; ++i) {
        if (neighbours >= 1) {
            i = slots[neighbours - 1];
            slots[neighbours - 1] = 0;
            neighbours -= 1;
        } else if (labels[i] == 0) {
            neighbours = 0;
            ++cluster;
            labels[i] = cluster;
        } else {
            continue;
        }
        for (int k = 0; k < n - 1; ++k) {
            if (i != k && labels[k] == 0) {
                const double dp = peaks[i] - peaks[k], di = 0.0;
                if (sqrt(dp * dp + di * di) < eps) {
                    labels[k] = cluster;
                    slots[k] = k;
                    ++neighbours;
                }
            }
        }
    }
    return cluster - 1;
}

/* --------------------------------------------------------------- DFT ---
 * 1-D DFT sweep (run_dft1d_classical, U-CUDA/parametric_engine.cpp:4809-5120;
 * kernels calculateDiscreteModelCUDA + DFT_custom, cudaLibrary.cu:1698-1777):
 * transient of (int)(transient / h) steps (with the loop's trailing pre
 * steps), a block of (int)(t_max / h / pre) samples of the sampled
 * variable; checker = the block's loop flag: -1 -> every a_k, b_k = -1,
 * 0 -> 0, else for each frequency f_k (linear, or the log axis formula)
 * a_k + i b_k = (1/N) sum_n w_n x_n (cos, sin)(2 pi dt f_k n) with dt =
 * h * pre, the phase advanced by a rotation and resynchronised every 1000
 * samples from cosf / sinf of the exact angle (float precision, as the
 * kernel's cosf). Window (build_window, :4787-4801): 0 none, 1 Hanning
 * 0.5 (1 - cos(2 pi n / (N - 1))), 2 Hamming 0.53836 - 0.46164 cos(...). */
void oc_dft_window(int window_type, int N, double* w) {
    if (window_type == 0) { for (int n = 0; n < N; ++n) w[n] = 1.0; return; }
    const double gamma = 2.0 * OC_PI / (double)(N - 1);
    for (int n = 0; n < N; ++n)
        w[n] = window_type == 2 ? 0.53836 - 0.46164 * cos(gamma * (double)n) : 0.5 * (1.0 - cos(gamma * (double)n));
}

void oc_dft_block(const double* x, int N, const double* w, int n_freq, double f_lo, double f_hi, int log_axis,
                  double dt, double* ak, double* bk) {
    const double f_step = (f_hi - f_lo) / (double)(n_freq - 1);
    const double psi = 2.0 * OC_PI * dt;
    for (int k = 0; k < n_freq; ++k) {
        double f_k;
        if (log_axis) {
            const double l0 = log10(f_lo), l1 = log10(f_hi);
            f_k = pow(10.0, l0 + (double)k * ((l1 - l0) / (double)(n_freq - 1)));
        } else {
            f_k = f_lo + (double)k * f_step;
        }
        const double c_t = cos(2.0 * OC_PI * dt * f_k), s_t = sin(2.0 * OC_PI * dt * f_k);
        double a = 0.0, b = 0.0, c_n = 1.0, s_n = 0.0;
        int reset = 1000;
        for (int n = 0; n < N; ++n) {
            const double wd = w[n] * x[n];
            a = a + wd * c_n;
            b = b + wd * s_n;
            const double nc = c_n * c_t - s_n * s_t;
            const double ns = s_n * c_t + c_n * s_t;
            c_n = nc;
            s_n = ns;
            if (--reset == 0) {
                reset = 1000;
                if (n + 1 < N) {
                    const double exact = psi * f_k * ((double)n + 1.0);
                    c_n = (double)cosf((float)exact);
                    s_n = (double)sinf((float)exact);
                }
            }
        }
        ak[k] = a / (double)N;
        bk[k] = b / (double)N;
    }
}

void oc_dft_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int n_freq, double f_lo,
                  double f_hi, int log_axis, int window_type, double* ak, double* bk, int* flags, int threads) {
    const long long N = (long long)(j->t_max / j->h / (double)j->pre_scaller);
    double* w = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    oc_dft_window(window_type, (int)N, w);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
#pragma omp for schedule(dynamic, 1)
        for (long long s = 0; s < n_sel; ++s) {
            double a[OC_MAXV], x[OC_MAXN], h;
            double* A = ak + s * n_freq;
            double* B = bk + s * n_freq;
            int flag = -1;
            if (oc_cell_setup(j, cells[s], a, x, &h)) {
                const long long skip = (long long)(j->transient_time / j->h);
                flag = oc_loop_model(step, x, a, h, skip, j->n_vars, j->pre_scaller, 0, j->max_value, NULL);
                if (flag == 1 || flag == -1)
                    flag = oc_loop_model(step, x, a, h, N, j->n_vars, j->pre_scaller, j->writable_var, j->max_value,
                                         data);
            }
            flags[s] = flag;
            if (flag == -1 || flag == 0) {
                for (int k = 0; k < n_freq; ++k) A[k] = B[k] = (double)flag;
                continue;
            }
            oc_dft_block(data, (int)N, w, n_freq, f_lo, f_hi, log_axis, h * (double)j->pre_scaller, A, B);
        }
        free(data);
    }
    free(w);
}

/* DFT under an h sweep (run_dft1d_cpu with sweep_over_h, no continuation,
 * U-CUDA/parametric_engine.cpp:1088-1225; its GPU twin dft1dHSweepKernel,
 * kernels/dft1d_hsweep.template.cu:85-205): per cell h from the grid,
 * block = (int)(t_max / h / pre) capped at ceil(t_max / min(h lo, h hi) /
 * pre); h <= 0 or a block of <= 2 samples -> -1; transient of
 * (int)(transient / h) single steps plus one trailing step, an escape there
 * (flag 0) -> -1; the block's loop flag -1 / 0 fills -1 / 0; otherwise the
 * block's DFT (cpu_dft_block, :811-851) with the window of the block's own
 * length (cpu_build_window, :794-803), dt = h * pre and the frequency axis
 * f_k = lo + k (hi - lo) / (n - 1), log: 10^(l0 + (l1 - l0) k / (n - 1))
 * (a single frequency sits at lo). */
static void dft_block_cpu(const double* x, int N, const double* w, int n_freq, double f_lo, double f_hi, int log_axis,
                          double dt, double* ak, double* bk) {
    const double f_step = n_freq > 1 ? (f_hi - f_lo) / (double)(n_freq - 1) : 0.0;
    const double psi = 2.0 * OC_PI * dt;
    for (int k = 0; k < n_freq; ++k) {
        const double f_k = log_axis
            ? pow(10.0, log10(f_lo) + (n_freq > 1 ? (log10(f_hi) - log10(f_lo)) * (double)k / (double)(n_freq - 1) : 0.0))
            : f_lo + (double)k * f_step;
        const double c_t = cos(2.0 * OC_PI * dt * f_k), s_t = sin(2.0 * OC_PI * dt * f_k);
        double a = 0.0, b = 0.0, c_n = 1.0, s_n = 0.0;
        int reset = 1000;
        for (int n = 0; n < N; ++n) {
            const double wd = w[n] * x[n];
            a = a + wd * c_n;
            b = b + wd * s_n;
            const double nc = c_n * c_t - s_n * s_t;
            const double ns = s_n * c_t + c_n * s_t;
            c_n = nc;
            s_n = ns;
            if (--reset == 0) {
                reset = 1000;
                if (n + 1 < N) {
                    const double exact = psi * f_k * (double)(n + 1);
                    c_n = (double)cosf((float)exact);
                    s_n = (double)sinf((float)exact);
                }
            }
        }
        ak[k] = a / (double)N;
        bk[k] = b / (double)N;
    }
}

void oc_dft_hsweep_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int n_freq,
                         double f_lo, double f_hi, int log_axis, int window_type, double* ak, double* bk, int* flags,
                         int threads) {
    int sax = 0;
    for (int ax = 0; ax < j->n_axes; ++ax)
        if (j->axis_kind[ax] == OC_AXIS_STEP) sax = ax;
    const double worst = j->axis_lo[sax] < j->axis_hi[sax] ? j->axis_lo[sax] : j->axis_hi[sax];
    const long long max_n = (long long)ceil(j->t_max / worst / (double)j->pre_scaller);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)(max_n > 0 ? max_n : 1));
        double* w = (double*)malloc(sizeof(double) * (size_t)(max_n > 0 ? max_n : 1));
#pragma omp for schedule(dynamic, 1)
        for (long long s = 0; s < n_sel; ++s) {
            double a[OC_MAXV], x[OC_MAXN], h;
            double* A = ak + s * n_freq;
            double* B = bk + s * n_freq;
            int flag = -1;
            long long n = 0;
            if (oc_cell_setup(j, cells[s], a, x, &h)) {
                n = (long long)(j->t_max / h / (double)j->pre_scaller);
                if (n > max_n) n = max_n;
                if (n > 2) {
                    flag = oc_loop_model(step, x, a, h, (long long)(j->transient_time / h), j->n_vars, 1, 0,
                                         j->max_value, NULL);
                    if (flag == 0)
                        flag = -1;
                    else
                        flag = oc_loop_model(step, x, a, h, n, j->n_vars, j->pre_scaller, j->writable_var,
                                             j->max_value, data);
                }
            }
            flags[s] = flag;
            if (flag == -1 || flag == 0) {
                for (int k = 0; k < n_freq; ++k) A[k] = B[k] = (double)flag;
                continue;
            }
            oc_dft_window(window_type, (int)n, w);
            dft_block_cpu(data, (int)n, w, n_freq, f_lo, f_hi, log_axis, h * (double)j->pre_scaller, A, B);
        }
        free(data);
        free(w);
    }
}

/* ------------------------------------------------------- continuation ---
 * Continuation sweeps (kernels/bifurcation1d_cont.template.cu:56-140,
 * lle1d_cont.template.cu:71-183, ls1d_cont.template.cu:57-180,
 * dft1d_cont.template.cu:72-175; host counts in
 * parametric_engine.cpp:4563-4568, 4163-4167): axis 0 (parameter or h) is
 * walked point by point with cont_sweep_value (:177-186), each point
 * starting from the previous point's state; axis 1, when present, gives
 * independent chains (cell = chain * n0 + point). Where a block escapes
 * part way, the reference goes on to read samples it never wrote; this
 * restatement (and the engine) report such a block as 0 instead, and the
 * DFT chain restarts after it. */
static double cont_value(const oc_job* j, int jj, int reverse) {
    const int n = j->axis_n[0];
    const double t = (double)jj / (double)(n > 1 ? n - 1 : 1);
    const double lo = j->axis_lo[0], hi = j->axis_hi[0];
    if (j->axis_log[0]) {
        const double l0 = log10(lo), l1 = log10(hi);
        return pow(10.0, reverse ? (l1 - (l1 - l0) * t) : (l0 + (l1 - l0) * t));
    }
    return reverse ? (hi - (hi - lo) * t) : (lo + (hi - lo) * t);
}

/* a[], x0 and h of chain `chain`: the base values with axis 1 applied */
static void chain_setup(const oc_job* j, long long chain, double* a, double* x0, double* h) {
    oc_job one = *j;
    one.n_axes = 0;
    oc_cell_setup(&one, chain, a, x0, h);  /* base values; random x0 keyed by the chain */
    if (j->n_axes > 1) {
        const double v = axis_at(j, 1, chain % j->axis_n[1]);
        if (j->axis_kind[1] == OC_AXIS_PARAM) a[j->axis_index[1]] = v;
        else if (j->axis_kind[1] == OC_AXIS_VAR) x0[j->axis_index[1]] = v;
        else *h = v;
    }
}

static double chain_point(const oc_job* j, int jj, int reverse, double* a, double h) {
    const double p = cont_value(j, jj, reverse);
    if (j->axis_kind[0] == OC_AXIS_STEP) return p;
    a[j->axis_index[0]] = p;
    return h;
}

static long long chains_of(const oc_job* j) { return j->n_axes > 1 ? j->axis_n[1] : 1; }

/* dft: 0 bifurcation chain (flags = peak counts, peaks / intervals
 * [max_peaks][cells]), 1 DFT chain (flags 1 / 0 / -1, blocks[cell][max_n],
 * block lengths lens[cell]) */
void oc_cont_bif(oc_step_fn step, const oc_job* j, int reverse, int dft, int* flags, double* peaks,
                 double* intervals, double* blocks, int* lens, int threads) {
    const int n0 = j->axis_n[0];
    const long long ncell = (long long)n0 * chains_of(j);
    double hmin = j->h;
    if (j->axis_kind[0] == OC_AXIS_STEP) hmin = j->axis_lo[0] < j->axis_hi[0] ? j->axis_lo[0] : j->axis_hi[0];
    else if (j->n_axes > 1 && j->axis_kind[1] == OC_AXIS_STEP) hmin = j->axis_lo[1] < j->axis_hi[1] ? j->axis_lo[1] : j->axis_hi[1];
    const long long max_n = (long long)ceil(j->t_max / hmin / (double)j->pre_scaller);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)(max_n + 8));
        double* pk = (double*)malloc(sizeof(double) * (size_t)(max_n + 8));
        double* tm = (double*)malloc(sizeof(double) * (size_t)(max_n + 8));
#pragma omp for schedule(dynamic, 1)
        for (long long chain = 0; chain < chains_of(j); ++chain) {
            double a[OC_MAXV], x0[OC_MAXN], x[OC_MAXN], hb;
            chain_setup(j, chain, a, x0, &hb);
            memcpy(x, x0, sizeof(double) * (size_t)j->n_vars);
            for (int jj = 0; jj < n0; ++jj) {
                const long long cell = chain * n0 + jj;
                const double h = chain_point(j, jj, reverse, a, hb);
                long long blk = h > 0.0 ? (long long)(j->t_max / h / (double)j->pre_scaller) : 0;
                if (blk > max_n) blk = max_n;
                const long long skip = h > 0.0 ? (long long)(j->transient_time / h) : 0;
                int flag;
                if (lens) lens[cell] = (int)blk;
                if (h <= 0.0 || blk <= (dft ? 2 : 0)) {
                    flag = -1;
                } else if (oc_loop_model(step, x, a, h, skip, j->n_vars, 1, 0, j->max_value, NULL) == 0) {
                    flag = -1;
                } else {
                    const int f = oc_loop_model(step, x, a, h, blk, j->n_vars, j->pre_scaller, j->writable_var,
                                                j->max_value, data);
                    if (f == -1 || f == 0) {
                        flag = f;
                    } else if (dft) {
                        flag = 1;
                        memcpy(blocks + cell * max_n, data, sizeof(double) * (size_t)blk);
                    } else {
                        int cnt = oc_peak_finder(data, blk, pk, tm, j->h * (double)j->pre_scaller);
                        if (cnt > j->max_peaks) cnt = j->max_peaks;
                        flag = cnt;
                        for (int k = 0; k < cnt; ++k) {
                            peaks[(long long)k * ncell + cell] = pk[k];
                            intervals[(long long)k * ncell + cell] = tm[k];
                        }
                    }
                }
                flags[cell] = flag;
                if (dft && flag != 1) memcpy(x, x0, sizeof(double) * (size_t)j->n_vars);
            }
        }
        free(data);
        free(pk);
        free(tm);
    }
}

/* DFT chain: the blocks of oc_cont_bif(dft = 1), then each block's DFT
 * (cpu_dft_block with the window of its length and dt = h pre) */
void oc_cont_dft(oc_step_fn step, const oc_job* j, int reverse, int n_freq, double f_lo, double f_hi, int log_axis,
                 int window_type, double* ak, double* bk, int* flags, int threads) {
    const int n0 = j->axis_n[0];
    const long long ncell = (long long)n0 * chains_of(j);
    double hmin = j->h;
    if (j->axis_kind[0] == OC_AXIS_STEP) hmin = j->axis_lo[0] < j->axis_hi[0] ? j->axis_lo[0] : j->axis_hi[0];
    else if (j->n_axes > 1 && j->axis_kind[1] == OC_AXIS_STEP) hmin = j->axis_lo[1] < j->axis_hi[1] ? j->axis_lo[1] : j->axis_hi[1];
    const long long max_n = (long long)ceil(j->t_max / hmin / (double)j->pre_scaller);
    double* blocks = (double*)malloc(sizeof(double) * (size_t)(ncell * max_n + 1));
    int* lens = (int*)malloc(sizeof(int) * (size_t)ncell);
    oc_cont_bif(step, j, reverse, 1, flags, NULL, NULL, blocks, lens, threads);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* w = (double*)malloc(sizeof(double) * (size_t)(max_n + 1));
#pragma omp for schedule(dynamic, 1)
        for (long long cell = 0; cell < ncell; ++cell) {
            double* A = ak + cell * n_freq;
            double* B = bk + cell * n_freq;
            if (flags[cell] != 1) {
                for (int k = 0; k < n_freq; ++k) A[k] = B[k] = (double)flags[cell];
                continue;
            }
            double a[OC_MAXV], x0[OC_MAXN], hb;
            chain_setup(j, cell / n0, a, x0, &hb);
            const double h = chain_point(j, (int)(cell % n0), reverse, a, hb);
            oc_dft_window(window_type, lens[cell], w);
            dft_block_cpu(blocks + cell * max_n, lens[cell], w, n_freq, f_lo, f_hi, log_axis,
                          h * (double)j->pre_scaller, A, B);
        }
        free(w);
    }
    free(blocks);
    free(lens);
}

/* Lyapunov chain, K = 1 or K = n_vars: values[K][cells], flags (1 / -1) */
static void cont_probes(unsigned long long* st, int K, int n, double* y) {
    for (int k = 0; k < K; ++k) {
        double zp = 0.0;
        for (int i = 0; i < n; ++i) {
            y[k * n + i] = (double)oc_rng_uniform(st) - 0.5;
            zp = zp + y[k * n + i] * y[k * n + i];
        }
        zp = sqrt(zp);
        for (int i = 0; i < n; ++i) y[k * n + i] = (zp == 0.0) ? 0.0 : y[k * n + i] / zp;
    }
}

void oc_cont_lyap(oc_step_fn step, const oc_job* j, int reverse, int K, double* values, int* flags, int threads) {
    const int n0 = j->axis_n[0], n = j->n_vars;
    const long long ncell = (long long)n0 * chains_of(j);
    const long long n_blocks = (long long)(j->t_max / j->NT);
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(dynamic, 1)
    for (long long chain = 0; chain < chains_of(j); ++chain) {
        double a[OC_MAXV], x0[OC_MAXN], x[OC_MAXN], hb;
        double y[OC_MAXN * OC_MAXN], z[OC_MAXN * OC_MAXN], den[OC_MAXN], sum[OC_MAXN];
        chain_setup(j, chain, a, x0, &hb);
        memcpy(x, x0, sizeof(double) * (size_t)n);
        unsigned long long st;
        oc_rng_init(j->probe_seed, 0ULL, 0ULL, &st);
        cont_probes(&st, K, n, y);
        int attached = 0;
        for (int jj = 0; jj < n0; ++jj) {
            const long long cell = chain * n0 + jj;
            const double h = chain_point(j, jj, reverse, a, hb);
            const long long nts = (h > 0.0 && n_blocks > 0) ? (long long)(j->NT / h) : 0;
            if (h <= 0.0 || n_blocks <= 0 || nts <= 0) {
                for (int k = 0; k < K; ++k) values[(long long)k * ncell + cell] = NAN;
                flags[cell] = -1;
                continue;
            }
            int alive = 1;
            long long settle = (long long)(j->transient_time / j->NT);
            if (!attached) {
                settle = 0;
                if (oc_loop_model(step, x, a, h, (long long)(j->transient_time / h), n, 1, 0, j->max_value, NULL) == 0) {
                    alive = 0;
                } else {
                    if (K == 1) {
                        for (int i = 0; i < n; ++i) y[i] = y[i] * j->eps + x[i];
                    } else {
                        oc_cgs(y, z, n, den);
                        for (int k = 0; k < K; ++k)
                            for (int i = 0; i < n; ++i) y[k * n + i] = z[k * n + i] * j->eps + x[i];
                    }
                    attached = 1;
                }
            }
            for (int k = 0; k < K; ++k) sum[k] = 0.0;
            for (long long b = 0; alive && b < settle + n_blocks; ++b) {
                if (oc_loop_model(step, x, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { alive = 0; break; }
                for (int k = 0; k < K; ++k)
                    if (oc_loop_model(step, y + k * n, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { alive = 0; break; }
                if (!alive) break;
                if (K == 1) {
                    double d = 0.0;
                    for (int l = 0; l < n; ++l) {
                        const double q = (x[l] - y[l]) / j->eps;
                        d = d + q * q;
                    }
                    d = sqrt(d);
                    if (d <= 1e-14) d = 1e-14;
                    if (b >= settle) sum[0] = sum[0] + log(d);
                    const double inv = 1.0 / d;
                    for (int l = 0; l < n; ++l) y[l] = x[l] - ((x[l] - y[l] + 1e-14) * inv);
                } else {
                    for (int k = 0; k < K; ++k)
                        for (int l = 0; l < n; ++l) y[k * n + l] = y[k * n + l] - x[l];
                    oc_cgs(y, z, n, den);
                    for (int k = 0; k < K; ++k) {
                        if (b >= settle) sum[k] = sum[k] + log(den[k] / j->eps);
                        for (int i = 0; i < n; ++i) y[k * n + i] = x[i] + z[k * n + i] * j->eps;
                    }
                }
            }
            if (alive) {
                for (int k = 0; k < K; ++k) values[(long long)k * ncell + cell] = sum[k] / j->t_max;
                flags[cell] = 1;
            } else {
                for (int k = 0; k < K; ++k) values[(long long)k * ncell + cell] = NAN;
                flags[cell] = -1;
                attached = 0;
                memcpy(x, x0, sizeof(double) * (size_t)n);
                oc_rng_init(j->probe_seed, (unsigned long long)jj + 1ULL, 0ULL, &st);
                cont_probes(&st, K, n, y);
            }
        }
    }
}

/* ------------------------------------------------------- fast synchro ---
 * FastSynchro (U-CUDA/cudaLibrary.cu): the coupled step
 * calculateDiscreteModelforFastSynchro (:424-452): N = K (S1 - X) from the
 * state before the step, X <- step(X, h forward / -h backward), X += h N;
 * the synchronisation loops loopCalculateDiscreteModelForFastSynchro_2
 * (:547-720, on grid) and loopCalculateDiscreteModelForFastSynchro
 * (:3630-3800, on attractor), which share the passes and estimators; the
 * master fill fillFSMasterTrajectory (kernels/fastsync_attr.template.cu:
 * 75-98) and the fixed-state transient fillFSTransientIC (:454-474), both
 * with K = 0 and S1 = X; run_fastsync's counts
 * (parametric_engine.cpp:6906-6916, 7075). */
static void fs_step(oc_step_fn step, int n, double* X, const double* S1, const double* K, const double* a, double h,
                    int fwd) {
    double N[OC_MAXN];
    for (int i = 0; i < n; ++i) N[i] = K[i] * (S1[i] - X[i]);
    step(X, a, fwd ? h : -h);
    for (int i = 0; i < n; ++i) X[i] = X[i] + h * N[i];
}

static double fs_sqdist(int n, const double* a, const double* b) {
    double r = 0.0;
    for (int j = 0; j < n; ++j) r = r + (a[j] - b[j]) * (a[j] - b[j]);
    return r;
}

/* master: rows [i * n + j] for the unidirectional mode; Xm: initial master
 * for the bidirectional one */
static double fs_sync(oc_step_fn step, const oc_fs* f, const double* master, double* Xm, double* Xs) {
    const int n = f->n_vars, N = f->n_iters;
    const double h = f->h;
    double rms = 0.0, prev[OC_MAXN];
    for (int m = 0; m < f->iter_of_synchr; ++m) {
        for (int i = 0; i < N - 1; ++i) {
            if (f->type_of_synch == 0)
                for (int j = 0; j < n; ++j) Xm[j] = master[(long long)i * n + j];
            if (f->error_estim == 0) {
                if (m == f->iter_of_synchr - 1) rms = rms + fs_sqdist(n, Xm, Xs);
            } else if (f->error_estim == 1) {
                if (i == N - 2) rms = fs_sqdist(n, Xm, Xs);
            } else if (f->error_estim == 2) {
                if (m == f->iter_of_synchr - 1 && i == N - 2) rms = fs_sqdist(n, Xm, Xs);
            } else {
                rms = sqrt(fs_sqdist(n, Xm, Xs));
                if (rms <= f->fs_error_trs) return (double)(i + 1) * h;
            }
            if (f->type_of_synch == 1) memcpy(prev, Xs, sizeof(double) * (size_t)n);
            fs_step(step, n, Xs, Xm, f->k_forward, f->a, h, 1);
            if (f->type_of_synch == 1) fs_step(step, n, Xm, prev, f->k_forward, f->a, h, 1);
        }
        if (f->error_estim == 1) {
            rms = sqrt(rms);
            if (rms <= f->fs_error_trs) return (double)(m + 1);
            rms = 0.0;
        }
        for (int i = N - 1; i > 0; --i) {
            if (f->type_of_synch == 0)
                for (int j = 0; j < n; ++j) Xm[j] = master[(long long)i * n + j];
            if (f->type_of_synch == 1) memcpy(prev, Xs, sizeof(double) * (size_t)n);
            fs_step(step, n, Xs, Xm, f->k_backward, f->a, h, 0);
            if (f->type_of_synch == 1) fs_step(step, n, Xm, prev, f->k_backward, f->a, h, 0);
        }
    }
    if (f->error_estim == 0) rms = sqrt(rms / (double)(N - 1));
    if (f->error_estim == 2) rms = sqrt(rms);
    rms = rms <= f->fs_error_trs ? log10(f->fs_error_trs) : log10(rms);
    if (isinf(rms) || isnan(rms) || rms >= 3.0) rms = 3.0;
    if (f->error_estim == 1) return (double)f->iter_of_synchr;
    if (f->error_estim == 3) return (double)N * h;
    return rms;
}

static void fs_free_self(oc_step_fn step, int n, double* X, const double* a, double h) {
    double s1[OC_MAXN], z[OC_MAXN];
    for (int i = 0; i < n; ++i) { s1[i] = X[i]; z[i] = 0.0; }
    fs_step(step, n, X, s1, z, a, h, 1);
}

long long oc_fs_attractor(oc_step_fn step, oc_fs* f, const double* ic_master, const double* ic_slave, double t_max,
                          double transient, double window, const long long* idx, long long n_sel, double* out,
                          double* traj_out, int threads) {
    const int n = f->n_vars;
    const int nt = (int)(window / f->h), ct = (int)(t_max / f->h), skip = (int)(transient / f->h);
    const long long n_pts = ct / f->pre_scaller, len = (long long)ct + nt;
    double* master = (double*)malloc(sizeof(double) * (size_t)(len > 0 ? len : 1) * (size_t)n);
    double X[OC_MAXN];
    for (int j = 0; j < n; ++j) X[j] = ic_master[j];
    for (int i = 0; i < skip; ++i) fs_free_self(step, n, X, f->a, f->h);
    for (long long i = 0; i < len; ++i) {
        for (int j = 0; j < n; ++j) master[i * n + j] = X[j];
        fs_free_self(step, n, X, f->a, f->h);
    }
    f->n_iters = nt;
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(dynamic, 1)
    for (long long s = 0; s < n_sel; ++s) {
        const double* w = master + idx[s] * f->pre_scaller * n;
        double Xm[OC_MAXN], Xs[OC_MAXN];
        for (int j = 0; j < n; ++j) { Xs[j] = ic_slave[j]; Xm[j] = w[j]; }
        if (traj_out)
            for (int j = 0; j < n; ++j) traj_out[s * n + j] = w[j];
        out[s] = fs_sync(step, f, w, Xm, Xs);
    }
    free(master);
    return n_pts;
}

void oc_fs_grid(oc_step_fn step, oc_fs* f, const double* ic_master, const double* ic_slave, double transient,
                double window, int n_pts, int axis_x_var, int axis_y_var, double x_lo, double x_hi, double y_lo,
                double y_hi, int swap, const long long* cells, long long n_sel, double* out, int threads) {
    const int n = f->n_vars;
    double icm[OC_MAXN], ics[OC_MAXN];
    for (int j = 0; j < n; ++j) { icm[j] = ic_master[j]; ics[j] = ic_slave[j]; }
    const int skip = (int)(transient / f->h);
    if (transient > 0.0 && skip > 0) {
        double* fixed = swap ? icm : ics;
        for (int i = 0; i < skip; ++i) fs_free_self(step, n, fixed, f->a, f->h);
    }
    const int nb = (int)(window / f->h / (double)f->pre_scaller);
    f->n_iters = nb;
    double* xv = (double*)malloc(sizeof(double) * (size_t)n_pts);
    double* yv = (double*)malloc(sizeof(double) * (size_t)n_pts);
    oc_axis_values(0, n_pts, x_lo, x_hi, xv);
    oc_axis_values(0, n_pts, y_lo, y_hi, yv);
#pragma omp parallel num_threads(threads > 1 ? threads : 1)
    {
        double* master = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
        for (long long s = 0; s < n_sel; ++s) {
            const long long c = cells[s];
            double Xm[OC_MAXN], Xs[OC_MAXN], z[OC_MAXN];
            for (int j = 0; j < n; ++j) { Xm[j] = icm[j]; Xs[j] = ics[j]; z[j] = 0.0; }
            double* X = swap ? Xs : Xm;
            X[axis_x_var] = xv[c % n_pts];
            X[axis_y_var] = yv[(c / n_pts) % n_pts];
            if (f->type_of_synch == 0) {
                double M[OC_MAXN];
                for (int j = 0; j < n; ++j) { master[j] = Xm[j]; M[j] = Xm[j]; }
                for (int i = 1; i < nb; ++i) {
                    fs_step(step, n, M, z, z, f->a, f->h, 1);
                    for (int j = 0; j < n; ++j) master[(long long)i * n + j] = M[j];
                }
            }
            out[s] = fs_sync(step, f, master, Xm, Xs);
        }
        free(master);
    }
    free(xv);
    free(yv);
}

static void lyap_counts(const oc_job* j, double h, long long* skip, long long* nts) {
    if (sweeps_step(j)) {
        *nts = (long long)ceil(j->NT / h);
        *skip = (long long)ceil(j->transient_time / h);
    } else {
        *nts = (long long)(j->NT / j->h);
        *skip = (long long)(j->transient_time / j->h);
    }
}

/* -------------------------------------------------------------- LLE FD ---
 * One cell of LLEKernelCUDA (cudaLibrary.cu:2486-2693; CPU twin
 * run_lle1d_cpu, parametric_engine.cpp:618-787). The distance is formed as
 * (1/eps)*(x - y) as in the kernel (the CPU twin divides, :697-700). */
static void lle_fd_cell(oc_step_fn step, const oc_job* j, long long cell, double* out, int* flag) {
    const int n = j->n_vars;
    double a[OC_MAXV], x[OC_MAXN], y[OC_MAXN], z[OC_MAXN], h;
    if (!oc_cell_setup(j, cell, a, x, &h)) { *out = NAN; *flag = -1; return; }
    long long skip, nts;
    lyap_counts(j, h, &skip, &nts);
    const long long blocks = (long long)(j->t_max / j->NT);
    if (oc_loop_model(step, x, a, h, skip, n, 1, 0, j->max_value, NULL) == 0) { *out = 999.0; *flag = -1; return; }
    probe_basis(j->probe_seed, cell, 1, n, z);
    for (int i = 0; i < n; ++i) y[i] = z[i] * j->eps + x[i];
    double sum = 0.0;
    for (long long b = 0; b < blocks; ++b) {
        if (oc_loop_model(step, x, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { *out = 999.0; *flag = -1; return; }
        if (oc_loop_model(step, y, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { *out = 999.0; *flag = -1; return; }
        double d = 0.0;
        for (int l = 0; l < n; ++l) {
            const double t2 = (1.0 / j->eps) * (x[l] - y[l]);
            d += t2 * t2;
        }
        d = sqrt(d);
        if (d <= 1e-14) d = 1e-14;
        sum += log(d);
        const double inv = 1.0 / d;
        for (int l = 0; l < n; ++l) y[l] = x[l] - ((x[l] - y[l] + 1e-14) * inv);
    }
    *out = sum / j->t_max;
    *flag = 1;
}

void oc_lle_fd_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, double* out,
                     int* flags, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1)
    for (long long s = 0; s < n_sel; ++s) lle_fd_cell(step, j, cells[s], &out[s], &flags[s]);
}

/* --------------------------------------------------------------- LS FD ---
 * One cell of LSKernelCUDA (cudaLibrary.cu:2878-3073; CPU twin
 * run_ls1d_cpu, parametric_engine.cpp:901-1074). out[k * stride]. */
static void ls_fd_cell(oc_step_fn step, const oc_job* j, long long cell, double* out, long long stride, int* flag) {
    const int n = j->n_vars;
    double a[OC_MAXV], x[OC_MAXN], h;
    double y[OC_MAXN * OC_MAXN], z[OC_MAXN * OC_MAXN], den[OC_MAXN], sum[OC_MAXN];
    if (!oc_cell_setup(j, cell, a, x, &h)) {
        for (int k = 0; k < n; ++k) out[k * stride] = NAN;
        *flag = -1;
        return;
    }
    long long skip, nts;
    lyap_counts(j, h, &skip, &nts);
    const long long blocks = (long long)(j->t_max / j->NT);
    probe_basis(j->probe_seed, cell, n, n, z);
    int alive = oc_loop_model(step, x, a, h, skip, n, 1, 0, j->max_value, NULL) != 0;
    if (alive) {
        oc_cgs(z, y, n, NULL);
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < n; ++i) y[k * n + i] = y[k * n + i] * j->eps + x[i];
        for (int k = 0; k < n; ++k) sum[k] = 0.0;
        for (long long b = 0; b < blocks && alive; ++b) {
            if (oc_loop_model(step, x, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { alive = 0; break; }
            for (int k = 0; k < n; ++k)
                if (oc_loop_model(step, y + k * n, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { alive = 0; break; }
            if (!alive) break;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) y[k * n + l] = y[k * n + l] - x[l];
            oc_cgs(y, z, n, den);
            for (int k = 0; k < n; ++k) {
                sum[k] += log(den[k] / j->eps);
                for (int q = 0; q < n; ++q) y[k * n + q] = x[q] + z[k * n + q] * j->eps;
            }
        }
    }
    for (int k = 0; k < n; ++k) out[k * stride] = alive ? sum[k] / j->t_max : 999.0;
    *flag = alive ? 1 : -1;
}

void oc_ls_fd_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, double* out,
                    int* flags, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1)
    for (long long s = 0; s < n_sel; ++s) ls_fd_cell(step, j, cells[s], out + s, n_sel, &flags[s]);
}

/* ------------------------------------------------------ variational ---
 * The engine's variational method on the reference block schedule: the
 * transient as in the FD kernels, then blocks of NT/h steps plus the step
 * loop's trailing step; K = 1 renormalises (floor 1e-14 like the FD
 * distance), K > 1 re-orthonormalises by modified Gram-Schmidt; exponents
 * are the log-norm sums over t_max. */
/* Rounding-noise model for reproducibility classification (noise > 0 only):
 * after every chunk of steps each state component is scaled by
 * (1 + noise * u), u uniform in [-1, 1) from a per-cell splitmix64 stream.
 * noise = 0 leaves the pinned path untouched. */
static double noise_u(unsigned long long* s) {
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static void noise_kick(double* x, int n, double noise, unsigned long long* s) {
    for (int i = 0; i < n; ++i) x[i] = x[i] * (1.0 + noise * noise_u(s));
}

static void lyap_var_cell(oc_rhs_fn rhs, oc_tan_fn tan, int K, const oc_job* j, long long cell, double* out,
                          long long stride, int* flag, double noise, unsigned long long noise_seed) {
    const int n = j->n_vars;
    double a[OC_MAXV], x[OC_MAXN], h, V[OC_MAXN * OC_MAXN], d[OC_MAXN], sum[OC_MAXN];
    if (!oc_cell_setup(j, cell, a, x, &h)) {
        for (int k = 0; k < K; ++k) out[k * stride] = NAN;
        *flag = -1;
        return;
    }
    long long skip, nts;
    lyap_counts(j, h, &skip, &nts);
    const long long blocks = (long long)(j->t_max / j->NT);
    tan_ctx ctx = {rhs, tan, V, K, n, 0.0};
    unsigned long long ns = noise_seed ^ ((unsigned long long)cell * 0xD1B54A32D192ED03ull);
    int alive;
    if (noise > 0.0) {
        /* chunks of 99 loop steps + the loop's trailing step = 100 steps each,
         * then the remainder with its trailing step: the same skip + 1 steps
         * and the same divergence-check steps as the single call below */
        alive = 1;
        long long done = 0;
        for (; skip - done > 100 && alive; done += 100) {
            alive = loop_generic(adv_rhs, &ctx, x, a, h, 99, n, 1, 0, j->max_value, NULL) != 0;
            noise_kick(x, n, noise, &ns);
        }
        if (alive) alive = loop_generic(adv_rhs, &ctx, x, a, h, skip - done, n, 1, 0, j->max_value, NULL) != 0;
    } else {
        alive = loop_generic(adv_rhs, &ctx, x, a, h, skip, n, 1, 0, j->max_value, NULL) != 0;
    }
    for (int k = 0; k < K; ++k) sum[k] = 0.0;
    if (alive) {
        probe_basis(j->probe_seed, cell, K, n, V);
        if (K > 1) oc_mgs(V, K, n, d);
        for (long long b = 0; b < blocks; ++b) {
            if (loop_generic(adv_tan, &ctx, x, a, h, nts, n, 1, 0, j->max_value, NULL) == 0) { alive = 0; break; }
            if (noise > 0.0) noise_kick(x, n, noise, &ns);
            if (K == 1) {
                double nrm = 0.0;
                for (int i = 0; i < n; ++i) nrm = nrm + V[i] * V[i];
                nrm = sqrt(nrm);
                if (nrm <= 1e-14) nrm = 1e-14;
                sum[0] = sum[0] + log(nrm);
                const double inv = 1.0 / nrm;
                for (int i = 0; i < n; ++i) V[i] = V[i] * inv;
            } else {
                oc_mgs(V, K, n, d);
                for (int k = 0; k < K; ++k) sum[k] = sum[k] + log(d[k]);
            }
        }
    }
    for (int k = 0; k < K; ++k) out[k * stride] = alive ? sum[k] / j->t_max : 999.0;
    *flag = alive ? 1 : -1;
}

void oc_lyap_var_cells(oc_rhs_fn rhs, oc_tan_fn tan, int K, const oc_job* j, const long long* cells, long long n_sel,
                       double* out, int* flags, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1)
    for (long long s = 0; s < n_sel; ++s) lyap_var_cell(rhs, tan, K, j, cells[s], out + s, n_sel, &flags[s], 0.0, 0);
}

void oc_lyap_var_cells_noisy(oc_rhs_fn rhs, oc_tan_fn tan, int K, const oc_job* j, const long long* cells,
                             long long n_sel, double* out, int* flags, int threads, double noise,
                             unsigned long long noise_seed) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1)
    for (long long s = 0; s < n_sel; ++s)
        lyap_var_cell(rhs, tan, K, j, cells[s], out + s, n_sel, &flags[s], noise, noise_seed);
}

/* ------------------------------------------------------------- basins ---
 * New method (the reference clusters DBSCAN features instead,
 * U-CUDA/parametric_engine.cpp:6035-6514): adaptive Dormand-Prince 5(4)
 * (Dormand & Prince 1980 tableau, Hairer's error norm and step controller)
 * with the step clipped to land on every Poincare section t = k*T; the
 * state is sampled on the last n_record sections, the smallest period
 * p <= max_period whose samples close within period_tol is found, and the
 * cycle point with the smallest (x0, x1) is the cell's canonical point
 * (p = 0: mean of the samples; p = -1: unbound; p = -2: unresolved). */
static const double DP_C[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
static const double DP_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
    {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
/* e = b5 - b4 */
static const double DP_E[7] = {71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0, -17253.0 / 339200.0,
                               22.0 / 525.0, -1.0 / 40.0};

typedef struct oc_basins_job {
    oc_job base;
    double rtol, atol, h_min, h_max, section_period, period_tol;
    int n_sections, n_record, max_period, label_vars;
    long long max_attempts;
} oc_basins_job;

/* one DP5(4) attempt: x (n) at local time t with step h; k1 = f(t, x) on
 * entry (FSAL); writes x5 and k7 = f(t + h, x5); returns the error norm
 * sqrt(mean_i (e_i / (atol + rtol * max(|x_i|, |x5_i|)))^2), e = h * sum_s E_s k_s. */
static double dp5_attempt(oc_rhs_fn rhs, const double* x, const double* a, double t, double h, int n,
                          const double* k1, double* x5, double* k7, double rtol, double atol, double* e_out) {
    double k[7][OC_MAXN], xs[OC_MAXN];
    for (int i = 0; i < n; ++i) k[0][i] = k1[i];
    for (int s = 1; s < 7; ++s) {
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int q = 0; q < s; ++q) acc = acc + DP_A[s][q] * k[q][i];
            xs[i] = x[i] + h * acc;
        }
        rhs(xs, a, t + DP_C[s] * h, k[s]);
    }
    /* stage 7 is evaluated at x5 itself (FSAL) */
    for (int i = 0; i < n; ++i) { x5[i] = xs[i]; k7[i] = k[6][i]; }
    double err = 0.0;
    for (int i = 0; i < n; ++i) {
        double e = 0.0;
        for (int s = 0; s < 7; ++s) e = e + DP_E[s] * k[s][i];
        e = h * e;
        if (e_out) e_out[i] = e;
        const double sc = atol + rtol * fmax(fabs(x[i]), fabs(x5[i]));
        err = err + (e / sc) * (e / sc);
    }
    return sqrt(err / (double)n);
}

/* Poincare signature of the trailing section samples buf[L][4] (buf[0] the
 * newest, D <= 4 leading state variables): the smallest period
 * q <= max_period, q < L, for which every sample r with r + q < L closes
 * |buf[r] - buf[r + q]| <= tol * (1 + |buf[r]|) in every component; the
 * cycle point among buf[0 .. q-1] with the smallest (x0, then x1) is the
 * canonical point. No closing period: 0 and the mean of the L samples. */
int oc_section_signature(const double* buf, int L, int D, int max_period, double tol, double* point) {
    int p = 0;
    for (int q = 1; q <= max_period && q < L && !p; ++q) {
        int closes = 1;
        for (int r = 0; r + q < L && closes; ++r)
            for (int i = 0; i < D; ++i)
                if (fabs(buf[r * 4 + i] - buf[(r + q) * 4 + i]) > tol * (1.0 + fabs(buf[r * 4 + i]))) { closes = 0; break; }
        if (closes) p = q;
    }
    if (p > 0) {
        int best = 0;
        for (int r = 1; r < p; ++r) {
            if (buf[r * 4] < buf[best * 4] || (buf[r * 4] == buf[best * 4] && D > 1 && buf[r * 4 + 1] < buf[best * 4 + 1]))
                best = r;
        }
        for (int i = 0; i < D; ++i) point[i] = buf[best * 4 + i];
    } else {
        for (int i = 0; i < D; ++i) {
            double m = 0.0;
            for (int r = 0; r < L; ++r) m += buf[r * 4 + i];
            point[i] = m / (double)L;
        }
    }
    return p;
}

/* One DP5(4) attempt from outside (tests): k1 = f(t, x) is evaluated here;
 * x5 gets the 5th-order solution, err_vec (optional) the raw embedded
 * error estimate e_i = h * sum_s E_s k_s; returns the controller's norm. */
double oc_dp5_step(oc_rhs_fn rhs, const double* x, const double* a, double t, double h, int n, double rtol,
                   double atol, double* x5, double* err_vec) {
    double k1[OC_MAXN], k7[OC_MAXN];
    rhs(x, a, t, k1);
    return dp5_attempt(rhs, x, a, t, h, n, k1, x5, k7, rtol, atol, err_vec);
}

static int basins_cell(oc_rhs_fn rhs, const oc_basins_job* bj, long long cell, double* point, int* period,
                       long long* steps) {
    const oc_job* j = &bj->base;
    const int n = j->n_vars, D = bj->label_vars, L = bj->n_record;
    double a[OC_MAXV], x[OC_MAXN], h;
    double buf[64][4];
    oc_cell_setup(j, cell, a, x, &h);
    const double T = bj->section_period;
    double k1[OC_MAXN], x5[OC_MAXN], k7[OC_MAXN];
    long long acc_steps = 0, rej_steps = 0, attempts = 0;
    for (int i = 0; i < D; ++i) point[i] = 0.0;
    for (int sec = 0; sec < bj->n_sections; ++sec) {
        double t = 0.0;
        rhs(x, a, t, k1);
        for (;;) {
            if (attempts >= bj->max_attempts) { *period = -2; steps[0] = acc_steps; steps[1] = rej_steps; return 0; }
            ++attempts;
            const int last = t + h >= T;
            const double hs = last ? T - t : h;
            const double err = dp5_attempt(rhs, x, a, t, hs, n, k1, x5, k7, bj->rtol, bj->atol, NULL);
            double fac = err == 0.0 ? 5.0 : 0.9 * pow(err, -0.2);
            fac = fac < 0.2 ? 0.2 : (fac > 5.0 ? 5.0 : fac);
            if (err <= 1.0) {
                ++acc_steps;
                for (int i = 0; i < n; ++i) { x[i] = x5[i]; k1[i] = k7[i]; }
                double chk = 0.0;
                for (int i = 0; i < n; ++i) chk += fabs(x[i]);
                if (isnan(chk) || isinf(chk) || (j->max_value != 0.0 && chk > j->max_value)) {
                    *period = -1; steps[0] = acc_steps; steps[1] = rej_steps; return 0;
                }
                if (last) {
                    /* keep the unclipped step for the next section */
                    const double hn = hs * fac;
                    if (hn > h) h = hn;
                    if (h > bj->h_max) h = bj->h_max;
                    break;
                }
                t = t + hs;
                h = hs * fac;
            } else {
                ++rej_steps;
                h = hs * fac;
            }
            if (h > bj->h_max) h = bj->h_max;
            if (h < bj->h_min) { *period = -2; steps[0] = acc_steps; steps[1] = rej_steps; return 0; }
        }
        for (int r = L - 1; r > 0; --r)
            for (int i = 0; i < D; ++i) buf[r][i] = buf[r - 1][i];
        for (int i = 0; i < D; ++i) buf[0][i] = x[i];
    }
    steps[0] = acc_steps;
    steps[1] = rej_steps;
    *period = oc_section_signature(&buf[0][0], L, D, bj->max_period, bj->period_tol, point);
    return 1;
}

void oc_basins_cells(oc_rhs_fn rhs, const oc_basins_job* bj, const long long* cells, long long n_sel, int* period,
                     double* point, long long* steps, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1)
    for (long long s = 0; s < n_sel; ++s) {
        double pt[4];
        long long st[2];
        basins_cell(rhs, bj, cells[s], pt, &period[s], st);
        for (int i = 0; i < bj->label_vars; ++i) point[(long long)i * n_sel + s] = pt[i];
        steps[s] = st[0];
        steps[n_sel + s] = st[1];
    }
}
