/* ucuda_oracle.h — FP64 CPU oracle for the ensemble ODE analysis path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the GPU engine is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it. The product library never links or calls it.
 *
 * It restates, in C with numb = double, the reference's CPU path:
 *   krs_cpu            step body -> native krs_step (U-CUDA/krs_cpu.cpp:293-332,405-479)
 *   cpu_loop_model     (U-CUDA/parametric_engine.cpp:59-101)
 *   cpu_peak_finder    (U-CUDA/parametric_engine.cpp:111-166)
 *   cpu_curand_*       (U-CUDA/parametric_engine.cpp:196-209)
 *   Gram-Schmidt       (U-CUDA/parametric_engine.cpp:858-885)
 *   bif / LLE / LS per-point drivers (U-CUDA/parametric_engine.cpp:374-521,
 *   618-787, 901-1074) with the GPU kernels' cell semantics where the two
 *   differ (U-CUDA/cudaLibrary.cu:944-1077, 1344-1396, 2486-2693, 2878-3073),
 * plus the variational Lyapunov method (new in this engine, same block
 * schedule as the reference).
 */
#ifndef UCUDA_ORACLE_H
#define UCUDA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*oc_step_fn)(double* X, const double* a, double h);
typedef void (*oc_rhs_fn)(const double* X, const double* a, double t, double* dX);
typedef void (*oc_tan_fn)(const double* X, const double* a, double t, const double* V, double* JV, int K);

#define OC_AXIS_NONE 0
#define OC_AXIS_PARAM 1
#define OC_AXIS_VAR 2
#define OC_AXIS_STEP 3

typedef struct oc_job {
    int n_vars, n_values;
    const double* ic;
    const double* params;
    int n_axes;
    int axis_kind[2], axis_index[2], axis_n[2], axis_log[2];
    double axis_lo[2], axis_hi[2];
    double h, transient_time, t_max, max_value;
    unsigned long long ic_seed;
    double ic_lo, ic_hi;
    int writable_var, pre_scaller, max_peaks;
    double NT, eps;
    unsigned long long probe_seed;
} oc_job;

/* RNG */
void oc_rng_init(unsigned long long seed, unsigned long long seq, unsigned long long off, unsigned long long* state);
float oc_rng_uniform(unsigned long long* state);

/* grid */
void oc_axis_values(int log_scale, int n, double lo, double hi, double* out);
int oc_cell_setup(const oc_job* j, long long cell, double* a, double* x0, double* h);

/* loop + peaks */
int oc_loop_model(oc_step_fn step, double* x, const double* a, double h, long long iters, int n, int pre,
                  int wvar, double max_value, double* data);
int oc_peak_finder(const double* data, long long n, double* peaks, double* times, double h);
/* 2-D bifurcation period: the reference's per-cell DBSCAN cluster count over
 * a cell's peaks (labels / slots: n ints of scratch) */
int oc_dbscan_count(const double* peaks, int n, double eps, int* labels, int* slots);
/* 1-D DFT sweep (window, one block, whole cells: ak / bk [n_sel][n_freq]) */
void oc_dft_window(int window_type, int N, double* w);
void oc_dft_block(const double* x, int N, const double* w, int n_freq, double f_lo, double f_hi, int log_axis,
                  double dt, double* ak, double* bk);
void oc_dft_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int n_freq, double f_lo,
                  double f_hi, int log_axis, int window_type, double* ak, double* bk, int* flags, int threads);
/* continuation chains along axis 0 (axis 1: independent chains) */
void oc_cont_bif(oc_step_fn step, const oc_job* j, int reverse, int dft, int* flags, double* peaks,
                 double* intervals, double* blocks, int* lens, int threads);
void oc_cont_dft(oc_step_fn step, const oc_job* j, int reverse, int n_freq, double f_lo, double f_hi, int log_axis,
                 int window_type, double* ak, double* bk, int* flags, int threads);
void oc_cont_lyap(oc_step_fn step, const oc_job* j, int reverse, int K, double* values, int* flags, int threads);
/* fast synchronisation: on attractor (returns the number of points; out /
 * traj_out for the selected ones) and on grid (selected cells) */
typedef struct oc_fs {
    int n_vars;
    const double* a;
    double h;
    int iter_of_synchr, pre_scaller, type_of_synch, error_estim;
    double fs_error_trs;
    const double* k_forward;
    const double* k_backward;
    int n_iters;   /* set by the calls */
} oc_fs;
long long oc_fs_attractor(oc_step_fn step, oc_fs* f, const double* ic_master, const double* ic_slave, double t_max,
                          double transient, double window, const long long* idx, long long n_sel, double* out,
                          double* traj_out, int threads);
void oc_fs_grid(oc_step_fn step, oc_fs* f, const double* ic_master, const double* ic_slave, double transient,
                double window, int n_pts, int axis_x_var, int axis_y_var, double x_lo, double x_hi, double y_lo,
                double y_hi, int swap, const long long* cells, long long n_sel, double* out, int threads);
/* the same under an h sweep (per-cell block length, window and dt) */
void oc_dft_hsweep_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int n_freq,
                         double f_lo, double f_hi, int log_axis, int window_type, double* ak, double* bk, int* flags,
                         int threads);

/* Gram-Schmidt */
void oc_cgs(const double* a, double* b, int n, double* den);
void oc_mgs(double* v, int K, int n, double* d);

/* one RK4 step of state (+ K tangents) with the engine's stage order */
void oc_rk4(oc_rhs_fn rhs, double* x, const double* a, double h, int n);
void oc_rk4_tan(oc_rhs_fn rhs, oc_tan_fn tan, double* x, double* V, int K, const double* a, double h, int n);

/* per-cell drivers over a selection of global cell indices (threads <= 1: serial) */
void oc_bif_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int* flags,
                  double* peaks, double* intervals, double* traj, int traj_len, int threads);
/* same with a relative state kick of size `noise` after every
 * `noise_every`-th step */
void oc_bif_cells_noisy(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, int* flags,
                        double* peaks, double* intervals, double* traj, int traj_len, int threads, double noise,
                        unsigned long long noise_seed, int noise_every);
void oc_lle_fd_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, double* out,
                     int* flags, int threads);
void oc_ls_fd_cells(oc_step_fn step, const oc_job* j, const long long* cells, long long n_sel, double* out,
                    int* flags, int threads);
void oc_lyap_var_cells(oc_rhs_fn rhs, oc_tan_fn tan, int K, const oc_job* j, const long long* cells, long long n_sel,
                       double* out, int* flags, int threads);
/* same with a relative state kick of size `noise` after every 100 transient
 * steps and every NT block (models per-step rounding of a lower precision or
 * of another operation order; used to classify reproducible cells) */
void oc_lyap_var_cells_noisy(oc_rhs_fn rhs, oc_tan_fn tan, int K, const oc_job* j, const long long* cells,
                             long long n_sel, double* out, int* flags, int threads, double noise,
                             unsigned long long noise_seed);

#ifdef __cplusplus
}
#endif
#endif
