/* ucuda.h — C ABI of the MI355X ensemble ODE analysis engine.
 *
 * Replaces the reference's NVRTC/driver layer (U-CUDA/nvrtc_engine.h:16-73,
 * U-CUDA/parametric_engine.cpp:1486-1652) and exposes the analysis entry
 * points that U-CUDA's ParametricEngine facade forwards to
 * (U-CUDA/parametric_engine.h:844-889). Plain C types only: no exceptions
 * cross this boundary, every call returns a status code and the engine keeps
 * the last error message (the reference's `Result.error` / NvrtcEngine::error()).
 *
 * Memory: request inputs are host pointers; outputs are caller-owned DEVICE
 * pointers laid out structure-of-arrays (plane-major, cell = iy * nx + ix).
 * Launches are enqueued on the caller's HIP stream (NULL = default stream).
 */
#ifndef UCUDA_H
#define UCUDA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define UC_OK 0
#define UC_ERR_ARGUMENT 1   /* request validation failed (message in uc_last_error) */
#define UC_ERR_PARSE 2      /* system text could not be parsed                        */
#define UC_ERR_COMPILE 3    /* hipRTC compilation failed (log in uc_last_error)       */
#define UC_ERR_HIP 4        /* HIP runtime error                                      */
#define UC_ERR_NO_DEVICE 5  /* engine created without a device                        */
#define UC_ERR_CANCELLED 6  /* serving: job cancelled                                 */
#define UC_ERR_BUSY 7       /* serving: job not finished yet                          */

/* ---- precision policies ---- */
#define UC_PREC_F64 0   /* everything in double                                   */
#define UC_PREC_F32C 1  /* float stages, Kahan-compensated state, double GS/logs   */

/* ---- sweep axis kinds (match UcJob) ---- */
#define UC_AXIS_PARAM 1
#define UC_AXIS_VAR 2
#define UC_AXIS_STEP 3

/* ---- Lyapunov methods ---- */
#define UC_METHOD_VARIATIONAL 0
#define UC_METHOD_FINITE_DIFF 1

/* ---- integrator schemes for uc_codegen_scheme (order of U-CUDA/codegen.hpp:21) ---- */
#define UC_SCHEME_EULER 0
#define UC_SCHEME_EULER_CROMER 1
#define UC_SCHEME_MIDPOINT 2
#define UC_SCHEME_RK4 3
#define UC_SCHEME_DOPRI78 4 /* Dormand-Prince 8(7), 13 stages, fixed step */
#define UC_SCHEME_CD 5      /* composition explicit / diagonally implicit halves, symmetry a[0] */

typedef struct uc_engine uc_engine;
typedef struct uc_system uc_system;
typedef struct uc_plan uc_plan;

/* System definition (U-CUDA/codegen.hpp:10-15). vars -> X[0..N-1],
 * params -> a[1..M], a[0] reserved. time_symbol: name of the independent
 * variable usable in rhs (NULL or "" for autonomous systems). */
typedef struct uc_system_desc {
    const char* const* vars;
    int n_vars;
    const char* const* params;
    int n_params;
    const char* const* rhs;
    int latex;
    const char* time_symbol;
} uc_system_desc;

/* One sweep axis. Values follow the reference kernel's grid
 * (getValueByIdx / getValueByIdx_log, U-CUDA/cudaLibrary.cu:1344-1396):
 * linear (1-t)*lo + t*hi with t = i/(n-1), n == 1 -> hi; log scale
 * 10^(log10 lo + i*(log10 hi - log10 lo)/(n-1)). */
typedef struct uc_axis {
    int kind;    /* UC_AXIS_*                                   */
    int index;   /* a[] index (1-based) or state index (0-based) */
    int n_pts;
    double lo, hi;
    int log_scale;
} uc_axis;

/* Fields shared by every analysis request (Bifurcation1DRequest,
 * LLE2DRequest, LS2DRequest: U-CUDA/parametric_engine.h:27-95,411-455,581-619). */
typedef struct uc_common {
    const double* initial_conditions; /* host, n_vars values                        */
    const double* base_values;        /* host, a[0..n_values-1]                     */
    int n_values;
    int n_axes;                       /* 1 or 2                                     */
    uc_axis axis[2];
    double h;
    double transient_time;
    double t_max;
    double max_value;                 /* 0 disables the magnitude check             */
    int precision;                    /* UC_PREC_*                                  */
    int strict;                       /* 1: no FMA contraction (bit-reproducible)   */
    /* optional per-cell random initial state X0 = lo + (hi-lo)*u, seed 0 = off */
    unsigned long long ic_seed;
    double ic_lo, ic_hi;
    /* optional contiguous range of the grid's cells (cell = iy * nx + ix):
     * cell_count > 0 runs cells [cell_begin, cell_begin + cell_count) with
     * the same values, seeds and results they have in the whole grid, and
     * the outputs hold those cells only (plane stride cell_count). 0 = the
     * whole grid. Honoured by uc_run_bif, uc_run_lle, uc_run_ls and
     * uc_run_lle_noalloc (a caller runs a grid in pieces, e.g. to check a
     * cancel flag or report progress between them); the other entry points
     * reject cell_count != 0. */
    long long cell_begin, cell_count;
} uc_common;

/* Bifurcation sweep (run_bifurcation_1d / _2d, parametric_engine.cpp:1674-2091). */
typedef struct uc_bif_request {
    uc_common c;
    int writable_var;   /* sampled state index, -1 = x0 + pi*x1 + e*x2 */
    int pre_scaller;
    int max_peaks;      /* stored peaks per cell (reference cap 2500) */
    /* precision of the recorded run when c.precision is UC_PREC_F32C:
     * UC_PREC_F64 (0, the default) continues the compensated state in
     * double for the recorded run, so the peaks are double-precision values
     * (float stages displace a weakly contracting cycle by ~1e-6 relative);
     * UC_PREC_F32C records in float stages as well (faster). Ignored for
     * UC_PREC_F64 requests. */
    int record_precision;
} uc_bif_request;

typedef struct uc_bif_output {
    int* flags;          /* [cells] peak count, -1 fixed point / bad step, 0 diverged */
    double* peaks;       /* [max_peaks][cells]                                         */
    double* intervals;   /* [max_peaks][cells] inter-peak intervals                   */
    double* traj;        /* optional [traj_len][cells] first recorded samples          */
    int traj_len;
    double* final_state; /* optional [n_vars][cells]                                   */
} uc_bif_output;

/* Lyapunov exponent maps (run_lle_1d/_2d, run_ls_1d/_2d). */
typedef struct uc_lyap_request {
    uc_common c;
    double NT;                       /* renormalisation interval (time units)  */
    double eps;                      /* finite-difference probe distance       */
    int method;                      /* UC_METHOD_*                            */
    unsigned long long probe_seed;   /* reference: 1234567891                  */
} uc_lyap_request;

typedef struct uc_lyap_output {
    double* values;      /* [n_exponents][cells]; 999 diverged, NaN bad step */
    int* flags;          /* [cells] 1 ok, -1 diverged / bad step             */
    double* final_state; /* optional [n_vars][cells]                          */
} uc_lyap_output;

/* Basins of attraction with adaptive Dormand-Prince 5(4) and Poincare-section
 * attractor signatures (new method; the reference clusters features with
 * DBSCAN, U-CUDA/parametric_engine.cpp:6035-6514). */
typedef struct uc_basins_request {
    uc_common c;                 /* axes sweep initial conditions; h = first step */
    double rtol, atol;
    double h_min, h_max;
    double section_period;       /* e.g. 2*pi/omega                               */
    int n_sections;              /* periods integrated in total                  */
    int n_record;                /* trailing sections used for the signature     */
    int max_period;              /* longest cycle searched                       */
    int label_vars;              /* leading state variables in the signature     */
    double period_tol;           /* closure tolerance                            */
    double quantum;              /* signature quantisation                        */
    long long max_attempts;      /* per-cell step cap (0 = default)              */
} uc_basins_request;

typedef struct uc_basins_output {
    int* labels;          /* [cells] attractor label: 1..K canonical order, 0 unbound / unresolved */
    int* periods;         /* [cells] period p >= 1, 0 no cycle found, -1 unbound, -2 unresolved  */
    double* cycle_point;  /* optional [label_vars][cells] canonical cycle point (device)          */
    long long* steps;     /* optional [2][cells] accepted / rejected steps (device)               */
    /* attractor catalogue (HOST arrays, capacity max_attractors): label k+1 is
     * attractor k with its period and the quantised canonical point of its
     * lowest key; clusters are ordered by (period, point), no-cycle last. */
    int max_attractors;
    int* n_attractors;
    int* attractor_period;
    double* attractor_point;  /* [max_attractors][2] */
} uc_basins_output;

/* ---- engine ---- */
/* device < 0 creates a compile-only engine (no HIP device is touched).
 * Batches and servers keep their engine alive: uc_engine_destroy may be
 * called before them and the engine is freed with the last of them. */
int uc_engine_create(int device, uc_engine** out);
void uc_engine_destroy(uc_engine* e);
const char* uc_last_error(const uc_engine* e);
const char* uc_version(void);

/* ---- systems / codegen (U-CUDA/codegen.hpp:25-48) ---- */
int uc_system_create(uc_engine* e, const uc_system_desc* desc, uc_system** out);
/* A caller-supplied KRS step body (the reference's custom schemes and any
 * codegen_scheme output, U-CUDA/system_record.h:10-13, krs_cpu.cpp:293-316):
 * statements over numb X[n_vars], const numb a[n_values], numb h. Usable by
 * uc_run_bif and by uc_run_lle / uc_run_ls with UC_METHOD_FINITE_DIFF.
 * Constant out-of-range X[k] / a[k] subscripts are rejected (UC_ERR_PARSE). */
int uc_system_create_krs(uc_engine* e, const char* krs_body, int n_vars, int n_values, uc_system** out);
void uc_system_destroy(uc_system* s);
int uc_system_n_vars(const uc_system* s);
int uc_system_n_values(const uc_system* s);
/* Copies generated source (NUL-terminated) into buf; *len gets the full size.
 * dialect 0: the complete hipRTC program; 1: the system functions as C. */
int uc_system_source(const uc_system* s, int dialect, char* buf, size_t cap, size_t* len);
/* Compiles the system's module with hipRTC for gfx950 (no device needed);
 * on an engine with a device the module is also loaded onto it.
 * *code_size receives the code-object size in bytes. */
int uc_system_compile(uc_engine* e, uc_system* s, int strict, size_t* code_size);
/* Algorithmic flop counts of one right-hand side / Jacobian evaluation. */
int uc_system_op_counts(const uc_system* s, int* rhs_flops, int* rhs_calls, int* jac_nonzeros, int* jac_flops);
/* codegen_scheme: KRS step body text for a scheme (UC_SCHEME_*). */
int uc_codegen_scheme(uc_engine* e, const uc_system_desc* desc, int scheme, char* buf, size_t cap, size_t* len);
/* Host evaluation of the right-hand side in double (SystemEvaluator::eval). */
int uc_system_eval_rhs(const uc_system* s, const double* X, const double* a, double t, double* dX);

/* ---- analyses: enqueue on `stream`, return when enqueued ---- */
int uc_run_bif(uc_engine* e, uc_system* s, const uc_bif_request* r, const uc_bif_output* o, void* stream);
/* 2-D bifurcation map (U-CUDA/parametric_engine.cpp:5487-5911, run_bif2d):
 * the bifurcation sweep of `r` (max_peaks ignored: every peak of the
 * recorded window is kept, in chunks of cells that fit a quarter of free
 * HBM), then each cell's peak values are clustered by the reference's
 * DBSCAN with radius eps (cudaLibrary.cu:2088-2166). periods[cells]
 * (device): -1 fixed point / bad step, 0 no peaks or diverged, 1 one peak,
 * else the cluster count (the period). Synchronises the stream. */
int uc_run_bif_periods(uc_engine* e, uc_system* s, const uc_bif_request* r, double eps_dbscan, int* periods,
                       void* stream);

/* 1-D DFT sweep (U-CUDA/parametric_engine.h:140-226, run_dft1d_classical
 * parametric_engine.cpp:4809-5120): per cell a transient of
 * (int)(transient_time / h) steps, a block of N = (int)(t_max / h / pre)
 * samples of the sampled variable, a window, and the coefficients
 * a_k + i b_k = (1/N) sum_n w_n x_n (cos, sin)(2 pi (h pre) f_k n) at n_freq
 * frequencies (DFT_custom, cudaLibrary.cu:1698-1777, including its
 * float-precision phase resync every 1000 samples). Axes may sweep
 * parameters or initial conditions (not h). Cells run in chunks whose blocks
 * fit a quarter of free HBM; the call synchronises the stream. */
typedef struct uc_dft_request {
    uc_common c;
    int writable_var;    /* sampled variable (-1: x0 + pi x1 + e x2)            */
    int pre_scaller;     /* steps per sample                                      */
    int n_freq;          /* >= 2                                                  */
    double freq_lo, freq_hi;
    int freq_log_scale;  /* 1: log-spaced frequencies (freq_lo, freq_hi > 0)      */
    int window_type;     /* 0 none, 1 Hanning, 2 Hamming                          */
} uc_dft_request;
typedef struct uc_dft_output {
    double* ak_cos;      /* [cells][n_freq] device; -1 / 0 for flagged blocks     */
    double* bk_sin;
    int* flags;          /* [cells] 1 ok, -1 fixed point / bad step, 0 diverged   */
} uc_dft_output;
int uc_run_dft(uc_engine* e, uc_system* s, const uc_dft_request* r, const uc_dft_output* o, void* stream);
/* Continuation sweeps (the reference's continuation mode of run_bifurcation_1d
 * / run_lle_1d / run_ls_1d / run_dft_1d, U-CUDA/parametric_engine.cpp:
 * 3983-4742; kernels/bifurcation1d_cont, lle1d_cont, ls1d_cont,
 * dft1d_cont .template.cu): axis 0 (a parameter or h) is visited point by
 * point, each point starting from the state (and probes) the previous one
 * left; t = j / (n - 1), point j = lo + (hi - lo) t, or hi - (hi - lo) t with
 * `reverse` (log: in log10). Outputs are indexed by visiting order, cell =
 * chain * n_pts + j. An optional axis 1 gives independent chains (one GPU
 * lane each; the reference runs one chain). FP64 only; synchronises.
 *  bif: block (int)(t_max / h / pre) samples (capped at ceil(t_max / h_min
 *    / pre)), transient (int)(transient / h) steps; flags = peak count, or -1
 *    for an empty block, an escaping transient (the state is carried on) or
 *    a block ending on a fixed point, 0 for a block that escapes part way;
 *    peak intervals use the base h.
 *  lle / ls: the reference's finite-difference Benettin chain with probes
 *    carried between points (method and precision fields are ignored);
 *    values [n_exponents][cells] = sum / t_max, NaN with flag -1 when the
 *    point escaped (the chain then restarts from the initial state).
 *  dft: a point whose block holds <= 2 samples, whose transient escapes or
 *    whose block ends on a fixed point is -1 everywhere and restarts the
 *    chain; a block that escapes part way is 0 everywhere and restarts it;
 *    the others get flag 1 and the coefficients of their block. */
int uc_run_bif_cont(uc_engine* e, uc_system* s, const uc_bif_request* r, int reverse, const uc_bif_output* o,
                    void* stream);
int uc_run_lle_cont(uc_engine* e, uc_system* s, const uc_lyap_request* r, int reverse, const uc_lyap_output* o,
                    void* stream);
int uc_run_ls_cont(uc_engine* e, uc_system* s, const uc_lyap_request* r, int reverse, const uc_lyap_output* o,
                   void* stream);
int uc_run_dft_cont(uc_engine* e, uc_system* s, const uc_dft_request* r, int reverse, const uc_dft_output* o,
                    void* stream);
int uc_run_lle(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream);
int uc_run_ls(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream);
int uc_run_basins(uc_engine* e, uc_system* s, const uc_basins_request* r, const uc_basins_output* o, void* stream);
/* uc_run_lle without allocating. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The launch tables go into a
 * static 512 KiB device array per GPU instead of a buffer from the
 * allocator, so the grid is limited to nx + ny <= 65,440 axis points (else
 * UC_ERR_ARGUMENT). The system's module must already be loaded on this
 * engine (uc_system_compile with the same strict flag; else
 * UC_ERR_ARGUMENT). Same kernels as uc_run_lle, same results. Every call on
 * one GPU shares that GPU's array, which keeps the last call's tables: calls
 * that overlap on one GPU race. */
int uc_run_lle_noalloc(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream);

/* The streaming peak finder over stored series (the reference's
 * peakFinderCUDA over d_data, U-CUDA/cudaLibrary.cu:1641-1674): device
 * data[n_samples][n_series]; counts[n_series], peaks / intervals
 * [max_peaks][n_series]; intervals scaled by time_step. */
int uc_peak_finder(uc_engine* e, uc_system* s, const double* data, long long n_samples, long long n_series,
                   double time_step, int max_peaks, int* counts, double* peaks, double* intervals, void* stream);

/* Fast synchronisation (U-CUDA/parametric_engine.h:767-842 FastSyncRequest /
 * FastSyncResult, run_fastsync parametric_engine.cpp:6833-7255): a slave copy
 * of the system is pulled towards a master by N = K (master - slave), over a
 * window forward with k_forward and back with k_backward, iter_of_synchr
 * times. FP64; `strict` as elsewhere.
 *  mode 0 (on attractor): the master is integrated from ic_master for
 *    (int)(transient / h) steps, then kept for (int)(t_max / h) +
 *    (int)(window / h) samples; point idx = 0 .. n_out - 1 with n_out =
 *    (int)(t_max / h) / pre_scaller synchronises the slave (from ic_slave)
 *    over the (int)(window / h) samples that start at sample idx *
 *    pre_scaller. sync_error[n_out]; traj (optional) [n_out][n_vars] the
 *    master at those starts.
 *  mode 1 (on grid): n_pts x n_pts cells (iy * n_pts + ix) set state
 *    variables axis_x_var = grid x, axis_y_var = grid y of the master
 *    (grid_swap_master_slave 0) or of the slave (1); the other, fixed
 *    initial state first runs (int)(transient / h) uncoupled steps; windows
 *    of (int)(window / h / pre_scaller) samples. sync_error[n_pts * n_pts].
 *  error_estim: 0 RMS over the last forward pass, 1 passes until the error
 *  at the window's end is <= fs_error_trs, 2 error at the window's end after
 *  the last pass, 3 time until the error is <= fs_error_trs; 0 / 2 report
 *  log10 (fs_error_trs floor, 3 for errors >= 1e3 or non-finite).
 *  type_of_synch: 0 unidirectional (stored master), 1 bidirectional. */
typedef struct uc_fastsync_request {
    const double* values;        /* a[] */
    int n_values;
    int mode;
    double h;
    int iter_of_synchr;
    int pre_scaller;
    const double* k_forward;     /* n_vars each */
    const double* k_backward;
    const double* ic_master;
    const double* ic_slave;
    double t_max;                /* mode 0 */
    double transient_time;
    double window;
    int axis_x_var, axis_y_var;  /* mode 1 */
    double axis_x_lo, axis_x_hi, axis_y_lo, axis_y_hi;
    int n_pts;
    int grid_swap_master_slave;
    int type_of_synch;
    int error_estim;
    double fs_error_trs;
    int strict;
} uc_fastsync_request;
typedef struct uc_fastsync_output {
    double* sync_error;
    double* traj;                /* mode 0 only, may be NULL */
} uc_fastsync_output;
/* Number of sync_error values a request produces (mode 0: (int)(t_max / h)
 * / pre_scaller; mode 1: n_pts^2); < 0 for an invalid request. */
long long uc_fastsync_points(const uc_fastsync_request* r);
int uc_run_fastsync(uc_engine* e, uc_system* s, const uc_fastsync_request* r, const uc_fastsync_output* o,
                    void* stream);

/* Periods from stored peaks (the reference's dbscanCUDA over its peak
 * buffers, U-CUDA/cudaLibrary.cu:2321-2345, with dbscan :2088-2166): device
 * peaks[max_peaks][n_cells] (peak-major, as uc_run_bif writes them) and
 * counts[n_cells] (-1 / 0 / number of peaks; a count above max_peaks is read
 * as max_peaks); periods[n_cells] as uc_run_bif_periods. Synchronises the
 * stream when max_peaks > 1024 (scratch for those cells is allocated). */
int uc_bif_dbscan_periods(uc_engine* e, uc_system* s, const double* peaks, const int* counts, long long n_cells,
                          int max_peaks, double eps, int* periods, void* stream);

/* ---- persistent heterogeneous batch ----
 * Jobs of any size, system and analysis, added with the same request /
 * output structs as the uc_run_* calls; jobs that share a compiled kernel
 * run as ONE persistent launch over a job table (workgroups take chunks of
 * cells from a global counter, longest jobs first), and all such launches
 * are captured into one hipGraph that uc_batch_launch replays. The modules
 * stay cached in the engine while the batch lives (LRU capacity 8). */
typedef struct uc_batch uc_batch;
int uc_batch_create(uc_engine* e, uc_batch** out);
void uc_batch_destroy(uc_batch* b);
int uc_batch_add_bif(uc_batch* b, uc_system* s, const uc_bif_request* r, const uc_bif_output* o, int* job_id);
int uc_batch_add_lyap(uc_batch* b, uc_system* s, const uc_lyap_request* r, int spectrum, const uc_lyap_output* o,
                      int* job_id);
/* Basin maps (as uc_run_basins): the adaptive integration runs in the
 * batch's graph; the attractor catalogue and the labels are computed after
 * each uc_batch_launch, which then synchronises its stream. The catalogue
 * arrays of `o` are host memory that must stay valid while the batch
 * lives; cycle_point may be NULL (the batch keeps a buffer). */
int uc_batch_add_basins(uc_batch* b, uc_system* s, const uc_basins_request* r, const uc_basins_output* o,
                        int* job_id);
/* A batch plan: the launch shape, the trajectory-steps each cell runs per
 * launch (0: whole runs; else jobs whose cells all step the same counts run
 * in stages of about that many steps, their cells' state kept in HBM
 * between stages) and the cells per lane (2: LLE maps that can run two
 * cells per lane do). Results do not depend on the plan. */
typedef struct uc_batch_plan {
    int block;             /* lanes per workgroup: 64, 128 or 256 */
    int waves_per_simd;    /* resident waves per SIMD, 1..8 */
    double slice_steps;    /* trajectory-steps per cell per launch, 0 = whole runs */
    int cells_per_lane;    /* 1 or 2 */
    int vector_lanes;      /* 1: spectrum maps run one lane per tangent vector
                              (N lanes per cell); 0: one lane per cell */
    double ms;             /* set by the tuner: the plan's measured time */
} uc_batch_plan;
int uc_batch_set_plan(uc_batch* b, const uc_batch_plan* p);
int uc_batch_get_plan(const uc_batch* b, uc_batch_plan* p);
/* autotuner: times the batch for each candidate plan (launch shapes with
 * whole runs, then steps per launch, cells per lane and lanes per tangent
 * vector at the best shape) and keeps the fastest */
int uc_batch_tune_plan(uc_batch* b, void* stream, uc_batch_plan* best);
/* launch shape only: block (64/128/256 lanes) and resident waves per SIMD (1..8) */
int uc_batch_set_launch(uc_batch* b, int block, int waves_per_simd);
/* uc_batch_tune_plan, returning the shape */
int uc_batch_tune(uc_batch* b, void* stream, int* best_block, int* best_waves_per_simd, double* best_ms);
/* one graph launch of the whole batch on a non-default stream (then the
 * basins jobs' catalogues, synchronously, when the batch has any) */
int uc_batch_launch(uc_batch* b, void* stream);
int uc_batch_info(const uc_batch* b, int* n_jobs, int* n_groups, int* block, int* waves_per_simd);

/* ---- serving: continuous batching ----
 * A scheduler thread admits jobs in arrival order while their outputs fit
 * the HBM budget and keeps up to six rounds in flight, one per stream:
 * each round takes cells of one duration class (the decade of the steps a
 * cell runs in the round) from the admitted jobs (a fair share per job,
 * then oldest first) up to the GPU's resident lanes and launches the
 * persistent kernels once per (system, analysis) group; a new round is
 * queued whenever a slot frees, and jobs submitted meanwhile join it. A job
 * of more than slice_min_cells cells and more than slice_steps steps per
 * cell runs in slices: each round advances its cells by about slice_steps
 * steps and keeps their state in HBM, so no round holds the GPU for long.
 * The waves of small grids run at a higher issue priority. The shortest
 * class goes first when none of its rounds is in flight, the others take
 * turns, and classes of >= 1e4 steps per cell hold at most three of the six
 * slots. Results are bit-identical to a direct uc_run_* of the same
 * request. Outputs stay on the device until uc_server_fetch copies them
 * out. Run the process with GPU_MAX_HW_QUEUES=8 (before the HIP runtime
 * starts): with HIP's default of 4 the six round streams share two
 * hardware queues and rounds on one queue run in turn. */
typedef struct uc_server uc_server;
typedef struct uc_server_config {
    double hbm_fraction;       /* admission budget as a fraction of free HBM (0 -> 0.5) */
    long long max_bytes;       /* explicit budget in bytes (overrides hbm_fraction if > 0) */
    double round_traj_steps;   /* optional cap on a round's trajectory-steps (0 -> none) */
    int block;                 /* lanes per workgroup (0 -> 256)                          */
    int waves_per_simd;        /* resident waves per SIMD (0 -> 8)                        */
    double slice_steps;        /* trajectory-steps per cell per round for jobs whose cells */
                               /* all step the same counts: longer jobs resume over several */
                               /* rounds, their cells' state kept in HBM (0 -> 1024,      */
                               /* < 0 -> whole runs)                                      */
    long long slice_min_cells; /* jobs of at most this many cells run whole (0 -> 65536,  */
                               /* < 0 -> any)                                             */
} uc_server_config;
typedef struct uc_server_stats_t {
    long long completed, rounds, queued, active;
    double p50_ms, p99_ms, max_ms, busy_s, wall_s, traj_steps, used_bytes, budget_bytes;
} uc_server_stats_t;
int uc_server_create(uc_engine* e, const uc_server_config* cfg, uc_server** out);
void uc_server_destroy(uc_server* s);
int uc_server_submit_lyap(uc_server* s, uc_system* sys, const uc_lyap_request* r, int spectrum, long long* job_id);
int uc_server_submit_bif(uc_server* s, uc_system* sys, const uc_bif_request* r, long long* job_id);
/* Basin maps: the adaptive integration runs in the rounds like the other
 * analyses; the attractor catalogue and the labels are computed when the
 * job's last cell is done. Fetch with uc_server_fetch_basins. */
int uc_server_submit_basins(uc_server* s, uc_system* sys, const uc_basins_request* r, long long* job_id);
/* state: 0 queued, 1 running, 2 done, 3 cancelled, 4 failed, 5 fetched */
int uc_server_poll(uc_server* s, long long job_id, int* state, double* latency_ms);
int uc_server_wait(uc_server* s, long long job_id, double timeout_s);
int uc_server_cancel(uc_server* s, long long job_id);
int uc_server_fetch(uc_server* s, long long job_id, int* flags, double* out0, double* out1);
/* host pointers: labels, periods (required), cycle_point, steps, catalogue (optional) */
int uc_server_fetch_basins(uc_server* s, long long job_id, const uc_basins_output* o);
int uc_server_stats(uc_server* s, uc_server_stats_t* stats);

/* Number of cells a request runs: its cell range, else the whole grid
 * (product of axis point counts). */
long long uc_grid_cells(const uc_common* c);
/* Axis values exactly as the kernels use them (host, n_pts doubles). */
int uc_axis_values(const uc_axis* ax, double* out);

#ifdef __cplusplus
}
#endif
#endif
