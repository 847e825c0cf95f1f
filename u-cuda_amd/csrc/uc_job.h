/* uc_job.h — launch descriptor shared by the host engine and the run-time
 * compiled HIP kernels (this text is pasted verbatim into every hipRTC
 * program, so it must stay plain C with fixed-size members only).
 *
 * One descriptor describes one ensemble: a grid of cells, each cell one
 * trajectory. Axis 0 varies fastest: cell = iy * nx + ix, the row-major
 * layout of the reference's 2-D results (values[iy * n + ix],
 * U-CUDA/parametric_engine.h:468-470) and of its kernel index mapping
 * (getValueByIdx, U-CUDA/cudaLibrary.cu:1344-1361).
 */
#ifndef UC_JOB_H
#define UC_JOB_H

#define UC_AXIS_NONE 0
#define UC_AXIS_PARAM 1   /* sweeps a[index] (1-based, a[0] reserved)   */
#define UC_AXIS_VAR 2     /* sweeps the initial condition X0[index]      */
#define UC_AXIS_STEP 3    /* sweeps the integration step h               */

#define UC_METHOD_VARIATIONAL 0  /* tangent-space variational equations    */
#define UC_METHOD_FINITE_DIFF 1  /* reference finite-difference Benettin   */

#define UC_LABEL_UNBOUND 0

typedef struct UcJob {
    /* ---- grid ---- */
    long long n_cells;            /* cells handled by this entry              */
    long long cell_offset;        /* global index of the entry's first cell   */
    long long stride;             /* plane stride of the SoA outputs          */
    int nx, ny;                   /* points on axis 0 / axis 1 (ny = 1: 1-D)  */
    int axis_kind[2];             /* UC_AXIS_*                                */
    int axis_index[2];
    const double* axis_vals[2];   /* device tables, nx resp. ny values        */
    const double* base_params;    /* a[0..n_values-1]                         */
    const double* base_ic;        /* X0[0..N-1]                               */
    int n_values;
    int pre_scaller;
    /* per-cell random initial state: X0[i] = ic_lo + (ic_hi - ic_lo) * u    */
    unsigned long long ic_seed;   /* 0 = use base_ic                          */
    double ic_lo, ic_hi;

    /* ---- integration ---- */
    double h;
    /* h / 2 and h / 6 in double, and (float)h, its half and sixth in float,
     * computed once on the host: kernels whose step is the same on every
     * lane read them as kernel arguments, i.e. from scalar registers       */
    double hh, h6;
    float hf, hhf, h6f;
    double max_value;             /* 0 disables the magnitude check           */
    float max_value_f;            /* the largest float <= max_value (float    */
                                  /* x > max_value_f iff (double)x > max_value) */
    long long skip_steps;         /* transient steps when no axis sweeps h    */
    long long n_iters;            /* bif: recorded samples; lle/ls: blocks    */
    long long nt_steps;           /* lle/ls: steps per renormalisation block  */
    double transient_time, t_max, nt;  /* per-lane recount under an h axis    */
    double eps;                   /* finite-difference probe distance         */
    unsigned long long probe_seed;/* tangent / probe basis seed               */
    int method;                   /* UC_METHOD_*                              */
    int writable_var;             /* bif: sampled variable (-1: x0+pi*x1+e*x2)*/
    int max_peaks;                /* bif: stored peaks per cell               */
    int traj_len;                 /* bif: samples copied to traj (0 = none)   */
    int record_f64;               /* bif, F32C: the recorded run in double    */
    int raw_flags;                /* bif: 0 flag = peak count; 1 flag = loop  */
                                  /* flag (1/-1/0); 2 the same under the DFT  */
                                  /* h-sweep rules (dft1dHSweepKernel)        */

    /* ---- adaptive integration / basins ---- */
    double rtol, atol;            /* Dormand-Prince 5(4) error control        */
    double h_min, h_max;
    double section_period;        /* Poincare section every section_period    */
    int n_sections;               /* sections integrated in total             */
    int n_record;                 /* trailing sections kept for the signature */
    int max_period;               /* longest cycle searched for               */
    int label_vars;               /* state variables used by the signature    */
    double period_tol;            /* cycle closure tolerance                  */
    double quantum;               /* signature quantisation step              */
    long long max_attempts;       /* per-cell cap on attempted steps          */

    /* ---- outputs (device pointers, structure of arrays) ---- */
    int* flag;                    /* per cell                                 */
    double* out0;                 /* bif peaks / lle / ls planes / basin sig. */
    double* out1;                 /* bif intervals / basin cycle point        */
    double* traj;                 /* bif: traj[i * n_cells + cell]            */
    double* final_state;          /* final_state[i * n_cells + cell]          */
    long long* counters;          /* per cell: accepted, rejected steps       */
    int* label;                   /* basins: int32 attractor signature hash   */

    /* ---- resumable slices (persistent kernels, wave-uniform counts) ---- */
    double* slab;                 /* cell state between launches,             */
                                  /* slab[w * stride + cell]; NULL: whole run */
    long long unit_begin;         /* work units run by this launch:           */
    long long unit_end;           /* [unit_begin, unit_end)                   */
    int priority;                 /* persistent kernels: issue priority of    */
                                  /* the entry's waves, 0..3 (s_setprio)      */
} UcJob;

/* Work units of a sliced cell: transient (and bifurcation record) steps in
 * chunks of UC_UNIT_STEPS, the magnitude-check interval, so the checks fall
 * on the same steps as in a whole run; one unit per renormalisation block
 * (LLE / LS) or per section (basins). Slab words per cell for N state
 * variables (and K tangent vectors): */
#define UC_UNIT_STEPS 100
#define UC_SLAB_LYAP_WORDS(N, K) (8 + 3 * (K) + 3 * (N) + 2 * (K) * (N))
#define UC_SLAB_BIF_WORDS(N) (10 + 2 * (N))
#define UC_SLAB_BAS_WORDS(N) (40 + 2 * (N))

/* Fast synchronisation (the reference's FastSynchro): a slave copy of the
 * system is pulled towards a master by N = K (master - slave), integrated
 * forward with k_forward over a window and back with k_backward,
 * iter_of_synchr times; the sync error is the map value. */
typedef struct UcFsJob {
    long long n_cells;            /* windows (on attractor) / grid cells      */
    long long cell_offset;        /* global index of the first grid cell      */
    long long stride;             /* plane stride of the per-cell master data */
    int n_pts;                    /* grid side                                */
    int axis_var[2];              /* state variable swept by each grid axis   */
    const double* axis_vals[2];   /* device tables, n_pts values each         */
    int swap;                     /* grid: 0 sweeps the master IC, 1 slave IC */
    const double* params;         /* a[] (UC_M slots)                         */
    const double* ic_master;      /* N values each                            */
    const double* ic_slave;
    const double* k_forward;
    const double* k_backward;
    double h;
    int n_iters;                  /* window samples (amountOfIterations)      */
    int iter_of_synchr;
    int pre_scaller;              /* on attractor: window start = idx * pre   */
    int type_of_synch;            /* 0 unidirectional, 1 bidirectional        */
    int error_estim;              /* 0 RMS on the last pass, 1 passes to      */
                                  /* reach fs_error_trs, 2 error at the last  */
                                  /* point, 3 time to reach fs_error_trs      */
    double fs_error_trs;
    const double* master;         /* on attractor: master[i * N + j]          */
    double* data;                 /* grid: master[(i * N + j) * stride + lane]*/
    double* out;                  /* sync error per cell                      */
    double* traj_out;             /* on attractor: decimated master [idx * N + j], optional */
} UcFsJob;

#endif
