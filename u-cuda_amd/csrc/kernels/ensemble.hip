// ensemble.hip — hand-written CDNA4 ensemble kernels, compiled at run time by
// hipRTC for gfx950 after the generated system functions (uc_rhs, uc_tangent,
// UC_N, UC_M) and the UcJob descriptor.
//
// One trajectory per lane. State, tangent vectors and parameters live in
// VGPRs (every array below is indexed by compile-time constants after
// unrolling); uniform parameters are read with scalar loads; results are
// written structure-of-arrays so a wavefront's stores are contiguous.
//
// Two precision policies:
//   F64  — everything in double.
//   F32C — Runge-Kutta stages in float; the state is a float pair (hi, c) with
//          Kahan-compensated updates; classification, peak interpolation,
//          Gram-Schmidt, norms and logarithm sums are done in double.
//
// The step loop, its divergence / fixed-point classification and the peak
// finder reproduce the reference semantics (U-CUDA/cudaLibrary.cu:782-895,
// 944-1077, 1503-1674, 2486-2693, 2832-3073); see each function.

#define UC_CHECK_INTERVAL 100              // CHECK_INTERVAL, U-CUDA/configCUDA.h:12
#define UC_EPS_FIXED_POINT 1e-8            // eps_fixed_point, configCUDA.h:84
#define UC_EPS_PEAK_DELTA 1e-14            // eps_peak_delta, configCUDA.h:85
#define UC_PEAK_THRESHOLD (-1e25)          // peak_threshold, configCUDA.h:86
#define UC_PI 3.1415926535897932384626433832795
#define UC_E 2.7182818284590452353602874713527
#define UC_BLOCK 256

typedef unsigned long long u64;

// ------------------------------------------------------------------ RNG ----
// splitmix64 seeding + 64-bit LCG draw, the generator the reference uses for
// its probe directions (U-CUDA/kernels/lle1d.template.cu:39-57).
struct UcRng {
    u64 s;
    __device__ __forceinline__ void init(u64 seed, u64 seq, u64 off) {
        u64 z = seed + seq * 0x9E3779B97F4A7C15ULL + off;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s = z ^ (z >> 31);
    }
    __device__ __forceinline__ float uniform() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (float)(((s >> 40) & 0xFFFFFFULL) + 1ULL) / 16777216.0f;
    }
};

// ------------------------------------------------------------ cell setup ---
__device__ __forceinline__ double uc_axis_value(const UcJob& j, int ax, long long cell) {
    long long i = (ax == 0) ? (cell % j.nx) : ((cell / j.nx) % j.ny);
    return j.axis_vals[ax][i];
}

// Fills the cell's parameters, initial state and step. Returns false when
// the cell's step is not positive (the reference marks such cells unusable).
__device__ __forceinline__ bool uc_setup(const UcJob& j, long long cell, double* a, double* x0, double& h) {
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = (i < j.n_values) ? j.base_params[i] : 0.0;
    if (j.ic_seed != 0ULL) {
        UcRng r;
        r.init(j.ic_seed, (u64)cell, 0ULL);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x0[i] = j.ic_lo + (j.ic_hi - j.ic_lo) * (double)r.uniform();
    } else {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x0[i] = j.base_ic[i];
    }
    h = j.h;
#pragma unroll
    for (int ax = 0; ax < 2; ++ax) {
        const int kind = j.axis_kind[ax];
        if (kind == UC_AXIS_NONE) continue;
        const double v = uc_axis_value(j, ax, cell);
        const int idx = j.axis_index[ax];
        if (kind == UC_AXIS_PARAM) {
#pragma unroll
            for (int i = 0; i < UC_M; ++i) a[i] = (i == idx) ? v : a[i];
        } else if (kind == UC_AXIS_VAR) {
#pragma unroll
            for (int i = 0; i < UC_N; ++i) x0[i] = (i == idx) ? v : x0[i];
        } else {
            h = v;
        }
    }
    return h > 0.0;
}

__device__ __forceinline__ bool uc_sweeps_step(const UcJob& j) {
    return j.axis_kind[0] == UC_AXIS_STEP || j.axis_kind[1] == UC_AXIS_STEP;
}

// --------------------------------------------------------- state policies --
// F64 state: plain doubles.
struct StateF64 {
    double x[UC_N];
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x[i] = x0[i];
    }
    __device__ __forceinline__ double get(int i) const { return x[i]; }
};

// F32C state: hi + Kahan compensation c (true value ~ hi - c).
struct StateF32C {
    float hi[UC_N], c[UC_N];
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) {
            hi[i] = (float)x0[i];
            c[i] = (float)((double)hi[i] - x0[i]);
        }
    }
    __device__ __forceinline__ double get(int i) const { return (double)hi[i] - (double)c[i]; }
};

__device__ __forceinline__ void uc_kahan_add(float& hi, float& c, float incr) {
    const float y = incr - c;
    const float t = hi + y;
    c = (t - hi) - y;
    hi = t;
}

// Classic RK4 with the reference's stage structure and update order
// (U-CUDA/codegen.cpp:597-606): X1 = X + 0.5*h*k for stages 0,1, X1 = X + h*k
// for stage 2, X += h*(k0 + 2*k1 + 2*k2 + k3)/6. Accumulating the bracket
// left to right keeps the reference's summation order.
__device__ __forceinline__ void uc_rk4(StateF64& s, const double* a, double h, double t) {
    double x1[UC_N], k[UC_N], acc[UC_N];
    const double hh = 0.5 * h;
    uc_rhs<double>(s.x, a, t, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.x[i] + hh * k[i]; }
    uc_rhs<double>(x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + hh * k[i]; }
    uc_rhs<double>(x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + h * k[i]; }
    uc_rhs<double>(x1, a, t + h, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
#if UC_STRICT
        s.x[i] = s.x[i] + h * acc[i] / 6.0;
#else
        s.x[i] = fma(h * (1.0 / 6.0), acc[i], s.x[i]);
#endif
    }
}

__device__ __forceinline__ void uc_rk4(StateF32C& s, const float* a, float h, float t) {
    float x1[UC_N], k[UC_N], acc[UC_N];
    const float hh = 0.5f * h;
    uc_rhs<float>(s.hi, a, t, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.hi[i] + hh * k[i]; }
    uc_rhs<float>(x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + hh * k[i]; }
    uc_rhs<float>(x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + h * k[i]; }
    uc_rhs<float>(x1, a, t + h, k);
    const float h6 = h * (1.0f / 6.0f);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
        uc_kahan_add(s.hi[i], s.c[i], h6 * acc[i]);
    }
}

// RK4 on the state together with K tangent vectors V (row-major K x N): the
// variational equations V' = J(X) V integrated with the same stages, which
// is exactly the derivative of the discrete RK4 map.
template <int K>
__device__ __forceinline__ void uc_rk4_tan(StateF64& s, double* V, const double* a, double h, double t) {
    double x1[UC_N], k[UC_N], acc[UC_N];
    double v1[K * UC_N], l[K * UC_N], lacc[K * UC_N];
    const double hh = 0.5 * h;
    uc_rhs<double>(s.x, a, t, k);
    uc_tangent<double, K>(s.x, a, t, V, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.x[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = l[i]; v1[i] = V[i] + hh * l[i]; }
    uc_rhs<double>(x1, a, t + hh, k);
    uc_tangent<double, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + hh * l[i]; }
    uc_rhs<double>(x1, a, t + hh, k);
    uc_tangent<double, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + h * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + h * l[i]; }
    uc_rhs<double>(x1, a, t + h, k);
    uc_tangent<double, K>(x1, a, t + h, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
#if UC_STRICT
        s.x[i] = s.x[i] + h * acc[i] / 6.0;
#else
        s.x[i] = fma(h * (1.0 / 6.0), acc[i], s.x[i]);
#endif
    }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) {
        lacc[i] = lacc[i] + l[i];
#if UC_STRICT
        V[i] = V[i] + h * lacc[i] / 6.0;
#else
        V[i] = fma(h * (1.0 / 6.0), lacc[i], V[i]);
#endif
    }
}

template <int K>
__device__ __forceinline__ void uc_rk4_tan(StateF32C& s, float* V, const float* a, float h, float t) {
    float x1[UC_N], k[UC_N], acc[UC_N];
    float v1[K * UC_N], l[K * UC_N], lacc[K * UC_N];
    const float hh = 0.5f * h;
    uc_rhs<float>(s.hi, a, t, k);
    uc_tangent<float, K>(s.hi, a, t, V, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.hi[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = l[i]; v1[i] = V[i] + hh * l[i]; }
    uc_rhs<float>(x1, a, t + hh, k);
    uc_tangent<float, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0f * l[i]; v1[i] = V[i] + hh * l[i]; }
    uc_rhs<float>(x1, a, t + hh, k);
    uc_tangent<float, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + h * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0f * l[i]; v1[i] = V[i] + h * l[i]; }
    uc_rhs<float>(x1, a, t + h, k);
    uc_tangent<float, K>(x1, a, t + h, v1, l);
    const float h6 = h * (1.0f / 6.0f);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
        uc_kahan_add(s.hi[i], s.c[i], h6 * acc[i]);
    }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + l[i]; V[i] = fmaf(h6, lacc[i], V[i]); }
}

// --------------------------------------------------- precision adaptors ---
template <class S> struct Prec;
template <> struct Prec<StateF64> {
    typedef double T;
    static __device__ __forceinline__ double to_t(double v) { return v; }
};
template <> struct Prec<StateF32C> {
    typedef float T;
    static __device__ __forceinline__ float to_t(double v) { return (float)v; }
};

// ------------------------------------------------------------ step loop ---
// The reference step loop (U-CUDA/cudaLibrary.cu:782-895): `iters`
// iterations, each optionally recording one sample before `pre` steps; the
// magnitude check sum|x| (NaN / Inf / > max_value -> 0) after iterations
// 0, 100, 200, ...; then `pre` trailing steps and the fixed-point test
// sum|x - x_prev| < 1e-8 -> -1, else 1.
struct NoObserver {
    static constexpr bool records = false;
    __device__ __forceinline__ void sample(long long, double) {}
};

template <class S>
__device__ __forceinline__ bool uc_diverged(const S& s, double max_value) {
    double chk = 0.0;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) chk = chk + fabs(s.get(i));
    if (isnan(chk) || isinf(chk)) return true;
    return max_value != 0.0 && fabs(chk) > max_value;
}

template <class S>
__device__ __forceinline__ int uc_fixed_point_flag(const S& s, const double* prev) {
    double d = 0.0;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) d = d + fabs(s.get(i) - prev[i]);
    return fabs(d) < UC_EPS_FIXED_POINT ? -1 : 1;
}

// Plain state loop.
template <class S, class Obs>
__device__ int uc_loop(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, long long iters, int pre,
                       int wvar, double max_value, Obs& obs) {
    int until_check = 0;
    for (long long i = 0; i < iters; ++i) {
        if (Obs::records) {
            double v;
            if (wvar < 0) {
                if (UC_N >= 3) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0) + UC_E * s.get(UC_N >= 3 ? 2 : 0);
                else if (UC_N == 2) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0);
                else v = s.get(0);
            } else {
                v = s.get(0);
#pragma unroll
                for (int q = 1; q < UC_N; ++q) v = (q == wvar) ? s.get(q) : v;
            }
            obs.sample(i, v);
            for (int p = 0; p < pre; ++p) uc_rk4(s, a, h, (typename Prec<S>::T)0);
        } else {
            uc_rk4(s, a, h, (typename Prec<S>::T)0);
        }
        if (until_check == 0) {
            if (uc_diverged(s, max_value)) return 0;
            until_check = UC_CHECK_INTERVAL;
        }
        --until_check;
    }
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    for (int p = 0; p < pre; ++p) uc_rk4(s, a, h, (typename Prec<S>::T)0);
    return uc_fixed_point_flag(s, prev);
}

// State + K tangent vectors advanced together for `iters` steps plus the
// trailing step, with the magnitude check on the state (variational
// counterpart of one reference renormalisation block).
template <int K, class S>
__device__ int uc_loop_tan(S& s, typename Prec<S>::T* V, const typename Prec<S>::T* a, typename Prec<S>::T h,
                           long long iters, double max_value) {
    int until_check = 0;
    for (long long i = 0; i < iters; ++i) {
        uc_rk4_tan<K>(s, V, a, h, (typename Prec<S>::T)0);
        if (until_check == 0) {
            if (uc_diverged(s, max_value)) return 0;
            until_check = UC_CHECK_INTERVAL;
        }
        --until_check;
    }
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    uc_rk4_tan<K>(s, V, a, h, (typename Prec<S>::T)0);
    return uc_fixed_point_flag(s, prev);
}

// ------------------------------------------------------ streaming peaks ---
// The reference peak finder (U-CUDA/cudaLibrary.cu:1503-1633) run as samples
// arrive instead of over a stored trajectory: candidate c needs d[c-1], d[c],
// d[c+1]; a plateau walk from c ends at the first j with d[j] < d[j+1] (no
// peak) or d[j] - d[j+1] > 1e-14 (peak at j, parabolic refinement); either
// way the next candidate is j + 2; a walk reaching n - 2 ends the search.
// The first peak only anchors the first interval (the reference's shift-left).
struct PeakObserver {
    static constexpr bool records = true;
    double w0, w1;          // d[k-2], d[k-1]
    long long n;            // samples in the record
    long long c_next;       // first admissible candidate
    int mode;               // 0 search, 1 plateau walk, 2 finished
    int raw;                // peaks found (before the shift)
    double last_t;
    double time_step;
    int max_peaks;
    long long stride;       // n_cells
    double* peaks;          // peaks[k * stride]
    double* intervals;
    double* traj;
    int traj_len;

    __device__ __forceinline__ void emit(double dm, double d0, double dp, long long j) {
        const double denom = dm - 2.0 * d0 + dp;
        double delta = 0.0;
        if (fabs(denom) > 1e-12) delta = 0.5 * (dm - dp) / denom;
        const double value = d0 - 0.25 * (dm - dp) * delta;
        const double tpk = (double)(j - 1) + delta;
        if (raw >= 1 && raw - 1 < max_peaks) {
            peaks[(long long)(raw - 1) * stride] = value;
            intervals[(long long)(raw - 1) * stride] = (tpk - last_t) * time_step;
        }
        last_t = tpk;
        ++raw;
    }

    __device__ __forceinline__ void sample(long long k, double v) {
        if (k < traj_len) traj[k * stride] = v;
        if (k >= 2 && mode != 2) {
            const long long c = k - 1;
            if (mode == 0 && c >= c_next && c < n - 2 &&
                w1 - w0 > UC_EPS_PEAK_DELTA && w1 > UC_PEAK_THRESHOLD && w1 >= v)
                mode = 1;
            if (mode == 1) {
                if (c < n - 2) {
                    if (w1 < v) { mode = 0; c_next = c + 2; }
                    else if (w1 - v > UC_EPS_PEAK_DELTA) { emit(w0, w1, v, c); mode = 0; c_next = c + 2; }
                } else {
                    mode = 2;
                }
            }
        }
        w0 = w1;
        w1 = v;
    }

    __device__ __forceinline__ int count() const {
        int c = raw > 1 ? raw - 1 : 0;
        return c > max_peaks ? max_peaks : c;
    }
};

// ================================================================ kernels ==
// Bifurcation sweep (calculateDiscreteModelCUDA + peakFinderCUDA fused):
// transient, recorded run with streaming peaks, flag = peak count / -1 / 0.
template <class S>
__device__ void uc_bif_body(const UcJob& j) {
    typedef typename Prec<S>::T T;
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= j.n_cells) return;
    const long long cell = j.cell_offset + lane;

    double ad[UC_M], x0[UC_N], hd;
    const bool ok = uc_setup(j, cell, ad, x0, hd);
    if (!ok) { j.flag[lane] = -1; return; }
    long long skip = j.skip_steps, iters = j.n_iters;
    if (uc_sweeps_step(j)) {
        skip = (long long)ceil(j.transient_time / hd);
        long long it = (long long)ceil(j.t_max / hd / (double)j.pre_scaller);
        iters = it < iters ? it : iters;
    }
    T a[UC_M];
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = Prec<S>::to_t(ad[i]);
    const T h = Prec<S>::to_t(hd);
    S s;
    s.load(x0);

    NoObserver none;
    int flag = uc_loop(s, a, h, skip, j.pre_scaller, 0, j.max_value, none);
    PeakObserver obs;
    obs.w0 = 0.0; obs.w1 = 0.0; obs.n = iters; obs.c_next = 2; obs.mode = 0; obs.raw = 0;
    obs.last_t = 0.0; obs.time_step = hd * (double)j.pre_scaller; obs.max_peaks = j.max_peaks;
    obs.stride = j.n_cells; obs.peaks = j.out0 + lane; obs.intervals = j.out1 + lane;
    obs.traj = j.traj ? j.traj + lane : nullptr; obs.traj_len = j.traj ? j.traj_len : 0;
    if (flag == 1 || flag == -1)
        flag = uc_loop(s, a, h, iters, j.pre_scaller, j.writable_var, j.max_value, obs);
    j.flag[lane] = (flag == 1) ? obs.count() : flag;
    if (j.final_state) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.n_cells + lane] = s.get(i);
    }
}

// Random unit probe vectors (U-CUDA/cudaLibrary.cu:2603-2626 for one vector,
// :2980-2997 for the basis): K vectors of N draws u - 0.5, each normalised.
template <int K>
__device__ __forceinline__ void uc_probe_basis(u64 seed, long long cell, double* z) {
    UcRng r;
    r.init(seed, (u64)cell, 0ULL);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double p = 0.0;
#pragma unroll
        for (int i = 0; i < UC_N; ++i) { z[k * UC_N + i] = (double)(r.uniform() - 0.5); p = p + z[k * UC_N + i] * z[k * UC_N + i]; }
        p = sqrt(p);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) z[k * UC_N + i] = z[k * UC_N + i] / p;
    }
}

// Modified Gram-Schmidt in double on K row vectors; d[k] = norms before
// normalisation (zero vectors stay zero, as in the reference).
template <int K>
__device__ __forceinline__ void uc_mgs(double* v, double* d) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int q = 0; q < k; ++q) {
            double r = 0.0;
#pragma unroll
            for (int i = 0; i < UC_N; ++i) r = r + v[k * UC_N + i] * v[q * UC_N + i];
#pragma unroll
            for (int i = 0; i < UC_N; ++i) v[k * UC_N + i] = v[k * UC_N + i] - r * v[q * UC_N + i];
        }
        double nrm = 0.0;
#pragma unroll
        for (int i = 0; i < UC_N; ++i) nrm = nrm + v[k * UC_N + i] * v[k * UC_N + i];
        nrm = sqrt(nrm);
        d[k] = nrm;
#pragma unroll
        for (int i = 0; i < UC_N; ++i) v[k * UC_N + i] = (nrm == 0.0) ? 0.0 : v[k * UC_N + i] / nrm;
    }
}

// Classical Gram-Schmidt exactly as U-CUDA/cudaLibrary.cu:2832-2874: the
// projections use the original vector a_i against the unnormalised b_j, then
// every b_i is normalised; den[i] = |b_i|.
template <int K>
__device__ __forceinline__ void uc_cgs(const double* a, double* b, double* den) {
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
        for (int q = 0; q < UC_N; ++q) b[i * UC_N + q] = a[i * UC_N + q];
#pragma unroll
        for (int jj = 0; jj < i; ++jj) {
            double num = 0.0, dd = 0.0;
#pragma unroll
            for (int q = 0; q < UC_N; ++q) { num = num + a[i * UC_N + q] * b[jj * UC_N + q]; dd = dd + b[jj * UC_N + q] * b[jj * UC_N + q]; }
            const double f = (dd == 0.0) ? 0.0 : num / dd;
#pragma unroll
            for (int q = 0; q < UC_N; ++q) b[i * UC_N + q] = b[i * UC_N + q] - f * b[jj * UC_N + q];
        }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double nrm = 0.0;
#pragma unroll
        for (int q = 0; q < UC_N; ++q) nrm = nrm + b[i * UC_N + q] * b[i * UC_N + q];
        nrm = sqrt(nrm);
#pragma unroll
        for (int q = 0; q < UC_N; ++q) b[i * UC_N + q] = (nrm == 0.0) ? 0.0 : b[i * UC_N + q] / nrm;
        den[i] = nrm;
    }
}

// Lyapunov exponents of one cell: K = 1 gives the largest exponent (LLE map),
// K = N the full spectrum. Variational method: state and tangent basis
// integrated together, renormalised (K = 1) or re-orthonormalised by MGS in
// double after every block of nt_steps + 1 steps (the reference block, which
// ends with the step loop's trailing step). Finite-difference method: the
// reference's Benettin scheme with probe copies at distance eps.
// Results: out0[k * n_cells + cell] = sum_k / t_max; flag 1, or -1 with 999
// (diverged) / NaN (non-positive step).
template <int K, class S>
__device__ void uc_lyap_body(const UcJob& j) {
    typedef typename Prec<S>::T T;
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= j.n_cells) return;
    const long long cell = j.cell_offset + lane;

    double ad[UC_M], x0[UC_N], hd;
    if (!uc_setup(j, cell, ad, x0, hd)) {
#pragma unroll
        for (int k = 0; k < K; ++k) j.out0[(long long)k * j.n_cells + lane] = __builtin_nan("");
        j.flag[lane] = -1;
        return;
    }
    long long skip = j.skip_steps, nts = j.nt_steps;
    if (uc_sweeps_step(j)) {
        nts = (long long)ceil(j.nt / hd);
        skip = (long long)ceil(j.transient_time / hd);
    }
    T a[UC_M];
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = Prec<S>::to_t(ad[i]);
    const T h = Prec<S>::to_t(hd);
    S s;
    s.load(x0);

    double sum[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sum[k] = 0.0;
    bool alive = true;

    NoObserver none;
    if (uc_loop(s, a, h, skip, 1, 0, j.max_value, none) == 0) alive = false;

    if (alive) {
        double z[K * UC_N];
        uc_probe_basis<K>(j.probe_seed, cell, z);
        if (j.method == UC_METHOD_VARIATIONAL) {
            double d[K];
            if (K > 1) uc_mgs<K>(z, d);
            T V[K * UC_N];
#pragma unroll
            for (int i = 0; i < K * UC_N; ++i) V[i] = Prec<S>::to_t(z[i]);
            for (long long b = 0; b < j.n_iters; ++b) {
                if (uc_loop_tan<K>(s, V, a, h, nts, j.max_value) == 0) { alive = false; break; }
                double v[K * UC_N];
#pragma unroll
                for (int i = 0; i < K * UC_N; ++i) v[i] = (double)V[i];
                if (K == 1) {
                    double nrm = 0.0;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) nrm = nrm + v[i] * v[i];
                    nrm = sqrt(nrm);
                    if (nrm <= 1e-14) nrm = 1e-14;
                    sum[0] = sum[0] + log(nrm);
                    const double inv = 1.0 / nrm;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) v[i] = v[i] * inv;
                } else {
                    uc_mgs<K>(v, d);
#pragma unroll
                    for (int k = 0; k < K; ++k) sum[k] = sum[k] + log(d[k]);
                }
#pragma unroll
                for (int i = 0; i < K * UC_N; ++i) V[i] = Prec<S>::to_t(v[i]);
            }
        } else {
            // Reference finite-difference Benettin (LLEKernelCUDA / LSKernelCUDA).
            const double eps = j.eps;
            S y[K];
            double q[K * UC_N], den[K];
            if (K > 1) uc_cgs<K>(z, q, den);
            else {
#pragma unroll
                for (int i = 0; i < UC_N; ++i) q[i] = z[i];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double yk[UC_N];
#pragma unroll
                for (int i = 0; i < UC_N; ++i) yk[i] = q[k * UC_N + i] * eps + s.get(i);
                y[k].load(yk);
            }
            for (long long b = 0; b < j.n_iters && alive; ++b) {
                if (uc_loop(s, a, h, nts, 1, 0, j.max_value, none) == 0) { alive = false; break; }
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (alive && uc_loop(y[k], a, h, nts, 1, 0, j.max_value, none) == 0) alive = false;
                if (!alive) break;
                if (K == 1) {
                    double d = 0.0;
#pragma unroll
                    for (int l = 0; l < UC_N; ++l) {
                        const double t2 = (1.0 / eps) * (s.get(l) - y[0].get(l));
                        d = d + t2 * t2;
                    }
                    d = sqrt(d);
                    if (d <= 1e-14) d = 1e-14;
                    sum[0] = sum[0] + log(d);
                    const double inv = 1.0 / d;
                    double yk[UC_N];
#pragma unroll
                    for (int l = 0; l < UC_N; ++l) yk[l] = s.get(l) - ((s.get(l) - y[0].get(l) + 1e-14) * inv);
                    y[0].load(yk);
                } else {
                    double dy[K * UC_N], zz[K * UC_N];
#pragma unroll
                    for (int k = 0; k < K; ++k)
#pragma unroll
                        for (int l = 0; l < UC_N; ++l) dy[k * UC_N + l] = y[k].get(l) - s.get(l);
                    uc_cgs<K>(dy, zz, den);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        sum[k] = sum[k] + log(den[k] / eps);
                        double yk[UC_N];
#pragma unroll
                        for (int l = 0; l < UC_N; ++l) yk[l] = s.get(l) + zz[k * UC_N + l] * eps;
                        y[k].load(yk);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) j.out0[(long long)k * j.n_cells + lane] = alive ? sum[k] / j.t_max : 999.0;
    j.flag[lane] = alive ? 1 : -1;
    if (j.final_state) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.n_cells + lane] = s.get(i);
    }
}

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bif_f64(UcJob j) { uc_bif_body<StateF64>(j); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bif_f32c(UcJob j) { uc_bif_body<StateF32C>(j); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_lle_f64(UcJob j) { uc_lyap_body<1, StateF64>(j); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_lle_f32c(UcJob j) { uc_lyap_body<1, StateF32C>(j); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_ls_f64(UcJob j) { uc_lyap_body<UC_N, StateF64>(j); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_ls_f32c(UcJob j) { uc_lyap_body<UC_N, StateF32C>(j); }
