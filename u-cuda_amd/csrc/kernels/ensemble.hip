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
#define UC_LLE_BLOCK 10                    // block length the LLE lane kernels unroll fully

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
// A value the optimiser must treat as computed, not loaded. Without it a
// chain of selects among state components becomes a load through a selected
// address: the state array is then indexed dynamically, stays in memory, and
// the backend moves it to LDS, so every step goes through LDS. A chain that
// starts from such a value has no pair of loads to fold.
template <typename T> __device__ __forceinline__ T uc_opaque(T v) {
    asm("" : "+v"(v));
    return v;
}

// Harmonic forcing in the fast modules: the generated source lists the
// right-hand side's cos / sin terms whose argument is w * t + phi (w, phi
// from the parameters; uc_harm_coef) and gives uc_rhs_h, the right-hand
// side with those terms as inputs hc[k] / hs[k]. Strict modules keep the
// reference's calls.
#if !UC_STRICT && defined(UC_HARMONICS) && UC_HARMONICS > 0 && UC_HARMONICS <= 8
#define UC_HARM UC_HARMONICS
#else
#define UC_HARM 0
#endif

// cos / sin of d: the Taylor series for |d| <= 0.8 (truncation below 1e-18
// relative in double, 2e-10 in float). Larger d (rare: a step longer than
// 0.8 / w) is reduced by multiples of 2 pi (two-part constant, exact
// enough while d's own ulp is far below 2 pi: |d| < 1e15 in double, 1e7 in
// float; beyond that, and for inf / NaN, the result is NaN), halved until
// |d| <= 0.8 (at most twice after the reduction) and doubled back, a few
// ulps and no library range reduction.
// [plain C++ from here to the matching end mark, compiled on the host too]
__device__ __forceinline__ void uc_sincos_poly(double d, double& c, double& s) {
    const double z = d * d;
    double p = 1.0 / 6402373705728000.0;  // 1/18!
    p = fma(p, z, -1.0 / 20922789888000.0);
    p = fma(p, z, 1.0 / 87178291200.0);
    p = fma(p, z, -1.0 / 479001600.0);
    p = fma(p, z, 1.0 / 3628800.0);
    p = fma(p, z, -1.0 / 40320.0);
    p = fma(p, z, 1.0 / 720.0);
    p = fma(p, z, -1.0 / 24.0);
    p = fma(p, z, 0.5);
    c = fma(-p, z, 1.0);
    double q = 1.0 / 355687428096000.0;  // 1/17!
    q = fma(q, z, -1.0 / 1307674368000.0);
    q = fma(q, z, 1.0 / 6227020800.0);
    q = fma(q, z, -1.0 / 39916800.0);
    q = fma(q, z, 1.0 / 362880.0);
    q = fma(q, z, -1.0 / 5040.0);
    q = fma(q, z, 1.0 / 120.0);
    q = fma(q, z, -1.0 / 6.0);
    s = fma(q * z, d, d);
}
__device__ __forceinline__ void uc_sincos_poly(float d, float& c, float& s) {
    const float z = d * d;
    float p = -1.0f / 3628800.0f;  // -1/10!
    p = fmaf(p, z, 1.0f / 40320.0f);
    p = fmaf(p, z, -1.0f / 720.0f);
    p = fmaf(p, z, 1.0f / 24.0f);
    p = fmaf(p, z, -0.5f);
    c = fmaf(p, z, 1.0f);
    float q = -1.0f / 39916800.0f;  // -1/11!
    q = fmaf(q, z, 1.0f / 362880.0f);
    q = fmaf(q, z, -1.0f / 5040.0f);
    q = fmaf(q, z, 1.0f / 120.0f);
    q = fmaf(q, z, -1.0f / 6.0f);
    s = fmaf(q * z, d, d);
}
template <typename T> struct UcTwoPi;
template <> struct UcTwoPi<double> {
    static constexpr double hi = 6.283185307179586, lo = 2.4492935982947064e-16, inv = 0.15915494309189535,
                            cap = 1e15;
};
template <> struct UcTwoPi<float> {
    static constexpr float hi = 6.2831854820251465f, lo = -1.7484555314695172e-07f, inv = 0.15915494f, cap = 1e7f;
};
template <typename T> __device__ __forceinline__ void uc_sincos_small(T d, T& c, T& s) {
    if (fabs(d) <= (T)0.8) {
        uc_sincos_poly(d, c, s);
        return;
    }
    if (!(fabs(d) < UcTwoPi<T>::cap)) {
        c = s = (T)__builtin_nan("");
        return;
    }
    const T n = rint(d * UcTwoPi<T>::inv);
    d = fma(-n, UcTwoPi<T>::hi, d);
    d = fma(-n, UcTwoPi<T>::lo, d);
    int halvings = 0;
    while (fabs(d) > (T)0.8) {
        d *= (T)0.5;
        ++halvings;
    }
    uc_sincos_poly(d, c, s);
    for (int i = 0; i < halvings; ++i) {
        const T s2 = (T)2 * s * c;
        c = (c - s) * (c + s);
        s = s2;
    }
}

// [end of plain C++]

#if UC_HARM
// cos / sin of every forcing phase at the state's own time, for the
// fixed-step RK4 loops: two rotations per step (to t + h/2, then t + h) by
// the cached half-step angle, in double. They are evaluated afresh from
// w t + phi (uc_sincos_small: reduction by 2 pi, Taylor series, a few ulps;
// the library's range reduction would be inlined into every unrolled step)
// when the caller's time or step is not the one tracked (a fresh state, one
// restored from the slab, a reset clock) and every 128 steps, so rounding
// never builds up over more than 128 steps.
struct UcPhase {
    double c[UC_HARM], s[UC_HARM];    // at time t
    double ch[UC_HARM], sh[UC_HARM];  // cos / sin(w hh)
    double t, hh;
    int n = -1;                       // steps since the exact evaluation; -1: none
};
#endif

struct StateF64 {
    double x[UC_N];
#if UC_HARM
    UcPhase ph;
#endif
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x[i] = x0[i];
#if UC_HARM
        ph.n = -1;
#endif
    }
    __device__ __forceinline__ double get(int i) const { return x[i]; }
    // component q (0 <= q < N) selected without indexing the array
    __device__ __forceinline__ double get_sel(int q) const {
        double v = uc_opaque(x[0]);
#pragma unroll
        for (int i = 1; i < UC_N; ++i) v = (i == q) ? x[i] : v;
        return v;
    }
};

// F32C state: hi + Kahan compensation c (true value ~ hi - c).
struct StateF32C {
    float hi[UC_N], c[UC_N];
#if UC_HARM
    UcPhase ph;
#endif
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) {
            hi[i] = (float)x0[i];
            c[i] = (float)((double)hi[i] - x0[i]);
        }
#if UC_HARM
        ph.n = -1;
#endif
    }
    __device__ __forceinline__ double get(int i) const { return (double)hi[i] - (double)c[i]; }
    // the pair is selected in float and converted once (same value as get(q))
    __device__ __forceinline__ double get_sel(int q) const {
        float h = uc_opaque(hi[0]), cc = uc_opaque(c[0]);
#pragma unroll
        for (int i = 1; i < UC_N; ++i) {
            h = (i == q) ? hi[i] : h;
            cc = (i == q) ? c[i] : cc;
        }
        return (double)h - (double)cc;
    }
};

// Forcing phases are evaluated afresh at the start of every unit a sliced
// launch can begin with (a 100-step chunk, a renormalisation block, 100
// recorded samples), in every kernel, so a cell resumed from the slab
// carries the same values as one run whole.
template <class S> __device__ __forceinline__ void uc_phase_mark(S& s) {
#if UC_HARM
    s.ph.n = -1;
#else
    (void)s;
#endif
}

// ------------------------------------------------------ resumable slices --
// A sliced cell (UcJob.slab set) keeps its state in the slab between
// launches: word w of the cell at slab[w * stride + lane]. Values are stored
// as bit patterns, so every one round-trips exactly, and saved and loaded by
// the same sequence of calls; words past the kind's count are dropped.
// Word 0 is the done mark: a cell whose results are written is skipped by
// later launches.
static_assert(UC_UNIT_STEPS == UC_CHECK_INTERVAL, "work units are magnitude-check intervals");
#define UC_UNITS_ALL (1LL << 50)  // unit_end of a whole run

struct UcSlab {
    double* p;
    long long stride;
    int w, cap;
    bool store;
    __device__ __forceinline__ UcSlab(const UcJob& j, long long lane, int cap_, bool store_)
        : p(j.slab + lane), stride(j.stride), w(0), cap(cap_), store(store_) {}
    __device__ __forceinline__ void seek(int w_) { w = w_; }
    __device__ __forceinline__ void u64w(u64& v) {
        if (w < cap) {
            double* q = p + (long long)w * stride;
            if (store) *q = __builtin_bit_cast(double, v);
            else v = __builtin_bit_cast(u64, *q);
        }
        ++w;
    }
    __device__ __forceinline__ void f64(double& v) { u64 b = __builtin_bit_cast(u64, v); u64w(b); v = __builtin_bit_cast(double, b); }
    __device__ __forceinline__ void f32(float& v) {
        u64 b = __builtin_bit_cast(unsigned int, v);
        u64w(b);
        v = __builtin_bit_cast(float, (unsigned int)b);
    }
    __device__ __forceinline__ void val(double& v) { f64(v); }
    __device__ __forceinline__ void val(float& v) { f32(v); }
    __device__ __forceinline__ void i64(long long& v) { u64 b = (u64)v; u64w(b); v = (long long)b; }
    __device__ __forceinline__ void i32(int& v) { long long b = v; i64(b); v = (int)b; }
    __device__ __forceinline__ void b8(bool& v) { int b = v ? 1 : 0; i32(b); v = b != 0; }
    __device__ __forceinline__ void st(StateF64& s) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) f64(s.x[i]);
    }
    __device__ __forceinline__ void st(StateF32C& s) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) f32(s.hi[i]);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) f32(s.c[i]);
    }
};
template <class S> struct UcWords { static constexpr int n = 2 * UC_N; };
template <> struct UcWords<StateF64> { static constexpr int n = UC_N; };

// Units of a chunked phase of `steps` steps (at least one: the trailing step).
__device__ __forceinline__ long long uc_chunks(long long steps) {
    return steps <= 0 ? 1 : (steps + UC_CHECK_INTERVAL - 1) / UC_CHECK_INTERVAL;
}
// Steps [i0, i1) of such a phase, starting at unit `base`, that launch units
// [u0, u1) cover; `last`: the launch reaches the phase's end (trailing step).
struct UcRange {
    long long i0, i1;
    bool last;
};
__device__ __forceinline__ UcRange uc_range(long long steps, long long base, long long u0, long long u1) {
    UcRange r;
    const long long a = u0 > base ? u0 - base : 0, b = u1 - base;
    r.i0 = a * UC_CHECK_INTERVAL;
    r.i1 = b * UC_CHECK_INTERVAL < steps ? b * UC_CHECK_INTERVAL : steps;
    r.last = b >= uc_chunks(steps);
    return r;
}

#define UC_LN2 0.6931471805599453
__device__ __forceinline__ float uc_ldexp(float x, int e) { return __builtin_ldexpf(x, e); }
__device__ __forceinline__ double uc_ldexp(double x, int e) { return __builtin_ldexp(x, e); }
__device__ __forceinline__ int uc_frexp_exp(float x) { int e; (void)__builtin_frexpf(x, &e); return e; }
__device__ __forceinline__ int uc_frexp_exp(double x) { int e; (void)__builtin_frexp(x, &e); return e; }

__device__ __forceinline__ void uc_kahan_add(float& hi, float& c, float incr) {
    const float y = incr - c;
    const float t = hi + y;
    c = (t - hi) - y;
    hi = t;
}

// The same update for the increment f * g: y = f * g - c is rounded once
// (one FMA instead of a multiply and a subtraction).
__device__ __forceinline__ void uc_kahan_fma(float& hi, float& c, float f, float g) {
    const float y = fmaf(f, g, -c);
    const float t = hi + y;
    c = (t - hi) - y;
    hi = t;
}

// Step factors h, h/2, h/6. A loop whose step is the same on every lane
// passes StepU, read from the job descriptor (host-computed, so they sit in
// scalar registers and the stage FMAs read one VGPR fewer); other loops pass
// the plain per-lane h. Both give the same values.
template <typename T> struct StepU { T h, hh, h6; };
__device__ __forceinline__ float uc_h(float h) { return h; }
__device__ __forceinline__ double uc_h(double h) { return h; }
__device__ __forceinline__ float uc_hh(float h) { return 0.5f * h; }
__device__ __forceinline__ double uc_hh(double h) { return 0.5 * h; }
__device__ __forceinline__ float uc_h6(float h) { return h * (1.0f / 6.0f); }
__device__ __forceinline__ double uc_h6(double h) { return h * (1.0 / 6.0); }
template <typename T> __device__ __forceinline__ T uc_h(const StepU<T>& s) { return s.h; }
template <typename T> __device__ __forceinline__ T uc_hh(const StepU<T>& s) { return s.hh; }
template <typename T> __device__ __forceinline__ T uc_h6(const StepU<T>& s) { return s.h6; }
__device__ __forceinline__ void uc_job_step(const UcJob& j, StepU<double>& s) { s.h = j.h; s.hh = j.hh; s.h6 = j.h6; }
__device__ __forceinline__ void uc_job_step(const UcJob& j, StepU<float>& s) { s.h = j.hf; s.hh = j.hhf; s.h6 = j.h6f; }

// UcRhs: the right-hand side of the four stages. With harmonic forcing the
// stage values come from the state's tracked phase (start):
// stage 1 at t, stages 2 and 3 at t + h/2, stage 4 at t + h.
template <typename T, class S> struct UcRhs {
#if UC_HARM
    T c0[UC_HARM], s0[UC_HARM], c1[UC_HARM], s1[UC_HARM], c2[UC_HARM], s2[UC_HARM];
#endif
    __device__ __forceinline__ void start(S& st, const T* a, double t, double h, double hh) {
#if UC_HARM
        UcPhase& ph = st.ph;
        if (ph.n < 0 || ph.n >= 128 || ph.t != t || ph.hh != hh) {
            T w[UC_HARM], phi[UC_HARM];
            uc_harm_coef<T>(a, w, phi);
#pragma unroll
            for (int k = 0; k < UC_HARM; ++k) {
                uc_sincos_small((double)w[k] * t + (double)phi[k], ph.c[k], ph.s[k]);
                uc_sincos_small((double)w[k] * hh, ph.ch[k], ph.sh[k]);
            }
            ph.hh = hh;
            ph.n = 0;
        }
#pragma unroll
        for (int k = 0; k < UC_HARM; ++k) {
            const double c = ph.c[k], s = ph.s[k], ch = ph.ch[k], sh = ph.sh[k];
            const double cm = c * ch - s * sh, sm = s * ch + c * sh;
            const double ce = cm * ch - sm * sh, se = sm * ch + cm * sh;
            c0[k] = (T)c; s0[k] = (T)s;
            c1[k] = (T)cm; s1[k] = (T)sm;
            c2[k] = (T)ce; s2[k] = (T)se;
            ph.c[k] = ce;
            ph.s[k] = se;
        }
        ph.t = t + h;
        ++ph.n;
#else
        (void)st; (void)a; (void)t; (void)h; (void)hh;
#endif
    }
    // stage q = 0..3 at time ts
    __device__ __forceinline__ void eval(int q, const T* x, const T* a, T ts, T* k) const {
#if UC_HARM
        uc_rhs_h<T>(x, a, ts, q == 0 ? c0 : q == 3 ? c2 : c1, q == 0 ? s0 : q == 3 ? s2 : s1, k);
#else
        (void)q;
        uc_rhs<T>(x, a, ts, k);
#endif
    }
};

#if UC_CUSTOM_KRS
// A caller-supplied step body (the reference's custom KRS schemes,
// U-CUDA/system_record.h:10-13) is the step map itself: numb = double, or
// float without compensation (the body forms X + increment internally).
template <class H> __device__ __forceinline__ void uc_rk4(StateF64& s, const double* a, H h, double) { uc_krs_step<double>(s.x, a, uc_h(h)); }
template <class H> __device__ __forceinline__ void uc_rk4(StateF32C& s, const float* a, H h, double) { uc_krs_step<float>(s.hi, a, uc_h(h)); }
#else
// Classic RK4 with the reference's stage structure and update order
// (U-CUDA/codegen.cpp:597-606): X1 = X + 0.5*h*k for stages 0,1, X1 = X + h*k
// for stage 2, X += h*(k0 + 2*k1 + 2*k2 + k3)/6. Accumulating the bracket
// left to right keeps the reference's summation order. t is the step's
// start time (double; the stages see it in their precision).
//
template <class H>
__device__ __forceinline__ void uc_rk4(StateF64& s, const double* a, H hs, double t) {
    double x1[UC_N], k[UC_N], acc[UC_N];
    const double h = uc_h(hs), hh = uc_hh(hs);
    UcRhs<double, StateF64> f;
    f.start(s, a, t, h, hh);
    f.eval(0, s.x, a, t, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.x[i] + hh * k[i]; }
    f.eval(1, x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + hh * k[i]; }
    f.eval(2, x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + h * k[i]; }
    f.eval(3, x1, a, t + h, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
#if UC_STRICT
        s.x[i] = s.x[i] + h * acc[i] / 6.0;
#else
        s.x[i] = fma(uc_h6(hs), acc[i], s.x[i]);
#endif
    }
}

template <class H>
__device__ __forceinline__ void uc_rk4(StateF32C& s, const float* a, H hs, double td) {
    float x1[UC_N], k[UC_N], acc[UC_N];
    const float h = uc_h(hs), hh = uc_hh(hs), t = (float)td;
    UcRhs<float, StateF32C> f;
    f.start(s, a, td, (double)h, (double)hh);
    f.eval(0, s.hi, a, t, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.hi[i] + hh * k[i]; }
    f.eval(1, x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + hh * k[i]; }
    f.eval(2, x1, a, t + hh, k);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + h * k[i]; }
    f.eval(3, x1, a, t + h, k);
    const float h6 = uc_h6(hs);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
        uc_kahan_fma(s.hi[i], s.c[i], h6, acc[i]);
    }
}

#endif

// RK4 on the state together with K tangent vectors V (row-major K x N): the
// variational equations V' = J(X) V integrated with the same stages, which
// is exactly the derivative of the discrete RK4 map.
template <int K, class H>
__device__ __forceinline__ void uc_rk4_tan(StateF64& s, double* V, const double* a, H hs, double t) {
    double x1[UC_N], k[UC_N], acc[UC_N];
    double v1[K * UC_N], l[K * UC_N], lacc[K * UC_N];
    const double h = uc_h(hs), hh = uc_hh(hs);
    UcRhs<double, StateF64> f;
    f.start(s, a, t, h, hh);
    f.eval(0, s.x, a, t, k);
    uc_tangent<double, K>(s.x, a, t, V, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.x[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = l[i]; v1[i] = V[i] + hh * l[i]; }
    f.eval(1, x1, a, t + hh, k);
    uc_tangent<double, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + hh * l[i]; }
    f.eval(2, x1, a, t + hh, k);
    uc_tangent<double, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0 * k[i]; x1[i] = s.x[i] + h * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0 * l[i]; v1[i] = V[i] + h * l[i]; }
    f.eval(3, x1, a, t + h, k);
    uc_tangent<double, K>(x1, a, t + h, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
#if UC_STRICT
        s.x[i] = s.x[i] + h * acc[i] / 6.0;
#else
        s.x[i] = fma(uc_h6(hs), acc[i], s.x[i]);
#endif
    }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) {
        lacc[i] = lacc[i] + l[i];
#if UC_STRICT
        V[i] = V[i] + h * lacc[i] / 6.0;
#else
        V[i] = fma(uc_h6(hs), lacc[i], V[i]);
#endif
    }
}

template <int K, class H>
__device__ __forceinline__ void uc_rk4_tan(StateF32C& s, float* V, const float* a, H hs, double td) {
    float x1[UC_N], k[UC_N], acc[UC_N];
    float v1[K * UC_N], l[K * UC_N], lacc[K * UC_N];
    const float h = uc_h(hs), hh = uc_hh(hs), t = (float)td;
    UcRhs<float, StateF32C> f;
    f.start(s, a, td, (double)h, (double)hh);
    f.eval(0, s.hi, a, t, k);
    uc_tangent<float, K>(s.hi, a, t, V, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = k[i]; x1[i] = s.hi[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = l[i]; v1[i] = V[i] + hh * l[i]; }
    f.eval(1, x1, a, t + hh, k);
    uc_tangent<float, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + hh * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0f * l[i]; v1[i] = V[i] + hh * l[i]; }
    f.eval(2, x1, a, t + hh, k);
    uc_tangent<float, K>(x1, a, t + hh, v1, l);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { acc[i] = acc[i] + 2.0f * k[i]; x1[i] = s.hi[i] + h * k[i]; }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + 2.0f * l[i]; v1[i] = V[i] + h * l[i]; }
    f.eval(3, x1, a, t + h, k);
    uc_tangent<float, K>(x1, a, t + h, v1, l);
    const float h6 = uc_h6(hs);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        acc[i] = acc[i] + k[i];
        uc_kahan_fma(s.hi[i], s.c[i], h6, acc[i]);
    }
#pragma unroll
    for (int i = 0; i < K * UC_N; ++i) { lacc[i] = lacc[i] + l[i]; V[i] = fmaf(h6, lacc[i], V[i]); }
}

// --------------------------------------------------- precision adaptors ---
template <class S> struct Prec;
// Fixed-point threshold of the trailing-step test: the reference's 1e-8
// (configCUDA.h:84) in double. With float stages the state cannot settle
// closer to an equilibrium than the float grid allows: one step there still
// moves it by about h*|J|*eps_f32*|x|, so the threshold is raised to
// 2^-16 * h * pre * sum|x| (128 float ulps of the state per unit |J|*h),
// far below the per-step motion of any oscillating orbit.
template <> struct Prec<StateF64> {
    typedef double T;
    static __device__ __forceinline__ double to_t(double v) { return v; }
    static __device__ __forceinline__ double fixed_point_eps(double, int, double) { return UC_EPS_FIXED_POINT; }
};
template <> struct Prec<StateF32C> {
    typedef float T;
    static __device__ __forceinline__ float to_t(double v) { return (float)v; }
    static __device__ __forceinline__ double fixed_point_eps(double h, int pre, double sum_abs) {
        return fmax(UC_EPS_FIXED_POINT, 1.52587890625e-05 * h * (double)pre * sum_abs);
    }
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

// F32C: the magnitude test sums the float state hi. The compensation c is
// below half an ulp of hi, so the sum moves by < 2^-24 relative, and the
// test stays in FP32 (no conversions to double on the step path).
__device__ __forceinline__ bool uc_diverged(const StateF32C& s, double max_value) {
    float chk = 0.0f;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) chk = chk + fabsf(s.hi[i]);
    if (!(chk <= __FLT_MAX__)) return true;  // NaN or Inf
    return max_value != 0.0 && (double)chk > max_value;
}

// The same check for loops whose bound is wave-uniform, without branches:
// F32C compares the float sum against max_value_f (the largest float <=
// max_value, so the verdict is the double comparison's).
__device__ __forceinline__ bool uc_diverged_u(const StateF64& s, const UcJob& j) {
    double chk = 0.0;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) chk = chk + fabs(s.x[i]);
    return !(chk <= __DBL_MAX__) || (j.max_value != 0.0 && chk > j.max_value);
}
__device__ __forceinline__ bool uc_diverged_u(const StateF32C& s, const UcJob& j) {
    float chk = 0.0f;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) chk = chk + fabsf(s.hi[i]);
    return !(chk <= __FLT_MAX__) || (j.max_value != 0.0 && chk > j.max_value_f);
}

// dst = take ? src : dst, member by member (selects, no branch)
__device__ __forceinline__ void uc_select(StateF64& dst, const StateF64& src, bool take) {
#pragma unroll
    for (int i = 0; i < UC_N; ++i) dst.x[i] = take ? src.x[i] : dst.x[i];
}
__device__ __forceinline__ void uc_select(StateF32C& dst, const StateF32C& src, bool take) {
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        dst.hi[i] = take ? src.hi[i] : dst.hi[i];
        dst.c[i] = take ? src.c[i] : dst.c[i];
    }
}

template <class S>
__device__ __forceinline__ int uc_fixed_point_flag(const S& s, const double* prev, double h, int pre) {
    double d = 0.0, m = 0.0;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { d = d + fabs(s.get(i) - prev[i]); m = m + fabs(prev[i]); }
    return fabs(d) < Prec<S>::fixed_point_eps(h, pre, m) ? -1 : 1;
}

// Plain state loop. `t` is the system time (used only by right-hand sides
// with a time symbol; dead code otherwise), advanced by h per step.
template <class S, class Obs>
__device__ int uc_loop(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t, long long iters, int pre,
                       int wvar, double max_value, Obs& obs) {
    typedef typename Prec<S>::T T;
    const double hd = (double)h;
    int until_check = 0;
    for (long long i = 0; i < iters; ++i) {
        if (Obs::records) {
            double v;
            if (wvar < 0) {
                if (UC_N >= 3) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0) + UC_E * s.get(UC_N >= 3 ? 2 : 0);
                else if (UC_N == 2) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0);
                else v = s.get(0);
            } else {
                v = s.get_sel(wvar);
            }
            obs.sample(i, v);
            for (int p = 0; p < pre; ++p) { uc_rk4(s, a, h, t); if (UC_USES_TIME) t += hd; }
        } else {
            uc_rk4(s, a, h, t);
            if (UC_USES_TIME) t += hd;
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
    for (int p = 0; p < pre; ++p) { uc_rk4(s, a, h, t); if (UC_USES_TIME) t += hd; }
    return uc_fixed_point_flag(s, prev, hd, pre);
}

// State + K tangent vectors advanced together for `iters` steps plus the
// trailing step, with the magnitude check on the state (variational
// counterpart of one reference renormalisation block).
template <int K, class S>
__device__ int uc_loop_tan(S& s, typename Prec<S>::T* V, const typename Prec<S>::T* a, typename Prec<S>::T h,
                           double& t, long long iters, double max_value) {
    typedef typename Prec<S>::T T;
    const double hd = (double)h;
    int until_check = 0;
    for (long long i = 0; i < iters; ++i) {
        uc_rk4_tan<K>(s, V, a, h, t);
        if (UC_USES_TIME) t += hd;
        if (until_check == 0) {
            if (uc_diverged(s, max_value)) return 0;
            until_check = UC_CHECK_INTERVAL;
        }
        --until_check;
    }
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    uc_rk4_tan<K>(s, V, a, h, t);
    if (UC_USES_TIME) t += hd;
    return uc_fixed_point_flag(s, prev, hd, 1);
}

// Wave-uniform variants of the two loops above, used when the step counts
// come from the job descriptor (every lane of the wave runs the same count;
// only sweeps over h make them per cell). The counters live in scalar
// registers and there is no per-step branch: iteration i with i % 100 == 0
// is followed by the magnitude check, then 99 unconditioned iterations run.
// A lane whose check fails is dead from that check on: its state and time at
// the check are kept, it records nothing more, and it keeps stepping
// (harmlessly) until the loop ends, which it does as soon as every lane of
// the wave is dead. Results equal those of the per-lane loops.
template <class S>
__device__ __forceinline__ double uc_sample_value(const S& s, int wvar) {
    double v;
    if (wvar < 0) {
        if (UC_N >= 3) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0) + UC_E * s.get(UC_N >= 3 ? 2 : 0);
        else if (UC_N == 2) v = s.get(0) + UC_PI * s.get(UC_N >= 2 ? 1 : 0);
        else v = s.get(0);
    } else {
        v = s.get_sel(wvar);
    }
    return v;
}

template <class S, class Obs>
__device__ __forceinline__ void uc_iter(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t,
                                        long long i, int pre, int wvar, bool dead, Obs& obs) {
    typedef typename Prec<S>::T T;
    if (Obs::records) {
        if (!dead) obs.sample(i, uc_sample_value(s, wvar));
        for (int p = 0; p < pre; ++p) { uc_rk4(s, a, h, t); if (UC_USES_TIME) t += (double)h; }
    } else {
        uc_rk4(s, a, h, t);
        if (UC_USES_TIME) t += (double)h;
    }
}

__device__ __forceinline__ bool uc_wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0ULL; }

// Iterations [i0, iters) (i0 a multiple of the check interval, so the
// checks fall where a run from 0 has them), then, when `trail`, the trailing
// steps and the fixed-point test; without it 1 (alive) or 0.
template <class S, class Obs>
__device__ int uc_loop_u(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t, long long i0,
                         long long iters, bool trail, int pre, int wvar, double max_value, Obs& obs) {
    typedef typename Prec<S>::T T;
    const double hd = (double)h;
    bool dead = false;
    S snap = s;
    double tsnap = t;
    for (long long i = i0; i < iters;) {
        uc_phase_mark(s);
        uc_iter(s, a, h, t, i, pre, wvar, dead, obs);
        if (!dead && uc_diverged(s, max_value)) { dead = true; snap = s; tsnap = t; }
        if (uc_wave_all(dead)) break;
        const long long stop = (iters - i > UC_CHECK_INTERVAL) ? i + UC_CHECK_INTERVAL : iters;
#pragma unroll 2
        for (++i; i < stop; ++i) uc_iter(s, a, h, t, i, pre, wvar, dead, obs);
    }
    if (dead) { s = snap; t = tsnap; return 0; }
    if (!trail) return 1;
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    for (int p = 0; p < pre; ++p) { uc_rk4(s, a, h, t); if (UC_USES_TIME) t += hd; }
    return uc_fixed_point_flag(s, prev, hd, pre);
}

template <class S, class Obs>
__device__ __forceinline__ int uc_loop_u(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t,
                                         long long iters, int pre, int wvar, double max_value, Obs& obs) {
    return uc_loop_u(s, a, h, t, 0, iters, true, pre, wvar, max_value, obs);
}

template <int K, class S>
__device__ int uc_loop_tan_u(S& s, typename Prec<S>::T* V, const typename Prec<S>::T* a, typename Prec<S>::T h,
                             double& t, long long iters, double max_value) {
    typedef typename Prec<S>::T T;
    const double hd = (double)h;
    bool dead = false;
    S snap = s;
    double tsnap = t;
    for (long long i = 0; i < iters;) {
        uc_phase_mark(s);
        uc_rk4_tan<K>(s, V, a, h, t);
        if (UC_USES_TIME) t += hd;
        if (!dead && uc_diverged(s, max_value)) { dead = true; snap = s; tsnap = t; }
        if (uc_wave_all(dead)) break;
        const long long stop = (iters - i > UC_CHECK_INTERVAL) ? i + UC_CHECK_INTERVAL : iters;
#pragma unroll 2
        for (++i; i < stop; ++i) {
            uc_rk4_tan<K>(s, V, a, h, t);
            if (UC_USES_TIME) t += hd;
        }
    }
    if (dead) { s = snap; t = tsnap; return 0; }
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    uc_rk4_tan<K>(s, V, a, h, t);
    if (UC_USES_TIME) t += hd;
    return uc_fixed_point_flag(s, prev, hd, 1);
}

// One block of uc_loop_tan_u for 1 <= iters <= UC_CHECK_INTERVAL, where the
// magnitude check falls on iteration 0 only: that step, the check, then
// `iters` more steps (the remaining iterations and the trailing step) in
// one loop with a 32-bit counter, run by the live lanes only, so a lane
// found diverged keeps its state and time at the check. Returns 0 for such
// a lane, else 1: callers that only need the divergence verdict skip the
// trailing fixed-point test, which cannot return 0. Same steps, states and
// dead-lane rule as uc_loop_tan_u.
template <int K, class S>
__device__ __forceinline__ int uc_block_tan_u(S& s, typename Prec<S>::T* V, const typename Prec<S>::T* a,
                                              typename Prec<S>::T h, double& t, int iters, double max_value) {
    typedef typename Prec<S>::T T;
    uc_phase_mark(s);
    uc_rk4_tan<K>(s, V, a, h, t);
    if (UC_USES_TIME) t += (double)h;
    const bool dead = uc_diverged(s, max_value);
    if (uc_wave_all(dead)) return 0;
    if (!dead) {
#pragma unroll 2
        for (int i = 0; i < iters; ++i) {
            uc_rk4_tan<K>(s, V, a, h, t);
            if (UC_USES_TIME) t += (double)h;
        }
    }
    return dead ? 0 : 1;
}

// ------------------------------------------------------ streaming peaks ---
// The reference peak finder (U-CUDA/cudaLibrary.cu:1503-1633) run as samples
// arrive instead of over a stored trajectory: candidate c needs d[c-1], d[c],
// d[c+1]; a plateau walk from c ends at the first j with d[j] < d[j+1] (no
// peak) or d[j] - d[j+1] > 1e-14 (peak at j, parabolic refinement); either
// way the next candidate is j + 2; a walk reaching n - 2 ends the search.
// The first peak only anchors the first interval (the reference's shift-left).
//
// Parabolic refinement of a peak at d0 between dm and dp: the offset of the
// vertex in samples (0 when the parabola is flat).
__device__ __forceinline__ double uc_peak_delta(double dm, double d0, double dp) {
    const double denom = dm - 2.0 * d0 + dp;
    double delta = 0.0;
#if UC_STRICT
    if (fabs(denom) > 1e-12) delta = 0.5 * (dm - dp) / denom;
#else
    // A wave runs this whenever any of its lanes confirms a peak (most
    // steps): the reciprocal and one Newton step replace the division
    // (the quotient moves by an ulp or two).
    if (fabs(denom) > 1e-12) {
        double r = __builtin_amdgcn_rcp(denom);
        r = fma(r, fma(-denom, r, 1.0), r);
        delta = 0.5 * (dm - dp) * r;
    }
#endif
    return delta;
}
__device__ __forceinline__ double uc_peak_value(double dm, double d0, double dp, double delta) {
    return d0 - 0.25 * (dm - dp) * delta;
}

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
    long long stride;       // output plane stride
    double* peaks;          // peaks[k * stride]
    double* intervals;
    double* traj;
    int traj_len;

    __device__ __forceinline__ void emit(double dm, double d0, double dp, long long j) {
        const double delta = uc_peak_delta(dm, d0, dp);
        const double value = uc_peak_value(dm, d0, dp, delta);
        const double tpk = (double)(j - 1) + delta;
        if (raw >= 1 && raw - 1 < max_peaks) {
            peaks[(long long)(raw - 1) * stride] = value;
            if (intervals) intervals[(long long)(raw - 1) * stride] = (tpk - last_t) * time_step;
        }
        last_t = tpk;
        ++raw;
    }

    __device__ __forceinline__ void sample(long long k, double v) {
        if (k < traj_len) traj[k * stride] = v;
        if (k >= 2) {
            // The search as selects, one branch (emit) left: a candidate
            // (mode 0 -> 1) can be confirmed or dropped in the same sample,
            // mode 2 (record end reached inside a plateau) is final.
            const long long c = k - 1;
            const bool inwin = c < n - 2;
            const bool start = mode == 0 && c >= c_next && inwin && w1 - w0 > UC_EPS_PEAK_DELTA &&
                               w1 > UC_PEAK_THRESHOLD && w1 >= v;
            const bool plateau = mode == 1 || start;
            const bool drop = plateau && inwin && w1 < v;
            const bool confirm = plateau && inwin && !(w1 < v) && w1 - v > UC_EPS_PEAK_DELTA;
            if (confirm) emit(w0, w1, v, c);
            mode = (plateau && !inwin) ? 2 : (drop || confirm) ? 0 : start ? 1 : mode;
            c_next = (drop || confirm) ? c + 2 : c_next;
        }
        w0 = w1;
        w1 = v;
    }

    __device__ __forceinline__ int count() const {
        int c = raw > 1 ? raw - 1 : 0;
        return c > max_peaks ? max_peaks : c;
    }
};

// uc_loop_u with a PeakObserver for the common record: one step per sample,
// a state component as the sample, no stored trajectory, fewer than 2^30
// samples, and an observer that is either fresh (search mode, c_next = 2)
// or resumed at i0 >= 2. Same iterations, magnitude checks, dead-lane rule,
// peaks, intervals and observer state as that loop; leaner per sample:
// 32-bit sample counters (scalar compares), the search as one bitwise
// expression (the branch around a confirmed peak is the only one), and the
// peak slots addressed by a running offset. uc_loop_u's `k >= 2` guard is
// implied: before sample 2 a fresh observer has c = k - 1 < c_next and is
// not in a plateau. A dead lane goes on updating the search but stores
// nothing; its observer is never read (its flag is 0 and it is not saved).
template <class S>
__device__ int uc_record_u1(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t, int i0,
                            int iters, bool trail, int wvar, double max_value, PeakObserver& o) {
    typedef typename Prec<S>::T T;
    const double hd = (double)h;
    double w0 = o.w0, w1 = o.w1, last_t = o.last_t;
    int cn = (int)o.c_next, mode = o.mode, raw = o.raw;
    const int n2 = (int)o.n - 2;
    const long long stride = o.stride;
    long long off = (long long)(raw - 1) * stride;  // slot of peak number raw - 1 (raw 0: the anchor)
    const unsigned maxp = (unsigned)o.max_peaks;
    double* const pk = o.peaks;
    double* const iv = o.intervals;
    const double ts = o.time_step;
    bool dead = false;
    S snap = s;
    double tsnap = t;
    // PeakObserver::sample's transitions, restated: `fall` (w1 - v > eps)
    // implies !(w1 < v), so a plateau in the window ends on `lt` (drop) or
    // `fall` (confirm), and otherwise goes on (mode 1); a plateau reaching
    // the record's end finishes the search (mode 2).
    auto iter = [&](int k) {
        const double v = s.get_sel(wvar);
        const int c = k - 1;
        const bool inwin = c < n2;
        const bool lt = w1 < v;
        const bool fall = w1 - v > UC_EPS_PEAK_DELTA;
        const bool start = (mode == 0) & (c >= cn) & inwin & (w1 - w0 > UC_EPS_PEAK_DELTA) &
                           (w1 > UC_PEAK_THRESHOLD) & (w1 >= v);
        const bool plateau = (mode == 1) | start;
        const bool ends = plateau & inwin & (lt | fall);
        if (plateau & inwin & fall & !dead) {
            const double delta = uc_peak_delta(w0, w1, v);
            const double value = uc_peak_value(w0, w1, v, delta);
            const double tpk = (double)(c - 1) + delta;
            if ((unsigned)(raw - 1) < maxp) {
                pk[off] = value;
                if (iv) iv[off] = (tpk - last_t) * ts;
            }
            last_t = tpk;
            ++raw;
            off += stride;
        }
        const int in_window = (lt | fall) ? 0 : 1;
        const int next = inwin ? in_window : 2;
        mode = plateau ? next : mode;
        cn = ends ? c + 2 : cn;
        w0 = w1;
        w1 = v;
        uc_rk4(s, a, h, t);
        if (UC_USES_TIME) t += hd;
    };
    for (int i = i0; i < iters;) {
        uc_phase_mark(s);
        iter(i);
        if (!dead && uc_diverged(s, max_value)) { dead = true; snap = s; tsnap = t; }
        if (uc_wave_all(dead)) break;
        const int stop = (iters - i > UC_CHECK_INTERVAL) ? i + UC_CHECK_INTERVAL : iters;
        for (++i; i + 1 < stop; i += 2) {  // two samples per trip: the window rotates without copies
            iter(i);
            iter(i + 1);
        }
        if (i < stop) iter(i++);
    }
    o.w0 = w0; o.w1 = w1; o.last_t = last_t; o.c_next = cn; o.mode = mode; o.raw = raw;
    if (dead) { s = snap; t = tsnap; return 0; }
    if (!trail) return 1;
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    uc_rk4(s, a, h, t);
    if (UC_USES_TIME) t += hd;
    return uc_fixed_point_flag(s, prev, hd, 1);
}

// uc_loop_u with a PeakObserver whose search nobody reads (no peak slots,
// raw loop flags: the recording of the DFT sweep), without the search: the
// same samples stored (while the lane is live), steps, magnitude checks,
// dead-lane rule and flag.
template <class S>
__device__ int uc_trace_u(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t, long long i0,
                          long long iters, bool trail, int pre, int wvar, double max_value, double* traj,
                          long long stride, long long traj_len) {
    const double hd = (double)h;
    bool dead = false;
    S snap = s;
    double tsnap = t;
    auto iter = [&](long long k) {
        const double v = uc_sample_value(s, wvar);
        if (!dead && k < traj_len) traj[k * stride] = v;
        for (int p = 0; p < pre; ++p) {
            uc_rk4(s, a, h, t);
            if (UC_USES_TIME) t += hd;
        }
    };
    for (long long i = i0; i < iters;) {
        uc_phase_mark(s);
        iter(i);
        if (!dead && uc_diverged(s, max_value)) { dead = true; snap = s; tsnap = t; }
        if (uc_wave_all(dead)) break;
        const long long stop = (iters - i > UC_CHECK_INTERVAL) ? i + UC_CHECK_INTERVAL : iters;
#pragma unroll 2
        for (++i; i < stop; ++i) iter(i);
    }
    if (dead) { s = snap; t = tsnap; return 0; }
    if (!trail) return 1;
    double prev[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) prev[i] = s.get(i);
    for (int p = 0; p < pre; ++p) { uc_rk4(s, a, h, t); if (UC_USES_TIME) t += hd; }
    return uc_fixed_point_flag(s, prev, hd, pre);
}

// The recorded iterations [i0, i1) of a wave-uniform record: uc_trace_u when
// only the samples are wanted, uc_record_u1 where it applies, uc_loop_u
// otherwise.
template <class S>
__device__ __forceinline__ int uc_record_u(S& s, const typename Prec<S>::T* a, typename Prec<S>::T h, double& t,
                                           long long i0, long long i1, bool trail, const UcJob& j, PeakObserver& o) {
    if (o.max_peaks == 0 && j.raw_flags != 0 && o.traj_len > 0)
        return uc_trace_u(s, a, h, t, i0, i1, trail, j.pre_scaller, j.writable_var, j.max_value, o.traj, o.stride,
                          o.traj_len);
    if (j.pre_scaller == 1 && j.writable_var >= 0 && o.traj_len == 0 && o.n < (1LL << 30) &&
        (i0 >= 2 || (o.mode == 0 && o.c_next >= 2)))
        return uc_record_u1(s, a, h, t, (int)i0, (int)i1, trail, j.writable_var, j.max_value, o);
    return uc_loop_u(s, a, h, t, i0, i1, trail, j.pre_scaller, j.writable_var, j.max_value, o);
}

// ================================================================ kernels ==
// The recorded run of a bifurcation cell (after a transient that ended with
// `flag`): streaming peaks, the cell's flag, the optional final state.
// Sliced: record units start at unit TU; an unfinished record saves the
// cell (done, flag, t, state) and the observer from word 3 + 2N.
template <bool SL, class S>
__device__ __forceinline__ void uc_bif_record(const UcJob& j, long long lane, S& s, const typename Prec<S>::T* a,
                                              typename Prec<S>::T h, double hd, double& t, bool uni, long long iters,
                                              int flag, long long u0, long long u1, long long TU) {
    constexpr int kCap = UC_SLAB_BIF_WORDS(UC_N);
    constexpr bool sliced = SL;
    PeakObserver obs;
    obs.w0 = 0.0; obs.w1 = 0.0; obs.n = iters; obs.c_next = 2; obs.mode = 0; obs.raw = 0;
    obs.last_t = 0.0; obs.time_step = hd * (double)j.pre_scaller; obs.max_peaks = j.max_peaks;
    obs.stride = j.stride; obs.peaks = j.out0 + lane; obs.intervals = j.out1 ? j.out1 + lane : nullptr;
    obs.traj = j.traj ? j.traj + lane : nullptr; obs.traj_len = j.traj ? j.traj_len : 0;
    auto cell_state = [&](bool store) {
        UcSlab io(j, lane, kCap, store);
        int done = 0;
        io.i32(done);
        io.i32(flag);
        io.f64(t);
        io.st(s);
        io.seek(3 + 2 * UC_N);
        io.f64(obs.w0);
        io.f64(obs.w1);
        io.i64(obs.c_next);
        io.i32(obs.mode);
        io.i32(obs.raw);
        io.f64(obs.last_t);
    };
    if (sliced && u0 >= TU) cell_state(false);  // a record in progress
    if (flag == 1 || flag == -1) {
        if (!sliced) {
            flag = uni ? uc_record_u(s, a, h, t, 0, j.n_iters, true, j, obs)
                       : uc_loop(s, a, h, t, iters, j.pre_scaller, j.writable_var, j.max_value, obs);
        } else {
            const UcRange r = uc_range(j.n_iters, TU, u0, u1);
            const int f = uc_record_u(s, a, h, t, r.i0, r.i1, r.last, j, obs);
            if (f != 0 && !r.last) { cell_state(true); return; }
            flag = f;
        }
    }
    j.flag[lane] = (flag == 1 && !j.raw_flags) ? obs.count() : flag;
    if (j.final_state) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.stride + lane] = s.get(i);
    }
    if (sliced) {
        UcSlab io(j, lane, kCap, true);
        int done = 1;
        io.i32(done);
    }
}

// Bifurcation sweep (calculateDiscreteModelCUDA + peakFinderCUDA fused):
// transient, recorded run with streaming peaks, flag = peak count / -1 / 0.
// F32C with record_f64: the transient in F32C, the recorded run in double.
// Sliced (SL, j.slab set): units are the transient's chunks, then the record's
// chunks of UC_CHECK_INTERVAL samples; the launch that ends the transient
// goes on into the record (and saves the record's state).
template <class S, bool SL = false>
__device__ void uc_bif_cell(const UcJob& j, long long lane) {
    typedef typename Prec<S>::T T;
    const long long cell = j.cell_offset + lane;
    constexpr bool sliced = SL;
    const long long u0 = sliced ? j.unit_begin : 0, u1 = sliced ? j.unit_end : UC_UNITS_ALL;
    constexpr int kCap = UC_SLAB_BIF_WORDS(UC_N);

    double ad[UC_M], x0[UC_N], hd;
    const bool ok = uc_setup(j, cell, ad, x0, hd);
    if (!ok) {
        if (u0 == 0) j.flag[lane] = -1;  // a sliced cell gets its flag from its first launch
        return;
    }
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
    double t = 0.0;
    if (j.raw_flags == 2) {
        // DFT block under an h sweep (dft1dHSweepKernel,
        // U-CUDA/kernels/dft1d_hsweep.template.cu:125-168): block of
        // (int)(t_max / h / pre) samples capped at n_iters, unusable (-1)
        // when it holds <= 2; transient of (int)(transient / h) single steps
        // plus one trailing step, an escape there (0) marks the cell -1; the
        // block's own loop flag otherwise. Never sliced.
        long long blk = (long long)(j.t_max / hd / (double)j.pre_scaller);
        if (blk > j.n_iters) blk = j.n_iters;
        if (blk <= 2) { j.flag[lane] = -1; return; }
        int f = uc_loop(s, a, h, t, (long long)(j.transient_time / hd), 1, 0, j.max_value, none);
        if (f == 0) { j.flag[lane] = -1; return; }
        PeakObserver rec;
        rec.w0 = 0.0; rec.w1 = 0.0; rec.n = blk; rec.c_next = 2; rec.mode = 2; rec.raw = 0;
        rec.last_t = 0.0; rec.time_step = 0.0; rec.max_peaks = 0; rec.stride = j.stride;
        rec.peaks = nullptr; rec.intervals = nullptr; rec.traj = j.traj + lane; rec.traj_len = (int)blk;
        j.flag[lane] = uc_loop(s, a, h, t, blk, j.pre_scaller, j.writable_var, j.max_value, rec);
        return;
    }
    const bool uni = !uc_sweeps_step(j);  // step counts are wave-uniform unless the sweep is over h
    int flag = 1;
    const long long TU = sliced ? uc_chunks(j.skip_steps) : 0;
    if (sliced && u0 > 0) {
        UcSlab io(j, lane, kCap, false);
        int done = 0;
        io.i32(done);
        if (done) return;
        if (u0 < TU) {  // inside the transient; the record loads its own state
            io.i32(flag);
            io.f64(t);
            io.st(s);
        }
    }
    if (!sliced) {
        flag = uni ? uc_loop_u(s, a, h, t, j.skip_steps, j.pre_scaller, 0, j.max_value, none)
                   : uc_loop(s, a, h, t, skip, j.pre_scaller, 0, j.max_value, none);
    } else if (u0 < TU) {
        const UcRange r = uc_range(j.skip_steps, 0, u0, u1);
        const int f = uc_loop_u(s, a, h, t, r.i0, r.i1, r.last, j.pre_scaller, 0, j.max_value, none);
        if (f != 0 && !r.last) {
            UcSlab io(j, lane, kCap, true);
            int done = 0;
            io.i32(done);
            io.i32(flag);
            io.f64(t);
            io.st(s);
            return;
        }
        flag = f;
    }
    if constexpr (sizeof(T) == 4) {
        if (j.record_f64) {
            // The recorded window in double from the compensated state, with
            // the double parameters and step: peaks are the output, and float
            // stages would displace a weakly contracting cycle by ~1e-6.
            StateF64 s64;
#pragma unroll
            for (int i = 0; i < UC_N; ++i) s64.x[i] = s.get(i);
            uc_bif_record<SL>(j, lane, s64, ad, hd, hd, t, uni, iters, flag, u0, u1, TU);
            return;
        }
    }
    uc_bif_record<SL>(j, lane, s, a, h, hd, t, uni, iters, flag, u0, u1, TU);
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
#if UC_STRICT
#pragma unroll
        for (int i = 0; i < UC_N; ++i) v[k * UC_N + i] = (nrm == 0.0) ? 0.0 : v[k * UC_N + i] / nrm;
#else
        // one division per vector (the quotients differ by at most an ulp)
        const double inv = (nrm == 0.0) ? 0.0 : 1.0 / nrm;
#pragma unroll
        for (int i = 0; i < UC_N; ++i) v[k * UC_N + i] = v[k * UC_N + i] * inv;
#endif
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
// Results: out0[k * stride + cell] = sum_k / t_max; flag 1, or -1 with 999
// (diverged) / NaN (non-positive step).
// Sliced (SL, j.slab set): units are the transient's chunks, then one per
// block.
// Slab: done, state, t, sum[K], then the block state from word
// 2 + words(S) + K (variational: V, Vs, start2, ksum, ref_mode, lsum,
// prod, esum; finite differences: the K probe states).
template <int K, class S, bool SL = false>
__device__ void uc_lyap_cell(const UcJob& j, long long lane) {
    typedef typename Prec<S>::T T;
    const long long cell = j.cell_offset + lane;
    constexpr bool sliced = SL;
    const long long u0 = sliced ? j.unit_begin : 0, u1 = sliced ? j.unit_end : UC_UNITS_ALL;
    constexpr int kCap = UC_SLAB_LYAP_WORDS(UC_N, K);
    constexpr int kBlockWord = 2 + UcWords<S>::n + K;

    double ad[UC_M], x0[UC_N], hd;
    if (!uc_setup(j, cell, ad, x0, hd)) {
        if (u0 == 0) {  // a sliced cell gets its results from its first launch
#pragma unroll
            for (int k = 0; k < K; ++k) j.out0[(long long)k * j.stride + lane] = __builtin_nan("");
            j.flag[lane] = -1;
        }
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
    double sum[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sum[k] = 0.0;
    bool alive = true;
    double t = 0.0;
    if (u0 == 0) {
        s.load(x0);
    } else {
        UcSlab io(j, lane, kCap, false);
        int done = 0;
        io.i32(done);
        if (done) return;
        io.st(s);
        io.f64(t);
#pragma unroll
        for (int k = 0; k < K; ++k) io.f64(sum[k]);
    }

    NoObserver none;
    // step counts are wave-uniform unless the sweep is over h (sliced cells
    // always are)
    const bool uni = !uc_sweeps_step(j);
    auto run = [&](S& st, double& tt, long long n_job, long long n_cell) -> int {
        return uni ? uc_loop_u(st, a, h, tt, n_job, 1, 0, j.max_value, none)
                   : uc_loop(st, a, h, tt, n_cell, 1, 0, j.max_value, none);
    };
    const long long TU = sliced ? uc_chunks(j.skip_steps) : 0;
    if (!sliced) {
        if (run(s, t, j.skip_steps, skip) == 0) alive = false;
    } else if (u0 < TU) {
        const UcRange r = uc_range(j.skip_steps, 0, u0, u1);
        if (uc_loop_u(s, a, h, t, r.i0, r.i1, r.last, 1, 0, j.max_value, none) == 0) alive = false;
    }
    // blocks [b0, b1) in this launch; the block state starts fresh in the
    // launch that ends the transient and is loaded after that
    const bool in_blocks = !sliced || u1 >= TU;
    const bool fresh = !sliced || u0 < TU;
    const long long b0 = (!sliced || u0 < TU) ? 0 : u0 - TU;
    const long long b1 = (!sliced || u1 - TU >= j.n_iters) ? j.n_iters : u1 - TU;
    const bool ends = b1 >= j.n_iters;

    if (alive && in_blocks) {
        double z[K * UC_N];
        uc_probe_basis<K>(j.probe_seed, cell, z);
        if (j.method == UC_METHOD_VARIATIONAL) {
            double d[K];
            if (K > 1) uc_mgs<K>(z, d);
            T V[K * UC_N];
#pragma unroll
            for (int i = 0; i < K * UC_N; ++i) V[i] = Prec<S>::to_t(z[i]);
#if !UC_STRICT
            // K = 1: the tangent equation is linear, so rescaling V between
            // blocks only telescopes the per-block log norms. V is rescaled
            // by a power of two (exact) to |V|^2 in [0.5, 2), the exponents
            // are summed as an integer and one log of the final |V| closes
            // the sum: sum_b log|v_b| = ksum * ln 2 + log|V_end|. A block
            // whose growth |V_end| / |V_start| would reach the reference's
            // 1e-14 floor switches the lane to the reference's unit
            // renormalisation (with the sum so far) for the rest of the run.
            // The norm and the exponent are taken in the state's precision
            // (FP32 for F32C): a power-of-two rescale is exact in either,
            // and every double-precision quantity of the sum is formed from
            // the rescaled vector itself, at the end or at the switch.
            // K > 1: Gram-Schmidt every block as the reference; the norms
            // are multiplied into a mantissa / exponent pair instead of one
            // log per norm and block.
            int ksum = 0;
            bool ref_mode = false;
            T start2 = (T)1.0;        // |V|^2 at the block start, for the floor test
            T Vs[UC_N];               // V at the block start, for the switch
            double lsum = 0.0;
            const T big = sizeof(T) == 4 ? (T)__FLT_MAX__ : (T)1e300;
#pragma unroll
            for (int i = 0; i < UC_N; ++i) Vs[i] = V[i];
            double prod[K];
            int esum[K];
#pragma unroll
            for (int k = 0; k < K; ++k) { prod[k] = 1.0; esum[k] = 0; }
            // one check per block (iteration 0) when the block is no longer
            // than the check interval and the counts are wave-uniform
            const bool one_check = K == 1 && uni && j.nt_steps >= 1 && j.nt_steps <= UC_CHECK_INTERVAL;
#else
            const bool one_check = false;
#endif
            auto block_state = [&](bool store) {
                UcSlab io(j, lane, kCap, store);
                io.seek(kBlockWord);
#pragma unroll
                for (int i = 0; i < K * UC_N; ++i) io.val(V[i]);
#if !UC_STRICT
#pragma unroll
                for (int i = 0; i < UC_N; ++i) io.val(Vs[i]);
                io.val(start2);
                io.i32(ksum);
                io.b8(ref_mode);
                io.f64(lsum);
#pragma unroll
                for (int k = 0; k < K; ++k) { io.f64(prod[k]); io.i32(esum[k]); }
#endif
            };
            if (!fresh) block_state(false);
            for (long long b = b0; b < b1; ++b) {
                const int r = one_check ? uc_block_tan_u<K>(s, V, a, h, t, (int)j.nt_steps, j.max_value)
                              : uni     ? uc_loop_tan_u<K>(s, V, a, h, t, j.nt_steps, j.max_value)
                                        : uc_loop_tan<K>(s, V, a, h, t, nts, j.max_value);
                if (r == 0) { alive = false; break; }
                double v[K * UC_N];
                if (K == 1) {
#if !UC_STRICT
                    T n2 = (T)0.0;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) n2 = n2 + V[i] * V[i];
                    if (!ref_mode && n2 > (T)1e-28 * start2 && n2 <= big) {
                        const int k = uc_frexp_exp(n2) >> 1;  // floor(e / 2)
#pragma unroll
                        for (int i = 0; i < UC_N; ++i) {
                            V[i] = uc_ldexp(V[i], -k);
                            Vs[i] = V[i];
                        }
                        ksum += k;
                        start2 = uc_ldexp(n2, -2 * k);
                        continue;
                    }
                    double inv0 = 1.0;
                    if (!ref_mode) {
                        // the reference's sum up to this block and its block-start unit vector
                        double s2 = 1.0;
                        if (b > 0) {
                            s2 = 0.0;
#pragma unroll
                            for (int i = 0; i < UC_N; ++i) s2 = s2 + (double)Vs[i] * (double)Vs[i];
                        }
                        lsum = (double)ksum * UC_LN2 + 0.5 * log(s2);
                        inv0 = 1.0 / sqrt(s2);
                        ref_mode = true;
                    }
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) v[i] = (double)V[i] * inv0;
#else
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) v[i] = (double)V[i];
#endif
                    double nrm = 0.0;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) nrm = nrm + v[i] * v[i];
                    nrm = sqrt(nrm);
                    if (nrm <= 1e-14) nrm = 1e-14;
#if UC_STRICT
                    sum[0] = sum[0] + log(nrm);
#else
                    lsum = lsum + log(nrm);
#endif
                    const double inv = 1.0 / nrm;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) v[i] = v[i] * inv;
                } else {
#pragma unroll
                    for (int i = 0; i < K * UC_N; ++i) v[i] = (double)V[i];
                    uc_mgs<K>(v, d);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
#if UC_STRICT
                        sum[k] = sum[k] + log(d[k]);
#else
                        int e;
                        prod[k] = __builtin_frexp(prod[k] * d[k], &e);
                        esum[k] += e;
#endif
                    }
                }
#pragma unroll
                for (int i = 0; i < K * UC_N; ++i) V[i] = Prec<S>::to_t(v[i]);
            }
            if (alive && !ends) {
                block_state(true);
            } else {
#if !UC_STRICT
                if (K == 1) {
                    double s2 = 1.0;
                    if (j.n_iters > 0) {
                        s2 = 0.0;
#pragma unroll
                        for (int i = 0; i < UC_N; ++i) s2 = s2 + (double)V[i] * (double)V[i];
                    }
                    sum[0] = ref_mode ? lsum : (double)ksum * UC_LN2 + 0.5 * log(s2);
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) sum[k] = log(prod[k]) + (double)esum[k] * UC_LN2;
                }
#endif
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
            auto probe_state = [&](bool store) {
                UcSlab io(j, lane, kCap, store);
                io.seek(kBlockWord);
#pragma unroll
                for (int k = 0; k < K; ++k) io.st(y[k]);
            };
            if (fresh) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    double yk[UC_N];
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) yk[i] = q[k * UC_N + i] * eps + s.get(i);
                    y[k].load(yk);
                }
            } else {
                probe_state(false);
            }
            for (long long b = b0; b < b1 && alive; ++b) {
                const double t0 = t;
                if (run(s, t, j.nt_steps, nts) == 0) { alive = false; break; }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    double tk = t0;
                    if (alive && run(y[k], tk, j.nt_steps, nts) == 0) alive = false;
                }
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
            if (alive && !ends) probe_state(true);
        }
    }
    if (alive && !(in_blocks && ends)) {
        // sliced, unfinished: the launch ended inside the transient or the blocks
        UcSlab io(j, lane, kCap, true);
        int done = 0;
        io.i32(done);
        io.st(s);
        io.f64(t);
#pragma unroll
        for (int k = 0; k < K; ++k) io.f64(sum[k]);
        return;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) j.out0[(long long)k * j.stride + lane] = alive ? sum[k] / j.t_max : 999.0;
    j.flag[lane] = alive ? 1 : -1;
    if (j.final_state) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.stride + lane] = s.get(i);
    }
    if (sliced) {
        UcSlab io(j, lane, kCap, true);
        int done = 1;
        io.i32(done);
    }
}

#if !UC_STRICT
// C cells per lane (C = 1 or 2; with two, cells `lane` and `lane + half`)
// for the fast single-vector variational LLE with wave-uniform counts and
// blocks of at most the check interval (the host picks this kernel only
// then). Each cell runs exactly the steps, magnitude checks, rescales and
// switch of uc_lyap_cell's fast path. Two cells' instruction streams are
// independent, which doubles the work a wave has in flight where the grid
// gives few waves per SIMD; one cell per lane is the same lean loop for
// grids with enough waves. A cell found diverged keeps stepping with the
// other (its results are discarded) and gets its state at the check back.
template <class S, int C>
__device__ void uc_lle_lanes(const UcJob& j, long long lane, long long half) {
    static_assert(C == 1 || C == 2, "one or two cells per lane");
    typedef typename Prec<S>::T T;
    long long cell[2] = {j.cell_offset + lane, j.cell_offset + lane + half};
    const bool has1 = C == 2 && lane + half < j.n_cells;
    if (!has1) cell[1] = cell[0];  // a stand-in copy of cell 0, never written
    double ad[UC_M], x0[UC_N], hd = 0.0;
    T a[2][UC_M];
    S s[2], snap[2];
    bool dead[2] = {true, true}, ok[2] = {false, false};
#pragma unroll
    for (int c = 0; c < C; ++c) {
        ok[c] = uc_setup(j, cell[c], ad, x0, hd);
#pragma unroll
        for (int i = 0; i < UC_M; ++i) a[c][i] = Prec<S>::to_t(ad[i]);
        s[c].load(x0);
        snap[c] = s[c];
        dead[c] = !ok[c] || (c == 1 && !has1);
    }
    // the step is the same for every lane (no sweep over h on this path)
    StepU<T> h;
    uc_job_step(j, h);
    const double hdd = (double)h.h;
    double t = 0.0;

    // transient: uc_loop_u with pre = 1 and no observer, per cell
    for (long long i = 0; i < j.skip_steps;) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uc_phase_mark(s[c]);
            uc_rk4(s[c], a[c], h, t);
        }
        if (UC_USES_TIME) t += hdd;
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (!dead[c] && uc_diverged(s[c], j.max_value)) { dead[c] = true; snap[c] = s[c]; }
        if (uc_wave_all(dead[0] && (C == 1 || dead[1]))) break;
        const long long stop = (j.skip_steps - i > UC_CHECK_INTERVAL) ? i + UC_CHECK_INTERVAL : j.skip_steps;
#pragma unroll 2
        for (++i; i < stop; ++i) {
#pragma unroll
            for (int c = 0; c < C; ++c) uc_rk4(s[c], a[c], h, t);
            if (UC_USES_TIME) t += hdd;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) uc_rk4(s[c], a[c], h, t);  // the trailing step
    if (UC_USES_TIME) t += hdd;

    T V[2][UC_N], Vs[2][UC_N], start2[2];
    int ksum[2];
    bool ref_mode[2];
    double lsum[2];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        double z[UC_N];
        uc_probe_basis<1>(j.probe_seed, cell[c], z);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) { V[c][i] = Prec<S>::to_t(z[i]); Vs[c][i] = V[c][i]; }
        start2[c] = (T)1.0;
        ksum[c] = 0;
        ref_mode[c] = false;
        lsum[c] = 0.0;
    }
    const T big = sizeof(T) == 4 ? (T)__FLT_MAX__ : (T)1e300;
    const int nts = (int)j.nt_steps;
    for (long long b = 0; b < j.n_iters; ++b) {
        // uc_block_tan_u for each cell, with the magnitude check and the
        // power-of-two rescale as selects: the per-cell branches of the
        // check cost more than the check at blocks of ~10 steps
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uc_phase_mark(s[c]);
            uc_rk4_tan<1>(s[c], V[c], a[c], h, t);
        }
        if (UC_USES_TIME) t += hdd;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const bool now = !dead[c] && uc_diverged_u(s[c], j);
            uc_select(snap[c], s[c], now);  // a lane found diverged keeps its state at the check
            dead[c] = dead[c] || now;
        }
        // the rest of the block; blocks of UC_LLE_BLOCK steps (h = NT / 10,
        // the usual renormalisation interval) run as straight-line code
        if (nts == UC_LLE_BLOCK) {
#pragma unroll
            for (int i = 0; i < UC_LLE_BLOCK; ++i) {
#pragma unroll
                for (int c = 0; c < C; ++c) uc_rk4_tan<1>(s[c], V[c], a[c], h, t);
                if (UC_USES_TIME) t += hdd;
            }
        } else {
#pragma unroll 4
            for (int i = 0; i < nts; ++i) {
#pragma unroll
                for (int c = 0; c < C; ++c) uc_rk4_tan<1>(s[c], V[c], a[c], h, t);
                if (UC_USES_TIME) t += hdd;
            }
        }
        // a wave whose cells have all diverged stops here: its extra steps
        // changed no result (each such cell's output comes from its check)
        if (uc_wave_all(dead[0] && (C == 1 || dead[1]))) break;
        // the fast path's renormalisation: |V|^2 back into [0.5, 2) by a
        // power of two (k = 0 leaves V as it is on the other lanes)
        bool need[2] = {false, false};
#pragma unroll
        for (int c = 0; c < C; ++c) {
            T n2 = (T)0.0;
#pragma unroll
            for (int i = 0; i < UC_N; ++i) n2 = n2 + V[c][i] * V[c][i];
            const bool fast = !ref_mode[c] && n2 > (T)1e-28 * start2[c] && n2 <= big;
            const int k = fast ? (uc_frexp_exp(n2) >> 1) : 0;
#pragma unroll
            for (int i = 0; i < UC_N; ++i) V[c][i] = uc_ldexp(V[c][i], -k);
            ksum[c] += k;
            start2[c] = fast ? uc_ldexp(n2, -2 * k) : start2[c];
            need[c] = !dead[c] && !fast;
        }
        // the rest switch to (or stay in) the reference's unit
        // renormalisation in double; Vs still holds their block-start vector
        if (!uc_wave_all(!need[0] && (C == 1 || !need[1]))) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (!need[c]) continue;
                double inv0 = 1.0;
                if (!ref_mode[c]) {
                    double s2 = 1.0;
                    if (b > 0) {
                        s2 = 0.0;
#pragma unroll
                        for (int i = 0; i < UC_N; ++i) s2 = s2 + (double)Vs[c][i] * (double)Vs[c][i];
                    }
                    lsum[c] = (double)ksum[c] * UC_LN2 + 0.5 * log(s2);
                    inv0 = 1.0 / sqrt(s2);
                    ref_mode[c] = true;
                }
                double v[UC_N], nrm = 0.0;
#pragma unroll
                for (int i = 0; i < UC_N; ++i) { v[i] = (double)V[c][i] * inv0; nrm = nrm + v[i] * v[i]; }
                nrm = sqrt(nrm);
                if (nrm <= 1e-14) nrm = 1e-14;
                lsum[c] = lsum[c] + log(nrm);
                const double inv = 1.0 / nrm;
#pragma unroll
                for (int i = 0; i < UC_N; ++i) V[c][i] = Prec<S>::to_t(v[i] * inv);
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int i = 0; i < UC_N; ++i) Vs[c][i] = V[c][i];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (c == 1 && !has1) break;
        const long long out = lane + (c == 1 ? half : 0);
        if (!ok[c]) {
            j.out0[out] = __builtin_nan("");
            j.flag[out] = -1;
            continue;
        }
        if (dead[c]) s[c] = snap[c];
        double sum = lsum[c];
        if (!ref_mode[c]) {
            double s2 = 1.0;
            if (j.n_iters > 0) {
                s2 = 0.0;
#pragma unroll
                for (int i = 0; i < UC_N; ++i) s2 = s2 + (double)V[c][i] * (double)V[c][i];
            }
            sum = (double)ksum[c] * UC_LN2 + 0.5 * log(s2);
        }
        j.out0[out] = dead[c] ? 999.0 : sum / j.t_max;
        j.flag[out] = dead[c] ? -1 : 1;
        if (j.final_state) {
#pragma unroll
            for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.stride + out] = s[c].get(i);
        }
    }
}
#endif

// ------------------------------------------- spectrum, one lane per vector --
// A value of lane `src` of the wave (every lane of the wave must be active).
__device__ __forceinline__ double uc_lane_f64(double v, int src) { return __shfl(v, src, 64); }

// The Lyapunov spectrum (K = N tangent vectors, variational method, counts
// the same for every cell) with a cell's vectors on N neighbouring lanes of
// one wave: lane k of the group steps the state (the same arithmetic, so the
// same values, on every lane of the group) and tangent vector k, and the
// modified Gram-Schmidt of each block runs across the group: for q = 0..K-1
// lane q normalises its vector and the group reads it, lanes k > q subtract
// their projection on it. Every value is the one uc_lyap_cell<N> computes
// (same operations in the same order), so the results are bit-identical to
// it, with N times the lanes and 1/N of the tangent work per lane: for grids
// with few waves per SIMD, where one lane per cell is latency-bound.
// `has`: the group holds a cell of the job (else it runs a stand-in copy of
// the cell and writes nothing); every lane of the wave stays active to the
// end, as the group reads need. A lane whose cell diverged keeps stepping
// (its group with it: the state, hence the verdict, is the group's) and its
// results come from its check.
template <class S>
__device__ void uc_ls_spread(const UcJob& j, long long lane, bool has, int k, int base) {
    constexpr int K = UC_N;
    static_assert(UC_N <= 64, "a cell's vectors on the lanes of one wave");
    typedef typename Prec<S>::T T;
    const long long cell = j.cell_offset + lane;
    double ad[UC_M], x0[UC_N], hd;
    const bool ok = uc_setup(j, cell, ad, x0, hd);
    T a[UC_M];
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = Prec<S>::to_t(ad[i]);
    const T h = Prec<S>::to_t(hd);
    S s;
    s.load(x0);
    double t = 0.0;
    NoObserver none;
    bool alive = ok && uc_loop_u(s, a, h, t, j.skip_steps, 1, 0, j.max_value, none) != 0;
    S fin = s;  // the state written out: at the check that found the cell diverged, else at the end

    // the cell's starting basis (the same on every lane of the group), row k
    double z[K * UC_N], d[K];
    uc_probe_basis<K>(j.probe_seed, cell, z);
    if (K > 1) uc_mgs<K>(z, d);
    T V[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        double zi = uc_opaque(z[i]);
#pragma unroll
        for (int q = 1; q < K; ++q) zi = (k == q) ? z[q * UC_N + i] : zi;
        V[i] = Prec<S>::to_t(zi);
    }
    double sum = 0.0, prod = 1.0;
    int esum = 0;
    for (long long b = 0; b < j.n_iters; ++b) {
        const int r = uc_loop_tan_u<1>(s, V, a, h, t, j.nt_steps, j.max_value);
        uc_select(fin, s, alive);  // a live cell's state; a cell that died here, at its check
        alive = alive && r != 0;
        if (uc_wave_all(!alive)) break;
        // uc_mgs<K> across the group
        double v[UC_N];
#pragma unroll
        for (int i = 0; i < UC_N; ++i) v[i] = (double)V[i];
        double dk = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (k == q) {
                double nrm = 0.0;
#pragma unroll
                for (int i = 0; i < UC_N; ++i) nrm = nrm + v[i] * v[i];
                nrm = sqrt(nrm);
                dk = nrm;
#if UC_STRICT
#pragma unroll
                for (int i = 0; i < UC_N; ++i) v[i] = (nrm == 0.0) ? 0.0 : v[i] / nrm;
#else
                const double inv = (nrm == 0.0) ? 0.0 : 1.0 / nrm;
#pragma unroll
                for (int i = 0; i < UC_N; ++i) v[i] = v[i] * inv;
#endif
            }
            if (q + 1 < K) {
                double vq[UC_N];
#pragma unroll
                for (int i = 0; i < UC_N; ++i) vq[i] = uc_lane_f64(v[i], base + q);
                if (k > q) {
                    double r2 = 0.0;
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) r2 = r2 + v[i] * vq[i];
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) v[i] = v[i] - r2 * vq[i];
                }
            }
        }
        if (alive) {
#if UC_STRICT
            sum = sum + log(dk);
#else
            int e;
            prod = __builtin_frexp(prod * dk, &e);
            esum += e;
#endif
        }
#pragma unroll
        for (int i = 0; i < UC_N; ++i) V[i] = Prec<S>::to_t(v[i]);
    }
#if !UC_STRICT
    sum = log(prod) + (double)esum * UC_LN2;
#endif
    if (!has) return;
    if (!ok) {
        j.out0[(long long)k * j.stride + lane] = __builtin_nan("");
        if (k == 0) j.flag[lane] = -1;
        return;
    }
    j.out0[(long long)k * j.stride + lane] = alive ? sum / j.t_max : 999.0;
    if (k == 0) {
        j.flag[lane] = alive ? 1 : -1;
        if (j.final_state) {
#pragma unroll
            for (int i = 0; i < UC_N; ++i) j.final_state[(long long)i * j.stride + lane] = fin.get(i);
        }
    }
}

// The same streaming finder over stored series: data[i * n_series + s].
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_peaks_series(const double* data, long long n_samples,
                                                                      long long n_series, double time_step,
                                                                      int max_peaks, int* counts, double* peaks,
                                                                      double* intervals) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= n_series) return;
    PeakObserver obs;
    obs.w0 = 0.0; obs.w1 = 0.0; obs.n = n_samples; obs.c_next = 2; obs.mode = 0; obs.raw = 0;
    obs.last_t = 0.0; obs.time_step = time_step; obs.max_peaks = max_peaks; obs.stride = n_series;
    obs.peaks = peaks + lane; obs.intervals = intervals + lane; obs.traj = nullptr; obs.traj_len = 0;
    for (long long i = 0; i < n_samples; ++i) obs.sample(i, data[i * n_series + lane]);
    counts[lane] = obs.count();
}

// ------------------------------------------------------- launch shapes ---
// Direct launch: one lane per cell of one job.
#define UC_DIRECT(NAME, CELL)                                                   \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(UcJob j) {      \
        const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x; \
        if (lane < j.n_cells) CELL(j, lane);                                    \
    }

// Persistent launch over a job table: each workgroup repeatedly takes the
// next chunk of blockDim cells from a global counter (chunks of all jobs in
// one index space, chunk_start = prefix sums, jobs ordered by cost so the
// longest work is taken first) until the counter passes the end; every
// workgroup reaches that exit, so the grid drains. A workgroup's lanes
// always belong to one job. The waves run at the entry's issue priority
// (the SIMD arbiter prefers higher-priority waves when several are ready),
// so a small job's few waves keep their pace next to a full round's.
__device__ __forceinline__ void uc_set_priority(int p) {
    if (p == 3) __builtin_amdgcn_s_setprio(3);
    else if (p == 2) __builtin_amdgcn_s_setprio(2);
    else if (p == 1) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}
#define UC_PERSISTENT(NAME, CELL)                                                                   \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(const UcJob* __restrict__ jobs,     \
                                                                const long long* __restrict__ start, \
                                                                int n_jobs, u64* counter) {          \
        __shared__ long long s_chunk;                                                               \
        const long long total = start[n_jobs];                                                      \
        for (;;) {                                                                                  \
            if (threadIdx.x == 0) s_chunk = (long long)atomicAdd((unsigned long long*)counter, 1ULL); \
            __syncthreads();                                                                        \
            const long long c = s_chunk;                                                            \
            __syncthreads();                                                                        \
            if (c >= total) break;                                                                  \
            int lo = 0, hi = n_jobs - 1;                                                            \
            while (lo < hi) {                                                                       \
                const int mid = (lo + hi + 1) >> 1;                                                 \
                if (start[mid] <= c) lo = mid; else hi = mid - 1;                                   \
            }                                                                                       \
            lo = __builtin_amdgcn_readfirstlane(lo); /* uniform: job fields become scalar loads */   \
            uc_set_priority(jobs[lo].priority);                                                     \
            const long long lane = (c - start[lo]) * (long long)blockDim.x + threadIdx.x;           \
            if (lane < jobs[lo].n_cells) CELL(jobs[lo], lane);                                      \
        }                                                                                           \
    }

// Period of a 2-D bifurcation cell: the reference's DBSCAN over the cell's
// peak values (dbscan, U-CUDA/cudaLibrary.cu:2088-2166, as compiled into its
// 2-D bifurcation template), the same steps in the same order because the
// cluster count is the map value. labels / slots start at 0; an unlabelled
// peak opens a cluster; every unlabelled peak k < n - 1 within eps (|dp|, the
// interval term being multiplied by 0) joins it with slot[k] = k; while the
// neighbour count is positive the loop variable is replaced by
// slot[count - 1], which is cleared, and the scan repeats from there.
// Result: -1 / 0 / 1 for -1 / 0 / 1 peaks, else clusters - 1.
// Peaks are peak-major ([k * stride + cell]).
//
// uc_bif_dbscan_wave: one wave per cell with n <= UC_DB_WAVE_MAX peaks.
// Within one scan the set of peaks that join does not depend on the order
// of k, so the 64 lanes test k = 64 j + lane at once (j < 16, peaks held in
// registers) and the wave adds the joins with ballots. Only whether a label
// or a slot is non-zero is ever read (slot[k] is k or 0, and slot 0 means
// k = 0 either way), so both are bit masks: bit j of `lab` / `slot` in lane
// l stands for peak 64 j + l. Walking over labelled peaks (neighbour count
// 0) jumps to the next unlabelled one by ballot. Cells with more peaks are
// left to uc_bif_dbscan (one lane per cell, scratch in HBM), launched with
// min_n = UC_DB_WAVE_MAX so that it takes exactly those.
#define UC_DB_WAVE_MAX 1024

__device__ __forceinline__ double uc_readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

//
// The neighbour test is the reference's sqrt(dp^2 + di^2) < eps with di = 0.
// In binary floating point with rounding to nearest, sqrt(fl(x * x)) = |x|
// whenever x * x neither underflows nor overflows (S. Boldo, "Stupid is as
// stupid does: taking the square root of the square of a floating-point
// number", 2015); for 2^-500 <= eps <= 2^500 the cases where it does
// (|dp| < 2^-511: both sides below eps; |dp| >= 2^511.99...: both sides not)
// give the same verdict too, so there the test is |dp| < eps, bit for bit.
//
// Workgroups are dealt round-robin over the chip's XCDs, each with its own
// L2; neighbouring cells share cache lines of the peak-major array, so the
// blocks are renumbered to give each XCD one contiguous range of cells.
#define UC_XCDS 8
__device__ __forceinline__ unsigned uc_xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned per = nb / UC_XCDS, rem = nb % UC_XCDS, x = b % UC_XCDS;
    return x * per + min(x, rem) + b / UC_XCDS;
}

template <bool ABS>
__device__ __forceinline__ bool uc_db_near(double dp, double eps) {
    if (ABS) return fabs(dp) < eps;
    const double di = 0.0;
    return sqrt(dp * dp + di * di) < eps;
}

// JM: the number of 64-peak chunks the cell's loops are unrolled over (a
// power of two >= the cell's chunk count); the chunks past the cell's
// peaks hold no candidates, so they never join.
template <bool ABS, int JM>
__device__ void uc_dbscan_wave_cell(const double* __restrict__ peaks, long long stride, long long c, int n, double eps,
                                    int* out);

template <bool ABS, bool WIDE>
__device__ __forceinline__ void uc_dbscan_wave_jm(const double* __restrict__ peaks, long long stride, long long c,
                                                  int n, double eps, int* out) {
    const int J = (n + 63) >> 6;
    if (WIDE) uc_dbscan_wave_cell<ABS, 16>(peaks, stride, c, n, eps, out);
    else if (J <= 1) uc_dbscan_wave_cell<ABS, 1>(peaks, stride, c, n, eps, out);
    else if (J <= 2) uc_dbscan_wave_cell<ABS, 2>(peaks, stride, c, n, eps, out);
    else if (J <= 4) uc_dbscan_wave_cell<ABS, 4>(peaks, stride, c, n, eps, out);
    else uc_dbscan_wave_cell<ABS, 8>(peaks, stride, c, n, eps, out);
}

// Cells of up to UC_DB_NARROW_MAX peaks (and the 0 / 1 / -1 counts) in
// uc_bif_dbscan_wave, the rest up to UC_DB_WAVE_MAX in uc_bif_dbscan_wave_wide:
// the 16-chunk form needs twice the registers, which would halve the
// waves per SIMD of the common one.
#define UC_DB_NARROW_MAX 512
template <bool WIDE>
__device__ __forceinline__ void uc_dbscan_wave_run(const double* __restrict__ peaks, const int* __restrict__ counts,
                                                   long long stride, long long n_cells, int max_n, double eps,
                                                   int* out) {
    const int lane = threadIdx.x & 63;
    const long long c = (long long)uc_xcd_block() * (UC_BLOCK / 64) + (threadIdx.x >> 6);
    if (c >= n_cells) return;  // wave-uniform
    const int n = __builtin_amdgcn_readfirstlane(min(counts[c], max_n));
    if (n <= 1) {
        if (!WIDE && lane == 0) out[c] = n;
        return;
    }
    if (WIDE ? (n <= UC_DB_NARROW_MAX || n > UC_DB_WAVE_MAX) : n > UC_DB_NARROW_MAX) return;
    if (eps >= 0x1p-500 && eps <= 0x1p500) uc_dbscan_wave_jm<true, WIDE>(peaks, stride, c, n, eps, out);
    else uc_dbscan_wave_jm<false, WIDE>(peaks, stride, c, n, eps, out);
}

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bif_dbscan_wave(const double* __restrict__ peaks,
                                                                            const int* __restrict__ counts,
                                                                            long long stride, long long n_cells,
                                                                            int max_n, double eps, int* out) {
    uc_dbscan_wave_run<false>(peaks, counts, stride, n_cells, max_n, eps, out);
}
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bif_dbscan_wave_wide(const double* __restrict__ peaks,
                                                                                 const int* __restrict__ counts,
                                                                                 long long stride, long long n_cells,
                                                                                 int max_n, double eps, int* out) {
    uc_dbscan_wave_run<true>(peaks, counts, stride, n_cells, max_n, eps, out);
}

template <bool ABS, int JM>
__device__ void uc_dbscan_wave_cell(const double* __restrict__ peaks, long long stride, long long c, int n, double eps,
                                    int* out) {
    const int lane = threadIdx.x & 63;
    const int J = (n + 63) >> 6;
    double p[JM];
    unsigned cand = 0;  // bit j: peak 64 j + lane exists and is < n - 1 (scanned as a neighbour)
#pragma unroll
    for (int j = 0; j < JM; ++j) {
        const int k = 64 * j + lane;
        p[j] = k < n ? peaks[(long long)k * stride + c] : 0.0;
        if (k < n - 1) cand |= 1u << j;
    }
    unsigned lab = 0, slot = 0;
    int cluster = 0, nb = 0, i = 0;
    while (i < n) {
        if (nb >= 1) {
            const int idx = nb - 1;
            const unsigned w = (unsigned)__builtin_amdgcn_readlane((int)slot, idx & 63);
            if (lane == (idx & 63)) slot &= ~(1u << (idx >> 6));
            i = ((w >> (idx >> 6)) & 1u) ? idx : 0;
            nb -= 1;
        } else {
            int next = n;
            for (int j = i >> 6; j < J; ++j) {
                const int k = 64 * j + lane;
                const u64 m = __builtin_amdgcn_ballot_w64(k >= i && k < n && !((lab >> j) & 1u));
                if (m) {
                    next = 64 * j + (int)__builtin_ctzll(m);
                    break;
                }
            }
            if (next >= n) break;
            i = next;
            ++cluster;
            if (lane == (i & 63)) lab |= 1u << (i >> 6);
        }
        double pi = 0.0;
#pragma unroll
        for (int j = 0; j < JM; ++j)
            if (j == (i >> 6)) pi = uc_readlane_f64(p[j], i & 63);
        const unsigned open = cand & ~lab;
        unsigned joined = 0;
#pragma unroll
        for (int j = 0; j < JM; ++j) {
            const bool hit = ((open >> j) & 1u) && (64 * j + lane) != i && uc_db_near<ABS>(pi - p[j], eps);
            nb += __builtin_popcountll(__builtin_amdgcn_ballot_w64(hit));
            if (hit) joined |= 1u << j;
        }
        lab |= joined;
        slot |= joined;
        ++i;
    }
    if (lane == 0) out[c] = cluster - 1;
}

// uc_bif_dbscan: one lane per cell with more than min_n peaks (the others
// are left alone); labels / slots are its scratch, peak-major like peaks.
// Both kernels read a count above max_n (the rows of peaks) as max_n.
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bif_dbscan(const double* __restrict__ peaks,
                                                                       const int* __restrict__ counts, long long stride,
                                                                       long long n_cells, int max_n, double eps,
                                                                       int min_n, int* labels, int* slots, int* out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int n = min(counts[c], max_n);
    if (n <= min_n) return;
    if (n <= 1) { out[c] = n; return; }
    int cluster = 0, neighbours = 0;
    for (int q = 0; q < n; ++q) { labels[(long long)q * stride + c] = 0; slots[(long long)q * stride + c] = 0; }
    for (long long i = 0; i < n; ++i) {
        if (neighbours >= 1) {
            i = slots[(long long)(neighbours - 1) * stride + c];
            slots[(long long)(neighbours - 1) * stride + c] = 0;
            neighbours -= 1;
        } else if (labels[i * stride + c] == 0) {
            neighbours = 0;
            ++cluster;
            labels[i * stride + c] = cluster;
        } else {
            continue;
        }
        const double pi = peaks[i * stride + c];
        for (int k = 0; k < n - 1; ++k) {
            if (i != k && labels[(long long)k * stride + c] == 0) {
                const double dp = pi - peaks[(long long)k * stride + c], di = 0.0;
                if (sqrt(dp * dp + di * di) < eps) {
                    labels[(long long)k * stride + c] = cluster;
                    slots[(long long)k * stride + c] = k;
                    ++neighbours;
                }
            }
        }
    }
    out[c] = cluster - 1;
}

// 1-D DFT of recorded blocks (DFT_custom, U-CUDA/cudaLibrary.cu:1698-1777):
// a_k + i b_k = (1/N) sum_n w_n x_n (cos, sin)(phase_n), the phase advanced
// by the per-sample rotation tab[2k], tab[2k+1] (cos, sin) and
// resynchronised after every 1000 samples (while samples remain) from
// resync[2 (k * n_res + r)], the float-precision cos / sin of the exact
// phase. The host computes tab and resync, so the kernel only multiplies
// and adds. Blocks flagged -1 / 0 give -1 / 0 for every coefficient. traj
// is sample-major ([n * stride + cell]); the coefficients are cell-major
// ([cell * n_freq + k]).
//
// Work split: one workgroup per (group of UC_DFT_C consecutive cells,
// UC_BLOCK consecutive frequencies), one thread per frequency. The rotation
// depends on the frequency alone, so each thread advances it once per
// sample for the whole group; each cell's sums see the same operations in
// the same order as a thread of its own would. Blocks of UC_DFT_U samples of
// the group's cells (and the window) are staged in LDS by one coalesced
// load per thread, then every thread reads them as LDS broadcasts.
#define UC_DFT_U 32
#define UC_DFT_C 8
static_assert(UC_DFT_U * UC_DFT_C <= UC_BLOCK, "one staged sample per thread");

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_dft(const double* __restrict__ traj, long long stride,
                                                                long long n_cells, int N, const int* __restrict__ flags,
                                                                const double* __restrict__ window, int n_freq,
                                                                const double* __restrict__ tab,
                                                                const double* __restrict__ resync, int n_res,
                                                                double* __restrict__ ak, double* __restrict__ bk) {
    __shared__ double xs[UC_DFT_U][UC_DFT_C];
    __shared__ double ws[UC_DFT_U];
    const int kblocks = (n_freq + UC_BLOCK - 1) / UC_BLOCK;
    const long long g = (long long)blockIdx.x / kblocks;
    const int k_raw = (int)((long long)blockIdx.x - g * kblocks) * UC_BLOCK + (int)threadIdx.x;
    const int k = min(k_raw, n_freq - 1);  // spare threads repeat the last frequency and store nothing
    const long long c0 = g * UC_DFT_C;
    const int nc = (int)min((long long)UC_DFT_C, n_cells - c0);
    const double ct = tab[2 * k], st = tab[2 * k + 1];
    double a[UC_DFT_C], b[UC_DFT_C];
#pragma unroll
    for (int c = 0; c < UC_DFT_C; ++c) a[c] = b[c] = 0.0;
    double cn = 1.0, sn = 0.0;
    int until = 1000, r = 0;
    // this thread's staging slot: sample lu of the block, cell lc of the group
    const int lu = (int)threadIdx.x / UC_DFT_C, lc = (int)threadIdx.x % UC_DFT_C;
    for (int n0 = 0; n0 < N; n0 += UC_DFT_U) {
        __syncthreads();  // the previous block has been consumed
        if (lu < UC_DFT_U) {
            const int n = n0 + lu;
            xs[lu][lc] = (n < N && lc < nc) ? traj[(long long)n * stride + c0 + lc] : 0.0;
            if (lc == 0) ws[lu] = n < N ? window[n] : 0.0;
        }
        __syncthreads();
        const int m = min(UC_DFT_U, N - n0);
        for (int u = 0; u < m; ++u) {
            const double w = ws[u];
#pragma unroll
            for (int c = 0; c < UC_DFT_C; ++c) {
                const double wd = w * xs[u][c];
                a[c] = a[c] + wd * cn;
                b[c] = b[c] + wd * sn;
            }
            const double nc_ = cn * ct - sn * st;
            const double ns_ = sn * ct + cn * st;
            cn = nc_;
            sn = ns_;
            if (--until == 0) {
                until = 1000;
                if (n0 + u + 1 < N) {
                    cn = resync[2 * ((long long)k * n_res + r)];
                    sn = resync[2 * ((long long)k * n_res + r) + 1];
                    ++r;
                }
            }
        }
    }
    if (k_raw >= n_freq) return;
#pragma unroll
    for (int c = 0; c < UC_DFT_C; ++c) {
        if (c < nc) {
            const long long cell = c0 + c;
            const int flag = flags[cell];
            const long long o = cell * n_freq + k;
            if (flag == -1 || flag == 0) {
                ak[o] = (double)flag;
                bk[o] = (double)flag;
            } else {
                ak[o] = a[c] / (double)N;
                bk[o] = b[c] / (double)N;
            }
        }
    }
}

// DFT of recorded blocks of per-cell length (the DFT under an h sweep,
// U-CUDA/kernels/dft1d_hsweep.template.cu:168-204): the arithmetic of
// uc_dft with everything per cell: N[cell] samples, its window at
// win[win_off[cell]], rotation tab[2 (cell * n_freq + k)] and resync values
// res[res_off[cell] + 2 (k * n_res + r)] with n_res = (N - 1) / 1000. One
// wave per (cell, 64 frequencies); samples and window through scalar loads.
// Loads are unguarded: traj holds ceil(max N / UC_DFTC_U) * UC_DFTC_U rows and
// each window row is padded to a multiple of UC_DFTC_U. UC_DFTC_U divides
// 1000, so a resync can only fall on a block's last sample.
#define UC_DFTC_U 4
static_assert(1000 % UC_DFTC_U == 0, "DFT sample blocks must tile the 1000-sample resync period");

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_dft_cells(
    const double* __restrict__ traj, long long stride, long long n_cells, const int* __restrict__ Ns,
    const int* __restrict__ flags, const double* __restrict__ win, const long long* __restrict__ win_off, int n_freq,
    const double* __restrict__ tab, const double* __restrict__ res, const long long* __restrict__ res_off,
    double* __restrict__ ak, double* __restrict__ bk) {
    const int kchunks = (n_freq + 63) >> 6;
    const long long wave = __builtin_amdgcn_readfirstlane(
        (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const long long cell = wave / kchunks;
    if (cell >= n_cells) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int k_raw = (int)(wave - cell * kchunks) * 64 + lane;
    const int k = min(k_raw, n_freq - 1);
    const int flag = flags[cell];
    if (flag == -1 || flag == 0) {
        if (k_raw < n_freq) {
            ak[cell * n_freq + k] = (double)flag;
            bk[cell * n_freq + k] = (double)flag;
        }
        return;
    }
    const int N = Ns[cell];
    const int n_res = (N - 1) / 1000;
    const double ct = tab[2 * (cell * n_freq + k)], st = tab[2 * (cell * n_freq + k) + 1];
    const double* w = win + win_off[cell];
    const double* x = traj + cell;
    const double* rs = res + res_off[cell] + 2LL * k * n_res;
    double a = 0.0, b = 0.0, cn = 1.0, sn = 0.0;
    int r = 0;
    for (int n0 = 0; n0 < N; n0 += UC_DFTC_U) {
        double xv[UC_DFTC_U], wv[UC_DFTC_U];
#pragma unroll
        for (int u = 0; u < UC_DFTC_U; ++u) {
            wv[u] = w[n0 + u];
            xv[u] = x[(long long)(n0 + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < UC_DFTC_U; ++u) {
            if (n0 + u < N) {
                const double wd = wv[u] * xv[u];
                a = a + wd * cn;
                b = b + wd * sn;
                const double nc_ = cn * ct - sn * st;
                const double ns_ = sn * ct + cn * st;
                cn = nc_;
                sn = ns_;
            }
        }
        if ((n0 + UC_DFTC_U) % 1000 == 0 && n0 + UC_DFTC_U < N) {
            cn = rs[2 * r];
            sn = rs[2 * r + 1];
            ++r;
        }
    }
    if (k_raw >= n_freq) return;
    ak[cell * n_freq + k] = a / (double)N;
    bk[cell * n_freq + k] = b / (double)N;
}

// ========================================================== continuation ==
// Continuation sweeps (the reference's bifurcation1d_cont / lle1d_cont /
// ls1d_cont / dft1d_cont templates, U-CUDA/kernels/*_cont.template.cu): the
// points of axis 0 are visited in order and each starts from the state the
// previous one left. One lane per chain: axis 1 (optional) gives independent
// chains, cell = chain * nx + point. axis_vals[0] holds the chain's points
// in visiting order (the host applies the reverse direction). FP64.
__device__ __forceinline__ void uc_chain_setup(const UcJob& J, long long chain, double* a, double* x0, double& h) {
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = (i < J.n_values) ? J.base_params[i] : 0.0;
    if (J.ic_seed != 0ULL) {
        UcRng r;
        r.init(J.ic_seed, (u64)chain, 0ULL);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x0[i] = J.ic_lo + (J.ic_hi - J.ic_lo) * (double)r.uniform();
    } else {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x0[i] = J.base_ic[i];
    }
    h = J.h;
    if (J.axis_kind[1] != UC_AXIS_NONE) {
        const double v = J.axis_vals[1][chain % J.ny];
        const int idx = J.axis_index[1];
        if (J.axis_kind[1] == UC_AXIS_PARAM) {
#pragma unroll
            for (int i = 0; i < UC_M; ++i) a[i] = (i == idx) ? v : a[i];
        } else if (J.axis_kind[1] == UC_AXIS_VAR) {
#pragma unroll
            for (int i = 0; i < UC_N; ++i) x0[i] = (i == idx) ? v : x0[i];
        } else {
            h = v;
        }
    }
}

// The chain's point jj: the parameter a[axis_index[0]] or the step.
__device__ __forceinline__ double uc_chain_point(const UcJob& J, long long jj, double* a, double h) {
    const double p = J.axis_vals[0][jj];
    if (J.axis_kind[0] == UC_AXIS_STEP) return p;
    const int idx = J.axis_index[0];
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = (i == idx) ? p : a[i];
    return h;
}

// Bifurcation chain (bifurcation1dContinuationKernel, :56-140, then
// peakFinderCUDA over each point's block with the base step h * pre as the
// time step): block of (int)(t_max / h / pre) samples capped at n_iters,
// -1 when empty or h <= 0; transient of (int)(transient / h) single steps
// plus one trailing step, an escape there -> -1 with the state carried on;
// flag = -1 when the block ends on a fixed point, else its peak count.
// raw_flags 1: the DFT chain (dft1dContinuationKernel, :72-175): a block of
// <= 2 samples, an escaping transient or a block ending on a fixed point
// marks the point -1 and restarts the chain from the initial state; other
// points get flag 1 and their block in traj. A block that escapes part way
// gets flag 0 in both (the reference goes on to read the block's unwritten
// samples there, which has no reproducible value); the DFT chain restarts
// after it as after its other unusable points.
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_cont_bif(UcJob J) {
    const long long chain = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= J.ny) return;
    double ad[UC_M], x0[UC_N], hb;
    uc_chain_setup(J, chain, ad, x0, hb);
    StateF64 s;
    s.load(x0);
    double t = 0.0;
    NoObserver none;
    const bool dft = J.raw_flags == 1;
    for (long long jj = 0; jj < J.nx; ++jj) {
        const long long cell = chain * J.nx + jj;
        const double h = uc_chain_point(J, jj, ad, hb);
        long long blk = h > 0.0 ? (long long)(J.t_max / h / (double)J.pre_scaller) : 0;
        if (blk > J.n_iters) blk = J.n_iters;
        const long long skip = h > 0.0 ? (long long)(J.transient_time / h) : 0;
        int flag;
        if (h <= 0.0 || blk <= (dft ? 2 : 0)) {
            flag = -1;
        } else {
            flag = uc_loop(s, ad, h, t, skip, 1, 0, J.max_value, none);
            if (flag == 0) {
                flag = -1;
            } else {
                PeakObserver obs;
                obs.w0 = 0.0; obs.w1 = 0.0; obs.n = blk; obs.c_next = 2; obs.mode = dft ? 2 : 0; obs.raw = 0;
                obs.last_t = 0.0; obs.time_step = J.h * (double)J.pre_scaller; obs.max_peaks = J.max_peaks;
                obs.stride = J.stride; obs.peaks = J.out0 ? J.out0 + cell : nullptr;
                obs.intervals = J.out1 ? J.out1 + cell : nullptr;
                obs.traj = J.traj ? J.traj + cell : nullptr; obs.traj_len = J.traj ? (int)blk : 0;
                const int f = uc_loop(s, ad, h, t, blk, J.pre_scaller, J.writable_var, J.max_value, obs);
                flag = (f == -1) ? -1 : (f == 0) ? 0 : (dft ? 1 : obs.count());
            }
        }
        J.flag[cell] = flag;
        if (dft && flag != 1) s.load(x0);
    }
}

// Finite-difference Lyapunov chain, K = 1 (lle1dContinuationKernel,
// :71-183) or K = N (ls1dContinuationKernel, :57-180): the probes y
// (absolute coordinates) ride with the state from point to point. A point
// that has no attached probes first runs the transient on the state alone
// and places the probes (K = 1: y eps + x; K = N: the classical
// Gram-Schmidt of the random basis, times eps, plus x); a point with probes
// attached runs (int)(transient / NT) extra blocks whose growth is not
// summed. Block: nt = (int)(NT / h) steps plus the trailing step for the
// state and each probe, an escape ends the point (NaN, flag -1, chain
// restarts from the initial state with a new probe basis from subsequence
// point + 1). K = 1: d = |(x - y) / eps| (floored at 1e-14), sum log d,
// y = x - (x - y + 1e-14) / d. K = N: y -= x, Gram-Schmidt, sum
// log(|b_k| / eps), y = x + eps b_k. Values sum / t_max.
template <int K>
__device__ void uc_cont_probes(UcRng& r, double* y) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double zp = 0.0;
#pragma unroll
        for (int i = 0; i < UC_N; ++i) {
            y[k * UC_N + i] = (double)r.uniform() - 0.5;
            zp = zp + y[k * UC_N + i] * y[k * UC_N + i];
        }
        zp = sqrt(zp);
#pragma unroll
        for (int i = 0; i < UC_N; ++i) y[k * UC_N + i] = (zp == 0.0) ? 0.0 : y[k * UC_N + i] / zp;
    }
}

// M independent trajectories stepped together for `iters` steps plus the
// trailing step, each with the step loop's magnitude checks (after
// iterations 0, 100, 200, ...): 0 when any of them escapes, else 1. The
// same operations per trajectory as M separate uc_loop calls; advancing
// them together gives a lone chain M independent dependency chains to
// interleave. (Only the escape matters to the chains, not the fixed-point
// test of the trailing step.)
template <int M>
__device__ int uc_loop_multi(StateF64* S, const double* a, double h, double& t, long long iters, double max_value) {
    int until_check = 0;
    for (long long i = 0; i < iters; ++i) {
#pragma unroll
        for (int m = 0; m < M; ++m) uc_rk4(S[m], a, h, t);
        if (UC_USES_TIME) t += h;
        if (until_check == 0) {
            bool esc = false;
#pragma unroll
            for (int m = 0; m < M; ++m) esc = esc || uc_diverged(S[m], max_value);
            if (esc) return 0;
            until_check = UC_CHECK_INTERVAL;
        }
        --until_check;
    }
#pragma unroll
    for (int m = 0; m < M; ++m) uc_rk4(S[m], a, h, t);
    if (UC_USES_TIME) t += h;
    return 1;
}

template <int K>
__device__ void uc_cont_lyap_chain(const UcJob& J) {
    const long long chain = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= J.ny) return;
    double ad[UC_M], x0[UC_N], hb;
    uc_chain_setup(J, chain, ad, x0, hb);
    StateF64 s;
    s.load(x0);
    double y[K * UC_N], z[K * UC_N], den[K], sum[K];
    UcRng r;
    r.init(J.probe_seed, 0ULL, 0ULL);
    uc_cont_probes<K>(r, y);
    bool attached = false;
    double t = 0.0;
    NoObserver none;
    const long long n_blocks = (long long)(J.t_max / J.nt);
    for (long long jj = 0; jj < J.nx; ++jj) {
        const long long cell = chain * J.nx + jj;
        const double h = uc_chain_point(J, jj, ad, hb);
        const long long nts = (h > 0.0 && n_blocks > 0) ? (long long)(J.nt / h) : 0;
        if (h <= 0.0 || n_blocks <= 0 || nts <= 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) J.out0[(long long)k * J.stride + cell] = __builtin_nan("");
            J.flag[cell] = -1;
            continue;
        }
        bool alive = true;
        long long settle = (long long)(J.transient_time / J.nt);
        if (!attached) {
            settle = 0;
            if (uc_loop(s, ad, h, t, (long long)(J.transient_time / h), 1, 0, J.max_value, none) == 0) {
                alive = false;
            } else {
                if (K == 1) {
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) y[i] = y[i] * J.eps + s.x[i];
                } else {
                    uc_cgs<K>(y, z, den);
#pragma unroll
                    for (int k = 0; k < K; ++k)
#pragma unroll
                        for (int i = 0; i < UC_N; ++i) y[k * UC_N + i] = z[k * UC_N + i] * J.eps + s.x[i];
                }
                attached = true;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) sum[k] = 0.0;
        for (long long b = 0; alive && b < settle + n_blocks; ++b) {
            // the state and its K probes over the same block, together
            StateF64 S[K + 1];
            S[0] = s;
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int i = 0; i < UC_N; ++i) S[k + 1].x[i] = y[k * UC_N + i];
            if (uc_loop_multi<K + 1>(S, ad, h, t, nts, J.max_value) == 0) { alive = false; break; }
            s = S[0];
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int i = 0; i < UC_N; ++i) y[k * UC_N + i] = S[k + 1].x[i];
            if (K == 1) {
                double d = 0.0;
#pragma unroll
                for (int l = 0; l < UC_N; ++l) {
                    const double q = (s.x[l] - y[l]) / J.eps;
                    d = d + q * q;
                }
                d = sqrt(d);
                if (d <= 1e-14) d = 1e-14;
                if (b >= settle) sum[0] = sum[0] + log(d);
                const double inv = 1.0 / d;
#pragma unroll
                for (int l = 0; l < UC_N; ++l) y[l] = s.x[l] - ((s.x[l] - y[l] + 1e-14) * inv);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int l = 0; l < UC_N; ++l) y[k * UC_N + l] = y[k * UC_N + l] - s.x[l];
                uc_cgs<K>(y, z, den);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (b >= settle) sum[k] = sum[k] + log(den[k] / J.eps);
#pragma unroll
                    for (int i = 0; i < UC_N; ++i) y[k * UC_N + i] = s.x[i] + z[k * UC_N + i] * J.eps;
                }
            }
        }
        if (alive) {
#pragma unroll
            for (int k = 0; k < K; ++k) J.out0[(long long)k * J.stride + cell] = sum[k] / J.t_max;
            J.flag[cell] = 1;
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) J.out0[(long long)k * J.stride + cell] = __builtin_nan("");
            J.flag[cell] = -1;
            attached = false;
            s.load(x0);
            r.init(J.probe_seed, (u64)jj + 1ULL, 0ULL);
            uc_cont_probes<K>(r, y);
        }
    }
}

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_cont_lle(UcJob J) { uc_cont_lyap_chain<1>(J); }
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_cont_ls(UcJob J) { uc_cont_lyap_chain<UC_N>(J); }

// ========================================================== fast synchro ==
// Fast synchronisation (the reference's FastSynchro,
// U-CUDA/cudaLibrary.cu:424-452 step, :547-720 grid loop, :3630-3800
// attractor loop), FP64. One coupled step: N = K (S1 - X) from the state
// before the step, X <- step(X, +h forward / -h backward), X += h N (the
// coupling always with +h). The loops follow the reference line by line,
// because the error estimators read the state at fixed points of them.
__device__ __forceinline__ void uc_fs_step(StateF64& X, const double* S1, const double* K, const double* a, double h,
                                           bool fwd) {
    double n[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) n[i] = K[i] * (S1[i] - X.x[i]);
    uc_rk4(X, a, fwd ? h : -h, 0.0);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) X.x[i] = X.x[i] + h * n[i];
}

// Uncoupled step as the reference writes it: K = 0 with S1 = X itself
// (self) or S1 = 0 (the grid's master fill); kept literal so that signed
// zeros and non-finite states propagate the same way.
__device__ __forceinline__ void uc_fs_free(StateF64& X, const double* a, double h, bool self) {
    double s1[UC_N], z[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { s1[i] = self ? X.x[i] : 0.0; z[i] = 0.0; }
    uc_fs_step(X, s1, z, a, h, true);
}

template <class Master>
__device__ double uc_fs_sync(const UcFsJob& J, const double* a, StateF64& Xm, StateF64& Xs, Master master) {
    const int n = J.n_iters;
    const double h = J.h;
    double kf[UC_N], kb[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) { kf[i] = J.k_forward[i]; kb[i] = J.k_backward[i]; }
    double rms = 0.0;
    double prev[UC_N];
    for (int m = 0; m < J.iter_of_synchr; ++m) {
        for (int i = 0; i < n - 1; ++i) {
            if (J.type_of_synch == 0) master(i, Xm);
            if (J.error_estim == 0) {
                if (m == J.iter_of_synchr - 1)
#pragma unroll
                    for (int j = 0; j < UC_N; ++j) rms = rms + (Xm.x[j] - Xs.x[j]) * (Xm.x[j] - Xs.x[j]);
            } else if (J.error_estim == 1) {
                if (i == n - 2) {
                    rms = 0.0;
#pragma unroll
                    for (int j = 0; j < UC_N; ++j) rms = rms + (Xm.x[j] - Xs.x[j]) * (Xm.x[j] - Xs.x[j]);
                }
            } else if (J.error_estim == 2) {
                if (m == J.iter_of_synchr - 1 && i == n - 2) {
                    rms = 0.0;
#pragma unroll
                    for (int j = 0; j < UC_N; ++j) rms = rms + (Xm.x[j] - Xs.x[j]) * (Xm.x[j] - Xs.x[j]);
                }
            } else {
                rms = 0.0;
#pragma unroll
                for (int j = 0; j < UC_N; ++j) rms = rms + (Xm.x[j] - Xs.x[j]) * (Xm.x[j] - Xs.x[j]);
                rms = sqrt(rms);
                if (rms <= J.fs_error_trs) return (double)(i + 1) * h;
            }
            if (J.type_of_synch == 1)
#pragma unroll
                for (int j = 0; j < UC_N; ++j) prev[j] = Xs.x[j];
            uc_fs_step(Xs, Xm.x, kf, a, h, true);
            if (J.type_of_synch == 1) uc_fs_step(Xm, prev, kf, a, h, true);
        }
        if (J.error_estim == 1) {
            rms = sqrt(rms);
            if (rms <= J.fs_error_trs) return (double)(m + 1);
            rms = 0.0;
        }
        for (int i = n - 1; i > 0; --i) {
            if (J.type_of_synch == 0) master(i, Xm);
            if (J.type_of_synch == 1)
#pragma unroll
                for (int j = 0; j < UC_N; ++j) prev[j] = Xs.x[j];
            uc_fs_step(Xs, Xm.x, kb, a, h, false);
            if (J.type_of_synch == 1) uc_fs_step(Xm, prev, kb, a, h, false);
        }
    }
    if (J.error_estim == 0) rms = sqrt(rms / (double)(n - 1));
    if (J.error_estim == 2) rms = sqrt(rms);
    rms = rms <= J.fs_error_trs ? log10(J.fs_error_trs) : log10(rms);
    if (isinf(rms) || isnan(rms) || rms >= 3.0) rms = 3.0;
    if (J.error_estim == 1) return (double)J.iter_of_synchr;
    if (J.error_estim == 3) return (double)n * h;
    return rms;
}

__device__ __forceinline__ void uc_fs_params(const UcFsJob& J, double* a) {
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = J.params[i];
}

// Master trajectory of the on-attractor mode (fillFSMasterTrajectory,
// U-CUDA/kernels/fastsync_attr.template.cu:75-98): one lane, `skip`
// uncoupled steps, then the state before each of `len` steps.
extern "C" __global__ void uc_fs_fill(UcFsJob J, int skip, int len, double* traj) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double a[UC_M];
    uc_fs_params(J, a);
    StateF64 X;
    X.load(J.ic_master);
    for (int i = 0; i < skip; ++i) uc_fs_free(X, a, J.h, true);
    for (int i = 0; i < len; ++i) {
#pragma unroll
        for (int j = 0; j < UC_N; ++j) traj[(long long)i * UC_N + j] = X.x[j];
        uc_fs_free(X, a, J.h, true);
    }
}

// Transient of the grid's fixed initial state (fillFSTransientIC,
// U-CUDA/cudaLibrary.cu:454-474): one lane, in place.
extern "C" __global__ void uc_fs_transient(UcFsJob J, int skip, double* x0) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double a[UC_M];
    uc_fs_params(J, a);
    StateF64 X;
    X.load(x0);
    for (int i = 0; i < skip; ++i) uc_fs_free(X, a, J.h, true);
#pragma unroll
    for (int j = 0; j < UC_N; ++j) x0[j] = X.x[j];
}

// On attractor (calculateDiscreteModelforFastSynchroCUDA,
// U-CUDA/cudaLibrary.cu:3586-3626): lane idx synchronises the slave, from
// ic_slave, over the master window starting at sample idx * pre.
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_fs_attr(UcFsJob J) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= J.n_cells) return;
    double a[UC_M];
    uc_fs_params(J, a);
    const double* w = J.master + idx * (long long)J.pre_scaller * UC_N;
    StateF64 Xm, Xs;
    Xs.load(J.ic_slave);
    if (J.type_of_synch == 1) Xm.load(w);
    if (J.traj_out)
#pragma unroll
        for (int j = 0; j < UC_N; ++j) J.traj_out[idx * UC_N + j] = w[j];
    J.out[idx] = uc_fs_sync(J, a, Xm, Xs, [&](int i, StateF64& m) {
#pragma unroll
        for (int j = 0; j < UC_N; ++j) m.x[j] = w[(long long)i * UC_N + j];
    });
}

// On grid (calculateDiscreteModelICCforFastSynchro,
// U-CUDA/cudaLibrary.cu:476-545): cell = iy * n_pts + ix sets state
// variables axis_var[0] / axis_var[1] of the master (swap 0) or the slave
// (swap 1); unidirectional runs first store the cell's uncoupled master
// window (S1 = 0 steps) in data.
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_fs_grid(UcFsJob J) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= J.n_cells) return;
    const long long cell = J.cell_offset + lane;
    double a[UC_M];
    uc_fs_params(J, a);
    double xm[UC_N], xs[UC_N];
#pragma unroll
    for (int j = 0; j < UC_N; ++j) { xm[j] = J.ic_master[j]; xs[j] = J.ic_slave[j]; }
    const double v0 = J.axis_vals[0][cell % J.n_pts], v1 = J.axis_vals[1][(cell / J.n_pts) % J.n_pts];
#pragma unroll
    for (int j = 0; j < UC_N; ++j) {
        double* x = J.swap ? xs : xm;
        if (j == J.axis_var[0]) x[j] = v0;
    }
#pragma unroll
    for (int j = 0; j < UC_N; ++j) {
        double* x = J.swap ? xs : xm;
        if (j == J.axis_var[1]) x[j] = v1;
    }
    StateF64 Xm, Xs;
    Xm.load(xm);
    Xs.load(xs);
    double* d = J.data ? J.data + lane : nullptr;  // unidirectional runs only
    if (J.type_of_synch == 0) {
#pragma unroll
        for (int j = 0; j < UC_N; ++j) d[(long long)j * J.stride] = Xm.x[j];
        StateF64 M = Xm;
        for (int i = 1; i < J.n_iters; ++i) {
            uc_fs_free(M, a, J.h, false);
#pragma unroll
            for (int j = 0; j < UC_N; ++j) d[((long long)i * UC_N + j) * J.stride] = M.x[j];
        }
    }
    // Fast modules: forward passes ask for the master samples 0, 1, ...,
    // n - 2 in order, and they are recomputed here from the window's first
    // state with the same uncoupled step that filled the stored row
    // (deterministic, so the same values). That halves the HBM reads of
    // this HBM-bound kernel (49.7 -> 40.8 ms at 1024^2). Strict modules,
    // without FMA contraction, are compute-heavier and read the row (the
    // recompute made them 10% slower). Backward passes walk the row in
    // reverse and read it.
    const StateF64 M0 = Xm;
    StateF64 R = M0;
    int last = -2;
    J.out[lane] = uc_fs_sync(J, a, Xm, Xs, [&](int i, StateF64& m) {
        if (UC_STRICT) {
#pragma unroll
            for (int j = 0; j < UC_N; ++j) m.x[j] = d[((long long)i * UC_N + j) * J.stride];
        } else if (i == 0) {
            R = M0;
            last = 0;
            m = R;
        } else if (i == last + 1) {
            uc_fs_free(R, a, J.h, false);
            last = i;
            m = R;
        } else {
            last = -2;
#pragma unroll
            for (int j = 0; j < UC_N; ++j) m.x[j] = d[((long long)i * UC_N + j) * J.stride];
        }
    });
}

// Zeroes the work counters of a batch (a kernel, not a memset, so the reset
// can live inside a captured graph).
extern "C" __global__ void uc_batch_reset(u64* counters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0ULL;
}

UC_DIRECT(uc_bif_f64, uc_bif_cell<StateF64>)
UC_DIRECT(uc_bif_f32c, uc_bif_cell<StateF32C>)
UC_DIRECT(uc_lle_f64, (uc_lyap_cell<1, StateF64>))
UC_DIRECT(uc_lle_f32c, (uc_lyap_cell<1, StateF32C>))
UC_DIRECT(uc_ls_f64, (uc_lyap_cell<UC_N, StateF64>))
UC_DIRECT(uc_ls_f32c, (uc_lyap_cell<UC_N, StateF32C>))
UC_PERSISTENT(uc_bif_f64_q, uc_bif_cell<StateF64>)
UC_PERSISTENT(uc_bif_f32c_q, uc_bif_cell<StateF32C>)
UC_PERSISTENT(uc_lle_f64_q, (uc_lyap_cell<1, StateF64>))
UC_PERSISTENT(uc_lle_f32c_q, (uc_lyap_cell<1, StateF32C>))
UC_PERSISTENT(uc_ls_f64_q, (uc_lyap_cell<UC_N, StateF64>))
UC_PERSISTENT(uc_ls_f32c_q, (uc_lyap_cell<UC_N, StateF32C>))
// Sliced persistent kernels: each launch runs its entries' units
// [unit_begin, unit_end) and keeps unfinished cells in the slab.
UC_PERSISTENT(uc_bif_f64_s, (uc_bif_cell<StateF64, true>))
UC_PERSISTENT(uc_bif_f32c_s, (uc_bif_cell<StateF32C, true>))
UC_PERSISTENT(uc_lle_f64_s, (uc_lyap_cell<1, StateF64, true>))
UC_PERSISTENT(uc_lle_f32c_s, (uc_lyap_cell<1, StateF32C, true>))
UC_PERSISTENT(uc_ls_f64_s, (uc_lyap_cell<UC_N, StateF64, true>))
UC_PERSISTENT(uc_ls_f32c_s, (uc_lyap_cell<UC_N, StateF32C, true>))
#if !UC_STRICT
// The lean LLE loop (uc_lle_lanes) with one cell per lane.
template <class S>
__device__ __forceinline__ void uc_lle_one(const UcJob& j, long long lane) { uc_lle_lanes<S, 1>(j, lane, 0); }
UC_DIRECT(uc_lle_f64_x1, uc_lle_one<StateF64>)
UC_DIRECT(uc_lle_f32c_x1, uc_lle_one<StateF32C>)
UC_PERSISTENT(uc_lle_f64_x1_q, uc_lle_one<StateF64>)
UC_PERSISTENT(uc_lle_f32c_x1_q, uc_lle_one<StateF32C>)
// Two cells per lane (uc_lle_lanes): lane l runs cells l and l + ceil(n / 2).
#define UC_DIRECT_PAIR(NAME, CELL)                                              \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(UcJob j) {      \
        const long long half = (j.n_cells + 1) / 2;                             \
        const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x; \
        if (lane < half) CELL(j, lane, half);                                   \
    }
UC_DIRECT_PAIR(uc_lle_f64_x2, (uc_lle_lanes<StateF64, 2>))
UC_DIRECT_PAIR(uc_lle_f32c_x2, (uc_lle_lanes<StateF32C, 2>))
// Persistent form (whole runs only): a chunk is 2 * blockDim cells, lane l
// of the chunk runs its cells l and l + blockDim.
#define UC_PERSISTENT_PAIR(NAME, CELL)                                                              \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(const UcJob* __restrict__ jobs,     \
                                                                const long long* __restrict__ start, \
                                                                int n_jobs, u64* counter) {          \
        __shared__ long long s_chunk;                                                               \
        const long long total = start[n_jobs];                                                      \
        for (;;) {                                                                                  \
            if (threadIdx.x == 0) s_chunk = (long long)atomicAdd((unsigned long long*)counter, 1ULL); \
            __syncthreads();                                                                        \
            const long long c = s_chunk;                                                            \
            __syncthreads();                                                                        \
            if (c >= total) break;                                                                  \
            int lo = 0, hi = n_jobs - 1;                                                            \
            while (lo < hi) {                                                                       \
                const int mid = (lo + hi + 1) >> 1;                                                 \
                if (start[mid] <= c) lo = mid; else hi = mid - 1;                                   \
            }                                                                                       \
            lo = __builtin_amdgcn_readfirstlane(lo);                                                \
            uc_set_priority(jobs[lo].priority);                                                     \
            const long long lane = (c - start[lo]) * 2LL * blockDim.x + threadIdx.x;               \
            if (lane < jobs[lo].n_cells) CELL(jobs[lo], lane, (long long)blockDim.x);              \
        }                                                                                           \
    }
UC_PERSISTENT_PAIR(uc_lle_f64_x2_q, (uc_lle_lanes<StateF64, 2>))
UC_PERSISTENT_PAIR(uc_lle_f32c_x2_q, (uc_lle_lanes<StateF32C, 2>))
#endif

// The spectrum with N lanes per cell (uc_ls_spread): a wave holds
// floor(64 / N) cells, cell c of the wave on lanes c * N .. c * N + N - 1;
// the lanes past the last group, and the groups past the grid, run stand-in
// copies of the grid's last cell.
__device__ __forceinline__ void uc_spread_lane(long long first, long long n_cells, long long& cell, bool& has,
                                               int& k, int& base) {
    constexpr int K = UC_N, PER = 64 / UC_N;
    const int w = (int)(threadIdx.x & 63);
    const int g = w / K < PER ? w / K : PER - 1;
    k = w / K < PER ? w % K : 0;
    base = g * K;
    cell = first + g;
    has = w / K < PER && cell < n_cells;
    if (cell >= n_cells) cell = n_cells - 1;
}
#define UC_DIRECT_SPREAD(NAME, CELL)                                                             \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(UcJob j) {                       \
        const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;          \
        long long cell;                                                                          \
        bool has;                                                                                \
        int k, base;                                                                             \
        if (wave * (64 / UC_N) >= j.n_cells) return; /* a whole wave past the grid */            \
        uc_spread_lane(wave * (64 / UC_N), j.n_cells, cell, has, k, base);                       \
        CELL(j, cell, has, k, base);                                                             \
    }
UC_DIRECT_SPREAD(uc_ls_f64_xn, uc_ls_spread<StateF64>)
UC_DIRECT_SPREAD(uc_ls_f32c_xn, uc_ls_spread<StateF32C>)
// Persistent form: a chunk is the workgroup's waves x floor(64 / N) cells.
#define UC_PERSISTENT_SPREAD(NAME, CELL)                                                            \
    extern "C" __global__ void __launch_bounds__(UC_BLOCK) NAME(const UcJob* __restrict__ jobs,     \
                                                                const long long* __restrict__ start, \
                                                                int n_jobs, u64* counter) {          \
        __shared__ long long s_chunk;                                                               \
        const long long total = start[n_jobs];                                                      \
        for (;;) {                                                                                  \
            if (threadIdx.x == 0) s_chunk = (long long)atomicAdd((unsigned long long*)counter, 1ULL); \
            __syncthreads();                                                                        \
            const long long c = s_chunk;                                                            \
            __syncthreads();                                                                        \
            if (c >= total) break;                                                                  \
            int lo = 0, hi = n_jobs - 1;                                                            \
            while (lo < hi) {                                                                       \
                const int mid = (lo + hi + 1) >> 1;                                                 \
                if (start[mid] <= c) lo = mid; else hi = mid - 1;                                   \
            }                                                                                       \
            lo = __builtin_amdgcn_readfirstlane(lo);                                                \
            uc_set_priority(jobs[lo].priority);                                                     \
            const long long per_chunk = (long long)(blockDim.x >> 6) * (64 / UC_N);                \
            long long cell;                                                                         \
            bool has;                                                                               \
            int k, base;                                                                            \
            const long long first = (c - start[lo]) * per_chunk + (long long)(threadIdx.x >> 6) * (64 / UC_N); \
            if (first < jobs[lo].n_cells) {                                                         \
                uc_spread_lane(first, jobs[lo].n_cells, cell, has, k, base);                        \
                CELL(jobs[lo], cell, has, k, base);                                                 \
            }                                                                                       \
        }                                                                                           \
    }
UC_PERSISTENT_SPREAD(uc_ls_f64_xn_q, uc_ls_spread<StateF64>)
UC_PERSISTENT_SPREAD(uc_ls_f32c_xn_q, uc_ls_spread<StateF32C>)

// ================================================================ basins ==
// Adaptive Dormand-Prince 5(4) (Dormand & Prince 1980 coefficients; Hairer's
// RMS error norm and step controller, FSAL) with every step clipped to land
// on the Poincare sections t = k*T. The system time seen by the right-hand
// side is local to the current section period (t in [0, T]), exact for
// T-periodic forcing. The last UC_REC section points of the first
// UC_LBL state variables are kept in a shift register; afterwards the
// smallest period p <= max_period whose points close within period_tol is
// found and the cycle point with the smallest (x0, x1) becomes the cell's
// canonical point (no cycle: p = 0 and the mean point; unbound: p = -1;
// step-size collapse or attempt cap: p = -2).
#define UC_REC 16
#define UC_MAXP 8
#define UC_LBL (UC_N < 2 ? UC_N : 2)

#define DPA21 (1.0 / 5.0)
#define DPA31 (3.0 / 40.0)
#define DPA32 (9.0 / 40.0)
#define DPA41 (44.0 / 45.0)
#define DPA42 (-56.0 / 15.0)
#define DPA43 (32.0 / 9.0)
#define DPA51 (19372.0 / 6561.0)
#define DPA52 (-25360.0 / 2187.0)
#define DPA53 (64448.0 / 6561.0)
#define DPA54 (-212.0 / 729.0)
#define DPA61 (9017.0 / 3168.0)
#define DPA62 (-355.0 / 33.0)
#define DPA63 (46732.0 / 5247.0)
#define DPA64 (49.0 / 176.0)
#define DPA65 (-5103.0 / 18656.0)
#define DPB1 (35.0 / 384.0)
#define DPB3 (500.0 / 1113.0)
#define DPB4 (125.0 / 192.0)
#define DPB5 (-2187.0 / 6784.0)
#define DPB6 (11.0 / 84.0)
#define DPE1 (71.0 / 57600.0)
#define DPE3 (-71.0 / 16695.0)
#define DPE4 (71.0 / 1920.0)
#define DPE5 (-17253.0 / 339200.0)
#define DPE6 (22.0 / 525.0)
#define DPE7 (-1.0 / 40.0)

// Hardware reciprocals for the fast modes' error norm (v_rcp_f32 / v_rcp_f64).
__device__ __forceinline__ float uc_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double uc_rcp(double v) { return __builtin_amdgcn_rcp(v); }

// One attempt from (t, x) with step h, k1 = f(t, x) on entry. Produces the
// 5th-order increment dx (x5 = x + dx), k7 = f(t + h, x5) and the error norm.
// The stage sums follow the tableau row order left to right, like the oracle.
// Strict modules return the RMS norm as the oracle computes it. Fast ones
// return its square, with the scaled errors formed in the stage precision by
// a hardware reciprocal: the norm only decides acceptance (err^2 <= 1 exactly
// when err <= 1) and sizes the next step, so it needs no correctly rounded
// division or square root.
//
// `rhs(xs, stage, time, out)` evaluates the right-hand side for stages 2..7
// (uc_dp5: the generated uc_rhs; the forced variant below supplies the
// harmonic forcing terms).
template <typename T, class R>
__device__ __forceinline__ double uc_dp5_core(const T* x, T t, T h, const T* k1, T* dx, T* k7, double rtol,
                                              double atol, R&& rhs) {
    T k2[UC_N], k3[UC_N], k4[UC_N], k5[UC_N], k6[UC_N], xs[UC_N];
#pragma unroll
    for (int i = 0; i < UC_N; ++i) xs[i] = x[i] + h * ((T)DPA21 * k1[i]);
    rhs(xs, 2, t + (T)(1.0 / 5.0) * h, k2);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) xs[i] = x[i] + h * ((T)DPA31 * k1[i] + (T)DPA32 * k2[i]);
    rhs(xs, 3, t + (T)(3.0 / 10.0) * h, k3);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) xs[i] = x[i] + h * ((T)DPA41 * k1[i] + (T)DPA42 * k2[i] + (T)DPA43 * k3[i]);
    rhs(xs, 4, t + (T)(4.0 / 5.0) * h, k4);
#pragma unroll
    for (int i = 0; i < UC_N; ++i)
        xs[i] = x[i] + h * ((T)DPA51 * k1[i] + (T)DPA52 * k2[i] + (T)DPA53 * k3[i] + (T)DPA54 * k4[i]);
    rhs(xs, 5, t + (T)(8.0 / 9.0) * h, k5);
#pragma unroll
    for (int i = 0; i < UC_N; ++i)
        xs[i] = x[i] + h * ((T)DPA61 * k1[i] + (T)DPA62 * k2[i] + (T)DPA63 * k3[i] + (T)DPA64 * k4[i] +
                            (T)DPA65 * k5[i]);
    rhs(xs, 6, t + h, k6);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        dx[i] = h * ((T)DPB1 * k1[i] + (T)0.0 * k2[i] + (T)DPB3 * k3[i] + (T)DPB4 * k4[i] + (T)DPB5 * k5[i] +
                     (T)DPB6 * k6[i]);
        xs[i] = x[i] + dx[i];
    }
    rhs(xs, 7, t + h, k7);
    double err = 0.0;
#pragma unroll
    for (int i = 0; i < UC_N; ++i) {
        const T e = h * ((T)DPE1 * k1[i] + (T)0.0 * k2[i] + (T)DPE3 * k3[i] + (T)DPE4 * k4[i] + (T)DPE5 * k5[i] +
                         (T)DPE6 * k6[i] + (T)DPE7 * k7[i]);
#if UC_STRICT
        const double sc = atol + rtol * fmax(fabs((double)x[i]), fabs((double)xs[i]));
        const double q = (double)e / sc;
#else
        const T sc = (T)atol + (T)rtol * fmax(fabs(x[i]), fabs(xs[i]));
        const double q = (double)(e * uc_rcp(sc));
#endif
        err = err + q * q;
    }
#if UC_STRICT
    return sqrt(err / (double)UC_N);
#else
    return err * (1.0 / (double)UC_N);
#endif
}

template <typename T>
__device__ __forceinline__ double uc_dp5(const T* x, const T* a, T t, T h, const T* k1, T* dx, T* k7, double rtol,
                                         double atol) {
    return uc_dp5_core<T>(x, t, h, k1, dx, k7, rtol, atol,
                          [&](const T* xs, int, T ts, T* out) { uc_rhs<T>(xs, a, ts, out); });
}

// Harmonic forcing (UC_HARM) in the adaptive integration: the basins loop
// keeps cos / sin of every phase at the current attempt's start time
// (exact at each section start, local time 0: cos / sin(phi)) and moves
// them to the stage times t + c h by the rotation (C, S) -> (C cd - S sd,
// S cd + C sd), cd / sd = cos / sin(w c h): a short Taylor series instead
// of the library's range reduction and polynomial, and one rotation for
// the two stages at t + h. An accepted step takes the t + h values on, so
// within a section the values carry the rounding of up to a few hundred
// rotations (~1e-14 relative in double), below the rounding the library
// call sees in its argument w * t itself.

// cos / sin of each harmonic phase at the current attempt's start time; the
// frequencies and phases are taken from the parameters where they are
// needed (no registers held for them).
template <typename T> struct UcForcing {
#if UC_HARM
    T c[UC_HARM], s[UC_HARM];
    __device__ __forceinline__ void reset(const T* a) {  // local time 0
        T w[UC_HARM], phi[UC_HARM];
        uc_harm_coef<T>(a, w, phi);
#pragma unroll
        for (int k = 0; k < UC_HARM; ++k) uc_sincos_small(phi[k], c[k], s[k]);
    }
#else
    __device__ __forceinline__ void reset(const T*) {}
#endif
};

#if UC_HARM
// uc_dp5 with the forcing rotated to the stage times; c1 / s1 receive the
// values at t + h (taken on when the attempt is accepted).
template <typename T>
__device__ __forceinline__ double uc_dp5_forced(const T* x, const T* a, T t, T h, const T* k1, T* dx, T* k7,
                                                double rtol, double atol, const UcForcing<T>& f, T* c1, T* s1) {
    T hc[UC_HARM], hs[UC_HARM];
    return uc_dp5_core<T>(x, t, h, k1, dx, k7, rtol, atol, [&](const T* xs, const int st, T ts, T* out) {
        if (st != 7) {  // stage 7 shares stage 6's time
            const T cst = st == 2 ? (T)(1.0 / 5.0) : st == 3 ? (T)(3.0 / 10.0) : st == 4 ? (T)(4.0 / 5.0)
                        : st == 5 ? (T)(8.0 / 9.0) : (T)1.0;
            T w[UC_HARM], phi[UC_HARM];
            uc_harm_coef<T>(a, w, phi);
#pragma unroll
            for (int k = 0; k < UC_HARM; ++k) {
                T cd, sd;
                uc_sincos_small(w[k] * (cst * h), cd, sd);
                hc[k] = f.c[k] * cd - f.s[k] * sd;
                hs[k] = f.s[k] * cd + f.c[k] * sd;
                if (st == 6) { c1[k] = hc[k]; s1[k] = hs[k]; }
            }
        }
        uc_rhs_h<T>(xs, a, ts, hc, hs, out);
    });
}
#endif

struct BasF64 {
    double x[UC_N];
    double t;
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) x[i] = x0[i];
    }
    __device__ __forceinline__ double get(int i) const { return x[i]; }
    __device__ __forceinline__ void reset_time() { t = 0.0; }
    __device__ __forceinline__ double time() const { return t; }
    __device__ __forceinline__ double remaining(double T) const { return T - t; }
    __device__ __forceinline__ void advance_time(double hs) { t = t + hs; }
};

struct BasF32C {
    float hi[UC_N], c[UC_N];
    float thi, tc;  // compensated local time
    __device__ __forceinline__ void load(const double* x0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) { hi[i] = (float)x0[i]; c[i] = (float)((double)hi[i] - x0[i]); }
    }
    __device__ __forceinline__ double get(int i) const { return (double)hi[i] - (double)c[i]; }
    __device__ __forceinline__ void reset_time() { thi = 0.0f; tc = 0.0f; }
    __device__ __forceinline__ double time() const { return (double)thi - (double)tc; }
    __device__ __forceinline__ double remaining(double T) const { return T - time(); }
    __device__ __forceinline__ void advance_time(double hs) { uc_kahan_add(thi, tc, (float)hs); }
};

// One accepted-or-rejected attempt for each state policy. Returns uc_dp5's
// norm (its square in fast modules); on acceptance (<= 1) the state moves
// and k1 <- k7 (FSAL).
// With harmonic forcing (fast modules) the forcing values move to t + h on
// acceptance as well.
template <typename T>
__device__ __forceinline__ double uc_bas_dp5(const T* x, const T* a, T t, T h, const T* k1, T* dx, T* k7,
                                             double rtol, double atol, UcForcing<T>& f, T* c1, T* s1) {
#if UC_HARM
    return uc_dp5_forced<T>(x, a, t, h, k1, dx, k7, rtol, atol, f, c1, s1);
#else
    (void)f; (void)c1; (void)s1;
    return uc_dp5<T>(x, a, t, h, k1, dx, k7, rtol, atol);
#endif
}
template <typename T> __device__ __forceinline__ void uc_bas_forcing_on(UcForcing<T>& f, const T* c1, const T* s1) {
#if UC_HARM
#pragma unroll
    for (int k = 0; k < UC_HARM; ++k) { f.c[k] = c1[k]; f.s[k] = s1[k]; }
#else
    (void)f; (void)c1; (void)s1;
#endif
}
__device__ __forceinline__ double uc_bas_attempt(BasF64& s, const double* a, double hs, double* k1, double rtol,
                                                 double atol, UcForcing<double>& f) {
    double dx[UC_N], k7[UC_N], c1[UC_HARM > 0 ? UC_HARM : 1], s1[UC_HARM > 0 ? UC_HARM : 1];
    const double err = uc_bas_dp5<double>(s.x, a, s.t, hs, k1, dx, k7, rtol, atol, f, c1, s1);
    if (err <= 1.0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) { s.x[i] = s.x[i] + dx[i]; k1[i] = k7[i]; }
        uc_bas_forcing_on(f, c1, s1);
    }
    return err;
}
__device__ __forceinline__ double uc_bas_attempt(BasF32C& s, const float* a, double hs, float* k1, double rtol,
                                                 double atol, UcForcing<float>& f) {
    float dx[UC_N], k7[UC_N], c1[UC_HARM > 0 ? UC_HARM : 1], s1[UC_HARM > 0 ? UC_HARM : 1];
    const double err = uc_bas_dp5<float>(s.hi, a, s.thi, (float)hs, k1, dx, k7, rtol, atol, f, c1, s1);
    if (err <= 1.0) {
#pragma unroll
        for (int i = 0; i < UC_N; ++i) { uc_kahan_add(s.hi[i], s.c[i], dx[i]); k1[i] = k7[i]; }
        uc_bas_forcing_on(f, c1, s1);
    }
    return err;
}
__device__ __forceinline__ void uc_bas_state(UcSlab& io, BasF64& s) {
#pragma unroll
    for (int i = 0; i < UC_N; ++i) io.f64(s.x[i]);
}
__device__ __forceinline__ void uc_bas_state(UcSlab& io, BasF32C& s) {
#pragma unroll
    for (int i = 0; i < UC_N; ++i) io.f32(s.hi[i]);
#pragma unroll
    for (int i = 0; i < UC_N; ++i) io.f32(s.c[i]);
}
// k1 at a section start (local time 0, forcing just reset)
template <typename T> __device__ __forceinline__ void uc_bas_rhs(const T* x, const T* a, T t, const UcForcing<T>& f,
                                                                T* k1) {
#if UC_HARM
    uc_rhs_h<T>(x, a, t, f.c, f.s, k1);
#else
    (void)f;
    uc_rhs<T>(x, a, t, k1);
#endif
}
__device__ __forceinline__ void uc_bas_k1(BasF64& s, const double* a, double* k1, const UcForcing<double>& f) {
    uc_bas_rhs<double>(s.x, a, s.t, f, k1);
}
__device__ __forceinline__ void uc_bas_k1(BasF32C& s, const float* a, float* k1, const UcForcing<float>& f) {
    uc_bas_rhs<float>(s.hi, a, s.thi, f, k1);
}

template <class S> struct BasPrec;
template <> struct BasPrec<BasF64> {
    typedef double T;
    static __device__ __forceinline__ double tol(double v) { return v; }
};
// float stages cannot resolve errors below ~1e-6 of the state
template <> struct BasPrec<BasF32C> {
    typedef float T;
    static __device__ __forceinline__ double tol(double v) { return fmax(v, 9.5367431640625e-07); }
};

// Sliced (SL, j.slab set): one unit per section; slab words done, h,
// accepted, rejected, state, then the section shift register from word
// 4 + 2N.
static_assert(UC_REC * UC_LBL <= 32, "basins slab layout");
template <class S, bool SL = false>
__device__ void uc_basins_cell(const UcJob& j, long long lane) {
    typedef typename BasPrec<S>::T T;
    const long long cell = j.cell_offset + lane;
    constexpr bool sliced = SL;
    constexpr int kCap = UC_SLAB_BAS_WORDS(UC_N);
    const long long sec0 = sliced ? j.unit_begin : 0;
    const long long sec1 = (sliced && j.unit_end < j.n_sections) ? j.unit_end : j.n_sections;
    double ad[UC_M], x0[UC_N], hd;
    uc_setup(j, cell, ad, x0, hd);
    T a[UC_M];
#pragma unroll
    for (int i = 0; i < UC_M; ++i) a[i] = (T)ad[i];
    const double rtol = BasPrec<S>::tol(j.rtol), atol = BasPrec<S>::tol(j.atol);
    const double T_sec = j.section_period;
    S s;
    s.load(x0);
    double h = hd;
    double buf[UC_REC][UC_LBL];
#pragma unroll
    for (int r = 0; r < UC_REC; ++r)
#pragma unroll
        for (int i = 0; i < UC_LBL; ++i) buf[r][i] = 0.0;
    long long acc = 0, rej = 0;
    int status = 0;  // 0 running, -1 unbound, -2 unresolved
    auto cell_state = [&](bool store) {
        UcSlab io(j, lane, kCap, store);
        int done = 0;
        io.i32(done);
        io.f64(h);
        io.i64(acc);
        io.i64(rej);
        uc_bas_state(io, s);
        io.seek(4 + 2 * UC_N);
#pragma unroll
        for (int r = 0; r < UC_REC; ++r)
#pragma unroll
            for (int i = 0; i < UC_LBL; ++i) io.f64(buf[r][i]);
    };
    if (sliced && sec0 > 0) {
        UcSlab io(j, lane, kCap, false);
        int done = 0;
        io.i32(done);
        if (done) return;
        cell_state(false);
    }
    T k1[UC_N];
    UcForcing<T> frc;
    for (long long sec = sec0; sec < sec1 && status == 0; ++sec) {
        s.reset_time();
        frc.reset(a);
        uc_bas_k1(s, a, k1, frc);
        for (;;) {
            if (acc + rej >= j.max_attempts) { status = -2; break; }
            const double rem = s.remaining(T_sec);
            const bool last = h >= rem;
            const double hs = last ? rem : h;
            const double err = uc_bas_attempt(s, a, hs, k1, rtol, atol, frc);
#if UC_STRICT
            double fac = err == 0.0 ? 5.0 : 0.9 * pow(err, -0.2);
#else
            // err is the squared norm here. The factor needs a few digits
            // only: (err^2)^(-1/10) from the FP32 hardware log2 / exp2
            // (v_log_f32, v_exp_f32) instead of the FP64 library pow, which
            // costs about as much as the stages' own arithmetic. Out-of-range
            // err (0, inf) is clamped below as before.
            double fac = err == 0.0 ? 5.0
                                    : 0.9 * (double)__builtin_amdgcn_exp2f(-0.1f * __builtin_amdgcn_logf((float)err));
#endif
            fac = fac < 0.2 ? 0.2 : (fac > 5.0 ? 5.0 : fac);
            if (err <= 1.0) {
                ++acc;
                if (uc_diverged(s, j.max_value)) { status = -1; break; }
                if (last) {
                    // keep the unclipped step for the next section
                    const double hn = hs * fac;
                    if (hn > h) h = hn;
                    if (h > j.h_max) h = j.h_max;
                    break;
                }
                s.advance_time(hs);
                h = hs * fac;
            } else {
                ++rej;
                h = hs * fac;
            }
            if (h > j.h_max) h = j.h_max;
            if (h < j.h_min) { status = -2; break; }
        }
        if (status != 0) break;
#pragma unroll
        for (int r = UC_REC - 1; r > 0; --r)
#pragma unroll
            for (int i = 0; i < UC_LBL; ++i) buf[r][i] = buf[r - 1][i];
#pragma unroll
        for (int i = 0; i < UC_LBL; ++i) buf[0][i] = s.get(i);
    }
    if (sliced && status == 0 && sec1 < j.n_sections) {
        cell_state(true);
        return;
    }
    int p = status;
    double pt[UC_LBL];
#pragma unroll
    for (int i = 0; i < UC_LBL; ++i) pt[i] = 0.0;
    if (status == 0) {
        const int L = j.n_record < UC_REC ? j.n_record : UC_REC;
        p = 0;
#pragma unroll
        for (int q = 1; q <= UC_MAXP; ++q) {
            if (p != 0 || q > j.max_period || q >= L) continue;
            bool closes = true;
#pragma unroll
            for (int r = 0; r + q < UC_REC; ++r) {
                if (r + q >= L) continue;
#pragma unroll
                for (int i = 0; i < UC_LBL; ++i)
                    closes = closes && fabs(buf[r][i] - buf[r + q][i]) <= j.period_tol * (1.0 + fabs(buf[r][i]));
            }
            if (closes) p = q;
        }
        if (p > 0) {
#pragma unroll
            for (int i = 0; i < UC_LBL; ++i) pt[i] = buf[0][i];
#pragma unroll
            for (int r = 1; r < UC_MAXP; ++r) {
                if (r >= p) continue;
                const bool less = buf[r][0] < pt[0] || (UC_LBL > 1 && buf[r][0] == pt[0] && buf[r][UC_LBL - 1] < pt[UC_LBL - 1]);
                if (less) {
#pragma unroll
                    for (int i = 0; i < UC_LBL; ++i) pt[i] = buf[r][i];
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < UC_REC; ++r) {
                if (r >= L) continue;
#pragma unroll
                for (int i = 0; i < UC_LBL; ++i) pt[i] = pt[i] + buf[r][i];
            }
#pragma unroll
            for (int i = 0; i < UC_LBL; ++i) pt[i] = pt[i] / (double)L;
        }
    }
    j.flag[lane] = p;
#pragma unroll
    for (int i = 0; i < UC_LBL; ++i) j.out0[(long long)i * j.stride + lane] = pt[i];
    if (j.counters) {
        j.counters[lane] = acc;
        j.counters[j.stride + lane] = rej;
    }
    if (sliced) {
        UcSlab io(j, lane, kCap, true);
        int done = 1;
        io.i32(done);
    }
}

#if UC_HARM
// The forced integration needs a few more registers than the library calls
// it replaces, just past the three waves per SIMD the FP64 map runs at;
// the register budget is held there.
#define UC_BAS_WAVES __attribute__((amdgpu_waves_per_eu(3)))
#else
#define UC_BAS_WAVES
#endif
extern "C" __global__ void __launch_bounds__(UC_BLOCK) UC_BAS_WAVES uc_bas_f64(UcJob j) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane < j.n_cells) uc_basins_cell<BasF64>(j, lane);
}
UC_DIRECT(uc_bas_f32c, uc_basins_cell<BasF32C>)
UC_PERSISTENT(uc_bas_f64_q, uc_basins_cell<BasF64>)
UC_PERSISTENT(uc_bas_f32c_q, uc_basins_cell<BasF32C>)
UC_PERSISTENT(uc_bas_f64_s, (uc_basins_cell<BasF64, true>))
UC_PERSISTENT(uc_bas_f32c_s, (uc_basins_cell<BasF32C, true>))

// ---- attractor catalogue: quantised signature keys in an open-addressing
// hash set (linear probing, bounded by the table size), then a per-cell
// lookup of the label the host assigned to each distinct key.
#define UC_KEY_EMPTY 0xFFFFFFFFFFFFFFFFULL

__device__ __forceinline__ u64 uc_sig_key(int p, const double* pt, double quantum) {
    // p in [0, 255]; two signed 28-bit quantised coordinates
    long long qx = 0, qv = 0;
    if (p > 0) {
        qx = (long long)floor(pt[0] / quantum);
        qv = UC_LBL > 1 ? (long long)floor(pt[UC_LBL - 1] / quantum) : 0;
    }
    const u64 mask = (1ULL << 28) - 1ULL;
    return ((u64)(p & 0xFF) << 56) | (((u64)qx & mask) << 28) | ((u64)qv & mask);
}

__device__ __forceinline__ u64 uc_hash(u64 k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
    return k;
}

// job fields used: n_cells, flag (periods), out0 (points), quantum;
// table: `cap` slots (a power of two); keys that find no slot count in *overflow.
extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bas_keys(UcJob j, u64* table, int cap, u64* overflow) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= j.n_cells) return;
    const int p = j.flag[lane];
    if (p < 0) return;
    double pt[UC_LBL];
#pragma unroll
    for (int i = 0; i < UC_LBL; ++i) pt[i] = j.out0[(long long)i * j.stride + lane];
    const u64 key = uc_sig_key(p, pt, j.quantum);
    unsigned slot = (unsigned)(uc_hash(key) & (u64)(cap - 1));
    for (int probe = 0; probe < cap; ++probe) {
        const u64 cur = __atomic_load_n(&table[slot], __ATOMIC_RELAXED);
        if (cur == key) return;
        if (cur == UC_KEY_EMPTY) {
            const u64 old = atomicCAS(&table[slot], UC_KEY_EMPTY, key);
            if (old == UC_KEY_EMPTY || old == key) return;
        }
        slot = (slot + 1) & (unsigned)(cap - 1);
    }
    atomicAdd((unsigned long long*)overflow, 1ULL);
}

extern "C" __global__ void __launch_bounds__(UC_BLOCK) uc_bas_labels(UcJob j, const u64* table, const int* slot_label,
                                                                    int cap) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= j.n_cells) return;
    const int p = j.flag[lane];
    if (p < 0) { j.label[lane] = UC_LABEL_UNBOUND; return; }
    double pt[UC_LBL];
#pragma unroll
    for (int i = 0; i < UC_LBL; ++i) pt[i] = j.out0[(long long)i * j.stride + lane];
    const u64 key = uc_sig_key(p, pt, j.quantum);
    unsigned slot = (unsigned)(uc_hash(key) & (u64)(cap - 1));
    int label = UC_LABEL_UNBOUND;
    for (int probe = 0; probe < cap; ++probe) {
        const u64 cur = table[slot];
        if (cur == key) { label = slot_label[slot]; break; }
        if (cur == UC_KEY_EMPTY) break;
        slot = (slot + 1) & (unsigned)(cap - 1);
    }
    j.label[lane] = label;
}
