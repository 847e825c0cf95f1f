// codegen.hpp — system definition -> integrator step text and HIP device code.
//
// Keeps the reference API (`System`, `Scheme`, `codegen_scheme`,
// `normalize_value`; U-CUDA/codegen.hpp:10-48) so callers that build a KRS
// ("step body") string keep working, and adds the emitters the MI355X engine
// needs: a right-hand side and a symbolic Jacobian-vector product, printed as
// templated device functions that hipRTC compiles for gfx950 together with the
// hand-written ensemble kernels.
#pragma once
#include <string>
#include <vector>

namespace ucuda {

// vars -> X[0..N-1]; params -> a[1..M] (a[0] is reserved for the CD scheme's
// symmetry coefficient); rhs[i] is dvars[i]/dt. Mirrors U-CUDA/codegen.hpp:10-15.
// `time_symbol` is an extension: when non-empty, that name may appear in rhs
// and denotes the independent variable t (needed for forced systems).
struct System {
    std::vector<std::string> vars;
    std::vector<std::string> params;
    std::vector<std::string> rhs;
    bool latex = false;
    std::string time_symbol;
};

// Same enumerators and order as U-CUDA/codegen.hpp:21.
enum class Scheme { Euler, EulerCromer, ExplicitMidpoint, RK4, DOPRI78, CD };

// KRS step body text in the reference's format (statements over X[], a[], h,
// `numb` locals). Throws std::runtime_error on parse errors.
std::string codegen_scheme(const System& s, Scheme sch);
Scheme scheme_from_name(const std::string& name);
std::string normalize_value(const std::string& value);

// Device-code dialects for the right-hand side / tangent emitters.
enum class Dialect {
    Hip,   // template<typename T> __device__ __forceinline__ ...
    C      // static inline C functions, T = double (CPU oracle)
};

// Emits, for the given system:
//   uc_rhs(X, a, t, dX)            dX[i] = f_i(X, a, t)
//   uc_tangent<K>(X, a, t, V, JV)  JV[k*N+i] = sum_j df_i/dx_j * V[k*N+j]
// The Jacobian entries are derived symbolically from the same expression
// trees as the right-hand side, constant-folded, and evaluated once per call
// (shared by the K tangent vectors). In the C dialect K is a run-time
// argument. Also emits UC_N / UC_M / UC_USES_TIME macros.
std::string emit_system_source(const System& s, Dialect d);

// Device source for a caller-supplied KRS step body (custom schemes): the
// body becomes uc_krs_step<numb>(X, a, h) with the reference prelude's
// names; right-hand side and Jacobian are absent (zero stubs).
std::string emit_custom_krs_source(const std::string& body, int n_vars, int n_values);
// Range check of constant X[k] / a[k] subscripts; empty string when clean.
std::string check_krs_indices(const std::string& body, int n_vars, int n_values);

// Number of nonzero Jacobian entries and right-hand-side operation counts,
// used for the roofline accounting (+,-,* and / each count 1; calls counted
// separately).
struct OpCounts { int rhs_flops = 0; int rhs_calls = 0; int jac_nonzeros = 0; int jac_flops = 0; int jac_calls = 0; };
OpCounts count_ops(const System& s);

// Evaluates the right-hand side in double through the parsed trees (host-side
// interpreter used for validation; mirrors SystemEvaluator, U-CUDA/codegen.cpp:907-948).
void eval_rhs(const System& s, const double* X, const double* a, double t, double* dX);

}  // namespace ucuda
