// C shim over the reference codegen (U-CUDA/codegen.hpp:25-74), compiled with
// the reference's own codegen.cpp into oracle/_ref/. TEST INFRASTRUCTURE ONLY:
// used in this container to write golden KRS-body text and right-hand-side
// values under tests/golden/.
#include <cstring>
#include <string>
#include <vector>

#include "codegen.hpp"

namespace {
System make_system(const char* const* vars, int n_vars, const char* const* params, int n_params,
                   const char* const* rhs, int latex) {
    System s;
    for (int i = 0; i < n_vars; ++i) s.vars.push_back(vars[i]);
    for (int i = 0; i < n_params; ++i) s.params.push_back(params[i]);
    for (int i = 0; i < n_vars; ++i) s.rhs.push_back(rhs[i]);
    s.latex = latex != 0;
    return s;
}
}  // namespace

extern "C" int ref_codegen_scheme(const char* const* vars, int n_vars, const char* const* params, int n_params,
                                  const char* const* rhs, int latex, int scheme, char* buf, size_t cap) {
    try {
        std::string t = codegen_scheme(make_system(vars, n_vars, params, n_params, rhs, latex), (Scheme)scheme);
        if (t.size() + 1 > cap) return -2;
        std::memcpy(buf, t.c_str(), t.size() + 1);
        return 0;
    } catch (...) {
        return -1;
    }
}

extern "C" int ref_eval_rhs(const char* const* vars, int n_vars, const char* const* params, int n_params,
                            const char* const* rhs, int latex, const double* X, const double* a, double* dX) {
    try {
        SystemEvaluator ev(make_system(vars, n_vars, params, n_params, rhs, latex));
        ev.eval(X, a, dX);
        return 0;
    } catch (...) {
        return -1;
    }
}

extern "C" int ref_normalize_value(const char* v, char* buf, size_t cap) {
    try {
        std::string t = normalize_value(v);
        if (t.size() + 1 > cap) return -2;
        std::memcpy(buf, t.c_str(), t.size() + 1);
        return 0;
    } catch (...) {
        return -1;
    }
}
