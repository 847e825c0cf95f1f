"""Writes tests/golden/codegen_golden.json from the REFERENCE codegen.

Runs only in a container where the reference tree is mounted: it loads
oracle/_ref/libref_codegen.so (the reference's own codegen.cpp compiled from
/root/reference by `make -C oracle ref`) and records, for a set of systems,
the KRS step text of each scheme (codegen_scheme), right-hand-side values
from the reference interpreter (SystemEvaluator::eval) at seeded points, and
normalize_value outputs. The fixture is data only; tests compare the
engine's codegen against it.
"""
import ctypes as C
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_codegen.so")

SYSTEMS = [
    dict(name="lorenz", vars=["x", "y", "z"], params=["sigma", "rho", "beta"],
         rhs=["sigma*(y-x)", "x*(rho-z)-y", "x*y-beta*z"], latex=False),
    dict(name="rossler", vars=["x", "y", "z"], params=["a", "b", "c"],
         rhs=["-y-z", "x+a*y", "b+z*(x-c)"], latex=False),
    dict(name="chen", vars=["x", "y", "z"], params=["a", "b", "c"],
         rhs=["a*(y-x)", "(c-a)*x-x*z+c*y", "x*y-b*z"], latex=False),
    dict(name="hyper_rossler", vars=["x", "y", "z", "w"], params=["a", "d"],
         rhs=["-y-z", "x+a*y+w", "3+x*z", "-0.5*z+d*w"], latex=False),
    dict(name="duffing_theta", vars=["x", "v", "th"], params=["delta", "gamma", "omega"],
         rhs=["v", "-delta*v+x-x^3+gamma*cos(th)", "omega"], latex=False),
    dict(name="functions", vars=["x", "y", "z"], params=["p", "q"],
         rhs=["sin(x)*exp(-y)+pow(z, 2.5)-atan2(x, y)", "sqrt(fabs(x))+log(1+y*y)/q-tanh(z)*p",
              "-x^2+8/3*y-1e-5*z+cbrt(x)*cosh(0.1*y)"], latex=False),
    dict(name="lorenz_latex", vars=["x", "y", "z"], params=["sigma", "rho", "beta"],
         rhs=["\\sigma( y-x )", "x \\cdot ( \\rho-z )-y", "x y-\\beta z"], latex=True),
    # an identifier directly followed by '(' is a call: the reference rejects it
    dict(name="latex_call_error", vars=["x", "y", "z"], params=["rho"],
         rhs=["y", "x ( \\rho-z )-y", "x"], latex=True),
    dict(name="latex_misc", vars=["x", "y"], params=["a", "b"],
         rhs=["\\frac{x}{1+y^{2}}-a \\cdot \\sin{y}", "\\left( x-b \\right) \\cdot y+\\exp{-x}"], latex=True),
    # CD classification cases: coef*var, var*coef, bare var, (a)*(b),
    # name*(...), powers, var*var, calls, absent variable
    dict(name="cd_linear", vars=["x", "y", "z"], params=["a", "b", "c"],
         rhs=["-a*x + y", "y*b - x*z", "z + 2*x - c"], latex=False),
    dict(name="cd_parens", vars=["x", "y", "z"], params=["a", "b"],
         rhs=["(a - b)*(y - x)", "a*(x - y + z)", "(x + y)*(z - 1)"], latex=False),
    dict(name="cd_nonlinear", vars=["u", "v", "w"], params=["k", "m"],
         rhs=["u^2 - v", "sin(v) + k*u", "w*w - m*exp(w) + u"], latex=False),
    dict(name="cd_absent", vars=["p", "q"], params=["g"],
         rhs=["q", "-g*sin(p) - 0.1*q"], latex=False),
    dict(name="cd_one_var", vars=["x"], params=["r"], rhs=["r*x*(1-x)"], latex=False),
]
NORMALIZE = ["8/3", "1e-5", "2", "8.5", "-0.5", "10", "0.1", "2^3", "3.0e2", "1/7"]
SCHEMES = [0, 1, 2, 3, 4, 5]  # Euler, Euler-Cromer, Explicit Midpoint, RK4, DOPRI78, CD


def strs(items):
    arr = (C.c_char_p * len(items))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


def main():
    L = C.CDLL(LIB)
    out = dict(systems=[], normalize={})
    rng = random.Random(20261203)
    for sy in SYSTEMS:
        v, p, r = strs(sy["vars"]), strs(sy["params"]), strs(sy["rhs"])
        rec = dict(sy, krs={}, evals=[])
        for sch in SCHEMES:
            buf = C.create_string_buffer(1 << 16)
            rc = L.ref_codegen_scheme(v, len(sy["vars"]), p, len(sy["params"]), r, int(sy["latex"]), sch, buf, 1 << 16)
            rec["krs"][str(sch)] = buf.value.decode() if rc == 0 else None
        for _ in range(5):
            X = [rng.uniform(0.2, 2.0) for _ in sy["vars"]]
            a = [0.5] + [rng.uniform(0.5, 3.0) for _ in sy["params"]]
            Xa = (C.c_double * len(X))(*X)
            aa = (C.c_double * len(a))(*a)
            dX = (C.c_double * len(X))()
            if L.ref_eval_rhs(v, len(sy["vars"]), p, len(sy["params"]), r, int(sy["latex"]), Xa, aa, dX) == 0:
                rec["evals"].append(dict(X=X, a=a, dX=list(dX)))
        out["systems"].append(rec)
    for val in NORMALIZE:
        buf = C.create_string_buffer(256)
        rc = L.ref_normalize_value(val.encode(), buf, 256)
        out["normalize"][val] = buf.value.decode() if rc == 0 else None
    for bad in ["2*pi", "x+1", "1+"]:
        buf = C.create_string_buffer(256)
        out["normalize"][bad] = buf.value.decode() if L.ref_normalize_value(bad.encode(), buf, 256) == 0 else None
    with open(os.path.join(HERE, "codegen_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", len(out["systems"]), "systems")


if __name__ == "__main__":
    main()
