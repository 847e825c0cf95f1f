"""Codegen against the reference codegen's own outputs (CPU only).

tests/golden/codegen_golden.json was written by tests/golden/make_golden.py
from the reference's codegen.cpp (compiled from its source): KRS step text
per scheme, right-hand-side values from its interpreter, normalize_value.
The Jacobian emitter has no reference counterpart; it is pinned by central
finite differences of the generated right-hand side.
"""
import json
import math
import os
import random

import numpy as np
import pytest

import oracle
import ucuda

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "codegen_golden.json")))


@pytest.mark.parametrize("sy", GOLDEN["systems"], ids=lambda s: s["name"])
def test_krs_text_matches_reference(cpu_engine, sy):
    for sch, text in sy["krs"].items():
        if text is None:
            with pytest.raises(ucuda.UcError):
                cpu_engine.codegen_scheme(sy["vars"], sy["params"], sy["rhs"], int(sch), latex=sy["latex"])
        else:
            assert cpu_engine.codegen_scheme(sy["vars"], sy["params"], sy["rhs"], int(sch), latex=sy["latex"]) == text


@pytest.mark.parametrize("sy", [s for s in GOLDEN["systems"] if s["evals"]], ids=lambda s: s["name"])
def test_rhs_values_match_reference_interpreter(cpu_engine, sy):
    s = ucuda.System(cpu_engine, sy["vars"], sy["params"], sy["rhs"], latex=sy["latex"])
    funcs = oracle.SystemFuncs(s.source(1))
    import ctypes as C
    for ev in sy["evals"]:
        host = s.eval_rhs(ev["X"], ev["a"])
        np.testing.assert_allclose(host, ev["dX"], rtol=1e-14, atol=1e-14)
        X = (C.c_double * len(ev["X"]))(*ev["X"])
        a = (C.c_double * len(ev["a"]))(*ev["a"])
        out = (C.c_double * len(ev["X"]))()
        fn = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                         C.POINTER(C.c_double))(funcs.rhs.value)
        fn(X, a, 0.0, out)
        # generated C uses the same expression trees; libm may differ by an ulp
        np.testing.assert_allclose(list(out), ev["dX"], rtol=1e-13, atol=1e-13)


def test_normalize_value_matches_reference():
    import ctypes as C
    L = ucuda.lib()
    for val, expect in GOLDEN["normalize"].items():
        eng = ucuda.Engine(-1)
        # normalize_value is exercised through a one-variable system whose rhs is the value
        if expect is None:
            continue
        s = eng.codegen_scheme(["x"], [], [val], ucuda.SCHEME_EULER_CROMER)
        assert s.strip() == "X[0] = X[0] + h * (%s);" % expect, (val, s)
    del L, C


def _tangent_fn(funcs):
    import ctypes as C
    P = C.POINTER(C.c_double)
    return C.CFUNCTYPE(None, P, P, C.c_double, P, P, C.c_int)(funcs.tan.value), \
        C.CFUNCTYPE(None, P, P, C.c_double, P)(funcs.rhs.value)


@pytest.mark.parametrize("name", ["lorenz", "rossler", "hyper_rossler", "duffing", "chen"])
def test_jacobian_matches_finite_differences(cpu_engine, name):
    import ctypes as C
    s = ucuda.preset(name, cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    tan, rhs = _tangent_fn(funcs)
    rng = random.Random(7)
    n = s.n_vars
    for _ in range(5):
        X = [rng.uniform(-2, 2) for _ in range(n)]
        a = s.values()
        t = rng.uniform(0, 6)
        V = [rng.uniform(-1, 1) for _ in range(n * 2)]
        Xa, aa = (C.c_double * n)(*X), (C.c_double * len(a))(*a)
        Va, JV = (C.c_double * (2 * n))(*V), (C.c_double * (2 * n))()
        tan(Xa, aa, t, Va, JV, 2)
        for k in range(2):
            e = 1e-6
            xp = (C.c_double * n)(*[X[i] + e * V[k * n + i] for i in range(n)])
            xm = (C.c_double * n)(*[X[i] - e * V[k * n + i] for i in range(n)])
            fp, fm = (C.c_double * n)(), (C.c_double * n)()
            rhs(xp, aa, t, fp)
            rhs(xm, aa, t, fm)
            fd = [(fp[i] - fm[i]) / (2 * e) for i in range(n)]
            np.testing.assert_allclose([JV[k * n + i] for i in range(n)], fd, rtol=1e-6, atol=1e-7)


def test_jacobian_of_functions_system(cpu_engine):
    sy = [s for s in GOLDEN["systems"] if s["name"] == "functions"][0]
    import ctypes as C
    s = ucuda.System(cpu_engine, sy["vars"], sy["params"], sy["rhs"])
    funcs = oracle.SystemFuncs(s.source(1))
    tan, rhs = _tangent_fn(funcs)
    X, a = [0.7, 1.3, 1.9], [0.5, 1.5, 2.5]
    for j in range(3):
        V = [1.0 if i == j else 0.0 for i in range(3)]
        Xa, aa, Va, JV = (C.c_double * 3)(*X), (C.c_double * 3)(*a), (C.c_double * 3)(*V), (C.c_double * 3)()
        tan(Xa, aa, 0.0, Va, JV, 1)
        e = 1e-6
        fp, fm = (C.c_double * 3)(), (C.c_double * 3)()
        rhs((C.c_double * 3)(*[X[i] + e * V[i] for i in range(3)]), aa, 0.0, fp)
        rhs((C.c_double * 3)(*[X[i] - e * V[i] for i in range(3)]), aa, 0.0, fm)
        np.testing.assert_allclose(list(JV), [(fp[i] - fm[i]) / (2 * e) for i in range(3)], rtol=1e-6, atol=1e-6)


def test_parse_errors(cpu_engine):
    with pytest.raises(ucuda.UcError):
        ucuda.System(cpu_engine, ["x"], [], ["foo(x)"])
    with pytest.raises(ucuda.UcError):
        ucuda.System(cpu_engine, ["x"], [], ["x + q"])
    with pytest.raises(ucuda.UcError):
        ucuda.System(cpu_engine, ["x"], [], ["(x + 1"])


def test_unary_minus_binds_tighter_than_power(cpu_engine):
    # reference grammar: -x^2 is (-x)^2
    s = ucuda.System(cpu_engine, ["x"], [], ["-x^2"])
    assert s.eval_rhs([3.0], [0.5])[0] == 9.0


@pytest.mark.parametrize("name", ["lorenz", "duffing"])
def test_hiprtc_compiles_for_gfx950_without_gpu(cpu_engine, name):
    s = ucuda.preset(name, cpu_engine)
    assert s.compile(strict=False) > 10000
    assert s.compile(strict=True) > 10000


def test_hiprtc_compiles_latex_and_function_systems(cpu_engine):
    for sy in GOLDEN["systems"]:
        if sy["name"] in ("functions", "latex_misc"):
            s = ucuda.System(cpu_engine, sy["vars"], sy["params"], sy["rhs"], latex=sy["latex"])
            assert s.compile() > 10000


def test_op_counts(cpu_engine):
    s = ucuda.preset("lorenz", cpu_engine)
    c = s.op_counts()
    # sigma*(y-x): 2; x*(rho-z)-y: 3; x*y-beta*z: 3
    assert c["rhs_flops"] == 8 and c["rhs_calls"] == 0
    assert c["jac_nonzeros"] == 8
    assert ucuda.rk4_flops(c, 3) == 4 * 8 + 36
    _ = math
