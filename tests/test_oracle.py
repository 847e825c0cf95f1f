"""Pins of the FP64 CPU oracle (CPU only).

Each check is chosen so a plausible slip in the oracle (dropped term, wrong
index or sign, swapped operand) fails it: closed forms (RK4 on linear
systems, exact Lyapunov exponents of diagonal linear systems, trace
invariants), independent known-answer vectors (generator integers computed
here from the published splitmix64 / LCG definition), hand-worked peak
sequences, and algebraic identities of Gram-Schmidt.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import ucuda

D = C.POINTER(C.c_double)


def _linear_system(eng, A):
    n = len(A)
    names = ["x%d" % i for i in range(n)]
    rhs = []
    for i in range(n):
        terms = ["%r*%s" % (float(A[i][j]), names[j]) for j in range(n) if A[i][j] != 0]
        rhs.append("+".join(terms) if terms else "0")
    return ucuda.System(eng, names, [], rhs)


# ------------------------------------------------------------------ RNG ----
def _ref_rng(seed, seq, count):
    """splitmix64 seeding + LCG draw, written from the published definition
    (U-CUDA/kernels/lle1d.template.cu:39-57) with Python integers."""
    M = (1 << 64) - 1
    z = (seed + seq * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    s = z ^ (z >> 31)
    out = []
    for _ in range(count):
        s = (s * 6364136223846793005 + 1442695040888963407) & M
        out.append(float(np.float32(((s >> 40) & 0xFFFFFF) + 1) / np.float32(16777216.0)))
    return out


def test_rng_known_answers():
    for seed, seq in ((1234567891, 0), (1234567891, 1), (1234567891, 262143), (20261203, 77)):
        assert oracle.rng_uniforms(seed, seq, 16) == _ref_rng(seed, seq, 16)
    # draws lie in (0, 1]
    u = oracle.rng_uniforms(1234567891, 5, 1000)
    assert min(u) > 0.0 and max(u) <= 1.0


# ----------------------------------------------------------------- grid ----
def test_axis_values_closed_forms():
    v = oracle.axis_values(False, 1024, 0.0, 250.0)
    assert v[0] == 0.0 and v[-1] == 250.0
    assert v[1] == 250.0 / 1023.0
    assert oracle.axis_values(False, 1, 3.0, 7.0) == [7.0]  # n == 1 -> hi (kernel convention)
    lg = oracle.axis_values(True, 5, 1e-3, 10.0)
    np.testing.assert_allclose(lg, [1e-3, 1e-2, 1e-1, 1.0, 10.0], rtol=1e-14)
    # engine host tables use the same formula
    assert ucuda.axis_values(ucuda.axis(1, 1, 1024, 0.0, 250.0)) == v
    assert ucuda.axis_values(ucuda.axis(1, 1, 5, 1e-3, 10.0, True)) == lg


# ------------------------------------------------------------------ RK4 ----
def test_rk4_step_is_the_taylor_polynomial_on_linear_systems(cpu_engine):
    A = np.array([[-0.3, 1.1, 0.0], [-1.2, -0.1, 0.4], [0.2, 0.0, -0.7]])
    s = _linear_system(cpu_engine, A)
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 1)
    h = 0.05
    x0 = np.array([0.3, -1.0, 2.0])
    x1 = np.array(step.step(list(x0), [0.5], h))
    hA = h * A
    P = np.eye(3) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    np.testing.assert_allclose(x1, P @ x0, rtol=1e-14, atol=1e-15)
    # the engine-order RK4 on the generated right-hand side is the same map
    funcs = oracle.SystemFuncs(s.source(1))
    xa = (C.c_double * 3)(*x0)
    oracle.lib().oc_rk4(funcs.rhs, xa, (C.c_double * 1)(0.5), h, 3)
    assert list(xa) == list(x1)


def test_rk4_tangent_is_the_derivative_of_the_step(cpu_engine):
    s = ucuda.preset("lorenz", cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    a = (C.c_double * 4)(*s.values())
    x = np.array([1.3, -0.4, 20.0])
    V = np.array([0.3, 0.5, -0.2])
    xa, Va = (C.c_double * 3)(*x), (C.c_double * 3)(*V)
    oracle.lib().oc_rk4_tan(funcs.rhs, funcs.tan, xa, Va, 1, a, 0.01, 3)
    e = 1e-6
    xp, xm = (C.c_double * 3)(*(x + e * V)), (C.c_double * 3)(*(x - e * V))
    oracle.lib().oc_rk4(funcs.rhs, xp, a, 0.01, 3)
    oracle.lib().oc_rk4(funcs.rhs, xm, a, 0.01, 3)
    fd = (np.array(list(xp)) - np.array(list(xm))) / (2 * e)
    np.testing.assert_allclose(list(Va), fd, rtol=1e-7, atol=1e-8)


# ----------------------------------------------------------------- loop ----
def test_loop_flags(cpu_engine):
    lor = ucuda.preset("lorenz", cpu_engine)
    step = oracle.KrsStep(cpu_engine.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    L = oracle.lib()
    x = (C.c_double * 3)(0.1, 0.1, 0.1)
    a = (C.c_double * 4)(*lor.values(rho=0.5))
    # rho < 1: the origin attracts -> fixed point
    assert L.oc_loop_model(step.ptr, x, a, 0.01, 20000, 3, 1, 0, 1e6, None) == -1
    x = (C.c_double * 3)(0.1, 0.1, 0.1)
    a = (C.c_double * 4)(*lor.values(rho=28.0))
    assert L.oc_loop_model(step.ptr, x, a, 0.01, 2000, 3, 1, 0, 1e6, None) == 1
    # magnitude check: sum|x| above max_value -> 0
    x = (C.c_double * 3)(0.1, 0.1, 0.1)
    assert L.oc_loop_model(step.ptr, x, a, 0.01, 2000, 3, 1, 0, 5.0, None) == 0
    # blow-up x' = x^2 -> inf/nan -> 0 even with max_value disabled
    bl = ucuda.System(cpu_engine, ["x"], [], ["x*x"])
    sb = oracle.KrsStep(cpu_engine.codegen_scheme(bl.vars, bl.params, bl.rhs), 1, 1)
    x = (C.c_double * 1)(1.0)
    assert L.oc_loop_model(sb.ptr, x, (C.c_double * 1)(0.5), 0.1, 1000, 1, 1, 0, 0.0, None) == 0


def test_loop_records_before_stepping(cpu_engine):
    lor = ucuda.preset("lorenz", cpu_engine)
    step = oracle.KrsStep(cpu_engine.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    a = lor.values()
    data = (C.c_double * 5)()
    x = (C.c_double * 3)(1.0, 2.0, 3.0)
    oracle.lib().oc_loop_model(step.ptr, x, (C.c_double * 4)(*a), 0.01, 5, 3, 1, -1, 1e6, data)
    assert data[0] == 1.0 + math.pi * 2.0 + math.e * 3.0  # x0 + pi*x1 + e*x2 of the initial state
    x2 = step.step([1.0, 2.0, 3.0], a, 0.01)
    assert data[1] == x2[0] + math.pi * x2[1] + math.e * x2[2]


# ---------------------------------------------------------- peak finder ----
def test_peak_finder_on_sampled_sine():
    h = 0.01
    t = np.arange(20000) * h
    pk, iv = oracle.peak_finder(np.sin(2.0 * t), h)
    # maxima at t = pi/4 + k*pi; the first is dropped; parabolic error O(h^2)
    assert len(pk) == int((t[-1] - math.pi / 4) / math.pi + 1) - 1
    np.testing.assert_allclose(pk, 1.0, atol=1e-9)
    np.testing.assert_allclose(iv, math.pi, atol=1e-6)


def test_peak_finder_hand_worked_sequences():
    # plateau of three equal samples: one peak at the LAST plateau sample,
    # refined with its neighbours (denominator 2 - 2*2 + 1 = -1, delta = 0.5)
    d = [0, 1, 2, 2, 2, 1, 0, 1, 3, 1, 0, 0, 5, 4, 0, 0]
    pk, iv = oracle.peak_finder(d, 1.0)
    # raw peaks: j=4 (value 2, delta=(0.5*(2-1))/(2-4+1)=-0.5 -> value 2-0.25*1*(-0.5)=2.125, t=3-0.5=2.5),
    #            j=8 (3: delta=0.5*(1-1)/(1-6+1)=0 -> 3, t=7), j=12 (5: delta=0.5*(0-4)/(0-10+4)=1/3 ->
    #            5-0.25*(-4)*(1/3)=5.3333, t=11+1/3); first dropped
    np.testing.assert_allclose(pk, [3.0, 5.0 + 1.0 / 3.0])
    np.testing.assert_allclose(iv, [7.0 - 2.5, 11.0 + 1.0 / 3.0 - 7.0])
    # rising plateau is not a peak; a single peak gives count 0 after the shift
    assert len(oracle.peak_finder([0, 1, 1, 2, 3, 2, 1, 0], 1.0)[0]) == 0
    assert len(oracle.peak_finder([0, 1, 2], 1.0)[0]) == 0
    # a walk that reaches the end of the record is no peak
    pk, _ = oracle.peak_finder([0, 1, 5, 1, 0, 1, 2, 2, 2, 2], 1.0)
    assert len(pk) == 0


# -------------------------------------------------------- Gram-Schmidt ----
@pytest.mark.parametrize("n", [2, 3, 4])
def test_gram_schmidt_identities(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    B = np.zeros(n * n)
    den = np.zeros(n)
    oracle.lib().oc_cgs(np.ascontiguousarray(A.ravel()).ctypes.data_as(D), B.ctypes.data_as(D), n,
                        den.ctypes.data_as(D))
    Q = B.reshape(n, n)
    np.testing.assert_allclose(Q @ Q.T, np.eye(n), atol=1e-12)
    R = np.linalg.qr(A.T)[1]  # rows of A are the vectors
    np.testing.assert_allclose(den, np.abs(np.diag(R)), rtol=1e-12)
    V = np.ascontiguousarray(A.ravel().copy())
    d2 = np.zeros(n)
    oracle.lib().oc_mgs(V.ctypes.data_as(D), n, n, d2.ctypes.data_as(D))
    np.testing.assert_allclose(d2, den, rtol=1e-12)
    np.testing.assert_allclose(V, B, atol=1e-12)


# ---------------------------------------------------- Lyapunov exponents ----
def _rk4_amp(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def test_lyapunov_on_diagonal_linear_system(cpu_engine):
    """diag(mu): every block multiplies volume by prod R(h mu)^(nts+1) exactly
    (R = RK4 amplification), so the SUM of exponents is exact; each exponent
    converges to its closed form up to the initial-alignment term
    log|c|/t_max of the random probe basis."""
    mus = [-0.5, -1.0, -2.0]
    s = _linear_system(cpu_engine, np.diag(mus))
    h, NT, t_max = 0.01, 0.1, 200.0
    ax = [ucuda.axis(ucuda.AXIS_VAR, 0, 2, 1.0, 2.0)]
    job = oracle.make_job(3, ax, [1.0, 1.0, 1.0], [0.5], h, 1.0, t_max, NT=NT, max_value=0.0)
    blocks, nts = int(t_max / NT), int(NT / h)
    expect = np.array([blocks * (nts + 1) * math.log(_rk4_amp(h * m)) / t_max for m in mus])
    funcs = oracle.SystemFuncs(s.source(1))
    ls = oracle.lyap_var_cells(funcs, 3, job, [0, 1])["values"]
    np.testing.assert_allclose(ls.sum(axis=0), expect.sum(), rtol=1e-12)
    np.testing.assert_allclose(ls[:, 0], expect, atol=0.03)
    lle = oracle.lyap_var_cells(funcs, 1, job, [0, 1])["values"][0]
    np.testing.assert_allclose(lle, expect[0], atol=0.03)
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 1)
    # the reference finite-difference method is exact on a linear system:
    # identical to the variational result up to rounding
    np.testing.assert_allclose(oracle.ls_fd_cells(step, job, [0, 1])["values"], ls, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(oracle.lle_fd_cells(step, job, [0, 1])["values"][0], lle, rtol=1e-6, atol=1e-9)


def test_lorenz_spectrum_invariants(cpu_engine):
    s = ucuda.preset("lorenz", cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    h, NT, t_max = 0.01, 0.1, 200.0
    ax = [ucuda.axis(ucuda.AXIS_PARAM, 2, 1, 28.0, 28.0)]
    job = oracle.make_job(3, ax, [0.1, 0.1, 0.1], s.values(), h, 20.0, t_max, NT=NT)
    ls = oracle.lyap_var_cells(funcs, 3, job, [0])["values"][:, 0]
    factor = int(t_max / NT) * (int(NT / h) + 1) * h / t_max  # steps actually taken / t_max
    # constant trace: sum of exponents = -(sigma + 1 + beta)
    np.testing.assert_allclose(ls.sum(), -(10 + 1 + 8 / 3) * factor, rtol=1e-3)
    # published values (sanity bounds): 0.906, 0, -14.57
    assert abs(ls[0] / factor - 0.906) < 0.06
    assert abs(ls[1] / factor) < 0.02


def test_variational_is_the_small_eps_limit_of_the_reference_method(cpu_engine):
    s = ucuda.preset("rossler", cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    ax = [ucuda.axis(ucuda.AXIS_PARAM, 3, 6, 4.0, 9.0)]
    job = oracle.make_job(3, ax, [1.0, 1.0, 1.0], s.values(), 0.01, 50.0, 200.0, NT=0.1)
    var = oracle.lyap_var_cells(funcs, 1, job, range(6))["values"][0]
    gaps = []
    for eps in (1e-4, 1e-5, 1e-6):
        job.eps = eps
        gaps.append(np.abs(oracle.lle_fd_cells(step, job, range(6))["values"][0] - var).max())
    assert gaps[1] < gaps[0] / 3 and gaps[2] < gaps[1] / 3 and gaps[2] < 1e-5


def test_krs_index_checker():
    assert oracle.check_indices("X[0] = X[2] + a[3];", 3, 4) == []
    d = oracle.check_indices("X[0] = X[3];\n a[4] = 1; // X[9]", 3, 4)
    assert [line for line, _ in d] == [1, 2]


def test_rounding_noise_model(cpu_engine):
    """The noisy variant takes exactly the plain path's steps: a kick size
    that rounds away (1 + 1e-300 == 1) gives bit-identical exponents. Under
    FP64-sized noise and parameter perturbation, a period-1 Rossler cell
    (c = 2.5) keeps its exponent to 1e-5 while a chaotic one (c = 5.7) moves
    by far more (over 550 time units a 1e-15 kick grows by e^(0.07 * 550))."""
    s = ucuda.preset("rossler", cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    ax = [ucuda.axis(ucuda.AXIS_PARAM, 3, 2, 2.5, 5.7)]
    job = oracle.make_job(3, ax, [1.0, 1.0, 1.0], s.values(), 0.01, 50.0, 500.0, NT=0.1)
    var = oracle.lyap_var_cells(funcs, 1, job, [0, 1])["values"]
    same = oracle.lyap_var_cells_noisy(funcs, 1, job, [0, 1], 1e-300, 7)["values"]
    assert np.array_equal(var, same)
    assert np.array_equal(oracle.lyap_var_cells(funcs, 1, oracle.perturbed_job(job, 0.0, (1, 1, 1, 1, 1)),
                                                [0, 1])["values"], var)
    spread = oracle.lyap_rounding_spread(funcs, 1, job, [0, 1], var, "f64")[0]
    assert spread[0] <= 1e-5 < 1e-3 < spread[1], spread


def test_bif_rounding_model(cpu_engine):
    """Bifurcation oracle with the rounding model: a kick that rounds away
    reproduces the plain run bit for bit; the period-1 Lorenz orbit at
    rho = 240 keeps its peaks to 1e-8 under FP64-sized noise, while the
    chaotic rho = 28 cell does not."""
    s = ucuda.preset("lorenz", cpu_engine)
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    ax = [ucuda.axis(ucuda.AXIS_PARAM, 2, 2, 240.0, 28.0)]
    job = oracle.make_job(3, ax, [0.1, 0.1, 0.1], s.values(), 0.01, 100.0, 100.0, max_peaks=256)
    a = oracle.bif_cells(step, job, [0, 1])
    b = oracle.bif_cells(step, job, [0, 1], noise=1e-300, noise_seed=3)
    for k in ("flags", "peaks", "intervals"):
        assert np.array_equal(a[k], b[k], equal_nan=True)
    stable = oracle.bif_rounding_stable(step, job, [0, 1], a, "f64", 1e-8)
    assert stable.tolist() == [True, False], (stable, a["flags"])


def test_hyperchaotic_rossler_spectrum_published(cpu_engine):
    """Rossler's hyperchaotic system (a, d) = (0.25, 0.05) from (-10, -6, 0,
    10.1): two positive exponents, one zero. Published spectrum (Wolf, Swift,
    Swinney & Vastano 1985): 0.112, 0.019, 0, -25.2; the oracle's exponents
    are sums over (NT/h + 1) steps per block divided by t_max, hence the
    20/21 factor."""
    s = ucuda.preset("hyper_rossler", cpu_engine)
    funcs = oracle.SystemFuncs(s.source(1))
    job = oracle.make_job(4, [ucuda.axis(ucuda.AXIS_PARAM, 2, 1, 0.05, 0.05)], [-10, -6, 0, 10.1], s.values(),
                          0.005, 0.0, 2000.0, NT=0.1)
    r = oracle.lyap_var_cells(funcs, 4, job, [0])
    assert r["flags"][0] == 1
    l1, l2, l3, l4 = r["values"][:, 0] * 20 / 21
    assert abs(l1 - 0.112) < 0.01 and abs(l2 - 0.019) < 0.005 and abs(l3) < 2e-3 and -30 < l4 < -20


def test_dopri78_and_cd_step_bodies(cpu_engine):
    """The DOPRI78 and CD step texts (byte-identical to the reference's)
    integrate as their methods say. Linear decay x' = -k x, y' = -m y:
    one CD step with symmetry s = a[0] is exactly
    (1 - s h k) / (1 + (1 - s) h k) (explicit then implicit half step;
    s = 1/2 is the trapezoidal rule); one DOPRI78 step matches exp(-h k) to
    the 13 significant digits its printed coefficients carry, far below
    RK4's h^5 / 120."""
    vars_, params, rhs = ["x", "y"], ["k", "m"], ["-k*x", "-m*y"]
    h, k, m = 0.1, 1.3, 0.4
    for s_sym in (0.5, 0.3):
        body = cpu_engine.codegen_scheme(vars_, params, rhs, ucuda.SCHEME_CD)
        step = oracle.KrsStep(body, 2, 3)
        X = (C.c_double * 2)(1.0, 2.0)
        step.fn(X, (C.c_double * 3)(s_sym, k, m), C.c_double(h))
        for got, x0, c in ((X[0], 1.0, k), (X[1], 2.0, m)):
            expect = x0 * (1 - s_sym * h * c) / (1 + (1 - s_sym) * h * c)
            assert abs(got - expect) <= 1e-15 * abs(expect)
    body = cpu_engine.codegen_scheme(vars_, params, rhs, ucuda.SCHEME_DOPRI78)
    step = oracle.KrsStep(body, 2, 3)
    X = (C.c_double * 2)(1.0, 2.0)
    step.fn(X, (C.c_double * 3)(0.5, k, m), C.c_double(h))
    rk4 = oracle.KrsStep(cpu_engine.codegen_scheme(vars_, params, rhs, ucuda.SCHEME_RK4), 2, 3)
    Y = (C.c_double * 2)(1.0, 2.0)
    rk4.fn(Y, (C.c_double * 3)(0.5, k, m), C.c_double(h))
    for got, ref4, x0, c in ((X[0], Y[0], 1.0, k), (X[1], Y[1], 2.0, m)):
        exact = x0 * math.exp(-h * c)
        assert abs(got - exact) < 1e-12 * x0
        assert abs(ref4 - exact) > 1e3 * abs(got - exact)


def test_bif2d_dbscan_counts():
    """2-D bifurcation period = the reference's DBSCAN cluster count of a
    cell's peaks: a clean period-p peak sequence (p levels further apart than
    eps) gives p for every length, a constant sequence 1, levels closer than
    eps merge (also through a chain of eps-neighbours), 0 / 1 peaks give 0 / 1."""
    for p in range(1, 7):
        levels = 3.0 + 0.7 * np.arange(p)
        for n in range(2 * p, 2 * p + 13):
            assert oracle.dbscan_count(np.tile(levels, 20)[:n], 0.1) == p, (p, n)
    assert all(oracle.dbscan_count([5.0] * n, 0.1) == 1 for n in range(2, 9))
    assert oracle.dbscan_count([1.0, 1.05] * 5, 0.1) == 1
    assert oracle.dbscan_count([1.0, 1.5] * 5, 0.1) == 2
    assert oracle.dbscan_count([0.0, 0.06, 0.12] * 3, 0.1) == 1
    assert oracle.dbscan_count([], 0.1) == 0 and oracle.dbscan_count([2.0], 0.1) == 1


def test_dft_closed_forms():
    """1-D DFT block: x_n = cos(2 pi f0 dt n) over a whole number of cycles of
    every probed frequency difference gives a_k = 1/2 at f0 and 0 elsewhere
    with the rectangular window, ~1/4 with Hanning (its mean is 1/2); below
    1000 samples (no float-precision resync) the rotation equals the direct
    sum; windows: Hanning 0 at both ends and 1 in the middle, Hamming 0.07672
    at the ends, both symmetric."""
    N, dt, f0 = 4000, 0.01, 2.0
    n = np.arange(N)
    x = np.cos(2 * np.pi * f0 * dt * n)
    ak, bk = oracle.dft_block(x, oracle.dft_window(0, N), 5, 1.0, 3.0, dt)
    np.testing.assert_allclose(ak, [0, 0, 0.5, 0, 0], atol=1e-9)
    assert np.abs(bk).max() < 1e-5  # float-precision resync of the phase every 1000 samples
    ak_h, _ = oracle.dft_block(x, oracle.dft_window(1, N), 5, 1.0, 3.0, dt)
    assert abs(ak_h[2] - 0.25) < 1e-3 and np.abs(ak_h[[0, 1, 3, 4]]).max() < 1e-6
    M = 999
    y = np.sin(0.37 * n[:M]) + 0.2 * np.cos(1.3 * n[:M])
    w = oracle.dft_window(2, M)
    freqs = np.linspace(0.1, 4.0, 7)
    a, b = oracle.dft_block(y, w, 7, 0.1, 4.0, dt)
    direct_a = [(w * y * np.cos(2 * np.pi * f * dt * n[:M])).sum() / M for f in freqs]
    direct_b = [(w * y * np.sin(2 * np.pi * f * dt * n[:M])).sum() / M for f in freqs]
    np.testing.assert_allclose(a, direct_a, atol=1e-12)
    np.testing.assert_allclose(b, direct_b, atol=1e-12)
    hann, hamm = oracle.dft_window(1, 101), oracle.dft_window(2, 101)
    assert hann[0] == 0.0 and abs(hann[-1]) < 1e-15 and abs(hann[50] - 1.0) < 1e-15
    np.testing.assert_allclose(hann, hann[::-1], atol=1e-15)
    np.testing.assert_allclose([hamm[0], hamm[-1]], [0.07672, 0.07672], atol=1e-15)
    np.testing.assert_allclose(hamm, hamm[::-1], atol=1e-15)


def test_dft_hsweep_pins(cpu_engine):
    """DFT under an h sweep: x'' = -(2 pi)^2 x from (1, 0) is cos(2 pi t), so
    every step of the sweep (each with its own block length, window and dt)
    puts amplitude 1/2 at 1 Hz and ~0 at the other whole-cycle frequencies
    of the 10 s block (rectangular window; the truncated block leaks
    O(h / t_max)); a block of <= 2 samples is unusable (-1); and a one-point
    h axis with pre_scaller 1 reproduces the fixed-h sweep bit for bit."""
    w = 2.0 * np.pi
    s = _linear_system(cpu_engine, np.array([[0.0, 1.0], [-w * w, 0.0]]))
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 2, 1)
    axes = [ucuda.axis(ucuda.AXIS_STEP, 0, 5, 0.004, 0.01)]
    job = oracle.make_job(2, axes, [1.0, 0.0], [0.5], 0.01, 1.0, 10.0)
    r = oracle.dft_hsweep_cells(step, job, range(5), 4, 0.5, 2.0, window_type=0)
    assert (r["flags"] == 1).all()
    amp = np.hypot(r["ak"], r["bk"])
    np.testing.assert_allclose(amp[:, 1], 0.5, atol=5e-3)
    assert amp[:, [0, 2, 3]].max() < 5e-3
    # h = 0.2 (50 samples), 2.6 (3 samples; RK4 unstable at omega h = 16: escapes, 0), 5 (2 samples: -1)
    job2 = oracle.make_job(2, [ucuda.axis(ucuda.AXIS_STEP, 0, 3, 0.2, 5.0)], [1.0, 0.0], [0.5], 0.01, 1.0, 10.0)
    r2 = oracle.dft_hsweep_cells(step, job2, range(3), 4, 0.5, 2.0)
    assert r2["flags"].tolist() == [1, 0, -1]
    assert (r2["ak"][1] == 0).all() and (r2["ak"][2] == -1).all() and (r2["bk"][2] == -1).all()
    h = 0.0075
    jh = oracle.make_job(2, [ucuda.axis(ucuda.AXIS_STEP, 0, 1, h, h)], [1.0, 0.0], [0.5], 0.01, 1.0, 10.0)
    jc = oracle.make_job(2, [ucuda.axis(ucuda.AXIS_VAR, 0, 1, 1.0, 1.0)], [1.0, 0.0], [0.5], h, 1.0, 10.0)
    rh = oracle.dft_hsweep_cells(step, jh, [0], 6, 0.2, 3.0, window_type=1)
    rc = oracle.dft_cells(step, jc, [0], 6, 0.2, 3.0, window_type=1)
    np.testing.assert_array_equal(rh["ak"], rc["ak"])
    np.testing.assert_array_equal(rh["bk"], rc["bk"])


def _rk4_R(z):
    return 1.0 + z + z * z / 2.0 + z ** 3 / 6.0 + z ** 4 / 24.0


def test_fastsync_closed_forms(cpu_engine):
    """Fast synchronisation on x' = c1 x, y' = c2 y (RK4): the coupled step
    N = K (Xm - Xs) before the step, then X += h N, makes the error e = Xm - Xs
    of each component obey e <- (R(h c) - h K) e with a stored master
    (unidirectional) and e <- (R(h c) - 2 h K) e when the master is coupled
    back (bidirectional). With one pass: estimator 2 = log10 |e| at the
    window's second-to-last sample, estimator 0 = log10 of the RMS over the
    pass (geometric sum), estimator 3 = (first i with |e_i| <= trs + 1) h,
    estimator 1 = 1 when the end error is below trs. On the attractor the
    window starting at sample idx * pre has e0 = R^(skip + idx pre) m0 - s0;
    on the grid the cell's value replaces the swept variable of the master
    (or of the slave with swap), and the fixed state first runs `skip`
    uncoupled steps."""
    c = np.array([-0.3, 0.2])
    s = _linear_system(cpu_engine, np.diag(c))
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 2, 1)
    h, K = 0.01, np.array([4.0, 6.0])
    R = _rk4_R(h * c)
    q1, q2 = R - h * K, R - 2 * h * K
    m0, s0 = np.array([1.0, -2.0]), np.array([0.3, 0.5])
    kw = dict(h=h, iter_of_synchr=1, pre_scaller=3, t_max=1.0, transient_time=0.05, window=0.5, fs_error_trs=1e-12)
    idx = [0, 7, 20]
    skip, N = 5, 50
    e0 = np.array([R ** (skip + i * 3) * m0 - s0 for i in idx])
    n, e2, traj = oracle.fs_attractor(step, 2, [0.5], m0, s0, K, K, idx, error_estim=2, **kw)
    assert n == 100 // 3
    np.testing.assert_allclose(traj, e0 + s0, rtol=1e-13)
    want = np.log10(np.sqrt(((q1 ** (N - 2) * e0) ** 2).sum(axis=1)))
    np.testing.assert_allclose(e2, want, rtol=1e-11)
    _, e0v, _ = oracle.fs_attractor(step, 2, [0.5], m0, s0, K, K, idx, error_estim=0, **kw)
    geo = ((e0 ** 2) * (1 - q1 ** (2 * (N - 1))) / (1 - q1 ** 2)).sum(axis=1)
    np.testing.assert_allclose(e0v, np.log10(np.sqrt(geo / (N - 1))), rtol=1e-11)
    _, eb, _ = oracle.fs_attractor(step, 2, [0.5], m0, s0, K, K, idx, error_estim=2, type_of_synch=1, **kw)
    np.testing.assert_allclose(eb, np.log10(np.sqrt(((q2 ** (N - 2) * e0) ** 2).sum(axis=1))), rtol=1e-11)
    trs = 0.5
    kw3 = dict(kw, fs_error_trs=trs)
    _, e3, _ = oracle.fs_attractor(step, 2, [0.5], m0, s0, K, K, idx, error_estim=3, **kw3)
    first = [next(i for i in range(N - 1) if np.sqrt(((q1 ** i * e) ** 2).sum()) <= trs) for e in e0]
    np.testing.assert_allclose(e3, [(i + 1) * h for i in first], rtol=1e-15)
    _, e1, _ = oracle.fs_attractor(step, 2, [0.5], m0, s0, K, K, idx, error_estim=1, **kw3)
    assert (e1 == 1.0).all()
    # on grid: 5 x 5 cells over (x, y) of the master, the slave after 5 uncoupled steps; swap sweeps the slave
    gkw = dict(h=h, iter_of_synchr=1, pre_scaller=2, transient_time=0.05, window=0.6, axis_x=(-1.0, 1.0),
               axis_y=(0.5, 2.5), n_pts=5, error_estim=2, fs_error_trs=1e-12)
    Ng = int(0.6 / h / 2)
    xv = np.array(oracle.axis_values(False, 5, -1.0, 1.0))
    yv = np.array(oracle.axis_values(False, 5, 0.5, 2.5))
    for swap in (False, True):
        g = oracle.fs_grid(step, 2, [0.5], m0, s0, K, K, range(25), grid_swap_master_slave=swap, **gkw)
        fixed = (R ** 5) * (m0 if swap else s0)
        grid = np.array([[xv[cc % 5], yv[cc // 5]] for cc in range(25)])
        e = (fixed - grid) if swap else (grid - fixed)
        np.testing.assert_allclose(g, np.log10(np.sqrt(((q1 ** (Ng - 2) * e) ** 2).sum(axis=1))), rtol=1e-11)
    # a slave started on the master's own window stays on it: zero error, the trs floor
    _, ez, _ = oracle.fs_attractor(step, 2, [0.5], m0, m0, K, K, [0], error_estim=0, **dict(kw, transient_time=0.0))
    assert ez[0] == -12.0


def _equilibrium(r, branch):
    """Real roots of r + x - x^3 = 0; branch -1 lowest, +1 highest."""
    roots = np.roots([-1.0, 0.0, 1.0, r])
    real = np.sort(roots[np.abs(roots.imag) < 1e-9].real)
    return real[0] if branch < 0 else real[-1]


def test_continuation_chains(cpu_engine):
    """Continuation chains. (1) x' = r + x - x^3 is bistable for |r| <
    2/(3 sqrt 3): walking r upwards the state stays on the lower branch,
    downwards on the upper one, so at r in {-0.25, 0, 0.25} the forward and
    reverse chains report the exponents 1 - 3 x*^2 of different equilibria
    (times the reference block factor (NT/h + 1) h / NT); outside the fold
    both agree. (2) The spectrum chain's exponents sum to the trace of the
    Jacobian at the equilibrium the chain sits on (pitchfork x' = r x - x^3
    plus y' = -y). (3) x'' = -w^2 x with w swept: the chain carries (x, v)
    from point to point, so every peak of point j is the amplitude
    sqrt(x^2 + (v / w)^2) of the RK4 map's matrix power over the steps the
    reference spends per point (transient + 1, block * pre + pre). (4) A
    one-point LLE chain is the ordinary finite-difference sweep of that
    point (same probe subsequence 0)."""
    # (1) hysteresis of the largest exponent
    s = ucuda.System(cpu_engine, ["x"], ["r"], ["r+x-x^3"])
    step = oracle.KrsStep(cpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 1, 2)
    rs = np.linspace(-1.0, 1.0, 9)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 1, 9, -1.0, 1.0)]
    job = oracle.make_job(1, axes, [0.0], [0.5, 0.0], 0.01, 20.0, 20.0, NT=0.1, eps=1e-4)
    fwd = oracle.cont_lyap(step, job, 1)["values"][0]
    rev = oracle.cont_lyap(step, job, 1, reverse=True)["values"][0][::-1]  # back in r order
    factor = 11 * 0.01 / 0.1
    for i, r in enumerate(rs):
        lo, hi = 1 - 3 * _equilibrium(r, -1) ** 2, 1 - 3 * _equilibrium(r, +1) ** 2
        if abs(r) < 2 / (3 * np.sqrt(3)):
            np.testing.assert_allclose([fwd[i], rev[i]], [lo * factor, hi * factor], rtol=2e-3)
        else:
            np.testing.assert_allclose(fwd[i], rev[i], rtol=1e-6)
    # (4) a one-point chain is the ordinary finite-difference sweep
    one = [ucuda.axis(ucuda.AXIS_PARAM, 1, 1, 0.3, 0.3)]
    j1 = oracle.make_job(1, one, [0.4], [0.5, 0.0], 0.01, 5.0, 10.0, NT=0.1, eps=1e-4)
    np.testing.assert_array_equal(oracle.cont_lyap(step, j1, 1)["values"][0], oracle.lle_fd_cells(step, j1, [0])["values"][0])
    # (2) spectrum chain: sum of exponents = trace at the chain's equilibrium
    s2 = ucuda.System(cpu_engine, ["x", "y"], ["r"], ["r*x-x^3", "-y"])
    step2 = oracle.KrsStep(cpu_engine.codegen_scheme(s2.vars, s2.params, s2.rhs), 2, 2)
    j2 = oracle.make_job(2, [ucuda.axis(ucuda.AXIS_PARAM, 1, 5, -1.0, 1.0)], [0.3, 0.2], [0.5, 0.0], 0.01, 30.0, 20.0,
                         NT=0.1, eps=1e-4)
    # walked downwards: r > 0 settles on x = sqrt(r) (trace -2 r - 1), r < 0 on 0 (trace r - 1); walked
    # upwards the chain would still sit on the unstable origin at r = 0.5 (x ~ 1e-33 after the r < 0
    # points: a delayed bifurcation, trace r - 1)
    ls = oracle.cont_lyap(step2, j2, 2, reverse=True)["values"].sum(axis=0)
    rv = np.linspace(1.0, -1.0, 5)
    trace = np.array([r - 1.0 if r < 0 else -2.0 * r - 1.0 for r in rv])
    keep = rv != 0.0  # r = 0: x decays algebraically
    np.testing.assert_allclose(ls[keep], trace[keep] * factor, rtol=2e-3)
    up = oracle.cont_lyap(step2, j2, 2)["values"].sum(axis=0)
    np.testing.assert_allclose(up[3], (0.5 - 1.0) * factor, rtol=2e-3)
    # (3) oscillator amplitudes carried along the chain
    s3 = ucuda.System(cpu_engine, ["x", "v"], ["w"], ["v", "-(w^2)*x"])  # the grammar reads -w^2 as (-w)^2
    step3 = oracle.KrsStep(cpu_engine.codegen_scheme(s3.vars, s3.params, s3.rhs), 2, 2)
    h, T0, T, pre = 0.01, 3.0, 25.0, 2
    for reverse in (False, True):
        ws = np.linspace(0.8, 2.0, 5)[::-1] if reverse else np.linspace(0.8, 2.0, 5)
        j3 = oracle.make_job(2, [ucuda.axis(ucuda.AXIS_PARAM, 1, 5, 0.8, 2.0)], [1.0, 0.0], [0.5, 0.0], h, T0, T,
                             pre_scaller=pre, max_peaks=64)
        r3 = oracle.cont_bif(step3, j3, reverse=reverse)
        st = np.array([1.0, 0.0])
        for jj, w in enumerate(ws):
            hA = h * np.array([[0.0, 1.0], [-w * w, 0.0]])
            P = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
            st = np.linalg.matrix_power(P, int(T0 / h) + 1) @ st
            amp = np.hypot(st[0], st[1] / w)
            blk = int(T / h / pre)
            m = r3["flags"][jj]
            assert m >= int(w * T / (2 * np.pi)) - 1
            np.testing.assert_allclose(r3["peaks"][:m, jj], amp, rtol=1e-6)
            st = np.linalg.matrix_power(P, (blk + 1) * pre) @ st


# --------------------------------------------------------------- basins ----
# The basins integrator is adaptive Dormand-Prince 5(4) (Dormand & Prince
# 1980) with steps clipped to land on every Poincare section, and a period /
# canonical-point search over the trailing section samples. Pins: the
# method's stability polynomials on y' = lam*y (closed forms of the published
# tableau), order 5 / order 4 of the solution / embedded estimate on a
# non-autonomous nonlinear problem with a known solution, global error
# proportional to the tolerance, the period search on synthetic section
# sequences, and the unforced double well (point attractors (+-1, 0), exact
# point symmetry).
def _funcs(eng, vars, params, rhs, time_symbol=None):
    return oracle.SystemFuncs(ucuda.System(eng, vars, params, rhs, time_symbol=time_symbol).source(1))


def test_dp5_step_is_the_stability_polynomial(cpu_engine):
    f = _funcs(cpu_engine, ["y"], ["lam"], ["lam*y"])
    for lam, h in ((-1.0, 0.1), (0.7, 0.5), (-3.0, 1.0), (2.0, 0.25), (-0.5, 2.0)):
        z = lam * h
        R5 = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24 + z**5 / 120 + z**6 / 600
        # embedded-pair difference: R5 - R4 of the published weights
        E = -97.0 / 120000 * z**5 + 13.0 / 40000 * z**6 - 1.0 / 24000 * z**7
        x5, e, _ = oracle.dp5_step(f.rhs, [1.5], [0.5, lam], 0.0, h)
        assert x5[0] == pytest.approx(1.5 * R5, rel=1e-14, abs=1e-15)
        assert e[0] == pytest.approx(1.5 * E, rel=1e-12, abs=1e-17)


def test_dp5_orders_on_a_nonautonomous_problem(cpu_engine):
    # u' = 3 cos(3t) u (1 - u), u = 1 / (1 + 2 exp(-sin 3t)): over a fixed
    # interval the 5th-order solution converges as h^5 and x5 - e (the
    # embedded 4th-order solution) as h^4; one step's estimate e scales as
    # h^5. A slipped A, b, c or E entry drops one of these orders.
    f = _funcs(cpu_engine, ["u", "w"], [], ["3*cos(3*t)*u*(1-u)", "u*w"], time_symbol="t")
    t0, span = 0.3, 0.8

    def exact(t):
        return 1.0 / (1.0 + 2.0 * math.exp(-math.sin(3 * t)))

    y0 = [exact(t0), 1.0]
    e5, e4, est = [], [], []
    for n in (8, 16, 32, 64):
        h = span / n
        y, y4 = list(y0), list(y0)
        for k in range(n):
            x5, _, _ = oracle.dp5_step(f.rhs, y, [0.5], t0 + k * h, h)
            y = list(x5)
            x5b, eb, _ = oracle.dp5_step(f.rhs, y4, [0.5], t0 + k * h, h)
            y4 = list(x5b - eb)
        e5.append(abs(y[0] - exact(t0 + span)))
        e4.append(abs(y4[0] - exact(t0 + span)))
        est.append(abs(oracle.dp5_step(f.rhs, y0, [0.5], t0, h)[1][0]))
    for k in range(3):
        assert 28 < e5[k] / e5[k + 1] < 42, e5     # 2^5
        assert 14 < e4[k] / e4[k + 1] < 19, e4     # 2^4
        assert 28 < est[k] / est[k + 1] < 44, est  # 2^5 (h = 0.1 is not yet asymptotic)


def _section_job(n_vars, ic, values, rtol, n_sections, T, **kw):
    axes = [dict(kind=oracle.AXIS_PARAM, index=1, n_pts=1, lo=values[1], hi=values[1], log_scale=0)]
    return oracle.make_basins_job(n_vars, axes, ic, values, 0.01, rtol, rtol, T, n_sections, **kw)


def test_basins_global_error_proportional_to_tolerance(cpu_engine):
    # x' = w v, v' = -w x: every section t = k 2pi/w returns to the start, so
    # the canonical point after 40 sections is the global error at t = 80 pi/w
    f = _funcs(cpu_engine, ["x", "v"], ["w"], ["w*v", "-w*x"])
    errs = []
    for tol in (1e-6, 1e-8, 1e-10):
        bj = _section_job(2, [1.2, -0.4], [0.5, 1.3], tol, 40, 2 * math.pi / 1.3)
        r = oracle.basins_cells(f, bj, [0])
        assert r["period"][0] == 1
        errs.append(float(np.hypot(r["point"][0, 0] - 1.2, r["point"][1, 0] + 0.4)))
        assert errs[-1] < 2e3 * tol, (tol, errs[-1])
    for k in range(2):
        ratio = errs[k] / errs[k + 1]
        assert 30 < ratio < 300, errs   # 100x tolerance -> ~100x error


def test_section_signature_on_synthetic_sequences():
    rng = np.random.default_rng(7)
    for p in range(1, 9):
        cyc = rng.uniform(-2, 2, size=(p, 2))
        for phase in range(p):
            seq = np.array([cyc[(r + phase) % p] for r in range(16)])
            seq += rng.uniform(-1e-6, 1e-6, size=seq.shape)          # far below the 1e-4 closure
            got, pt = oracle.section_signature(seq, max_period=8, tol=1e-4)
            assert got == p, (p, got)
            best = min(range(p), key=lambda r: (seq[r, 0], seq[r, 1]))
            assert np.array_equal(pt, seq[best])                     # lexicographic minimum of the cycle
    # ties in x0 are broken by x1
    seq = np.array([[0.5, 2.0], [0.5, -1.0], [3.0, 0.0]] * 6)[:16]
    assert oracle.section_signature(seq)[0] == 3
    assert np.array_equal(oracle.section_signature(seq)[1], [0.5, -1.0])
    # period 2 also closes at 4 and 6: the smallest is reported
    seq = np.array([[1.0, 0.0], [-1.0, 0.0]] * 8)
    assert oracle.section_signature(seq)[0] == 2
    # period 9 is beyond max_period 8; no closure -> 0 and the mean
    cyc = rng.uniform(-2, 2, size=(9, 2))
    seq = np.array([cyc[r % 9] for r in range(16)])
    p, pt = oracle.section_signature(seq, max_period=8)
    assert p == 0 and np.allclose(pt, seq.mean(axis=0), rtol=0, atol=1e-15)
    # closure is relative: |d| <= tol (1 + |x|); one sample 3 tol off breaks period 1
    seq = np.ones((16, 2)) * 10.0
    seq[7, 1] += 3e-4 * 11.0
    assert oracle.section_signature(seq, tol=1e-4)[0] == 0
    seq[7, 1] = 10.0 + 0.9e-4 * 11.0
    assert oracle.section_signature(seq, tol=1e-4)[0] == 1


def test_basins_unforced_double_well(cpu_engine):
    # x'' + 0.15 x' - x + x^3 = 0: point attractors (+-1, 0); the flow is odd,
    # so mirrored initial states give exactly mirrored sections and steps
    f = _funcs(cpu_engine, ["x", "v"], ["delta", "gamma", "omega"],
               ["v", "-delta*v+x-x^3+gamma*cos(omega*t)"], time_symbol="t")
    rng = np.random.default_rng(11)
    ics = rng.uniform(-2.5, 2.5, size=(24, 2))
    for x0, v0 in ics:
        res = []
        for sgn in (1.0, -1.0):
            bj = _section_job(2, [sgn * x0, sgn * v0], [0.5, 0.15, 0.0, 1.0], 1e-9, 60, 2 * math.pi)
            res.append(oracle.basins_cells(f, bj, [0]))
        a, b = res
        assert a["period"][0] == 1 and b["period"][0] == 1
        assert abs(abs(a["point"][0, 0]) - 1.0) < 1e-6 and abs(a["point"][1, 0]) < 1e-6
        assert np.array_equal(b["point"][:, 0], -a["point"][:, 0])
        assert np.array_equal(a["steps"], b["steps"])
    # inside a well near its bottom the cell stays in that well
    for x0 in (0.8, 1.2, -0.9):
        bj = _section_job(2, [x0, 0.05], [0.5, 0.15, 0.0, 1.0], 1e-9, 60, 2 * math.pi)
        r = oracle.basins_cells(f, bj, [0])
        assert r["point"][0, 0] == pytest.approx(math.copysign(1.0, x0), abs=1e-6)
