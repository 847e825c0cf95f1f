"""GPU parity: HIP kernels (through the C ABI) vs the FP64 CPU oracle.

Tolerances:
  * FP64 strict mode (no FMA contraction): trajectories, flags and peaks are
    expected bit-identical to the oracle; asserted at 1e-12 relative.
  * FP64 fast mode: short-horizon trajectories 1e-10 relative (before one
    Lyapunov time), peaks in periodic windows 1e-6 relative.
  * Lyapunov exponents: 1e-3 absolute against the oracle's reference method
    (finite-difference Benettin), 1e-9 against the oracle's variational method.
  * FP32-compensated: short-horizon trajectories 1e-5 relative, periodic
    peaks 1e-6 relative, LLE 1e-3 absolute (on non-chaotic cells; the error
    distribution on chaotic cells is reported).
"""
import numpy as np
import pytest

import oracle
import ucuda

pytestmark = pytest.mark.gpu


def _lorenz_case(eng, n_rho=64, lo=0.0, hi=250.0, transient=100.0, t_max=100.0, max_peaks=256):
    s = ucuda.preset("lorenz", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n_rho, lo, hi)]
    vals = s.values()
    step = oracle.KrsStep(eng.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, transient, t_max, max_peaks=max_peaks)
    return s, axes, vals, step, job


def _periodic(peaks, count, tol=1e-7, max_period=16, tail=64):
    """True when the last `tail` peaks repeat with some period <= max_period."""
    if count < tail + max_period:
        return False
    p = peaks[count - tail - max_period:count]
    scale = max(1.0, np.abs(p).max())
    for P in range(1, max_period + 1):
        if np.abs(p[P:] - p[:-P]).max() < tol * scale:
            return True
    return False


def test_bif_lorenz_f64_strict_bit_exact(gpu_engine):
    print("case: Lorenz bif 64 rho, 1e4+1e4 steps, F64 strict", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256, strict=True,
                        traj_len=500)
    ref = oracle.bif_cells(step, job, range(64), traj_len=500)
    flags = gpu["flags"].cpu().numpy()
    np.testing.assert_array_equal(flags, ref["flags"])
    ok = ref["flags"] != 0
    tr = gpu["traj"].cpu().numpy()
    np.testing.assert_allclose(tr[:, ok], ref["traj"][:, ok], rtol=1e-12, atol=0)
    pk, iv = gpu["peaks"].cpu().numpy(), gpu["intervals"].cpu().numpy()
    for c in range(64):
        n = ref["flags"][c]
        if n > 0:
            np.testing.assert_allclose(pk[:n, c], ref["peaks"][:n, c], rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(iv[:n, c], ref["intervals"][:n, c], rtol=1e-12, atol=1e-12)


def test_bif_lorenz_f64_fast(gpu_engine):
    print("case: Lorenz bif 64 rho, F64 fast (FMA)", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256, traj_len=100)
    ref = oracle.bif_cells(step, job, range(64), traj_len=100)
    flags = gpu["flags"].cpu().numpy()
    pk = gpu["peaks"].cpu().numpy()
    # periodic windows: peaks to 1e-6 relative
    n_periodic = 0
    for c in range(64):
        n = ref["flags"][c]
        if n > 0 and _periodic(ref["peaks"][:, c], n):
            n_periodic += 1
            assert flags[c] > 0
            m = min(n, flags[c])
            # compare the attractor's peak set (orbit phase may shift by whole periods)
            a = np.sort(np.unique(np.round(ref["peaks"][n - 32:n, c], 5)))
            b = np.sort(np.unique(np.round(pk[m - 32:m, c], 5)))
            assert len(a) == len(b)
            np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-6)
    assert n_periodic >= 4
    # short horizon from a common state: first samples of the recorded run
    ok = (ref["flags"] != 0) & (flags != 0)
    tr = gpu["traj"].cpu().numpy()
    # the transient (1e4 steps) exceeds a Lyapunov time in chaos: compare from a fresh start
    s2 = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4, traj_len=100)
    job2 = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4)
    r2 = oracle.bif_cells(step, job2, range(64), traj_len=100)
    np.testing.assert_allclose(s2["traj"].cpu().numpy(), r2["traj"], rtol=1e-10, atol=1e-12)
    assert ok.sum() > 40 and tr.shape == (100, 64)


def test_bif_lorenz_f32c(gpu_engine):
    print("case: Lorenz bif 64 rho, F32 compensated", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256,
                        precision=ucuda.PREC_F32C)
    ref = oracle.bif_cells(step, job, range(64))
    flags = gpu["flags"].cpu().numpy()
    pk = gpu["peaks"].cpu().numpy()
    n_periodic = 0
    for c in range(64):
        n = ref["flags"][c]
        if n > 0 and _periodic(ref["peaks"][:, c], n):
            n_periodic += 1
            m = flags[c]
            assert m > 32
            a = np.sort(np.unique(np.round(ref["peaks"][n - 32:n, c], 4)))
            b = np.sort(np.unique(np.round(pk[m - 32:m, c], 4)))
            assert len(a) == len(b), (c, a, b)
            np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-6)
    assert n_periodic >= 4
    # fixed-point / divergence classification identical
    np.testing.assert_array_equal(flags <= 0, ref["flags"] <= 0)
    s2 = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4, traj_len=100,
                       precision=ucuda.PREC_F32C)
    job2 = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4)
    r2 = oracle.bif_cells(step, job2, range(64), traj_len=100)
    np.testing.assert_allclose(s2["traj"].cpu().numpy(), r2["traj"], rtol=1e-5, atol=1e-6)


def _rossler_lle_case(eng, n=16):
    s = ucuda.preset("rossler", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, n, 1.0, 20.0)]
    return s, axes, s.values()


def test_lle_rossler_f64(gpu_engine):
    print("case: Rossler LLE 16x16, h=0.01, NT=0.1, 5e3 + 5e4 steps", flush=True)
    s, axes, vals = _rossler_lle_case(gpu_engine)
    args = (s, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, 0.1)
    var_s = ucuda.run_lyap(*args, strict=True)["values"].cpu().numpy()[0]
    fd_s = ucuda.run_lyap(*args, strict=True, method=ucuda.METHOD_FINITE_DIFF)["values"].cpu().numpy()[0]
    var_f = ucuda.run_lyap(*args)["values"].cpu().numpy()[0]
    funcs = oracle.SystemFuncs(s.source(1))
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, NT=0.1)
    cells = range(256)
    r_var = oracle.lyap_var_cells(funcs, 1, job, cells, threads=8)["values"][0]
    r_fd = oracle.lle_fd_cells(step, job, cells, threads=8)["values"][0]
    np.testing.assert_allclose(var_s, r_var, atol=1e-9, rtol=0)
    np.testing.assert_allclose(fd_s, r_fd, atol=1e-9, rtol=0)
    ok = r_fd != 999.0
    assert np.abs(var_s[ok] - r_fd[ok]).max() < 1e-3
    assert np.abs(var_f[ok] - r_fd[ok]).max() < 1e-3
    np.testing.assert_array_equal(var_f == 999.0, r_fd == 999.0)


def test_lle_rossler_f32c(gpu_engine):
    print("case: Rossler LLE 16x16, F32 compensated vs FP64 oracle", flush=True)
    s, axes, vals = _rossler_lle_case(gpu_engine)
    g = ucuda.run_lyap(s, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, 0.1,
                       precision=ucuda.PREC_F32C)["values"].cpu().numpy()[0]
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, NT=0.1)
    r = oracle.lle_fd_cells(step, job, range(256), threads=8)["values"][0]
    ok = r != 999.0
    err = np.abs(g[ok] - r[ok])
    print("F32C LLE |err| median %.2e p90 %.2e max %.2e" % (np.median(err), np.percentile(err, 90), err.max()))
    np.testing.assert_array_equal(g == 999.0, r == 999.0)
    # non-chaotic cells (lambda <= 0.01): within 1e-3
    calm = ok & (r < 0.01)
    assert np.abs(g[calm] - r[calm]).max() < 1e-3
    assert np.median(err) < 1e-3


def test_ls_lorenz_trace_and_oracle(gpu_engine):
    print("case: Lorenz LS 8 rho values, F64 strict vs oracle; sum = -(sigma+1+beta)", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 20.0, 200.0)]
    vals = s.values()
    out = ucuda.run_lyap(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 20.0, 100.0, 0.1, spectrum=True, strict=True)
    g = out["values"].cpu().numpy()
    funcs = oracle.SystemFuncs(s.source(1))
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 20.0, 100.0, NT=0.1)
    r = oracle.lyap_var_cells(funcs, 3, job, range(8))["values"]
    np.testing.assert_allclose(g, r, atol=1e-9, rtol=0)
    # trace of the Jacobian is constant: sum of exponents over the actual
    # integration time ((NT/h + 1) steps per block) is exactly -(sigma+1+beta)
    blocks, nts = int(100.0 / 0.1), int(0.1 / 0.01)
    factor = blocks * (nts + 1) * 0.01 / 100.0
    np.testing.assert_allclose(g.sum(axis=0), -(10 + 1 + 8 / 3) * factor, rtol=1e-6)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    rf = oracle.ls_fd_cells(step, job, range(8))["values"]
    assert np.abs(g - rf).max() < 1e-3 * max(1.0, np.abs(rf).max())


def test_divergence_and_fixed_points(gpu_engine):
    print("case: Lorenz flags: rho < 1 fixed point, tiny max_value diverges", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 0.1, 0.9)]
    vals = s.values()
    g = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 50.0, max_peaks=8)
    assert (g["flags"].cpu().numpy() == -1).all()
    g2 = ucuda.run_bif(s, [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 20.0, 200.0)], [0.1, 0.1, 0.1], vals, 0.01, 10.0,
                       10.0, max_peaks=8, max_value=5.0)
    assert (g2["flags"].cpu().numpy() == 0).all()
    l = ucuda.run_lyap(s, [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 20.0, 200.0)], [0.1, 0.1, 0.1], vals, 0.01, 10.0,
                       10.0, 0.1, max_value=5.0)
    assert (l["values"].cpu().numpy() == 999.0).all() and (l["flags"].cpu().numpy() == -1).all()
