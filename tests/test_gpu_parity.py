"""GPU parity: HIP kernels (through the C ABI) vs the FP64 CPU oracle.

Tolerances:
  * FP64 strict mode (no FMA contraction, the reference's operation order):
    trajectories, flags and peaks are expected bit-identical to the oracle;
    asserted at 1e-12 relative. Exponents differ from the oracle by at most a
    few 1e-17: the device log and glibc's log round differently in the last
    bit.
  * FP64 fast mode: short-horizon trajectories 1e-10 relative (before one
    Lyapunov time), peaks in periodic windows 1e-6 relative.
  * Lyapunov exponents, strict FP64: 1e-9 against the oracle for both the
    variational method and the reference finite-difference method (same
    trajectories). The variational result is the eps -> 0 limit of the
    reference method (their gap shrinks linearly with eps).
  * Lyapunov exponents, fast FP64 and FP32-compensated: 1e-3 absolute on
    cells whose FP64 exponent is reproducible: three oracle runs with state
    noise and parameter perturbations of the size of the rounding each mode
    injects (1e-15 for FP64, ~1e-7 for FP32) move it by <= 1e-5. On chaotic
    cells any non-identical rounding changes the finite-time exponent by
    ~1e-2 (the FP64 oracle's own spread); the error there is reported, not
    asserted.
  * Peaks are asserted on cells whose oracle peaks survive state noise and
    parameter perturbations of the mode's rounding size (oracle.ROUNDING):
    FP64 1e-6 relative; FP32-compensated 1e-6 relative (values) and 1e-5
    (intervals): with the recorded window in double (the default) on the
    whole 1024-rho sweep, on the cells that hold 5e-7 under float-sized noise
    in the transient and double-sized noise in the record; with float
    recording on the cells that hold 1e-6 under float-sized noise.
  * FP32-compensated short-horizon trajectories: within 4x the deviation
    that float-sized per-step noise produces in the FP64 oracle (the float
    stages are amplified by the flow; ~3e-4 relative at t = 1 for rho = 250).
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


def _stable_cells(step, job, cells, precision, rel):
    """FP64 oracle peaks and the cells whose peak count and values (within
    `rel`) survive state noise and parameter perturbations of the precision's
    rounding size (periodic windows with enough contraction)."""
    a = oracle.bif_cells(step, job, cells, threads=8)
    return a, oracle.bif_rounding_stable(step, job, cells, a, precision, rel, threads=8)


def _axes_of(job):
    return [dict(kind=job.axis_kind[i], index=job.axis_index[i], n_pts=job.axis_n[i], lo=job.axis_lo[i],
                 hi=job.axis_hi[i], log_scale=bool(job.axis_log[i])) for i in range(job.n_axes)]


def _check_periodic_peaks(gpu, ref, stable, rel, rel_intervals=None):
    flags = gpu["flags"].cpu().numpy()
    pk, iv = gpu["peaks"].cpu().numpy(), gpu["intervals"].cpu().numpy()
    for c in np.nonzero(stable)[0]:
        n = ref["flags"][c]
        assert flags[c] == n, (c, flags[c], n)
        np.testing.assert_allclose(pk[:n, c], ref["peaks"][:n, c], rtol=rel, atol=rel)
        ri = rel if rel_intervals is None else rel_intervals
        np.testing.assert_allclose(iv[:n, c], ref["intervals"][:n, c], rtol=ri, atol=ri)


def test_bif_lorenz_f64_fast(gpu_engine):
    print("case: Lorenz bif 64 rho, F64 fast (FMA)", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256)
    ref, stable = _stable_cells(step, job, range(64), "f64", 1e-8)
    assert stable.sum() >= 8, stable.sum()
    _check_periodic_peaks(gpu, ref, stable, 1e-6)
    flags = gpu["flags"].cpu().numpy()
    np.testing.assert_array_equal(flags <= 0, ref["flags"] <= 0)
    # short horizon from the initial state (no transient): 1e-10 relative
    s2 = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4, traj_len=100)
    job2 = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4)
    r2 = oracle.bif_cells(step, job2, range(64), traj_len=100)
    np.testing.assert_allclose(s2["traj"].cpu().numpy(), r2["traj"], rtol=1e-10, atol=1e-12)


def test_bif_lorenz_f32c_full_sweep(gpu_engine):
    """The reference's CPU case whole (1024 rho on [0, 250], 1e4 + 1e4 RK4
    steps, <= 256 peaks) in FP32C with the recorded window in double (the
    default record_precision). Checked cells: those whose FP64 oracle peaks
    move by <= 5e-7 (half the 1e-6 bar) when the transient gets float-sized
    state noise and parameter perturbations and the record double-sized
    noise, i.e. the rounding this method injects; there every peak must
    match to 1e-6 relative."""
    print("case: Lorenz bif 1024 rho (config 1), F32C transient + F64 record", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine, n_rho=1024)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256,
                        precision=ucuda.PREC_F32C)
    ref, stable = _stable_cells(step, job, range(1024), "f32_rec64", 5e-7)
    assert stable.sum() >= 40, stable.sum()
    flags = gpu["flags"].cpu().numpy()
    pk = gpu["peaks"].cpu().numpy()
    errs = [np.abs(pk[:n, c] - ref["peaks"][:n, c]).max() / np.abs(ref["peaks"][:n, c]).max()
            for c, n in zip(np.nonzero(stable)[0], ref["flags"][stable]) if flags[c] == n]
    print("FP32C + F64 record: %d stable cells, worst relative peak error %.2e" % (stable.sum(), max(errs)))
    _check_periodic_peaks(gpu, ref, stable, 1e-6, rel_intervals=1e-5)
    # flag classes (periodic / fixed point / diverged) on the whole sweep
    assert ((flags <= 0) == (ref["flags"] <= 0)).mean() >= 0.99


def test_bif_lorenz_f32c(gpu_engine):
    print("case: Lorenz bif 64 rho, F32 compensated with the record in float too", flush=True)
    s, axes, vals, step, job = _lorenz_case(gpu_engine)
    gpu = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 100.0, 100.0, max_peaks=256,
                        precision=ucuda.PREC_F32C, record_precision=ucuda.PREC_F32C)
    # float-sized noise moves the peaks of most periodic Lorenz cells by
    # 1e-6..1e-5 (weak contraction along the orbit); the cells that hold 1e-6
    # under it are checked at 1e-6, the others are reported
    ref, stable = _stable_cells(step, job, range(64), "f32", 1e-6)
    assert stable.sum() >= 2, stable.sum()
    periodic = (ref["flags"] > 8) & (gpu["flags"].cpu().numpy() == ref["flags"])
    pk = gpu["peaks"].cpu().numpy()
    worst = max(np.abs(pk[:n, c] - ref["peaks"][:n, c]).max() / np.abs(ref["peaks"][:n, c]).max()
                for c, n in zip(np.nonzero(periodic)[0], ref["flags"][periodic])) if periodic.any() else 0.0
    print("FP32C peaks: %d rounding-stable cells; worst relative error over %d equal-count cells %.1e"
          % (stable.sum(), periodic.sum(), worst))
    # peak values 1e-6; intervals come from the parabolic refinement of float
    # samples (a second difference divided by ~h^2): 1e-5
    _check_periodic_peaks(gpu, ref, stable, 1e-6, rel_intervals=1e-5)
    flags = gpu["flags"].cpu().numpy()
    np.testing.assert_array_equal(flags <= 0, ref["flags"] <= 0)
    # short horizon: float stages inject ~6e-8 relative per step and at rho
    # ~ 250 the origin's unstable eigenvalue (~45) amplifies that to ~3e-4
    # relative by t = 1. Bound: the running maximum of |gpu - oracle| stays
    # within 4x the running maximum deviation that 6e-8 per-step state noise
    # plus float-sized parameter perturbations produce in the FP64 oracle,
    # plus 1e-6 of the trajectory's scale (a CPU float emulation of the step
    # reaches 0.74x of that bound).
    s2 = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4, traj_len=100,
                       precision=ucuda.PREC_F32C, record_precision=ucuda.PREC_F32C)
    job2 = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.0, 1.0, max_peaks=4)
    r2 = oracle.bif_cells(step, job2, range(64), traj_len=100)["traj"]
    spread = np.zeros_like(r2)
    for k, signs in enumerate(oracle._PERTURB_SIGNS):
        rn = oracle.bif_cells(step, oracle.perturbed_job(job2, 6e-8, signs), range(64), traj_len=100, noise=6e-8,
                              noise_seed=3000 + k, noise_every=1)["traj"]
        spread = np.maximum(spread, np.abs(rn - r2))
    err = np.maximum.accumulate(np.abs(s2["traj"].cpu().numpy() - r2), axis=0)
    bound = 4.0 * (np.maximum.accumulate(spread, axis=0) + 1e-6 * np.abs(r2).max(axis=0))
    print("FP32C short horizon: max relative error %.1e; max err / bound %.2f"
          % ((err / np.abs(r2).max(axis=0)).max(), (err / bound).max()))
    assert (err <= bound).all()


def _rossler_lle_case(eng, n=16):
    s = ucuda.preset("rossler", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, n, 1.0, 20.0)]
    return s, axes, s.values()


def _lle_oracles(eng, s, axes, vals, n_cells, ic=(1.0, 1.0, 1.0), precision="f64"):
    """FP64 variational oracle, and its spread over three runs with state
    noise and parameter perturbations of the precision's rounding size
    (1e-15 for FP64, ~1e-7 for FP32-compensated, oracle.ROUNDING). A cell is
    reproducible when the spread is <= 1e-5 (chaotic cells move by ~1e-2)."""
    funcs = oracle.SystemFuncs(s.source(1))
    step = oracle.KrsStep(eng.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, list(ic), vals, 0.01, 50.0, 500.0, NT=0.1)
    cells = range(n_cells)
    r_var = oracle.lyap_var_cells(funcs, 1, job, cells, threads=8)["values"][0]
    spread = oracle.lyap_rounding_spread(funcs, 1, job, cells, r_var, precision, threads=8)[0]
    repro = (r_var != 999.0) & (spread <= 1e-5)
    return step, job, r_var, repro, spread


def _check_sensitive(err, spread, bounded, frac=0.98, c=6.0):
    """The precision contract where the FP64 oracle itself is not
    reproducible: on at least `frac` of the bounded cells the error is within
    max(c x the oracle's rounding spread, 1e-3) (the spread is the largest
    move of three noisy oracle runs; a run that rounds differently behaves
    like a fourth), and the median error does not exceed the median spread."""
    within = err[bounded] <= np.maximum(c * spread[bounded], 1e-3)
    sens = bounded & (spread > 1e-5)
    assert within.mean() >= frac, (within.mean(), np.sort(err[bounded] / np.maximum(spread[bounded], 1e-300))[-5:])
    if sens.sum() >= 8:
        assert np.median(err[sens]) <= np.median(spread[sens]), (np.median(err[sens]), np.median(spread[sens]))
    return within.mean()


def test_lle_rossler_f64(gpu_engine):
    print("case: Rossler LLE 16x16, h=0.01, NT=0.1, 5e3 + 5e4 steps, F64", flush=True)
    s, axes, vals = _rossler_lle_case(gpu_engine)
    args = (s, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, 0.1)
    var_s = ucuda.run_lyap(*args, strict=True)["values"].cpu().numpy()[0]
    fd_s = ucuda.run_lyap(*args, strict=True, method=ucuda.METHOD_FINITE_DIFF)["values"].cpu().numpy()[0]
    var_f = ucuda.run_lyap(*args)["values"].cpu().numpy()[0]
    step, job, r_var, repro, _ = _lle_oracles(gpu_engine, s, axes, vals, 256)
    r_fd = oracle.lle_fd_cells(step, job, range(256), threads=8)["values"][0]
    # strict: identical trajectories -> both methods equal the oracle
    np.testing.assert_allclose(var_s, r_var, atol=1e-9, rtol=0)
    np.testing.assert_allclose(fd_s, r_fd, atol=1e-9, rtol=0)
    ok = r_fd != 999.0
    np.testing.assert_array_equal(var_f == 999.0, ~ok)
    # variational = eps -> 0 limit of the reference method: the gap shrinks with eps
    job.eps = 1e-6
    r_fd6 = oracle.lle_fd_cells(step, job, range(256), threads=8)["values"][0]
    assert np.abs(var_s[ok] - r_fd6[ok]).max() < 1e-4
    assert np.abs(var_s[ok] - r_fd6[ok]).max() < np.abs(var_s[ok] - r_fd[ok]).max()
    # fast (FMA) mode: 1e-3 on cells whose FP64 exponent is reproducible,
    # the oracle's own rounding spread on the others
    assert repro.sum() > 64
    assert np.abs(var_f[repro] - r_var[repro]).max() < 1e-3
    _, _, _, _, spread = _lle_oracles(gpu_engine, s, axes, vals, 256)
    w = _check_sensitive(np.abs(var_f - r_var), spread, ok)
    print("fast FP64: %.4f of bounded cells within max(6 x spread, 1e-3)" % w)


def test_lle_rossler_f32c(gpu_engine):
    print("case: Rossler LLE 16x16, F32 compensated vs FP64 oracle", flush=True)
    s, axes, vals = _rossler_lle_case(gpu_engine)
    g = ucuda.run_lyap(s, axes, [1.0, 1.0, 1.0], vals, 0.01, 50.0, 500.0, 0.1,
                       precision=ucuda.PREC_F32C)["values"].cpu().numpy()[0]
    step, job, r_var, repro, spread = _lle_oracles(gpu_engine, s, axes, vals, 256, precision="f32")
    ok = r_var != 999.0
    np.testing.assert_array_equal(g == 999.0, ~ok)
    err = np.abs(g - r_var)
    print("reproducible cells %d: max |err| %.2e; sensitive cells: median |err| %.2e vs FP64 spread %.2e"
          % (repro.sum(), err[repro].max(), np.median(err[ok & ~repro]), np.median(spread[ok & ~repro])))
    assert repro.sum() > 64
    assert err[repro].max() < 1e-3
    w = _check_sensitive(err, spread, ok)
    print("FP32C: %.4f of bounded cells within max(6 x spread, 1e-3)" % w)


def test_lle_two_cells_per_lane(gpu_engine):
    """A 511 x 511 map (odd cell count, 1.99 pair-lane waves per SIMD) takes
    the two-cells-per-lane kernels in fast FP64 and FP32C. Over a short
    horizon chaotic growth stays small, so both agree cell by cell with the
    strict one-cell-per-lane FP64 kernel, including the cells that fail the
    magnitude check (max_value 6, ~2.5% of the cells by the oracle on every
    61st cell): same flags, state at the check."""
    print("case: Rossler LLE 511x511 two cells per lane vs strict one cell per lane", flush=True)
    s = ucuda.preset("rossler", gpu_engine)
    n = 511
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, n, 1.0, 20.0)]
    args = (s, axes, [1.0, 1.0, 1.0], s.values(a=0.2), 0.01, 2.0, 5.0, 0.1)
    ref = ucuda.run_lyap(*args, strict=True, max_value=6.0, want_final_state=True)
    rv, rf = ref["values"].cpu().numpy()[0], ref["flags"].cpu().numpy()
    rs = ref["final_state"].cpu().numpy()
    assert 100 < (rf == -1).sum() < rf.size // 2
    for prec, tol, stol in ((ucuda.PREC_F64, 1e-9, 1e-9), (ucuda.PREC_F32C, 1e-3, 1e-3)):
        g = ucuda.run_lyap(*args, precision=prec, max_value=6.0, want_final_state=True)
        gv, gf = g["values"].cpu().numpy()[0], g["flags"].cpu().numpy()
        gs = g["final_state"].cpu().numpy()
        same = gf == rf
        live = same & (rf == 1)
        dead = same & (rf == -1)
        err = np.abs(gv[live] - rv[live])
        serr = np.abs(gs - rs).max(axis=0) / np.maximum(1.0, np.abs(rs).max(axis=0))
        print("precision %d: flags equal on %.5f, max |LLE err| %.2e, max state err live %.2e dead %.2e"
              % (prec, same.mean(), err.max(), serr[live].max(), serr[dead].max()))
        assert same.mean() > 0.999
        assert err.max() < tol
        assert np.all(gv[dead] == 999.0)
        assert serr[live].max() < stol and serr[dead].max() < stol
    # a right-hand side with a time symbol (forced Duffing): the two cells of
    # a lane share the system time
    print("case: forced Duffing LLE 511x511 over (x0, v0), two cells per lane vs strict", flush=True)
    d = ucuda.preset("duffing", gpu_engine)
    daxes = [ucuda.axis(ucuda.AXIS_VAR, 0, n, -2.5, 2.5), ucuda.axis(ucuda.AXIS_VAR, 1, n, -2.5, 2.5)]
    dargs = (d, daxes, [0.0, 0.0], d.values(), 0.01, 2.0, 10.0, 0.1)
    ref = ucuda.run_lyap(*dargs, strict=True)
    rv, rf = ref["values"].cpu().numpy()[0], ref["flags"].cpu().numpy()
    assert (rf == 1).all()
    for prec, tol in ((ucuda.PREC_F64, 1e-9), (ucuda.PREC_F32C, 1e-3)):
        g = ucuda.run_lyap(*dargs, precision=prec)
        gv, gf = g["values"].cpu().numpy()[0], g["flags"].cpu().numpy()
        print("precision %d: max |LLE err| %.2e" % (prec, np.abs(gv - rv).max()))
        assert (gf == rf).all() and np.abs(gv - rv).max() < tol


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
    # (the RK4 tangent map's determinant matches exp(h tr J) to O(h^5) per step)
    np.testing.assert_allclose(g.sum(axis=0), -(10 + 1 + 8 / 3) * factor, rtol=1e-3)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    rf = oracle.ls_fd_cells(step, job, range(8))["values"]
    assert np.abs(g - rf).max() < 1e-3 * max(1.0, np.abs(rf).max())
    # fast FP64 and FP32C over 64 rho in [20, 200]: 1e-3 on the cells whose
    # oracle spectrum holds 1e-4 under the mode's rounding noise, the spread
    # contract on every exponent elsewhere (>= 95%: 64 cells)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 64, 20.0, 200.0)]
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 20.0, 100.0, NT=0.1)
    r = oracle.lyap_var_cells(funcs, 3, job, range(64), threads=8)["values"]
    ok = r[0] != 999.0
    for prec, name, min_repro in ((ucuda.PREC_F64, "f64", 5), (ucuda.PREC_F32C, "f32", 2)):
        gp = ucuda.run_lyap(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 20.0, 100.0, 0.1, spectrum=True,
                            precision=prec)["values"].cpu().numpy()
        sp = oracle.lyap_rounding_spread(funcs, 3, job, range(64), r, name, threads=8)
        err = np.abs(gp - r)
        repro = ok & (sp.max(axis=0) <= 1e-4)
        assert repro.sum() >= min_repro, repro.sum()
        assert err[:, repro].max() < 1e-3
        within = [_check_sensitive(err[k], sp[k], ok, frac=0.95) for k in range(3)]
        print("Lorenz %s spectrum: %d reproducible cells; within max(6 x spread, 1e-3) %s"
              % (name, repro.sum(), np.round(within, 4)))


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


def test_streaming_peak_finder_matches_batch_on_plateaus(gpu_engine):
    print("case: streaming peak finder vs batch oracle, 512 integer series x 200 samples", flush=True)
    import torch
    s = ucuda.preset("lorenz", gpu_engine)
    rng = np.random.default_rng(11)
    n_s, n_series = 200, 512
    data = rng.integers(0, 5, size=(n_s, n_series)).astype(np.float64)
    data[:, 0] = [0, 1, 2, 2, 2, 1, 0, 1, 3, 1, 0, 0, 5, 4, 0, 0] * 12 + [0] * 8
    data[:, 1] = np.sin(0.37 * np.arange(n_s))
    data[:, 2] = 1.0  # constant
    g = ucuda.peak_finder(s, torch.from_numpy(data).cuda(), 0.01, max_peaks=100)
    counts, pk, iv = g["counts"].cpu().numpy(), g["peaks"].cpu().numpy(), g["intervals"].cpu().numpy()
    for c in range(n_series):
        rp, ri = oracle.peak_finder(data[:, c], 0.01)
        n = min(len(rp), 100)
        assert counts[c] == n, (c, counts[c], len(rp))
        np.testing.assert_array_equal(pk[:n, c], rp[:n])
        np.testing.assert_array_equal(iv[:n, c], ri[:n])
    # short series: fewer than 5 samples never produce peaks
    g2 = ucuda.peak_finder(s, torch.from_numpy(data[:4].copy()).cuda(), 0.01, max_peaks=4)
    assert (g2["counts"].cpu().numpy() == 0).all()


def test_sweep_kinds_ic_step_log_and_writable(gpu_engine):
    print("case: IC axis x step axis, log-scale param axis, writable_var=-1, pre_scaller=2", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    vals = s.values()
    cases = [
        [ucuda.axis(ucuda.AXIS_VAR, 0, 6, -5.0, 5.0), ucuda.axis(ucuda.AXIS_STEP, 0, 4, 0.005, 0.01)],
        [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 1.0, 300.0, log_scale=True)],
        [ucuda.axis(ucuda.AXIS_PARAM, 1, 1, 3.0, 10.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 5, 1.0, 3.0)],
    ]
    for axes in cases:
        for wv, pre in ((0, 1), (-1, 2), (2, 3)):
            n = ucuda.n_cells(axes)
            gpu = ucuda.run_bif(s, axes, [0.1, 0.2, 0.3], vals, 0.01, 2.0, 5.0, writable_var=wv, pre_scaller=pre,
                                max_peaks=64, strict=True, traj_len=50)
            job = oracle.make_job(3, axes, [0.1, 0.2, 0.3], vals, 0.01, 2.0, 5.0, writable_var=wv, pre_scaller=pre,
                                  max_peaks=64)
            ref = oracle.bif_cells(step, job, range(n), traj_len=50)
            np.testing.assert_array_equal(gpu["flags"].cpu().numpy(), ref["flags"])
            ok = ref["flags"] != 0
            np.testing.assert_array_equal(gpu["traj"].cpu().numpy()[:, ok], ref["traj"][:, ok])
            for c in np.nonzero(ref["flags"] > 0)[0]:
                m = ref["flags"][c]
                np.testing.assert_array_equal(gpu["peaks"].cpu().numpy()[:m, c], ref["peaks"][:m, c])


def test_ls_hyper_rossler_random_ics(gpu_engine):
    print("case: hyper-Rossler 4-exponent spectrum: random ICs in [-1,1]^4 over (d, a), and an IC box "
          "around the hyperchaotic attractor; F64 strict + F32C", flush=True)
    s = ucuda.preset("hyper_rossler", gpu_engine)
    vals = s.values()
    funcs = oracle.SystemFuncs(s.source(1))
    # (1) the sweep over (d, a) with random ICs: trajectories from [-1,1]^4
    # escape (z' = 3 + xz), about half of them by t = 15; flags must agree
    # cell for cell
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 0.0, 0.1), ucuda.axis(ucuda.AXIS_PARAM, 1, 8, 0.2, 0.3)]
    kw = dict(spectrum=True, ic_seed=20261203, ic_lo=-1.0, ic_hi=1.0)
    g = ucuda.run_lyap(s, axes, None, vals, 0.005, 0.0, 15.0, 0.1, strict=True, **kw)["values"].cpu().numpy()
    job = oracle.make_job(4, axes, None, vals, 0.005, 0.0, 15.0, NT=0.1, ic_seed=20261203, ic_lo=-1.0, ic_hi=1.0)
    r = oracle.lyap_var_cells(funcs, 4, job, range(64), threads=8)["values"]
    ok = r[0] != 999.0
    assert 16 < ok.sum() < 48, ok.sum()
    np.testing.assert_array_equal(g[0] == 999.0, ~ok)
    np.testing.assert_allclose(g[:, ok], r[:, ok], atol=1e-9, rtol=0)
    # (2) ICs (x0, w0) around the attractor at (a, d) = (0.25, 0.05), 16 x 16:
    # most cells stay bounded; strict FP64 equals the oracle on every cell.
    # Fast FP64 and FP32C: 1e-3 on every exponent of the cells whose oracle
    # spectrum moves by <= 1e-4 under the mode's rounding noise, and the
    # spread contract on each exponent elsewhere; bounded / escaped agrees on
    # >= 99% (an escape decided by rounding can fall inside the last block,
    # after its magnitude check: NaN exponents with flag 1, as the
    # reference's block-start check gives)
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, 16, -12.0, -8.0), ucuda.axis(ucuda.AXIS_VAR, 3, 16, 8.1, 12.1)]
    ic = [-10.0, -6.0, 0.0, 10.1]
    g = ucuda.run_lyap(s, axes, ic, vals, 0.005, 5.0, 50.0, 0.1, strict=True, spectrum=True)["values"].cpu().numpy()
    job = oracle.make_job(4, axes, ic, vals, 0.005, 5.0, 50.0, NT=0.1)
    r = oracle.lyap_var_cells(funcs, 4, job, range(256), threads=8)["values"]
    ok = r[0] != 999.0
    assert ok.sum() > 128
    np.testing.assert_array_equal(g[0] == 999.0, ~ok)
    np.testing.assert_allclose(g[:, ok], r[:, ok], atol=1e-9, rtol=0)
    for prec, name, min_repro in ((ucuda.PREC_F64, "f64", 128), (ucuda.PREC_F32C, "f32", 20)):
        gp = ucuda.run_lyap(s, axes, ic, vals, 0.005, 5.0, 50.0, 0.1, precision=prec,
                            spectrum=True)["values"].cpu().numpy()
        sp = oracle.lyap_rounding_spread(funcs, 4, job, range(256), r, name, threads=8)
        fin = np.isfinite(gp).all(axis=0) & (gp[0] != 999.0)
        assert (fin == ok).mean() >= 0.99, (fin == ok).mean()
        err = np.where(np.isfinite(gp), np.abs(gp - r), np.inf)
        repro = ok & fin & (sp.max(axis=0) <= 1e-4)
        assert repro.sum() >= min_repro, repro.sum()
        assert err[:, repro].max() < 1e-3, err[:, repro].max(axis=1)
        within = [_check_sensitive(err[k], sp[k], ok) for k in range(4)]
        print("%s spectrum: %d reproducible cells, max |err| there %s; within max(6 x spread, 1e-3) %s; "
              "median |err| %s" % (name, repro.sum(), err[:, repro].max(axis=1), np.round(within, 4),
                                   np.median(err[:, ok & fin], axis=1)))


def test_ls_spread_lanes_equal_one_lane(gpu_engine):
    """Spectrum maps on grids of few waves run each cell's N tangent vectors
    on N lanes (uc_ls_*_xn, Gram-Schmidt across the lanes). A 256 x 256 map
    (one wave per SIMD: one lane per cell) and small ranges of it (the N-lane
    kernel) give the same bits: exponents, flags and final states, with
    cells that fail the magnitude check, in strict FP64, fast FP64 and
    FP32C; hyper-Rossler (N = 4, 16 cells per wave) and Lorenz (N = 3, 21
    cells per wave and one idle lane)."""
    cases = [("hyper_rossler", [ucuda.axis(ucuda.AXIS_VAR, 0, 256, -12.0, -8.0),
                                ucuda.axis(ucuda.AXIS_VAR, 3, 256, 8.1, 12.1)], [-10.0, -6.0, 0.0, 10.1],
              0.005, 1.0, 4.0, 40.0),
             ("lorenz", [ucuda.axis(ucuda.AXIS_PARAM, 2, 256, 20.0, 200.0),
                         ucuda.axis(ucuda.AXIS_PARAM, 1, 256, 5.0, 16.0)], [0.1, 0.1, 0.1], 0.01, 2.0, 6.0, 200.0)]
    for name, axes, ic, h, tt, tmax, maxv in cases:
        s = ucuda.preset(name, gpu_engine)
        for kw in (dict(strict=True), dict(precision=ucuda.PREC_F64), dict(precision=ucuda.PREC_F32C)):
            print("case: %s spectrum 256x256 whole vs ranges, %s" % (name, kw), flush=True)
            args = (s, axes, ic, s.values(), h, tt, tmax, 0.1)
            opt = dict(spectrum=True, max_value=maxv, want_final_state=True, **kw)
            whole = ucuda.run_lyap(*args, **opt)
            wv = whole["values"].cpu().numpy()
            wf = whole["flags"].cpu().numpy()
            ws = whole["final_state"].cpu().numpy()
            dead = 0
            for b, c in ((0, 1000), (1000, 37), (40000, 4099), (65536 - 21, 21)):
                part = ucuda.run_lyap(*args, cells=(b, c), **opt)
                pv = part["values"].cpu().numpy()
                pf = part["flags"].cpu().numpy()
                np.testing.assert_array_equal(pf, wf[b:b + c])
                np.testing.assert_array_equal(pv, wv[:, b:b + c])
                np.testing.assert_array_equal(part["final_state"].cpu().numpy(), ws[:, b:b + c])
                dead += int((pf == -1).sum())
            # the oracle's flags on every 7th compared cell: ~0.4% (hyper-Rossler, max
            # value 40) and ~56% (Lorenz, 200) of the cells diverge
            print("  %d cells compared, %d of them diverged" % (1000 + 37 + 4099 + 21, dead))
            assert 0 < dead < 5000


def test_custom_krs_bodies(gpu_engine):
    print("case: custom KRS step bodies (Euler, midpoint, RK4, DOPRI78, CD text) -> bif + FD LLE vs oracle",
          flush=True)
    lor = ucuda.preset("lorenz", gpu_engine)
    vals = lor.values()
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 10.0, 200.0)]
    for scheme in (ucuda.SCHEME_EULER, ucuda.SCHEME_MIDPOINT, ucuda.SCHEME_RK4, ucuda.SCHEME_DOPRI78,
                   ucuda.SCHEME_CD):
        body = gpu_engine.codegen_scheme(lor.vars, lor.params, lor.rhs, scheme)
        ks = ucuda.KrsSystem(gpu_engine, body, 3, 4)
        g = ucuda.run_bif(ks, axes, [0.1, 0.1, 0.1], vals, 0.005, 5.0, 10.0, max_peaks=32, strict=True, traj_len=40)
        step = oracle.KrsStep(body, 3, 4)
        job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.005, 5.0, 10.0, max_peaks=32)
        r = oracle.bif_cells(step, job, range(16), traj_len=40)
        np.testing.assert_array_equal(g["flags"].cpu().numpy(), r["flags"])
        ok = r["flags"] != 0
        np.testing.assert_array_equal(g["traj"].cpu().numpy()[:, ok], r["traj"][:, ok])
    # the RK4 step text integrates exactly like the built-in RK4 (same operation order)
    body = gpu_engine.codegen_scheme(lor.vars, lor.params, lor.rhs, ucuda.SCHEME_RK4)
    ks = ucuda.KrsSystem(gpu_engine, body, 3, 4)
    a = ucuda.run_lyap(ks, axes, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 20.0, 0.1, method=ucuda.METHOD_FINITE_DIFF,
                       strict=True)["values"].cpu().numpy()
    b = ucuda.run_lyap(lor, axes, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 20.0, 0.1, method=ucuda.METHOD_FINITE_DIFF,
                       strict=True)["values"].cpu().numpy()
    np.testing.assert_array_equal(a, b)
    with pytest.raises(ucuda.UcError):
        ucuda.run_lyap(ks, axes, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 20.0, 0.1)  # variational needs a Jacobian


def test_bif2d_periods(gpu_engine):
    print("case: 2-D bifurcation periods (Lorenz rho x sigma 24x12, DBSCAN eps 0.1), strict vs oracle", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 24, 20.0, 250.0), ucuda.axis(ucuda.AXIS_PARAM, 1, 12, 5.0, 16.0)]
    vals = s.values()
    g = ucuda.run_bif_periods(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 30.0, 30.0, eps_dbscan=0.1,
                              strict=True).cpu().numpy()
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 30.0, 30.0)
    r = oracle.bif_periods(step, job, range(24 * 12), 0.1, threads=8)
    np.testing.assert_array_equal(g, r)
    assert len(set(r.tolist())) >= 4  # fixed points, periodic windows and chaos all present
    g32 = ucuda.run_bif_periods(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 30.0, 30.0, eps_dbscan=0.1,
                                precision=ucuda.PREC_F32C).cpu().numpy()
    print("F32C periods equal to FP64 on %.3f of the cells" % (g32 == r).mean())
    assert ((g32 <= 0) == (r <= 0)).mean() > 0.95


def test_bif2d_periods_many_peaks(gpu_engine):
    print("case: 2-D periods with > 1024 peaks per cell (two oscillators, x + pi v + e y, 250 s): full-capacity path",
          flush=True)
    s = ucuda.System(gpu_engine, ["x", "v", "y", "u"], ["w", "c"], ["v", "-(w^2)*x", "u", "-(c*w^2)*y"])
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 1, 4, 2.0, 31.4), ucuda.axis(ucuda.AXIS_PARAM, 2, 2, 2.25, 6.25)]
    ic, vals = [1.0, 0.0, 3.0, 0.0], [0.5, 1.0, 1.0]
    g = ucuda.run_bif_periods(s, axes, ic, vals, 0.01, 1.0, 250.0, eps_dbscan=0.1, writable_var=-1,
                              strict=True).cpu().numpy()
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 4, 3)
    r = oracle.bif_periods(step, oracle.make_job(4, axes, ic, vals, 0.01, 1.0, 250.0, writable_var=-1), range(8),
                           0.1, threads=8)
    np.testing.assert_array_equal(g, r)
    fl = oracle.bif_cells(step, oracle.make_job(4, axes, ic, vals, 0.01, 1.0, 250.0, writable_var=-1,
                                                max_peaks=20000), range(8))
    assert fl["flags"].max() > 1024 and fl["flags"].min() < 1024  # both kernels took part
    assert len(set(r.tolist())) >= 6


def test_dft_hsweep(gpu_engine):
    print("case: DFT over h (Lorenz, 40 steps in [0.002, 0.06], per-cell blocks), strict vs oracle", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    vals = s.values()
    for axes, kw, pre in (([ucuda.axis(ucuda.AXIS_STEP, 0, 40, 0.002, 0.06)],
                           dict(n_freq=48, freq_lo=0.0, freq_hi=4.0, window_type=1), 1),
                          ([ucuda.axis(ucuda.AXIS_STEP, 0, 24, 0.004, 0.03, True),
                            ucuda.axis(ucuda.AXIS_PARAM, 2, 3, 20.0, 160.0)],
                           dict(n_freq=70, freq_lo=0.05, freq_hi=6.0, freq_log_scale=True, window_type=2), 2)):
        g = ucuda.run_dft(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 12.0, strict=True, pre_scaller=pre, **kw)
        job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 12.0, pre_scaller=pre)
        n = ucuda.n_cells(axes)
        r = oracle.dft_hsweep_cells(step, job, range(n), kw["n_freq"], kw["freq_lo"], kw["freq_hi"],
                                    log_axis=kw.get("freq_log_scale", False), window_type=kw["window_type"], threads=8)
        np.testing.assert_array_equal(g["flags"].cpu().numpy(), r["flags"])
        assert (r["flags"] == 1).sum() >= n // 2
        np.testing.assert_array_equal(g["ak"].cpu().numpy(), r["ak"])
        np.testing.assert_array_equal(g["bk"].cpu().numpy(), r["bk"])
    g3 = ucuda.run_dft(s, [ucuda.axis(ucuda.AXIS_STEP, 0, 40, 0.002, 0.06)], [0.1, 0.1, 0.1], vals, 0.01, 5.0, 12.0,
                       precision=ucuda.PREC_F32C, n_freq=48, freq_lo=0.0, freq_hi=4.0)
    print("F32C h sweep flags equal on %.2f of the cells" % (g3["flags"].cpu().numpy() == 1).mean())


def _dbscan_cases(rng):
    """Peak sets for the clustering kernels: sizes around the 64-lane words
    and the 1024-peak wave / lane boundary, uniform values, levels on a grid
    of eps (distances at the eps boundary in floating point), tight chains,
    repeated values, period tiles, and the -1 / 0 / 1 counts."""
    sizes = [2, 3, 5, 63, 64, 65, 66, 127, 128, 129, 200, 256, 300, 500, 700, 960, 1023, 1024, 1025, 1026, 1300]
    cases = []
    for n in sizes:
        cases.append(rng.uniform(0.0, 3.0, n))
        cases.append(rng.integers(0, 9, n) * 0.1 + 2.0)
        cases.append(np.cumsum(rng.uniform(0.0, 0.11, n)))
        cases.append(np.round(rng.normal(5.0, 0.2, n), 2))
        cases.append(np.tile(3.0 + 0.5 * np.arange(1 + n % 7), n)[:n])
    return cases


@pytest.mark.parametrize("eps", [0.1, 1e-160, 1e160])
def test_dbscan_periods_stored_peaks(gpu_engine, eps):
    """eps 0.1: the wave kernel's |dp| < eps test; 1e-160 and 1e160 lie
    outside [2^-500, 2^500], where it keeps sqrt(dp^2) < eps (the peak sets
    then also hold values near 1e200, whose squares overflow)."""
    print("case: DBSCAN periods of stored peak sets (wave kernel n <= 1024, lane kernel above) vs oracle, eps %g"
          % eps, flush=True)
    import torch
    s = ucuda.preset("lorenz", gpu_engine)
    rng = np.random.default_rng(20261203)
    cases = _dbscan_cases(rng)
    if eps != 0.1:
        cases += [np.concatenate([c[:len(c) // 2], 1e200 * c[len(c) // 2:]]) for c in cases[::5]]
        cases += [c * 1e-162 for c in cases[1::5]]
    max_peaks = max(len(c) for c in cases)
    n_cells = len(cases) + 3
    peaks = np.full((max_peaks, n_cells), np.nan)
    counts = np.zeros(n_cells, dtype=np.int32)
    for c, v in enumerate(cases):
        peaks[:len(v), c] = v
        counts[c] = len(v)
    counts[-3:] = [-1, 0, 1]
    peaks[0, -1] = 4.2
    expect = np.array([oracle.dbscan_count(peaks[:counts[c], c], eps) if counts[c] > 1 else counts[c]
                       for c in range(n_cells)], dtype=np.int32)
    assert len(set(expect.tolist())) > (20 if eps == 0.1 else 3)
    for mp in (max_peaks, 1024):  # all cells; then only the wave kernel (longer counts read as 1024)
        g = ucuda.dbscan_periods(s, torch.from_numpy(np.ascontiguousarray(peaks[:mp])).cuda(),
                                 torch.from_numpy(counts).cuda(), eps).cpu().numpy()
        exp = expect if mp == max_peaks else np.array(
            [oracle.dbscan_count(peaks[:min(counts[c], mp), c], eps) if counts[c] > 1 else counts[c]
             for c in range(n_cells)], dtype=np.int32)
        np.testing.assert_array_equal(g, exp)


def test_dft_sweep(gpu_engine):
    print("case: 1-D DFT sweep (Lorenz 16 rho, 2000-sample blocks, 32 freqs, Hanning), strict vs oracle", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.5, 250.0)]
    vals = s.values()
    kw = dict(n_freq=32, freq_lo=0.0, freq_hi=5.0, window_type=1)
    g = ucuda.run_dft(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0, strict=True, **kw)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0)
    r = oracle.dft_cells(step, job, range(16), 32, 0.0, 5.0, window_type=1, threads=8)
    np.testing.assert_array_equal(g["flags"].cpu().numpy(), r["flags"])
    assert (r["flags"] == 1).sum() >= 8 and (r["flags"] == -1).any()  # rho < 1: fixed point
    np.testing.assert_array_equal(g["ak"].cpu().numpy(), r["ak"])
    np.testing.assert_array_equal(g["bk"].cpu().numpy(), r["bk"])
    # log frequency axis, Hamming, sampled combination x0 + pi x1 + e x2, pre_scaller 2
    kw2 = dict(n_freq=24, freq_lo=0.05, freq_hi=8.0, freq_log_scale=True, window_type=2, writable_var=-1,
               pre_scaller=2)
    g2 = ucuda.run_dft(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 30.0, strict=True, **kw2)
    job2 = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 30.0, writable_var=-1, pre_scaller=2)
    r2 = oracle.dft_cells(step, job2, range(16), 24, 0.05, 8.0, log_axis=True, window_type=2, threads=8)
    np.testing.assert_array_equal(g2["ak"].cpu().numpy(), r2["ak"])
    np.testing.assert_array_equal(g2["bk"].cpu().numpy(), r2["bk"])
    # FP32-compensated: the spectrum's dominant line matches
    g3 = ucuda.run_dft(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0, precision=ucuda.PREC_F32C, **kw)
    ok = r["flags"] == 1
    amp = np.hypot(r["ak"], r["bk"])[ok]
    amp3 = np.hypot(g3["ak"].cpu().numpy(), g3["bk"].cpu().numpy())[ok]
    print("F32C: max |amplitude error| / max amplitude = %.1e" % (np.abs(amp3 - amp).max() / amp.max()))


def test_cell_ranges_equal_the_whole_grid(gpu_engine):
    """uc_common.cell_begin / cell_count: a map run in pieces gives exactly
    the whole run's values (strict FP64, and fast F32C through the same
    kernels), bifurcation and LLE, for pieces that split rows and pairs."""
    print("case: Rossler LLE 40x30 and Lorenz bif 50 rho in cell ranges vs whole", flush=True)
    ros = ucuda.preset("rossler", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 40, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 30, 1.0, 20.0)]
    for kw in (dict(strict=True), dict(precision=ucuda.PREC_F32C)):
        args = (ros, axes, [1.0, 1.0, 1.0], ros.values(), 0.01, 2.0, 20.0, 0.1)
        whole = ucuda.run_lyap(*args, **kw)
        parts = [ucuda.run_lyap(*args, cells=(b, c), **kw) for b, c in ((0, 7), (7, 500), (507, 693))]
        v = np.concatenate([p["values"].cpu().numpy()[0] for p in parts])
        f = np.concatenate([p["flags"].cpu().numpy() for p in parts])
        np.testing.assert_array_equal(f, whole["flags"].cpu().numpy())
        np.testing.assert_array_equal(v, whole["values"].cpu().numpy()[0])
    lor = ucuda.preset("lorenz", gpu_engine)
    bax = [ucuda.axis(ucuda.AXIS_PARAM, 2, 50, 0.0, 250.0)]
    bargs = (lor, bax, [0.1, 0.1, 0.1], lor.values(), 0.01, 10.0, 10.0)
    whole = ucuda.run_bif(*bargs, max_peaks=16, strict=True)
    parts = [ucuda.run_bif(*bargs, max_peaks=16, strict=True, cells=(b, c)) for b, c in ((0, 13), (13, 37))]
    np.testing.assert_array_equal(np.concatenate([p["flags"].cpu().numpy() for p in parts]),
                                  whole["flags"].cpu().numpy())
    pk = np.concatenate([p["peaks"].cpu().numpy() for p in parts], axis=1)
    wf = whole["flags"].cpu().numpy()
    wp = whole["peaks"].cpu().numpy()
    for c in range(50):
        if wf[c] > 0:
            np.testing.assert_array_equal(pk[:min(wf[c], 16), c], wp[:min(wf[c], 16), c])
    # beyond the grid, and on an entry point that runs whole grids only
    with pytest.raises(ucuda.UcError, match="beyond the grid"):
        ucuda.run_lyap(ros, axes, [1.0, 1.0, 1.0], ros.values(), 0.01, 2.0, 20.0, 0.1, cells=(1000, 201))


@pytest.mark.parametrize("t_max,wvar", [(10.01, 0), (10.02, 2)])
def test_bif_record_odd_lengths_strict(gpu_engine, t_max, wvar):
    """The recorded loop takes two samples per trip and ends a chunk on a
    single one when its length is odd: records of 1,001 / 1,002 samples
    after a 57-step transient, the x or z component, a cap of 8 peaks;
    strict against the oracle bit for bit (flags, peaks, intervals)."""
    print("case: Lorenz bif 64 rho, record t_max %g (var %d), transient 0.57, strict vs oracle" % (t_max, wvar),
          flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 64, 0.0, 250.0)]
    vals = s.values()
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    g = ucuda.run_bif(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.57, t_max, writable_var=wvar, max_peaks=8,
                      strict=True)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 0.57, t_max, writable_var=wvar, max_peaks=8)
    r = oracle.bif_cells(step, job, range(64), threads=8)
    gf, gp, gi = g["flags"].cpu().numpy(), g["peaks"].cpu().numpy(), g["intervals"].cpu().numpy()
    np.testing.assert_array_equal(gf, r["flags"])
    assert (gf == 8).any() and (gf >= 1).sum() > 32  # cells at the cap and many with peaks
    for c in range(64):
        k = min(int(gf[c]), 8)
        if k > 0:
            np.testing.assert_array_equal(gp[:k, c], r["peaks"][:k, c])
            np.testing.assert_array_equal(gi[:k, c], r["intervals"][:k, c])
