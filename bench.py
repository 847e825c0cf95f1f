"""Headline benchmark: Rossler largest-Lyapunov-exponent map on one MI355X.

Workload (512 x 512 cells): dx=-y-z, dy=x+a*y, dz=b+z*(x-c), a=0.2,
b in [0.05, 2] x c in [1, 20], X0=(1,1,1), RK4 h=0.01, 5e3 transient steps,
t_max=500 averaged with renormalisation every NT=0.1 (10 steps + the step
loop's trailing step per block, as the reference schedules it), tangent
basis seeded by the reference generator with subsequence = cell index.

One "step" = one whole LLE map. Both precision policies run: FP64 and
FP32-compensated (float RK stages, Kahan state, FP64 renormalisation and
log sums). `value` is the FP32-compensated throughput in trajectory-steps/s
(accepted base-state RK4 steps per second over the whole grid, inputs
resident in HBM). The FP64 CPU oracle (the reference's finite-difference
Benettin method) runs on a fixed subsample of cells on this host, both as
the CPU baseline and as the reference for the LLE error of each GPU mode.
"""
import argparse
import json
import os
import sys
import time

# The serving measurement keeps six rounds in flight on six HIP streams.
# HIP maps a process's streams onto at most GPU_MAX_HW_QUEUES hardware queues
# (default 4, some taken by other streams), and rounds sharing a queue run
# one after another; eight queues give each round stream its own. Must be
# set before the HIP runtime starts; a value set by the caller wins.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

_T0 = time.perf_counter()


def progress(msg):
    print("[bench %6.1fs] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "u-cuda_amd"))

METRIC = "RK4 trajectory-steps/s, FP32+comp vs FP64, at LLE err≤1e-3; p99 job latency"
PEAK_FP32_VALU = 157.3e12   # MI355X published FP32 vector peak
PEAK_FP64_VALU = 78.6e12    # MI355X published FP64 vector peak


def valu_roofline(kernel, flops, ms, fp64, bound="valu", note=None):
    """Roofline entry of an analysis: algorithmic flops of the launch(es)
    (+, -, *, / = 1, FMA = 2; transcendental calls, compares and integer
    work not counted) over the measured time, against the MI355X vector
    peak of the precision the stages compute in."""
    peak = PEAK_FP64_VALU if fp64 else PEAK_FP32_VALU
    achieved = flops / (ms * 1e-3)
    d = dict(bound=bound, kernel=kernel, achieved=achieved / 1e12, peak=peak / 1e12, unit="TFLOP/s",
             frac=achieved / peak, flops_per_launch=float(flops), ms=ms)
    if note:
        d["note"] = note
    return d


def dp5_flops(ops, n):
    """One Dormand-Prince 5(4) attempt: 6 new right-hand sides (FSAL), the
    stage sums (1 + 2 + 3 + 4 + 5 terms), the 5th-order solution (5 terms)
    and the error estimate (6 terms) as FMAs per component, and the scaled
    RMS norm (5 flop per component)."""
    return 6 * ops["rhs_flops"] + (2 * (15 + 5 + 6) + 5) * n


def workload(n):
    import ucuda
    return dict(axes=[ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, n, 1.0, 20.0)],
                ic=[1.0, 1.0, 1.0], h=0.01, transient=50.0, t_max=500.0, NT=0.1)


def steps_per_cell(w):
    skip = int(w["transient"] / w["h"])
    blocks = int(w["t_max"] / w["NT"])
    nts = int(w["NT"] / w["h"])
    return skip + 1 + blocks * (nts + 1)


def tangent_steps_per_cell(w):
    """Steps that also advance the tangent vector: the renormalisation
    blocks. The transient (and its trailing step) advances the state only."""
    return int(w["t_max"] / w["NT"]) * (int(w["NT"] / w["h"]) + 1)


def headline_kernel(cells):
    """Kernel the engine launches for the FP32C map: the lean loop with two
    cells per lane (uc_lle_f32c_x2) when that gives 1.5..2 lane-waves per
    SIMD on the 256 CUs of an MI355X (the engine's measured rule), else with
    one (uc_lle_f32c_x1)."""
    pw = ((cells + 1) // 2) / (256 * 4 * 64)
    return "uc_lle_f32c_x2" if 1.5 < pw <= 2.0 else "uc_lle_f32c_x1"


def profiled_traffic():
    """HBM bytes per launch of the headline kernel as measured by the PMC
    passes recorded in profiles/ (FETCH_SIZE and WRITE_SIZE, each scaled by
    its reading for a plain copy of known size); None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", "lle_map_rossler_512_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def run_gpu(sysm, w, precision, steps, warmup, strict=False):
    import torch
    import ucuda
    n = ucuda.n_cells(w["axes"])
    dev = torch.device("cuda", 0)
    out = dict(values=torch.empty((1, n), dtype=torch.float64, device=dev),
               flags=torch.empty(n, dtype=torch.int32, device=dev), final_state=None)
    args = (sysm, w["axes"], w["ic"], sysm.values(a=0.2), w["h"], w["transient"], w["t_max"], w["NT"])
    for _ in range(warmup):
        ucuda.run_lyap(*args, precision=precision, strict=strict, out=out)
    stream = torch.cuda.current_stream()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ev in events:
        ev[0].record(stream)
        ucuda.run_lyap(*args, precision=precision, strict=strict, out=out, stream=stream, sync=False)
        ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]
    return dict(wall=wall, kernel_ms=kernel_ms, values=out["values"][0].cpu().numpy(),
                flags=out["flags"].cpu().numpy())


def cpu_baseline(eng, sysm, w, n, stride, threads_all):
    """Times the reference CPU path (finite-difference Benettin, FP64, compiled
    step body) single-threaded and on all cores over a fixed subsample, then
    computes on the same cells the FP64 variational oracle and its spread
    under rounding-sized state noise and parameter perturbations, for FP64
    (1e-15) and FP32 (~1e-7), which separates reproducible cells from
    sensitive (chaotic) ones."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    step = oracle.KrsStep(eng.codegen_scheme(sysm.vars, sysm.params, sysm.rhs), 3, 4)
    vals = sysm.values(a=0.2)
    job = oracle.make_job(3, w["axes"], w["ic"], vals, w["h"], w["transient"], w["t_max"], NT=w["NT"])
    idx = np.arange(0, n, stride)
    cells = (idx[:, None] * n + idx[None, :]).ravel()  # iy * n + ix, every stride-th point per axis
    progress("cpu baseline: %d cells, single thread" % len(cells))
    t0 = time.perf_counter()
    r1 = oracle.lle_fd_cells(step, job, cells, threads=1)
    t1 = time.perf_counter() - t0
    progress("cpu baseline: single thread %.1fs; now %d threads" % (t1, threads_all))
    t0 = time.perf_counter()
    rN = oracle.lle_fd_cells(step, job, cells, threads=threads_all)
    tN = time.perf_counter() - t0
    assert (r1["values"] == rN["values"]).all()
    progress("cpu baseline: %d threads %.1fs" % (threads_all, tN))
    funcs = oracle.SystemFuncs(sysm.source(1))
    var = oracle.lyap_var_cells(funcs, 1, job, cells, threads=threads_all)["values"][0]
    spreads = {p: oracle.lyap_rounding_spread(funcs, 1, job, cells, var, p, threads=threads_all)[0]
               for p in ("f64", "f32")}
    progress("oracle variational + rounding spreads done")
    return dict(cells=cells, fd=r1["values"][0], var=var, spreads=spreads, t1=t1, tN=tN)


def lle_errors(g, cpu, rounding):
    """Error of a GPU LLE map against the FP64 oracle on the sampled cells.
    Reproducible cells: the oracle's exponent moves by <= 1e-5 under state
    noise and parameter perturbations of the `rounding` precision's size."""
    import numpy as np
    g = g[cpu["cells"]]
    var, spread, fd = cpu["var"], cpu["spreads"][rounding], cpu["fd"]
    ok = var != 999.0
    repro = ok & (spread <= 1e-5)
    sens = ok & ~repro
    e = np.abs(g - var)
    efd = np.abs(g - fd)
    return dict(reproducible_cells=int(repro.sum()), sensitive_cells=int(sens.sum()),
                max_err_reproducible=float(e[repro].max()) if repro.any() else None,
                ok_le_1e3=bool(repro.any() and e[repro].max() <= 1e-3),
                frac_all_le_1e3=float((e[ok] <= 1e-3).mean()),
                median_err_sensitive=float(np.median(e[sens])) if sens.any() else None,
                median_fp64_spread_sensitive=float(np.median(spread[sens])) if sens.any() else None,
                vs_reference_fd_eps1e4=dict(max=float(efd[ok].max()), median=float(np.median(efd[ok]))),
                flag_mismatch=int(((g == 999.0) != ~ok).sum()))


def strict_check(g, cpu):
    """Strict FP64 (no FMA contraction, reference operation order) against the
    oracle on the sampled cells: expected bit-identical on every cell,
    chaotic ones included."""
    import numpy as np
    g = g[cpu["cells"]]
    return dict(cells=int(len(g)), bit_identical=int((g == cpu["var"]).sum()),
                max_abs_err=float(np.abs(g - cpu["var"]).max()))


def extra_bif_config1(eng, threads_all):
    """The reference's CPU-runnable case, whole: Lorenz bifurcation sweep,
    sigma = 10, beta = 8/3, 1024 rho values on [0, 250], X0 = (0.1, 0.1, 0.1),
    RK4 h = 0.01, 1e4 transient + 1e4 recorded steps, maxima of x, <= 256
    peaks. GPU time from HIP events (strict FP64, fast FP64, FP32C), the
    FP64 oracle (the reference's CPU algorithm) timed single-threaded on the
    full sweep, and parity: strict must equal the oracle bit for bit; fast
    FP64 and FP32C peaks are compared on the cells whose oracle peaks
    survive rounding-sized noise of that precision."""
    import numpy as np
    import torch
    import ucuda
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    lor = ucuda.preset("lorenz", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 1024, 0.0, 250.0)]
    args = (lor, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0)
    n = 1024
    traj_steps = n * (10000 + 1 + 10000 + 1)
    out, gpu = {}, {}
    stream = torch.cuda.current_stream()
    for name, kw in (("f64_strict", dict(strict=True)), ("f64", {}), ("f32c", dict(precision=ucuda.PREC_F32C)),
                     ("f32c_float_record", dict(precision=ucuda.PREC_F32C, record_precision=ucuda.PREC_F32C))):
        ucuda.run_bif(*args, max_peaks=256, **kw)  # compile + warm-up
        ms = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            r = ucuda.run_bif(*args, max_peaks=256, stream=stream, sync=False, **kw)
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        gpu[name] = {k: v.cpu().numpy() for k, v in r.items() if v is not None}
        out[name] = dict(ms=float(np.mean(ms)), traj_steps_per_s=traj_steps / (np.mean(ms) / 1e3))
    step = oracle.KrsStep(eng.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0, max_peaks=256)
    t0 = time.perf_counter()
    ref = oracle.bif_cells(step, job, range(n), threads=1)
    cpu_s = time.perf_counter() - t0
    out["cpu_oracle"] = dict(s=cpu_s, traj_steps_per_s=traj_steps / cpu_s, cores=1)
    sf = ucuda.rk4_flops(lor.op_counts(), 3)
    out["roofline"] = {
        name: valu_roofline("uc_bif_" + name, traj_steps * sf, out[name]["ms"], name == "f64", bound="latency",
                            note="1,024 cells = 16 waves on 1,024 SIMDs: each lane's dependent step chain sets the "
                                 "time; f32c: transient in FP32, recorded window in FP64; peak tests not counted")
        for name in ("f64", "f32c")}
    g = gpu["f64_strict"]
    same = np.array_equal(g["flags"], ref["flags"])
    for c in range(n):
        m = ref["flags"][c]
        if m > 0:
            same &= np.array_equal(g["peaks"][:m, c], ref["peaks"][:m, c])
            same &= np.array_equal(g["intervals"][:m, c], ref["intervals"][:m, c])
    out["f64_strict"]["identical_to_oracle"] = bool(same)
    # checked cells: oracle peaks that hold `rel` under the rounding each
    # mode injects (f32_rec64: float-sized noise and parameter perturbations
    # in the transient, double-sized noise in the double recorded window)
    for name, prec, rel in (("f64", "f64", 1e-8), ("f32c", "f32_rec64", 5e-7), ("f32c_float_record", "f32", 1e-6)):
        stable = oracle.bif_rounding_stable(step, job, range(n), ref, prec, rel, threads=threads_all)
        g = gpu[name]
        errs = []
        for c in np.nonzero(stable)[0]:
            m = ref["flags"][c]
            errs.append(np.inf if g["flags"][c] != m else
                        float(np.abs(g["peaks"][:m, c] - ref["peaks"][:m, c]).max() / np.abs(ref["peaks"][:m, c]).max()))
        out[name].update(stable_cells=int(stable.sum()), stable_model=prec, stable_rel=rel,
                         max_rel_peak_err_stable=max(errs) if errs else None,
                         ok_le_1e6=bool(errs) and max(errs) <= 1e-6,
                         flag_class_agreement=float(((g["flags"] <= 0) == (ref["flags"] <= 0)).mean()))
    out["f32c"]["record"] = "transient FP32C, recorded window FP64 from the compensated state"
    return out


def _timed_calls(fn, reps=3):
    """Wall time of whole calls (host set-up, temporary buffers, every
    kernel), synchronised on both sides; returns (ms list, last result)."""
    import torch
    fn()  # compile + warm-up
    torch.cuda.synchronize()
    ms, r = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    return ms, r


def extra_bif2d(eng, n=512, stride=8):
    """2-D bifurcation period map: Lorenz rho in [20, 250] x sigma in [5, 16]
    (n x n), beta = 8/3, X0 = (0.1, 0.1, 0.1), RK4 h = 0.01, 1e4 transient +
    1e4 recorded steps, maxima of x clustered by DBSCAN with eps 0.1 (the
    reference's default eps). GPU: whole-call wall time (sweep + per-cell
    DBSCAN). Parity: strict FP64 against the oracle on every stride-th point
    per axis, whose single-thread time is the CPU rate."""
    import numpy as np
    import ucuda
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    lor = ucuda.preset("lorenz", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n, 20.0, 250.0), ucuda.axis(ucuda.AXIS_PARAM, 1, n, 5.0, 16.0)]
    args = (lor, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0)
    traj_steps = n * n * (10000 + 1 + 10000 + 1)
    out, maps = {}, {}
    for name, kw in (("f64_strict", dict(strict=True)), ("f64", {}), ("f32c", dict(precision=ucuda.PREC_F32C)),
                     ("f32c_float_record", dict(precision=ucuda.PREC_F32C, record_precision=ucuda.PREC_F32C))):
        ms, r = _timed_calls(lambda: ucuda.run_bif_periods(*args, eps_dbscan=0.1, **kw))
        maps[name] = r.cpu().numpy()
        out[name] = dict(ms=float(np.mean(ms)), traj_steps_per_s=traj_steps / (np.mean(ms) / 1e3))
        progress("bif2d %s: %.1f ms per map" % (name, np.mean(ms)))
    sf = ucuda.rk4_flops(lor.op_counts(), 3)
    out["roofline"] = {
        name: valu_roofline("uc_bif_%s + uc_bif_dbscan_wave" % name, traj_steps * sf, out[name]["ms"], name == "f64",
                            note="whole call: sweep with streaming peaks, then per-cell DBSCAN (its work not "
                                 "counted); f32c: transient in FP32, recorded window in FP64")
        for name in ("f64", "f32c")}
    idx = np.arange(0, n, stride)
    cells = (idx[:, None] * n + idx[None, :]).ravel()
    step = oracle.KrsStep(eng.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0)
    t0 = time.perf_counter()
    ref = oracle.bif_periods(step, job, cells, 0.1, threads=1)
    cpu_s = time.perf_counter() - t0
    sample_steps = len(cells) * (10000 + 1 + 10000 + 1)
    out["cpu_oracle"] = dict(cells=int(len(cells)), s=cpu_s, traj_steps_per_s=sample_steps / cpu_s, cores=1)
    out["f64_strict"]["identical_to_oracle"] = bool(np.array_equal(maps["f64_strict"][cells], ref))
    # cells whose first 256 peaks survive rounding-sized noise of the mode (the
    # rest are chaotic: any rounding difference changes their peak set)
    threads_all = min(16, os.cpu_count() or 1)
    jobp = oracle.make_job(3, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0, max_peaks=256)
    base = oracle.bif_cells(step, jobp, cells, threads=threads_all)
    for name, prec in (("f64", "f64"), ("f32c", "f32")):
        st = oracle.bif_rounding_stable(step, jobp, cells, base, prec, 1e-6, threads=threads_all)
        out[name]["period_agreement_vs_oracle"] = float((maps[name][cells] == ref).mean())
        out[name]["stable_cells"] = int(st.sum())
        out[name]["period_agreement_on_stable_cells"] = float((maps[name][cells][st] == ref[st]).mean()) if st.any() else None
    out["periods_histogram"] = {int(k): int(v) for k, v in zip(*np.unique(maps["f64_strict"], return_counts=True))
                                if k <= 8 or v > n * n // 100}
    return out


def extra_dft(eng, n=16384, n_freq=200, stride=64):
    """1-D DFT sweep: Lorenz rho in [0.5, 250] (n points), X0 = (0.1, 0.1,
    0.1), RK4 h = 0.01, 1e4 transient steps, blocks of 1e4 samples of x,
    Hanning window, n_freq frequencies on [0, 10] (the reference's default
    frequency axis). GPU: whole-call wall time (sweep recording the blocks,
    then the DFT kernel). Parity: strict FP64 coefficients against the oracle
    on every stride-th point, whose single-thread time is the CPU rate."""
    import numpy as np
    import ucuda
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    lor = ucuda.preset("lorenz", eng)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.5, 250.0)]
    args = (lor, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0)
    kw = dict(n_freq=n_freq, freq_lo=0.0, freq_hi=10.0, window_type=1)
    N = 10000
    out, res = {}, {}
    for name, extra in (("f64_strict", dict(strict=True)), ("f64", {}), ("f32c", dict(precision=ucuda.PREC_F32C))):
        ms, r = _timed_calls(lambda: ucuda.run_dft(*args, **kw, **extra))
        res[name] = {k: v.cpu().numpy() for k, v in r.items()}
        out[name] = dict(ms=float(np.mean(ms)), cells_per_s=n / (np.mean(ms) / 1e3),
                         dft_terms_per_s=n * n_freq * N / (np.mean(ms) / 1e3))
        progress("dft %s: %.1f ms per sweep" % (name, np.mean(ms)))
    # recording: 1e4 + 1 transient and 1e4 + 1 block steps per cell; DFT:
    # 11 flop per (cell, frequency, sample)
    flops = n * (2 * (N + 1)) * ucuda.rk4_flops(lor.op_counts(), 3) + 11.0 * n * n_freq * N
    out["roofline"] = {
        name: valu_roofline("uc_bif_%s (raw record) + uc_dft" % name, flops, out[name]["ms"], True,
                            note="whole call; the DFT runs in FP64 for both modes, so the FP64 peak")
        for name in ("f64", "f32c")}
    cells = np.arange(0, n, stride)
    step = oracle.KrsStep(eng.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0)
    t0 = time.perf_counter()
    ref = oracle.dft_cells(step, job, cells, n_freq, 0.0, 10.0, window_type=1, threads=1)
    cpu_s = time.perf_counter() - t0
    out["cpu_oracle"] = dict(cells=int(len(cells)), s=cpu_s, cells_per_s=len(cells) / cpu_s, cores=1)
    g = res["f64_strict"]
    out["f64_strict"]["identical_to_oracle"] = bool(
        np.array_equal(g["flags"][cells], ref["flags"]) and np.array_equal(g["ak"][cells], ref["ak"])
        and np.array_equal(g["bk"][cells], ref["bk"]))
    ok = ref["flags"] == 1
    amp = np.hypot(ref["ak"], ref["bk"])
    threads_all = min(16, os.cpu_count() or 1)
    jobp = oracle.make_job(3, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 100.0, 100.0, max_peaks=256)
    base = oracle.bif_cells(step, jobp, cells, threads=threads_all)
    for name, prec in (("f64", "f64"), ("f32c", "f32")):
        a = np.hypot(res[name]["ak"][cells], res[name]["bk"][cells])
        err = np.abs(a - amp).max(axis=1) / np.maximum(amp.max(axis=1), 1e-300)
        st = ok & oracle.bif_rounding_stable(step, jobp, cells, base, prec, 1e-6, threads=threads_all)
        out[name]["flag_agreement"] = float((res[name]["flags"][cells] == ref["flags"]).mean())
        out[name]["median_amp_err_rel_max"] = float(np.median(err[ok]))
        out[name]["stable_cells"] = int(st.sum())
        out[name]["max_amp_err_rel_on_stable_cells"] = float(err[st].max()) if st.any() else None
    out["config"] = dict(cells=n, n_freq=n_freq, samples_per_block=N, window="hanning")
    return out


def extra_fastsync(eng, n=1024, stride=64):
    """Fast synchronisation, Lorenz (10, 28, 8/3), RK4 h = 0.01, coupling
    k_forward = (1, 0, 0), k_backward = (15, 15, 15), 10 forward/backward
    passes, unidirectional, error at the window's end (estimator 2):
    * on grid: n x n master initial states x in [-15, 15], z in [5, 40],
      slave (3, -2, 20) after 1 s uncoupled, windows of 500 samples ->
      500 + 10 * 2 * 499 coupled steps per cell;
    * on attractor: 10,000 windows of 500 samples along a master trajectory
      of 100 s (the reference's default t_max) after a 2 s transient
      (pre_scaller 1); one lane per window, so 157 waves: latency-bound.
    GPU: whole-call wall time. Parity: strict FP64 against the oracle on a
    sample (every stride-th point per axis / every 100th window, 1e-15
    relative: the states are bit-identical, log10 may differ in the last
    bit); the oracle's single-thread time on that sample is the CPU rate."""
    import numpy as np
    import ucuda
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    lor = ucuda.preset("lorenz", eng)
    vals = lor.values()
    m0, s0, kf, kb = [0.1, 0.1, 0.1], [3.0, -2.0, 20.0], [1.0, 0.0, 0.0], [15.0, 15.0, 15.0]
    step = oracle.KrsStep(eng.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    out = {}
    gkw = dict(h=0.01, iter_of_synchr=10, pre_scaller=1, transient_time=1.0, window=5.0, axis_x_var=0, axis_y_var=2,
               axis_x=(-15.0, 15.0), axis_y=(5.0, 40.0), n_pts=n, error_estim=2)
    steps_cell = 499 + 10 * 2 * 499
    res = {}
    for name, kw in (("f64_strict", dict(strict=True)), ("f64", {})):
        ms, r = _timed_calls(lambda: ucuda.run_fastsync(lor, vals, m0, s0, kf, kb, mode=1, **gkw, **kw))
        res[name] = r["sync_error"].cpu().numpy()
        out["grid_" + name] = dict(ms=float(np.mean(ms)), coupled_steps_per_s=n * n * steps_cell / (np.mean(ms) / 1e3))
        progress("fastsync grid %s: %.1f ms" % (name, np.mean(ms)))
    # a coupled step: one RK4 of the slave plus the coupling term (3 N)
    out["roofline"] = dict(grid_f64=valu_roofline(
        "uc_fs_grid", n * n * steps_cell * (ucuda.rk4_flops(lor.op_counts(), 3) + 9), out["grid_f64"]["ms"], True,
        note="whole call; the backward passes read the stored master window (8 N B per step) and the fast "
             "module recomputes the forward master (not counted)"))
    idx = np.arange(0, n, stride)
    cells = (idx[:, None] * n + idx[None, :]).ravel()
    t0 = time.perf_counter()
    ref = oracle.fs_grid(step, 3, vals, m0, s0, kf, kb, cells, threads=1, **gkw)
    cpu_s = time.perf_counter() - t0
    out["grid_cpu_oracle"] = dict(cells=int(len(cells)), s=cpu_s, coupled_steps_per_s=len(cells) * steps_cell / cpu_s,
                                  cores=1)
    g = res["f64_strict"][cells]
    out["grid_f64_strict"]["max_rel_err_vs_oracle"] = float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-300)))
    out["grid_f64"]["median_abs_diff_vs_strict"] = float(np.median(np.abs(res["f64"] - res["f64_strict"])))
    akw = dict(h=0.01, iter_of_synchr=10, pre_scaller=1, t_max=100.0, transient_time=2.0, window=5.0, error_estim=2)
    ms, r = _timed_calls(lambda: ucuda.run_fastsync(lor, vals, m0, s0, kf, kb, mode=0, strict=True, **akw))
    a = r["sync_error"].cpu().numpy()
    out["attractor_f64_strict"] = dict(ms=float(np.mean(ms)), windows=int(len(a)),
                                       coupled_steps_per_s=len(a) * steps_cell / (np.mean(ms) / 1e3))
    sel = np.arange(0, len(a), 100)
    t0 = time.perf_counter()
    _, ra, _ = oracle.fs_attractor(step, 3, vals, m0, s0, kf, kb, sel, threads=1, **akw)
    out["attractor_cpu_oracle"] = dict(windows=int(len(sel)), s=time.perf_counter() - t0, cores=1)
    out["attractor_f64_strict"]["max_rel_err_vs_oracle"] = float(
        np.max(np.abs(a[sel] - ra) / np.maximum(np.abs(ra), 1e-300)))
    progress("fastsync on attractor: %.1f ms" % np.mean(ms))
    return out


def extra_continuation(eng, n_pts=128, n_chains=1024):
    """Continuation LLE chains, Lorenz (sigma, 8/3), rho walked upwards over
    [10, 200] in n_pts points, X0 = (0.1, 0.1, 0.1), RK4 h = 0.01, transient
    5 (first point) then (int)(5 / NT) settle blocks per carried point,
    t_max 20, NT = 0.1, finite-difference probe eps 1e-4. One chain is the
    reference's case (one GPU thread there, one lane here: latency-bound);
    n_chains independent chains over sigma in [6, 14] is the GPU-shaped use.
    GPU: whole-call wall time, strict FP64; parity of the single chain against
    the oracle (1e-12: the device log may round the last bit differently);
    CPU: the oracle's single chain, one thread."""
    import numpy as np
    import ucuda
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    lor = ucuda.preset("lorenz", eng)
    vals = lor.values()
    one = [ucuda.axis(ucuda.AXIS_PARAM, 2, n_pts, 10.0, 200.0)]
    many = one + [ucuda.axis(ucuda.AXIS_PARAM, 1, n_chains, 6.0, 14.0)]
    args = ([0.1, 0.1, 0.1], vals, 0.01, 5.0, 20.0, 0.1)
    # state + probe steps per chain: 500 + 1 transient, then per point (settle + 200 blocks) x 11 x 2
    steps = 501 + n_pts * 200 * 11 * 2 + (n_pts - 1) * 50 * 11 * 2
    out = {}
    ms, r1 = _timed_calls(lambda: ucuda.run_lyap_cont(lor, one, *args, strict=True), reps=1)
    out["one_chain"] = dict(ms=float(np.mean(ms)), steps_per_s=steps / (np.mean(ms) / 1e3))
    progress("continuation: one chain %.1f ms" % np.mean(ms))
    ms, rm = _timed_calls(lambda: ucuda.run_lyap_cont(lor, many, *args, strict=True), reps=1)
    out["%d_chains" % n_chains] = dict(ms=float(np.mean(ms)), steps_per_s=n_chains * steps / (np.mean(ms) / 1e3))
    progress("continuation: %d chains %.1f ms" % (n_chains, np.mean(ms)))
    sf = ucuda.rk4_flops(lor.op_counts(), 3)
    out["roofline"] = {
        "one_chain": valu_roofline("uc_cont_lle", steps * sf, out["one_chain"]["ms"], True, bound="latency",
                                   note="one lane: a lone wave issues one instruction per ~4-5 cycles"),
        "%d_chains" % n_chains: valu_roofline("uc_cont_lle", n_chains * steps * sf,
                                              out["%d_chains" % n_chains]["ms"], True, bound="latency",
                                              note="1,024 chains = 16 waves on 1,024 SIMDs")}
    step = oracle.KrsStep(eng.codegen_scheme(lor.vars, lor.params, lor.rhs), 3, 4)
    job = oracle.make_job(3, one, [0.1, 0.1, 0.1], vals, 0.01, 5.0, 20.0, NT=0.1, eps=1e-4)
    t0 = time.perf_counter()
    ref = oracle.cont_lyap(step, job, 1, threads=1)
    cpu_s = time.perf_counter() - t0
    out["cpu_oracle_one_chain"] = dict(s=cpu_s, steps_per_s=steps / cpu_s, cores=1)
    g = r1["values"].cpu().numpy()[0]
    out["one_chain"]["max_abs_err_vs_oracle"] = float(np.nanmax(np.abs(g - ref["values"][0])))
    out["one_chain"]["flags_equal"] = bool(np.array_equal(r1["flags"].cpu().numpy(), ref["flags"]))
    return out


def extra_ls(eng, n=1024, steps_scale=1.0):
    """hyper-Rossler 4-exponent spectrum maps, h=0.005, Gram-Schmidt every 20
    steps, 1e5 steps, FP32-compensated and FP64:
      sweep: n x n over d in [0, 0.1] x a in [0.2, 0.3], random X0 in
        [-1, 1]^4 (seed 20261203). From there trajectories escape
        (z' = 3 + xz) before t = 20, so this measures little work.
      attractor: n x n initial conditions (x0, w0) in [-12, -8] x [8.1, 12.1]
        (y0 = -6, z0 = 0) at the hyperchaotic point (a, d) = (0.25, 0.05);
        ~2/3 of the cells stay bounded for the whole run. Its traj-steps/s
        counts the bounded cells' steps only (a lower bound on the work)."""
    import torch
    import ucuda
    s = ucuda.preset("hyper_rossler", eng)
    t_max = 500.0 * steps_scale
    blocks, nts = int(t_max / 0.1), int(0.1 / 0.005)
    spc = 1 + blocks * (nts + 1)
    cases = {
        "sweep": dict(axes=[ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.0, 0.1), ucuda.axis(ucuda.AXIS_PARAM, 1, n, 0.2, 0.3)],
                      ic=None, kw=dict(ic_seed=20261203, ic_lo=-1.0, ic_hi=1.0)),
        "attractor": dict(axes=[ucuda.axis(ucuda.AXIS_VAR, 0, n, -12.0, -8.0), ucuda.axis(ucuda.AXIS_VAR, 3, n, 8.1, 12.1)],
                          ic=[-10.0, -6.0, 0.0, 10.1], kw={}),
    }
    out = {}
    for prec in (ucuda.PREC_F32C, ucuda.PREC_F64):  # compile + first-launch costs outside the timings
        ucuda.run_lyap(s, [ucuda.axis(ucuda.AXIS_PARAM, 2, 2, 0.0, 0.1)], [-10.0, -6.0, 0.0, 10.1], s.values(), 0.005,
                       0.0, 0.1, 0.1, spectrum=True, precision=prec)
    for case, c in cases.items():
        for name, prec in (("f32c", ucuda.PREC_F32C), ("f64", ucuda.PREC_F64)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ucuda.run_lyap(s, c["axes"], c["ic"], s.values(), 0.005, 0.0, t_max, 0.1, spectrum=True,
                               precision=prec, **c["kw"])
            dt = time.perf_counter() - t0
            ok = int((r["flags"].cpu().numpy() == 1).sum())
            out["%s_%s" % (case, name)] = dict(s=dt, cells=n * n, steps_per_cell=spc, bounded_cells=ok,
                                               traj_steps_per_s_bounded=ok * spc / dt)
            progress("LS %s %s: %.2fs, %d bounded cells" % (case, name, dt, ok))
    a32, a64 = out["attractor_f32c"], out["attractor_f64"]
    out["attractor_f32c_over_f64"] = a64["s"] / a32["s"]
    # bounded cells only: one state-only step, then every step advances the
    # state and 4 tangent vectors (the Gram-Schmidt per block not counted)
    sf4 = ucuda.rk4_flops(s.op_counts(), 4, tangents=4)
    out["roofline"] = {
        name: valu_roofline("uc_ls_" + name, out["attractor_" + name]["bounded_cells"] * spc * sf4,
                            out["attractor_" + name]["s"] * 1e3, name == "f64",
                            note="attractor map; bounded cells' work only (a lower bound); whole call")
        for name in ("f32c", "f64")}
    # full-size parity and CPU baseline on a fixed subsample of the attractor
    # map: every 64th point per axis, the whole 1e5-step horizon
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    c = cases["attractor"]
    strict = ucuda.run_lyap(s, c["axes"], c["ic"], s.values(), 0.005, 0.0, t_max, 0.1, spectrum=True,
                            strict=True)["values"].cpu().numpy()
    f32 = ucuda.run_lyap(s, c["axes"], c["ic"], s.values(), 0.005, 0.0, t_max, 0.1, spectrum=True,
                         precision=ucuda.PREC_F32C)["values"].cpu().numpy()
    idx = np.arange(0, n, 64)
    cells = (idx[:, None] * n + idx[None, :]).ravel()
    funcs = oracle.SystemFuncs(s.source(1))
    job = oracle.make_job(4, c["axes"], c["ic"], s.values(), 0.005, 0.0, t_max, NT=0.1)
    t0 = time.perf_counter()
    r1 = oracle.lyap_var_cells(funcs, 4, job, cells[::4], threads=1)
    cpu_s = time.perf_counter() - t0
    bounded1 = int((r1["flags"] == 1).sum())
    ref = oracle.lyap_var_cells(funcs, 4, job, cells, threads=16)["values"]
    ok = ref[0] != 999.0
    spread32 = oracle.lyap_rounding_spread(funcs, 4, job, cells, ref, "f32", threads=16)
    progress("LS oracle: %d cells single-thread %.1fs, %d cells on 16 threads" % (len(cells) // 4, cpu_s, len(cells)))
    out["subsample"] = dict(cells=int(len(cells)), bounded=int(ok.sum()),
                            strict_f64_max_abs_err=float(np.abs(strict[:, cells][:, ok] - ref[:, ok]).max()),
                            strict_flags_equal=bool(((strict[0, cells] == 999.0) == ~ok).all()),
                            f32c_median_abs_err=[float(np.median(np.abs(f32[k, cells][ok] - ref[k, ok])))
                                                 for k in range(4)],
                            fp64_median_spread_under_f32_rounding=[float(np.median(spread32[k, ok])) for k in range(4)],
                            # the precision contract: 1e-3 where the oracle
                            # holds 1e-4 under float-sized rounding, within
                            # max(6 x its spread, 1e-3) elsewhere
                            f32c_reproducible_cells=int((ok & (spread32.max(axis=0) <= 1e-4)).sum()),
                            f32c_frac_within_max_6spread_1e3=[
                                float((np.abs(f32[k, cells][ok] - ref[k, ok]) <=
                                       np.maximum(6.0 * spread32[k, ok], 1e-3)).mean()) for k in range(4)],
                            cpu_oracle=dict(s=cpu_s, cells=int(len(cells) // 4), bounded=bounded1,
                                            traj_steps_per_s_bounded=bounded1 * spc / cpu_s, cores=1,
                                            kind="port", method="variational, 4 tangents"))
    return out


def extra_basins(eng, n=2048):
    """forced Duffing basins: n x n ICs over [-2.5, 2.5]^2, DP5(4) rtol=atol=1e-9,
    200 forcing periods, Poincare signatures -> int32 labels."""
    import math
    import numpy as np
    import torch
    import ucuda
    s = ucuda.preset("duffing", eng)
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, n, -2.5, 2.5), ucuda.axis(ucuda.AXIS_VAR, 1, n, -2.5, 2.5)]
    out = {}
    labels = {}
    for prec in (ucuda.PREC_F64, ucuda.PREC_F32C):  # compile + first-launch costs outside the timings
        ucuda.run_basins(s, [ucuda.axis(ucuda.AXIS_VAR, 0, 2, -1.0, 1.0)], [0.0, 0.0], s.values(), n_sections=16,
                         precision=prec, want_points=False)
    for name, prec in (("f64", ucuda.PREC_F64), ("f32c", ucuda.PREC_F32C)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = ucuda.run_basins(s, axes, [0.0, 0.0], s.values(), n_sections=200, precision=prec, want_points=False)
        dt = time.perf_counter() - t0
        st = r["steps"].cpu().numpy()
        labels[name] = (r["labels"].cpu().numpy(), r["attractors"])
        out[name] = dict(s=dt, cells=n * n, accepted_steps=int(st[0].sum()), rejected_steps=int(st[1].sum()),
                         traj_steps_per_s=float(st[0].sum()) / dt, attractors=r["attractors"])
    # float stages cannot estimate errors below 2^-20 of the state: the
    # FP32C integration runs at max(requested, 2^-20)
    out["f32c_tolerance"] = dict(requested_rtol=1e-9, requested_atol=1e-9, rtol=max(1e-9, 2.0 ** -20),
                                 atol=max(1e-9, 2.0 ** -20), rule="max(requested, 2^-20)")
    # FP64 at the tolerance FP32C actually runs, so that the two precisions
    # are also compared at equal accuracy (the 1e-9 FP64 map takes ~4x the
    # steps of either)
    tol = max(1e-9, 2.0 ** -20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = ucuda.run_basins(s, axes, [0.0, 0.0], s.values(), rtol=tol, atol=tol, n_sections=200,
                         precision=ucuda.PREC_F64, want_points=False)
    dt = time.perf_counter() - t0
    st = r["steps"].cpu().numpy()
    lab_eq = r["labels"].cpu().numpy()
    out["f64_at_f32c_tolerance"] = dict(s=dt, rtol=tol, atol=tol, accepted_steps=int(st[0].sum()),
                                        rejected_steps=int(st[1].sum()), traj_steps_per_s=float(st[0].sum()) / dt,
                                        attractors=r["attractors"])
    out["f32c_over_f64_at_equal_tolerance"] = dt / out["f32c"]["s"]
    df = dp5_flops(s.op_counts(), s.n_vars)
    out["roofline"] = {
        name: valu_roofline("uc_bas_" + name, (out[name]["accepted_steps"] + out[name]["rejected_steps"]) * df,
                            out[name]["s"] * 1e3, name == "f64", bound="valu+transcendental",
                            note="DP5(4) attempts x %d flop; the forcing cos in each right-hand side (6 per "
                                 "attempt), the step-factor log2/exp2 and the catalogue not counted; whole call"
                                 % df)
        for name in ("f64", "f32c")}
    # label agreement FP32C vs FP64 (catalogues mapped by period + point)
    l64, a64 = labels["f64"]
    l32, a32 = labels["f32c"]
    m = {}
    for a in a32:
        for b in a64:
            if a["period"] == b["period"] and (a["period"] == 0 or (abs(a["point"][0] - b["point"][0]) <= 3e-3 and
                                                                    abs(a["point"][1] - b["point"][1]) <= 3e-3)):
                m[a["label"]] = b["label"]
    l32m = np.array([m.get(int(v), 0 if v == 0 else -99) for v in l32])
    out["f32c_label_agreement"] = float((l32m == l64).mean())
    # the FP64 map at 2^-20 against the 1e-9 one (same catalogue numbering
    # when the attractors match by period and point)
    meq = {}
    for a in out["f64_at_f32c_tolerance"]["attractors"]:
        for b in a64:
            if a["period"] == b["period"] and (a["period"] == 0 or (abs(a["point"][0] - b["point"][0]) <= 3e-3 and
                                                                    abs(a["point"][1] - b["point"][1]) <= 3e-3)):
                meq[a["label"]] = b["label"]
    lut = np.full(int(lab_eq.max()) + 1, -99, dtype=np.int64)
    lut[0] = 0
    for k, v in meq.items():
        lut[k] = v
    leqm = lut[lab_eq]
    out["f64_at_f32c_tolerance"]["label_agreement_with_f64_1e9"] = float((leqm == l64).mean())
    # full-size parity and CPU baseline on every 32nd point per axis
    # (64 x 64 = 4096 cells) of the same 2048^2 grid; boundary cells are
    # those whose FP64 label map changes within 2 grid spacings
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    idx = np.arange(0, n, 32)
    cells = (idx[:, None] * n + idx[None, :]).ravel()
    funcs = oracle.SystemFuncs(s.source(1))
    bj = oracle.make_basins_job(2, axes, [0.0, 0.0], s.values(), 0.01, 1e-9, 1e-9, 2 * math.pi, 200)
    t0 = time.perf_counter()
    r1 = oracle.basins_cells(funcs, bj, cells[::4], threads=1)
    cpu_s = time.perf_counter() - t0
    ref = oracle.basins_cells(funcs, bj, cells, threads=16)
    progress("basins oracle: %d cells single-thread %.1fs, %d cells on 16 threads" % (len(cells) // 4, cpu_s, len(cells)))
    lab2 = l64.reshape(n, n)
    inner = np.ones_like(lab2, dtype=bool)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            sh = np.roll(np.roll(lab2, dy, axis=0), dx, axis=1)
            inner &= sh == lab2
    inner[:2, :] = inner[-2:, :] = False
    inner[:, :2] = inner[:, -2:] = False
    inner = inner.ravel()[cells]
    per_ref, pt_ref = ref["period"], ref["point"]
    lab_ref = np.zeros(len(cells), dtype=np.int64)
    for i in range(len(cells)):
        p = per_ref[i]
        if p < 0:
            continue
        lab_ref[i] = -99
        for a in labels["f64"][1]:
            if a["period"] == p and (p == 0 or (abs(pt_ref[0, i] - a["point"][0]) <= 3e-3 and
                                                abs(pt_ref[1, i] - a["point"][1]) <= 3e-3)):
                lab_ref[i] = a["label"]
                break
    lg = l64[cells]
    out["subsample"] = dict(cells=int(len(cells)), interior=int(inner.sum()),
                            f64_label_agreement_interior=float((lg[inner] == lab_ref[inner]).mean()),
                            f64_label_agreement_all=float((lg == lab_ref).mean()),
                            cpu_oracle=dict(s=cpu_s, cells=int(len(cells) // 4), cores=1, kind="port",
                                            accepted_steps_per_s=float(r1["steps"][0].sum()) / cpu_s))
    return out


def extra_batch(eng, n_jobs=64, seed=0):
    """64 heterogeneous jobs in one persistent hipGraph launch after the
    launch-shape tuner, systems uniform over Lorenz / Rossler / Chen /
    Duffing, grid side round(64*16^U), transient round(10^(3+2U)) steps, job
    seed 1000+j. Lorenz / Rossler / Chen jobs are LLE maps (2e3 averaged
    steps); Duffing jobs are forced-Duffing basin maps over [-2.5, 2.5]^2 at
    rtol = atol = 1e-9 with max(16, steps / 500) forcing periods, their
    attractor catalogues and labels included in the timed launch. All
    FP32-compensated."""
    import math
    import numpy as np
    import torch
    import ucuda
    rng = np.random.default_rng(seed)
    systems = {k: ucuda.preset(k, eng) for k in ("lorenz", "rossler", "chen", "duffing")}
    boxes = {"lorenz": (2, 20.0, 200.0, 3, 1.0, 4.0), "rossler": (2, 0.05, 2.0, 3, 1.0, 20.0),
             "chen": (1, 30.0, 45.0, 3, 20.0, 30.0)}
    b = ucuda.Batch(eng)
    lle_steps = 0
    lle_flops = 0.0
    basins_out = []
    kinds = {"lle": 0, "basins": 0}
    for j in range(n_jobs):
        name = list(systems)[int(rng.integers(0, 4))]
        sysm = systems[name]
        side = int(round(64 * 16 ** rng.uniform()))
        transient_steps = int(round(10 ** (3 + 2 * rng.uniform())))
        if name == "duffing":
            axes = [ucuda.axis(ucuda.AXIS_VAR, 0, side, -2.5, 2.5), ucuda.axis(ucuda.AXIS_VAR, 1, side, -2.5, 2.5)]
            basins_out.append(b.add_basins(sysm, axes, [0.0, 0.0], sysm.values(),
                                           n_sections=max(16, transient_steps // 500), section_period=2 * math.pi,
                                           precision=ucuda.PREC_F32C, want_points=False))
            kinds["basins"] += 1
            continue
        h = 0.0025 if name == "chen" else 0.01
        p1, lo1, hi1, p2, lo2, hi2 = boxes[name]
        axes = [ucuda.axis(ucuda.AXIS_PARAM, p1, side, lo1, hi1), ucuda.axis(ucuda.AXIS_PARAM, p2, side, lo2, hi2)]
        ic = [0.1] * sysm.n_vars
        t_max = 2000 * h
        b.add_lyap(sysm, axes, ic, sysm.values(), h, transient_steps * h, t_max, 10 * h,
                   precision=ucuda.PREC_F32C, probe_seed=1000 + j)
        lle_steps += side * side * (transient_steps + 1 + int(t_max / (10 * h)) * 11)
        ops = sysm.op_counts()
        lle_flops += side * side * ((transient_steps + 1) * ucuda.rk4_flops(ops, sysm.n_vars) +
                                    int(t_max / (10 * h)) * 11 * ucuda.rk4_flops(ops, sysm.n_vars, tangents=1))
        kinds["lle"] += 1
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        tuned = b.tune(stream)
        b.launch(stream)  # graph build + first replay
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.launch(stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    bas_acc = sum(int(o["steps"][0].sum()) for o in basins_out)
    bas_att = sum(int(o["steps"].sum()) for o in basins_out)
    bas_cells = sum(int(o["labels"].numel()) for o in basins_out)
    flops = lle_flops + bas_att * dp5_flops(systems["duffing"].op_counts(), systems["duffing"].n_vars)
    roofline = valu_roofline("uc_lle_f32c_*_q + uc_bas_f32c_q", flops, dt * 1e3, False,
                             note="LLE maps (state-only transient, state + tangent blocks) and basin DP5 attempts; "
                                  "one graph launch plus the basins catalogues")
    return dict(jobs=n_jobs, jobs_by_kind=kinds, s=dt, roofline=roofline, lle_traj_steps=lle_steps,
                basins_cells=bas_cells,
                basins_accepted_steps=bas_acc, traj_steps_per_s=(lle_steps + bas_acc) / dt,
                basins_attractors=[len(o["attractors"]) for o in basins_out], tuned=tuned, **b.info())


def extra_serving(eng, duration_s=8.0, rate=None, seed=1, capacity=None, dump_jobs=None, server_kw=None):
    """Poisson arrivals of mixed requests, a third each: Lorenz bifurcation
    sweeps (side^2 rho values), Rossler LLE maps (side x side) and forced
    Duffing basin maps (side x side initial conditions, rtol = atol = 1e-9),
    side log-uniform in [32, 2048], step count log-uniform in [1e3, 1e5]
    (basins: steps / 500 forcing periods, at least 16), FP32-compensated.
    Arrivals at 50 jobs/s, the survey's serving rate. The offered
    utilisation is estimated against `capacity` (LLE traj-steps/s),
    weighting a bifurcation step 0.5 and a basin section 410 LLE steps (the
    measured FP32C rates of the three kernels); if it would reach 0.8 the
    rate is lowered to 70% of capacity.
    The server runs jobs of more than 65,536 cells and 1,024 steps per cell
    in slices of about 1,024 steps (their cells resume over several rounds)
    and gives the waves of small grids issue priority.
    Reports throughput and p50 / p99 job latency, the slowdown (latency over
    the same request's latency alone on an idle server, measured after the
    stream for every job) and p99 latency of the smaller and larger half of
    the jobs by size (cells x steps); 1% of the jobs of each kind are run
    again alone in FP64 and compared. `dump_jobs`: a .json path for the
    per-job records (kind, side, steps, submit time, latency in the stream and
    alone); `server_kw`: ucuda.Server options."""
    import numpy as np
    import ucuda
    rng = np.random.default_rng(seed)
    ros, lor, duf = ucuda.preset("rossler", eng), ucuda.preset("lorenz", eng), ucuda.preset("duffing", eng)
    specs = []
    for _ in range(4000):
        side = int(round(32 * 64 ** rng.uniform()))
        steps = int(round(10 ** (3 + 2 * rng.uniform())))
        kind = ("lle", "bif", "basins")[int(rng.integers(0, 3))]
        specs.append((kind, side, steps))
    weight = {"lle": 1.0, "bif": 0.5}

    def cost(kind, side, steps):
        if kind == "basins":
            return side * side * max(16, steps // 500) * 410.0
        return side * side * steps * weight[kind]
    est = float(np.mean([cost(*sp) for sp in specs]))
    if rate is None:
        rate = 50.0
        if capacity and rate * est / capacity >= 0.8:
            rate = 0.7 * capacity / est
    server_kw = dict(server_kw or {})
    srv = ucuda.Server(eng, **server_kw)
    ids = []
    t_sub = []
    t = 0.0
    t0 = time.perf_counter()
    i = 0

    def request(kind, side, steps):
        if kind == "lle":
            axes = [ucuda.axis(1, 2, side, 0.05, 2.0), ucuda.axis(1, 3, side, 1.0, 20.0)]
            blocks = max(1, steps // 11)
            return axes, dict(fn="lyap", args=(ros, axes, [1, 1, 1], ros.values(), 0.01, 0.0, blocks * 0.1, 0.1))
        if kind == "bif":
            axes = [ucuda.axis(1, 2, side * side, 0.0, 250.0)]
            return axes, dict(fn="bif", args=(lor, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, steps * 0.005,
                                              steps * 0.005), kw=dict(max_peaks=8))
        axes = [ucuda.axis(2, 0, side, -2.5, 2.5), ucuda.axis(2, 1, side, -2.5, 2.5)]
        return axes, dict(fn="basins", args=(duf, axes, [0.0, 0.0], duf.values()),
                          kw=dict(n_sections=max(16, steps // 500)))

    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        kind, side, steps = specs[i % len(specs)]
        i += 1
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        axes, rq = request(kind, side, steps)
        kw = dict(rq.get("kw", {}), precision=ucuda.PREC_F32C)
        if rq["fn"] == "lyap":
            jid = srv.submit_lyap(*rq["args"], **kw)
        elif rq["fn"] == "bif":
            jid = srv.submit_bif(*rq["args"], **kw)
        else:
            jid = srv.submit_basins(*rq["args"], **kw)
        ids.append((jid, kind, side, steps))
        t_sub.append(time.perf_counter() - t0)
    progress("serving: %d jobs submitted, draining" % len(ids))
    pending, last = [jid for jid, *_ in ids], time.perf_counter()
    deadline = last + 300.0
    while pending and time.perf_counter() < deadline:
        pending = [jid for jid in pending if srv.poll(jid)[0] not in (2, 3, 4, 5)]
        if pending:
            time.sleep(0.02)
        if time.perf_counter() - last > 15.0:
            progress("serving: %d jobs still running" % len(pending))
            last = time.perf_counter()
    if pending:
        raise RuntimeError("serving: %d jobs unfinished after 300 s" % len(pending))
    wall = time.perf_counter() - t0
    st = srv.stats()
    failed = [jid for jid, *_ in ids if srv.poll(jid)[0] != 2]
    # audit: 1% of the jobs of each kind (every 100th, from the first)
    # recomputed alone in FP64
    audit = []
    sample = [x for kind_ in ("lle", "bif", "basins") for x in [y for y in ids if y[1] == kind_][::100]]
    for jid, kind, side, steps in sample:
        got = srv.fetch(jid)
        axes, rq = request(kind, side, steps)
        if kind == "lle":
            ref = ucuda.run_lyap(*rq["args"])
            rv = ref["values"].cpu().numpy()[0]
            ok = (got["flags"] == 1) & (ref["flags"].cpu().numpy() == 1)
            e = np.abs(got["values"][0][ok] - rv[ok])
            audit.append(dict(kind=kind, cells=side * side, median_abs_diff=float(np.median(e)) if e.size else None,
                              frac_le_1e3=float((e <= 1e-3).mean()) if e.size else None))
        elif kind == "bif":
            ref = ucuda.run_bif(*rq["args"], **rq["kw"])
            rf = ref["flags"].cpu().numpy()
            audit.append(dict(kind=kind, cells=side * side, frac_same_peak_count=float((got["flags"] == rf).mean())))
        else:
            ref = ucuda.run_basins(*rq["args"], **rq["kw"])
            m = {}
            for a32 in got["attractors"]:
                for a64 in ref["attractors"]:
                    if a32["period"] == a64["period"] and abs(a32["point"][0] - a64["point"][0]) <= 3e-3 and \
                            abs(a32["point"][1] - a64["point"][1]) <= 3e-3:
                        m[a32["label"]] = a64["label"]
            mapped = np.array([m.get(int(v), 0 if v == 0 else -99) for v in got["labels"]])
            audit.append(dict(kind=kind, cells=side * side,
                              label_agreement=float((mapped == ref["labels"].cpu().numpy()).mean())))
    lat = np.array([srv.poll(jid)[1] for jid, *_ in ids])
    srv.close()
    # each job's latency alone: the same request on an idle server (same
    # admission, rounds and kernels), one job at a time
    progress("serving: %d jobs again, each alone" % len(ids))
    solo = ucuda.Server(eng, **server_kw)
    solo_ms = []
    for jid, kind, side, steps in ids:
        axes, rq = request(kind, side, steps)
        kw = dict(rq.get("kw", {}), precision=ucuda.PREC_F32C)
        sub = {"lyap": solo.submit_lyap, "bif": solo.submit_bif, "basins": solo.submit_basins}[rq["fn"]]
        j2 = sub(*rq["args"], **kw)
        solo.wait(j2, 120.0)
        solo_ms.append(solo.poll(j2)[1])
        solo.fetch(j2)
    solo.close()
    solo_ms = np.array(solo_ms)
    costs = np.array([cost(kind, side, steps) for _, kind, side, steps in ids])
    if dump_jobs:
        with open(dump_jobs, "w") as f:
            json.dump([dict(kind=k, side=sd, steps=sp, t_submit=ts, latency_ms=float(l), solo_ms=float(so))
                       for (_, k, sd, sp), ts, l, so in zip(ids, t_sub, lat, solo_ms)], f)
    sizes = {}
    if len(ids):
        slow = lat / solo_ms
        small = costs <= np.median(costs)
        sizes = dict(solo_ms_p50=float(np.percentile(solo_ms, 50)), solo_ms_p99=float(np.percentile(solo_ms, 99)),
                     slowdown_p50=float(np.percentile(slow, 50)), slowdown_p99=float(np.percentile(slow, 99)),
                     slowdown_max=float(slow.max()),
                     p50_ms_smaller_half=float(np.percentile(lat[small], 50)),
                     p99_ms_smaller_half=float(np.percentile(lat[small], 99)),
                     p99_ms_larger_half=float(np.percentile(lat[~small], 99)))
    kinds = {k: sum(1 for _, kk, _, _ in ids if kk == k) for k in ("lle", "bif", "basins")}
    return dict(arrival_rate_per_s=rate, offered_utilisation_est=rate * est / capacity if capacity else None,
                server=dict(server_kw, slice_steps=server_kw.get("slice_steps", 1024)),
                jobs=len(ids), jobs_by_kind=kinds, completed=st["completed"], failed=len(failed),
                p50_ms=st["p50_ms"], p99_ms=st["p99_ms"], max_ms=st["max_ms"], **sizes,
                fixed_step_traj_steps_per_s=st["traj_steps"] / wall, rounds=st["rounds"],
                roofline=dict(bound="arrivals", gpu_busy_frac=st["busy_s"] / wall,
                              note="throughput is the offered load (50 jobs/s), not a kernel bound; the kernels' "
                                   "own rooflines are the other extras'"),
                gpu_busy_frac=st["busy_s"] / wall, wall_s=wall, audit_fp64=audit)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512, help="grid side")
    ap.add_argument("--cpu-stride", type=int, default=8, help="CPU subsample: every k-th point per axis")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the LS / basins / batch / serving measurements")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--extras", default="all",
                    help="comma-separated subset of the extra measurements (keys of the JSON line), or 'all'")
    ap.add_argument("--only", choices=["f32c", "f64"], default=None, help="time one precision only (profiling)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after timing, run one device-to-device copy of this size (HBM counter calibration)")
    ap.add_argument("--dump-cells", default="", help="write the sampled GPU and oracle exponents to this .npz")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine's path runs on a single MI355X "
              "(across GPUs the cells would shard by contiguous ranges, one process per GPU). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ucuda

    eng = ucuda.Engine(0)
    sysm = ucuda.preset("rossler", eng)
    w = workload(args.n)
    n_cells = args.n * args.n
    spc = steps_per_cell(w)
    traj_steps = n_cells * spc
    ops = sysm.op_counts()
    # algorithmic flops per launch: state-only transient steps, state +
    # tangent steps in the blocks
    state_flops = ucuda.rk4_flops(ops, 3, tangents=0)
    tangent_step_flops = ucuda.rk4_flops(ops, 3, tangents=1)
    tsteps = tangent_steps_per_cell(w)
    flops_per_launch = n_cells * ((spc - tsteps) * state_flops + tsteps * tangent_step_flops)
    flops_per_step = flops_per_launch / traj_steps

    res = {}
    for name, prec in (("f64", ucuda.PREC_F64), ("f32c", ucuda.PREC_F32C)):
        if args.only and name != args.only:
            continue
        res[name] = run_gpu(sysm, w, prec, args.steps, args.warmup)
        progress("%s LLE map: %.2f ms per map" % (name, sum(res[name]["kernel_ms"]) / len(res[name]["kernel_ms"])))

    cpu = None
    if not args.no_cpu and not args.only:
        strict = run_gpu(sysm, w, ucuda.PREC_F64, 1, 0, strict=True)
        progress("f64 strict LLE map: %.2f ms" % strict["kernel_ms"][0])
        cb = cpu_baseline(eng, sysm, w, args.n, args.cpu_stride, min(16, os.cpu_count() or 1))
        errs = {"f64": lle_errors(res["f64"]["values"], cb, "f64"), "f32c": lle_errors(res["f32c"]["values"], cb, "f32"),
                "f64_strict": strict_check(strict["values"], cb)}
        cpu = dict(cells=len(cb["cells"]), t1=cb["t1"], tN=cb["tN"], errs=errs)
        if args.dump_cells:
            np.savez(args.dump_cells, cells=cb["cells"], var=cb["var"], fd=cb["fd"], spread_f64=cb["spreads"]["f64"],
                     spread_f32=cb["spreads"]["f32"], f64=res["f64"]["values"][cb["cells"]],
                     f32c=res["f32c"]["values"][cb["cells"]], f64_strict=strict["values"][cb["cells"]])

    if args.calibrate_copy_mib:
        src = torch.empty(args.calibrate_copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        progress("copy calibration: %d MiB read + %d MiB written" % (args.calibrate_copy_mib, args.calibrate_copy_mib))
    if args.only:
        r = res[args.only]
        print(json.dumps({"only": args.only, "ms_per_map": float(np.mean(r["kernel_ms"])),
                          "traj_steps_per_s": traj_steps / (float(np.mean(r["kernel_ms"])) / 1e3)}))
        return
    f32 = res["f32c"]
    ms = float(np.mean(f32["kernel_ms"]))
    value = traj_steps / (ms / 1e3)
    f64_ms = float(np.mean(res["f64"]["kernel_ms"]))
    achieved = flops_per_launch / (ms / 1e3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "traj-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+kahan (f64 renorm/log)",
        "data": "synthetic",
        "config": {"workload": "rossler_lle_map_%dx%d" % (args.n, args.n), "cells": n_cells,
                   "steps_per_cell": spc, "h": w["h"], "NT": w["NT"], "transient_steps": int(w["transient"] / w["h"]),
                   "averaged_steps": int(w["t_max"] / w["h"]), "method": "variational", "parallelism": "none"},
        "fp64": {"value": traj_steps / (f64_ms / 1e3), "ms_per_step": f64_ms},
        "f32c_over_f64": f64_ms / ms,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32_VALU / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU, "flops_per_traj_step": flops_per_step,
                     "flops_per_launch": flops_per_launch,
                     "flops_per_step_by_phase": {"transient_state_only": state_flops,
                                                 "blocks_state_and_tangent": tangent_step_flops},
                     "kernel": headline_kernel(n_cells), "traffic": profiled_traffic()},
        "wall_s_per_step": f32["wall"] / args.steps,
    }
    if not args.no_extra:
        cap = value  # traj-steps/s of the headline kernel sets the serving load
        threads_all = min(16, os.cpu_count() or 1)
        extras = [
            ("bif_lorenz_1024", "Lorenz bifurcation, 1024 rho (the reference's CPU case)",
             lambda: extra_bif_config1(eng, threads_all)),
            ("bif2d_lorenz_512", "2-D period map (Lorenz 512^2)", lambda: extra_bif2d(eng)),
            ("dft_lorenz_16384", "DFT sweep (Lorenz 16384 rho x 200 freqs)", lambda: extra_dft(eng)),
            ("fastsync_lorenz", "fast synchro (Lorenz grid 1024^2, 10000 windows on the attractor)",
             lambda: extra_fastsync(eng)),
            ("continuation_lorenz", "continuation LLE chains (Lorenz, 128 rho points; 1 and 1024 chains)",
             lambda: extra_continuation(eng)),
            ("ls_hyper_rossler_1024", "LS spectrum map (hyper-Rossler 1024^2)", lambda: extra_ls(eng)),
            ("basins_duffing_2048", "basins (Duffing 2048^2)", lambda: extra_basins(eng)),
            ("batch_64_jobs", "batch (64 jobs)", lambda: extra_batch(eng)),
            ("serving", "serving (%.0f s of Poisson arrivals)" % args.serve_seconds,
             lambda: extra_serving(eng, duration_s=args.serve_seconds, capacity=cap)),
        ]
        wanted = None if args.extras == "all" else set(args.extras.split(","))
        for key, what, fn in extras:
            if wanted is not None and key not in wanted:
                continue
            progress("%s start" % what)
            line[key] = fn()
            progress("%s done" % what)
        if "serving" in line:
            line["p99_job_latency_ms"] = line["serving"]["p99_ms"]
    if cpu:
        line["cpu_baseline"] = {"value": cpu["cells"] * spc / cpu["t1"], "unit": "traj-steps/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d cells (every %dth point per axis), full horizon, FP64 finite-difference "
                                          "Benettin as the reference CPU path" % (cpu["cells"], args.cpu_stride),
                                "value_all_cores": cpu["cells"] * spc / cpu["tN"],
                                "cores_all": min(16, os.cpu_count() or 1)}
        line["lle_err_vs_fp64_oracle"] = cpu["errs"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
