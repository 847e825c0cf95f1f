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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "u-cuda_amd"))

METRIC = "RK4 trajectory-steps/s, FP32+comp vs FP64, at LLE err≤1e-3; p99 job latency"
PEAK_FP32_VALU = 157.3e12   # MI355X published FP32 vector peak (packed)
PEAK_FP64_VALU = 78.6e12    # MI355X published FP64 vector peak


def workload(n):
    import ucuda
    return dict(axes=[ucuda.axis(ucuda.AXIS_PARAM, 2, n, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, n, 1.0, 20.0)],
                ic=[1.0, 1.0, 1.0], h=0.01, transient=50.0, t_max=500.0, NT=0.1)


def steps_per_cell(w):
    skip = int(w["transient"] / w["h"])
    blocks = int(w["t_max"] / w["NT"])
    nts = int(w["NT"] / w["h"])
    return skip + 1 + blocks * (nts + 1)


def run_gpu(sysm, w, precision, steps, warmup):
    import torch
    import ucuda
    n = ucuda.n_cells(w["axes"])
    dev = torch.device("cuda", 0)
    out = dict(values=torch.empty((1, n), dtype=torch.float64, device=dev),
               flags=torch.empty(n, dtype=torch.int32, device=dev), final_state=None)
    args = (sysm, w["axes"], w["ic"], sysm.values(a=0.2), w["h"], w["transient"], w["t_max"], w["NT"])
    for _ in range(warmup):
        ucuda.run_lyap(*args, precision=precision, out=out)
    stream = torch.cuda.current_stream()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ev in events:
        ev[0].record(stream)
        ucuda.run_lyap(*args, precision=precision, out=out, stream=stream, sync=False)
        ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]
    return dict(wall=wall, kernel_ms=kernel_ms, values=out["values"][0].cpu().numpy(),
                flags=out["flags"].cpu().numpy())


def cpu_baseline(eng, sysm, w, n, stride, threads_all):
    """Times the reference CPU path (finite-difference Benettin, FP64, compiled
    step body) single-threaded and on all cores over a fixed subsample, and
    computes the FP64 variational oracle on the same cells twice (X0 and X0
    perturbed by 1e-9) to separate reproducible cells from sensitive ones."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    step = oracle.KrsStep(eng.codegen_scheme(sysm.vars, sysm.params, sysm.rhs), 3, 4)
    vals = sysm.values(a=0.2)
    job = oracle.make_job(3, w["axes"], w["ic"], vals, w["h"], w["transient"], w["t_max"], NT=w["NT"])
    idx = np.arange(0, n, stride)
    cells = (idx[:, None] * n + idx[None, :]).ravel()  # iy * n + ix, every stride-th point per axis
    t0 = time.perf_counter()
    r1 = oracle.lle_fd_cells(step, job, cells, threads=1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rN = oracle.lle_fd_cells(step, job, cells, threads=threads_all)
    tN = time.perf_counter() - t0
    assert (r1["values"] == rN["values"]).all()
    funcs = oracle.SystemFuncs(sysm.source(1))
    var = oracle.lyap_var_cells(funcs, 1, job, cells, threads=threads_all)["values"][0]
    spread = np.zeros_like(var)
    for d in ((1e-9, 0, 0), (0, 0, 1e-9), (0, -1e-9, 0)):
        icp = [w["ic"][i] + d[i] for i in range(3)]
        jobp = oracle.make_job(3, w["axes"], icp, vals, w["h"], w["transient"], w["t_max"], NT=w["NT"])
        varp = oracle.lyap_var_cells(funcs, 1, jobp, cells, threads=threads_all)["values"][0]
        spread = np.maximum(spread, np.abs(varp - var))
    return dict(cells=cells, fd=r1["values"][0], var=var, spread=spread, t1=t1, tN=tN)


def lle_errors(g, cpu):
    """Error of a GPU LLE map against the FP64 oracle on the sampled cells."""
    import numpy as np
    g = g[cpu["cells"]]
    var, spread, fd = cpu["var"], cpu["spread"], cpu["fd"]
    ok = var != 999.0
    # reproducible: three 1e-9 perturbations of X0 move the FP64 exponent by <= 1e-5
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512, help="grid side")
    ap.add_argument("--cpu-stride", type=int, default=8, help="CPU subsample: every k-th point per axis")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine's path runs on a single MI355X "
              "(cells shard by contiguous ranges across GPUs; see DESIGN.md). Use --gpus 1.", file=sys.stderr)
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
    flops_per_step = ucuda.rk4_flops(ops, 3, tangents=1)

    res = {}
    for name, prec in (("f64", ucuda.PREC_F64), ("f32c", ucuda.PREC_F32C)):
        res[name] = run_gpu(sysm, w, prec, args.steps, args.warmup)

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(eng, sysm, w, args.n, args.cpu_stride, min(16, os.cpu_count() or 1))
        errs = {name: lle_errors(res[name]["values"], cb) for name in res}
        cpu = dict(cells=len(cb["cells"]), t1=cb["t1"], tN=cb["tN"], errs=errs)

    f32 = res["f32c"]
    ms = float(np.mean(f32["kernel_ms"]))
    value = traj_steps / (ms / 1e3)
    f64_ms = float(np.mean(res["f64"]["kernel_ms"]))
    achieved = traj_steps * flops_per_step / (ms / 1e3)
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
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU,
                     "frac_unpacked_fp32": achieved / PEAK_FP64_VALU, "flops_per_traj_step": flops_per_step,
                     "kernel": "uc_lle_f32c", "traffic": None},
        "wall_s_per_step": f32["wall"] / args.steps,
    }
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
