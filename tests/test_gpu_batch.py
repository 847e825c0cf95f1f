"""Persistent heterogeneous batch: every job's result from the captured-graph
batch launch is bit-identical to the same job run alone (same kernel code
per cell, placement does not matter), across systems, analyses (LLE, LS,
bifurcation, basins with their attractor catalogue), precisions and grid
sizes; graph replays reproduce the first launch."""
import numpy as np
import pytest
import torch

import ucuda

pytestmark = pytest.mark.gpu


def _jobs(eng):
    lor, ros, chen, hr, duf = (ucuda.preset(n, eng) for n in ("lorenz", "rossler", "chen", "hyper_rossler", "duffing"))
    J = []
    J.append(("lyap", ros, dict(axes=[ucuda.axis(1, 2, 37, 0.05, 2.0), ucuda.axis(1, 3, 29, 1.0, 20.0)],
                                ic=[1, 1, 1], values=ros.values(), h=0.01, transient_time=20.0, t_max=60.0, NT=0.1,
                                precision=ucuda.PREC_F32C)))
    J.append(("lyap", lor, dict(axes=[ucuda.axis(1, 2, 300, 20.0, 200.0)], ic=[0.1, 0.1, 0.1], values=lor.values(),
                                h=0.01, transient_time=5.0, t_max=30.0, NT=0.1)))
    J.append(("lyap", hr, dict(axes=[ucuda.axis(1, 2, 12, 0.0, 0.1), ucuda.axis(1, 1, 11, 0.2, 0.3)], ic=None,
                               values=hr.values(), h=0.005, transient_time=2.0, t_max=20.0, NT=0.1, spectrum=True,
                               ic_seed=20261203)))
    J.append(("lyap", duf, dict(axes=[ucuda.axis(2, 0, 33, -2.5, 2.5), ucuda.axis(2, 1, 17, -2.5, 2.5)],
                                ic=[0.0, 0.0], values=duf.values(), h=0.01, transient_time=10.0, t_max=40.0, NT=0.1)))
    J.append(("bif", chen, dict(axes=[ucuda.axis(1, 3, 500, 20.0, 30.0)], ic=[0.1, 0.1, 0.1], values=chen.values(),
                                h=0.0025, transient_time=5.0, t_max=10.0, max_peaks=32)))
    J.append(("bif", lor, dict(axes=[ucuda.axis(1, 2, 7, 100.0, 200.0)], ic=[0.1, 0.1, 0.1], values=lor.values(),
                               h=0.01, transient_time=20.0, t_max=20.0, max_peaks=32, precision=ucuda.PREC_F32C)))
    J.append(("basins", duf, dict(axes=[ucuda.axis(2, 0, 40, -2.5, 2.5), ucuda.axis(2, 1, 36, -2.5, 2.5)],
                                  ic=[0.0, 0.0], values=duf.values(), n_sections=60)))
    J.append(("basins", duf, dict(axes=[ucuda.axis(2, 0, 24, -2.0, 2.0), ucuda.axis(2, 1, 20, -2.0, 2.0)],
                                  ic=[0.0, 0.0], values=duf.values(), n_sections=40, precision=ucuda.PREC_F32C)))
    return J


def _alone(kind, sysm, kw):
    kw = dict(kw)
    if kind == "basins":
        return ucuda.run_basins(sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), **kw)
    if kind == "lyap":
        return ucuda.run_lyap(sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), kw.pop("h"),
                              kw.pop("transient_time"), kw.pop("t_max"), kw.pop("NT"), **kw)
    return ucuda.run_bif(sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), kw.pop("h"), kw.pop("transient_time"),
                         kw.pop("t_max"), **kw)


def _written(o, key):
    """Host copy of output `key` with the slots a kernel does not write set to
    zero: bifurcation rows k >= flags[c] of peaks / intervals (a cell writes
    only its own peak count; the rest of the buffer is whatever the allocator
    held)."""
    v = o[key].cpu().clone()
    if key in ("peaks", "intervals"):
        n = o["flags"].cpu().clamp(min=0).to(torch.int64)
        v[torch.arange(v.shape[0])[:, None] >= n[None, :]] = 0
    return v


def _same(a, b):
    if isinstance(a, list):
        return a == b
    return a.shape == b.shape and bool(torch.all((a == b) | (torch.isnan(a) & torch.isnan(b)))) \
        if a.is_floating_point() else torch.equal(a, b)


def _add(b, kind, sysm, kw):
    kw2 = dict(kw)
    if kind == "basins":
        return b.add_basins(sysm, kw2.pop("axes"), kw2.pop("ic"), kw2.pop("values"), **kw2)
    args = (sysm, kw2.pop("axes"), kw2.pop("ic"), kw2.pop("values"), kw2.pop("h"), kw2.pop("transient_time"),
            kw2.pop("t_max"))
    if kind == "lyap":
        return b.add_lyap(*args, kw2.pop("NT"), **kw2)
    return b.add_bif(*args, **kw2)


def _snapshot(o):
    return {k: (list(v) if k == "attractors" else _written(o, k)) for k, v in o.items() if v is not None}


_ALONE = {}


# plans: whole runs; stages of 512 and 2048 traj-steps per cell (every job
# is sliced, the basins maps one or two sections per stage); two cells per
# lane for the LLE maps that can; one lane per tangent vector for the
# spectrum map (alone, its 132 cells take the same N-lanes kernel)
PLANS = [None, dict(slice_steps=512.0), dict(block=64, waves_per_simd=2, slice_steps=2048.0),
         dict(cells_per_lane=2), dict(block=128, vector_lanes=1)]


@pytest.mark.parametrize("plan", PLANS, ids=["whole", "slice512", "slice2048_b64", "pair", "vector_lanes"])
def test_batch_equals_alone_and_replays(gpu_engine, plan):
    print("case: 8 heterogeneous jobs (5 systems, lle/ls/bif/basins, f64/f32c) in one graph launch, plan %s"
          % plan, flush=True)
    jobs = _jobs(gpu_engine)
    b = ucuda.Batch(gpu_engine)
    outs = [_add(b, kind, sysm, kw) for kind, sysm, kw in jobs]
    if plan:
        b.set_plan(**plan)
        got = b.plan()
        for k, v in plan.items():
            assert got[k] == v
    info = b.info()
    assert info["n_jobs"] == 8 and info["n_groups"] >= 7
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.launch(stream)
    torch.cuda.synchronize()
    first = [_snapshot(o) for o in outs]
    for i, ((kind, sysm, kw), f) in enumerate(zip(jobs, first)):
        if i not in _ALONE:
            _ALONE[i] = _snapshot(_alone(kind, sysm, kw))
        ref = _ALONE[i]
        assert set(f) == set(ref), (kind, sysm.name)
        for key in f:
            assert _same(f[key], ref[key]), (kind, sysm.name, key)
        if kind == "basins":
            print(sysm.name, "attractors:", f["attractors"])
            assert len(f["attractors"]) >= 2
    # replays with eager work in between reproduce the first launch
    for _ in range(3):
        for o in outs:
            o.pop("attractors", None)
            for v in o.values():
                if v is not None:
                    v.fill_(-7)
        with torch.cuda.stream(stream):
            b.launch(stream)
        torch.cuda.synchronize()
        for o, f in zip(outs, first):
            g = _snapshot(o)
            for key in f:
                assert _same(g[key], f[key]), key


def test_batch_tuner_and_shapes(gpu_engine):
    print("case: tuner over launch shapes, steps per launch, cells per lane and lanes per tangent vector; "
          "every plan gives identical "
          "results", flush=True)
    ros = ucuda.preset("rossler", gpu_engine)
    b = ucuda.Batch(gpu_engine)
    outs = [b.add_lyap(ros, [ucuda.axis(1, 2, 64, 0.05, 2.0), ucuda.axis(1, 3, n, 1.0, 20.0)], [1, 1, 1],
                       ros.values(), 0.01, 5.0, 200.0, 0.1, precision=ucuda.PREC_F32C) for n in (3, 50, 200)]
    hr = ucuda.preset("hyper_rossler", gpu_engine)
    outs.append(b.add_lyap(hr, [ucuda.axis(2, 0, 40, -12.0, -8.0), ucuda.axis(2, 3, 40, 8.1, 12.1)],
                           [-10.0, -6.0, 0.0, 10.1], hr.values(), 0.005, 2.0, 20.0, 0.1, spectrum=True,
                           precision=ucuda.PREC_F32C))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        best = b.tune(stream)
        torch.cuda.synchronize()
        print("tuned plan:", best)
        ref = [o["values"].clone() for o in outs]
        assert best["block"] in (64, 256) and 1 <= best["waves_per_simd"] <= 8 and best["ms"] > 0
        assert best["slice_steps"] in (0.0, 16384.0, 4096.0) and best["cells_per_lane"] in (1, 2)
        assert best["vector_lanes"] in (0, 1)
        assert b.plan() == best
        for blk, w, sl, cpl, vl in ((64, 2, 0.0, 1, 0), (256, 8, 1000.0, 1, 0), (128, 1, 0.0, 2, 0),
                                    (256, 4, 300.0, 1, 0), (64, 4, 0.0, 2, 1), (256, 2, 0.0, 1, 1)):
            b.set_plan(blk, w, sl, cpl, vl)
            b.launch(stream)
            for o, r in zip(outs, ref):  # NaN exponents (an escape inside the last block) compare equal
                assert torch.equal(o["values"].nan_to_num(nan=-777.0), r.nan_to_num(nan=-777.0)), (blk, w, sl, cpl, vl)
