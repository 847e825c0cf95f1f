"""Continuous-batching server: jobs submitted while others run are admitted
into later rounds, share rounds with them, and every result (bifurcation,
Lyapunov, basins with their attractor catalogue) is bit-identical to the
same request run alone, whole or resumed over several rounds (slices);
cancellation, admission by memory budget and latency statistics."""
import time

import numpy as np
import pytest

import ucuda

pytestmark = pytest.mark.gpu


def _mix(eng):
    ros, lor, hr = ucuda.preset("rossler", eng), ucuda.preset("lorenz", eng), ucuda.preset("hyper_rossler", eng)
    jobs = []
    for i in range(6):
        n = 16 + 13 * i
        jobs.append(("lyap", ros, dict(axes=[ucuda.axis(1, 2, n, 0.05, 2.0), ucuda.axis(1, 3, n, 1.0, 20.0)],
                                       ic=[1, 1, 1], values=ros.values(), h=0.01, transient_time=5.0 + i,
                                       t_max=20.0 + 10 * i, NT=0.1,
                                       precision=ucuda.PREC_F32C if i % 2 else ucuda.PREC_F64)))
    jobs.append(("bif", lor, dict(axes=[ucuda.axis(1, 2, 333, 0.0, 250.0)], ic=[0.1, 0.1, 0.1], values=lor.values(),
                                  h=0.01, transient_time=20.0, t_max=20.0, max_peaks=32)))
    jobs.append(("lyap", hr, dict(axes=[ucuda.axis(1, 2, 9, 0.0, 0.1), ucuda.axis(1, 1, 7, 0.2, 0.3)], ic=None,
                                  values=hr.values(), h=0.005, transient_time=2.0, t_max=10.0, NT=0.1, spectrum=True,
                                  ic_seed=20261203)))
    duf = ucuda.preset("duffing", eng)
    for prec in (ucuda.PREC_F64, ucuda.PREC_F32C):
        jobs.append(("basins", duf, dict(axes=[ucuda.axis(2, 0, 24, -2.5, 2.5), ucuda.axis(2, 1, 20, -2.5, 2.5)],
                                         ic=[0.0, 0.0], values=duf.values(), n_sections=40, precision=prec)))
    return jobs


def _submit(srv, kind, sysm, kw):
    kw = dict(kw)
    if kind == "basins":
        return srv.submit_basins(sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), **kw)
    args = (sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), kw.pop("h"), kw.pop("transient_time"),
            kw.pop("t_max"))
    if kind == "lyap":
        return srv.submit_lyap(*args, kw.pop("NT"), **kw)
    return srv.submit_bif(*args, **kw)


def _alone(kind, sysm, kw):
    kw = dict(kw)
    if kind == "basins":
        out = ucuda.run_basins(sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), **kw)
        return dict(labels=out["labels"].cpu().numpy(), periods=out["periods"].cpu().numpy(),
                    points=out["points"].cpu().numpy(), steps=out["steps"].cpu().numpy(),
                    attractors=out["attractors"])
    args = (sysm, kw.pop("axes"), kw.pop("ic"), kw.pop("values"), kw.pop("h"), kw.pop("transient_time"),
            kw.pop("t_max"))
    if kind == "lyap":
        out = ucuda.run_lyap(*args, kw.pop("NT"), **kw)
        return dict(values=out["values"].cpu().numpy(), flags=out["flags"].cpu().numpy())
    out = ucuda.run_bif(*args, **kw)
    return dict(flags=out["flags"].cpu().numpy(), peaks=out["peaks"].cpu().numpy(),
                intervals=out["intervals"].cpu().numpy())


_ALONE = {}


@pytest.mark.parametrize("slice_steps", [0.0, 300.0, -1.0], ids=["slice1024", "slice300", "whole"])
def test_results_identical_alone_and_in_the_mix(gpu_engine, slice_steps):
    print("case: 10 staggered submissions (lle f64/f32c, bif, ls, basins f64/f32c) with small rounds, "
          "slice_steps %g" % slice_steps, flush=True)
    jobs = _mix(gpu_engine)
    if not _ALONE:
        _ALONE.update({i: _alone(k, s, kw) for i, (k, s, kw) in enumerate(jobs)})
    alone = [_ALONE[i] for i in range(len(jobs))]
    srv = ucuda.Server(gpu_engine, round_traj_steps=2e7, slice_steps=slice_steps, slice_min_cells=-1)
    ids = []
    for i, (k, s, kw) in enumerate(jobs):
        ids.append(_submit(srv, k, s, kw))
        if i % 3 == 2:
            time.sleep(0.01)  # later jobs arrive while earlier ones are running
    for jid in ids:
        srv.wait(jid, 120.0)
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["rounds"] > 1
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    for (k, s, kw), jid, ref in zip(jobs, ids, alone):
        got = srv.fetch(jid)
        if k == "lyap":
            np.testing.assert_array_equal(got["flags"], ref["flags"])
            np.testing.assert_array_equal(got["values"], ref["values"])
        elif k == "basins":
            for key in ("labels", "periods", "points", "steps"):
                np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
            assert got["attractors"] == ref["attractors"]
            assert (got["labels"] > 0).any()
        else:
            np.testing.assert_array_equal(got["flags"], ref["flags"])
            for c in range(len(ref["flags"])):
                m = ref["flags"][c]
                if m > 0:
                    np.testing.assert_array_equal(got["peaks"][:m, c], ref["peaks"][:m, c])
                    np.testing.assert_array_equal(got["intervals"][:m, c], ref["intervals"][:m, c])
    srv.close()


def test_admission_budget_and_cancel(gpu_engine):
    """Outputs held on the device count against the budget until fetched:
    with room for two 120 kB jobs, the third waits while the first two are
    done but unfetched, and is admitted once one is fetched."""
    print("case: 300 kB HBM budget, three 120 kB jobs, fetch frees room; cancel a queued job", flush=True)
    import time
    ros = ucuda.preset("rossler", gpu_engine)
    srv = ucuda.Server(gpu_engine, max_bytes=300000, round_traj_steps=1e7)
    axes = [ucuda.axis(1, 2, 100, 0.05, 2.0), ucuda.axis(1, 3, 100, 1.0, 20.0)]  # 10k cells = 120 kB each
    ids = [srv.submit_lyap(ros, axes, [1, 1, 1], ros.values(), 0.01, 5.0, 20.0, 0.1) for _ in range(4)]
    srv.cancel(ids[3])
    for jid in ids[:2]:
        srv.wait(jid, 120.0)
    time.sleep(0.2)
    assert srv.poll(ids[2])[0] == 0                 # still queued: no room while 0 and 1 hold results
    assert srv.stats()["used_bytes"] <= 300000
    a = srv.fetch(ids[0])["values"]                 # frees 120 kB
    srv.wait(ids[2], 120.0)
    assert srv.stats()["used_bytes"] <= 300000
    b = srv.fetch(ids[1])["values"]
    c = srv.fetch(ids[2])["values"]
    assert srv.poll(ids[3])[0] == 3
    st = srv.stats()
    assert st["completed"] == 3 and st["used_bytes"] == 0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
    srv.close()


def test_short_jobs_do_not_wait_for_long_cells(gpu_engine):
    """Short jobs submitted right after a job of long cells finish long
    before it: rounds hold one duration class (decade of steps per cell), so
    they never share the long job's round, and with the GPU not full (the
    long job is 4,096 cells, run whole) their rounds run beside it."""
    print("case: one 64x64 LLE job of 110,001 steps per cell, then five 16x16 jobs of 111", flush=True)
    ros = ucuda.preset("rossler", gpu_engine)
    vals = ros.values()
    big = [ucuda.axis(1, 2, 64, 0.05, 2.0), ucuda.axis(1, 3, 64, 1.0, 20.0)]
    small = [ucuda.axis(1, 2, 16, 0.05, 2.0), ucuda.axis(1, 3, 16, 1.0, 20.0)]
    # warm module and allocator outside the timed part
    ucuda.run_lyap(ros, small, [1, 1, 1], vals, 0.01, 0.0, 1.0, 0.1, precision=ucuda.PREC_F32C)
    srv = ucuda.Server(gpu_engine)
    try:
        long_id = srv.submit_lyap(ros, big, [1, 1, 1], vals, 0.01, 0.0, 1000.0, 0.1, precision=ucuda.PREC_F32C)
        short_ids = [srv.submit_lyap(ros, small, [1, 1, 1], vals, 0.01, 0.0, 1.0, 0.1, precision=ucuda.PREC_F32C)
                     for _ in range(5)]
        for jid in short_ids + [long_id]:
            srv.wait(jid, timeout_s=120.0)
        lat_long = srv.poll(long_id)[1]
        lat_short = [srv.poll(j)[1] for j in short_ids]
        print("long job %.1f ms; short jobs %s ms" % (lat_long, " ".join("%.1f" % v for v in lat_short)))
        assert all(srv.poll(j)[0] == 2 for j in short_ids + [long_id])
        assert max(lat_short) < 0.5 * lat_long
    finally:
        srv.close()


def test_short_jobs_pass_a_job_that_fills_the_gpu(gpu_engine):
    """A long job with more cells than the GPU holds resident (1024^2) runs
    in slices, so short jobs submitted after it get into the next rounds
    instead of waiting for its cells to finish."""
    print("case: one 1024x1024 LLE job of 55,001 steps per cell, then five 16x16 jobs of 111", flush=True)
    ros = ucuda.preset("rossler", gpu_engine)
    vals = ros.values()
    big = [ucuda.axis(1, 2, 1024, 0.05, 2.0), ucuda.axis(1, 3, 1024, 1.0, 20.0)]
    small = [ucuda.axis(1, 2, 16, 0.05, 2.0), ucuda.axis(1, 3, 16, 1.0, 20.0)]
    ucuda.run_lyap(ros, small, [1, 1, 1], vals, 0.01, 0.0, 1.0, 0.1, precision=ucuda.PREC_F32C)
    srv = ucuda.Server(gpu_engine)
    try:
        long_id = srv.submit_lyap(ros, big, [1, 1, 1], vals, 0.01, 0.0, 500.0, 0.1, precision=ucuda.PREC_F32C)
        time.sleep(0.005)
        short_ids = [srv.submit_lyap(ros, small, [1, 1, 1], vals, 0.01, 0.0, 1.0, 0.1, precision=ucuda.PREC_F32C)
                     for _ in range(5)]
        for jid in short_ids + [long_id]:
            srv.wait(jid, timeout_s=120.0)
        lat_long = srv.poll(long_id)[1]
        lat_short = [srv.poll(j)[1] for j in short_ids]
        print("long job %.1f ms; short jobs %s ms" % (lat_long, " ".join("%.1f" % v for v in lat_short)))
        assert all(srv.poll(j)[0] == 2 for j in short_ids + [long_id])
        assert max(lat_short) < 0.3 * lat_long
    finally:
        srv.close()
