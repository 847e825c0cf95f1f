"""Continuation sweeps through the C ABI against the FP64 oracle (strict
FP64: states bit-identical to the oracle; bifurcation peaks and DFT
coefficients compared exactly; Lyapunov sums, whose per-block log may round
differently on the device, at 1e-12 absolute)."""
import numpy as np
import pytest

import oracle
import ucuda

pytestmark = pytest.mark.gpu


def _lorenz(eng):
    s = ucuda.preset("lorenz", eng)
    return s, oracle.KrsStep(eng.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)


def test_bif_chains(gpu_engine):
    s, step = _lorenz(gpu_engine)
    vals = s.values()
    cases = (([ucuda.axis(ucuda.AXIS_PARAM, 2, 40, 20.0, 200.0)], False),
             ([ucuda.axis(ucuda.AXIS_PARAM, 2, 40, 20.0, 200.0)], True),
             ([ucuda.axis(ucuda.AXIS_STEP, 0, 24, 0.004, 0.012)], False),
             ([ucuda.axis(ucuda.AXIS_PARAM, 2, 30, 30.0, 170.0), ucuda.axis(ucuda.AXIS_PARAM, 1, 5, 6.0, 14.0)], True))
    for axes, rev in cases:
        print("case: bifurcation chain %s reverse=%d" % ([a["kind"] for a in axes], rev), flush=True)
        g = ucuda.run_bif_cont(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0, reverse=rev, max_peaks=64,
                               pre_scaller=2, strict=True)
        job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0, pre_scaller=2, max_peaks=64)
        r = oracle.cont_bif(step, job, reverse=rev, threads=8)
        gf = g["flags"].cpu().numpy()
        np.testing.assert_array_equal(gf, r["flags"])
        gp, gi = g["peaks"].cpu().numpy(), g["intervals"].cpu().numpy()
        for c in np.nonzero(r["flags"] > 0)[0]:
            m = r["flags"][c]
            np.testing.assert_array_equal(gp[:m, c], r["peaks"][:m, c])
            np.testing.assert_array_equal(gi[:m, c], r["intervals"][:m, c])
        assert (r["flags"] > 0).sum() >= len(gf) // 2


def test_lyap_chains(gpu_engine):
    s, step = _lorenz(gpu_engine)
    vals = s.values()
    for spectrum, axes, rev in ((False, [ucuda.axis(ucuda.AXIS_PARAM, 2, 24, 10.0, 200.0)], False),
                                (False, [ucuda.axis(ucuda.AXIS_PARAM, 2, 24, 10.0, 200.0)], True),
                                (True, [ucuda.axis(ucuda.AXIS_PARAM, 2, 12, 20.0, 120.0),
                                        ucuda.axis(ucuda.AXIS_PARAM, 3, 3, 1.0, 3.0)], False),
                                (False, [ucuda.axis(ucuda.AXIS_STEP, 0, 10, 0.005, 0.02)], True)):
        print("case: Lyapunov chain spectrum=%d reverse=%d axes=%d" % (spectrum, rev, len(axes)), flush=True)
        g = ucuda.run_lyap_cont(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 4.0, 8.0, 0.1, spectrum=spectrum, reverse=rev,
                                strict=True)
        job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 4.0, 8.0, NT=0.1, eps=1e-4)
        r = oracle.cont_lyap(step, job, 3 if spectrum else 1, reverse=rev, threads=8)
        np.testing.assert_array_equal(g["flags"].cpu().numpy(), r["flags"])
        np.testing.assert_allclose(g["values"].cpu().numpy(), r["values"], rtol=0, atol=1e-12)
        assert (r["flags"] == 1).all()


def test_dft_chains(gpu_engine):
    s, step = _lorenz(gpu_engine)
    vals = s.values()
    for axes, rev, kw in (([ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.5, 250.0)], False,
                           dict(n_freq=40, freq_lo=0.0, freq_hi=5.0, window_type=1)),
                          ([ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.5, 250.0)], True,
                           dict(n_freq=40, freq_lo=0.05, freq_hi=5.0, freq_log_scale=True, window_type=2)),
                          ([ucuda.axis(ucuda.AXIS_STEP, 0, 12, 0.004, 0.02)], False,
                           dict(n_freq=33, freq_lo=0.0, freq_hi=4.0, window_type=0))):
        print("case: DFT chain axis kind %d reverse=%d" % (axes[0]["kind"], rev), flush=True)
        g = ucuda.run_dft_cont(s, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0, reverse=rev, strict=True, **kw)
        job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], vals, 0.01, 10.0, 20.0)
        r = oracle.cont_dft(step, job, kw["n_freq"], kw["freq_lo"], kw["freq_hi"], reverse=rev,
                            log_axis=kw.get("freq_log_scale", False), window_type=kw["window_type"], threads=8)
        np.testing.assert_array_equal(g["flags"].cpu().numpy(), r["flags"])
        np.testing.assert_array_equal(g["ak"].cpu().numpy(), r["ak"])
        np.testing.assert_array_equal(g["bk"].cpu().numpy(), r["bk"])
        assert (r["flags"] == 1).sum() >= 8
        if axes[0]["kind"] == ucuda.AXIS_PARAM:
            assert (r["flags"] == -1).sum() == 1  # rho = 0.5 ends on the origin's fixed point


def test_hysteresis_and_errors(gpu_engine):
    print("case: fold hysteresis x' = r + x - x^3, forward vs reverse chains; request checks", flush=True)
    s = ucuda.System(gpu_engine, ["x"], ["r"], ["r+x-x^3"])
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 1, 2)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 1, 9, -1.0, 1.0)]
    fwd = ucuda.run_lyap_cont(s, axes, [0.0], [0.5, 0.0], 0.01, 20.0, 20.0, 0.1)["values"].cpu().numpy()[0]
    rev = ucuda.run_lyap_cont(s, axes, [0.0], [0.5, 0.0], 0.01, 20.0, 20.0, 0.1, reverse=True)["values"].cpu().numpy()[0]
    job = oracle.make_job(1, axes, [0.0], [0.5, 0.0], 0.01, 20.0, 20.0, NT=0.1, eps=1e-4)
    np.testing.assert_allclose(fwd, oracle.cont_lyap(step, job, 1)["values"][0], atol=1e-9)
    # r = -0.25 and 0.25 lie inside the fold: the two walks sit on different branches (at r = 0 the
    # branches x* = -1 and 1 have the same exponent)
    assert np.abs(fwd[[3, 5]] - rev[::-1][[3, 5]]).min() > 0.5
    with pytest.raises(ucuda.UcError):  # continuation walks a parameter or h, not an initial state
        ucuda.run_lyap_cont(s, [ucuda.axis(ucuda.AXIS_VAR, 0, 4, 0.0, 1.0)], [0.0], [0.5, 0.0], 0.01, 1.0, 2.0, 0.1)
