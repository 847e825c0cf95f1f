"""Harmonic forcing in the fixed-step RK4 kernels of the fast modules.

Each cell state tracks cos / sin of its forcing phases and rotates them to
the RK4 stage times; strict modules evaluate cos / sin in every right-hand
side like the oracle. Over horizons short enough that chaotic growth stays
small, fast FP64 follows strict FP64 to rounding-scale differences and
FP32C to its own precision on the bulk of the cells, on the wave-uniform
loops (phases evaluated afresh at every chunk and block start) and on the
per-lane loops of a sweep over h (afresh only every 128 steps).
"""
import numpy as np
import pytest

import ucuda

pytestmark = pytest.mark.gpu

RHS = ["v", "-d*v + x - x^3 + g*cos(w*t) + 0.05*sin(2*w*t + f)"]
VALS = [0.5, 0.3, 0.2, 1.0, 0.7]


def _system(eng):
    s = ucuda.System(eng, ["x", "v"], ["d", "g", "w", "f"], RHS, time_symbol="t")
    assert "#define UC_HARMONICS 2" in s.source(0)
    return s


def test_lle_map_fast_vs_strict(gpu_engine):
    print("case: two-harmonic forced Duffing LLE 256x256 over (x0, v0), fast vs strict, t = 10", flush=True)
    s = _system(gpu_engine)
    n = 256
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, n, -2.0, 2.0), ucuda.axis(ucuda.AXIS_VAR, 1, n, -2.0, 2.0)]
    args = (s, axes, [0.0, 0.0], VALS, 0.01, 2.0, 10.0, 0.1)
    ref = ucuda.run_lyap(*args, strict=True)
    rv, rf = ref["values"].cpu().numpy()[0], ref["flags"].cpu().numpy()
    assert (rf == 1).all()
    for prec, tol in ((ucuda.PREC_F64, 1e-9), (ucuda.PREC_F32C, 1e-3)):
        g = ucuda.run_lyap(*args, precision=prec)
        gv, gf = g["values"].cpu().numpy()[0], g["flags"].cpu().numpy()
        print("precision %d: max |LLE err| %.2e" % (prec, np.abs(gv - rv).max()))
        assert (gf == rf).all() and np.abs(gv - rv).max() < tol


@pytest.mark.parametrize("axis", ["gamma", "h"])
def test_bifurcation_fast_vs_strict(gpu_engine, axis):
    """Sweep over g (wave-uniform loops, lean recorded loop) and over h
    (per-lane loops): 1024 cells, 15 + 25 time units."""
    print("case: two-harmonic forced Duffing bifurcation over %s, fast vs strict" % axis, flush=True)
    s = _system(gpu_engine)
    if axis == "gamma":
        axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 1024, 0.05, 0.5)]
    else:
        axes = [ucuda.axis(ucuda.AXIS_STEP, 0, 1024, 0.004, 0.02)]
    args = (s, axes, [0.1, 0.1], VALS, 0.01, 15.0, 25.0)
    ref = ucuda.run_bif(*args, strict=True, max_peaks=32)
    rf, rp = ref["flags"].cpu().numpy(), ref["peaks"].cpu().numpy()
    assert (rf >= 2).sum() > 512
    for prec, tol in ((ucuda.PREC_F64, 1e-7), (ucuda.PREC_F32C, 1e-4)):
        g = ucuda.run_bif(*args, precision=prec, max_peaks=32)
        gf, gp = g["flags"].cpu().numpy(), g["peaks"].cpu().numpy()
        same = gf == rf
        both = same & (rf >= 1)
        k = int(rf[both].min()) if both.any() else 0
        err = np.abs(gp[:k, both] - rp[:k, both]) / np.maximum(1.0, np.abs(rp[:k, both]))
        cell_err = err.max(axis=0)
        print("precision %d: flags equal on %.4f, rel peak err max %.2e median %.2e, share <= %g: %.4f (%d peaks)"
              % (prec, same.mean(), cell_err.max(), np.median(cell_err), tol, (cell_err <= tol).mean(), k))
        assert k >= 1
        if prec == ucuda.PREC_F64:
            assert same.mean() > 0.999 and cell_err.max() < tol
        else:
            # float stages through the transient: cells sensitive to their
            # rounding drift further; the bulk stays at float precision
            assert same.mean() > 0.98 and (cell_err <= tol).mean() > 0.95 and np.median(cell_err) < 1e-5
