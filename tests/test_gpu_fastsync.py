"""Fast synchronisation (both modes) through the C ABI against the FP64
oracle. Strict FP64 keeps every state bit-identical to the oracle; the
reported log10 errors may differ in the last bit (the device's log10 and
glibc's round differently), so estimators 0 / 2 are compared at 1e-15
relative and estimators 1 / 3 (pass counts, times) exactly."""
import numpy as np
import pytest

import oracle
import ucuda

pytestmark = pytest.mark.gpu

LOR_K = [5.0, 0.0, 0.0]


def _check(g, r, estim):
    if estim in (1, 3):
        np.testing.assert_array_equal(g, r)
    else:
        np.testing.assert_allclose(g, r, rtol=1e-15, atol=0)


def test_fastsync_on_attractor(gpu_engine):
    s = ucuda.preset("lorenz", gpu_engine)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    vals = s.values()
    m0, s0 = [0.1, 0.1, 0.1], [3.0, -2.0, 20.0]
    strong = ([5.0, 5.0, 5.0], [20.0, 20.0, 20.0], 1.5)  # the backward pass needs K above Lorenz's contraction
    weak = ([1.0, 0.0, 0.0], [15.0, 15.0, 15.0], 1.0)
    for estim, typ, trs, (kf, kb, win) in ((2, 0, 1e-12, strong), (0, 0, 1e-12, strong), (1, 0, 1e-6, weak),
                                           (3, 0, 1e-2, strong), (2, 1, 1e-12, weak), (1, 1, 1e-6, weak),
                                           (3, 1, 1e-2, strong)):
        print("case: fast synchro on attractor, Lorenz, estim %d type %d" % (estim, typ), flush=True)
        kw = dict(h=0.01, iter_of_synchr=4, pre_scaller=2, t_max=4.0, transient_time=2.0, window=win,
                  type_of_synch=typ, error_estim=estim, fs_error_trs=trs)
        g = ucuda.run_fastsync(s, vals, m0, s0, kf, kb, mode=0, strict=True, **kw)
        n, r, traj = oracle.fs_attractor(step, 3, vals, m0, s0, kf, kb, range(200), threads=8, **kw)
        assert n == 200 and g["sync_error"].shape[0] == 200
        np.testing.assert_array_equal(g["traj"].cpu().numpy(), traj)
        _check(g["sync_error"].cpu().numpy(), r, estim)
        assert len(np.unique(r)) > 1


def test_fastsync_on_grid(gpu_engine):
    s = ucuda.preset("lorenz", gpu_engine)
    step = oracle.KrsStep(gpu_engine.codegen_scheme(s.vars, s.params, s.rhs), 3, 4)
    vals = s.values()
    m0, s0 = [0.1, 0.1, 0.1], [3.0, -2.0, 20.0]
    kf, kb = [1.0, 0.0, 0.0], [15.0, 15.0, 15.0]
    for estim, typ, swap in ((2, 0, False), (2, 0, True), (0, 1, False), (3, 0, True), (1, 1, False)):
        print("case: fast synchro on grid 14x14, Lorenz, estim %d type %d swap %d" % (estim, typ, swap), flush=True)
        kw = dict(h=0.01, iter_of_synchr=3, pre_scaller=1, transient_time=1.0, window=1.2, axis_x_var=0,
                  axis_y_var=2, axis_x=(-15.0, 15.0), axis_y=(5.0, 40.0), n_pts=14, grid_swap_master_slave=swap,
                  type_of_synch=typ, error_estim=estim,
                  fs_error_trs=1e-2 if estim == 3 else (1e-6 if estim == 1 else 1e-12))
        g = ucuda.run_fastsync(s, vals, m0, s0, kf, kb, mode=1, strict=True, **kw)
        r = oracle.fs_grid(step, 3, vals, m0, s0, kf, kb, range(196), threads=8, **kw)
        _check(g["sync_error"].cpu().numpy(), r, estim)
        assert len(np.unique(r)) > 1


def test_fastsync_fast_mode_and_errors(gpu_engine):
    print("case: fast synchro fast FP64 vs strict; request validation", flush=True)
    s = ucuda.preset("lorenz", gpu_engine)
    vals = s.values()
    kw = dict(h=0.01, iter_of_synchr=4, pre_scaller=2, t_max=4.0, transient_time=2.0, window=1.5, error_estim=2)
    kf, kb = [5.0, 5.0, 5.0], [20.0, 20.0, 20.0]
    a = ucuda.run_fastsync(s, vals, [0.1, 0.1, 0.1], [3.0, -2.0, 20.0], kf, kb, **kw)["sync_error"].cpu().numpy()
    b = ucuda.run_fastsync(s, vals, [0.1, 0.1, 0.1], [3.0, -2.0, 20.0], kf, kb, strict=True,
                           **kw)["sync_error"].cpu().numpy()
    print("fast vs strict: max |diff| %.2e over %d windows" % (np.abs(a - b).max(), len(a)))
    with pytest.raises(ucuda.UcError):
        ucuda.run_fastsync(s, vals, [0.1] * 3, [0.1] * 3, LOR_K, LOR_K, mode=1, axis_x_var=1, axis_y_var=1, n_pts=4)
    with pytest.raises(ucuda.UcError):
        ucuda.run_fastsync(s, vals, [0.1] * 3, [0.1] * 3, LOR_K, LOR_K, mode=0, window=0.0)
