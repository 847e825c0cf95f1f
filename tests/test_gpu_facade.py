"""The C++ drop-in facade (ParametricEngine with the reference's Request /
Result structs) returns exactly what the C ABI returns for the same request."""
import json
import os
import subprocess

import numpy as np
import pytest

import ucuda

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_matches_c_abi(gpu_engine):
    print("case: facade_demo (bif 16 rho, LLE2D 8x8, LS2D 4x4) vs C ABI", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    g = ucuda.run_bif(lor, [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.0, 250.0)], [0.1, 0.1, 0.1], lor.values(), 0.01,
                      20.0, 20.0, max_peaks=64)
    flags = g["flags"].cpu().numpy()
    assert res["bif_flags"] == flags.tolist()
    first = np.where(flags > 0, g["peaks"].cpu().numpy()[0], 0.0)
    assert res["bif_first_peaks"] == first.tolist()
    ros = ucuda.preset("rossler", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 8, 1.0, 20.0)]
    l = ucuda.run_lyap(ros, axes, [1.0, 1.0, 1.0], ros.values(), 0.01, 5.0, 50.0, 0.1)["values"].cpu().numpy()[0]
    # separate processes compile the module separately: equal to the last bits
    np.testing.assert_allclose(res["lle2d"], l, rtol=1e-12, atol=0)
    axes4 = [ucuda.axis(ucuda.AXIS_PARAM, 2, 4, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 4, 1.0, 20.0)]
    ls = ucuda.run_lyap(ros, axes4, [1.0, 1.0, 1.0], ros.values(), 0.01, 5.0, 50.0, 0.1, spectrum=True)
    s = ls["values"].cpu().numpy()
    np.testing.assert_allclose(res["ls2d_sum"], s[0] + s[1] + s[2], rtol=0, atol=1e-15)


def _rows(path):
    with open(path) as f:
        return [line.rstrip("\n") for line in f]


def test_facade_csv_files(gpu_engine, tmp_path):
    print("case: facade_demo writes the reference's CSV files (bif1d, lle2d, ls2d, basins)", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    g15 = lambda v: "%.15g" % v  # the reference prints with 15 significant digits
    # bifurcation config: exact lines
    assert _rows(tmp_path / "bif1d.csv_config.csv") == [
        "1D classical bifurcation", "Parameter estimation",
        "a[4] = { 0.5, 10, 28, %s }" % g15(8.0 / 3.0), "X0[3] = { 0.1, 0.1, 0.1 }",
        "CT = 20", "TT = 20", "h = 0.01", "decimator = 1", "indexVar for peakfinder = 0",
        "indexPar for estimation = 2", "start value = 0, stop value = 250"]
    # bifurcation rows: "param, peak, interval" per peak, "param, 0, -1" fixed point, "param, 0, 0" none
    lor = ucuda.preset("lorenz", gpu_engine)
    g = ucuda.run_bif(lor, [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.0, 250.0)], [0.1, 0.1, 0.1], lor.values(), 0.01,
                      20.0, 20.0, max_peaks=64)
    flags = g["flags"].cpu().numpy()
    pk, iv = g["peaks"].cpu().numpy(), g["intervals"].cpu().numpy()
    expect = []
    for c in range(16):
        p = g15(0.0 + 250.0 * c / 15)
        if flags[c] > 0:
            expect += ["%s, %s, %s" % (p, g15(pk[k, c]), g15(iv[k, c])) for k in range(flags[c])]
        else:
            expect.append("%s, 0, %s" % (p, "-1" if flags[c] == -1 else "0"))
    assert _rows(tmp_path / "bif1d.csv") == expect
    # LLE 2-D grid: 8 rows of 8 values, row = iy (same numbers as the JSON summary)
    grid = np.array([[float(x) for x in r.split(", ")] for r in _rows(tmp_path / "lle2d.csv")])
    assert grid.shape == (8, 8)
    grid = np.where(np.isnan(grid), 999.0, grid)  # the summary writes diverged cells as 999
    np.testing.assert_allclose(grid.ravel(), res["lle2d"], rtol=1e-14, atol=0)
    cfg = _rows(tmp_path / "lle2d.csv_config.csv")
    assert cfg[:2] == ["2D LLE", "Parameter estimation"] and "n_pts = 8x8" in cfg
    # LS 2-D: one row per cell with its 3 exponents
    ls = np.array([[float(x) for x in r.split(", ")] for r in _rows(tmp_path / "ls2d.csv")])
    assert ls.shape == (16, 3)
    np.testing.assert_allclose(ls.sum(axis=1), res["ls2d_sum"], rtol=1e-13, atol=1e-13)
    assert "exponents = 3" in _rows(tmp_path / "ls2d.csv_config.csv")
    # basins: range lines, then the label grid; _1, _2, _3 and _config alongside
    b = _rows(tmp_path / "basins.csv")
    assert b[:2] == ["-2.5 2.5", "-2.5 2.5"]
    labels = [int(x) for r in b[2:] for x in r.split(", ")]
    assert labels == res["basins"]
    for suffix in ("_1.csv", "_2.csv", "_3.csv", "_config.csv"):
        assert (tmp_path / ("basins.csv" + suffix)).exists(), suffix
    assert _rows(tmp_path / "basins.csv_config.csv")[0] == "Basins of attraction"


def test_facade_bif2d(gpu_engine, tmp_path):
    print("case: facade run_bifurcation_2d (Lorenz rho x sigma 12x12) vs C ABI run_bif_periods + CSV", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 12, 20.0, 250.0), ucuda.axis(ucuda.AXIS_PARAM, 1, 12, 5.0, 16.0)]
    per = ucuda.run_bif_periods(lor, axes, [0.1, 0.1, 0.1], lor.values(), 0.01, 30.0, 30.0, eps_dbscan=0.1)
    per = per.cpu().numpy()
    expect = np.where(per < 0, -1.0, per.astype(np.float64))
    # the facade process compiles with the system hipRTC, this one with torch's bundled hipRTC: chaotic
    # cells can differ (last-bit differences grow), the map as a whole agrees
    got = np.array(res["bif2d"])
    print("facade vs C ABI periods equal on %.3f of the cells" % (got == expect).mean())
    assert (got == expect).mean() > 0.85
    grid = np.array([[float(x) for x in r.split(", ")] for r in _rows(tmp_path / "bif2d.csv")])
    np.testing.assert_array_equal(grid.ravel(), got)
    cfg = _rows(tmp_path / "bif2d.csv_config.csv")
    assert cfg[:2] == ["2D bifurcation (DBSCAN)", "Parameter estimation"]
    assert "eps_DBSCAN = 0.1" in cfg and "indices = 2, 1" in cfg and "n_pts = 12x12" in cfg


def test_facade_dft(gpu_engine, tmp_path):
    print("case: facade run_dft_1d (Lorenz 16 rho, 32 freqs) CSV files and flags vs C ABI", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    g = ucuda.run_dft(lor, [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.5, 250.0)], [0.1, 0.1, 0.1], lor.values(), 0.01,
                      10.0, 20.0, n_freq=32, freq_lo=0.0, freq_hi=5.0, window_type=1)
    assert res["dft_flags"] == g["flags"].cpu().numpy().tolist()
    for name in ("dft1d_AkCOS.csv", "dft1d_BkSIN.csv"):
        rows = _rows(tmp_path / name)
        assert rows[:2] == ["0.5, 250", "0, 5"]
        m = np.array([[float(x) for x in r.split(", ")] for r in rows[2:]])
        assert m.shape == (16, 32)
        assert (m[0] == -1.0).all()  # rho = 0.5 settles on the origin: fixed point, every coefficient -1
    cfg = _rows(tmp_path / "dft1d_config.csv")
    assert cfg[0] == "1D bifurcation DFT" and "n_freq = 32" in cfg and "window = Hanning" in cfg
    assert "freq_lo = 0, freq_hi = 5" in cfg and "indexVar for DFT = 0" in cfg


def test_facade_fastsync(gpu_engine, tmp_path):
    print("case: facade run_fastsync (grid 10x10 and on attractor) CSV files vs C ABI", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    kw = dict(h=0.01, iter_of_synchr=3, transient_time=1.0, axis_x_var=0, axis_y_var=2, axis_x=(-15.0, 15.0),
              axis_y=(5.0, 40.0), n_pts=10)
    g = ucuda.run_fastsync(lor, lor.values(), [0.1, 0.1, 0.1], [3.0, -2.0, 20.0], [1.0, 0.0, 0.0], [15.0] * 3,
                           mode=1, window=1.2, **kw)["sync_error"].cpu().numpy()
    # facade (system hipRTC) and this process (torch's hipRTC) differ in the last bits of fast-mode code
    np.testing.assert_allclose(res["fs_grid"], g, atol=1e-6)
    rows = _rows(tmp_path / "fastsync_grid.csv")
    assert rows[:2] == ["-15 15", "5 40"] and len(rows) == 12
    np.testing.assert_allclose([float(v) for r in rows[2:] for v in r.split(",")], res["fs_grid"], rtol=1e-14)
    cfg = _rows(tmp_path / "fastsync_grid.csv_config.csv")
    assert cfg[0] == "Fast Synchro (On Grid)" and "IC_master[3] = { 0.1, 0.1, 0.1 }" in cfg
    assert "n_pts = 10x10" in cfg and "grid_swap_master_slave = 0" in cfg
    assert "type_of_synch = Unidirectional, error_estim = RMS at last point, fs_error_trs = 1e-12" in cfg
    a = _rows(tmp_path / "fastsync_attr.csv")
    assert a[0] == "x,y,z,sync_error" and len(a) == 1 + 200 // 4
    assert len(res["fs_attr"]) == 50
    assert _rows(tmp_path / "fastsync_attr.csv_config.csv")[0] == "Fast Synchro (On Attractor)"


def test_facade_continuation(gpu_engine, tmp_path):
    print("case: facade run_lle_1d with continuation (fold x' = r + x - x^3, both directions) vs C ABI", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    s = ucuda.System(gpu_engine, ["x"], ["r"], ["r+x-x^3"])
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 1, 9, -1.0, 1.0)]
    for key, rev in (("cont_fwd", False), ("cont_rev", True)):
        g = ucuda.run_lyap_cont(s, axes, [0.0], [0.5, 0.0], 0.01, 20.0, 20.0, 0.1, reverse=rev)["values"].cpu().numpy()[0]
        np.testing.assert_allclose(res[key], g, atol=1e-9)
    assert res["cont_rev_flag"] == 1
    assert abs(res["cont_fwd"][3] - res["cont_rev"][5]) > 0.5  # r = -0.25 on different branches


def test_facade_krs_body_only_dopri78(gpu_engine):
    """A session that passes only its step body (the DOPRI78 text of
    codegen_scheme, no right-hand side) gets that step integrated: the
    bifurcation sweep and the reference's finite-difference LLE equal the
    oracle stepping the same body (rho in [5, 20]: no chaos, so the fast
    mode's FMA contraction stays at rounding level)."""
    import oracle
    print("case: facade krs_body-only DOPRI78 bif 1-D and LLE 1-D vs the oracle", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    body = gpu_engine.codegen_scheme(lor.vars, lor.params, lor.rhs, ucuda.SCHEME_DOPRI78)
    step = oracle.KrsStep(body, 3, 4)
    axes = [dict(kind=oracle.AXIS_PARAM, index=2, n_pts=16, lo=5.0, hi=20.0, log_scale=0)]
    job = oracle.make_job(3, axes, [0.1, 0.1, 0.1], [0.5, 10.0, 28.0, 8.0 / 3.0], 0.01, 2.0, 10.0, max_peaks=32,
                          NT=0.1)
    ref = oracle.bif_cells(step, job, range(16))
    assert res["krs_bif_flags"] == ref["flags"].tolist()
    for c in range(16):
        n = ref["flags"][c]
        if n > 0:
            np.testing.assert_allclose(res["krs_bif_peaks"][c], ref["peaks"][:n, c], rtol=1e-9, atol=1e-12)
    lle = oracle.lle_fd_cells(step, job, range(16))["values"][0]
    np.testing.assert_allclose(res["krs_lle"], lle, rtol=0, atol=1e-8)
    assert res["krs_snapshot_index"] == 2


def test_facade_cancel_between_pieces(gpu_engine):
    """A 1024 x 1024 LLE map runs in four pieces of the grid; a cancel set
    after the first piece reported progress ends the run before the last:
    the result says cancelled and progress stopped short of 1."""
    print("case: facade cancel between pieces (Rossler LLE 1024^2)", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo"), "--cancel"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    print(res)
    assert res["cancelled"] == 1 and res["ok"] == 0 and res["error"] == "Cancelled by user"
    assert 0.0 < res["seen"] < 1.0 and res["progress"] < 1.0
