"""The CSV writers against the reference's file layout, byte for byte.

tests/csv_writers.cpp calls every writer of u-cuda_amd/csrc/data_export.cpp
with fixed inputs; the expected text here is built independently from the
reference writers' layout (U-CUDA/data_export.cpp:16-58 bif1d, 96-152 DFT,
178-266 LLE1D / LS1D, 300-364 the 2-D helpers and Bif2D, 384-453 LLE2D /
LS2D, 473-522 basins, 553-657 fast synchronisation; param_value_at,
U-CUDA/data_export.h:239-243), with numbers printed as C++ streams print a
double at std::setprecision(15) (set_precision, U-CUDA/configCUDA.h), that
is printf's %.15g. The reference's writer is not built here: its
translation unit includes the GUI's session and plot headers.
"""
import math
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

VALUES = [10.0, 28.0, 8.0 / 3.0, -0.5, 1e-20]
IC = [0.1, 1.0 / 3.0, 2.5e10]
T_MAX, TRANSIENT, H, NT, EPS = 100.0, 50.5, 0.01, 0.1, 1e-4


def g(x):
    return "%.15g" % x


def vec(name, v):
    return "%s[%d] = { %s }\n" % (name, len(v), ", ".join(g(x) for x in v))


def param_value_at(i, n, lo, hi):
    return lo if n <= 1 else lo + (hi - lo) * float(i) / float(n - 1)


def par_or_var_line(pv):
    return {1: "Parameter estimation\n", 0: "Initial conditions estimation\n", 2: "Mixed: x=IC, y=parameter\n"}[pv]


def expected():
    e = {}
    e["bif1d_config"] = ("1D classical bifurcation\nParameter estimation\n" + vec("a", VALUES) + vec("X0", IC) +
                         "CT = %s\nTT = %s\nh = %s\ndecimator = 2\nindexVar for peakfinder = 1\n"
                         "indexPar for estimation = 2\nstart value = 0, stop value = 250\n" % (g(T_MAX), g(TRANSIENT),
                                                                                            g(H)))
    pk, tm = [12.345678901234567, -0.0], [0.7, 1e-7]
    p0, p3, p4 = (param_value_at(i, 5, 0.0, 250.0) for i in (0, 3, 4))
    e["bif1d_rows"] = ("%s, 0, 0\n" % g(p0) + "%s, 0, -1\n" % g(p3) +
                       "".join("%s, %s, %s\n" % (g(p4), g(a), g(b)) for a, b in zip(pk, tm)) +
                       "%s, 0, 0\n" % g(param_value_at(0, 1, 7.5, 250.0)))
    for w, name in ((0, "None"), (1, "Hanning"), (2, "Hamming")):
        e["dft1d_config_%d" % w] = ("1D bifurcation DFT\nParameter estimation\n" + vec("a", VALUES) + vec("X0", IC) +
                                    "CT = %s\nTT = %s\nh = %s\ndecimator = 1\nindexVar for DFT = 0\n"
                                    "indexPar for estimation = 3\nstart value = -1, stop value = 1\n"
                                    "n_freq = 3\nfreq_lo = 0.5, freq_hi = %s\nwindow = %s\n"
                                    % (g(T_MAX), g(TRANSIENT), g(H), g(123456789012345678.0), name))
    e["dft1d_header"] = "-1, 1\n0.5, %s\n" % g(123456789012345678.0)
    m = [1.0 / 7.0, 2e-300, 3.0, -4.5e15, 0.0, 1e21]
    e["dft1d_matrix"] = "".join(", ".join(g(x) for x in m[r * 3:r * 3 + 3]) + "\n" for r in range(2))
    for key, title in (("lle1d_config", "1D LLE"), ("ls1d_config", "1D LS")):
        e[key] = (title + "\nParameter estimation\n" + vec("a", VALUES) + vec("X0", IC) +
                  "CT = %s\nNT = %s\nTT = %s\nh = %s\neps = %s\nindexPar = 1\nstart value = 0.05, stop value = 2\n"
                  % (g(T_MAX), g(NT), g(TRANSIENT), g(H), g(EPS)))
    e["lle1d_rows"] = ("%s, %s\n" % (g(param_value_at(1, 3, 0.05, 2.0)), g(0.0713)) +
                       "%s, %s\n" % (g(param_value_at(2, 3, 0.05, 2.0)), g(-1.0 / 3.0)))
    ex = [0.0713, 1e-12, -5.394]
    e["ls1d_rows"] = "0.05, " + ", ".join(g(x) for x in ex) + "\n" + "2, " + g(ex[0]) + "\n"
    for pv in range(3):
        e["bif2d_config_%d" % pv] = ("2D bifurcation (DBSCAN)\n" + par_or_var_line(pv) + vec("a", VALUES) +
                                     vec("X0", IC) + "CT = %s\nTT = %s\nh = %s\ndecimator = 3\n"
                                     "eps_DBSCAN = 0.1\nindexVar for peakfinder = 2\nindices = 1, 2\n"
                                     "axis1: -2.5 .. 2.5\naxis2: 0.05 .. 20\nn_pts = 3x3\n"
                                     % (g(T_MAX), g(TRANSIENT), g(H)))
    common = "CT = %s\nNT = %s\nTT = %s\nh = %s\neps = %s\n" % (g(T_MAX), g(NT), g(TRANSIENT), g(H), g(EPS))
    e["lle2d_config"] = ("2D LLE\n" + par_or_var_line(1) + vec("a", VALUES) + vec("X0", IC) + common +
                         "indices = 2, 3\naxis1: 0.05 .. 2\naxis2: 1 .. 20\nn_pts = 512x512\n")
    e["ls2d_config"] = ("2D LS\n" + par_or_var_line(2) + vec("a", VALUES) + vec("X0", IC) + common +
                        "indices = 0, 3\naxis1: -2.5 .. 2.5\naxis2: 1 .. 20\nn_pts = 4x4\nexponents = 3\n")
    gv = [0.1, -1.0 / 3.0, 1e300, 7.0, 0.0, -0.0, 999.0, 2.5e-8, 1234567.0]
    e["grid"] = "".join(", ".join(g(x) for x in gv[r * 3:r * 3 + 3]) + "\n" for r in range(3))
    v = [0.1, 0.2, 0.3, 0.4, -0.1, -0.2, -0.3, -0.4, 1.0 / 3.0, 2.0 / 3.0, 1e-9, 5.0]
    e["ls2d_cells"] = "".join(", ".join(g(v[k * 4 + c]) for k in range(3)) + "\n" for c in range(4))
    e["basins_config"] = ("Basins of attraction\n" + vec("a", VALUES) + vec("X0", IC) +
                          "CT = %s\nTT = %s\nh = %s\ndecimator = 1\neps_DBSCAN = 0.25\n"
                          "indexVar for peakfinder = 0\nindexVar for estimation = 0, 1\n"
                          "start value_1 = -2.5, stop value_1 = 2.5\nstart value_2 = -2, stop value_2 = 2\n"
                          % (g(T_MAX), g(TRANSIENT), g(H)))
    ranges = "-2.5 2.5\n-2 2\n"
    e["basins_labels"] = ranges + "0, 1\n-1, 12\n"
    fin = [x if math.isfinite(x) else 999.0 for x in (math.nan, math.inf, -math.inf, 2.5)]
    e["basins_finite"] = ranges + "%s, %s\n%s, %s\n" % tuple(g(x) for x in fin)
    icm, ics, kf, kb = [1.0, 2.0, 3.0], [1.5, -2.0, 1.0 / 3.0], [0.5, 0.0, 0.0], [0.0, 0.25, 1e-3]

    def fs(mode, tos, estim):
        s = ("Fast Synchro (On Attractor)\n" if mode == 0 else "Fast Synchro (On Grid)\n")
        s += vec("a", VALUES) + vec("IC_master", icm) + vec("IC_slave", ics) + vec("k_forward", kf)
        s += vec("k_backward", kb)
        s += "h = %s\niter_of_synchr = 40\ndecimator = 2\nwindow = 5.5\n" % g(H)
        s += "type_of_synch = %s, error_estim = %s, fs_error_trs = 1e-06\n" % (
            "Bidirectional" if tos == 1 else "Unidirectional",
            {1: "# iters to reach FS_error_trs", 2: "RMS at last point"}.get(estim, "RMS on last iter"))
        return s
    e["fastsync_config_0"] = fs(0, 0, 1) + "CT = %s\nTT = %s\n" % (g(T_MAX), g(TRANSIENT))
    e["fastsync_config_1"] = (fs(1, 1, 2) + "TT = %s\naxis_x_var = 0, axis_y_var = 2\naxis_x: -20 .. 20\n"
                              "axis_y: 0 .. 50\nn_pts = 128x128\ngrid_swap_master_slave = 1\n" % g(TRANSIENT))
    e["fastsync_config_estim0"] = fs(0, 0, 0) + "CT = %s\nTT = %s\n" % (g(T_MAX), g(TRANSIENT))
    traj, err = [1.0, 2.0, 1.0 / 3.0, -4.0, 5e-5, 6.0], [0.125, 1e-17]
    e["fastsync_attractor"] = "x,y,x2,sync_error\n" + "".join(
        "".join(g(traj[i * 3 + j]) + "," for j in range(3)) + g(err[i]) + "\n" for i in range(2))
    hm = [1.0, 2.0 / 3.0, 40.0, -1.0]
    e["fastsync_grid"] = "-20 20\n0 50\n" + "%s,%s\n%s,%s\n" % tuple(g(x) for x in hm)
    return e


def test_every_csv_writer_matches_the_reference_layout(tmp_path):
    exe = str(tmp_path / "csv_writers")
    src = os.path.join(ROOT, "u-cuda_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", src,
                    os.path.join(ROOT, "tests", "csv_writers.cpp"), os.path.join(src, "data_export.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    out = tmp_path / "out"
    out.mkdir()
    subprocess.run([exe, str(out)], check=True, timeout=60)
    exp = expected()
    written = sorted(p.name for p in out.iterdir())
    assert written == sorted(exp), (written, sorted(exp))
    for name, text in exp.items():
        got = (out / name).read_text()
        assert got == text, (name, got, text)


def test_planted_format_slips_are_caught():
    """The expected text is not a copy of the writer's output: a slipped
    separator, precision or label in the builder fails the comparison."""
    e = expected()
    assert "2.66666666666667" in e["bif1d_config"]          # 15 significant digits
    assert "1.23456789012346e+17" in e["dft1d_header"]      # exponent form past 15 digits
    assert e["basins_finite"].splitlines()[2:] == ["999, 999", "999, 2.5"]
    assert e["grid"].splitlines()[1] == "7, 0, -0"
    assert "%.6g" % (8.0 / 3.0) not in e["bif1d_config"].replace("2.66666666666667", "")
