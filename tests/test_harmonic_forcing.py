"""Harmonic forcing in the fast modules.

Codegen: every cos / sin of the right-hand side whose argument is linear in
time, w * t + phi with w and phi free of the state and of t, becomes an
input hc[k] / hs[k] of the device source's uc_rhs_h, with uc_harm_coef
giving w[k] and phi[k]; the C text the oracle compiles and the strict
modules keep the calls.

Kernels: the basins integration rotates cos / sin of each phase to the
stage times with uc_sincos_small. Its text is cut from the kernel source
between its plain-C++ marks and compiled on the host with g++, and compared
with long-double sin / cos over the whole range it serves: the Taylor
branch (|d| <= 0.8), the reduction by 2 pi with halving and doubling above
it, and the NaN answer for inf, NaN and |d| beyond the reduction's cap.
"""
import math
import os
import re
import subprocess

import numpy as np

import ucuda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNEL = os.path.join(ROOT, "u-cuda_amd", "csrc", "kernels", "ensemble.hip")


def _harmonic_block(src):
    i = src.index("#define UC_HARMONICS")
    j = src.index("template <typename T, int K>")
    return src[i:j]


def test_duffing_forcing_becomes_an_input():
    eng = ucuda.Engine(device=-1)
    s = ucuda.preset("duffing", eng)
    b = _harmonic_block(s.source(0))
    assert "#define UC_HARMONICS 1\n" in b
    assert "    w[0] = a[3];\n    phi[0] = (T)0;\n" in b
    assert "dX[1] = -a[1] * X[1] + X[0] - (X[0] * X[0] * X[0]) + a[2] * hc[0];" in b
    # the uc_rhs the strict modules use and the oracle's C text keep the call
    assert "a[2] * cos(a[3] * t);" in s.source(0)
    assert "uc_rhs_h" not in s.source(1) and "cos(a[3] * t)" in s.source(1)


def test_linear_arguments_and_what_is_left_alone():
    eng = ucuda.Engine(device=-1)
    rhs = ["v", "-d*v - x^3 + g*cos(w*t) + g*sin(2*w*t + f) - cos(t/w) + sin(x*t) + cos(w)*t"
                " + sin(-(w*t) - f) + cos(f - w*t) + g*cos(w*t)"]
    s = ucuda.System(eng, ["x", "v"], ["d", "g", "w", "f"], rhs, time_symbol="t")
    b = _harmonic_block(s.source(0))
    assert "#define UC_HARMONICS 5\n" in b
    coef = re.findall(r"    (w|phi)\[(\d)\] = (.*);", b)
    assert coef == [("w", "0", "a[3]"), ("phi", "0", "(T)0"),
                    ("w", "1", "((T)2.0) * a[3]"), ("phi", "1", "a[4]"),
                    ("w", "2", "((T)1.0) / a[3]"), ("phi", "2", "(T)0"),
                    ("w", "3", "-a[3]"), ("phi", "3", "-a[4]"),
                    ("w", "4", "-a[3]"), ("phi", "4", "a[4]")]
    # a state-dependent argument and a time-free one stay calls; a repeated
    # harmonic reuses its input
    assert ("dX[1] = -a[1] * X[1] - (X[0] * X[0] * X[0]) + a[2] * hc[0] + a[2] * hs[1] - hc[2] + sin(X[0] * t)"
            " + cos(a[3]) * t + hs[3] + hc[4] + a[2] * hc[0];") in b


def test_unforced_systems_have_no_harmonics():
    eng = ucuda.Engine(device=-1)
    for name in ("lorenz", "rossler", "hyper_rossler", "chen"):
        b = _harmonic_block(ucuda.preset(name, eng).source(0))
        assert b.startswith("#define UC_HARMONICS 0\n") and "uc_rhs_h" not in b


def _sincos_source():
    text = open(KERNEL).read()
    a = text.index("// [plain C++ from here to the matching end mark, compiled on the host too]")
    b = text.index("// [end of plain C++]")
    return text[a:b].replace("__device__ __forceinline__", "inline")


DRIVER = r"""
#include <cmath>
#include <cstdio>
#include <cstdlib>
using std::fabs; using std::fma; using std::rint; using std::fmaf;
%s
int main(int argc, char** argv) {
    FILE* f = fopen(argv[1], "rb");
    FILE* o = fopen(argv[2], "wb");
    double d;
    while (fread(&d, sizeof d, 1, f) == 1) {
        double c, s;
        uc_sincos_small(d, c, s);
        float cf, sf;
        uc_sincos_small((float)d, cf, sf);
        double r[4] = {c, s, (double)cf, (double)sf};
        fwrite(r, sizeof r, 1, o);
    }
    return 0;
}
"""


def test_small_angle_sincos_against_long_double(tmp_path):
    src = tmp_path / "sc.cpp"
    src.write_text(DRIVER % _sincos_source())
    exe = str(tmp_path / "sc")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", str(src), "-o", exe], check=True,
                   capture_output=True, text=True)
    rng = np.random.default_rng(11)
    small = np.concatenate([rng.uniform(-0.8, 0.8, 200000), [0.0, -0.0, 1e-300, -5e-324, 0.8, -0.8],
                            np.nextafter([0.8, -0.8], 0.0), np.ldexp(1.0, np.arange(-60, 0))])
    large = np.concatenate([np.nextafter([0.8, -0.8], [2.0, -2.0]), rng.uniform(-10.0, 10.0, 100000),
                            rng.uniform(-1e4, 1e4, 50000), np.exp(rng.uniform(0.0, math.log(9e14), 20000)),
                            [math.pi, 2 * math.pi, 100 * math.pi, -1e3 * math.pi]])
    bad = np.array([np.inf, -np.inf, np.nan, 1e15, -2e15, 1e300])
    d = np.concatenate([small, large, bad])
    fin = tmp_path / "d.bin"
    fout = tmp_path / "r.bin"
    d.astype(np.float64).tofile(fin)
    subprocess.run([exe, str(fin), str(fout)], check=True, timeout=60)
    r = np.fromfile(fout, dtype=np.float64).reshape(-1, 4)
    ns, nl = len(small), len(large)

    import mpmath  # exact references for the sample (mpmath ships with this image's sympy stack)
    mpmath.mp.prec = 120
    ref = np.array([[float(mpmath.cos(mpmath.mpf(float(x)))), float(mpmath.sin(mpmath.mpf(float(x))))]
                    for x in d[:ns + nl:7]])
    got = r[:ns + nl:7]
    idx = np.arange(0, ns + nl, 7)
    ulp = np.finfo(np.float64).eps
    err = np.abs(got[:, :2] - ref)
    on_series = idx < ns
    # Taylor branch: within 2 ulps of 1; reduction branch: a few ulps of the reduced angle
    assert err[on_series].max() <= 2 * ulp, err[on_series].max()
    assert err[~on_series].max() <= 16 * ulp, err[~on_series].max()
    # float version on the float-rounded argument
    xf = d[:ns + nl:7].astype(np.float32).astype(np.float64)
    reff = np.array([[float(mpmath.cos(mpmath.mpf(x))), float(mpmath.sin(mpmath.mpf(x)))] for x in xf])
    okf = np.abs(xf) < 1e7
    errf = np.abs(got[okf][:, 2:] - reff[okf])
    assert errf[on_series[okf]].max() <= 2 * np.finfo(np.float32).eps, errf[on_series[okf]].max()
    assert errf[~on_series[okf]].max() <= 16 * np.finfo(np.float32).eps, errf[~on_series[okf]].max()
    # inf, NaN and beyond the reduction's cap: NaN out (double and float)
    assert np.isnan(r[ns + nl:]).all()
