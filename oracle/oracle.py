"""ctypes glue for the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The engine never imports it.

`KrsStep` restates krs_cpu (U-CUDA/krs_cpu.cpp:21-99, 293-332, 405-479) for
Linux: the step body is wrapped in the reference prelude with numb = double,
compiled by g++ with -ffp-contract=off into a cached shared object keyed by
FNV-1a over (body, nx, nv, prelude version, sizeof numb), and `krs_step` is
loaded with dlopen. `SystemFuncs` does the same for the C right-hand side and
tangent functions generated by the engine's codegen.
"""
import ctypes as C
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PRELUDE_VERSION = 3          # kPreludeVersion, U-CUDA/krs_cpu.cpp:274
NUMB_BYTES = 8               # numb = double

AXIS_NONE, AXIS_PARAM, AXIS_VAR, AXIS_STEP = 0, 1, 2, 3


class OcFs(C.Structure):
    _fields_ = [("n_vars", C.c_int), ("a", C.POINTER(C.c_double)), ("h", C.c_double),
                ("iter_of_synchr", C.c_int), ("pre_scaller", C.c_int), ("type_of_synch", C.c_int),
                ("error_estim", C.c_int), ("fs_error_trs", C.c_double), ("k_forward", C.POINTER(C.c_double)),
                ("k_backward", C.POINTER(C.c_double)), ("n_iters", C.c_int)]


class OcJob(C.Structure):
    _fields_ = [("n_vars", C.c_int), ("n_values", C.c_int),
                ("ic", C.POINTER(C.c_double)), ("params", C.POINTER(C.c_double)),
                ("n_axes", C.c_int),
                ("axis_kind", C.c_int * 2), ("axis_index", C.c_int * 2), ("axis_n", C.c_int * 2),
                ("axis_log", C.c_int * 2), ("axis_lo", C.c_double * 2), ("axis_hi", C.c_double * 2),
                ("h", C.c_double), ("transient_time", C.c_double), ("t_max", C.c_double),
                ("max_value", C.c_double),
                ("ic_seed", C.c_ulonglong), ("ic_lo", C.c_double), ("ic_hi", C.c_double),
                ("writable_var", C.c_int), ("pre_scaller", C.c_int), ("max_peaks", C.c_int),
                ("NT", C.c_double), ("eps", C.c_double), ("probe_seed", C.c_ulonglong)]


_lib = None
_P = C.c_void_p
_LL = C.POINTER(C.c_longlong)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.oc_rng_init.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_ulonglong, C.POINTER(C.c_ulonglong)]
        L.oc_rng_uniform.argtypes = [C.POINTER(C.c_ulonglong)]
        L.oc_rng_uniform.restype = C.c_float
        L.oc_axis_values.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _D]
        L.oc_cell_setup.argtypes = [C.POINTER(OcJob), C.c_longlong, _D, _D, _D]
        L.oc_loop_model.argtypes = [_P, _D, _D, C.c_double, C.c_longlong, C.c_int, C.c_int, C.c_int,
                                    C.c_double, _D]
        L.oc_peak_finder.argtypes = [_D, C.c_longlong, _D, _D, C.c_double]
        L.oc_dbscan_count.argtypes = [_D, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.oc_dft_window.argtypes = [C.c_int, C.c_int, _D]
        L.oc_dft_block.argtypes = [_D, C.c_int, _D, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, _D, _D]
        L.oc_dft_cells.argtypes = [_P, C.POINTER(OcJob), _LL, C.c_longlong, C.c_int, C.c_double, C.c_double, C.c_int,
                                   C.c_int, _D, _D, _I, C.c_int]
        L.oc_dft_hsweep_cells.argtypes = L.oc_dft_cells.argtypes
        L.oc_cont_bif.argtypes = [_P, C.POINTER(OcJob), C.c_int, C.c_int, _I, _D, _D, _D, _I, C.c_int]
        L.oc_cont_dft.argtypes = [_P, C.POINTER(OcJob), C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                  _D, _D, _I, C.c_int]
        L.oc_cont_lyap.argtypes = [_P, C.POINTER(OcJob), C.c_int, C.c_int, _D, _I, C.c_int]
        L.oc_fs_attractor.argtypes = [_P, C.POINTER(OcFs), _D, _D, C.c_double, C.c_double, C.c_double, _LL,
                                      C.c_longlong, _D, _D, C.c_int]
        L.oc_fs_attractor.restype = C.c_longlong
        L.oc_fs_grid.argtypes = [_P, C.POINTER(OcFs), _D, _D, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _LL, C.c_longlong, _D,
                                 C.c_int]
        L.oc_cgs.argtypes = [_D, _D, C.c_int, _D]
        L.oc_mgs.argtypes = [_D, C.c_int, C.c_int, _D]
        L.oc_rk4.argtypes = [_P, _D, _D, C.c_double, C.c_int]
        L.oc_rk4_tan.argtypes = [_P, _P, _D, _D, C.c_int, _D, C.c_double, C.c_int]
        L.oc_bif_cells.argtypes = [_P, C.POINTER(OcJob), _LL, C.c_longlong, _I, _D, _D, _D, C.c_int, C.c_int]
        L.oc_bif_cells_noisy.argtypes = [_P, C.POINTER(OcJob), _LL, C.c_longlong, _I, _D, _D, _D, C.c_int, C.c_int,
                                         C.c_double, C.c_ulonglong, C.c_int]
        L.oc_lle_fd_cells.argtypes = [_P, C.POINTER(OcJob), _LL, C.c_longlong, _D, _I, C.c_int]
        L.oc_ls_fd_cells.argtypes = [_P, C.POINTER(OcJob), _LL, C.c_longlong, _D, _I, C.c_int]
        L.oc_lyap_var_cells.argtypes = [_P, _P, C.c_int, C.POINTER(OcJob), _LL, C.c_longlong, _D, _I, C.c_int]
        L.oc_lyap_var_cells_noisy.argtypes = [_P, _P, C.c_int, C.POINTER(OcJob), _LL, C.c_longlong, _D, _I, C.c_int,
                                              C.c_double, C.c_ulonglong]
        _lib = L
    return _lib


def _cache_root():
    root = os.path.join(tempfile.gettempdir(), "ucuda-krs")
    os.makedirs(root, exist_ok=True)
    return root


def fnv1a_key(body, nx, nv):
    """hash_key, U-CUDA/krs_cpu.cpp:318-332 (FNV-1a over body, nx, nv, version, sizeof numb)."""
    h = 1469598103934665603
    data = body.encode() + int(nx).to_bytes(4, "little", signed=True) + int(nv).to_bytes(4, "little", signed=True) \
        + PRELUDE_VERSION.to_bytes(4, "little") + NUMB_BYTES.to_bytes(4, "little")
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def check_indices(body, nx, nv):
    """krs_cpu_check_indices, U-CUDA/krs_cpu.cpp:21-99: constant X[k] / a[k]
    subscripts must be in range. Returns a list of (line, message)."""
    diags = []
    i, line, n = 0, 1, len(body)
    while i < n:
        c = body[i]
        if c == "\n":
            line += 1
            i += 1
            continue
        if body.startswith("//", i):
            while i < n and body[i] != "\n":
                i += 1
            continue
        if body.startswith("/*", i):
            i += 2
            while i + 1 < n and not body.startswith("*/", i):
                if body[i] == "\n":
                    line += 1
                i += 1
            i = i + 2 if i + 1 < n else n
            continue
        if c in "\"'":
            q = c
            i += 1
            while i < n and body[i] != q:
                if body[i] == "\\":
                    i += 1
                elif body[i] == "\n":
                    line += 1
                i += 1
            i += 1
            continue
        if not (c.isalpha() or c == "_"):
            i += 1
            continue
        s = i
        while i < n and (body[i].isalnum() or body[i] == "_"):
            i += 1
        ident = body[s:i]
        if ident not in ("X", "a"):
            continue
        m = re.match(r"[ \t]*\[[ \t]*(\d+)[ \t]*\]", body[i:])
        if not m:
            continue
        idx = int(m.group(1))
        if ident == "X" and not (0 <= idx < nx):
            diags.append((line, "X[%d]: the system has %d variables" % (idx, nx)))
        if ident == "a" and not (0 <= idx < nv):
            diags.append((line, "a[%d]: a[] has %d entries" % (idx, nv)))
    return diags


def make_source(body, nx):
    """make_source, U-CUDA/krs_cpu.cpp:293-316 (Linux export, numb = double)."""
    return ("#include <cmath>\n#include <cstdlib>\nusing std::abs;\ntypedef double numb;\n"
            "static inline numb min(numb x, numb y) { return x < y ? x : y; }\n"
            "static inline numb max(numb x, numb y) { return x > y ? x : y; }\n"
            "#define AMOUNTOFX %d\n"
            "static const numb pi    = (numb)3.1415926535897932384626433832795;\n"
            "static const numb euler = (numb)2.7182818284590452353602874713527;\n"
            "extern \"C\" __attribute__((visibility(\"default\")))\n"
            "void krs_step(numb* X, const numb* a, numb h) {\n#line 1 \"krs\"\n%s\n}\n") % (nx, body)


def _compile(source, suffix, compiler, key):
    d = os.path.join(_cache_root(), "%016x" % key)
    so = os.path.join(d, "unit.so")
    if not os.path.exists(so):
        os.makedirs(d, exist_ok=True)
        src = os.path.join(d, "unit" + suffix)
        with open(src, "w") as f:
            f.write(source)
        tmp = so + ".%d.tmp" % os.getpid()
        subprocess.run([compiler, "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                        "-o", tmp, src, "-lm"], check=True, capture_output=True)
        os.replace(tmp, so)
    return C.CDLL(so)


class KrsStep:
    """A KRS step body compiled to native `krs_step(numb*, const numb*, numb)`."""

    def __init__(self, body, n_vars, n_values):
        diags = check_indices(body, n_vars, n_values)
        if diags:
            raise ValueError("CPU KRS: " + "; ".join("line %d: %s" % d for d in diags))
        self.so = _compile(make_source(body, n_vars), ".cpp", "g++", fnv1a_key(body, n_vars, n_values))
        self.fn = self.so.krs_step
        self.ptr = C.cast(self.fn, C.c_void_p)
        self.n_vars = n_vars

    def step(self, x, a, h, n_steps=1):
        xa = (C.c_double * len(x))(*x)
        aa = (C.c_double * len(a))(*a)
        self.fn.argtypes = [_D, _D, C.c_double]
        for _ in range(n_steps):
            self.fn(xa, aa, float(h))
        return list(xa)


class SystemFuncs:
    """uc_rhs / uc_tangent from the engine codegen's C dialect, compiled by gcc."""

    def __init__(self, c_source):
        src = "#include <math.h>\n" + c_source + (
            "\n__attribute__((visibility(\"default\"))) void oc_sys_rhs(const double* X, const double* a, double t,"
            " double* dX) { uc_rhs(X, a, t, dX); }\n"
            "__attribute__((visibility(\"default\"))) void oc_sys_tan(const double* X, const double* a, double t,"
            " const double* V, double* JV, int K) { uc_tangent(X, a, t, V, JV, K); }\n")
        key = int(hashlib.sha1(src.encode()).hexdigest()[:16], 16)
        self.so = _compile(src, ".c", "gcc", key)
        self.rhs = C.cast(self.so.oc_sys_rhs, C.c_void_p)
        self.tan = C.cast(self.so.oc_sys_tan, C.c_void_p)


def make_job(n_vars, axes, ic, values, h, transient_time, t_max, max_value=1e6, writable_var=0, pre_scaller=1,
             max_peaks=256, NT=1.0, eps=1e-4, probe_seed=1234567891, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
    j = OcJob()
    ic = list(ic) if ic is not None else [0.0] * n_vars
    j._ic = (C.c_double * max(1, len(ic)))(*ic)
    j._params = (C.c_double * len(values))(*values)
    j.n_vars, j.n_values = n_vars, len(values)
    j.ic, j.params = j._ic, j._params
    j.n_axes = len(axes)
    for i, ax in enumerate(axes):
        j.axis_kind[i], j.axis_index[i], j.axis_n[i] = ax["kind"], ax["index"], ax["n_pts"]
        j.axis_log[i], j.axis_lo[i], j.axis_hi[i] = int(ax["log_scale"]), ax["lo"], ax["hi"]
    j.h, j.transient_time, j.t_max, j.max_value = h, transient_time, t_max, max_value
    j.ic_seed, j.ic_lo, j.ic_hi = ic_seed, ic_lo, ic_hi
    j.writable_var, j.pre_scaller, j.max_peaks = writable_var, pre_scaller, max_peaks
    j.NT, j.eps, j.probe_seed = NT, eps, probe_seed
    return j


def _cells(cells):
    cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int64))
    return cells, cells.ctypes.data_as(_LL)


def bif_cells(step, job, cells, traj_len=0, threads=1, noise=0.0, noise_seed=0, noise_every=100):
    """Bifurcation cells; noise > 0 adds a relative state kick of that size
    after every noise_every-th step (rounding model, see
    lyap_rounding_spread)."""
    cells, cp = _cells(cells)
    n = len(cells)
    mp = max(1, job.max_peaks)
    flags = np.zeros(n, dtype=np.int32)
    peaks = np.full((mp, n), np.nan)
    intervals = np.full((mp, n), np.nan)
    traj = np.full((traj_len, n), np.nan) if traj_len else None
    lib().oc_bif_cells_noisy(step.ptr, C.byref(job), cp, n, flags.ctypes.data_as(_I), peaks.ctypes.data_as(_D),
                             intervals.ctypes.data_as(_D), traj.ctypes.data_as(_D) if traj is not None else None,
                             traj_len, threads, C.c_double(noise), C.c_ulonglong(noise_seed), int(noise_every))
    return dict(flags=flags, peaks=peaks, intervals=intervals, traj=traj)


def bif_cells_phased(step, job, job_rec, cells, noise, noise_rec, noise_seed, threads=1, noise_every=100):
    """Bifurcation cells whose transient runs `job` with relative kicks of
    size `noise` and whose recorded run uses `job_rec`'s parameters with
    kicks of size `noise_rec` (a transient in one precision, the record in
    another)."""
    cells, cp = _cells(cells)
    n = len(cells)
    mp = max(1, job.max_peaks)
    flags = np.zeros(n, dtype=np.int32)
    peaks = np.full((mp, n), np.nan)
    intervals = np.full((mp, n), np.nan)
    L = lib()
    L.oc_bif_cells_phased.argtypes = [_P, C.POINTER(OcJob), C.POINTER(OcJob), _LL, C.c_longlong, _I, _D, _D, _D,
                                      C.c_int, C.c_int, C.c_double, C.c_double, C.c_ulonglong, C.c_int]
    L.oc_bif_cells_phased(step.ptr, C.byref(job), C.byref(job_rec), cp, n, flags.ctypes.data_as(_I),
                          peaks.ctypes.data_as(_D), intervals.ctypes.data_as(_D), None, 0, threads,
                          float(noise), float(noise_rec), noise_seed, int(noise_every))
    return dict(flags=flags, peaks=peaks, intervals=intervals)


def lle_fd_cells(step, job, cells, threads=1):
    cells, cp = _cells(cells)
    out = np.zeros((1, len(cells)))
    flags = np.zeros(len(cells), dtype=np.int32)
    lib().oc_lle_fd_cells(step.ptr, C.byref(job), cp, len(cells), out.ctypes.data_as(_D), flags.ctypes.data_as(_I),
                          threads)
    return dict(values=out, flags=flags)


def ls_fd_cells(step, job, cells, threads=1):
    cells, cp = _cells(cells)
    out = np.zeros((job.n_vars, len(cells)))
    flags = np.zeros(len(cells), dtype=np.int32)
    lib().oc_ls_fd_cells(step.ptr, C.byref(job), cp, len(cells), out.ctypes.data_as(_D), flags.ctypes.data_as(_I),
                         threads)
    return dict(values=out, flags=flags)


def lyap_var_cells(funcs, K, job, cells, threads=1):
    cells, cp = _cells(cells)
    out = np.zeros((K, len(cells)))
    flags = np.zeros(len(cells), dtype=np.int32)
    lib().oc_lyap_var_cells(funcs.rhs, funcs.tan, K, C.byref(job), cp, len(cells), out.ctypes.data_as(_D),
                            flags.ctypes.data_as(_I), threads)
    return dict(values=out, flags=flags)


def lyap_var_cells_noisy(funcs, K, job, cells, noise, noise_seed, threads=1):
    """Variational exponents with a relative state kick of size `noise` after
    every 100 transient steps and every NT block; used to find the cells whose
    finite-time exponent survives rounding-sized noise."""
    cells, cp = _cells(cells)
    out = np.zeros((K, len(cells)))
    flags = np.zeros(len(cells), dtype=np.int32)
    lib().oc_lyap_var_cells_noisy(funcs.rhs, funcs.tan, K, C.byref(job), cp, len(cells), out.ctypes.data_as(_D),
                                  flags.ctypes.data_as(_I), threads, C.c_double(noise), C.c_ulonglong(noise_seed))
    return dict(values=out, flags=flags)


# Relative rounding sizes of the two engine precisions, used as state noise
# (per 100 transient steps / per NT block) and as parameter perturbation:
# FP64 ~ a few ulp of double, FP32-compensated ~ one float rounding (the
# stage arithmetic and the parameters are float there).
ROUNDING = {"f64": (1e-15, 1e-15), "f32": (1e-7, 6e-8)}
_PERTURB_SIGNS = ((1, -1, 1, -1, 1), (-1, 1, -1, 1, -1), (1, 1, -1, -1, 1))


def perturbed_job(job, dp, signs):
    """Copy of `job` with the axis ends (lo0, hi0, lo1, hi1) and the base
    values scaled by (1 + dp * sign)."""
    axes = [dict(kind=job.axis_kind[i], index=job.axis_index[i], n_pts=job.axis_n[i], log_scale=bool(job.axis_log[i]),
                 lo=job.axis_lo[i] * (1.0 + dp * signs[2 * i]), hi=job.axis_hi[i] * (1.0 + dp * signs[2 * i + 1]))
            for i in range(job.n_axes)]
    values = [job._params[i] * (1.0 + dp * signs[4]) for i in range(job.n_values)]
    return make_job(job.n_vars, axes, [job._ic[i] for i in range(job.n_vars)], values, job.h, job.transient_time,
                    job.t_max, max_value=job.max_value, writable_var=job.writable_var, pre_scaller=job.pre_scaller,
                    max_peaks=job.max_peaks, NT=job.NT, eps=job.eps, probe_seed=job.probe_seed, ic_seed=job.ic_seed,
                    ic_lo=job.ic_lo, ic_hi=job.ic_hi)


def lyap_rounding_spread(funcs, K, job, cells, base, precision, threads=1):
    """Largest change of the variational exponents `base` over three runs
    that each add state noise and a parameter perturbation of the given
    precision's rounding size (ROUNDING), with distinct seeds and signs. A
    cell whose spread is small has a finite-time exponent that any
    implementation rounding at that size must reproduce; chaotic cells move
    by ~1e-2."""
    noise, dp = ROUNDING[precision]
    spread = np.zeros_like(np.asarray(base, dtype=np.float64))
    for k, signs in enumerate(_PERTURB_SIGNS):
        r = lyap_var_cells_noisy(funcs, K, perturbed_job(job, dp, signs), cells, noise, 1000 + k, threads=threads)
        spread = np.maximum(spread, np.abs(r["values"] - base))
    return spread


def bif_rounding_stable(step, job, cells, base, precision, rel, min_peaks=9, threads=1):
    """Cells of `base` (bif_cells output) with more than min_peaks - 1 peaks
    whose peak count and values (within `rel` of the largest |peak|) survive
    three runs with state noise and parameter perturbations of the given
    precision's rounding size (ROUNDING). precision "f32_rec64" models a
    float transient followed by a double recorded window: float-sized noise
    and parameter perturbations in the transient, double-sized noise and
    the exact parameters in the record."""
    n = base["flags"]
    stable = n >= min_peaks
    for k, signs in enumerate(_PERTURB_SIGNS):
        if precision == "f32_rec64":
            noise, dp = ROUNDING["f32"]
            r = bif_cells_phased(step, perturbed_job(job, dp, signs), job, cells, noise, ROUNDING["f64"][0],
                                 2000 + k, threads=threads)
        else:
            noise, dp = ROUNDING[precision]
            r = bif_cells(step, perturbed_job(job, dp, signs), cells, threads=threads, noise=noise,
                          noise_seed=2000 + k)
        for c in np.nonzero(stable)[0]:
            m = n[c]
            if r["flags"][c] != m:
                stable[c] = False
                continue
            d = np.abs(base["peaks"][:m, c] - r["peaks"][:m, c]).max()
            stable[c] = bool(d <= rel * max(1.0, np.abs(base["peaks"][:m, c]).max()))
    return stable


def dbscan_count(peaks, eps):
    """The reference's per-cell peak clustering (2-D bifurcation period)."""
    a = np.ascontiguousarray(np.asarray(peaks, dtype=np.float64))
    n = len(a)
    lab = (C.c_int * max(1, n))()
    sl = (C.c_int * max(1, n))()
    return lib().oc_dbscan_count(a.ctypes.data_as(_D), n, float(eps), lab, sl)


def bif_periods(step, job, cells, eps, threads=1):
    """2-D bifurcation periods: every peak of the recorded window, then the
    per-cell DBSCAN count (-1 / 0 / 1 for -1 / 0 / 1 peaks)."""
    iters = int(np.ceil(job.t_max / job.h / job.pre_scaller))
    keep = job.max_peaks
    job.max_peaks = iters // 2 + 2
    r = bif_cells(step, job, cells, threads=threads)
    job.max_peaks = keep
    out = np.zeros(len(r["flags"]), dtype=np.int32)
    for c, n in enumerate(r["flags"]):
        out[c] = n if n <= 1 else dbscan_count(r["peaks"][:n, c], eps)
    return out


def dft_window(window_type, N):
    w = np.zeros(max(1, N))
    lib().oc_dft_window(int(window_type), int(N), w.ctypes.data_as(_D))
    return w[:N]


def dft_block(x, w, n_freq, f_lo, f_hi, dt, log_axis=False):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    ak, bk = np.zeros(n_freq), np.zeros(n_freq)
    lib().oc_dft_block(x.ctypes.data_as(_D), len(x), w.ctypes.data_as(_D), int(n_freq), float(f_lo), float(f_hi),
                       int(log_axis), float(dt), ak.ctypes.data_as(_D), bk.ctypes.data_as(_D))
    return ak, bk


def dft_cells(step, job, cells, n_freq, f_lo, f_hi, log_axis=False, window_type=1, threads=1):
    """1-D DFT sweep: ak, bk [cells, n_freq] and the block's loop flag."""
    cells, cp = _cells(cells)
    n = len(cells)
    ak, bk = np.zeros((n, n_freq)), np.zeros((n, n_freq))
    flags = np.zeros(n, dtype=np.int32)
    lib().oc_dft_cells(step.ptr, C.byref(job), cp, n, int(n_freq), float(f_lo), float(f_hi), int(log_axis),
                       int(window_type), ak.ctypes.data_as(_D), bk.ctypes.data_as(_D), flags.ctypes.data_as(_I),
                       int(threads))
    return dict(ak=ak, bk=bk, flags=flags)


def dft_hsweep_cells(step, job, cells, n_freq, f_lo, f_hi, log_axis=False, window_type=1, threads=1):
    """1-D DFT sweep over h (per-cell block length, window and dt): ak, bk
    [cells, n_freq] and flags."""
    cells, cp = _cells(cells)
    n = len(cells)
    ak, bk = np.zeros((n, n_freq)), np.zeros((n, n_freq))
    flags = np.zeros(n, dtype=np.int32)
    lib().oc_dft_hsweep_cells(step.ptr, C.byref(job), cp, n, int(n_freq), float(f_lo), float(f_hi), int(log_axis),
                              int(window_type), ak.ctypes.data_as(_D), bk.ctypes.data_as(_D),
                              flags.ctypes.data_as(_I), int(threads))
    return dict(ak=ak, bk=bk, flags=flags)


def _n_cells(job):
    n = 1
    for i in range(job.n_axes):
        n *= job.axis_n[i]
    return n


def cont_bif(step, job, reverse=False, threads=1):
    """Continuation bifurcation chains: flags, peaks / intervals [max_peaks, cells]."""
    n = _n_cells(job)
    mp = max(1, job.max_peaks)
    flags = np.zeros(n, dtype=np.int32)
    peaks, intervals = np.full((mp, n), np.nan), np.full((mp, n), np.nan)
    lib().oc_cont_bif(step.ptr, C.byref(job), int(reverse), 0, flags.ctypes.data_as(_I), peaks.ctypes.data_as(_D),
                      intervals.ctypes.data_as(_D), None, None, int(threads))
    return dict(flags=flags, peaks=peaks, intervals=intervals)


def cont_dft(step, job, n_freq, f_lo, f_hi, reverse=False, log_axis=False, window_type=1, threads=1):
    """Continuation DFT chains: ak, bk [cells, n_freq], flags."""
    n = _n_cells(job)
    ak, bk = np.zeros((n, n_freq)), np.zeros((n, n_freq))
    flags = np.zeros(n, dtype=np.int32)
    lib().oc_cont_dft(step.ptr, C.byref(job), int(reverse), int(n_freq), float(f_lo), float(f_hi), int(log_axis),
                      int(window_type), ak.ctypes.data_as(_D), bk.ctypes.data_as(_D), flags.ctypes.data_as(_I),
                      int(threads))
    return dict(ak=ak, bk=bk, flags=flags)


def cont_lyap(step, job, K=1, reverse=False, threads=1):
    """Continuation finite-difference Lyapunov chains: values [K, cells], flags."""
    n = _n_cells(job)
    values = np.zeros((K, n))
    flags = np.zeros(n, dtype=np.int32)
    lib().oc_cont_lyap(step.ptr, C.byref(job), int(reverse), int(K), values.ctypes.data_as(_D),
                       flags.ctypes.data_as(_I), int(threads))
    return dict(values=values, flags=flags)


def _fs(n_vars, values, h, iter_of_synchr, pre_scaller, type_of_synch, error_estim, fs_error_trs, kf, kb):
    f = OcFs()
    f._a = (C.c_double * max(1, len(values)))(*values)
    f._kf = (C.c_double * n_vars)(*kf)
    f._kb = (C.c_double * n_vars)(*kb)
    f.n_vars, f.a, f.h = n_vars, f._a, h
    f.iter_of_synchr, f.pre_scaller, f.type_of_synch, f.error_estim = iter_of_synchr, pre_scaller, type_of_synch, error_estim
    f.fs_error_trs, f.k_forward, f.k_backward = fs_error_trs, f._kf, f._kb
    return f


def _dvec(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    return a, a.ctypes.data_as(_D)


def fs_attractor(step, n_vars, values, ic_master, ic_slave, k_forward, k_backward, idx, h=0.01, iter_of_synchr=100,
                 pre_scaller=1, t_max=100.0, transient_time=0.0, window=50.0, type_of_synch=0, error_estim=2,
                 fs_error_trs=1e-12, threads=1):
    """Fast synchronisation on the attractor: sync error and master point of
    the selected window starts idx; returns (n_points, errors, traj)."""
    f = _fs(n_vars, values, h, iter_of_synchr, pre_scaller, type_of_synch, error_estim, fs_error_trs, k_forward,
            k_backward)
    idx, ip = _cells(idx)
    out = np.zeros(len(idx))
    traj = np.zeros((len(idx), n_vars))
    m, mp = _dvec(ic_master)
    sl, sp = _dvec(ic_slave)
    n = lib().oc_fs_attractor(step.ptr, C.byref(f), mp, sp, float(t_max), float(transient_time), float(window), ip,
                              len(idx), out.ctypes.data_as(_D), traj.ctypes.data_as(_D), int(threads))
    return n, out, traj


def fs_grid(step, n_vars, values, ic_master, ic_slave, k_forward, k_backward, cells, h=0.01, iter_of_synchr=100,
            pre_scaller=1, transient_time=0.0, window=50.0, axis_x_var=0, axis_y_var=1, axis_x=(-10.0, 10.0),
            axis_y=(-10.0, 10.0), n_pts=200, grid_swap_master_slave=False, type_of_synch=0, error_estim=2,
            fs_error_trs=1e-12, threads=1):
    """Fast synchronisation on an initial-state grid: errors of the selected cells."""
    f = _fs(n_vars, values, h, iter_of_synchr, pre_scaller, type_of_synch, error_estim, fs_error_trs, k_forward,
            k_backward)
    cells, cp = _cells(cells)
    out = np.zeros(len(cells))
    m, mp = _dvec(ic_master)
    sl, sp = _dvec(ic_slave)
    lib().oc_fs_grid(step.ptr, C.byref(f), mp, sp, float(transient_time), float(window), int(n_pts), int(axis_x_var),
                     int(axis_y_var), float(axis_x[0]), float(axis_x[1]), float(axis_y[0]), float(axis_y[1]),
                     int(grid_swap_master_slave), cp, len(cells), out.ctypes.data_as(_D), int(threads))
    return out


def peak_finder(data, h):
    d = np.ascontiguousarray(np.asarray(data, dtype=np.float64))
    pk = np.zeros(max(1, len(d)))
    tm = np.zeros(max(1, len(d)))
    n = lib().oc_peak_finder(d.ctypes.data_as(_D), len(d), pk.ctypes.data_as(_D), tm.ctypes.data_as(_D), float(h))
    return pk[:n].copy(), tm[:n].copy()


def rng_uniforms(seed, seq, count):
    st = C.c_ulonglong()
    lib().oc_rng_init(seed, seq, 0, C.byref(st))
    return [lib().oc_rng_uniform(C.byref(st)) for _ in range(count)]


def axis_values(log_scale, n, lo, hi):
    out = (C.c_double * n)()
    lib().oc_axis_values(int(log_scale), n, float(lo), float(hi), out)
    return list(out)


class OcBasinsJob(C.Structure):
    _fields_ = [("base", OcJob), ("rtol", C.c_double), ("atol", C.c_double), ("h_min", C.c_double),
                ("h_max", C.c_double), ("section_period", C.c_double), ("period_tol", C.c_double),
                ("n_sections", C.c_int), ("n_record", C.c_int), ("max_period", C.c_int), ("label_vars", C.c_int),
                ("max_attempts", C.c_longlong)]


def make_basins_job(n_vars, axes, ic, values, h0, rtol, atol, section_period, n_sections, n_record=16,
                    max_period=8, label_vars=2, period_tol=1e-4, h_min=1e-12, h_max=None, max_value=1e6,
                    max_attempts=10 ** 7):
    bj = OcBasinsJob()
    base = make_job(n_vars, axes, ic, values, h0, 0.0, 1.0, max_value=max_value)
    bj._keep = base
    bj.base = base
    bj.rtol, bj.atol, bj.h_min = rtol, atol, h_min
    bj.h_max = h_max if h_max is not None else section_period / 4.0
    bj.section_period, bj.period_tol = section_period, period_tol
    bj.n_sections, bj.n_record, bj.max_period, bj.label_vars = n_sections, n_record, max_period, label_vars
    bj.max_attempts = max_attempts
    return bj


def basins_cells(funcs, bj, cells, threads=1):
    """Per cell: period (p >= 1 cycle, 0 none found, -1 unbound, -2 unresolved),
    canonical point [label_vars, n], accepted / rejected steps [2, n]."""
    cells, cp = _cells(cells)
    n = len(cells)
    period = np.zeros(n, dtype=np.int32)
    point = np.zeros((bj.label_vars, n))
    steps = np.zeros((2, n), dtype=np.int64)
    L = lib()
    L.oc_basins_cells.argtypes = [_P, C.POINTER(OcBasinsJob), _LL, C.c_longlong, _I, _D, _LL, C.c_int]
    L.oc_basins_cells(funcs.rhs, C.byref(bj), cp, n, period.ctypes.data_as(_I), point.ctypes.data_as(_D),
                      steps.ctypes.data_as(_LL), threads)
    return dict(period=period, point=point, steps=steps)


def dp5_step(rhs_ptr, x, a, t, h, rtol=0.0, atol=1.0):
    """One Dormand-Prince 5(4) attempt (the basins integrator's step):
    returns (x5, e, norm) with e the raw embedded error estimate."""
    L = lib()
    L.oc_dp5_step.argtypes = [_P, _D, _D, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _D, _D]
    L.oc_dp5_step.restype = C.c_double
    n = len(x)
    xa = (C.c_double * n)(*x)
    aa = (C.c_double * max(1, len(a)))(*a)
    x5 = (C.c_double * n)()
    e = (C.c_double * n)()
    nrm = L.oc_dp5_step(rhs_ptr, xa, aa, float(t), float(h), n, float(rtol), float(atol), x5, e)
    return np.array(list(x5)), np.array(list(e)), nrm


def section_signature(samples, max_period=8, tol=1e-4):
    """Period and canonical point of trailing Poincare samples (row 0 the
    newest, columns = leading state variables, at most 4)."""
    s = np.asarray(samples, dtype=np.float64)
    L_, D_ = s.shape
    buf = np.zeros((L_, 4))
    buf[:, :D_] = s
    buf = np.ascontiguousarray(buf)
    pt = (C.c_double * 4)()
    L = lib()
    L.oc_section_signature.argtypes = [_D, C.c_int, C.c_int, C.c_int, C.c_double, _D]
    L.oc_section_signature.restype = C.c_int
    p = L.oc_section_signature(buf.ctypes.data_as(_D), L_, D_, int(max_period), float(tol), pt)
    return p, np.array(list(pt)[:D_])
