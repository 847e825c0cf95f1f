"""Python host mirror of the engine's C ABI (include/ucuda.h).

Thin ctypes layer: systems are defined as in the reference (vars, params,
rhs strings; U-CUDA/codegen.hpp:10-15), requests mirror the reference's
Request structs (U-CUDA/parametric_engine.h:27-126, 411-482, 581-648) and
results come back as torch tensors on the GPU. torch only supplies device
memory and the stream; every computation runs in libucuda.so's HIP kernels.
There is no fallback: if libucuda.so is missing or a launch fails, the call
raises.
"""
import atexit
import ctypes as C
import math
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libucuda.so")

UC_OK = 0
PREC_F64, PREC_F32C = 0, 1
AXIS_PARAM, AXIS_VAR, AXIS_STEP = 1, 2, 3
METHOD_VARIATIONAL, METHOD_FINITE_DIFF = 0, 1
SCHEME_EULER, SCHEME_EULER_CROMER, SCHEME_MIDPOINT, SCHEME_RK4, SCHEME_DOPRI78, SCHEME_CD = 0, 1, 2, 3, 4, 5
REFERENCE_PROBE_SEED = 1234567891  # U-CUDA/cudaLibrary.cu:2603


class UcError(RuntimeError):
    pass


class SystemDesc(C.Structure):
    _fields_ = [("vars", C.POINTER(C.c_char_p)), ("n_vars", C.c_int),
                ("params", C.POINTER(C.c_char_p)), ("n_params", C.c_int),
                ("rhs", C.POINTER(C.c_char_p)), ("latex", C.c_int),
                ("time_symbol", C.c_char_p)]


class Axis(C.Structure):
    _fields_ = [("kind", C.c_int), ("index", C.c_int), ("n_pts", C.c_int),
                ("lo", C.c_double), ("hi", C.c_double), ("log_scale", C.c_int)]


class Common(C.Structure):
    _fields_ = [("initial_conditions", C.POINTER(C.c_double)),
                ("base_values", C.POINTER(C.c_double)), ("n_values", C.c_int),
                ("n_axes", C.c_int), ("axis", Axis * 2),
                ("h", C.c_double), ("transient_time", C.c_double), ("t_max", C.c_double),
                ("max_value", C.c_double), ("precision", C.c_int), ("strict", C.c_int),
                ("ic_seed", C.c_ulonglong), ("ic_lo", C.c_double), ("ic_hi", C.c_double),
                ("cell_begin", C.c_longlong), ("cell_count", C.c_longlong)]


class BifRequest(C.Structure):
    _fields_ = [("c", Common), ("writable_var", C.c_int), ("pre_scaller", C.c_int),
                ("max_peaks", C.c_int), ("record_precision", C.c_int)]


class BifOutput(C.Structure):
    _fields_ = [("flags", C.c_void_p), ("peaks", C.c_void_p), ("intervals", C.c_void_p),
                ("traj", C.c_void_p), ("traj_len", C.c_int), ("final_state", C.c_void_p)]


class LyapRequest(C.Structure):
    _fields_ = [("c", Common), ("NT", C.c_double), ("eps", C.c_double), ("method", C.c_int),
                ("probe_seed", C.c_ulonglong)]


class LyapOutput(C.Structure):
    _fields_ = [("values", C.c_void_p), ("flags", C.c_void_p), ("final_state", C.c_void_p)]


class DftRequest(C.Structure):
    _fields_ = [("c", Common), ("writable_var", C.c_int), ("pre_scaller", C.c_int), ("n_freq", C.c_int),
                ("freq_lo", C.c_double), ("freq_hi", C.c_double), ("freq_log_scale", C.c_int),
                ("window_type", C.c_int)]


class DftOutput(C.Structure):
    _fields_ = [("ak_cos", C.c_void_p), ("bk_sin", C.c_void_p), ("flags", C.c_void_p)]


class FastSyncRequest(C.Structure):
    _fields_ = [("values", C.c_void_p), ("n_values", C.c_int), ("mode", C.c_int), ("h", C.c_double),
                ("iter_of_synchr", C.c_int), ("pre_scaller", C.c_int), ("k_forward", C.c_void_p),
                ("k_backward", C.c_void_p), ("ic_master", C.c_void_p), ("ic_slave", C.c_void_p),
                ("t_max", C.c_double), ("transient_time", C.c_double), ("window", C.c_double),
                ("axis_x_var", C.c_int), ("axis_y_var", C.c_int), ("axis_x_lo", C.c_double),
                ("axis_x_hi", C.c_double), ("axis_y_lo", C.c_double), ("axis_y_hi", C.c_double),
                ("n_pts", C.c_int), ("grid_swap_master_slave", C.c_int), ("type_of_synch", C.c_int),
                ("error_estim", C.c_int), ("fs_error_trs", C.c_double), ("strict", C.c_int)]


class FastSyncOutput(C.Structure):
    _fields_ = [("sync_error", C.c_void_p), ("traj", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_fraction", C.c_double), ("max_bytes", C.c_longlong), ("round_traj_steps", C.c_double),
                ("block", C.c_int), ("waves_per_simd", C.c_int), ("slice_steps", C.c_double),
                ("slice_min_cells", C.c_longlong)]


class BatchPlan(C.Structure):
    _fields_ = [("block", C.c_int), ("waves_per_simd", C.c_int), ("slice_steps", C.c_double),
                ("cells_per_lane", C.c_int), ("vector_lanes", C.c_int), ("ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_longlong), ("rounds", C.c_longlong), ("queued", C.c_longlong),
                ("active", C.c_longlong), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("busy_s", C.c_double), ("wall_s", C.c_double), ("traj_steps", C.c_double),
                ("used_bytes", C.c_double), ("budget_bytes", C.c_double)]


class BasinsRequest(C.Structure):
    _fields_ = [("c", Common), ("rtol", C.c_double), ("atol", C.c_double),
                ("h_min", C.c_double), ("h_max", C.c_double), ("section_period", C.c_double),
                ("n_sections", C.c_int), ("n_record", C.c_int), ("max_period", C.c_int),
                ("label_vars", C.c_int), ("period_tol", C.c_double), ("quantum", C.c_double),
                ("max_attempts", C.c_longlong)]


class BasinsOutput(C.Structure):
    _fields_ = [("labels", C.c_void_p), ("periods", C.c_void_p), ("cycle_point", C.c_void_p),
                ("steps", C.c_void_p), ("max_attractors", C.c_int), ("n_attractors", C.POINTER(C.c_int)),
                ("attractor_period", C.POINTER(C.c_int)), ("attractor_point", C.POINTER(C.c_double))]


# Every symbol include/ucuda.h declares, with its ctypes signature.
_SIGNATURES = {
    "uc_version": (C.c_char_p, []),
    "uc_engine_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "uc_engine_destroy": (None, [C.c_void_p]),
    "uc_last_error": (C.c_char_p, [C.c_void_p]),
    "uc_system_create": (C.c_int, [C.c_void_p, C.POINTER(SystemDesc), C.POINTER(C.c_void_p)]),
    "uc_system_create_krs": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "uc_system_destroy": (None, [C.c_void_p]),
    "uc_system_n_vars": (C.c_int, [C.c_void_p]),
    "uc_system_n_values": (C.c_int, [C.c_void_p]),
    "uc_system_source": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "uc_system_compile": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_size_t)]),
    "uc_system_op_counts": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 4),
    "uc_codegen_scheme": (C.c_int, [C.c_void_p, C.POINTER(SystemDesc), C.c_int, C.c_char_p, C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
    "uc_system_eval_rhs": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                     C.POINTER(C.c_double)]),
    "uc_run_bif": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BifRequest), C.POINTER(BifOutput), C.c_void_p]),
    "uc_run_lle": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.POINTER(LyapOutput), C.c_void_p]),
    "uc_run_lle_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.POINTER(LyapOutput),
                                     C.c_void_p]),
    "uc_run_ls": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.POINTER(LyapOutput), C.c_void_p]),
    "uc_run_basins": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BasinsRequest), C.POINTER(BasinsOutput),
                                C.c_void_p]),
    "uc_run_bif_periods": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BifRequest), C.c_double, C.c_void_p,
                                     C.c_void_p]),
    "uc_run_dft": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DftRequest), C.POINTER(DftOutput), C.c_void_p]),
    "uc_run_bif_cont": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BifRequest), C.c_int, C.POINTER(BifOutput),
                                  C.c_void_p]),
    "uc_run_lle_cont": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.c_int, C.POINTER(LyapOutput),
                                  C.c_void_p]),
    "uc_run_ls_cont": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.c_int, C.POINTER(LyapOutput),
                                 C.c_void_p]),
    "uc_run_dft_cont": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DftRequest), C.c_int, C.POINTER(DftOutput),
                                  C.c_void_p]),
    "uc_fastsync_points": (C.c_longlong, [C.POINTER(FastSyncRequest)]),
    "uc_run_fastsync": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FastSyncRequest), C.POINTER(FastSyncOutput),
                                  C.c_void_p]),
    "uc_peak_finder": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_longlong, C.c_double,
                                 C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uc_bif_dbscan_periods": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int,
                                        C.c_double, C.c_void_p, C.c_void_p]),
    "uc_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "uc_batch_destroy": (None, [C.c_void_p]),
    "uc_batch_add_bif": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BifRequest), C.POINTER(BifOutput),
                                   C.POINTER(C.c_int)]),
    "uc_batch_add_lyap": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.c_int, C.POINTER(LyapOutput),
                                    C.POINTER(C.c_int)]),
    "uc_batch_add_basins": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BasinsRequest), C.POINTER(BasinsOutput),
                                      C.POINTER(C.c_int)]),
    "uc_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "uc_batch_set_plan": (C.c_int, [C.c_void_p, C.POINTER(BatchPlan)]),
    "uc_batch_get_plan": (C.c_int, [C.c_void_p, C.POINTER(BatchPlan)]),
    "uc_batch_tune_plan": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BatchPlan)]),
    "uc_batch_tune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]),
    "uc_batch_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "uc_batch_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 4),
    "uc_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "uc_server_destroy": (None, [C.c_void_p]),
    "uc_server_submit_lyap": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LyapRequest), C.c_int,
                                        C.POINTER(C.c_longlong)]),
    "uc_server_submit_bif": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BifRequest), C.POINTER(C.c_longlong)]),
    "uc_server_poll": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "uc_server_wait": (C.c_int, [C.c_void_p, C.c_longlong, C.c_double]),
    "uc_server_cancel": (C.c_int, [C.c_void_p, C.c_longlong]),
    "uc_server_fetch": (C.c_int, [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uc_server_submit_basins": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BasinsRequest), C.POINTER(C.c_longlong)]),
    "uc_server_fetch_basins": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(BasinsOutput)]),
    "uc_server_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "uc_grid_cells": (C.c_longlong, [C.POINTER(Common)]),
    "uc_axis_values": (C.c_int, [C.POINTER(Axis), C.POINTER(C.c_double)]),
}

_lib = None


def lib():
    """Loads libucuda.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UcError("libucuda.so not built: run __graft_entry__.build()")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7
        # (same soname as the system one libucuda.so links), so torch is
        # loaded first and libucuda.so binds to the runtime torch mapped;
        # loading libucuda.so first would map a second runtime that torch's
        # own device initialisation then fails on.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def _strs(items):
    arr = (C.c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


class Engine:
    """Wraps uc_engine. device < 0: compile-only engine (no GPU touched)."""

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        rc = lib().uc_engine_create(device, C.byref(h))
        if rc != UC_OK:
            raise UcError("uc_engine_create failed (rc=%d)" % rc)
        self.h = h

    def error(self):
        return lib().uc_last_error(self.h).decode(errors="replace")

    def check(self, rc, what):
        if rc != UC_OK:
            raise UcError("%s failed (rc=%d): %s" % (what, rc, self.error()))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().uc_engine_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def codegen_scheme(self, vars, params, rhs, scheme=SCHEME_RK4, latex=False):
        d, keep = _desc(vars, params, rhs, latex, None)
        n = C.c_size_t()
        self.check(lib().uc_codegen_scheme(self.h, C.byref(d), scheme, None, 0, C.byref(n)), "codegen_scheme")
        buf = C.create_string_buffer(n.value)
        self.check(lib().uc_codegen_scheme(self.h, C.byref(d), scheme, buf, n.value, C.byref(n)), "codegen_scheme")
        del keep
        return buf.value.decode()


def _desc(vars, params, rhs, latex, time_symbol):
    v, p, r = _strs(vars), _strs(params), _strs(rhs)
    d = SystemDesc(v, len(vars), p, len(params), r, 1 if latex else 0,
                   time_symbol.encode() if time_symbol else None)
    return d, (v, p, r)


class System:
    """A parsed system (uc_system). values() builds the a[] vector."""

    def __init__(self, engine, vars, params, rhs, latex=False, time_symbol=None, defaults=None, name=""):
        self.engine = engine
        self.vars, self.params, self.rhs = list(vars), list(params), list(rhs)
        self.latex, self.time_symbol = latex, time_symbol
        self.defaults = dict(defaults or {})
        self.name = name
        d, keep = _desc(self.vars, self.params, self.rhs, latex, time_symbol)
        h = C.c_void_p()
        engine.check(lib().uc_system_create(engine.h, C.byref(d), C.byref(h)), "uc_system_create")
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().uc_system_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def n_vars(self):
        return len(self.vars)

    @property
    def n_values(self):
        return 1 + len(self.params)

    def values(self, **kw):
        """a[] with a[0] = 0.5 (the reference sessions' CD symmetry default,
        U-CUDA/analysis_session.cpp:491-492) and params in alphabet order."""
        vals = dict(self.defaults)
        vals.update(kw)
        return [0.5] + [float(vals[p]) for p in self.params]

    def source(self, dialect=0):
        n = C.c_size_t()
        self.engine.check(lib().uc_system_source(self.h, dialect, None, 0, C.byref(n)), "uc_system_source")
        buf = C.create_string_buffer(n.value)
        self.engine.check(lib().uc_system_source(self.h, dialect, buf, n.value, C.byref(n)), "uc_system_source")
        return buf.value.decode()

    def compile(self, strict=False):
        n = C.c_size_t()
        self.engine.check(lib().uc_system_compile(self.engine.h, self.h, 1 if strict else 0, C.byref(n)),
                          "uc_system_compile")
        return n.value

    def op_counts(self):
        v = [C.c_int() for _ in range(4)]
        self.engine.check(lib().uc_system_op_counts(self.h, *[C.byref(x) for x in v]), "op_counts")
        return dict(zip(("rhs_flops", "rhs_calls", "jac_nonzeros", "jac_flops"), [x.value for x in v]))

    def eval_rhs(self, X, a, t=0.0):
        Xa = (C.c_double * len(X))(*X)
        aa = (C.c_double * len(a))(*a)
        out = (C.c_double * len(X))()
        self.engine.check(lib().uc_system_eval_rhs(self.h, Xa, aa, float(t), out), "eval_rhs")
        return list(out)


class KrsSystem(System):
    """A system given as a KRS step body (custom scheme) instead of rhs strings."""

    def __init__(self, engine, body, n_vars, n_values, name="krs"):
        self.engine = engine
        self.vars = ["x%d" % i for i in range(n_vars)]
        self.params = ["p%d" % i for i in range(n_values - 1)]
        self.rhs, self.latex, self.time_symbol, self.defaults, self.name = [], False, None, {}, name
        self.body = body
        h = C.c_void_p()
        engine.check(lib().uc_system_create_krs(engine.h, body.encode(), n_vars, n_values, C.byref(h)),
                     "uc_system_create_krs")
        self.h = h


PRESETS = {
    # library_template/<name>/system.json (alphabet order = param order)
    "lorenz": dict(vars=["x", "y", "z"], params=["sigma", "rho", "beta"],
                   rhs=["sigma*(y-x)", "x*(rho-z)-y", "x*y-beta*z"],
                   defaults=dict(sigma=10.0, rho=28.0, beta=8.0 / 3.0)),
    "rossler": dict(vars=["x", "y", "z"], params=["a", "b", "c"],
                    rhs=["-y-z", "x+a*y", "b+z*(x-c)"],
                    defaults=dict(a=0.2, b=0.2, c=5.7)),
    "chen": dict(vars=["x", "y", "z"], params=["a", "b", "c"],
                 rhs=["a*(y-x)", "(c-a)*x-x*z+c*y", "x*y-b*z"],
                 defaults=dict(a=40.0, b=3.0, c=28.0)),
    # 4-D hyperchaotic Rossler: dx=-y-z, dy=x+a*y+w, dz=3+x*z, dw=-0.5*z+d*w
    "hyper_rossler": dict(vars=["x", "y", "z", "w"], params=["a", "d"],
                          rhs=["-y-z", "x+a*y+w", "3+x*z", "-0.5*z+d*w"],
                          defaults=dict(a=0.25, d=0.05)),
    # forced double-well Duffing, x'' + delta x' - x + x^3 = gamma cos(omega t)
    "duffing": dict(vars=["x", "v"], params=["delta", "gamma", "omega"],
                    rhs=["v", "-delta*v+x-x^3+gamma*cos(omega*t)"], time_symbol="t",
                    defaults=dict(delta=0.15, gamma=0.3, omega=1.0)),
}


def preset(name, engine):
    p = dict(PRESETS[name])
    return System(engine, p["vars"], p["params"], p["rhs"], time_symbol=p.get("time_symbol"),
                  defaults=p.get("defaults"), name=name)


def axis(kind, index, n_pts, lo, hi, log_scale=False):
    return dict(kind=kind, index=index, n_pts=int(n_pts), lo=float(lo), hi=float(hi), log_scale=bool(log_scale))


def axis_values(ax):
    a = Axis(ax["kind"], ax["index"], ax["n_pts"], ax["lo"], ax["hi"], int(ax["log_scale"]))
    out = (C.c_double * ax["n_pts"])()
    rc = lib().uc_axis_values(C.byref(a), out)
    if rc != UC_OK:
        raise UcError("uc_axis_values rc=%d" % rc)
    return list(out)


def _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
            ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
    c = Common()
    ic = list(ic) if ic is not None else [0.0] * sysm.n_vars
    icarr = (C.c_double * len(ic))(*ic)
    varr = (C.c_double * len(values))(*values)
    c.initial_conditions = icarr
    c.base_values = varr
    c.n_values = len(values)
    c.n_axes = len(axes)
    for i, ax in enumerate(axes):
        c.axis[i] = Axis(ax["kind"], ax["index"], ax["n_pts"], ax["lo"], ax["hi"], int(ax["log_scale"]))
    c.h, c.transient_time, c.t_max, c.max_value = float(h), float(transient_time), float(t_max), float(max_value)
    c.precision, c.strict = int(precision), 1 if strict else 0
    c.ic_seed, c.ic_lo, c.ic_hi = int(ic_seed), float(ic_lo), float(ic_hi)
    return c, (icarr, varr)


def n_cells(axes):
    n = 1
    for ax in axes:
        n *= ax["n_pts"]
    return n


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def run_bif(sysm, axes, ic, values, h, transient_time, t_max, writable_var=0, pre_scaller=1,
            max_peaks=256, max_value=1e6, precision=PREC_F64, strict=False, traj_len=0,
            want_final_state=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0, out=None, stream=None, sync=True,
            record_precision=PREC_F64, cells=None):
    """Bifurcation sweep; returns dict(flags, peaks[max_peaks, cells], intervals, traj, final_state).
    record_precision (F32C sweeps): PREC_F64 records in double after the
    F32C transient, PREC_F32C records in float stages too.
    cells=(begin, count): run only that range of the grid's cells."""
    import torch
    eng = sysm.engine
    n = n_cells(axes) if cells is None else int(cells[1])
    dev = torch.device("cuda", eng.device)
    if out is None:
        out = dict(flags=torch.empty(n, dtype=torch.int32, device=dev),
                   peaks=torch.empty((max(max_peaks, 1), n), dtype=torch.float64, device=dev),
                   intervals=torch.empty((max(max_peaks, 1), n), dtype=torch.float64, device=dev),
                   traj=torch.empty((traj_len, n), dtype=torch.float64, device=dev) if traj_len else None,
                   final_state=torch.empty((sysm.n_vars, n), dtype=torch.float64, device=dev)
                   if want_final_state else None)
    r = BifRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                        ic_seed, ic_lo, ic_hi)
    r.writable_var, r.pre_scaller, r.max_peaks = writable_var, pre_scaller, max_peaks
    r.record_precision = record_precision
    if cells is not None:
        r.c.cell_begin, r.c.cell_count = int(cells[0]), int(cells[1])
    o = BifOutput(_ptr(out["flags"]).value, _ptr(out["peaks"]).value, _ptr(out["intervals"]).value,
                  _ptr(out.get("traj")).value, traj_len if out.get("traj") is not None else 0,
                  _ptr(out.get("final_state")).value)
    eng.check(lib().uc_run_bif(eng.h, sysm.h, C.byref(r), C.byref(o), _stream(stream)), "uc_run_bif")
    if sync:
        torch.cuda.synchronize(dev)
    del keep
    return out


def run_bif_periods(sysm, axes, ic, values, h, transient_time, t_max, eps_dbscan=0.1, writable_var=0,
                    pre_scaller=1, max_value=1e6, precision=PREC_F64, strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0,
                    stream=None, record_precision=PREC_F64):
    """2-D bifurcation map: periods[cells] (int32 CUDA tensor) = DBSCAN cluster
    count of each cell's peaks (-1 fixed point, 0 none / diverged)."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    periods = torch.empty(n, dtype=torch.int32, device=torch.device("cuda", eng.device))
    r = BifRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                        ic_seed, ic_lo, ic_hi)
    r.writable_var, r.pre_scaller, r.max_peaks = writable_var, pre_scaller, 0
    r.record_precision = record_precision
    eng.check(lib().uc_run_bif_periods(eng.h, sysm.h, C.byref(r), float(eps_dbscan), _ptr(periods).value,
                                       _stream(stream)), "uc_run_bif_periods")
    del keep
    return periods


def run_dft(sysm, axes, ic, values, h, transient_time, t_max, n_freq=200, freq_lo=0.0, freq_hi=10.0,
            freq_log_scale=False, window_type=1, writable_var=0, pre_scaller=1, max_value=1e6, precision=PREC_F64,
            strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0, stream=None):
    """1-D DFT sweep: dict(ak, bk [cells, n_freq], flags [cells]) CUDA tensors."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    dev = torch.device("cuda", eng.device)
    out = dict(ak=torch.empty((n, n_freq), dtype=torch.float64, device=dev),
               bk=torch.empty((n, n_freq), dtype=torch.float64, device=dev),
               flags=torch.empty(n, dtype=torch.int32, device=dev))
    r = DftRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                        ic_seed, ic_lo, ic_hi)
    r.writable_var, r.pre_scaller, r.n_freq = writable_var, pre_scaller, n_freq
    r.freq_lo, r.freq_hi, r.freq_log_scale, r.window_type = float(freq_lo), float(freq_hi), int(freq_log_scale), window_type
    o = DftOutput(_ptr(out["ak"]).value, _ptr(out["bk"]).value, _ptr(out["flags"]).value)
    eng.check(lib().uc_run_dft(eng.h, sysm.h, C.byref(r), C.byref(o), _stream(stream)), "uc_run_dft")
    del keep
    return out


def run_bif_cont(sysm, axes, ic, values, h, transient_time, t_max, reverse=False, writable_var=0, pre_scaller=1,
                 max_peaks=256, max_value=1e6, strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
    """Continuation bifurcation sweep along axes[0] (axes[1]: independent
    chains); dict(flags, peaks[max_peaks, cells], intervals)."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    dev = torch.device("cuda", eng.device)
    out = dict(flags=torch.empty(n, dtype=torch.int32, device=dev),
               peaks=torch.empty((max(max_peaks, 1), n), dtype=torch.float64, device=dev),
               intervals=torch.empty((max(max_peaks, 1), n), dtype=torch.float64, device=dev))
    r = BifRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, PREC_F64, strict,
                        ic_seed, ic_lo, ic_hi)
    r.writable_var, r.pre_scaller, r.max_peaks = writable_var, pre_scaller, max_peaks
    o = BifOutput(_ptr(out["flags"]).value, _ptr(out["peaks"]).value, _ptr(out["intervals"]).value, None, 0, None)
    eng.check(lib().uc_run_bif_cont(eng.h, sysm.h, C.byref(r), int(reverse), C.byref(o), None), "uc_run_bif_cont")
    del keep
    return out


def run_lyap_cont(sysm, axes, ic, values, h, transient_time, t_max, NT, spectrum=False, reverse=False, eps=1e-4,
                  probe_seed=REFERENCE_PROBE_SEED, max_value=1e6, strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
    """Continuation LLE (spectrum=False, values[1, cells]) or spectrum
    (values[N, cells]) along axes[0]; dict(values, flags)."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    k = sysm.n_vars if spectrum else 1
    dev = torch.device("cuda", eng.device)
    out = dict(values=torch.empty((k, n), dtype=torch.float64, device=dev),
               flags=torch.empty(n, dtype=torch.int32, device=dev))
    r = LyapRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, PREC_F64, strict,
                        ic_seed, ic_lo, ic_hi)
    r.NT, r.eps, r.method, r.probe_seed = float(NT), float(eps), METHOD_FINITE_DIFF, int(probe_seed)
    o = LyapOutput(_ptr(out["values"]).value, _ptr(out["flags"]).value, None)
    fn = lib().uc_run_ls_cont if spectrum else lib().uc_run_lle_cont
    eng.check(fn(eng.h, sysm.h, C.byref(r), int(reverse), C.byref(o), None), "uc_run_lyap_cont")
    del keep
    return out


def run_dft_cont(sysm, axes, ic, values, h, transient_time, t_max, reverse=False, n_freq=200, freq_lo=0.0,
                 freq_hi=10.0, freq_log_scale=False, window_type=1, writable_var=0, pre_scaller=1, max_value=1e6,
                 strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
    """Continuation DFT sweep along axes[0]; dict(ak, bk [cells, n_freq], flags)."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    dev = torch.device("cuda", eng.device)
    out = dict(ak=torch.empty((n, n_freq), dtype=torch.float64, device=dev),
               bk=torch.empty((n, n_freq), dtype=torch.float64, device=dev),
               flags=torch.empty(n, dtype=torch.int32, device=dev))
    r = DftRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, PREC_F64, strict,
                        ic_seed, ic_lo, ic_hi)
    r.writable_var, r.pre_scaller, r.n_freq = writable_var, pre_scaller, n_freq
    r.freq_lo, r.freq_hi, r.freq_log_scale, r.window_type = float(freq_lo), float(freq_hi), int(freq_log_scale), window_type
    o = DftOutput(_ptr(out["ak"]).value, _ptr(out["bk"]).value, _ptr(out["flags"]).value)
    eng.check(lib().uc_run_dft_cont(eng.h, sysm.h, C.byref(r), int(reverse), C.byref(o), None), "uc_run_dft_cont")
    del keep
    return out


def fastsync_request(sysm, values, ic_master, ic_slave, k_forward, k_backward, mode=0, h=0.01, iter_of_synchr=100,
                     pre_scaller=1, t_max=100.0, transient_time=0.0, window=50.0, axis_x_var=0, axis_y_var=1,
                     axis_x=(-10.0, 10.0), axis_y=(-10.0, 10.0), n_pts=200, grid_swap_master_slave=False,
                     type_of_synch=0, error_estim=2, fs_error_trs=1e-12, strict=False):
    """uc_fastsync_request with its host arrays kept alive (returns (request, keep))."""
    n = sysm.n_vars
    arrs = [(C.c_double * max(1, len(v)))(*v) for v in (values, k_forward, k_backward, ic_master, ic_slave)]
    for v in (k_forward, k_backward, ic_master, ic_slave):
        assert len(v) == n, "k_forward / k_backward / ic_master / ic_slave need n_vars values"
    r = FastSyncRequest()
    r.values, r.k_forward, r.k_backward, r.ic_master, r.ic_slave = [C.cast(a, C.c_void_p) for a in arrs]
    r.n_values, r.mode, r.h, r.iter_of_synchr, r.pre_scaller = len(values), mode, h, iter_of_synchr, pre_scaller
    r.t_max, r.transient_time, r.window = t_max, transient_time, window
    r.axis_x_var, r.axis_y_var, r.n_pts = axis_x_var, axis_y_var, n_pts
    r.axis_x_lo, r.axis_x_hi = axis_x
    r.axis_y_lo, r.axis_y_hi = axis_y
    r.grid_swap_master_slave, r.type_of_synch, r.error_estim = int(grid_swap_master_slave), type_of_synch, error_estim
    r.fs_error_trs, r.strict = fs_error_trs, int(strict)
    return r, arrs


def run_fastsync(sysm, values, ic_master, ic_slave, k_forward, k_backward, with_traj=True, stream=None, **kw):
    """Fast synchronisation; returns dict(sync_error, traj) CUDA tensors
    (traj: on-attractor master points, [n_out, n_vars])."""
    import torch
    eng = sysm.engine
    r, keep = fastsync_request(sysm, values, ic_master, ic_slave, k_forward, k_backward, **kw)
    n = lib().uc_fastsync_points(C.byref(r))
    if n < 0:
        raise UcError("invalid fast synchro request")
    dev = torch.device("cuda", eng.device)
    out = dict(sync_error=torch.empty(n, dtype=torch.float64, device=dev),
               traj=torch.empty((n, sysm.n_vars), dtype=torch.float64, device=dev)
               if (with_traj and r.mode == 0) else None)
    o = FastSyncOutput(_ptr(out["sync_error"]).value, _ptr(out.get("traj")).value)
    eng.check(lib().uc_run_fastsync(eng.h, sysm.h, C.byref(r), C.byref(o), _stream(stream)), "uc_run_fastsync")
    torch.cuda.synchronize(dev)
    del keep
    return out


def run_lyap(sysm, axes, ic, values, h, transient_time, t_max, NT, spectrum=False, eps=1e-4,
             method=METHOD_VARIATIONAL, probe_seed=REFERENCE_PROBE_SEED, max_value=1e6,
             precision=PREC_F64, strict=False, want_final_state=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0,
             out=None, stream=None, sync=True, noalloc=False, cells=None):
    """LLE map (spectrum=False, values[1, cells]) or full spectrum (values[N, cells]).
    noalloc=True: uc_run_lle_noalloc (tables in the static device scratch;
    the system's module must be loaded: sysm.compile(strict) first).
    cells=(begin, count): run only that range of the grid's cells."""
    import torch
    eng = sysm.engine
    n = n_cells(axes) if cells is None else int(cells[1])
    k = sysm.n_vars if spectrum else 1
    dev = torch.device("cuda", eng.device)
    if out is None:
        out = dict(values=torch.empty((k, n), dtype=torch.float64, device=dev),
                   flags=torch.empty(n, dtype=torch.int32, device=dev),
                   final_state=torch.empty((sysm.n_vars, n), dtype=torch.float64, device=dev)
                   if want_final_state else None)
    r = LyapRequest()
    r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                        ic_seed, ic_lo, ic_hi)
    r.NT, r.eps, r.method, r.probe_seed = float(NT), float(eps), int(method), int(probe_seed)
    if cells is not None:
        r.c.cell_begin, r.c.cell_count = int(cells[0]), int(cells[1])
    o = LyapOutput(_ptr(out["values"]).value, _ptr(out["flags"]).value, _ptr(out.get("final_state")).value)
    if noalloc:
        if spectrum:
            raise ValueError("noalloc covers the LLE map only")
        fn, name = lib().uc_run_lle_noalloc, "uc_run_lle_noalloc"
    else:
        fn, name = (lib().uc_run_ls, "uc_run_ls") if spectrum else (lib().uc_run_lle, "uc_run_lle")
    eng.check(fn(eng.h, sysm.h, C.byref(r), C.byref(o), _stream(stream)), name)
    if sync:
        torch.cuda.synchronize(dev)
    del keep
    return out


def _basins_request(sysm, axes, ic, values, h0, rtol, atol, section_period, n_sections, n_record, max_period,
                    period_tol, quantum, h_min, h_max, max_attempts, max_value, precision, strict):
    r = BasinsRequest()
    r.c, keep = _common(sysm, axes, ic, values, h0, 0.0, 1.0, max_value, precision, strict)
    r.rtol, r.atol, r.h_min = float(rtol), float(atol), float(h_min)
    r.h_max = float(h_max) if h_max is not None else float(section_period) / 4.0
    r.section_period, r.n_sections, r.n_record = float(section_period), int(n_sections), int(n_record)
    r.max_period, r.label_vars, r.period_tol, r.quantum = (int(max_period), min(sysm.n_vars, 2), float(period_tol),
                                                           float(quantum))
    r.max_attempts = int(max_attempts)
    return r, keep


def run_basins(sysm, axes, ic, values, h0=0.01, rtol=1e-9, atol=1e-9, section_period=2 * math.pi,
               n_sections=200, n_record=16, max_period=8, period_tol=1e-4, quantum=1e-3, h_min=1e-12,
               h_max=None, max_attempts=0, max_value=1e6, precision=PREC_F64, strict=False,
               want_points=True, want_steps=True, max_attractors=64, stream=None):
    """Basins of attraction: labels[cells] (1..K canonical, 0 unbound), periods,
    canonical points [2, cells], step counts [2, cells], attractor catalogue."""
    import torch
    eng = sysm.engine
    n = n_cells(axes)
    dev = torch.device("cuda", eng.device)
    lbl = min(sysm.n_vars, 2)
    out = dict(labels=torch.empty(n, dtype=torch.int32, device=dev),
               periods=torch.empty(n, dtype=torch.int32, device=dev),
               points=torch.empty((lbl, n), dtype=torch.float64, device=dev) if want_points else None,
               steps=torch.empty((2, n), dtype=torch.int64, device=dev) if want_steps else None)
    r, keep = _basins_request(sysm, axes, ic, values, h0, rtol, atol, section_period, n_sections, n_record,
                              max_period, period_tol, quantum, h_min, h_max, max_attempts, max_value, precision, strict)
    na = C.c_int(0)
    ap = (C.c_int * max_attractors)()
    apt = (C.c_double * (2 * max_attractors))()
    o = BasinsOutput(_ptr(out["labels"]).value, _ptr(out["periods"]).value, _ptr(out["points"]).value,
                     _ptr(out["steps"]).value, max_attractors, C.pointer(na), ap, apt)
    eng.check(lib().uc_run_basins(eng.h, sysm.h, C.byref(r), C.byref(o), _stream(stream)), "uc_run_basins")
    torch.cuda.synchronize(dev)
    k = min(na.value, max_attractors)
    out["attractors"] = [dict(label=i + 1, period=ap[i], point=(apt[2 * i], apt[2 * i + 1])) for i in range(k)]
    del keep
    return out


def peak_finder(sysm, data, time_step, max_peaks=2500, stream=None):
    """Streaming peak finder over stored series: data is a [n_samples, n_series]
    float64 CUDA tensor. Returns counts, peaks[max_peaks, n_series], intervals."""
    import torch
    eng = sysm.engine
    ns, nser = data.shape
    dev = data.device
    counts = torch.empty(nser, dtype=torch.int32, device=dev)
    peaks = torch.full((max(1, max_peaks), nser), float("nan"), dtype=torch.float64, device=dev)
    intervals = torch.full((max(1, max_peaks), nser), float("nan"), dtype=torch.float64, device=dev)
    data = data.contiguous()
    eng.check(lib().uc_peak_finder(eng.h, sysm.h, _ptr(data), ns, nser, float(time_step), max_peaks, _ptr(counts),
                                   _ptr(peaks), _ptr(intervals), _stream(stream)), "uc_peak_finder")
    torch.cuda.synchronize(dev)
    return dict(counts=counts, peaks=peaks, intervals=intervals)


def dbscan_periods(sysm, peaks, counts, eps_dbscan=0.1, stream=None):
    """Periods from stored peaks: peaks [max_peaks, n_cells] float64 and
    counts [n_cells] int32 CUDA tensors (peak-major, as run_bif returns them).
    Returns periods [n_cells] int32 as run_bif_periods."""
    import torch
    eng = sysm.engine
    peaks = peaks.contiguous()
    counts = counts.contiguous()
    max_peaks, n = peaks.shape
    assert counts.shape == (n,) and counts.dtype == torch.int32 and peaks.dtype == torch.float64
    periods = torch.empty(n, dtype=torch.int32, device=peaks.device)
    eng.check(lib().uc_bif_dbscan_periods(eng.h, sysm.h, _ptr(peaks), _ptr(counts), n, max_peaks, float(eps_dbscan),
                                          _ptr(periods), _stream(stream)), "uc_bif_dbscan_periods")
    return periods


class Batch:
    """Persistent heterogeneous batch (uc_batch): add jobs, tune, launch the
    captured graph. Outputs are torch tensors allocated at add time."""

    def __init__(self, engine):
        self.engine = engine
        h = C.c_void_p()
        engine.check(lib().uc_batch_create(engine.h, C.byref(h)), "uc_batch_create")
        self.h = h
        self.jobs = []
        self._keep = []
        self._basins = []  # (out, n_attractors, periods, points, max_attractors)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().uc_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def add_lyap(self, sysm, axes, ic, values, h, transient_time, t_max, NT, spectrum=False, eps=1e-4,
                 method=METHOD_VARIATIONAL, probe_seed=REFERENCE_PROBE_SEED, max_value=1e6, precision=PREC_F64,
                 strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
        import torch
        n = n_cells(axes)
        k = sysm.n_vars if spectrum else 1
        dev = torch.device("cuda", self.engine.device)
        out = dict(values=torch.empty((k, n), dtype=torch.float64, device=dev),
                   flags=torch.empty(n, dtype=torch.int32, device=dev))
        r = LyapRequest()
        r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                            ic_seed, ic_lo, ic_hi)
        r.NT, r.eps, r.method, r.probe_seed = float(NT), float(eps), int(method), int(probe_seed)
        o = LyapOutput(_ptr(out["values"]).value, _ptr(out["flags"]).value, None)
        jid = C.c_int()
        self.engine.check(lib().uc_batch_add_lyap(self.h, sysm.h, C.byref(r), 1 if spectrum else 0, C.byref(o),
                                                  C.byref(jid)), "uc_batch_add_lyap")
        self._keep.append((keep, sysm))
        self.jobs.append(out)
        return out

    def add_bif(self, sysm, axes, ic, values, h, transient_time, t_max, writable_var=0, pre_scaller=1,
                max_peaks=64, max_value=1e6, precision=PREC_F64, strict=False, record_precision=PREC_F64):
        import torch
        n = n_cells(axes)
        dev = torch.device("cuda", self.engine.device)
        out = dict(flags=torch.empty(n, dtype=torch.int32, device=dev),
                   peaks=torch.empty((max(1, max_peaks), n), dtype=torch.float64, device=dev),
                   intervals=torch.empty((max(1, max_peaks), n), dtype=torch.float64, device=dev))
        r = BifRequest()
        r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict)
        r.writable_var, r.pre_scaller, r.max_peaks = writable_var, pre_scaller, max_peaks
        r.record_precision = record_precision
        o = BifOutput(_ptr(out["flags"]).value, _ptr(out["peaks"]).value, _ptr(out["intervals"]).value, None, 0, None)
        jid = C.c_int()
        self.engine.check(lib().uc_batch_add_bif(self.h, sysm.h, C.byref(r), C.byref(o), C.byref(jid)),
                          "uc_batch_add_bif")
        self._keep.append((keep, sysm))
        self.jobs.append(out)
        return out

    def add_basins(self, sysm, axes, ic, values, h0=0.01, rtol=1e-9, atol=1e-9, section_period=2 * math.pi,
                   n_sections=200, n_record=16, max_period=8, period_tol=1e-4, quantum=1e-3, h_min=1e-12,
                   h_max=None, max_attempts=0, max_value=1e6, precision=PREC_F64, strict=False,
                   want_points=True, want_steps=True, max_attractors=64):
        """A basin map as run_basins; out["attractors"] is filled by launch()."""
        import torch
        n = n_cells(axes)
        dev = torch.device("cuda", self.engine.device)
        lbl = min(sysm.n_vars, 2)
        out = dict(labels=torch.empty(n, dtype=torch.int32, device=dev),
                   periods=torch.empty(n, dtype=torch.int32, device=dev),
                   points=torch.empty((lbl, n), dtype=torch.float64, device=dev) if want_points else None,
                   steps=torch.empty((2, n), dtype=torch.int64, device=dev) if want_steps else None)
        r, keep = _basins_request(sysm, axes, ic, values, h0, rtol, atol, section_period, n_sections, n_record,
                                  max_period, period_tol, quantum, h_min, h_max, max_attempts, max_value, precision,
                                  strict)
        na = C.c_int(0)
        ap = (C.c_int * max_attractors)()
        apt = (C.c_double * (2 * max_attractors))()
        o = BasinsOutput(_ptr(out["labels"]).value, _ptr(out["periods"]).value, _ptr(out["points"]).value,
                         _ptr(out["steps"]).value, max_attractors, C.pointer(na), ap, apt)
        jid = C.c_int()
        self.engine.check(lib().uc_batch_add_basins(self.h, sysm.h, C.byref(r), C.byref(o), C.byref(jid)),
                          "uc_batch_add_basins")
        self._keep.append((keep, sysm, o))
        self._basins.append((out, na, ap, apt, max_attractors))
        self.jobs.append(out)
        return out

    def set_launch(self, block, waves_per_simd):
        self.engine.check(lib().uc_batch_set_launch(self.h, block, waves_per_simd), "uc_batch_set_launch")

    def set_plan(self, block=256, waves_per_simd=4, slice_steps=0.0, cells_per_lane=1, vector_lanes=0):
        """Launch shape, trajectory-steps per cell per launch (0: whole runs),
        cells per lane (1, or 2 for LLE maps that can) and vector_lanes (1:
        spectrum maps run one lane per tangent vector)."""
        p = BatchPlan(block, waves_per_simd, float(slice_steps), cells_per_lane, vector_lanes, 0.0)
        self.engine.check(lib().uc_batch_set_plan(self.h, C.byref(p)), "uc_batch_set_plan")

    def plan(self):
        p = BatchPlan()
        self.engine.check(lib().uc_batch_get_plan(self.h, C.byref(p)), "uc_batch_get_plan")
        return dict(block=p.block, waves_per_simd=p.waves_per_simd, slice_steps=p.slice_steps,
                    cells_per_lane=p.cells_per_lane, vector_lanes=p.vector_lanes, ms=p.ms)

    def tune(self, stream=None):
        """Times the candidate plans and keeps the fastest (all five knobs)."""
        p = BatchPlan()
        self.engine.check(lib().uc_batch_tune_plan(self.h, _stream(stream), C.byref(p)), "uc_batch_tune_plan")
        return dict(block=p.block, waves_per_simd=p.waves_per_simd, slice_steps=p.slice_steps,
                    cells_per_lane=p.cells_per_lane, vector_lanes=p.vector_lanes, ms=p.ms)

    def launch(self, stream=None, sync=True):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        self.engine.check(lib().uc_batch_launch(self.h, _stream(stream)), "uc_batch_launch")
        for out, na, ap, apt, cap in self._basins:
            k = min(na.value, cap)
            out["attractors"] = [dict(label=i + 1, period=ap[i], point=(apt[2 * i], apt[2 * i + 1])) for i in range(k)]
        if sync:
            torch.cuda.synchronize()

    def info(self):
        v = [C.c_int() for _ in range(4)]
        self.engine.check(lib().uc_batch_info(self.h, *[C.byref(x) for x in v]), "uc_batch_info")
        return dict(zip(("n_jobs", "n_groups", "block", "waves_per_simd"), [x.value for x in v]))


_LIVE_SERVERS = weakref.WeakSet()


@atexit.register
def _close_servers():
    # stop scheduler threads while the HIP runtime is still fully up
    for srv in list(_LIVE_SERVERS):
        try:
            srv.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (uc_server). submit_* return a job id;
    fetch() returns numpy results once the job is done. The server keeps
    six rounds in flight on six streams; run the process with
    GPU_MAX_HW_QUEUES=8 (set before the HIP runtime starts) so that each
    stream has a hardware queue of its own: with HIP's default of 4 the
    round streams share two queues and rounds on one queue run in turn."""

    def __init__(self, engine, hbm_fraction=0.5, max_bytes=0, round_traj_steps=0.0, block=0, waves_per_simd=0,
                 slice_steps=0.0, slice_min_cells=0):
        """slice_steps: trajectory-steps per cell per round for jobs whose
        cells all step the same counts (0: 1024; < 0: whole runs);
        slice_min_cells: jobs of at most this many cells run whole
        (0: 65536; < 0: any job may be sliced)."""
        self.engine = engine
        cfg = ServerConfig(hbm_fraction, max_bytes, round_traj_steps, block, waves_per_simd, float(slice_steps),
                           int(slice_min_cells))
        h = C.c_void_p()
        engine.check(lib().uc_server_create(engine.h, C.byref(cfg), C.byref(h)), "uc_server_create")
        self.h = h
        self.meta = {}
        _LIVE_SERVERS.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().uc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_lyap(self, sysm, axes, ic, values, h, transient_time, t_max, NT, spectrum=False, eps=1e-4,
                    method=METHOD_VARIATIONAL, probe_seed=REFERENCE_PROBE_SEED, max_value=1e6, precision=PREC_F64,
                    strict=False, ic_seed=0, ic_lo=-1.0, ic_hi=1.0):
        r = LyapRequest()
        r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict,
                            ic_seed, ic_lo, ic_hi)
        r.NT, r.eps, r.method, r.probe_seed = float(NT), float(eps), int(method), int(probe_seed)
        jid = C.c_longlong()
        self.engine.check(lib().uc_server_submit_lyap(self.h, sysm.h, C.byref(r), 1 if spectrum else 0,
                                                      C.byref(jid)), "uc_server_submit_lyap")
        self.meta[jid.value] = ("lyap", n_cells(axes), sysm.n_vars if spectrum else 1, 0)
        del keep
        return jid.value

    def submit_bif(self, sysm, axes, ic, values, h, transient_time, t_max, writable_var=0, pre_scaller=1,
                   max_peaks=64, max_value=1e6, precision=PREC_F64, strict=False, record_precision=PREC_F64):
        r = BifRequest()
        r.c, keep = _common(sysm, axes, ic, values, h, transient_time, t_max, max_value, precision, strict)
        r.writable_var, r.pre_scaller, r.max_peaks = writable_var, pre_scaller, max_peaks
        r.record_precision = record_precision
        jid = C.c_longlong()
        self.engine.check(lib().uc_server_submit_bif(self.h, sysm.h, C.byref(r), C.byref(jid)), "uc_server_submit_bif")
        self.meta[jid.value] = ("bif", n_cells(axes), max(1, max_peaks), 1)
        del keep
        return jid.value

    def submit_basins(self, sysm, axes, ic, values, h0=0.01, rtol=1e-9, atol=1e-9, section_period=2 * math.pi,
                      n_sections=200, n_record=16, max_period=8, period_tol=1e-4, quantum=1e-3, h_min=1e-12,
                      h_max=None, max_attempts=0, max_value=1e6, precision=PREC_F64, strict=False):
        r, keep = _basins_request(sysm, axes, ic, values, h0, rtol, atol, section_period, n_sections, n_record,
                                  max_period, period_tol, quantum, h_min, h_max, max_attempts, max_value, precision,
                                  strict)
        jid = C.c_longlong()
        self.engine.check(lib().uc_server_submit_basins(self.h, sysm.h, C.byref(r), C.byref(jid)),
                          "uc_server_submit_basins")
        self.meta[jid.value] = ("basins", n_cells(axes), min(sysm.n_vars, 2), 0)
        del keep
        return jid.value

    def poll(self, jid):
        st, lat = C.c_int(), C.c_double()
        self.engine.check(lib().uc_server_poll(self.h, jid, C.byref(st), C.byref(lat)), "uc_server_poll")
        return st.value, lat.value

    def wait(self, jid, timeout_s=120.0):
        rc = lib().uc_server_wait(self.h, jid, float(timeout_s))
        self.engine.check(rc, "uc_server_wait")

    def cancel(self, jid):
        self.engine.check(lib().uc_server_cancel(self.h, jid), "uc_server_cancel")

    def fetch(self, jid, max_attractors=64):
        import numpy as np
        kind, n, k, two = self.meta[jid]
        if kind == "basins":
            labels = np.zeros(n, dtype=np.int32)
            periods = np.zeros(n, dtype=np.int32)
            points = np.zeros((k, n))
            steps = np.zeros((2, n), dtype=np.int64)
            na = C.c_int(0)
            ap = (C.c_int * max_attractors)()
            apt = (C.c_double * (2 * max_attractors))()
            o = BasinsOutput(labels.ctypes.data, periods.ctypes.data, points.ctypes.data, steps.ctypes.data,
                             max_attractors, C.pointer(na), ap, apt)
            self.engine.check(lib().uc_server_fetch_basins(self.h, jid, C.byref(o)), "uc_server_fetch_basins")
            att = [dict(label=i + 1, period=ap[i], point=(apt[2 * i], apt[2 * i + 1]))
                   for i in range(min(na.value, max_attractors))]
            return dict(labels=labels, periods=periods, points=points, steps=steps, attractors=att)
        flags = np.zeros(n, dtype=np.int32)
        out0 = np.zeros((k, n))
        out1 = np.zeros((k, n)) if two else None
        self.engine.check(lib().uc_server_fetch(self.h, jid, flags.ctypes.data_as(C.c_void_p),
                                                out0.ctypes.data_as(C.c_void_p),
                                                out1.ctypes.data_as(C.c_void_p) if two else None), "uc_server_fetch")
        if kind == "lyap":
            return dict(values=out0, flags=flags)
        return dict(flags=flags, peaks=out0, intervals=out1)

    def stats(self):
        st = ServerStats()
        self.engine.check(lib().uc_server_stats(self.h, C.byref(st)), "uc_server_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}


def rk4_flops(ops, n, tangents=0):
    """Algorithmic flops of one RK4 step: 4 rhs + 12 N for the state
    (convention +,-,*,/ = 1, FMA = 2), plus 4 JV products + 12 N per tangent."""
    f = 4 * ops["rhs_flops"] + 12 * n
    if tangents:
        jv = ops["jac_flops"] + tangents * (2 * ops["jac_nonzeros"] - n)
        f += 4 * jv + 12 * n * tangents
    return f


__all__ = [n for n in dir() if not n.startswith("_")]
_ = math
