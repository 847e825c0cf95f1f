"""The C ABI library (CPU only): loads, exports every symbol include/ucuda.h
declares, validates requests before touching a device, and reports errors
through status codes + uc_last_error."""
import ctypes as C
import os
import re
import subprocess

import pytest

import ucuda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ucuda.h")).read()
    return sorted(set(re.findall(r"\b(uc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", ucuda.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(line.split()[-1] for line in out.stdout.splitlines() if line.strip())
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    # the Python binding knows every declared entry point
    assert set(decl) == set(ucuda.exported_symbols())


def test_no_fused_sincos_in_host_tables():
    """The DFT rotation / resync tables (engine) and their restatement
    (oracle) must call the C library's sin and cos separately: glibc's sincos
    rounds differently from cos on ~0.1% of arguments, so a compiler that
    fuses the pair in one library and not the other breaks bit-exact parity."""
    for path in (ucuda.LIB_PATH, os.path.join(ROOT, "oracle", "liboracle.so")):
        out = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True, text=True, check=True)
        imported = set(line.split()[-1].split("@")[0] for line in out.stdout.splitlines() if line.strip())
        assert not imported & {"sincos", "sincosf"}, (path, imported & {"sincos", "sincosf"})


def test_compile_only_engine_and_errors(cpu_engine):
    s = ucuda.preset("lorenz", cpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 0.0, 10.0)]
    r = ucuda.BifRequest()
    r.c, keep = ucuda._common(s, axes, [0.1, 0.1, 0.1], s.values(), 0.01, 1.0, 1.0, 1e6, 0, 0)
    r.writable_var, r.pre_scaller, r.max_peaks = 0, 1, 8
    o = ucuda.BifOutput(1, 1, 1, 0, 0, 0)  # never dereferenced: no device
    L = ucuda.lib()
    # a valid request on an engine without a device -> UC_ERR_NO_DEVICE
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 5
    # argument errors are reported before any device work
    r.c.axis[0].index = 9
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    assert "param_index" in cpu_engine.error()
    r.c.axis[0].index = 2
    r.pre_scaller = 0
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.pre_scaller = 1
    r.c.n_axes = 0
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.c.n_axes = 1
    r.c.h = -1.0
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.c.h = 0.01
    r.max_peaks = 3000
    assert L.uc_run_bif(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    del keep


def test_lyap_validation(cpu_engine):
    s = ucuda.preset("rossler", cpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 4, 0.1, 0.3), ucuda.axis(ucuda.AXIS_STEP, 0, 4, 0.005, 0.01)]
    r = ucuda.LyapRequest()
    r.c, keep = ucuda._common(s, axes, [1, 1, 1], s.values(), 0.01, 1.0, 10.0, 1e6, 0, 0)
    r.NT, r.eps, r.method, r.probe_seed = 0.1, 1e-4, 0, 1234567891
    o = ucuda.LyapOutput(1, 1, 0)
    L = ucuda.lib()
    assert L.uc_run_lle(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 5
    r.NT = 0.0
    assert L.uc_run_lle(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.NT = 0.1
    r.method = 7
    assert L.uc_run_ls(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.method = 0
    r.c.axis[0].kind = ucuda.AXIS_STEP  # two step axes
    assert L.uc_run_lle(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    del keep


def test_basins_validation(cpu_engine):
    s = ucuda.preset("duffing", cpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, 4, -1, 1), ucuda.axis(ucuda.AXIS_VAR, 1, 4, -1, 1)]
    r = ucuda.BasinsRequest()
    r.c, keep = ucuda._common(s, axes, [0, 0], s.values(), 0.01, 0.0, 1.0, 1e6, 0, 0)
    r.rtol = r.atol = 1e-9
    r.h_min, r.h_max, r.section_period = 1e-12, 1.0, 6.283185307179586
    r.n_sections, r.n_record, r.max_period, r.label_vars, r.period_tol, r.quantum = 10, 16, 8, 2, 1e-4, 1e-3
    o = ucuda.BasinsOutput(1, 1, 0, 0, 0, None, None, None)
    L = ucuda.lib()
    assert L.uc_run_basins(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 5
    r.max_period = 16
    assert L.uc_run_basins(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    r.max_period = 8
    r.n_record = 40
    assert L.uc_run_basins(cpu_engine.h, s.h, C.byref(r), C.byref(o), None) == 1
    del keep


def test_version_and_sources(cpu_engine):
    assert ucuda.lib().uc_version().decode().startswith("ucuda")
    s = ucuda.preset("duffing", cpu_engine)
    src = s.source(0)
    assert "uc_bif_f64" in src and "uc_bas_f32c" in src and "#define UC_USES_TIME 1" in src
    assert "cos(a[3] * t)" in s.source(1)


def test_code_object_cache(tmp_path, monkeypatch):
    """Compiled modules are kept on disk under UC_CACHE_DIR: one entry per
    (program, options, hipRTC version) with its header; a fresh engine loads
    it instead of compiling; a damaged entry is recompiled and rewritten;
    UC_CACHE_DIR set but empty writes nothing."""
    cache = tmp_path / "kcache"
    monkeypatch.setenv("UC_CACHE_DIR", str(cache))

    def compile_once():
        eng = ucuda.Engine(device=-1)  # own in-memory module cache
        return ucuda.preset("lorenz", eng).compile(strict=True)

    n = compile_once()
    files = sorted(cache.iterdir())
    assert len(files) == 1 and files[0].suffix == ".hsaco"
    raw = files[0].read_bytes()
    assert raw.startswith(b"UCKC1 ") and raw[57:61] == b"\x7fELF" and len(raw) == 57 + n
    assert int(raw[23:39], 16) == n
    assert compile_once() == n and files[0].read_bytes() == raw  # loaded, not rewritten
    for bad in (raw[: len(raw) // 2], raw[:-1] + bytes([raw[-1] ^ 1])):  # truncated, one bit flipped
        files[0].write_bytes(bad)
        assert compile_once() == n
        assert files[0].read_bytes() == raw  # recompiled and rewritten
    monkeypatch.setenv("UC_CACHE_DIR", "")
    other = tmp_path / "unused"
    assert compile_once() == n and not other.exists() and len(list(cache.iterdir())) == 1


def test_engine_usable_while_it_compiles(monkeypatch):
    """A hipRTC compile (seconds) runs outside the engine lock: while one
    thread compiles a new module, other calls on the same engine return at
    once, and a compile error on one thread leaves another thread's last
    error untouched."""
    import threading
    import time
    monkeypatch.setenv("UC_CACHE_DIR", "")  # no disk cache: a real compile
    eng = ucuda.Engine(device=-1)
    s = ucuda.System(eng, ["p", "q"], ["k"], ["q", "-k*p - 0.01*q^3"])
    done = {}

    def compile_it():
        t0 = time.perf_counter()
        s.compile(strict=True)
        done["s"] = time.perf_counter() - t0

    th = threading.Thread(target=compile_it)
    th.start()
    time.sleep(0.3)
    waits = []
    while th.is_alive() and len(waits) < 5:
        t0 = time.perf_counter()
        ucuda.System(eng, ["x"], ["a"], ["-a*x"])     # takes the engine lock
        waits.append(time.perf_counter() - t0)
        time.sleep(0.2)
    th.join()
    assert done["s"] > 1.0, done                     # the compile really ran
    assert waits and max(waits) < 0.5, (waits, done)
    # per-thread error copies
    with pytest.raises(ucuda.UcError):
        ucuda.System(eng, ["x"], [], ["x +* 2"])
    msg = eng.error()
    assert msg
    other = {}
    t = threading.Thread(target=lambda: other.setdefault("m", eng.error()))
    t.start()
    t.join()
    assert other["m"] == msg


def test_session_layer_compiles_and_validates(tmp_path):
    """A translation unit that builds every request the way the reference's
    analysis sessions do (krs_body from codegen_scheme, the same fields,
    snapshot and recluster use) compiles against the facade header, links
    against the library, and on a device-less engine gets every request
    through validation to the device stage; the host-only recluster labels
    synthetic features (three attractors, one spread over two touching boxes,
    unbound cells) in canonical order."""
    import json
    exe = str(tmp_path / "session_fields")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "u-cuda_amd", "csrc"),
                    os.path.join(ROOT, "tests", "session_fields.cpp"), "-o", exe,
                    "-L", os.path.join(ROOT, "u-cuda_amd"), "-lucuda",
                    "-Wl,-rpath," + os.path.join(ROOT, "u-cuda_amd"), "-Wl,-rpath,/opt/rocm/lib",
                    "-Wl,-rpath-link,/opt/rocm/lib"], check=True, capture_output=True, text=True)
    out = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120, check=True)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    for k in ("bif_error", "lle1d_error", "lle2d_error", "basins_error"):
        assert res[k] == "engine was created without a device", (k, res[k])
    assert "right-hand side" in res["basins_body_only_error"]
    assert res["recluster_ok"] == 1
    assert res["recluster_labels"] == [3, 3, 3, 1, 1, 0, 2, 2, 0]
    assert res["recluster_n"] == 3 and res["snapshot_eps"] == 0.1
