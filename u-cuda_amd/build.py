"""Builds the engine library in-tree.

libucuda.so is the C-ABI engine (codegen + hipRTC + launches). The ensemble
kernels (csrc/kernels/ensemble.hip) and the launch descriptor (csrc/uc_job.h)
are embedded into it as text; hipRTC compiles them for gfx950 together with
each system's generated functions. `check_kernels()` runs that compilation
for the bundled systems without a GPU, so a broken kernel fails the build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libucuda.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _raw_literal(name, text):
    delim = "UCSRC"
    assert ")" + delim + '"' not in text
    # split into chunks: some compilers limit a single literal's length
    parts = [text[i:i + 30000] for i in range(0, len(text), 30000)] or [""]
    body = "\n".join('R"%s(%s)%s"' % (delim, p, delim) for p in parts)
    return "static const char* %s =\n%s;\n" % (name, body)


def write_embed():
    with open(os.path.join(CSRC, "kernels", "ensemble.hip")) as f:
        ens = f.read()
    with open(os.path.join(CSRC, "uc_job.h")) as f:
        job = f.read()
    out = "// generated by u-cuda_amd/build.py — do not edit\n"
    out += _raw_literal("kEnsembleSource", ens) + _raw_literal("kJobHeader", job)
    path = os.path.join(CSRC, "kernels_embed.inc")
    old = open(path).read() if os.path.exists(path) else None
    if old != out:
        with open(path, "w") as f:
            f.write(out)


def build_scratch_object(verbose=False):
    """scratch.hip holds device code (the static scratch array of
    uc_run_lle_noalloc), so it is compiled for gfx950 on its own."""
    obj = os.path.join(HERE, "scratch.o")
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
           "-c", os.path.join(CSRC, "scratch.hip"), "-o", obj]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return obj


def build_library(verbose=False):
    write_embed()
    scratch_obj = build_scratch_object(verbose)
    # host sin / cos stay separate calls (a fused sincos rounds differently
    # from cos on some arguments; the DFT tables must match the C library's
    # sin and cos)
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "-O2", "-std=c++17", "-fPIC", "-shared",
           "-Wall", "-Wno-unused-function",
           "-fno-builtin-sin", "-fno-builtin-cos", "-fno-builtin-sinf", "-fno-builtin-cosf",
           "-o", LIB + ".tmp",
           os.path.join(CSRC, "engine.cpp"), os.path.join(CSRC, "codegen.cpp"),
           os.path.join(CSRC, "parametric_engine.cpp"), os.path.join(CSRC, "data_export.cpp"), "-x", "none", scratch_obj,
           "-L" + os.path.join(ROCM, "lib"), "-lhiprtc", "-lamdhip64", "-ldl",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    # C++ drop-in facade demo, linked against the library like a session layer would be
    demo = os.path.join(HERE, "facade_demo")
    subprocess.run([os.path.join(ROCM, "bin", "hipcc"), "-O2", "-std=c++17", "-o", demo + ".tmp",
                    os.path.join(CSRC, "facade_demo.cpp"), "-L" + HERE, "-lucuda", "-Wl,-rpath," + HERE,
                    "-Wl,-rpath," + os.path.join(ROCM, "lib")], check=True)
    os.replace(demo + ".tmp", demo)
    return LIB


def check_kernels():
    """hipRTC-compiles the kernel library for bundled systems (no GPU needed),
    one engine per system on its own thread (the compiles release the GIL).
    The code objects land in the engine's on-disk cache (kcache/ beside the
    library), so processes that use these systems later load them instead of
    compiling; ucuda binds torch's HIP runtime first, as a GPU process does,
    so the entries carry the hipRTC version those processes will look for."""
    from concurrent.futures import ThreadPoolExecutor
    import shutil
    import ucuda  # noqa: E402  (sys.path holds this directory)

    # every build compiles: drop the entries of earlier builds first
    shutil.rmtree(os.path.join(HERE, "kcache"), ignore_errors=True)

    def one(name):
        eng = ucuda.Engine(device=-1)
        sysm = ucuda.preset(name, eng)
        return {(name, strict): sysm.compile(strict=strict) for strict in (False, True)}

    sizes = {}
    names = ("lorenz", "rossler", "hyper_rossler", "duffing")
    with ThreadPoolExecutor(max_workers=len(names)) as pool:
        for part in pool.map(one, names):
            sizes.update(part)
    return sizes


if __name__ == "__main__":
    build_library(verbose=True)
