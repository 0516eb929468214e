"""Build recipe of the fargo3d_amd shared library (gfx950).

Produces ``fargo3d_amd/lib/libfargo3d_amd.so``: the HIP kernels and their C
ABI (``include/fargo3d_hip.h``), the C host layer (mesh, parameters, planets,
initial conditions) and the C step driver (``f3d_driver.h``), plus the
``fargo3d_amd/bin/fargo3d_amd`` command-line runner.

Flags that matter for results: ``-ffp-contract=off`` (no fused multiply-add,
so the arithmetic rounds like the sequential CPU build) and the packed-FP32
feature turned off (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("F3D_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["host/f3d_mesh.c", "host/f3d_param.c", "host/f3d_planets.c", "host/f3d_setup.c",
             "host/f3d_driver.c"]
HIP_SRCS = ["hip/f3d_runtime.hip", "hip/f3d_routines.hip", "hip/f3d_fused.hip", "hip/f3d_fused32.hip"]

CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Werror=implicit-function-declaration"]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
# The reduced-precision build is not bit-matched to anything: it may fuse
# multiply-adds.  Its divisions stay IEEE-rounded: the hardware reciprocal's
# error is biased, and over thousands of steps of a quasi-steady disk a bias
# of the same sign in every cell acts like a source (measured: the density
# error after 10 orbits of the adiabatic disk halves without it).
EXTRA = {"hip/f3d_fused32.hip": ["-ffp-contract=fast"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    # hipcc's host pass warns about the unknown packed-fp32 feature: harmless
    return r.stdout


def _stale(out, srcs):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(s) > t for s in srcs)


def build(force=True, verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    headers = [os.path.join(CSRC, "host", "f3d_host.h"), os.path.join(CSRC, "host", "f3d_driver.h"),
               os.path.join(ROOT, "include", "fargo3d_hip.h")]
    headers += [os.path.join(CSRC, "hip", h) for h in os.listdir(os.path.join(CSRC, "hip"))
                if h.endswith(".h")]
    objs, jobs = [], []
    for s in HOST_SRCS:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        if force or _stale(obj, [src] + headers):
            jobs.append(["gcc"] + CFLAGS + ["-I" + os.path.join(ROOT, "include"), "-c", src, "-o", obj])
        objs.append(obj)
    for s in HIP_SRCS:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        if force or _stale(obj, [src] + headers):
            jobs.append([HIPCC] + HIPFLAGS + EXTRA.get(s, []) + ["-c", src, "-o", obj])
        objs.append(obj)
    # the translation units are independent: compile them side by side
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(len(jobs), 8) or 1) as pool:
        list(pool.map(_run, jobs))
    lib = os.path.join(LIBDIR, "libfargo3d_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, lib)
    main_src = os.path.join(CSRC, "host", "f3d_main.c")
    if os.path.exists(main_src):
        exe = os.path.join(BINDIR, "fargo3d_amd")
        _run(["gcc"] + CFLAGS + ["-I" + os.path.join(ROOT, "include"), main_src, "-o", exe,
                                 "-L" + LIBDIR, "-lfargo3d_amd", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    return lib


def clean():
    shutil.rmtree(OBJDIR, ignore_errors=True)


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
