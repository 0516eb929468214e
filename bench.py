"""Benchmark of FARGO3D's hydro time step on one MI355X.

Workload: the 3-D spherical isothermal disk with an embedded planet
(256 azimuth x 96 radius x 32 colatitude cells, FP64, FARGO orbital
advection, alpha viscosity, Stockholm damping), synthetic initial conditions
from the setup (vertical hydrostatic equilibrium), the fused HIP path.

One "step" is one pass of the DT loop body (src/main.c:361-448): CFL time
step, sources, transport, damping and ghost fill, including the one device->
host read of dt the host needs for the planet integrator.  Fields are device
resident before the timed region starts.

Prints one JSON line: the FP64 step's cell-updates/s; the HBM roofline of the
stage with the largest time (the step's algorithmic bytes per cell-update,
B_alg, times the cell-updates of one launch, over its HIP-event duration;
beside it the same with the kernel's own algorithmic bytes, and the traffic
rocprofv3 measured for it); the same for the whole step; the per-stage
averages; the precision form (speed and per-field error against the FP64
path) and the batched form (64 uneven planet-disk runs over 1000 steps,
against looping the FP64 path, with its error); the other workloads at their
full sizes (the 2-D 384x128 disk, the 2-D adiabatic 4096x1024 disk in FP64 and
in the precision form, gas + 4 dust 1024x512 in both), each with its own
roofline; and the CPU baseline (the oracle on one host core).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
os.chdir(ROOT)

HBM_PEAK_GBS = 8000.0
# the dominant stage's HIP events bracket every EVENT_EVERY-th launch in the
# timed region (an event pair adds about 6 us to the step it brackets)
EVENT_EVERY = 8
PAR = {"c1": "fargo3d_amd/setups/c1_fargo2d.par", "c2": "fargo3d_amd/setups/c2_p3d.par",
       "c3": "fargo3d_amd/setups/c3_adiabatic2d.par", "c4": "fargo3d_amd/setups/c4_multifluid2d.par"}
WORKLOAD = {
    "c1": "2-D polar isothermal disk 384x128, FP64, FARGO, nu=1e-5, 1e-3 Mstar planet",
    "c2": "3-D spherical isothermal disk 256x96x32, FP64, FARGO, alpha=1e-3, 3e-4 Mstar planet",
    "c3": "2-D polar adiabatic disk 4096x1024, FP64, FARGO, gamma=1.4, nu=1e-5, 1e-3 Mstar planet",
    "c4": "2-D polar disk, gas + 4 dust fluids 1024x512, FP64, FARGO, implicit drag, dust diffusion, 1e-3 Mstar "
          "planet (fluid-cell-updates)",
}
# Algorithmic bytes per active cell and launch of each fused stage: every
# 3-D array the stage must read once and write once, at 8 bytes (FP64);
# 1-D/2-D tables and ghost cells are not counted.  dz: 3-D, ad: adiabatic
# (the energy is one more transported array), nu: constant viscosity (alpha
# viscosity also reads the sound speed in the stress stage).
def stage_arrays(stage, dz, ad, alpha):
    nv = 3 if dz else 2                  # velocity components
    nm = 2 * nv                          # staggered momenta
    ns = 6 if dz else 3                  # stress components
    table = {
        "cfl": 1 + nv + ad,                                  # cs or e, v (, rho)
        "sources_a": (1 + nv + alpha) + ns,                  # rho, v (, cs) -> stresses
        "sources_b": (3 + nv + ns) + (nv + ad),              # rho, cs/e, pot, v, stresses -> v_temp (, e); q on the fly
        "substep3": (1 + nv) + 1,                            # e, v_temp -> e
        "sweep_z": (1 + nv + ad) + (1 + nm + ad),            # rho, v_temp (, e) -> rho, momenta (, e)
        "sweep_y": ((1 + nm + ad + 1) if dz else (1 + nv + ad)) + (1 + nm + ad),
        "sweep_x": (1 + nm + 1 + ad) + (2 + nm - 2 + ad),    # rho, momenta, vx_temp (, e) -> rho, vx, y/z momenta (, e)
        "finish": (1 + nm - 2) + nv,                         # rho, y/z momenta -> rho, vy (, vz)
        "potential": 1,                                      # -> pot
    }
    return table.get(stage, 0)


# the kernels each stage's events bracket (names without namespace and template arguments)
STAGE_KERNEL = {"cfl": ["kf_cfl", "kf_cfl_fluids", "kf_vmed_rings", "kf_cfl_fin"], "sources_a": ["kf_src_a"],
                "sources_b": ["kf_src_b", "kf_src_b_fluids", "kf_src_fused2"], "sweep_z": ["kf_sweep_z"], "sweep_y": ["kf_sweep_y2", "kf_sweep_y2v", "kf_sweep_y3"],
                "sweep_x": ["kf_sweep_xw", "kf_sweep_xtw", "kf_sweep_xr_fluids", "kf_sweep_x_tile"], "finish": ["kf_finish"],
                "potential": ["k_potential", "k_potential_fin", "kf_potential"], "substep3": ["kf_substep3"]}
# rocprofv3 summary with the measured HBM traffic of each kernel of this workload
PROFILE = {"c2": "profiles/fused_step_c2_3d_256x96x32.md", "c1": "profiles/fused_step_c1_2d_384x128.md",
           "c3": "profiles/fused_step_c3_2d_adiabatic_4096x1024.md",
           "c4": "profiles/fused_step_c4_multifluid_1024x512.md"}
PROFILE32 = {"c2": "profiles/fused_step_fp32_c2_3d_256x96x32.md",
             "c3": "profiles/fused_step_fp32_c3_2d_adiabatic_4096x1024.md"}
# stages whose one launch covers every fluid
FLUID_BATCHED = {"cfl", "sweep_x", "finish", "sweep_y", "sources_b"}


def measured_traffic(config, stage, esize=8):
    """HBM bytes per launch of the stage's kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, (PROFILE if esize == 8 else PROFILE32).get(config, ""))
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch = "):
            table = json.loads(line.split("=", 1)[1])
            # every kernel the stage's events bracket (the CFL: VxMed pass, CFL, MIN pass)
            hits = [v for name, v in table.items()
                    if name.split("<")[0].split("::")[-1] in STAGE_KERNEL.get(stage, [])]
            return sum(hits) if hits else None
    return None


def cpu_baseline(parfile, budget_s=12.0):
    """The CPU restatement (oracle) on one host core, a bounded sample."""
    import oracle
    o = oracle.Oracle(parfile)
    t0 = time.perf_counter()
    n = 0
    while True:
        o.step()
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    return n, el


def copy_calibration(L, sim, nbytes):
    """Three copies of nbytes between two device buffers (outside the timed region)."""
    L.f3d_hip_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.f3d_hip_free.argtypes = [ctypes.c_void_p]
    L.f3d_hip_memcpy_d2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.f3d_hip_memset.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if L.f3d_hip_malloc(ctypes.byref(a), nbytes) or L.f3d_hip_malloc(ctypes.byref(b), nbytes):
        raise RuntimeError("copy calibration: allocation failed")
    L.f3d_hip_memset(sim.ctx, a, 1, nbytes)
    for _ in range(3):
        if L.f3d_hip_memcpy_d2d(sim.ctx, b, a, nbytes):
            raise RuntimeError("copy calibration: copy failed")
    sim.sync()
    L.f3d_hip_free(a)
    L.f3d_hip_free(b)


def _read_stages(L, sim, names):
    nst = len(names)
    ms = (ctypes.c_double * nst)()
    cnt = (ctypes.c_long * nst)()
    L.f3d_hip_timing_read(sim.ctx, ms, cnt)
    return {names[st]: {"avg_us": round(1e3 * ms[st] / cnt[st], 3), "launches": int(cnt[st])}
            for st in range(nst) if cnt[st]}


def timed_run(fa, L, names, parfile, overrides, arch, steps, warmup, probe=20):
    """Warm up, take every stage's HIP-event average over `probe` steps, then
    time exactly `steps` steps (synchronised on both sides) with events around
    the dominant stage only, on every EVENT_EVERY-th of its launches (each
    event pair adds a gap between launches).  Returns the wall time, the
    per-stage averages (the dominant stage's from the timed region) and the
    simulation."""
    # the handle is a pointer: every timing entry point gets its C signature
    L.f3d_hip_timing_mask.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    L.f3d_hip_timing_read.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.f3d_hip_timing_every.argtypes = [ctypes.c_void_p, ctypes.c_int]
    sim = fa.Simulation(parfile, overrides, arch=arch)
    for _ in range(warmup):
        sim.step()
    sim.sync()
    L.f3d_hip_timing_mask(sim.ctx, (1 << len(names)) - 1)
    for _ in range(probe):
        sim.step()
    stages = _read_stages(L, sim, names)
    dom = max(stages, key=lambda n: stages[n]["avg_us"] * stages[n]["launches"]) if stages else None
    L.f3d_hip_timing_mask(sim.ctx, (1 << names.index(dom)) if dom else 0)
    # the event pair goes around every EVENT_EVERY-th launch of the dominant
    # stage only: each pair puts a gap of a few microseconds into the step
    L.f3d_hip_timing_every(sim.ctx, EVENT_EVERY)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.sync()
    elapsed = time.perf_counter() - t0
    if dom:
        timed = _read_stages(L, sim, names)[dom]
        stages[dom] = dict(timed, dominant=True, probe_avg_us=stages[dom]["avg_us"])
    L.f3d_hip_timing_mask(sim.ctx, 0)
    L.f3d_hip_timing_every(sim.ctx, 1)
    return elapsed, stages, sim


def b_alg_arrays(dimz, ad, nfl=1):
    """Arrays per cell-update of the compulsory state traffic: rho, v (, e)
    read and written, the sound speed read (isothermal); with several fluids,
    per fluid-cell update, the gas's arrays and each pressureless dust
    fluid's rho, v read and written, averaged."""
    nv = 3 if dimz else 2
    n = 2 * (1 + nv + ad) + (0 if ad else 1)
    if nfl > 1:
        n = (n + (nfl - 1) * 2 * (1 + nv)) / nfl
    return n


def roofline_of(stages, cells, dimz, ad, alpha, esize, config, b_alg, nfl=1):
    """Roofline figures of the stage with the largest total time: achieved =
    B_alg (the step's algorithmic bytes per cell-update) x the cell-updates
    one launch performs / its average HIP-event duration; the same with the
    kernel's own algorithmic bytes (every array it must read and write once)
    beside it; traffic = its HBM bytes per launch from the committed rocprofv3
    counters."""
    if not stages:
        return None
    doms = [n for n in stages if stages[n].get("dominant")]
    dom = doms[0] if doms else max(stages, key=lambda n: stages[n]["avg_us"] * stages[n]["launches"])
    avg_s = stages[dom]["avg_us"] * 1e-6
    units = cells * nfl if dom in FLUID_BATCHED else cells
    achieved = b_alg * units / avg_s / 1e9 if avg_s > 0 else 0.0
    per_cell = esize * stage_arrays(dom, dimz, ad, alpha)
    k_achieved = per_cell * units / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(config, dom, esize),
            "bytes_per_cell_update": b_alg, "cell_updates_per_launch": units, "avg_us": stages[dom]["avg_us"],
            "kernel_bytes_per_cell": per_cell, "kernel_achieved": round(k_achieved, 1),
            "kernel_frac": round(k_achieved / HBM_PEAK_GBS, 4)}


def workload_run(fa, L, names, config, overrides, steps, warmup, esize):
    """cell-updates/s of one workload (fluid-cell-updates with several
    fluids), its step roofline under B_alg and the dominant stage's roofline."""
    el, st, sim = timed_run(fa, L, names, PAR[config], overrides, {}, steps, warmup)
    cells = sim.nx * sim.ny * (sim.nz if sim.dimz else 1)
    nfl, ad, dimz = sim.nfluids, 1 if sim.adiabatic else 0, sim.dimz
    sim.close()
    v = cells * nfl * steps / el
    b_alg = esize * b_alg_arrays(dimz, ad, nfl)
    return {"workload": WORKLOAD[config], "value": round(v, 1),
            "unit": "fluid-cell-updates/s" if nfl > 1 else "cell-updates/s", "cells": cells, "fluids": nfl,
            "steps": steps, "ms_per_step": round(1e3 * el / steps, 4),
            "step_roofline": {"bytes_per_cell_update": b_alg, "achieved": round(v * b_alg / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(v * b_alg / 1e9 / HBM_PEAK_GBS, 4)},
            "roofline": roofline_of(st, cells, dimz, ad, 1 if config == "c2" else 0, esize, config, b_alg, nfl),
            "stages": st}


def fp32_errors(fa, parfile, steps=None, orbits=None):
    """Relative L1 and L-infinity errors of every field of the FP32 form against
    the FP64 path after `steps` steps, or after `orbits` orbits at r = 1, on one
    sequence of time steps (each FP32 step takes the FP64 step before the
    output-time clamp, so both clamp alike)."""
    import numpy as np
    s64 = fa.Simulation(parfile)
    s32 = fa.Simulation(parfile, "Precision=32")
    n = 0
    t_end = 2 * np.pi * orbits if orbits else None
    while (n < steps) if steps else (s64.scalars()["PhysicalTime"] < t_end * (1 - 1e-12)):
        s64.step()
        s32.force_dt(s64.step_time())
        s32.step()
        n += 1
    fields = ["Density", "Vx", "Vy"] + (["Vz"] if s64.dimz else []) + (["Energy"] if s64.adiabatic else [])
    out = {"steps": n, "time": round(s64.scalars()["PhysicalTime"], 6)}
    if orbits:
        out["orbits"] = orbits
    for f in fields:
        a, b = s32.field(f), s64.field(f)
        a = a[3:-3, 3:-3, :] if s64.dimz else a[:, 3:-3, :]
        b = b[3:-3, 3:-3, :] if s64.dimz else b[:, 3:-3, :]
        d = np.abs(a - b)
        out[f] = {"L1": float("%.3e" % (d.sum() / np.abs(b).sum())), "Linf": float("%.3e" % (d.max() / np.abs(b).max()))}
    s64.close()
    s32.close()
    return out


def batched_form(fa, n, steps, loop_steps, err_every=4):
    """Aggregate cell-updates/s of n disks advanced together (FP32, one graph
    replay per step, autotuned tiles, planets integrated on the device) and of
    looping the plain FP64 path over the same disks; then the FP32 form's
    error against the FP64 path on every err_every-th disk at equal step
    count (each run on its own time steps)."""
    import numpy as np
    specs = fa.sweep_specs(n)
    b = fa.Batch(specs)
    cells = sum(b.cells)
    b.step(5)
    b.sync()
    rows_y, tile = b.tune(reps=3)
    b.step(5)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    el = time.perf_counter() - t0
    v = cells * steps / el
    done = 10 + steps
    errs = {"Density": [], "Vx": []}
    tdiff = []
    for k in range(0, n, err_every):
        s = fa.Simulation(*specs[k])
        for _ in range(done):
            s.step()
        t64 = s.scalars()["PhysicalTime"]
        tdiff.append(abs(b.clock(k)["PhysicalTime"] - t64) / t64)
        for name in errs:
            a = b.instances[k].field(name)[:, 3:-3, :]
            c = s.field(name)[:, 3:-3, :]
            errs[name].append((float(np.abs(a - c).sum() / np.abs(c).sum()), float(np.abs(a - c).max() / np.abs(c).max())))
        s.close()
    b.close()
    # the plain FP64 path, one disk after the other
    t_loop = 0.0
    for par, ov in specs:
        s = fa.Simulation(par, ov)
        for _ in range(3):
            s.step()
        s.sync()
        t1 = time.perf_counter()
        for _ in range(loop_steps):
            s.step()
        s.sync()
        t_loop += time.perf_counter() - t1
        s.close()
    v_loop = cells * loop_steps / t_loop
    error = {"steps": done, "instances": len(tdiff), "time_rel_max": float("%.2e" % max(tdiff))}
    for name, e in errs.items():
        l1 = [x[0] for x in e]
        error[name] = {"L1_max": float("%.3e" % max(l1)), "L1_median": float("%.3e" % float(np.median(l1))),
                       "Linf_max": float("%.3e" % max(x[1] for x in e))}
    return {"value": round(v, 1), "unit": "cell-updates/s", "instances": n, "cells": cells, "steps": steps,
            "ms_per_step": round(1e3 * el / steps, 4), "dtype": "f32 (f64 sums, residual, cfl, planets)",
            "tuned": {"rows_y": rows_y, "x_tile": tile},
            "step_roofline": {"bytes_per_cell_update": 28, "achieved": round(v * 28 / 1e9, 1), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(v * 28 / 1e9 / HBM_PEAK_GBS, 4)},
            "fp64_loop": {"value": round(v_loop, 1), "steps_per_instance": loop_steps},
            "speedup_vs_fp64_loop": round(v / v_loop, 2),
            "error_vs_fp64": error,
            "workload": "%d 2-D isothermal planet-disk runs, Nx in {256,384,768,1536} x Ny in {128,192,384}, seeded "
                        "planet mass, viscosity and aspect ratio; planets and frame integrated on the device" % n}


def output_rate(sim, reps=3):
    """Time one WriteOutputs(ALL) of every fluid from device state (merged
    raw files into a temporary folder): device-to-host copy plus the file
    writes, best of reps; not part of the timed step."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="f3d_out_")
    try:
        best = None
        for r in range(reps):
            sim.sync()
            t0 = time.perf_counter()
            sim.write_outputs(r, d)
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d)
                     if f.endswith("0.dat") and not f.startswith("planet"))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"what": "planet files + every fluid's dens/energy/vx/vy(/vz) files, merged raw format",
            "bytes": nbytes, "ms": round(1e3 * best, 3), "GB_s": round(nbytes / best / 1e9, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(PAR))
    ap.add_argument("--routines", action="store_true", help="one kernel per routine instead of the fused path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-fp32", action="store_true", help="skip the reduced-precision form")
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched form (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-workloads", dest="workloads", action="store_false",
                    help="skip the other workloads (2-D disk, large adiabatic disk, gas + dust)")
    ap.add_argument("--workload-steps", type=int, default=200)
    ap.add_argument("--error-orbits", type=float, default=10.0,
                    help="orbits at r = 1 of the paired FP64/FP32 runs of the 3-D disk and the large adiabatic "
                         "disk that measure the precision form's error (0: skip them)")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="steps of the paired FP64/FP32 run that measures the FP32 form's error")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run three device-to-device copies of MIB MiB on the library stream (a copy "
                         "kernel of known size that scales the hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU decomposition is a design only "
              "(radial bands with halo exchange), so --gpus %d is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    import fargo3d_amd as fa
    L = fa.lib()
    L.f3d_hip_timing_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.f3d_hip_timing_mask.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    L.f3d_hip_timing_read.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.f3d_hip_stage_name.restype = ctypes.c_char_p
    nst = L.f3d_hip_stage_count()
    names = [L.f3d_hip_stage_name(s).decode() for s in range(nst)]

    parfile = PAR[args.config]
    arch = {}
    if args.routines:
        arch = {"FusedCfl": "no", "FusedSources": "no", "FusedTransport": "no"}
    elapsed, stages, sim = timed_run(fa, L, names, parfile, "", arch, args.steps, args.warmup)
    if args.copy_calibration > 0:
        copy_calibration(L, sim, args.copy_calibration << 20)
    cells = sim.nx * sim.ny * (sim.nz if sim.dimz else 1)  # per fluid: what one stage launch processes
    nfl = sim.nfluids
    ad = 1 if sim.adiabatic else 0
    alpha = 1 if args.config == "c2" else 0
    dimz = sim.dimz
    outputs = output_rate(sim)
    sim.close()

    value = cells * nfl * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    n_arrays = b_alg_arrays(dimz, ad, nfl)
    b_alg = 8 * n_arrays
    roofline = roofline_of(stages, cells, dimz, ad, alpha, 8, args.config, b_alg, nfl)
    step_gbs = value * b_alg / 1e9

    # the reduced-precision form on the same workload (FP32 fields, FP64 sums,
    # residual/shift split and CFL) and its error against the FP64 path
    forms = None
    if not args.routines and not args.no_fp32:
        e32, st32, s32 = timed_run(fa, L, names, parfile, "Precision=32", {}, args.steps, args.warmup)
        s32.close()
        v32 = cells * nfl * args.steps / e32
        forms = {"fp32": {"value": round(v32, 1), "ms_per_step": round(1e3 * e32 / args.steps, 4),
                          "speedup_vs_fp64": round(v32 / value, 3), "dtype": "f32 (f64 sums, residual, cfl)",
                          "step_roofline": {"bytes_per_cell_update": 4 * n_arrays,
                                            "achieved": round(v32 * 4 * n_arrays / 1e9, 1), "peak": HBM_PEAK_GBS,
                                            "unit": "GB/s", "frac": round(v32 * 4 * n_arrays / 1e9 / HBM_PEAK_GBS, 4)},
                          "roofline": roofline_of(st32, cells, dimz, ad, alpha, 4, args.config,
                                                  4 * n_arrays, nfl),
                          "stages": st32,
                          "error_vs_fp64": fp32_errors(fa, parfile, steps=args.error_steps)}}
        if args.error_orbits > 0:
            forms["fp32"]["error_vs_fp64_orbits"] = fp32_errors(fa, parfile, orbits=args.error_orbits)
    if not args.routines and args.batch > 0:
        forms = forms or {}
        forms["batched"] = batched_form(fa, args.batch, args.batch_steps, 10)

    # the other single-disk workloads at their full sizes: the 2-D disk, the
    # large adiabatic disk (FP64 and the precision form) and gas + 4 dust
    workloads = None
    if not args.routines and args.workloads:
        workloads = {}
        for cfg in ("c1", "c3", "c4"):
            if cfg == args.config:
                continue
            workloads[cfg] = {"fp64": workload_run(fa, L, names, cfg, "", args.workload_steps, args.warmup, 8)}
            if cfg in ("c3", "c4"):
                w32 = workload_run(fa, L, names, cfg, "Precision=32", args.workload_steps, args.warmup, 4)
                w32["speedup_vs_fp64"] = round(w32["value"] / workloads[cfg]["fp64"]["value"], 3)
                if cfg == "c3" and args.error_orbits > 0:
                    # the large adiabatic disk's own horizon: ten orbits at r = 1
                    w32["error_vs_fp64_orbits"] = fp32_errors(fa, PAR[cfg], orbits=args.error_orbits)
                workloads[cfg]["fp32"] = w32

    cpu = None
    if not args.no_cpu_baseline:
        n, el = cpu_baseline(parfile, args.cpu_budget)
        cpu = {"value": round(cells * n / el, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": "%d steps of the same workload by the CPU restatement (oracle/), one core, %.1f s"
                         % (n, el)}

    line = {
        "metric": "cell-updates/s on 1 MI355X (FP64 vs FP32+FP64-fix vs batched), % HBM roofline",
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (setup initial conditions)",
        "config": {"workload": WORKLOAD[args.config], "form": "fp64 routines" if args.routines else "fp64 fused",
                   "cells": cells, "fluids": nfl, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": {"bytes_per_cell_update": b_alg, "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "forms": forms,
        "workloads": workloads,
        "outputs": outputs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
