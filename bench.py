"""Benchmark of FARGO3D's hydro time step on one MI355X.

Workload: the 3-D spherical isothermal disk with an embedded planet
(256 azimuth x 96 radius x 32 colatitude cells, FP64, FARGO orbital
advection, alpha viscosity, Stockholm damping), synthetic initial conditions
from the setup (vertical hydrostatic equilibrium), the fused HIP path.

One "step" is one pass of the DT loop body (src/main.c:361-448): CFL time
step, sources, transport, damping and ghost fill, including the one device->
host read of dt the host needs for the planet integrator.  Fields are device
resident before the timed region starts.

Prints one JSON line (see DESIGN.md for the figures it reports).
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
PAR = {"c1": "fargo3d_amd/setups/c1_fargo2d.par", "c2": "fargo3d_amd/setups/c2_p3d.par",
       "c3": "fargo3d_amd/setups/c3_adiabatic2d.par"}
WORKLOAD = {
    "c1": "2-D polar isothermal disk 384x128, FP64, FARGO, nu=1e-5, 1e-3 Mstar planet",
    "c2": "3-D spherical isothermal disk 256x96x32, FP64, FARGO, alpha=1e-3, 3e-4 Mstar planet",
    "c3": "2-D polar adiabatic disk 4096x1024, FP64, FARGO, gamma=1.4, nu=1e-5, 1e-3 Mstar planet",
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
        "sources_a": (1 + nv + alpha) + (ns + nv),           # rho, v (, cs) -> stresses + q
        "sources_b": (3 + nv + ns + nv) + (nv + ad),         # rho, cs/e, pot, v, stresses, q -> v_temp (, e)
        "substep3": (1 + nv) + 1,                            # e, v_temp -> e
        "sweep_z": (1 + nv + ad) + (1 + nm + ad),            # rho, v_temp (, e) -> rho, momenta (, e)
        "sweep_y": ((1 + nm + ad + 1) if dz else (1 + nv + ad)) + (1 + nm + ad),
        "sweep_x": (1 + nm + 1 + ad) + (2 + nm - 2 + ad),    # rho, momenta, vx_temp (, e) -> rho, vx, y/z momenta (, e)
        "finish": (1 + nm - 2) + nv,                         # rho, y/z momenta -> rho, vy (, vz)
        "potential": 1,                                      # -> pot
    }
    return table.get(stage, 0)


# kernel behind each stage (names in the rocprofv3 summaries)
STAGE_KERNEL = {"cfl": ["kf_cfl"], "sources_a": ["kf_src_a"], "sources_b": ["kf_src_b"], "sweep_z": ["kf_sweep_z"],
                "sweep_y": ["kf_sweep_y2", "kf_sweep_y3"], "sweep_x": ["kf_sweep_x", "kf_sweep_x_tile", "kf_vmed"],
                "finish": ["kf_finish"], "potential": ["k_potential"], "substep3": ["kf_substep3"]}
# rocprofv3 summary with the measured HBM traffic of each kernel of this workload
PROFILE = {"c2": "profiles/fused_step_c2_3d_256x96x32.md", "c1": "profiles/fused_step_c1_2d_384x128.md",
           "c3": "profiles/fused_step_c3_2d_adiabatic_4096x1024.md"}


def measured_traffic(config, stage):
    """HBM bytes per launch of the stage's kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, PROFILE.get(config, ""))
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch = "):
            table = json.loads(line.split("=", 1)[1])
            # every kernel the stage launches (e.g. the VxMed pass of the tiled X sweep)
            hits = [v for name, v in table.items()
                    if name.split("<")[0] in STAGE_KERNEL.get(stage, [])]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(PAR))
    ap.add_argument("--routines", action="store_true", help="one kernel per routine instead of the fused path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
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
    L.f3d_hip_timing_read.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.f3d_hip_stage_name.restype = ctypes.c_char_p
    nst = L.f3d_hip_stage_count()
    names = [L.f3d_hip_stage_name(s).decode() for s in range(nst)]

    parfile = PAR[args.config]
    arch = {}
    if args.routines:
        arch = {"FusedCfl": "no", "FusedSources": "no", "FusedTransport": "no"}
    sim = fa.Simulation(parfile, arch=arch)
    cells = sim.nx * sim.ny * (sim.nz if sim.dimz else 1)
    ad = 1 if sim.adiabatic else 0
    alpha = 1 if args.config == "c2" else 0
    for _ in range(args.warmup):
        sim.step()
    sim.sync()
    L.f3d_hip_timing_enable(sim.ctx, 1)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
    sim.sync()
    elapsed = time.perf_counter() - t0
    ms = (ctypes.c_double * nst)()
    cnt = (ctypes.c_long * nst)()
    L.f3d_hip_timing_read(sim.ctx, ms, cnt)
    L.f3d_hip_timing_enable(sim.ctx, 0)
    if args.copy_calibration > 0:
        copy_calibration(L, sim, args.copy_calibration << 20)

    value = cells * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    # bytes per cell-update of the state (FP64): rho, v (, e) read and written,
    # the sound speed read (isothermal)
    nv = 3 if sim.dimz else 2
    b_alg = 8 * (2 * (1 + nv + ad) + (0 if ad else 1))
    stages = {}
    for s in range(nst):
        if cnt[s]:
            stages[names[s]] = {"avg_us": round(1e3 * ms[s] / cnt[s], 3), "launches": int(cnt[s])}
    roofline = None
    if stages:
        dom = max(stages, key=lambda n: stages[n]["avg_us"] * stages[n]["launches"])
        per_cell = 8 * stage_arrays(dom, sim.dimz, ad, alpha)
        avg_s = stages[dom]["avg_us"] * 1e-6
        achieved = per_cell * cells / avg_s / 1e9 if avg_s > 0 else 0.0
        roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(args.config, dom), "bytes_per_cell": per_cell,
                    "bytes_per_launch": per_cell * cells}
    step_gbs = value * b_alg / 1e9

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
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": {"bytes_per_cell_update": b_alg, "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
