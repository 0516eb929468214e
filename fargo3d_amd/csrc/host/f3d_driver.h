/*
 * f3d_driver.h -- C step driver of the MI355X build (the product's host side).
 *
 * Keeps FARGO3D's host structure: the .par-driven set-up (src/main.c:176-316),
 * the DT loop (src/main.c:361-448), Sources/Transport (src/algogas.c:30-159,
 * src/transport.c:69-193) and the per-routine function-pointer table selected
 * by a func_arch.cfg-style file (src/change_arch.c:146-570).  Every routine
 * in the table runs on the GPU through the C ABI of include/fargo3d_hip.h;
 * fields are device-resident for the whole run.
 */
#ifndef F3D_DRIVER_H
#define F3D_DRIVER_H

#include "fargo3d_hip.h"
#include "f3d_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct f3d_sim f3d_sim;

f3d_sim *f3d_sim_create(const char *parfile, const char *overrides, char *err, int errlen);
void f3d_sim_destroy(f3d_sim *S);
const char *f3d_sim_error(f3d_sim *S);

/* One pass of the DT loop body; *dt_out receives the time step used. */
int f3d_sim_step(f3d_sim *S, double *dt_out);
/* Use dt (> 0) as the next step's CFL time step instead of the computed one
 * (no counterpart in the reference: lets two runs, e.g. FP64 and FP32, share
 * one dt sequence so that their difference measures arithmetic alone). */
int f3d_sim_force_dt(f3d_sim *S, double dt);
int f3d_sim_run(f3d_sim *S, int nsteps);

/* func_arch table: group names as std/func_arch.cfg; modes "GPU"/"HIP"
 * (one kernel per routine) or, for the stage groups "FusedCfl",
 * "FusedSources", "FusedTransport", "yes"/"no".  "CPU" is rejected. */
int f3d_sim_set_arch(f3d_sim *S, const char *group, const char *mode);
int f3d_sim_read_arch(f3d_sim *S, const char *path);

/* field access (host copies of the device arrays) */
long f3d_sim_field_size(f3d_sim *S, const char *name);
int f3d_sim_get_field(f3d_sim *S, const char *name, double *host);
int f3d_sim_set_field(f3d_sim *S, const char *name, const double *host);
int f3d_sim_get_nshift(f3d_sim *S, int *host);
int f3d_sim_set_nshift(f3d_sim *S, const int *host);
/* planets as 8 doubles each: x y z vx vy vz mass (unused) */
int f3d_sim_set_planets(f3d_sim *S, const double *in, int n);
double *f3d_sim_device_field(f3d_sim *S, const char *name);
void f3d_sim_get_scalars(f3d_sim *S, double out[8]);
void f3d_sim_set_scalars(f3d_sim *S, const double in[8]);
void f3d_sim_get_planets(f3d_sim *S, double *out, int maxn);
void f3d_sim_shape(f3d_sim *S, int out[4]); /* nx, ny, nz, dimz */
int f3d_sim_eos(f3d_sim *S);                 /* 1 isothermal, 2 adiabatic */
/* SelectFluid (src/LowTasks.c:565): field access, set_arch-independent calls and
 * f3d_sim_ctx's bound fields then refer to fluid n (0 = gas). */
int f3d_sim_nfluids(f3d_sim *S);
int f3d_sim_select_fluid(f3d_sim *S, int n);
double f3d_sim_dt_out(f3d_sim *S);           /* DT, the output interval the step clamps to */
f3d_hip_ctx *f3d_sim_ctx(f3d_sim *S);
/* read-only views of the set-up, the planetary system and the clocks (the
   batched form copies them to the device) */
const f3d_params *f3d_sim_params(f3d_sim *S);
const f3d_psys *f3d_sim_psys(f3d_sim *S);
const f3d_clock *f3d_sim_clock(f3d_sim *S);
/* refresh the kernels' copy of the run parameters and host scalars */
int f3d_sim_push_phys(f3d_sim *S);

/* ---- outputs and restart ----
 * Files in FARGO3D's merged raw format (WriteMerging, src/output.c:283-315):
 * active cells, x fastest, native-endian doubles, no header.  The precision
 * form writes and reads doubles too. */
int f3d_sim_write_field(f3d_sim *S, const char *name, const char *path);
/* WritePlanetSystemFile + WriteOutputs(ALL) of every fluid (src/main.c:328-334,
 * src/output.c:551-700): <dir>/planet<k>.dat gets the line of output n, and
 * <dir>/<fluid>{dens,energy,vx,vy,vz}<n>.dat are written (no energy for dust,
 * vz in 3-D only; fluids "gas", "dust1", ...).  tqwk<k>.dat is not written. */
int f3d_sim_write_outputs(f3d_sim *S, int n, const char *dir);
/* start of a run (src/main.c:262-266, src/init_stockholm.c:106-124): empties
 * planet<k>.dat and writes the damping references <fluid><q>0_2d.dat */
int f3d_sim_write_init(f3d_sim *S, const char *dir);
/* restart from output n of <dir> (src/main.c:249-260, RestartSimulation
 * src/LowTasks.c:816-881 with Restart_Full, RestartPlanetarySystem
 * src/output.c:112-123): the active cells of every fluid's primitive fields,
 * PhysicalTime, OMEGAFRAME and the planets from planet<k>.dat, then the ghost
 * fill and init_stockholm (references from the 0_2d files when present).
 * Returns begin_i = n * NINTERM, or -1. */
int f3d_sim_restart(f3d_sim *S, int n, const char *dir);
/* one pass of the DT loop: steps until the output interval DT is reached;
 * returns the number of steps, or -1 */
int f3d_sim_advance(f3d_sim *S);
/* monitored quantities: the F3D_MON_* bits of f3d_host.h */
/* MonitorGlobal of every fluid (src/monitor.c:236-254): each quantity of
 * m2d | myraw | mscalar, per planet for TORQ, summed on the device into
 * Reduction2D (ring sums in sequential i order) and from there on the host
 * in the reference's order, written under <dir>/monitor/<fluid>/: the scalar
 * line "<time>\t<sum>" of <q>[_planet_<k>].dat, the raw Y profile appended to
 * <q>_1d_Y_raw[_planet_<k>].dat, the 2-D map FG<timestep>/<q>_2d_<counter>.dat.
 * The precision form computes the cell values and every sum in FP64. */
int f3d_sim_monitor(f3d_sim *S, int m2d, int myraw, int mscalar, int timestep, const char *dir);
/* the main loop from begin_i to NTOT with outputs every NINTERM passes
 * (src/main.c:324-356) */
int f3d_sim_main_loop(f3d_sim *S, int begin_i, const char *dir);
/* the file layout on plain host arrays (pitch, stride in elements; a 2-D
 * grid passes nz = 1, nghz = 0); read fills the active cells only and fails
 * unless the file holds exactly nx*ny*nz doubles */
int f3d_write_active(const double *h, int nx, int ny, int nz, int nghy, int nghz, long pitch, long stride,
                     const char *path);
int f3d_read_active(double *h, int nx, int ny, int nz, int nghy, int nghz, long pitch, long stride,
                    const char *path);

/* ---- batched form (no counterpart in the reference): n independent 2-D
 * isothermal disks, each its own .par set-up, run in the FP32 form; every
 * stage of the step is one launch over all instances and one step is one
 * hipGraph replay, time steps included (per-instance CFL and output-time
 * clamp on the device), and so is each instance's planetary system (FP64 on
 * the device: the potential from the planets' positions, RK5, the frame that
 * follows the first planet, CorrectVtheta).  Instances are read through
 * f3d_batch_instance + f3d_sim_get_field; their planets and frame rate through
 * f3d_batch_planets (the instance's host-side planets stay at their initial
 * state). */
typedef struct f3d_batch f3d_batch;
f3d_batch *f3d_batch_create(int n, const char *const *parfiles, const char *const *overrides, char *err,
                            int errlen);
void f3d_batch_destroy(f3d_batch *B);
const char *f3d_batch_error(f3d_batch *B);
int f3d_batch_size(f3d_batch *B);
f3d_sim *f3d_batch_instance(f3d_batch *B, int i);
int f3d_batch_step(f3d_batch *B, int nsteps);
int f3d_batch_sync(f3d_batch *B);                        /* also reports a null time step */
int f3d_batch_clock(f3d_batch *B, int i, double out[3]); /* dt, dtemp, PhysicalTime */
/* planets of instance i as f3d_sim_get_planets lays them out (8 doubles per
 * planet: x, y, z, vx, vy, vz, mass, count), at most maxn; *omegaframe gets
 * the frame rate.  Returns the planet count, or -1/-2 on error. */
int f3d_batch_planets(f3d_batch *B, int i, double *out, int maxn, double *omegaframe);
/* tuning knobs (rows per thread of the Y sweep; 64/128/256-cell X tiles) and
 * the autotuner that times every combination on the batch itself, restoring
 * the state afterwards */
int f3d_batch_set_tiling(f3d_batch *B, int rows_y, int tile);
int f3d_batch_tune(f3d_batch *B, int reps, int *best_rows_y, int *best_tile);

#ifdef __cplusplus
}
#endif
#endif
