/*
 * hip_bind.c -- the binding INTEGRATION.md describes, complete for the hydro
 * step, compiled against FARGO3D's pointer types (fargo3d_api.h restates
 * src/global.h:303-416).  TEST INFRASTRUCTURE: the binding test builds it
 * into tests/bind/lib with the miniature host of fargo_host.c.
 *
 * Each *_hip wrapper has the reference pointer's exact type and calls the
 * library entry point of the same name with the fields' device pointers
 * (Field->field_gpu).  A failed call records the library's message in
 * HipError, where FARGO3D's build would call prs_error (src/LowTasks.c:3-6),
 * the fatal path check_errors takes after a failed launch (src/lowtasks.cu:6-15).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "fargo3d_api.h"
#include "fargo3d_hip.h"

static f3d_hip_ctx *H;
char HipError[512];

static void hip_fail(const char *what) {
  if (!HipError[0]) snprintf(HipError, sizeof(HipError), "fargo3d_amd: %s: %s", what, H ? f3d_hip_last_error(H) : "");
}
#define HIPCALL(name, call) \
  do {                      \
    if (call) hip_fail(name); \
  } while (0)

/* after InitSpace: mesh sizes and 1-D arrays (src/light_global_dev.c:5-23) */
void HipInit(void) {
  f3d_mesh_desc d;
  memset(&d, 0, sizeof(d));
  d.nx = Nx; d.ny = Ny; d.nz = Nz; d.nghy = NGHY; d.nghz = NGHZ;
  d.dimz = GRID.dimz;
  d.geometry = GRID.geometry;
  d.pitch = Pitch_gpu; d.stride = Stride_gpu;
  d.XMIN = XMIN; d.XMAX = XMAX; d.mindx = Mindx;
  d.xmin = Xmin; d.ymin = Ymin; d.zmin = Zmin; d.xmed = Xmed; d.ymed = Ymed; d.zmed = Zmed;
  d.sxi = Sxi; d.sxj = Sxj; d.syj = Syj; d.szj = Szj; d.sxk = Sxk; d.syk = Syk; d.szk = Szk;
  d.invvj = InvVj; d.invdiffxmed = InvDiffXmed;
  /* the cos/sin tables, evaluated on the host with the CPU build's libm */
  d.cosxmed = GRID.cosxmed; d.sinxmed = GRID.sinxmed;
  d.sinzmed = GRID.sinzmed; d.coszmed = GRID.coszmed; d.sinzmin = GRID.sinzmin; d.coszmin = GRID.coszmin;
  if (f3d_hip_create(&d, &H)) snprintf(HipError, sizeof(HipError), "fargo3d_amd: no HIP device");
}

/* the host/device copies of src/fresh.c's INPUT/OUTPUT on the GPU build */
void HipToDevice(void *dst, const void *src, size_t bytes) {
  HIPCALL("copy to device", f3d_hip_memcpy_h2d(H, dst, src, bytes));
}
void HipToHost(void *dst, const void *src, size_t bytes) { HIPCALL("copy to host", f3d_hip_memcpy_d2h(H, dst, src, bytes)); }

void HipFini(void) {
  if (H) f3d_hip_destroy(H);
  H = NULL;
}

/* SelectFluid (src/LowTasks.c:565-593) calls this */
void HipBindFluid(void) {
  f3d_fields f;
  memset(&f, 0, sizeof(f));
  f.Density = Density->field_gpu; f.Energy = Energy->field_gpu;
  f.Vx = Vx->field_gpu; f.Vy = Vy->field_gpu; f.Vx_temp = Vx_temp->field_gpu; f.Vy_temp = Vy_temp->field_gpu;
  if (GRID.dimz) {
    f.Vz = Vz->field_gpu;
    f.Vz_temp = Vz_temp->field_gpu;
  }
  f.Mpx = Mpx->field_gpu; f.Mmx = Mmx->field_gpu; f.Mpy = Mpy->field_gpu; f.Mmy = Mmy->field_gpu;
  f.Mpz = Mpz->field_gpu; f.Mmz = Mmz->field_gpu; f.Pot = Pot->field_gpu; f.Slope = Slope->field_gpu;
  f.DivRho = DivRho->field_gpu; f.DensStar = DensStar->field_gpu; f.Qs = Qs->field_gpu;
  f.Pressure = Pressure->field_gpu; f.Total_Density = Total_Density->field_gpu;
  f.QL = QL->field_gpu; f.QR = QR->field_gpu;
  f.Density0 = Density0->field_gpu; f.Energy0 = Energy0->field_gpu; f.Vx0 = Vx0->field_gpu;
  f.Vy0 = Vy0->field_gpu; f.Vz0 = Vz0->field_gpu;
  f.Reduction2D = Reduction2D->field_gpu; f.VxMed = VxMed->field_gpu; f.Nshift = Nshift->field_gpu;
  HIPCALL("bind", f3d_hip_bind(H, &f));
}

/* the run parameters and the host scalars the kernels read: param.h values,
   OMEGAFRAME, the boundary factors and the planets of Sys */
static void HipPhys(void) {
  f3d_phys p;
  memset(&p, 0, sizeof(p));
  p.fluidtype = Fluidtype; p.eos = P.eos; p.visc = P.visc; p.potential = P.potential; p.fargo = P.fargo;
  p.stockholm = P.stockholm; p.indirectterm = P.indirectterm; p.periodicz = P.periodicz;
  p.periodicy = P.periodicy; p.bc = P.bc; p.energy_bc = P.energy_bc;
  p.nu = P.nu; p.alpha = P.alpha; p.gamma = P.gamma; p.cfl = P.cfl;
  p.OMEGAFRAME = OMEGAFRAME; p.OMEGAFRAME0 = OMEGAFRAME0; p.PhysicalTime = PhysicalTime;
  p.ymin_par = P.ymin; p.ymax_par = P.ymax; p.zmin_par = P.zmin; p.zmax_par = P.zmax; p.taudamp = P.taudamp;
  p.Y_inf = DAMP.Y_inf; p.Y_sup = DAMP.Y_sup; p.Z_inf = DAMP.Z_inf; p.Z_sup = DAMP.Z_sup;
  for (int q = 0; q < 3; q++) {
    p.ymin_dens[q] = BCT.ymin_dens[q]; p.ymax_dens[q] = BCT.ymax_dens[q];
    p.ymin_vsq[q] = BCT.ymin_vsq[q]; p.ymax_vsq[q] = BCT.ymax_vsq[q];
    p.ymin_vfac[q] = BCT.ymin_vfac[q]; p.ymax_vfac[q] = BCT.ymax_vfac[q];
    p.zmin_vfac[q] = BCT.zmin_vfac[q]; p.zmax_vfac[q] = BCT.zmax_vfac[q];
  }
  p.ymin_dexp = BCT.ymin_dexp; p.ymax_dexp = BCT.ymax_dexp; p.zmin_dens = BCT.zmin_dens; p.zmax_dens = BCT.zmax_dens;
  const double taper = f3d_planet_taper(&P, PhysicalTime);
  p.nplanets = Sys.nb;
  for (int n = 0; n < Sys.nb && n < 16; n++) {
    p.px[n] = Sys.x[n]; p.py[n] = Sys.y[n]; p.pz[n] = Sys.z[n];
    p.pmass[n] = Sys.mass[n] * taper;
    p.psmooth2[n] = f3d_planet_smoothing2(&Sys, &P, n, taper);
  }
  HIPCALL("set_phys", f3d_hip_set_phys(H, &p));
}

/* ---- one wrapper per pointer (the data is already on the device) ---- */
static void ComputePressureFieldIso_hip() { HIPCALL("ComputePressureFieldIso", f3d_ComputePressureFieldIso(H)); }
static void ComputePressureFieldAd_hip() { HipPhys(); HIPCALL("ComputePressureFieldAd", f3d_ComputePressureFieldAd(H)); }
static void SubStep1_x_hip(real dt) { HipPhys(); HIPCALL("SubStep1_x", f3d_SubStep1_x(H, dt)); }
static void SubStep1_y_hip(real dt) { HIPCALL("SubStep1_y", f3d_SubStep1_y(H, dt)); }
static void SubStep1_z_hip(real dt) { HIPCALL("SubStep1_z", f3d_SubStep1_z(H, dt)); }
static void SubStep2_a_hip(real dt) { HIPCALL("SubStep2_a", f3d_SubStep2_a(H, dt)); }
static void SubStep2_b_hip(real dt) { HIPCALL("SubStep2_b", f3d_SubStep2_b(H, dt)); }
static void SubStep3_hip(real dt) { HIPCALL("SubStep3", f3d_SubStep3(H, dt)); }
static void DivideByRho_hip(Field *Q) { HIPCALL("DivideByRho", f3d_DivideByRho(H, Q->field_gpu)); }
static void VanLeerX_a_hip(Field *Q) { HIPCALL("VanLeerX_a", f3d_VanLeerX_a(H, Q->field_gpu)); }
static void VanLeerX_b_hip(real dt, Field *Q, Field *Qs_, Field *Vx_t) {
  HIPCALL("VanLeerX_b", f3d_VanLeerX_b(H, dt, Q->field_gpu, Qs_->field_gpu, Vx_t->field_gpu));
}
static void VanLeerY_a_hip(Field *Q) { HIPCALL("VanLeerY_a", f3d_VanLeerY_a(H, Q->field_gpu)); }
static void VanLeerY_b_hip(real dt, Field *Q, Field *Qs_) {
  HIPCALL("VanLeerY_b", f3d_VanLeerY_b(H, dt, Q->field_gpu, Qs_->field_gpu));
}
static void VanLeerZ_a_hip(Field *Q) { HIPCALL("VanLeerZ_a", f3d_VanLeerZ_a(H, Q->field_gpu)); }
static void VanLeerZ_b_hip(real dt, Field *Q, Field *Qs_) {
  HIPCALL("VanLeerZ_b", f3d_VanLeerZ_b(H, dt, Q->field_gpu, Qs_->field_gpu));
}
static void momenta_x_hip() { HipPhys(); HIPCALL("momenta_x", f3d_momenta_x(H)); }
static void momenta_y_hip() { HIPCALL("momenta_y", f3d_momenta_y(H)); }
static void momenta_z_hip() { HIPCALL("momenta_z", f3d_momenta_z(H)); }
static void reduction_SUM_hip(Field *F, int ymin, int ymax, int zmin, int zmax) {
  HIPCALL("reduction_SUM", f3d_reduction_SUM(H, F->field_gpu, ymin, ymax, zmin, zmax));
}
static void reduction_MIN_hip(Field *F, int ymin, int ymax, int zmin, int zmax) {
  HIPCALL("reduction_MIN", f3d_reduction_MIN(H, F->field_gpu, ymin, ymax, zmin, zmax));
}
static void UpdateX_hip(real dt, Field *Q, Field *Qs_, Field *Vx_t) {
  HIPCALL("UpdateX", f3d_UpdateX(H, dt, Q->field_gpu, Qs_->field_gpu, Vx_t->field_gpu));
}
static void UpdateY_hip(real dt, Field *Q, Field *Qs_) { HIPCALL("UpdateY", f3d_UpdateY(H, dt, Q->field_gpu, Qs_->field_gpu)); }
static void UpdateZ_hip(real dt, Field *Q, Field *Qs_) { HIPCALL("UpdateZ", f3d_UpdateZ(H, dt, Q->field_gpu, Qs_->field_gpu)); }
static void UpdateDensityX_hip(real dt, Field *Q, Field *Vx_t) {
  HIPCALL("UpdateDensityX", f3d_UpdateDensityX(H, dt, Q->field_gpu, Vx_t->field_gpu));
}
static void UpdateDensityY_hip(real dt, Field *Q) { HIPCALL("UpdateDensityY", f3d_UpdateDensityY(H, dt, Q->field_gpu)); }
static void UpdateDensityZ_hip(real dt, Field *Q) { HIPCALL("UpdateDensityZ", f3d_UpdateDensityZ(H, dt, Q->field_gpu)); }
static void NewVelocity_x_hip() { HipPhys(); HIPCALL("NewVelocity_x", f3d_NewVelocity_x(H)); }
static void NewVelocity_y_hip() { HIPCALL("NewVelocity_y", f3d_NewVelocity_y(H)); }
static void NewVelocity_z_hip() { HIPCALL("NewVelocity_z", f3d_NewVelocity_z(H)); }
static void AdvectSHIFT_hip(Field *F, FieldInt2D *NS) { HIPCALL("AdvectSHIFT", f3d_AdvectSHIFT(H, F->field_gpu, NS->field_gpu)); }
static void ComputeResidual_hip(real dt) { HIPCALL("ComputeResidual", f3d_ComputeResidual(H, dt)); }
static void ComputeVweight_hip(Field *V, Field *Q) { HIPCALL("ComputeVweight", f3d_ComputeVweight(H, V->field_gpu, Q->field_gpu)); }
static void Potential_hip() { HipPhys(); HIPCALL("Potential", f3d_Potential(H)); }
static void CorrectVtheta_hip(real domega) { HIPCALL("CorrectVtheta", f3d_CorrectVtheta(H, domega)); }
static void cfl_hip(void) {
  double m;
  HipPhys();
  HIPCALL("cfl", f3d_cfl(H, &m));
  Min[FluidIndex] = m; /* the value cfl_b leaves in Min[] (src/cfl_b.c:6-37) */
}
static void ComputeForce_hip(real x, real y, real z, real rsm, real mass) {
  HIPCALL("_ComputeForce", f3d_ComputeForce(H, x, y, z, rsm, mass, localforce));
}
static void copy_velocities_hip(int option) { HIPCALL("copy_velocities", f3d_copy_velocities(H, option)); }
static void VanLeerX_PPA_a_hip(Field *Q) { HIPCALL("VanLeerX_PPA_a", f3d_VanLeerX_PPA_a(H, Q->field_gpu)); }
static void VanLeerX_PPA_b_hip(Field *Q) { HIPCALL("VanLeerX_PPA_b", f3d_VanLeerX_PPA_b(H, Q->field_gpu)); }
static void VanLeerX_PPA_c_hip(Field *Q) { HIPCALL("VanLeerX_PPA_c", f3d_VanLeerX_PPA_c(H, Q->field_gpu)); }
static void VanLeerX_PPA_d_hip(real dt, Field *Q, Field *Qs_, Field *Vx_t) {
  HIPCALL("VanLeerX_PPA_d", f3d_VanLeerX_PPA_d(H, dt, Q->field_gpu, Qs_->field_gpu, Vx_t->field_gpu));
}
static void Reset_field_hip(Field *F) { HIPCALL("Reset_field", f3d_Reset_field(H, F->field_gpu)); }
static void ComputeTotalDensity_hip() { HIPCALL("ComputeTotalDensity", f3d_ComputeTotalDensity(H)); }
static void StockholmBoundary_hip(real dt) { HipPhys(); HIPCALL("StockholmBoundary", f3d_StockholmBoundary(H, dt)); }
static void visctensor_cyl_hip() { HipPhys(); HIPCALL("visctensor_cyl", f3d_visctensor_cyl(H)); }
static void addviscosity_cyl_hip(real dt) { HIPCALL("addviscosity_cyl", f3d_addviscosity_cyl(H, dt)); }
static void visctensor_sph_hip() { HipPhys(); HIPCALL("visctensor_sph", f3d_visctensor_sph(H)); }
static void addviscosity_sph_hip(real dt) { HIPCALL("addviscosity_sph", f3d_addviscosity_sph(H, dt)); }
/* boundaries() (src/boundary.c:2-23) calls zmin, zmax, ymin, ymax in turn;
   f3d_boundaries fills the four sides in that order in one call, so it is
   bound to the last of them and the others are empty */
static void boundary_none_hip() {}
static void boundary_all_hip() { HipPhys(); HIPCALL("boundaries", f3d_boundaries(H)); }

/* ChangeArch with every hydro group of std/func_arch.cfg set to HIP
   (src/change_arch.c:181-567) */
void ChangeArchHip(void) {
  ComputePressureFieldIso = ComputePressureFieldIso_hip;
  ComputePressureFieldAd = ComputePressureFieldAd_hip;
  SubStep1_x = SubStep1_x_hip; SubStep1_y = SubStep1_y_hip; SubStep1_z = SubStep1_z_hip;
  SubStep2_a = SubStep2_a_hip; SubStep2_b = SubStep2_b_hip; SubStep3 = SubStep3_hip;
  DivideByRho = DivideByRho_hip;
  VanLeerX_a = VanLeerX_a_hip; VanLeerX_b = VanLeerX_b_hip;
  VanLeerY_a = VanLeerY_a_hip; VanLeerY_b = VanLeerY_b_hip;
  VanLeerZ_a = VanLeerZ_a_hip; VanLeerZ_b = VanLeerZ_b_hip;
  momenta_x = momenta_x_hip; momenta_y = momenta_y_hip; momenta_z = momenta_z_hip;
  reduction_SUM = reduction_SUM_hip; reduction_MIN = reduction_MIN_hip;
  UpdateX = UpdateX_hip; UpdateY = UpdateY_hip; UpdateZ = UpdateZ_hip;
  UpdateDensityX = UpdateDensityX_hip; UpdateDensityY = UpdateDensityY_hip; UpdateDensityZ = UpdateDensityZ_hip;
  NewVelocity_x = NewVelocity_x_hip; NewVelocity_y = NewVelocity_y_hip; NewVelocity_z = NewVelocity_z_hip;
  AdvectSHIFT = AdvectSHIFT_hip; ComputeResidual = ComputeResidual_hip; ComputeVweight = ComputeVweight_hip;
  Potential = Potential_hip; CorrectVtheta = CorrectVtheta_hip; cfl = cfl_hip; _ComputeForce = ComputeForce_hip;
  copy_velocities = copy_velocities_hip;
  VanLeerX_PPA_a = VanLeerX_PPA_a_hip; VanLeerX_PPA_b = VanLeerX_PPA_b_hip;
  VanLeerX_PPA_c = VanLeerX_PPA_c_hip; VanLeerX_PPA_d = VanLeerX_PPA_d_hip;
  Reset_field = Reset_field_hip; ComputeTotalDensity = ComputeTotalDensity_hip;
  StockholmBoundary = StockholmBoundary_hip;
  visctensor_cyl = visctensor_cyl_hip; addviscosity_cyl = addviscosity_cyl_hip;
  visctensor_sph = visctensor_sph_hip; addviscosity_sph = addviscosity_sph_hip;
  boundary_zmin[0] = boundary_none_hip; boundary_zmax[0] = boundary_none_hip;
  boundary_ymin[0] = boundary_none_hip; boundary_ymax[0] = boundary_all_hip;
}
