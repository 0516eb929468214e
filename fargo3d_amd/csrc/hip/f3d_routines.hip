// f3d_routines.hip -- one gfx950 kernel per routine of FARGO3D's hydro step.
//
// These are the drop-in entry points: each C function below has the name and
// arguments of the reference routine it replaces (cited per function) and
// computes the same cells over the same loop ranges, with the same
// arithmetic in the same order (see f3d_device.h), so that results match
// the sequential CPU build bit for bit.  Threads map to cells with i
// (azimuth, contiguous in memory) fastest, so every wavefront reads 64
// consecutive doubles of a ring.  The fused, fewer-pass kernels of the fast
// path live in f3d_fused.hip and reuse the per-cell arithmetic written here.
#include "f3d_internal.h"
#include "f3d_cell.h"
#include "../host/f3d_host.h"

using namespace f3d;

namespace {

constexpr int BX = 128;
// cells per thread of the potential kernels: four on rings of 1024 cells or
// more, else one (shorter rings would leave threads idle)
inline int pot_cells(const f3d_hip_ctx* c) { return c->m.nx >= 8 * BX ? 4 : 1; }

struct Range {
  int jb, je, kb, ke;
};

inline dim3 grid_of(const f3d_hip_ctx* c, Range r) {
  return dim3((c->m.nx + BX - 1) / BX, (unsigned)(r.je - r.jb), (unsigned)(r.ke - r.kb));
}
inline bool empty(Range r) { return r.je <= r.jb || r.ke <= r.kb; }

// k range of a routine: [kb, size_z) with Z, the single plane k = 0 without.
inline Range rng(const f3d_hip_ctx* c, int jb, int je, int kb, int ke) {
  if (!c->m.dimz) return Range{jb, je, 0, 1};
  return Range{jb, je, kb, ke};
}

}  // namespace

#define CELL_INDEX                                       \
  const int i = blockIdx.x * blockDim.x + threadIdx.x;   \
  if (i >= m.nx) return;                                 \
  const int j = r.jb + (int)blockIdx.y;                  \
  const int k = r.kb + (int)blockIdx.z;                  \
  const int l = i + j * m.pitch + k * m.stride;          \
  (void)l

#define DISPATCH(ctx, KERN, r, ...) DISPATCH_X(ctx, KERN, r, 1, __VA_ARGS__)
// the same with cpt cells per thread along x (a workgroup covers cpt * BX)
#define DISPATCH_X(ctx, KERN, r, cpt, ...)                                                       \
  do {                                                                                           \
    if (empty(r)) break;                                                                         \
    dim3 g_ = grid_of(ctx, r);                                                                   \
    g_.x = ((ctx)->m.nx + (cpt) * BX - 1) / ((cpt) * BX);                                        \
    if ((ctx)->m.geom == SPH) {                                                                  \
      if ((ctx)->m.dimz)                                                                         \
        hipLaunchKernelGGL((KERN<SPH, 1>), g_, dim3(BX), 0, (ctx)->stream, (ctx)->m, r, __VA_ARGS__); \
      else                                                                                       \
        hipLaunchKernelGGL((KERN<SPH, 0>), g_, dim3(BX), 0, (ctx)->stream, (ctx)->m, r, __VA_ARGS__); \
    } else {                                                                                     \
      if ((ctx)->m.dimz)                                                                         \
        hipLaunchKernelGGL((KERN<CYL, 1>), g_, dim3(BX), 0, (ctx)->stream, (ctx)->m, r, __VA_ARGS__); \
      else                                                                                       \
        hipLaunchKernelGGL((KERN<CYL, 0>), g_, dim3(BX), 0, (ctx)->stream, (ctx)->m, r, __VA_ARGS__); \
    }                                                                                            \
    F3D_LAUNCH_CHECK(ctx);                                                                       \
  } while (0)

#define NEED(ptr, name) F3D_REQUIRE(ctx, (ptr) != nullptr, "field " name " is not bound")

// ------------------------------------------------------------------ kernels

template <int G, int DZ>
__global__ void k_pres_iso(dmesh m, Range r, const double* dens, const double* cs, double* pres) {
  CELL_INDEX;
  pres[l] = dens[l] * cs[l] * cs[l];
}

template <int G, int DZ>
__global__ void k_pres_ad(dmesh m, Range r, const double* e, double* pres, double gamma) {
  CELL_INDEX;
  pres[l] = (gamma - 1.0) * e[l];
}

template <int G, int DZ>
__global__ void k_totaldens(dmesh m, Range r, const double* dens, double* td) {
  CELL_INDEX;
  td[l] += dens[l];
}

template <int G, int DZ>
__global__ void k_fill(dmesh m, Range r, double* f, double v) {
  CELL_INDEX;
  f[l] = v;
}

template <int G, int DZ>
__global__ void k_substep1_x(dmesh m, Range r, double dt, const double* p, const double* rho,
                             const double* pot, const double* vx, double* vx_temp, int dust, int potential) {
  CELL_INDEX;
  vx_temp[l] = cell::substep1_x<G>(m, i, j, k, dt, cell::gl(rho, m, l, i), cell::gl(p, m, l, i),
                                   cell::gl(pot, m, l, i), cell::gl(vx, m, l, i), dust, potential);
}

template <int G, int DZ>
__global__ void k_substep1_y(dmesh m, Range r, double dt, const double* p, const double* rho,
                             const double* pot, const double* vx, const double* vy, const double* vz,
                             double* vy_temp, int dust, int potential, double omegaframe) {
  CELL_INDEX;
  vy_temp[l] = cell::substep1_y<G, DZ>(m, i, j, k, dt, cell::gl(rho, m, l, i), cell::gl(p, m, l, i),
                                       cell::gl(pot, m, l, i), cell::gl(vx, m, l, i), cell::gl(vy, m, l, i),
                                       cell::gl(DZ ? vz : vy, m, l, i), dust, potential, omegaframe);
}

template <int G, int DZ>
__global__ void k_substep1_z(dmesh m, Range r, double dt, const double* p, const double* rho,
                             const double* pot, const double* vx, const double* vz, double* vz_temp,
                             int dust, int potential, double omegaframe) {
  CELL_INDEX;
  vz_temp[l] = cell::substep1_z<G>(m, i, j, k, dt, cell::gl(rho, m, l, i), cell::gl(p, m, l, i),
                                   cell::gl(pot, m, l, i), cell::gl(vx, m, l, i), cell::gl(vz, m, l, i), dust,
                                   potential, omegaframe);
}

template <int G, int DZ>
__global__ void k_substep2_a(dmesh m, Range r, const double* rho, const double* vx, const double* vy,
                             const double* vz, double* px, double* py, double* pz) {
  CELL_INDEX;
  const int lxp = lxp_of(l, i, m.nx);
  px[l] = cell::art_q(rho[l], vx[lxp] - vx[l]);
  py[l] = cell::art_q(rho[l], vy[l + m.pitch] - vy[l]);
  if (DZ) pz[l] = cell::art_q(rho[l], vz[l + m.stride] - vz[l]);
}

template <int G, int DZ>
__global__ void k_substep2_b(dmesh m, Range r, double dt, const double* rho, const double* vx,
                             const double* vy, const double* vz, double* vxt, double* vyt, double* vzt,
                             const double* px, const double* py, const double* pz, double* e, int adiabatic) {
  CELL_INDEX;
  double ax, ay, az, de;
  const double* vzs = DZ ? vz : vy;
  const double* pzs = DZ ? pz : py;
  cell::substep2_b<G, DZ>(m, i, j, k, dt, cell::gl(rho, m, l, i), cell::gl(vx, m, l, i), cell::gl(vy, m, l, i),
                          cell::gl(vzs, m, l, i), cell::gl(px, m, l, i), cell::gl(py, m, l, i),
                          cell::gl(pzs, m, l, i), adiabatic, vxt[l], vyt[l], DZ ? vzt[l] : 0.0,
                          adiabatic ? e[l] : 0.0, ax, ay, az, de);
  vxt[l] = ax;
  vyt[l] = ay;
  if (DZ) vzt[l] = az;
  if (adiabatic) e[l] = de;
}

template <int G, int DZ>
__global__ void k_substep3(dmesh m, Range r, double dt, double* e, const double* vx, const double* vy,
                           const double* vz, double gamma) {
  CELL_INDEX;
  const double* vzs = DZ ? vz : vy;
  e[l] = cell::substep3<DZ>(m, i, j, k, dt, e[l], cell::gl(vx, m, l, i), cell::gl(vy, m, l, i),
                            cell::gl(vzs, m, l, i), gamma);
}

template <int G, int DZ>
__global__ void k_visctensor(dmesh m, Range r, const double* rho, const double* energy, const double* vx,
                             const double* vy, const double* vz, double* txx, double* tyy, double* tzz,
                             double* txz, double* tyx, double* tzy, cell::ViscPar vp) {
  CELL_INDEX;
  const double* vzs = DZ ? vz : vy;
  const double* es = energy ? energy : rho;
  cell::Taus t = cell::visctensor<G, DZ>(m, i, j, k, cell::gl(rho, m, l, i), cell::gl(es, m, l, i),
                                         cell::gl(vx, m, l, i), cell::gl(vy, m, l, i), cell::gl(vzs, m, l, i), vp);
  txx[l] = t.xx;
  tyy[l] = t.yy;
  tyx[l] = t.yx;
  if (DZ) {
    tzz[l] = t.zz;
    txz[l] = t.xz;
    tzy[l] = t.zy;
  }
}

template <int G, int DZ>
__global__ void k_addviscosity(dmesh m, Range r, double dt, const double* rho, double* vx, double* vy,
                               double* vz, const double* txx, const double* tyy, const double* tzz,
                               const double* txz, const double* tyx, const double* tzy) {
  CELL_INDEX;
  double ax, ay, az;
  const double* tzzs = DZ ? tzz : txx;
  const double* txzs = DZ ? txz : txx;
  const double* tzys = DZ ? tzy : txx;
  cell::addviscosity<G, DZ>(m, i, j, k, dt, cell::gl(rho, m, l, i), cell::gl(txx, m, l, i), cell::gl(tyy, m, l, i),
                            cell::gl(tzzs, m, l, i), cell::gl(txzs, m, l, i), cell::gl(tyx, m, l, i),
                            cell::gl(tzys, m, l, i), vx[l], vy[l], DZ ? vz[l] : 0.0, ax, ay, az);
  vx[l] = ax;
  vy[l] = ay;
  if (DZ) vz[l] = az;
}

template <int G, int DZ>
__global__ void __launch_bounds__(BX) k_potential(dmesh m, Range r, double* pot, cell::PlanetSet ps, int cpt) {
  // cpt (1 or 4) cells per thread, BX apart (a workgroup covers cpt * BX
  // cells of one ring): what depends on the ring alone is evaluated once
  const int i0 = (int)blockIdx.x * BX * cpt + (int)threadIdx.x;
  const int j = r.jb + (int)blockIdx.y, k = r.kb + (int)blockIdx.z;
  const int l0 = i0 + j * m.pitch + k * m.stride;
  if (cpt == 4) cell::potential_cells<G, 4>(m, i0, BX, j, k, ps, pot, l0);
  else cell::potential_cells<G, 1>(m, i0, BX, j, k, ps, pot, l0);
}
// the same, its first workgroup also taking the queued fused CFL's MIN pass
// (which would otherwise be a launch of its own; the potential needs no dt)
template <int G, int DZ>
__global__ void __launch_bounds__(BX) k_potential_fin(dmesh m, Range r, double* pot, cell::PlanetSet ps,
                                                      f3d_hip_ctx::CflFin fin, double* host_out,
                                                      unsigned long long seq, int cpt) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) {
    __shared__ double sh[BX / 64];
    cfl_min_publish(fin.ringmin, fin.n, fin.result, fin.slot, host_out, seq, sh);
  }
  const int i0 = (int)blockIdx.x * BX * cpt + (int)threadIdx.x;
  const int j = r.jb + (int)blockIdx.y, k = r.kb + (int)blockIdx.z;
  const int l0 = i0 + j * m.pitch + k * m.stride;
  if (cpt == 4) cell::potential_cells<G, 4>(m, i0, BX, j, k, ps, pot, l0);
  else cell::potential_cells<G, 1>(m, i0, BX, j, k, ps, pot, l0);
}

template <int G, int DZ>
__global__ void k_correct_vtheta(dmesh m, Range r, double* vt, double domega) {
  CELL_INDEX;
  if (G == SPH)
    vt[l] -= domega * m.ymed[j] * m.sinzmed[k];
  else
    vt[l] -= domega * m.ymed[j];
}

template <int G, int DZ>
__global__ void k_force(dmesh m, Range r, const double* dens, double* fxi, double* fyi, double* fzi,
                        double* fxhi, double* fyhi, double* fzhi, double x, double y, double z, double rs) {
  CELL_INDEX;
  double cellmass = vol(m, i, j, k) * dens[l];
  double dx, dy, dz;
  if (G == SPH) {
    dx = m.ymed[j] * m.cosxmed[i] * m.sinzmed[k] - x;
    dy = m.ymed[j] * m.sinxmed[i] * m.sinzmed[k] - y;
    dz = m.ymed[j] * m.coszmed[k] - z;
  } else {
    dx = m.ymed[j] * m.cosxmed[i] - x;
    dy = m.ymed[j] * m.sinxmed[i] - y;
    dz = m.zmed[k] - z;
  }
  double dist2 = dx * dx + dy * dy + dz * dz;
  dist2 += rs * rs;
  double distance = sqrt(dist2);
  double InvDist3 = 1.0 / (dist2 * distance);
  InvDist3 *= 1.0 * cellmass;
  fxi[l] = dx * InvDist3;
  fyi[l] = dy * InvDist3;
  fzi[l] = dz * InvDist3;
  fxhi[l] = 0.0;
  fyhi[l] = 0.0;
  fzhi[l] = 0.0;
}

template <int G, int DZ>
__global__ void k_momenta_x(dmesh m, Range r, const double* rho, const double* vx, double* mmx, double* mpx,
                            double omegaframe) {
  CELL_INDEX;
  double a, b;
  cell::momenta_x<G>(m, i, j, k, rho[l], vx[l], vx[lxp_of(l, i, m.nx)], omegaframe, a, b);
  mmx[l] = a;
  mpx[l] = b;
}

template <int G, int DZ>
__global__ void k_momenta_y(dmesh m, Range r, const double* rho, const double* vy, double* mmy, double* mpy) {
  CELL_INDEX;
  mmy[l] = vy[l] * rho[l];
  mpy[l] = vy[l + m.pitch] * rho[l];
}

template <int G, int DZ>
__global__ void k_momenta_z(dmesh m, Range r, const double* rho, const double* vz, double* mmz, double* mpz) {
  CELL_INDEX;
  double p = vz[l + m.stride] * rho[l];
  double q = vz[l] * rho[l];
  if (G == SPH) {
    p *= m.ymed[j];
    q *= m.ymed[j];
  }
  mpz[l] = p;
  mmz[l] = q;
}

template <int G, int DZ>
__global__ void k_divrho(dmesh m, Range r, const double* q, const double* rho, double* out) {
  CELL_INDEX;
  out[l] = q[l] / rho[l];
}

template <int G, int DZ>
__global__ void k_vly_a(dmesh m, Range r, const double* q, double* slope) {
  CELL_INDEX;
  double dqm = (q[l] - q[l - m.pitch]) / zone_size_y(m, j);
  double dqp = (q[l + m.pitch] - q[l]) / zone_size_y(m, j + 1);
  slope[l] = vanleer_slope(dqm, dqp);
}

template <int G, int DZ>
__global__ void k_vly_b(dmesh m, Range r, double dt, const double* q, double* qs, const double* slope,
                        const double* vy) {
  CELL_INDEX;
  const int lm = l - m.pitch;
  qs[l] = vanleer_face(vy[l], dt, q[lm], slope[lm], zone_size_y(m, j - 1), q[l], slope[l], zone_size_y(m, j));
}

template <int G, int DZ>
__global__ void k_vlz_a(dmesh m, Range r, const double* q, double* slope) {
  CELL_INDEX;
  double dqm = (q[l] - q[l - m.stride]) / zone_size_z<G>(m, j, k);
  double dqp = (q[l + m.stride] - q[l]) / zone_size_z<G>(m, j, k + 1);
  slope[l] = vanleer_slope(dqm, dqp);
}

template <int G, int DZ>
__global__ void k_vlz_b(dmesh m, Range r, double dt, const double* q, double* qs, const double* slope,
                        const double* vz) {
  CELL_INDEX;
  const int lm = l - m.stride;
  qs[l] = vanleer_face(vz[l], dt, q[lm], slope[lm], zone_size_z<G>(m, j, k - 1), q[l], slope[l],
                       zone_size_z<G>(m, j, k));
}

template <int G, int DZ>
__global__ void k_vlx_a(dmesh m, Range r, const double* q, double* slope) {
  CELL_INDEX;
  const int lxm = lxm_of(l, i, m.nx), lxp = lxp_of(l, i, m.nx);
  double dqm = (q[l] - q[lxm]) / zone_size_x<G>(m, i, j, k);
  double dqp = (q[lxp] - q[l]) / zone_size_x<G>(m, ixp_of(i, m.nx), j, k);
  slope[l] = vanleer_slope(dqm, dqp);
}

template <int G, int DZ>
__global__ void k_vlx_b(dmesh m, Range r, double dt, const double* q, double* qs, const double* slope,
                        const double* vx) {
  CELL_INDEX;
  const int lxm = lxm_of(l, i, m.nx);
  qs[l] = vanleer_face(vx[l], dt, q[lxm], slope[lxm], zone_size_x<G>(m, ixm_of(i, m.nx), j, k), q[l], slope[l],
                       zone_size_x<G>(m, i, j, k));
}

template <int G, int DZ>
__global__ void k_ppa_a(dmesh m, Range r, const double* q, double* slope) {
  CELL_INDEX;
  slope[l] = cell::ppa_slope(q[lxm_of(l, i, m.nx)], q[l], q[lxp_of(l, i, m.nx)]);
}

template <int G, int DZ>
__global__ void k_ppa_b(dmesh m, Range r, const double* q, const double* slope, double* qL, double* qR) {
  CELL_INDEX;
  const int lxp = lxp_of(l, i, m.nx);
  double temp = cell::ppa_face(q[l], q[lxp], slope[l], slope[lxp]);
  qR[l] = temp;
  qL[lxp] = temp;
}

template <int G, int DZ>
__global__ void k_ppa_c(dmesh m, Range r, const double* q, double* qL, double* qR) {
  CELL_INDEX;
  double a = qL[l], b = qR[l];
  cell::ppa_limit(q[l], a, b);
  qL[l] = a;
  qR[l] = b;
}

template <int G, int DZ>
__global__ void k_ppa_d(dmesh m, Range r, double dt, const double* q, double* qs, const double* qL,
                        const double* qR, const double* vx) {
  CELL_INDEX;
  const int lxm = lxm_of(l, i, m.nx);
  qs[l] = cell::ppa_flux_value(vx[l], dt, zone_size_x<G>(m, i, j, k), q[lxm], qL[lxm], qR[lxm], q[l], qL[l], qR[l]);
}

template <int G, int DZ>
__global__ void k_update_x(dmesh m, Range r, double dt, double* qb, const double* qs, const double* vx,
                           const double* rho_s) {
  CELL_INDEX;
  const int lxp = lxp_of(l, i, m.nx);
  qb[l] += cell::flux_x(m, i, j, k, dt, vx[l] * qs[l] * rho_s[l], vx[lxp] * qs[lxp] * rho_s[lxp]);
}

template <int G, int DZ>
__global__ void k_update_density_x(dmesh m, Range r, double dt, double* qb, const double* vx, const double* rho_s) {
  CELL_INDEX;
  const int lxp = lxp_of(l, i, m.nx);
  qb[l] += cell::flux_x(m, i, j, k, dt, vx[l] * rho_s[l], vx[lxp] * rho_s[lxp]);
}

template <int G, int DZ>
__global__ void k_update_y(dmesh m, Range r, double dt, double* qb, const double* qs, const double* vy,
                           const double* rho_s) {
  CELL_INDEX;
  const int lyp = l + m.pitch;
  qb[l] += cell::flux_y(m, i, j, k, dt, vy[l] * qs[l] * rho_s[l], vy[lyp] * qs[lyp] * rho_s[lyp]);
}

template <int G, int DZ>
__global__ void k_update_density_y(dmesh m, Range r, double dt, double* qb, const double* vy, const double* rho_s) {
  CELL_INDEX;
  const int lyp = l + m.pitch;
  qb[l] += cell::flux_y(m, i, j, k, dt, vy[l] * rho_s[l], vy[lyp] * rho_s[lyp]);
}

template <int G, int DZ>
__global__ void k_update_z(dmesh m, Range r, double dt, double* qb, const double* qs, const double* vz,
                           const double* rho_s) {
  CELL_INDEX;
  const int lzp = l + m.stride;
  qb[l] += cell::flux_z(m, i, j, k, dt, vz[l] * qs[l] * rho_s[l], vz[lzp] * qs[lzp] * rho_s[lzp]);
}

template <int G, int DZ>
__global__ void k_update_density_z(dmesh m, Range r, double dt, double* qb, const double* vz, const double* rho_s) {
  CELL_INDEX;
  const int lzp = l + m.stride;
  qb[l] += cell::flux_z(m, i, j, k, dt, vz[l] * rho_s[l], vz[lzp] * rho_s[lzp]);
}

template <int G, int DZ>
__global__ void k_newvel_x(dmesh m, Range r, const double* rho, const double* mmx, const double* mpx, double* vx,
                           double omegaframe) {
  CELL_INDEX;
  const int lxm = lxm_of(l, i, m.nx);
  vx[l] = cell::newvel_x<G>(m, i, j, k, rho[l], rho[lxm], mmx[l], mpx[lxm], omegaframe);
}

template <int G, int DZ>
__global__ void k_newvel_y(dmesh m, Range r, const double* rho, const double* mmy, const double* mpy, double* vy) {
  CELL_INDEX;
  const int lym = l - m.pitch;
  vy[l] = cell::newvel_y(m, i, j, k, rho[l], rho[lym], mmy[l], mpy[lym]);
}

template <int G, int DZ>
__global__ void k_newvel_z(dmesh m, Range r, const double* rho, const double* mmz, const double* mpz, double* vz) {
  CELL_INDEX;
  const int lzm = l - m.stride;
  // "mmz*Vol(i,j,k)" associates as ((mmz*Syk)*Sxi)/InvVj (src/define.h:229)
  double v = (mul_vol(m, mmz[l], i, j, k) + mul_vol(m, mpz[lzm], i, j, k - 1)) /
      (mul_vol(m, rho[l], i, j, k) + mul_vol(m, rho[lzm], i, j, k - 1));
  if (G == SPH) v /= m.ymed[j];
  vz[l] = v;
}

template <int G, int DZ>
__global__ void k_vweight(dmesh m, Range r, const double* v, double* q) {
  CELL_INDEX;
  q[l] = v[l] * (m.xmin[i + 1] - m.xmin[i]);
}

template <int G, int DZ>
__global__ void k_residual(dmesh m, Range r, double dt, double* vx_temp, double* vx, const double* vxmed,
                           int* nshift) {
  CELL_INDEX;
  const int l2 = j + k * m.sy;
  double zs = zone_size_x<G>(m, i, j, k);
  double ntilde = vxmed[l2] * dt / zs;
  double nround = floor(ntilde + 0.5);
  if (i == 0) nshift[l2] = (int)nround;
  vx[l] = vx_temp[l] - vxmed[l2];
  vx_temp[l] = (ntilde - nround) * zs / dt;
}

template <int G, int DZ>
__global__ void k_shift_gather(dmesh m, Range r, const double* f, double* aux, const int* nshift) {
  CELL_INDEX;
  int it = i - nshift[j + k * m.sy];
  it %= m.nx;  // the reference wraps with while loops; the result is the same residue
  if (it < 0) it += m.nx;
  aux[l] = f[l - i + it];
}

template <int G, int DZ>
__global__ void k_stockholm(dmesh m, Range r, double dt, double* rho, double* vx, double* vy, double* vz,
                            const double* rho0, const double* vx0, const double* vy0, const double* vz0,
                            cell::DampPar dp) {
  CELL_INDEX;
  cell::stockholm<G, DZ>(m, j, k, l, dt, rho, vx, vy, vz, rho0, vx0, vy0, vz0, dp);
}

// cfl per cell (src/cfl.c:174-253) into DensStar
template <int G, int DZ>
__global__ void k_cfl(dmesh m, Range r, const double* cs, const double* rho, const double* vx, const double* vy,
                      const double* vz, const double* vxmed, double* dtime, cell::CflPar cp) {
  CELL_INDEX;
  dtime[l] = cell::cfl_cell<G, DZ>(m, i, j, k, l, cs, rho, vx, vy, vz, vxmed[j + k * m.sy], cp);
}

// ---- reductions (src/reduction_generic.c:5-23, sequential i order per ring)
__global__ void k_ring_sum(dmesh m, const double* f, double* red2d, int jb, int je, int kb, int ke, int is_min) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nj = je - jb;
  if (t >= nj * (ke - kb)) return;
  int j = jb + t % nj, k = kb + t / nj;
  const double* row = f + (size_t)j * m.pitch + (size_t)k * m.stride;
  double acc = is_min ? 1e20 : 0.0;
  if (is_min) {
    for (int i = 0; i < m.nx; i++) {
      double b = row[i];
      acc = (acc < b) ? acc : b;
    }
  } else {
    for (int i = 0; i < m.nx; i++) acc = acc + row[i];
  }
  red2d[j + k * m.sy] = acc;
}

// final (j,k) pass of reduction_full (src/reduction_full_generic.c:11-16), one thread
__global__ void k_full_finish(dmesh m, const double* red2d, double* out, int jb, int je, int kb, int ke, int is_min) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double result = is_min ? 1e20 : 0.0;
  for (int k = kb; k < ke; k++)
    for (int j = jb; j < je; j++) {
      double a = red2d[j + k * m.sy];
      result = is_min ? ((a < result) ? a : result) : a + result;
    }
  *out = result;
}

// cfl_b shear limit (src/cfl_b.c:14-37) folded into the MIN of the per-ring minima
__global__ void k_cfl_finish(dmesh m, const double* red2d, const double* vxmed, double* out, int fargo, double cfl) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int kb = m.dimz ? m.nghz : 0, ke = m.dimz ? m.nz + m.nghz : 1;
  double result = 1e20;
  for (int k = kb; k < ke; k++)
    for (int j = m.nghy; j < m.ny + m.nghy; j++) {
      double a = red2d[j + k * m.sy];
      result = (a < result) ? a : result;
    }
  if (fargo) {
    for (int k = kb; k < ke; k++)
      for (int j = m.nghy; j < m.ny + m.nghy; j++) {
        int l2 = j + k * m.sy;
        double shearlimit = cfl * m.mindx / fabs(vxmed[l2] / m.ymed[j] - vxmed[l2 + 1] / m.ymed[j + 1]);
        if (shearlimit < result) result = shearlimit;
      }
  }
  *out = result;
}

__global__ void k_vmed_div(dmesh m, const double* red2d, double* vxmed, int n2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n2) return;
  vxmed[t] = red2d[t] / (m.XMAX - m.XMIN);
}

// ---- boundaries (std/boundary_template.c via scripts/boundparser.py:237-363)
// Y ghost fill of side side0 (0: ymin) and, when gridDim.y = 2 nghy, also of
// side 1 in the same launch.  With ny > nghy the two sides write disjoint rows
// and neither reads a row the other writes, so the reference's order (ymin
// first, src/boundary.c:2-23) does not change any value.
__global__ void k_bound_y(dmesh m, int side0, double* dens, double* vx, double* vy, double* vz, double* e, cell::BcPar bp0,
                          cell::BcPar bp1, const double* dexp0, const double* dexp1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int side = side0 + ((int)blockIdx.y >= m.nghy);
  const int mm = (int)blockIdx.y - (side - side0) * m.nghy, k = blockIdx.z;
  cell::bound_y_cell(m, side, i, mm, k, dens, vx, vy, vz, e, side ? bp1 : bp0, side ? dexp1 : dexp0);
}

// Z ghost fill, the same arrangement as k_bound_y (both sides when nz > nghz).
__global__ void k_bound_z(dmesh m, int side0, double* dens, double* vx, double* vy, double* vz, double* e, cell::BcPar bp0,
                          cell::BcPar bp1, const double* zdens0, const double* zdens1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int side = side0 + ((int)blockIdx.z >= m.nghz);
  const int mm = (int)blockIdx.z - (side - side0) * m.nghz, j = blockIdx.y;
  cell::bound_z_cell(m, side, i, mm, j, dens, vx, vy, vz, e, side ? bp1 : bp0, side ? zdens1 : zdens0);
}

// CondInit on the device (f3d_CondInit): the jump-ahead of glibc's drand48
// and the per-cell arithmetic of the host's f3d_ic_cell, operation for
// operation (FP64, no contraction), so the host and the device give the same
// bits -- except the log-normal set-up, whose exp/log/cos are the device's
// own (within a few ulp of glibc's).
struct IcJump {
  unsigned long long x0;                  // state before the first draw
  unsigned long long A[48], C[48];        // X -> A X + C: 2^b * ndraw draws ahead
};
struct IcPar {
  int mode, ndraw, adiabatic;
  double noise, h, gamma;
};
template <typename T>
__global__ void k_condinit(dmesh m, IcPar d, IcJump jp, const double* __restrict__ R, const double* __restrict__ V,
                           const double* __restrict__ CS, const double* __restrict__ E, T* __restrict__ dens,
                           T* __restrict__ en, T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                           const double* __restrict__ vb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const size_t l = (size_t)i + (size_t)j * m.pitch + (size_t)k * m.stride, l2 = (size_t)j + (size_t)k * m.sy;
  constexpr unsigned long long MASK = (1ULL << 48) - 1;
  double u[3] = {0.0, 0.0, 0.0};
  if (d.ndraw) {
    // the cell's place in the set-ups' loop order (x fastest, then y, then z)
    unsigned long long n = (unsigned long long)l2 * (unsigned long long)m.nx + (unsigned long long)i;
    unsigned long long X = jp.x0;
    for (int b = 0; n; b++, n >>= 1)
      if (n & 1) X = (jp.A[b] * X + jp.C[b]) & MASK;
    for (int q = 0; q < 3; q++) {
      if (q >= d.ndraw) break;
      X = (0x5DEECE66DULL * X + 0xBULL) & MASK;
      u[q] = (double)X * 0x1p-48;  // erand48: the 48 bits as the fraction of 1 + x, minus 1
    }
  }
  const double Rv = R[l2], Vv = V[l2];
  double r = Rv, x = Vv, y = 0.0, e;
  if (d.mode == F3D_IC_PLAIN) {
    e = E[l2];
  } else {
    if (d.mode == F3D_IC_KEPLER) {
      r = Rv * (1.0 + d.noise * (u[0] - .5));
    } else if (d.mode == F3D_IC_FARGO) {
      r = Rv * (1.0 + d.noise * (u[0] - .5));
      x = Vv * (1. + d.h * d.noise * (u[1] - .5));
      y = CS[l2] * d.noise * (u[2] - .5);
    } else if (d.noise != 0.0) {
      double u1 = u[0];
      const double u2 = u[1];
      if (u1 < 1e-300) u1 = 1e-300;
      r *= exp(d.noise * (sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)));
    }
    e = d.adiabatic ? E[l2] * r / (d.gamma - 1.0) : CS[l2];
  }
  dens[l] = (T)r;
  en[l] = (T)e;
  vx[l] = vb ? (T)(x - vb[l2]) : (T)x;
  vy[l] = (T)y;
  if (vz) vz[l] = (T)0.0;
}

__global__ void k_init_stockholm(dmesh m, const double* rho, const double* vx, const double* vy, const double* vz,
                                 const double* e, double* rho0, double* vx0, double* vy0, double* vz0, double* e0,
                                 int adiabatic) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nk = m.dimz ? m.sz : 1;
  if (t >= m.sy * nk) return;
  int j = t % m.sy, k = t / m.sy;
  int l = j * m.pitch + k * m.stride, l2 = j + k * m.sy;
  if (adiabatic) e0[l2] = e[l];
  vx0[l2] = vx[l];
  vy0[l2] = vy[l];
  if (m.dimz) vz0[l2] = vz[l];
  rho0[l2] = rho[l];
}

// mon_dens, mon_momx/y/z and mon_torque (src/mon_dens.c:10-68, src/mon_momx.c,
// src/mon_momy.c, src/mon_momz.c, src/mon_torque.c): the monitored quantity of
// each active cell into an FP64 array, with the reference's association
// (Vol(i,j,k) expands to Syk(k)*Sxi(i)/InvVj(j) inside the product).  T is the
// field type: double, or float for the precision form, whose values are
// widened before any arithmetic so that its monitors are FP64 sums; vb is the
// precision form's ring profile of the azimuthal velocity (its Vx holds the
// deviation from it), null for the FP64 form.
enum { MON_DENS, MON_MOMX, MON_MOMY, MON_MOMZ, MON_TORQ };

template <int G, int DZ, int Q, typename T>
__global__ void k_mon(dmesh m, Range r, const T* __restrict__ dens, const T* __restrict__ v,
                      double* __restrict__ out, double omegaframe, double xp, double yp, double zp, double rsm2,
                      const double* __restrict__ vb) {
  CELL_INDEX;
  const double rho = (double)dens[l];
  double q;
  if (Q == MON_DENS) {
    q = mul_vol(m, rho, i, j, k);
  } else if (Q == MON_MOMX) {
    double v0 = (double)v[l], v1 = (double)v[lxp_of(l, i, m.nx)];
    if (vb) {
      // the full velocities, as a host read-back of the field forms them
      v0 += vb[j + k * m.sy];
      v1 += vb[j + k * m.sy];
    }
    const double vm = .5 * (v0 + v1);
    if (G == SPH) {
      const double rcyl = m.ymed[j] * m.sinzmed[k];
      q = mul_vol(m, ((vm + rcyl * omegaframe) * rcyl) * rho, i, j, k);
    } else {
      q = mul_vol(m, ((vm + m.ymed[j] * omegaframe) * m.ymed[j]) * rho, i, j, k);
    }
  } else if (Q == MON_MOMY) {
    q = mul_vol(m, (rho * .5) * ((double)v[l] + (double)v[l + m.pitch]), i, j, k);
  } else if (Q == MON_MOMZ) {
    q = mul_vol(m, (rho * .5) * ((double)v[l] + (double)v[l + m.stride]), i, j, k);
    if (G == SPH) q *= m.ymed[j];
  } else {
    const double cellmass = vol(m, i, j, k) * rho;
    double dx, dy, dz = 0.0;
    if (G == SPH) {
      dx = m.ymed[j] * m.cosxmed[i] * m.sinzmed[k] - xp;
      dy = m.ymed[j] * m.sinxmed[i] * m.sinzmed[k] - yp;
      if (DZ) dz = m.ymed[j] * m.coszmed[k] - zp;
    } else {
      dx = m.ymed[j] * m.cosxmed[i] - xp;
      dy = m.ymed[j] * m.sinxmed[i] - yp;
      if (DZ) dz = m.zmed[k] - zp;
    }
    double dist2 = dx * dx + dy * dy + dz * dz;
    dist2 += rsm2;
    const double distance = sqrt(dist2);
    double InvDist3 = 1.0 / (dist2 * distance);
    InvDist3 *= 1.0 * cellmass;
    const double fxi = dx * InvDist3;
    const double fyi = dy * InvDist3;
    q = xp * fyi - yp * fxi;
  }
  out[l] = q;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int f3d_ComputePressureFieldIso(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Density, "Density"); NEED(ctx->f.Energy, "Energy"); NEED(ctx->f.Pressure, "Pressure");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_pres_iso, r, ctx->f.Density, ctx->f.Energy, ctx->f.Pressure);
  return 0;
}

int f3d_ComputePressureFieldAd(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Energy, "Energy"); NEED(ctx->f.Pressure, "Pressure");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_pres_ad, r, ctx->f.Energy, ctx->f.Pressure, ctx->ph.gamma);
  return 0;
}

int f3d_ComputeTotalDensity(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Density, "Density"); NEED(ctx->f.Total_Density, "Total_Density");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_totaldens, r, ctx->f.Density, ctx->f.Total_Density);
  return 0;
}

int f3d_Reset_field(f3d_hip_ctx* ctx, double* F) {
  F3D_REQUIRE(ctx, F != nullptr, "Reset_field: null field");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_fill, r, F, 0.0);
  return 0;
}

int f3d_SubStep1_x(f3d_hip_ctx* ctx, double dt) {
  NEED(ctx->f.Pressure, "Pressure"); NEED(ctx->f.Density, "Density"); NEED(ctx->f.Pot, "Pot");
  NEED(ctx->f.Vx, "Vx"); NEED(ctx->f.Vx_temp, "Vx_temp");
  Range r = rng(ctx, 1, ctx->m.sy - 1, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep1_x, r, dt, ctx->f.Pressure, ctx->f.Density, ctx->f.Pot, ctx->f.Vx, ctx->f.Vx_temp,
           (int)(ctx->ph.fluidtype == 2), ctx->ph.potential);
  return 0;
}

int f3d_SubStep1_y(f3d_hip_ctx* ctx, double dt) {
  NEED(ctx->f.Pressure, "Pressure"); NEED(ctx->f.Density, "Density"); NEED(ctx->f.Pot, "Pot");
  NEED(ctx->f.Vx, "Vx"); NEED(ctx->f.Vy, "Vy"); NEED(ctx->f.Vy_temp, "Vy_temp");
  if (ctx->m.dimz) NEED(ctx->f.Vz, "Vz");
  Range r = rng(ctx, 1, ctx->m.sy - 1, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep1_y, r, dt, ctx->f.Pressure, ctx->f.Density, ctx->f.Pot, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz,
           ctx->f.Vy_temp, (int)(ctx->ph.fluidtype == 2), ctx->ph.potential, ctx->ph.OMEGAFRAME);
  return 0;
}

int f3d_SubStep1_z(f3d_hip_ctx* ctx, double dt) {
  if (!ctx->m.dimz) return 0;
  NEED(ctx->f.Vz, "Vz"); NEED(ctx->f.Vz_temp, "Vz_temp");
  Range r = rng(ctx, 1, ctx->m.sy - 1, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep1_z, r, dt, ctx->f.Pressure, ctx->f.Density, ctx->f.Pot, ctx->f.Vx, ctx->f.Vz,
           ctx->f.Vz_temp, (int)(ctx->ph.fluidtype == 2), ctx->ph.potential, ctx->ph.OMEGAFRAME);
  return 0;
}

int f3d_SubStep2_a(f3d_hip_ctx* ctx, double dt) {
  (void)dt;
  NEED(ctx->f.Mpx, "Mpx"); NEED(ctx->f.Mpy, "Mpy");
  if (ctx->m.dimz) NEED(ctx->f.Mpz, "Mpz");
  Range r = rng(ctx, 0, ctx->m.sy - 1, 0, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep2_a, r, ctx->f.Density, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.Mpx, ctx->f.Mpy, ctx->f.Mpz);
  return 0;
}

int f3d_SubStep2_b(f3d_hip_ctx* ctx, double dt) {
  Range r = rng(ctx, 1, ctx->m.sy - 1, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep2_b, r, dt, ctx->f.Density, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.Vx_temp,
           ctx->f.Vy_temp, ctx->f.Vz_temp, ctx->f.Mpx, ctx->f.Mpy, ctx->f.Mpz, ctx->f.Energy,
           (int)(ctx->ph.eos == 2));
  return 0;
}

int f3d_SubStep3(f3d_hip_ctx* ctx, double dt) {
  Range r = rng(ctx, 0, ctx->m.sy - 1, 0, ctx->m.sz - 1);
  DISPATCH(ctx, k_substep3, r, dt, ctx->f.Energy, ctx->f.Vx_temp, ctx->f.Vy_temp, ctx->f.Vz_temp, ctx->ph.gamma);
  return 0;
}

static cell::ViscPar visc_par(const f3d_hip_ctx* ctx) {
  cell::ViscPar v;
  v.model = ctx->ph.visc;
  v.nu = ctx->ph.nu;
  v.alpha = ctx->ph.alpha;
  v.gamma = ctx->ph.gamma;
  v.iso = (ctx->ph.eos != 2);
  return v;
}

static int visctensor_any(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Mmx, "Mmx"); NEED(ctx->f.Mmy, "Mmy"); NEED(ctx->f.Mpy, "Mpy");
  Range r = rng(ctx, 1, ctx->m.sy - 1, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_visctensor, r, ctx->f.Density, ctx->f.Energy, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.Mmx,
           ctx->f.Mmy, ctx->f.Mmz, ctx->f.Mpx, ctx->f.Mpy, ctx->f.Mpz, visc_par(ctx));
  return 0;
}

static int addviscosity_any(f3d_hip_ctx* ctx, double dt) {
  Range r = rng(ctx, 1, ctx->m.sy - 2, 1, ctx->m.sz - 2);
  DISPATCH(ctx, k_addviscosity, r, dt, ctx->f.Density, ctx->f.Vx_temp, ctx->f.Vy_temp, ctx->f.Vz_temp, ctx->f.Mmx,
           ctx->f.Mmy, ctx->f.Mmz, ctx->f.Mpx, ctx->f.Mpy, ctx->f.Mpz);
  return 0;
}

int f3d_visctensor_cyl(f3d_hip_ctx* ctx) {
  F3D_REQUIRE(ctx, ctx->m.geom == CYL, "visctensor_cyl on a spherical mesh");
  return visctensor_any(ctx);
}
int f3d_addviscosity_cyl(f3d_hip_ctx* ctx, double dt) {
  F3D_REQUIRE(ctx, ctx->m.geom == CYL, "addviscosity_cyl on a spherical mesh");
  return addviscosity_any(ctx, dt);
}
int f3d_visctensor_sph(f3d_hip_ctx* ctx) {
  F3D_REQUIRE(ctx, ctx->m.geom == SPH, "visctensor_sph on a cylindrical mesh");
  return visctensor_any(ctx);
}
int f3d_addviscosity_sph(f3d_hip_ctx* ctx, double dt) {
  F3D_REQUIRE(ctx, ctx->m.geom == SPH, "addviscosity_sph on a cylindrical mesh");
  return addviscosity_any(ctx, dt);
}

static cell::PlanetSet planet_set(const f3d_hip_ctx* ctx) {
  cell::PlanetSet ps;
  ps.n = ctx->ph.nplanets < cell::kMaxPlanets ? ctx->ph.nplanets : cell::kMaxPlanets;
  ps.indirect = ctx->ph.indirectterm;
  for (int n = 0; n < ps.n; n++) {
    ps.x[n] = ctx->ph.px[n];
    ps.y[n] = ctx->ph.py[n];
    ps.z[n] = ctx->ph.pz[n];
    ps.mp[n] = ctx->ph.pmass[n];
    ps.smooth2[n] = ctx->ph.psmooth2[n];
  }
  return ps;
}

int f3d_Potential(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Pot, "Pot");
  F3D_REQUIRE(ctx, ctx->ph.nplanets <= cell::kMaxPlanets, "Potential: too many planets");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  f3d_stage_begin(ctx, F3D_ST_POTENTIAL);
  const int cpt = pot_cells(ctx);
  if (ctx->fin_pending && !empty(r)) {
    // the fused CFL queued before this launch left its MIN pass to it
    ctx->fin_pending = 0;
    DISPATCH_X(ctx, k_potential_fin, r, cpt, ctx->f.Pot, planet_set(ctx), ctx->fin, ctx->dhres, ctx->cfl_seq, cpt);
  } else {
    DISPATCH_X(ctx, k_potential, r, cpt, ctx->f.Pot, planet_set(ctx), cpt);
  }
  f3d_stage_end(ctx, F3D_ST_POTENTIAL);
  return 0;
}

int f3d_CorrectVtheta(f3d_hip_ctx* ctx, double domega) {
  NEED(ctx->f.Vx, "Vx");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_correct_vtheta, r, ctx->f.Vx, domega);
  return 0;
}

int f3d_copy_velocities(f3d_hip_ctx* ctx, int option) {
  size_t b = ctx->n3 * sizeof(double);
  const f3d_fields& f = ctx->f;
  if (option == 1) {
    F3D_CHECK(ctx, hipMemcpyAsync(f.Vx, f.Vx_temp, b, hipMemcpyDeviceToDevice, ctx->stream));
    F3D_CHECK(ctx, hipMemcpyAsync(f.Vy, f.Vy_temp, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->m.dimz) F3D_CHECK(ctx, hipMemcpyAsync(f.Vz, f.Vz_temp, b, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    F3D_CHECK(ctx, hipMemcpyAsync(f.Vx_temp, f.Vx, b, hipMemcpyDeviceToDevice, ctx->stream));
    F3D_CHECK(ctx, hipMemcpyAsync(f.Vy_temp, f.Vy, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->m.dimz) F3D_CHECK(ctx, hipMemcpyAsync(f.Vz_temp, f.Vz, b, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

static cell::BcPar bc_par(const f3d_hip_ctx* ctx, int side) {
  cell::BcPar b;
  b.p3d = (ctx->ph.bc == 2);
  b.energy = ctx->ph.energy_bc && ctx->f.Energy != nullptr;
  b.of = ctx->ph.OMEGAFRAME;
  for (int q = 0; q < 3; q++) {
    b.dens[q] = side ? ctx->ph.ymax_dens[q] : ctx->ph.ymin_dens[q];
    b.vsq[q] = side ? ctx->ph.ymax_vsq[q] : ctx->ph.ymin_vsq[q];
    b.vfac[q] = side ? ctx->ph.ymax_vfac[q] : ctx->ph.ymin_vfac[q];
    b.zvfac[q] = side ? ctx->ph.zmax_vfac[q] : ctx->ph.zmin_vfac[q];
  }
  return b;
}

int f3d_boundaries(f3d_hip_ctx* ctx) {
  return f3d_boundaries_on(ctx, ctx->f.Density, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz);
}

// The ghost fill on explicit buffers (the fused path fills V_temp directly).
int f3d_boundaries_on(f3d_hip_ctx* ctx, double* dens, double* vx, double* vy, double* vz) {
  const dmesh& m = ctx->m;
  f3d_fields f = ctx->f;
  f.Density = dens;
  f.Vx = vx;
  f.Vy = vy;
  f.Vz = vz;
  dim3 blk(BX);
  // zmin, zmax, then ymin, ymax (src/boundary.c:2-23); the Y fill covers the
  // Z ghost planes, so it runs after the Z fill
  if (!ctx->ph.periodicz && m.dimz) {
    F3D_REQUIRE(ctx, ctx->ph.bc == 2, "3-D boundaries need the spherical-disk family");
    const cell::BcPar b0 = bc_par(ctx, 0), b1 = bc_par(ctx, 1);
    if (m.nz > m.nghz) {
      dim3 g((m.nx + BX - 1) / BX, m.sy, 2 * m.nghz);
      hipLaunchKernelGGL(k_bound_z, g, blk, 0, ctx->stream, m, 0, f.Density, f.Vx, f.Vy, f.Vz, f.Energy, b0, b1,
                         ctx->d_zmin_dens, ctx->d_zmax_dens);
      F3D_LAUNCH_CHECK(ctx);
    } else {
      for (int side = 0; side < 2; side++) {
        dim3 g((m.nx + BX - 1) / BX, m.sy, m.nghz);
        const cell::BcPar b = side ? b1 : b0;
        const double* t = side ? ctx->d_zmax_dens : ctx->d_zmin_dens;
        hipLaunchKernelGGL(k_bound_z, g, blk, 0, ctx->stream, m, side, f.Density, f.Vx, f.Vy, f.Vz, f.Energy, b, b, t, t);
        F3D_LAUNCH_CHECK(ctx);
      }
    }
  }
  if (!ctx->ph.periodicy) {
    const cell::BcPar b0 = bc_par(ctx, 0), b1 = bc_par(ctx, 1);
    if (m.ny > m.nghy) {
      dim3 g((m.nx + BX - 1) / BX, 2 * m.nghy, m.dimz ? m.sz : 1);
      hipLaunchKernelGGL(k_bound_y, g, blk, 0, ctx->stream, m, 0, f.Density, f.Vx, f.Vy, f.Vz, f.Energy, b0, b1,
                         ctx->d_ymin_dexp, ctx->d_ymax_dexp);
      F3D_LAUNCH_CHECK(ctx);
    } else {
      for (int side = 0; side < 2; side++) {
        dim3 g((m.nx + BX - 1) / BX, m.nghy, m.dimz ? m.sz : 1);
        const cell::BcPar b = side ? b1 : b0;
        const double* t = side ? ctx->d_ymax_dexp : ctx->d_ymin_dexp;
        hipLaunchKernelGGL(k_bound_y, g, blk, 0, ctx->stream, m, side, f.Density, f.Vx, f.Vy, f.Vz, f.Energy, b, b, t, t);
        F3D_LAUNCH_CHECK(ctx);
      }
    }
  }
  return 0;
}

int f3d_momenta_x(f3d_hip_ctx* ctx) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_momenta_x, r, ctx->f.Density, ctx->f.Vx_temp, ctx->f.Mmx, ctx->f.Mpx, ctx->ph.OMEGAFRAME);
  return 0;
}

int f3d_momenta_y(f3d_hip_ctx* ctx) {
  Range r = rng(ctx, 0, ctx->m.sy - 1, 0, ctx->m.sz);
  DISPATCH(ctx, k_momenta_y, r, ctx->f.Density, ctx->f.Vy_temp, ctx->f.Mmy, ctx->f.Mpy);
  return 0;
}

int f3d_momenta_z(f3d_hip_ctx* ctx) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz - 1);
  DISPATCH(ctx, k_momenta_z, r, ctx->f.Density, ctx->f.Vz_temp, ctx->f.Mmz, ctx->f.Mpz);
  return 0;
}

int f3d_DivideByRho(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_divrho, r, Q, ctx->f.Density, ctx->f.DivRho);
  return 0;
}

int f3d_VanLeerY_a(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 1, ctx->m.sy - 1, 0, ctx->m.sz);
  DISPATCH(ctx, k_vly_a, r, Q, ctx->f.Slope);
  return 0;
}

int f3d_VanLeerY_b(f3d_hip_ctx* ctx, double dt, const double* Q, double* Qs) {
  Range r = rng(ctx, 1, ctx->m.sy - 1, 0, ctx->m.sz);
  DISPATCH(ctx, k_vly_b, r, dt, Q, Qs, ctx->f.Slope, ctx->f.Vy_temp);
  return 0;
}

int f3d_VanLeerZ_a(f3d_hip_ctx* ctx, const double* Q) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 0, ctx->m.sy, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_vlz_a, r, Q, ctx->f.Slope);
  return 0;
}

int f3d_VanLeerZ_b(f3d_hip_ctx* ctx, double dt, const double* Q, double* Qs) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 0, ctx->m.sy, 1, ctx->m.sz - 1);
  DISPATCH(ctx, k_vlz_b, r, dt, Q, Qs, ctx->f.Slope, ctx->f.Vz_temp);
  return 0;
}

int f3d_VanLeerX_a(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_vlx_a, r, Q, ctx->f.Slope);
  return 0;
}

int f3d_VanLeerX_b(f3d_hip_ctx* ctx, double dt, const double* Q, double* Qs, const double* Vx_t) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_vlx_b, r, dt, Q, Qs, ctx->f.Slope, Vx_t);
  return 0;
}

int f3d_VanLeerX_PPA_a(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_ppa_a, r, Q, ctx->f.Slope);
  return 0;
}

int f3d_VanLeerX_PPA_b(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_ppa_b, r, Q, ctx->f.Slope, ctx->f.QL, ctx->f.QR);
  return 0;
}

int f3d_VanLeerX_PPA_c(f3d_hip_ctx* ctx, const double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_ppa_c, r, Q, ctx->f.QL, ctx->f.QR);
  return 0;
}

int f3d_VanLeerX_PPA_d(f3d_hip_ctx* ctx, double dt, const double* Q, double* Qs, const double* Vx_t) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_ppa_d, r, dt, Q, Qs, ctx->f.QL, ctx->f.QR, Vx_t);
  return 0;
}

int f3d_UpdateX(f3d_hip_ctx* ctx, double dt, double* Q, const double* Qs, const double* Vx_t) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_update_x, r, dt, Q, Qs, Vx_t, ctx->f.DensStar);
  return 0;
}

int f3d_UpdateY(f3d_hip_ctx* ctx, double dt, double* Q, const double* Qs) {
  Range r = rng(ctx, 0, ctx->m.sy - 1, 0, ctx->m.sz);
  DISPATCH(ctx, k_update_y, r, dt, Q, Qs, ctx->f.Vy_temp, ctx->f.DensStar);
  return 0;
}

int f3d_UpdateZ(f3d_hip_ctx* ctx, double dt, double* Q, const double* Qs) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz - 1);
  DISPATCH(ctx, k_update_z, r, dt, Q, Qs, ctx->f.Vz_temp, ctx->f.DensStar);
  return 0;
}

int f3d_UpdateDensityX(f3d_hip_ctx* ctx, double dt, double* Q, const double* Vx_t) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_update_density_x, r, dt, Q, Vx_t, ctx->f.DensStar);
  return 0;
}

int f3d_UpdateDensityY(f3d_hip_ctx* ctx, double dt, double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy - 1, 0, ctx->m.sz);
  DISPATCH(ctx, k_update_density_y, r, dt, Q, ctx->f.Vy_temp, ctx->f.DensStar);
  return 0;
}

int f3d_UpdateDensityZ(f3d_hip_ctx* ctx, double dt, double* Q) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz - 1);
  DISPATCH(ctx, k_update_density_z, r, dt, Q, ctx->f.Vz_temp, ctx->f.DensStar);
  return 0;
}

int f3d_NewVelocity_x(f3d_hip_ctx* ctx) {
  Range r = rng(ctx, 1, ctx->m.sy, 1, ctx->m.sz);
  DISPATCH(ctx, k_newvel_x, r, ctx->f.Density, ctx->f.Mmx, ctx->f.Mpx, ctx->f.Vx, ctx->ph.OMEGAFRAME);
  return 0;
}

int f3d_NewVelocity_y(f3d_hip_ctx* ctx) {
  Range r = rng(ctx, 1, ctx->m.sy, 1, ctx->m.sz);
  DISPATCH(ctx, k_newvel_y, r, ctx->f.Density, ctx->f.Mmy, ctx->f.Mpy, ctx->f.Vy);
  return 0;
}

int f3d_NewVelocity_z(f3d_hip_ctx* ctx) {
  if (!ctx->m.dimz) return 0;
  Range r = rng(ctx, 1, ctx->m.sy, 1, ctx->m.sz);
  DISPATCH(ctx, k_newvel_z, r, ctx->f.Density, ctx->f.Mmz, ctx->f.Mpz, ctx->f.Vz);
  return 0;
}

int f3d_ComputeVweight(f3d_hip_ctx* ctx, const double* V, double* Q) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_vweight, r, V, Q);
  return 0;
}

static int ring_reduce(f3d_hip_ctx* ctx, const double* F, double* red2d, int jb, int je, int kb, int ke, int is_min) {
  int n = (je - jb) * (ke - kb);
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_ring_sum, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, ctx->m, F, red2d, jb, je, kb, ke,
                     is_min);
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

int f3d_ComputeVmed(f3d_hip_ctx* ctx, const double* V) {
  NEED(ctx->f.Qs, "Qs"); NEED(ctx->f.Reduction2D, "Reduction2D"); NEED(ctx->f.VxMed, "VxMed");
  if (f3d_ComputeVweight(ctx, V, ctx->f.Qs)) return -1;
  int ke = ctx->m.dimz ? ctx->m.sz : 1;
  if (ring_reduce(ctx, ctx->f.Qs, ctx->f.Reduction2D, 0, ctx->m.sy, 0, ke, 0)) return -1;
  int n2 = ctx->m.sy * ke;
  hipLaunchKernelGGL(k_vmed_div, dim3((n2 + 255) / 256), dim3(256), 0, ctx->stream, ctx->m, ctx->f.Reduction2D,
                     ctx->f.VxMed, n2);
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

int f3d_ComputeResidual(f3d_hip_ctx* ctx, double dt) {
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_residual, r, dt, ctx->f.Vx_temp, ctx->f.Vx, ctx->f.VxMed, ctx->f.Nshift);
  return 0;
}

int f3d_AdvectSHIFT(f3d_hip_ctx* ctx, double* F, const int* Nshift) {
  NEED(ctx->f.Pressure, "Pressure");
  Range r = Range{0, ctx->m.sy, 0, ctx->m.dimz ? ctx->m.sz : 1};
  DISPATCH(ctx, k_shift_gather, r, F, ctx->f.Pressure, Nshift);
  F3D_CHECK(ctx, hipMemcpyAsync(F, ctx->f.Pressure, ctx->n3 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

int f3d_reduction_SUM(f3d_hip_ctx* ctx, const double* F, int ymin, int ymax, int zmin, int zmax) {
  NEED(ctx->f.Reduction2D, "Reduction2D");
  if (!ctx->m.dimz) { zmin = 0; zmax = 1; }
  return ring_reduce(ctx, F, ctx->f.Reduction2D, ymin, ymax, zmin, zmax, 0);
}

int f3d_reduction_MIN(f3d_hip_ctx* ctx, const double* F, int ymin, int ymax, int zmin, int zmax) {
  NEED(ctx->f.Reduction2D, "Reduction2D");
  if (!ctx->m.dimz) { zmin = 0; zmax = 1; }
  return ring_reduce(ctx, F, ctx->f.Reduction2D, ymin, ymax, zmin, zmax, 1);
}

static int full_reduce(f3d_hip_ctx* ctx, const double* F, int ymin, int ymax, int zmin, int zmax, int is_min,
                       double* result) {
  NEED(ctx->f.Reduction2D, "Reduction2D");
  if (!ctx->m.dimz) { zmin = 0; zmax = 1; }
  if (ring_reduce(ctx, F, ctx->f.Reduction2D, ymin, ymax, zmin, zmax, is_min)) return -1;
  hipLaunchKernelGGL(k_full_finish, dim3(1), dim3(64), 0, ctx->stream, ctx->m, ctx->f.Reduction2D, ctx->dred,
                     ymin, ymax, zmin, zmax, is_min);
  F3D_LAUNCH_CHECK(ctx);
  F3D_CHECK(ctx, hipMemcpyAsync(ctx->hres, ctx->dred, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *result = ctx->hres[0];
  return 0;
}

int f3d_reduction_full_SUM(f3d_hip_ctx* ctx, const double* F, int ymin, int ymax, int zmin, int zmax, double* result) {
  return full_reduce(ctx, F, ymin, ymax, zmin, zmax, 0, result);
}

int f3d_reduction_full_MIN(f3d_hip_ctx* ctx, const double* F, int ymin, int ymax, int zmin, int zmax, double* result) {
  return full_reduce(ctx, F, ymin, ymax, zmin, zmax, 1, result);
}

int f3d_ComputeForce(f3d_hip_ctx* ctx, double x, double y, double z, double rsmoothing, double mass,
                     double lf[12]) {
  (void)mass;
  NEED(ctx->f.Total_Density, "Total_Density");
  const dmesh& m = ctx->m;
  Range r = rng(ctx, m.nghy, m.ny + m.nghy, m.nghz, m.nz + m.nghz);
  DISPATCH(ctx, k_force, r, ctx->f.Total_Density, ctx->f.Mmx, ctx->f.Mpx, ctx->f.Mmy, ctx->f.Mpy, ctx->f.Vx_temp,
           ctx->f.Vy_temp, x, y, z, rsmoothing);
  // split into inner/outer sums at the planet's ring (src/compute_force.c:150-217)
  double aa = sqrt(x * x + y * y + z * z);
  // the host mesh copy is needed for Ymin; it lives on the device, read two values
  double ylo, yhi;
  F3D_CHECK(ctx, hipMemcpyAsync(&ylo, m.ymin + m.nghy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipMemcpyAsync(&yhi, m.ymin + m.ny + m.nghy + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  int index = (int)(((aa - ylo) / (yhi - ylo)) * (double)m.ny) + m.nghy;
  double* flds[6] = {ctx->f.Mmx, ctx->f.Mpx, ctx->f.Mmy, ctx->f.Mpy, ctx->f.Vx_temp, ctx->f.Vy_temp};
  int zb = m.nghz, ze = m.nz + m.nghz;
  for (int c = 0; c < 12; c++) lf[c] = 0.0;
  for (int c = 0; c < 6; c++) {
    if (index >= m.nghy) {
      if (index < m.ny + m.nghy) {
        if (full_reduce(ctx, flds[c], m.nghy, index, zb, ze, 0, &lf[c])) return -1;
        if (full_reduce(ctx, flds[c], index, m.ny + m.nghy, zb, ze, 0, &lf[6 + c])) return -1;
      } else {
        if (full_reduce(ctx, flds[c], m.nghy, m.ny + m.nghy, zb, ze, 0, &lf[c])) return -1;
      }
    } else {
      if (full_reduce(ctx, flds[c], m.nghy, m.ny + m.nghy, zb, ze, 0, &lf[6 + c])) return -1;
    }
  }
  return 0;
}

static cell::CflPar cfl_par(const f3d_hip_ctx* ctx) {
  cell::CflPar c;
  c.cfl = ctx->ph.cfl;
  c.nu = ctx->ph.nu;
  c.alpha = ctx->ph.alpha;
  c.gamma = ctx->ph.gamma;
  c.iso = (ctx->ph.eos != 2);
  c.alpha_visc = (ctx->ph.visc == 2);
  c.fargo = ctx->ph.fargo;
  return c;
}

int f3d_cfl(f3d_hip_ctx* ctx, double* fluid_min) {
  NEED(ctx->f.DensStar, "DensStar");
  const dmesh& m = ctx->m;
  Range r = rng(ctx, m.nghy, m.ny + m.nghy, m.nghz, m.nz + m.nghz);
  DISPATCH(ctx, k_cfl, r, ctx->f.Energy, ctx->f.Density, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.VxMed,
           ctx->f.DensStar, cfl_par(ctx));
  int kb = m.dimz ? m.nghz : 0, ke = m.dimz ? m.nz + m.nghz : 1;
  if (ring_reduce(ctx, ctx->f.DensStar, ctx->f.Reduction2D, m.nghy, m.ny + m.nghy, kb, ke, 1)) return -1;
  hipLaunchKernelGGL(k_cfl_finish, dim3(1), dim3(64), 0, ctx->stream, m, ctx->f.Reduction2D, ctx->f.VxMed, ctx->dred,
                     ctx->ph.fargo, ctx->ph.cfl);
  F3D_LAUNCH_CHECK(ctx);
  F3D_CHECK(ctx, hipMemcpyAsync(ctx->hres, ctx->dred, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *fluid_min = ctx->hres[0];
  return 0;
}

static cell::DampPar damp_par(const f3d_hip_ctx* ctx) {
  cell::DampPar d;
  d.Y_inf = ctx->ph.Y_inf;
  d.Y_sup = ctx->ph.Y_sup;
  d.Z_inf = ctx->ph.Z_inf;
  d.Z_sup = ctx->ph.Z_sup;
  d.y_min = ctx->ph.ymin_par;
  d.y_max = ctx->ph.ymax_par;
  d.z_min = ctx->ph.zmin_par;
  d.z_max = ctx->ph.zmax_par;
  d.ds = ctx->ph.taudamp;
  d.of = ctx->ph.OMEGAFRAME;
  d.of0 = ctx->ph.OMEGAFRAME0;
  d.periodic_z = ctx->ph.periodicz;
  return d;
}

int f3d_StockholmBoundary(f3d_hip_ctx* ctx, double dt) {
  NEED(ctx->f.Density0, "Density0"); NEED(ctx->f.Vx0, "Vx0"); NEED(ctx->f.Vy0, "Vy0");
  Range r = rng(ctx, 0, ctx->m.sy, 0, ctx->m.sz);
  DISPATCH(ctx, k_stockholm, r, dt, ctx->f.Density, ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.Density0, ctx->f.Vx0,
           ctx->f.Vy0, ctx->f.Vz0, damp_par(ctx));
  return 0;
}

// the device CondInit of either form (fp32: FP32 fields, vx as the deviation
// from the handle's profile)
static int condinit(f3d_hip_ctx* ctx, const f3d_ic_desc* d, void* dens, void* energy, void* vx, void* vy, void* vz,
                    int fp32) {
  F3D_REQUIRE(ctx, d && d->R && d->V && d->CS && d->E, "CondInit: missing profiles");
  F3D_REQUIRE(ctx, dens && energy && vx && vy && (vz || !ctx->m.dimz), "CondInit: missing fields");
  F3D_REQUIRE(ctx, d->ndraw >= 0 && d->ndraw <= 3, "CondInit: at most three draws per cell");
  F3D_REQUIRE(ctx, !fp32 || ctx->vbase, "CondInit: the precision form's profile is not set");
  const dmesh& m = ctx->m;
  const size_t n2 = (size_t)m.sy * m.sz;
  double* prof = nullptr;
  F3D_CHECK(ctx, hipMalloc(&prof, 4 * n2 * sizeof(double)));
  const double* src[4] = {d->R, d->V, d->CS, d->E};
  for (int a = 0; a < 4; a++) {
    const hipError_t e = hipMemcpyAsync(prof + a * n2, src[a], n2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(prof);
      F3D_CHECK(ctx, e);
    }
  }
  IcPar p{d->mode, d->ndraw, d->adiabatic, d->noise, d->aspectratio, d->gamma};
  IcJump jp;
  jp.x0 = d->x0 & ((1ULL << 48) - 1);
  f3d_drand48_jumps(d->ndraw ? d->ndraw : 1, jp.A, jp.C);
  const dim3 g((m.nx + BX - 1) / BX, m.sy, m.sz);
  if (fp32)
    hipLaunchKernelGGL(k_condinit<float>, g, dim3(BX), 0, ctx->stream, m, p, jp, prof, prof + n2, prof + 2 * n2,
                       prof + 3 * n2, (float*)dens, (float*)energy, (float*)vx, (float*)vy, (float*)vz, ctx->vbase);
  else
    hipLaunchKernelGGL(k_condinit<double>, g, dim3(BX), 0, ctx->stream, m, p, jp, prof, prof + n2, prof + 2 * n2,
                       prof + 3 * n2, (double*)dens, (double*)energy, (double*)vx, (double*)vy, (double*)vz,
                       (const double*)nullptr);
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipFree(prof);
  F3D_CHECK(ctx, le);
  F3D_CHECK(ctx, se);
  return 0;
}

int f3d_CondInit(f3d_hip_ctx* ctx, const f3d_ic_desc* d, double* dens, double* energy, double* vx, double* vy,
                 double* vz) {
  return condinit(ctx, d, dens, energy, vx, vy, vz, 0);
}

int f3d32_CondInit(f3d_hip_ctx* ctx, const f3d_ic_desc* d, double* dens, double* energy, double* vx, double* vy,
                   double* vz) {
  return condinit(ctx, d, dens, energy, vx, vy, vz, 1);
}

int f3d_init_stockholm(f3d_hip_ctx* ctx) {
  NEED(ctx->f.Density0, "Density0"); NEED(ctx->f.Vx0, "Vx0"); NEED(ctx->f.Vy0, "Vy0");
  int n = ctx->m.sy * (ctx->m.dimz ? ctx->m.sz : 1);
  hipLaunchKernelGGL(k_init_stockholm, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->m, ctx->f.Density,
                     ctx->f.Vx, ctx->f.Vy, ctx->f.Vz, ctx->f.Energy, ctx->f.Density0, ctx->f.Vx0, ctx->f.Vy0,
                     ctx->f.Vz0, ctx->f.Energy0, (int)(ctx->ph.eos == 2));
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}


// the monitored quantity q (MON_*) of the bound fluid over the active cells
static int mon_launch(f3d_hip_ctx* ctx, int q, int fp32, double xp, double yp, double zp, double rsmoothing,
                      double* out) {
  F3D_REQUIRE(ctx, out != nullptr, "monitor: null output array");
  NEED(ctx->f.Density, "Density");
  const double* v = nullptr;
  if (q == MON_MOMX) { NEED(ctx->f.Vx, "Vx"); v = ctx->f.Vx; }
  if (q == MON_MOMY) { NEED(ctx->f.Vy, "Vy"); v = ctx->f.Vy; }
  if (q == MON_MOMZ) {
    F3D_REQUIRE(ctx, ctx->m.dimz, "mon_momz: the mesh has no Z dimension");
    NEED(ctx->f.Vz, "Vz");
    v = ctx->f.Vz;
  }
  const dmesh& m = ctx->m;
  Range r = rng(ctx, m.nghy, m.nghy + m.ny, m.nghz, m.nghz + m.nz);
  if (empty(r)) return 0;
  const dim3 g = grid_of(ctx, r);
  const double of = ctx->ph.OMEGAFRAME, rsm2 = rsmoothing * rsmoothing;
#define MON_Q(G_, DZ_, Q_)                                                                                    \
  do {                                                                                                        \
    if (fp32)                                                                                                 \
      hipLaunchKernelGGL((k_mon<G_, DZ_, Q_, float>), g, dim3(BX), 0, ctx->stream, m, r,                      \
                         reinterpret_cast<const float*>(ctx->f.Density), reinterpret_cast<const float*>(v),   \
                         out, of, xp, yp, zp, rsm2, ctx->vbase);                                              \
    else                                                                                                      \
      hipLaunchKernelGGL((k_mon<G_, DZ_, Q_, double>), g, dim3(BX), 0, ctx->stream, m, r, ctx->f.Density, v, \
                         out, of, xp, yp, zp, rsm2, nullptr);                                                 \
  } while (0)
#define MON_G(G_, DZ_)                                  \
  do {                                                  \
    switch (q) {                                        \
      case MON_DENS: MON_Q(G_, DZ_, MON_DENS); break;   \
      case MON_MOMX: MON_Q(G_, DZ_, MON_MOMX); break;   \
      case MON_MOMY: MON_Q(G_, DZ_, MON_MOMY); break;   \
      case MON_MOMZ: MON_Q(G_, DZ_, MON_MOMZ); break;   \
      default: MON_Q(G_, DZ_, MON_TORQ); break;         \
    }                                                   \
  } while (0)
  if (m.geom == SPH) {
    if (m.dimz) MON_G(SPH, 1); else MON_G(SPH, 0);
  } else {
    if (m.dimz) MON_G(CYL, 1); else MON_G(CYL, 0);
  }
#undef MON_G
#undef MON_Q
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

int f3d_mon_dens(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_DENS, 0, 0, 0, 0, 0, out); }
int f3d_mon_momx(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMX, 0, 0, 0, 0, 0, out); }
int f3d_mon_momy(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMY, 0, 0, 0, 0, 0, out); }
int f3d_mon_momz(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMZ, 0, 0, 0, 0, 0, out); }
int f3d_mon_torq(f3d_hip_ctx* ctx, double xp, double yp, double zp, double rsmoothing, double* out) {
  return mon_launch(ctx, MON_TORQ, 0, xp, yp, zp, rsmoothing, out);
}
int f3d32_mon_dens(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_DENS, 1, 0, 0, 0, 0, out); }
int f3d32_mon_momx(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMX, 1, 0, 0, 0, 0, out); }
int f3d32_mon_momy(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMY, 1, 0, 0, 0, 0, out); }
int f3d32_mon_momz(f3d_hip_ctx* ctx, double* out) { return mon_launch(ctx, MON_MOMZ, 1, 0, 0, 0, 0, out); }
int f3d32_mon_torq(f3d_hip_ctx* ctx, double xp, double yp, double zp, double rsmoothing, double* out) {
  return mon_launch(ctx, MON_TORQ, 1, xp, yp, zp, rsmoothing, out);
}

}  // extern "C"
