// f3d_cell.h -- per-cell arithmetic of the hydro step, shared by the
// one-kernel-per-routine entry points and the fused kernels.
//
// Each function evaluates one cell of one reference routine.  Field values
// are read through "loaders": callables L(di, dj, dk) returning the field at
// the cell offset (di, dj, dk) from the current cell, x periodic.  A global
// loader reads HBM, a tile loader reads LDS, so both kernel families run the
// identical sequence of floating-point operations.  Every expression keeps
// the association of the reference statement it restates (cited per
// function); the build uses -ffp-contract=off.
#pragma once

#include "f3d_device.h"

namespace f3d {
namespace cell {

constexpr double CVNR = 1.41;  // src/define.h:9
constexpr int kMaxPlanets = 8;

// Loader over a dense global array with x wrap (src/define.h:391-392).
struct GL {
  const double* p;
  int l, i, nx, pitch, stride;
  __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
    int off = di;
    int ii = i + di;
    if (ii < 0) off += nx;
    else if (ii >= nx) off -= nx;
    return p[l + off + dj * pitch + dk * stride];
  }
};

__device__ __forceinline__ GL gl(const double* p, const dmesh& m, int l, int i) {
  return GL{p, l, i, m.nx, m.pitch, m.stride};
}

// ---------------------------------------------------------------- SubStep1
// src/substep1_x.c:112-120
template <int G, class LR, class LP, class LPOT, class LV>
__device__ __forceinline__ double substep1_x(const dmesh& m, int i, int j, int k, double dt, LR rho, LP p,
                                             LPOT pot, LV vx, int dust, int potential) {
  const int im = ixm_of(i, m.nx);
  double dxmed = 0.5 * (m.sxi[i] + m.sxi[im]);
  double dtOVERrhom = 2. * dt / (rho(0, 0, 0) * m.sxi[i] + rho(-1, 0, 0) * m.sxi[im]) * dxmed;
  double v = vx(0, 0, 0);
  const double izs = inv_zone_size_xmed<G>(m, i, j, k);
  if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(-1, 0, 0)) * izs;
  if (potential) v -= (pot(0, 0, 0) - pot(-1, 0, 0)) * dt * izs;
  return v;
}

// src/substep1_y.c:313-316, 352-355 (CYL), 379-383 (SPH), 408-410
template <int G, int DZ, class LR, class LP, class LPOT, class LVX, class LVY, class LVZ>
__device__ __forceinline__ double substep1_y(const dmesh& m, int i, int j, int k, double dt, LR rho, LP p,
                                             LPOT pot, LVX vx, LVY vy, LVZ vz, int dust, int potential,
                                             double omegaframe) {
  double dtOVERrhom = 2.0 * dt / (rho(0, 0, 0) * (m.ymin[j + 1] - m.ymin[j]) + rho(0, -1, 0) * (m.ymin[j] - m.ymin[j - 1])) *
                      (m.ymed[j] - m.ymed[j - 1]);
  double v = vy(0, 0, 0);
  if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]);
  double vphi = .25 * (vx(0, 0, 0) + vx(1, 0, 0) + vx(0, -1, 0) + vx(1, -1, 0));
  if (G == SPH) {
    vphi += m.ymin[j] * m.sinzmed[k] * omegaframe;
    double vzz = DZ ? .25 * (vz(0, 0, 0) + vz(0, 0, 1) + vz(0, -1, 0) + vz(0, -1, 1)) : 0.0;
    v += (vphi * vphi + vzz * vzz) / m.ymin[j] * dt;
  } else {
    vphi += m.ymin[j] * omegaframe;
    v += vphi * vphi / m.ymin[j] * dt;
  }
  if (potential) v -= (pot(0, 0, 0) - pot(0, -1, 0)) * dt / (m.ymed[j] - m.ymed[j - 1]);
  return v;
}

// src/substep1_z.c:120-183 (CYLINDRICAL and SPHERICAL branches)
template <int G, class LR, class LP, class LPOT, class LVX, class LVZ>
__device__ __forceinline__ double substep1_z(const dmesh& m, int i, int j, int k, double dt, LR rho, LP p,
                                             LPOT pot, LVX vx, LVZ vz, int dust, int potential,
                                             double omegaframe) {
  double dtOVERrhom = 2.0 * dt / (rho(0, 0, 0) * (m.zmin[k + 1] - m.zmin[k]) + rho(0, 0, -1) * (m.zmin[k] - m.zmin[k - 1])) *
                      (m.zmed[k] - m.zmed[k - 1]);
  double v;
  if (G == SPH) {
    double vphi = .25 * (vx(0, 0, 0) + vx(1, 0, 0) + vx(0, 0, -1) + vx(1, 0, -1));
    vphi += m.ymed[j] * m.sinzmin[k] * omegaframe;
    v = vz(0, 0, 0);
    if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(0, 0, -1)) / (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1]));
    v += vphi * vphi * m.coszmin[k] / (m.sinzmin[k] * m.ymed[j]) * dt;
    if (potential) v -= (pot(0, 0, 0) - pot(0, 0, -1)) * dt / (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1]));
  } else {
    v = vz(0, 0, 0);
    if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(0, 0, -1)) / (m.zmed[k] - m.zmed[k - 1]);
    if (potential) v -= (pot(0, 0, 0) - pot(0, 0, -1)) * dt / (m.zmed[k] - m.zmed[k - 1]);
  }
  return v;
}

// ---------------------------------------------------------------- SubStep2
// artificial pressure (src/substep2_a.c:103-113)
__device__ __forceinline__ double art_q(double rho, double dv) {
  if (dv < 0.0) return CVNR * CVNR * rho * dv * dv;
  return 0.0;
}

// src/substep2_b.c:145-190; inputs are the current v_temp (and e) values
template <int G, int DZ, class LR, class LVX, class LVY, class LVZ, class LQX, class LQY, class LQZ>
__device__ __forceinline__ void substep2_b(const dmesh& m, int i, int j, int k, double dt, LR rho, LVX vx, LVY vy,
                                           LVZ vz, LQX px, LQY py, LQZ pz, int adiabatic, double vxt, double vyt,
                                           double vzt, double e, double& oxt, double& oyt, double& ozt,
                                           double& oe) {
  const int im = ixm_of(i, m.nx);
  double dxrho1 = m.sxi[i];
  double dxrho2 = m.sxi[im];
  double dxmed = 0.5 * (m.sxi[i] + m.sxi[im]);
  vxt += -2.0 * dt * dxmed * (px(0, 0, 0) - px(-1, 0, 0)) / (rho(0, 0, 0) * dxrho1 + rho(-1, 0, 0) * dxrho2) *
         inv_zone_size_xmed<G>(m, i, j, k);
  if (adiabatic) e += -dt * (px(0, 0, 0) * (vx(1, 0, 0) - vx(0, 0, 0)) / zone_size_x<G>(m, i, j, k));
  double dyrho1 = m.ymin[j + 1] - m.ymin[j];
  double dyrho2 = m.ymin[j] - m.ymin[j - 1];
  vyt += -2.0 * (py(0, 0, 0) - py(0, -1, 0)) / (rho(0, 0, 0) * dyrho1 + rho(0, -1, 0) * dyrho2) * dt;
  if (adiabatic) e += -dt * (py(0, 0, 0) * (vy(0, 1, 0) - vy(0, 0, 0)) / zone_size_y(m, j));
  if (DZ) {
    double dzmed = m.zmed[k] - m.zmed[k - 1];
    double dzrho1 = m.zmin[k + 1] - m.zmin[k];
    double dzrho2 = m.zmin[k] - m.zmin[k - 1];
    if (G == SPH)
      vzt += -2.0 * dzmed * (pz(0, 0, 0) - pz(0, 0, -1)) / (rho(0, 0, 0) * dzrho1 + rho(0, 0, -1) * dzrho2) * dt /
             (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1]));
    else
      vzt += -2.0 * (pz(0, 0, 0) - pz(0, 0, -1)) / (rho(0, 0, 0) * dzrho1 + rho(0, 0, -1) * dzrho2) * dt;
    if (adiabatic) e += -dt * (pz(0, 0, 0) * (vz(0, 0, 1) - vz(0, 0, 0)) / zone_size_z<G>(m, j, k));
  }
  oxt = vxt;
  oyt = vyt;
  ozt = vzt;
  oe = e;
}

// velocity divergence with the SurfX/SurfY/SurfZ forms (src/substep3.c:98-105)
template <int DZ, class LVX, class LVY, class LVZ>
__device__ __forceinline__ double div_v(const dmesh& m, int i, int j, int k, LVX vx, LVY vy, LVZ vz) {
  double d = 0.0;
  d += mul_surfx(m, vx(1, 0, 0) - vx(0, 0, 0), j, k);
  d += (mul_surfy(m, vy(0, 1, 0), i, j + 1, k) - mul_surfy(m, vy(0, 0, 0), i, j, k));
  if (DZ) d += (mul_surfz(m, vz(0, 0, 1), i, j, k + 1) - mul_surfz(m, vz(0, 0, 0), i, j, k));
  return d;
}

// src/substep3.c:106-110
template <int DZ, class LVX, class LVY, class LVZ>
__device__ __forceinline__ double substep3(const dmesh& m, int i, int j, int k, double dt, double e, LVX vx,
                                           LVY vy, LVZ vz, double gamma) {
  double dv = div_v<DZ>(m, i, j, k, vx, vy, vz);
  double term = mul_invvol(m, 0.5 * dt * (gamma - 1.) * dv, i, j, k);
  e *= (1.0 - term) / (1.0 + term);
  return e;
}

// ---------------------------------------------------------------- viscosity
struct ViscPar {
  int model;  // 1 NU, 2 ALPHA
  int iso;
  double nu, alpha, gamma;
};

struct Taus {
  double xx, yy, zz, xz, yx, zy;
};

// src/visctensor_cyl.c:140-198 and src/visctensor_sph.c:126-206
template <int G, int DZ, class LR, class LE, class LVX, class LVY, class LVZ>
__device__ __forceinline__ Taus visctensor(const dmesh& m, int i, int j, int k, LR rho, LE energy, LVX vx, LVY vy,
                                           LVZ vz, const ViscPar& vp) {
  double visc, viscm, visczm = 0.0, visczmym = 0.0;
  if (vp.model == 2) {
    const double sq_med = sqrt(m.ymed[j] * m.ymed[j] * m.ymed[j] / (1.0 * 1.0));
    const double sq_min = sqrt(m.ymin[j] * m.ymin[j] * m.ymin[j] / (1.0 * 1.0));
    if (vp.iso) {
      visc = vp.alpha * energy(0, 0, 0) * energy(0, 0, 0) * sq_med;
      double s = energy(0, 0, 0) + energy(-1, 0, 0) + energy(0, -1, 0) + energy(-1, -1, 0);
      viscm = vp.alpha * 0.0625 * s * s * sq_min;
      if (DZ) {
        double szm = energy(0, 0, 0) + energy(0, 0, -1) + energy(-1, 0, 0) + energy(-1, 0, -1);
        visczm = vp.alpha * 0.0625 * szm * szm * sq_med;
        double szy = (G == SPH) ? energy(0, 0, 0) + energy(0, 0, -1) + energy(0, -1, 0) + energy(0, -1, -1)
                                : energy(0, 0, 0) + energy(0, -1, 0) + energy(0, 0, -1) + energy(0, -1, -1);
        visczmym = vp.alpha * 0.0625 * szy * szy * sq_min;
      }
    } else {
      const double gg = vp.alpha * vp.gamma * (vp.gamma - 1.0);
      visc = gg * energy(0, 0, 0) / rho(0, 0, 0) * sq_med;
      viscm = gg * (energy(0, 0, 0) + energy(-1, 0, 0) + energy(0, -1, 0) + energy(-1, -1, 0)) /
              (rho(0, 0, 0) + rho(-1, 0, 0) + rho(0, -1, 0) + rho(-1, -1, 0)) * sq_min;
      if (DZ) {
        visczm = gg * (energy(0, 0, 0) + energy(0, 0, -1) + energy(-1, 0, 0) + energy(-1, 0, -1)) /
                 (rho(0, 0, 0) + rho(0, 0, -1) + rho(-1, 0, 0) + rho(-1, 0, -1)) * sq_med;
        visczmym = gg * (energy(0, 0, 0) + energy(0, -1, 0) + energy(0, 0, -1) + energy(0, -1, -1)) /
                   (rho(0, 0, 0) + rho(0, -1, 0) + rho(0, 0, -1) + rho(0, -1, -1)) * sq_min;
      }
    }
  } else {
    visc = viscm = visczm = visczmym = vp.nu;
  }
  double dv = div_v<DZ>(m, i, j, k, vx, vy, vz);
  dv *= mul_invvol(m, 2.0 / 3.0, i, j, k);
  const double r0 = rho(0, 0, 0);
  Taus t;
  t.xx = visc * r0 * (2.0 * (vx(1, 0, 0) - vx(0, 0, 0)) / zone_size_x<G>(m, i, j, k) - dv);
  t.xx += visc * r0 * (vy(0, 1, 0) + vy(0, 0, 0)) / m.ymed[j];
  if (G == SPH && DZ) t.xx += visc * r0 * (vz(0, 0, 1) + vz(0, 0, 0)) * m.coszmed[k] / (m.sinzmed[k] * m.ymed[j]);
  t.yy = visc * r0 * (2.0 * (vy(0, 1, 0) - vy(0, 0, 0)) / (m.ymin[j + 1] - m.ymin[j]) - dv);
  t.zz = t.xz = t.zy = 0.0;
  if (DZ) {
    if (G == SPH) {
      t.zz = visc * r0 * (2.0 * (vz(0, 0, 1) - vz(0, 0, 0)) / (m.ymed[j] * (m.zmin[k + 1] - m.zmin[k])) - dv);
      t.zz += visc * r0 * (vy(0, 0, 0) + vy(0, 1, 0)) / m.ymed[j];
      t.xz = visczm * .25 * (r0 + rho(0, 0, -1) + rho(-1, 0, 0) + rho(-1, 0, -1)) *
             ((vx(0, 0, 0) / m.sinzmed[k] - vx(0, 0, -1) / m.sinzmed[k - 1]) * m.sinzmin[k] /
                  (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1])) +
              ((vz(0, 0, 0) - vz(-1, 0, 0)) * m.invdiffxmed[i] / (m.sinzmin[k] * m.ymed[j])));
    } else {
      t.zz = visc * r0 * (2.0 * (vz(0, 0, 1) - vz(0, 0, 0)) / (m.zmin[k + 1] - m.zmin[k]) - dv);
      t.xz = visczm * .25 * (r0 + rho(0, 0, -1) + rho(-1, 0, 0) + rho(-1, 0, -1)) *
             ((vx(0, 0, 0) - vx(0, 0, -1)) / (m.zmed[k] - m.zmed[k - 1]) +
              (vz(0, 0, 0) - vz(-1, 0, 0)) * m.invdiffxmed[i] / m.ymed[j]);
    }
  }
  if (G == SPH)
    t.yx = viscm * .25 * (r0 + rho(-1, 0, 0) + rho(0, -1, 0) + rho(-1, -1, 0)) *
           ((vy(0, 0, 0) - vy(-1, 0, 0)) * m.invdiffxmed[i] / (m.ymin[j] * m.sinzmed[k]) +
            (vx(0, 0, 0) - vx(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - .5 * (vx(0, 0, 0) + vx(0, -1, 0)) / m.ymin[j]);
  else
    t.yx = viscm * .25 * (r0 + rho(-1, 0, 0) + rho(0, -1, 0) + rho(-1, -1, 0)) *
           ((vy(0, 0, 0) - vy(-1, 0, 0)) * m.invdiffxmed[i] / m.ymin[j] +
            (vx(0, 0, 0) - vx(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - .5 * (vx(0, 0, 0) + vx(0, -1, 0)) / m.ymin[j]);
  if (DZ) {
    if (G == SPH)
      t.zy = visczmym * .25 * (r0 + rho(0, -1, 0) + rho(0, 0, -1) + rho(0, -1, -1)) *
             ((vz(0, 0, 0) - vz(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - .5 * (vz(0, 0, 0) + vz(0, -1, 0)) / m.ymin[j] +
              (vy(0, 0, 0) - vy(0, 0, -1)) / (m.ymin[j] * (m.zmed[k] - m.zmed[k - 1])));
    else
      t.zy = visczmym * .25 * (r0 + rho(0, -1, 0) + rho(0, 0, -1) + rho(0, -1, -1)) *
             ((vz(0, 0, 0) - vz(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) +
              (vy(0, 0, 0) - vy(0, 0, -1)) / (m.zmed[k] - m.zmed[k - 1]));
  }
  return t;
}

// src/addviscosity_cyl.c:97-131 and src/addviscosity_sph.c:97-143
template <int G, int DZ, class LR, class LXX, class LYY, class LZZ, class LXZ, class LYX, class LZY>
__device__ __forceinline__ void addviscosity(const dmesh& m, int i, int j, int k, double dt, LR rho, LXX txx,
                                             LYY tyy, LZZ tzz, LXZ txz, LYX tyx, LZY tzy, double vx, double vy,
                                             double vz, double& ox, double& oy, double& oz) {
  const double r0 = rho(0, 0, 0), rxm = rho(-1, 0, 0), rym = rho(0, -1, 0);
  const double ymn = m.ymin[j], ymp = m.ymin[j + 1], yd = m.ymed[j], ydm = m.ymed[j - 1];
  vx += 2.0 * (txx(0, 0, 0) - txx(-1, 0, 0)) * inv_zone_size_xmed<G>(m, i, j, k) / ((r0 + rxm)) * dt;
  vx += 2.0 * (ymp * ymp * tyx(0, 1, 0) - ymn * ymn * tyx(0, 0, 0)) / ((ymp - ymn) * yd * yd * (rxm + r0)) * dt;
  if (G == SPH) {
    vx += (tyx(0, 1, 0) + tyx(0, 0, 0)) / ((rxm + r0) * yd) * dt;
    if (DZ) {
      vx += 2.0 * (txz(0, 0, 1) - txz(0, 0, 0)) / (zone_size_z<G>(m, j, k) * (rxm + r0)) * dt;
      vx += 2.0 * (txz(0, 0, 1) + txz(0, 0, 0)) * m.coszmed[k] / (m.sinzmed[k] * yd * (rxm + r0)) * dt;
    }
    vy += 2.0 * (yd * yd * tyy(0, 0, 0) - ydm * ydm * tyy(0, -1, 0)) / ((yd - ydm) * (r0 + rym) * ymn * ymn) * dt;
    vy += 2.0 * (tyx(1, 0, 0) - tyx(0, 0, 0)) / ((m.xmin[i + 1] - m.xmin[i]) * ymn * m.sinzmed[k] * (r0 + rym)) * dt;
    vy -= (txx(0, 0, 0) + txx(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
    if (DZ) {
      const double rzm = rho(0, 0, -1);
      vy -= (tzz(0, 0, 0) + tzz(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
      vy += 2.0 * (tzy(0, 0, 1) * m.sinzmin[k + 1] - tzy(0, 0, 0) * m.sinzmin[k]) /
            ((m.zmin[k + 1] - m.zmin[k]) * ymn * m.sinzmed[k] * (r0 + rym)) * dt;
      vz += 2.0 * (tzz(0, 0, 0) * m.sinzmed[k] - tzz(0, 0, -1) * m.sinzmed[k - 1]) /
            ((m.zmed[k] - m.zmed[k - 1]) * yd * m.sinzmin[k] * (r0 + rzm)) * dt;
      vz += 2.0 * (txz(1, 0, 0) - txz(0, 0, 0)) / ((m.sxi[i] * yd * m.sinzmin[k]) * (r0 + rzm)) * dt;
      vz += 2.0 * (ymp * ymp * ymp * tzy(0, 1, 0) - ymn * ymn * ymn * tzy(0, 0, 0)) /
            ((ymp - ymn) * yd * yd * yd * (r0 + rzm)) * dt;
      vz -= (txx(0, 0, 0) + txx(0, 0, -1)) * m.coszmin[k] / (yd * m.sinzmin[k] * (r0 + rzm)) * dt;
    }
  } else {
    if (DZ) vx += 2.0 * (txz(0, 0, 1) - txz(0, 0, 0)) / ((m.zmin[k + 1] - m.zmin[k]) * (rxm + r0)) * dt;
    vy += 2.0 * (yd * tyy(0, 0, 0) - ydm * tyy(0, -1, 0)) / ((yd - ydm) * (r0 + rym) * ymn) * dt;
    vy += 2.0 * (tyx(1, 0, 0) - tyx(0, 0, 0)) / ((m.xmin[i + 1] - m.xmin[i]) * ymn * (r0 + rym)) * dt;
    vy -= (txx(0, 0, 0) + txx(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
    if (DZ) {
      const double rzm = rho(0, 0, -1);
      vy += 2.0 * (tzy(0, 0, 1) - tzy(0, 0, 0)) / ((m.zmin[k + 1] - m.zmin[k]) * (r0 + rym)) * dt;
      vz += 2.0 * (tzz(0, 0, 0) - tzz(0, 0, -1)) / ((m.zmed[k] - m.zmed[k - 1]) * (r0 + rzm)) * dt;
      vz += 2.0 * (txz(1, 0, 0) - txz(0, 0, 0)) / (zone_size_x<G>(m, i, j, k) * (r0 + rzm)) * dt;
      vz += 2.0 * (ymp * tzy(0, 1, 0) - ymn * tzy(0, 0, 0)) / ((ymp - ymn) * yd * (r0 + rzm)) * dt;
    }
  }
  ox = vx;
  oy = vy;
  oz = vz;
}

// ---------------------------------------------------------------- potential
struct PlanetSet {
  int n, indirect;
  double x[kMaxPlanets], y[kMaxPlanets], z[kMaxPlanets], mp[kMaxPlanets], smooth2[kMaxPlanets];
};

// src/potential.c:100-196 (default star, planets, planet indirect term);
// planet mass already tapered, smoothing^2 evaluated on the host.
template <int G>
__device__ __forceinline__ double potential(const dmesh& m, int i, int j, int k, const PlanetSet& ps) {
  double XC, YC, ZC, pot;
  if (G == SPH) {
    XC = m.ymed[j] * m.cosxmed[i] * m.sinzmed[k];
    YC = m.ymed[j] * m.sinxmed[i] * m.sinzmed[k];
    ZC = m.ymed[j] * m.coszmed[k];
    pot = -1.0 * 1.0 / m.ymed[j];
  } else {
    XC = m.ymed[j] * m.cosxmed[i];
    YC = m.ymed[j] * m.sinxmed[i];
    ZC = m.zmed[k];
    pot = -1.0 * 1.0 / sqrt(m.ymed[j] * m.ymed[j] + ZC * ZC);
  }
  for (int n = 0; n < ps.n; n++) {
    const double mp = ps.mp[n];
    const double xp = ps.x[n], yp = ps.y[n], zp = ps.z[n];
    double planetdistance = sqrt(xp * xp + yp * yp + zp * zp);
    double dist = ((XC - xp) * (XC - xp) + (YC - yp) * (YC - yp) + (ZC - zp) * (ZC - zp));
    if (ps.indirect)
      pot += 1.0 * mp * (XC * xp + YC * yp + ZC * zp) / (planetdistance * planetdistance * planetdistance);
    pot += -1.0 * mp / sqrt(dist + ps.smooth2[n]);
  }
  return pot;
}

// ---------------------------------------------------------------- momenta / new velocities
// src/momenta_x.c:76-84
template <int G>
__device__ __forceinline__ void momenta_x(const dmesh& m, int i, int j, int k, double rho, double vx, double vxp,
                                          double omegaframe, double& mmx, double& mpx) {
  if (G == SPH) {
    double rcyl = m.ymed[j] * m.sinzmed[k];
    mmx = (vx + rcyl * omegaframe) * rcyl * rho;
    mpx = (vxp + rcyl * omegaframe) * rcyl * rho;
  } else {
    mmx = (vx + m.ymed[j] * omegaframe) * m.ymed[j] * rho;
    mpx = (vxp + m.ymed[j] * omegaframe) * m.ymed[j] * rho;
  }
}

// src/newvel_x.c:78-86
template <int G>
__device__ __forceinline__ double newvel_x(const dmesh& m, int i, int j, int k, double rho, double rhoxm, double mmx,
                                           double mpxm, double omegaframe) {
  const int im = ixm_of(i, m.nx);
  double r = (G == SPH) ? m.ymed[j] * m.sinzmed[k] : m.ymed[j];
  return (mul_vol(m, mmx, i, j, k) + mul_vol(m, mpxm, im, j, k)) /
             ((mul_vol(m, rho, i, j, k) + mul_vol(m, rhoxm, im, j, k)) * r) -
         omegaframe * r;
}

// src/newvel_y.c:67
__device__ __forceinline__ double newvel_y(const dmesh& m, int i, int j, int k, double rho, double rhoym, double mmy,
                                           double mpym) {
  return (mul_vol(m, mmy, i, j, k) + mul_vol(m, mpym, i, j - 1, k)) /
         (mul_vol(m, rho, i, j, k) + mul_vol(m, rhoym, i, j - 1, k));
}

// ---------------------------------------------------------------- PPA (src/fargo_ppa_*.c)
// monotonised-centred slope (src/fargo_ppa_a.c:58-69)
__device__ __forceinline__ double ppa_slope(double qm, double q, double qp) {
  double dqm = (q - qm);
  double dqp = (qp - q);
  if (dqp * dqm <= 0.0) return 0.0;
  double s = 0.5 * (qp - qm);
  double work = fabs(s);
  if (2.0 * fabs(dqm) < work) work = 2.0 * fabs(dqm);
  if (2.0 * fabs(dqp) < work) work = 2.0 * fabs(dqp);
  return (s < 0) ? -work : work;
}
// interface value q_{i+1/2} (src/fargo_ppa_b.c:52-55)
__device__ __forceinline__ double ppa_face(double q, double qp, double s, double sp) {
  return q + 0.5 * (qp - q) - 1.0 / 6.0 * (sp - s);
}
// Colella-Woodward monotonicity (src/fargo_ppa_c.c:52-63)
__device__ __forceinline__ void ppa_limit(double q, double& qL, double& qR) {
  if ((qR - q) * (q - qL) < 0.0) {
    qL = q;
    qR = q;
  }
  double diff = qR - qL;
  double cord = q - 0.5 * (qL + qR);
  if (6.0 * diff * cord > diff * diff) qL = 3.0 * q - 2.0 * qR;
  if (-diff * diff > 6.0 * diff * cord) qR = 3.0 * q - 2.0 * qL;
}
// parabolic average over the swept zone (src/fargo_ppa_d.c:62-70)
__device__ __forceinline__ double ppa_flux_value(double v, double dt, double zsx, double qm, double qLm, double qRm,
                                                 double q, double qL, double qR) {
  double qs;
  if (v > 0.0) {
    double ksi = v * dt / zsx;
    qs = qRm + ksi * (qm - qRm);
    qs += ksi * (1.0 - ksi) * (2.0 * qm - qRm - qLm);
  } else {
    double ksi = -v * dt / zsx;
    qs = qL + ksi * (q - qL);
    qs += ksi * (1.0 - ksi) * (2.0 * q - qR - qL);
  }
  return qs;
}

// ---------------------------------------------------------------- conservative updates
// src/update_x.c:69-71: ((a - b)*SurfX*dt*InvVol)
__device__ __forceinline__ double flux_x(const dmesh& m, int i, int j, int k, double dt, double a, double b) {
  return mul_invvol(m, mul_surfx(m, a - b, j, k) * dt, i, j, k);
}
// src/update_y.c:68-70: (a*SurfY(j) - b*SurfY(j+1))*dt*InvVol
__device__ __forceinline__ double flux_y(const dmesh& m, int i, int j, int k, double dt, double a, double b) {
  return mul_invvol(m, (mul_surfy(m, a, i, j, k) - mul_surfy(m, b, i, j + 1, k)) * dt, i, j, k);
}
__device__ __forceinline__ double flux_z(const dmesh& m, int i, int j, int k, double dt, double a, double b) {
  return mul_invvol(m, (mul_surfz(m, a, i, j, k) - mul_surfz(m, b, i, j, k + 1)) * dt, i, j, k);
}

// ---------------------------------------------------------------- cfl (src/cfl.c:174-253)
struct CflPar {
  double cfl, nu, alpha, gamma;
  int iso, alpha_visc, fargo;
};

template <int G, int DZ>
__device__ __forceinline__ double cfl_cell(const dmesh& m, int i, int j, int k, int l, const double* cs,
                                           const double* rho, const double* vx, const double* vy, const double* vz,
                                           double vxmed, const CflPar& cp) {
  const int lxp = lxp_of(l, i, m.nx), lyp = l + m.pitch, lzp = l + m.stride;
  double vxx, vxxp;
  if (!cp.fargo) {
    vxx = vx[l];
    vxxp = vx[lxp];
  } else {
    vxx = vx[l] - vxmed;
    vxxp = vx[lxp] - vxmed;
  }
  double soundspeed2, viscosity;
  if (cp.iso) soundspeed2 = cs[l] * cs[l];
  else soundspeed2 = cp.gamma * (cp.gamma - 1) * cs[l] / rho[l];
  if (cp.alpha_visc) {
    const double sq = sqrt(m.ymed[j] * m.ymed[j] * m.ymed[j] / (1.0 * 1.0));
    if (cp.iso) viscosity = cp.alpha * cs[l] * cs[l] * sq;
    else viscosity = cp.alpha * cp.gamma * (cp.gamma - 1.0) * cs[l] / rho[l] * sq;
  } else {
    viscosity = cp.nu;
  }
  double soundspeed = sqrt(soundspeed2);
  const double zx = zone_size_x<G>(m, i, j, k), zy = zone_size_y(m, j);
  const double zz = DZ ? zone_size_z<G>(m, j, k) : 1.0;
  double cfl1_a = soundspeed / zx;
  double cfl1_b = soundspeed / zy;
  double cfl1_c = DZ ? soundspeed / zz : 0.0;
  double t1 = cfl1_a > cfl1_b ? cfl1_a : cfl1_b;
  double cfl1 = t1 > cfl1_c ? t1 : cfl1_c;
  double ax = fabs(vxx), axp = fabs(vxxp);
  double cfl2 = (ax > axp ? ax : axp) / zx;
  double ay = fabs(vy[l]), ayp = fabs(vy[lyp]);
  double cfl3 = (ay > ayp ? ay : ayp) / zy;
  double cfl4 = 0.0;
  if (DZ) {
    double az = fabs(vz[l]), azp = fabs(vz[lzp]);
    cfl4 = (az > azp ? az : azp) / zz;
  }
  double cfl5_a = fabs(vx[lxp] - vx[l]) / zx;
  double cfl5_b = fabs(vy[lyp] - vy[l]) / zy;
  double cfl5_c = DZ ? fabs(vz[lzp] - vz[l]) / zz : 0.0;
  double t5 = cfl5_a > cfl5_b ? cfl5_a : cfl5_b;
  double cfl5 = (t5 > cfl5_c ? t5 : cfl5_c) * 4.0 * CVNR;
  double cfl7_a = 1.0 / zx;
  double cfl7_b = 1.0 / zy;
  double cfl7_c = DZ ? 1.0 / zz : 0.0;
  double t7 = cfl7_a > cfl7_b ? cfl7_a : cfl7_b;
  double m7 = t7 > cfl7_c ? t7 : cfl7_c;
  double cfl7 = 4.0 * viscosity * (m7 * m7);
  // the reference sums ten squares; the MHD/strong-shock terms (cfl6, cfl8-10)
  // are exact zeros here and adding +0.0 does not change the sum
  double s = cfl1 * cfl1 + cfl2 * cfl2 + cfl3 * cfl3 + cfl4 * cfl4 + cfl5 * cfl5 + 0.0 + cfl7 * cfl7;
  return cp.cfl / sqrt(s);
}

// ---------------------------------------------------------------- Stockholm damping
struct DampPar {
  double Y_inf, Y_sup, Z_inf, Z_sup, y_min, y_max, z_min, z_max, ds, of, of0;
  int periodic_z;
};

// src/stockholm.c:129-184
template <int G, int DZ>
__device__ __forceinline__ void stockholm(const dmesh& m, int j, int k, int l, double dt, double* rho, double* vx,
                                          double* vy, double* vz, const double* rho0, const double* vx0,
                                          const double* vy0, const double* vz0, const DampPar& d) {
  const double ym = m.ymed[j];
  double rampy = 0.0, rampz = 0.0, rampzz = 0.0;
  if (ym > d.Y_sup) rampy = (ym - d.Y_sup) / (d.y_max - d.Y_sup);
  if (ym < d.Y_inf) rampy = (d.Y_inf - ym) / (d.Y_inf - d.y_min);
  rampy *= rampy;
  if (DZ) {
    const double zm = m.zmed[k], zn = m.zmin[k];
    if (zm > d.Z_sup) rampz = (zm - d.Z_sup) / (d.z_max - d.Z_sup);
    if (zm < d.Z_inf) rampz = (d.Z_inf - zm) / (d.Z_inf - d.z_min);
    rampz = rampz * rampz;
    if (zn > d.Z_sup) rampzz = (zn - d.Z_sup) / (d.z_max - d.Z_sup);
    if (zn < d.Z_inf) rampzz = (d.Z_inf - zn) / (d.Z_inf - d.z_min);
    rampzz = rampzz * rampzz;
  }
  if (d.periodic_z) {
    rampz = 0.0;
    rampzz = 0.0;
  }
  double ramp = rampy + rampz;
  double rampi = rampy + rampzz;
  double tau = d.ds * sqrt(ym * ym * ym / 1.0 / 1.0);
  const int l2 = j + k * m.sy;
  if (ramp > 0.0) {
    double taud = tau / ramp;
    rho[l] = (rho[l] * taud + rho0[l2] * dt) / (dt + taud);
    double vx0_target = vx0[l2];
    double radius = ym;
    if (G == SPH) radius *= m.sinzmed[k];
    vx0_target -= (d.of - d.of0) * radius;
    vx[l] = (vx[l] * taud + vx0_target * dt) / (dt + taud);
    vy[l] = (vy[l] * taud + vy0[l2] * dt) / (dt + taud);
  }
  if (DZ) {
    if (rampi > 0.0) {
      double taud = tau / rampi;
      vz[l] = (vz[l] * taud + vz0[l2] * dt) / (dt + taud);
    }
  }
}

// ---------------------------------------------------------------- boundaries
struct BcPar {
  int p3d;
  int energy;  // Energy ghosts SYMMETRIC (centered |a|a|, std/boundaries.txt:55-57)
  double of;
  double dens[3], vsq[3], vfac[3], zvfac[3];
};

// Y sides (scripts/boundparser.py:323-340 indices; std/boundaries.txt:35-40,
// 62-66, 81-85 formulas; field lists setups/fargo/fargo.bound.0,
// setups/p3disof/p3disof.bound.0)
__device__ __forceinline__ void bound_y_cell(const dmesh& m, int side, int i, int mm, int k, double* dens, double* vx,
                                             double* vy, double* vz, double* e, const BcPar& b, const double* dexp) {
  const int pitch = m.pitch, stride = m.stride, ny = m.ny, nghy = m.nghy;
  int lgh, lghs, lact, lacts, lnull, jgh, jact;
  if (!side) {
    lgh = lghs = i + mm * pitch + k * stride;
    lact = i + (2 * nghy - mm - 1) * pitch + k * stride;
    lacts = i + (2 * nghy - mm) * pitch + k * stride;
    lnull = i + nghy * pitch + k * stride;
    jgh = mm;
    jact = 2 * nghy - mm - 1;
  } else {
    lgh = i + (ny + nghy + mm) * pitch + k * stride;
    lghs = i + (ny + nghy + 1 + mm) * pitch + k * stride;
    lact = lacts = i + (ny + nghy - 1 - mm) * pitch + k * stride;
    lnull = i + (ny + nghy) * pitch + k * stride;
    jgh = ny + nghy + mm;
    jact = ny + nghy - 1 - mm;
  }
  const double of = b.of;
  if (!b.p3d) {
    dens[lgh] = dens[lact] * b.dens[mm];
    vx[lgh] = (vx[lact] + m.ymed[jact] * of) * b.vsq[mm] - m.ymed[jgh] * of;
  } else {
    dens[lgh] = dens[lact] * b.dens[mm] * dexp[mm * m.sz + k];
    vx[lgh] = (vx[lact] + of * m.ymed[jact] * m.sinzmed[k]) * b.vsq[mm] * b.vfac[mm] - m.ymed[jgh] * of * m.sinzmed[k];
    vz[lgh] = vz[lact];
  }
  if (!side || mm < nghy - 1) vy[lghs] = -vy[lacts];
  vy[lnull] = 0;
  if (b.energy) e[lgh] = e[lact];
}

// Z sides (scripts/boundparser.py:341-358; std/boundaries.txt:31-40, 87-91)
__device__ __forceinline__ void bound_z_cell(const dmesh& m, int side, int i, int mm, int j, double* dens, double* vx,
                                             double* vy, double* vz, double* e, const BcPar& b, const double* zdens) {
  const int pitch = m.pitch, stride = m.stride, nz = m.nz, nghz = m.nghz;
  int lgh, lghs, lact, lacts, lnull, kgh, kact;
  if (!side) {
    lgh = lghs = i + j * pitch + mm * stride;
    lact = i + j * pitch + (2 * nghz - mm - 1) * stride;
    lacts = i + j * pitch + (2 * nghz - mm) * stride;
    lnull = i + j * pitch + nghz * stride;
    kgh = mm;
    kact = 2 * nghz - mm - 1;
  } else {
    lgh = i + j * pitch + (nz + nghz + mm) * stride;
    lghs = i + j * pitch + (nz + nghz + 1 + mm) * stride;
    lact = lacts = i + j * pitch + (nz + nghz - 1 - mm) * stride;
    lnull = i + j * pitch + (nz + nghz) * stride;
    kgh = nz + nghz + mm;
    kact = nz + nghz - 1 - mm;
  }
  const double of = b.of;
  dens[lgh] = dens[lact] * zdens[mm * m.sy + j];
  vx[lgh] = (vx[lact] + of * m.ymed[j] * m.sinzmed[kact]) * b.zvfac[mm] - m.ymed[j] * of * m.sinzmed[kgh];
  vy[lgh] = vy[lact];
  if (!side || mm < nghz - 1) vz[lghs] = -vz[lacts];
  vz[lnull] = 0;
  if (b.energy) e[lgh] = e[lact];
}

}  // namespace cell
}  // namespace f3d
