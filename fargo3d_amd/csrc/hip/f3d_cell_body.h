// f3d_cell_body.h -- per-cell arithmetic of the hydro step, shared by the
// one-kernel-per-routine entry points and the fused kernels.
//
// Each function evaluates one cell of one reference routine.  Field values
// are read through "loaders": callables L(di, dj, dk) returning the field at
// the cell offset (di, dj, dk) from the current cell, x periodic.  A global
// loader reads HBM, a tile loader reads LDS, so both kernel families run the
// identical sequence of floating-point operations.  Every expression keeps
// the association of the reference statement it restates (cited per
// function); the build uses -ffp-contract=off.
// Written once for the type `real` of a build (see f3d_device_body.h).

namespace F3D_NS {
namespace cell {

constexpr real CVNR = real(1.41);  // src/define.h:9
constexpr int kMaxPlanets = 8;

// Loader over a dense global array with x wrap (src/define.h:391-392).
template <class T>
struct GLT {
  const T* p;
  int l, i, nx, pitch, stride;
  __device__ __forceinline__ T operator()(int di, int dj, int dk) const {
    int off = di;
    int ii = i + di;
    if (ii < 0) off += nx;
    else if (ii >= nx) off -= nx;
    return p[l + off + dj * pitch + dk * stride];
  }
};
using GL = GLT<real>;

__device__ __forceinline__ GL gl(const real* p, const dmesh& m, int l, int i) {
  return GL{p, l, i, m.nx, m.pitch, m.stride};
}
// the gravitational potential is an FP64 array in every build
__device__ __forceinline__ GLT<double> gld(const double* p, const dmesh& m, int l, int i) {
  return GLT<double>{p, l, i, m.nx, m.pitch, m.stride};
}

// In the FP32 build the terms that cancel in a rotating disk -- the
// centrifugal acceleration against gravity, radially and (3-D) vertically --
// are summed in FP64 from the FP64 potential and the FP64 mesh, and only
// their sum (of the order of the pressure support) is rounded to FP32.
constexpr bool kFix64 = sizeof(real) == 4;

// The FP32 build also stores the azimuthal velocity as its deviation from an
// FP64 profile of the rings, vx = V0(j,k) + vx' (the fields Vx, Vx_temp and
// the damping reference Vx0 hold vx'), and the azimuthal momenta
// rho (vx + r Omega_f) r as their deviation rho vx' r from rho l0(j,k), with
// l0 = (V0 + r Omega_f) r.  A disk in near-Keplerian rotation has |vx'| of
// the order of its pressure support and perturbations, so the FP32 rounding
// of the stored and transported values is that much smaller than on the
// full rotation, whose rounding acts as a torque that never averages out.
// The azimuthal X sweep is invariant under a constant added along a ring
// (its slopes, faces, limiters and fluxes see only differences), so it runs
// on the deviations unchanged; the terms that need the full velocity (the
// centrifugal acceleration, the viscous stress, the ghost formulas, the CFL,
// the ring means) add V0 in FP64.  In the FP64 build these return 0 and no
// code path reads them.
__device__ __forceinline__ double vbase(const dmesh& m, int j, int k) {
  if constexpr (kFix64) return m.vb[j + k * m.sy];
  return 0.0;
}
// the cylindrical radius of ring (j,k) in FP64
template <int G>
__device__ __forceinline__ double rcyl64(const dmesh& m, int j, int k) {
  return (G == SPH) ? m.ymed64[j] * m.sinzmed64[k] : m.ymed64[j];
}
// l0(j,k) = (V0 + r Omega_f) r, the specific angular momentum of the profile
template <int G>
__device__ __forceinline__ double lbase(const dmesh& m, int j, int k, double omegaframe) {
  const double r = rcyl64<G>(m, j, k);
  return (vbase(m, j, k) + r * omegaframe) * r;
}

// ---------------------------------------------------------------- SubStep1
// src/substep1_x.c:112-120
template <int G, class LR, class LP, class LPOT, class LV>
__device__ __forceinline__ real substep1_x(const dmesh& m, int i, int j, int k, real dt, LR rho, LP p,
                                             LPOT pot, LV vx, int dust, int potential) {
  const int im = ixm_of(i, m.nx);
  real dxmed = real(0.5) * (m.sxi[i] + m.sxi[im]);
  real dtOVERrhom = real(2.) * dt / (rho(0, 0, 0) * m.sxi[i] + rho(-1, 0, 0) * m.sxi[im]) * dxmed;
  real v = vx(0, 0, 0);
  const real izs = inv_zone_size_xmed<G>(m, i, j, k);
  if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(-1, 0, 0)) * izs;
  if (potential) v -= (pot(0, 0, 0) - pot(-1, 0, 0)) * dt * izs;
  return v;
}

// src/substep1_y.c:313-316, 352-355 (CYL), 379-383 (SPH), 408-410
template <int G, int DZ, class LR, class LP, class LPOT, class LVX, class LVY, class LVZ>
__device__ __forceinline__ real substep1_y(const dmesh& m, int i, int j, int k, real dt, LR rho, LP p,
                                             LPOT pot, LVX vx, LVY vy, LVZ vz, int dust, int potential,
                                             double omegaframe) {
  real dtOVERrhom = real(2.0) * dt / (rho(0, 0, 0) * (m.ymin[j + 1] - m.ymin[j]) + rho(0, -1, 0) * (m.ymin[j] - m.ymin[j - 1])) *
                      (m.ymed[j] - m.ymed[j - 1]);
  real v = vy(0, 0, 0);
  if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]);
  if constexpr (kFix64) {
    double vphi = 0.25 * ((double)vx(0, 0, 0) + (double)vx(1, 0, 0) + (double)vx(0, -1, 0) + (double)vx(1, -1, 0)) +
                  0.5 * (vbase(m, j, k) + vbase(m, j - 1, k));
    double acc;
    if (G == SPH) {
      vphi += m.ymin64[j] * m.sinzmed64[k] * (double)omegaframe;
      const double vzz = DZ ? 0.25 * ((double)vz(0, 0, 0) + (double)vz(0, 0, 1) + (double)vz(0, -1, 0) +
                                      (double)vz(0, -1, 1))
                            : 0.0;
      acc = (vphi * vphi + vzz * vzz) / m.ymin64[j];
    } else {
      vphi += m.ymin64[j] * (double)omegaframe;
      acc = vphi * vphi / m.ymin64[j];
    }
    if (potential) acc -= ((double)pot(0, 0, 0) - (double)pot(0, -1, 0)) / (m.ymed64[j] - m.ymed64[j - 1]);
    return v + (real)(acc * (double)dt);
  } else {
    real vphi = real(.25) * (vx(0, 0, 0) + vx(1, 0, 0) + vx(0, -1, 0) + vx(1, -1, 0));
    if (G == SPH) {
      vphi += m.ymin[j] * m.sinzmed[k] * omegaframe;
      real vzz = DZ ? real(.25) * (vz(0, 0, 0) + vz(0, 0, 1) + vz(0, -1, 0) + vz(0, -1, 1)) : real(0.0);
      v += (vphi * vphi + vzz * vzz) / m.ymin[j] * dt;
    } else {
      vphi += m.ymin[j] * omegaframe;
      v += vphi * vphi / m.ymin[j] * dt;
    }
    if (potential) v -= (pot(0, 0, 0) - pot(0, -1, 0)) * dt / (m.ymed[j] - m.ymed[j - 1]);
    return v;
  }
}

// src/substep1_z.c:120-183 (CYLINDRICAL and SPHERICAL branches)
template <int G, class LR, class LP, class LPOT, class LVX, class LVZ>
__device__ __forceinline__ real substep1_z(const dmesh& m, int i, int j, int k, real dt, LR rho, LP p,
                                             LPOT pot, LVX vx, LVZ vz, int dust, int potential,
                                             double omegaframe) {
  real dtOVERrhom = real(2.0) * dt / (rho(0, 0, 0) * (m.zmin[k + 1] - m.zmin[k]) + rho(0, 0, -1) * (m.zmin[k] - m.zmin[k - 1])) *
                      (m.zmed[k] - m.zmed[k - 1]);
  real v;
  if (G == SPH && kFix64) {
    double vphi = 0.25 * ((double)vx(0, 0, 0) + (double)vx(1, 0, 0) + (double)vx(0, 0, -1) + (double)vx(1, 0, -1)) +
                  0.5 * (vbase(m, j, k) + vbase(m, j, k - 1));
    const double smin = m.sinzmin64[k], r = m.ymed64[j];
    vphi += r * smin * (double)omegaframe;
    v = vz(0, 0, 0);
    if (!dust) v -= dtOVERrhom * (p(0, 0, 0) - p(0, 0, -1)) / (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1]));
    double acc = vphi * vphi * m.coszmin64[k] / (smin * r);
    if (potential) acc -= ((double)pot(0, 0, 0) - (double)pot(0, 0, -1)) / (r * (m.zmed64[k] - m.zmed64[k - 1]));
    v += (real)(acc * (double)dt);
  } else if (G == SPH) {
    real vphi = real(.25) * (vx(0, 0, 0) + vx(1, 0, 0) + vx(0, 0, -1) + vx(1, 0, -1));
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
__device__ __forceinline__ real art_q(real rho, real dv) {
  if (dv < real(0.0)) return CVNR * CVNR * rho * dv * dv;
  return real(0.0);
}

// src/substep2_b.c:145-190; inputs are the current v_temp (and e) values
template <int G, int DZ, class LR, class LVX, class LVY, class LVZ, class LQX, class LQY, class LQZ>
__device__ __forceinline__ void substep2_b(const dmesh& m, int i, int j, int k, real dt, LR rho, LVX vx, LVY vy,
                                           LVZ vz, LQX px, LQY py, LQZ pz, int adiabatic, real vxt, real vyt,
                                           real vzt, real e, real& oxt, real& oyt, real& ozt,
                                           real& oe) {
  const int im = ixm_of(i, m.nx);
  real dxrho1 = m.sxi[i];
  real dxrho2 = m.sxi[im];
  real dxmed = real(0.5) * (m.sxi[i] + m.sxi[im]);
  vxt += -real(2.0) * dt * dxmed * (px(0, 0, 0) - px(-1, 0, 0)) / (rho(0, 0, 0) * dxrho1 + rho(-1, 0, 0) * dxrho2) *
         inv_zone_size_xmed<G>(m, i, j, k);
  if (adiabatic) e += -dt * (px(0, 0, 0) * (vx(1, 0, 0) - vx(0, 0, 0)) / zone_size_x<G>(m, i, j, k));
  real dyrho1 = m.ymin[j + 1] - m.ymin[j];
  real dyrho2 = m.ymin[j] - m.ymin[j - 1];
  vyt += -real(2.0) * (py(0, 0, 0) - py(0, -1, 0)) / (rho(0, 0, 0) * dyrho1 + rho(0, -1, 0) * dyrho2) * dt;
  if (adiabatic) e += -dt * (py(0, 0, 0) * (vy(0, 1, 0) - vy(0, 0, 0)) / zone_size_y(m, j));
  if (DZ) {
    real dzmed = m.zmed[k] - m.zmed[k - 1];
    real dzrho1 = m.zmin[k + 1] - m.zmin[k];
    real dzrho2 = m.zmin[k] - m.zmin[k - 1];
    if (G == SPH)
      vzt += -real(2.0) * dzmed * (pz(0, 0, 0) - pz(0, 0, -1)) / (rho(0, 0, 0) * dzrho1 + rho(0, 0, -1) * dzrho2) * dt /
             (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1]));
    else
      vzt += -real(2.0) * (pz(0, 0, 0) - pz(0, 0, -1)) / (rho(0, 0, 0) * dzrho1 + rho(0, 0, -1) * dzrho2) * dt;
    if (adiabatic) e += -dt * (pz(0, 0, 0) * (vz(0, 0, 1) - vz(0, 0, 0)) / zone_size_z<G>(m, j, k));
  }
  oxt = vxt;
  oyt = vyt;
  ozt = vzt;
  oe = e;
}

// velocity divergence with the SurfX/SurfY/SurfZ forms (src/substep3.c:98-105)
template <int DZ, class LVX, class LVY, class LVZ>
__device__ __forceinline__ real div_v(const dmesh& m, int i, int j, int k, LVX vx, LVY vy, LVZ vz) {
  real d = real(0.0);
  d += mul_surfx(m, vx(1, 0, 0) - vx(0, 0, 0), j, k);
  d += (mul_surfy(m, vy(0, 1, 0), i, j + 1, k) - mul_surfy(m, vy(0, 0, 0), i, j, k));
  if (DZ) d += (mul_surfz(m, vz(0, 0, 1), i, j, k + 1) - mul_surfz(m, vz(0, 0, 0), i, j, k));
  return d;
}

// src/substep3.c:106-110
template <int DZ, class LVX, class LVY, class LVZ>
__device__ __forceinline__ real substep3(const dmesh& m, int i, int j, int k, real dt, real e, LVX vx,
                                           LVY vy, LVZ vz, real gamma) {
  real dv = div_v<DZ>(m, i, j, k, vx, vy, vz);
  real term = mul_invvol(m, real(0.5) * dt * (gamma - real(1.)) * dv, i, j, k);
  if constexpr (kFix64) {
    // e (1 - t)/(1 + t) = e - e 2t/(1 + t): the change of e formed directly.
    // In FP32 the rounding of 1 - t is the same for every cell of a ring
    // with the same t and would heat or cool the ring by a fixed amount
    // each step; here it is relative to the change only.
    e -= e * (real(2.0) * term / (real(1.0) + term));
    return e;
  }
  e *= (real(1.0) - term) / (real(1.0) + term);
  return e;
}

// ---------------------------------------------------------------- viscosity
struct ViscPar {
  int model;  // 1 NU, 2 ALPHA
  int iso;
  real nu, alpha, gamma;
};

struct Taus {
  real xx, yy, zz, xz, yx, zy;
};

// src/visctensor_cyl.c:140-198 and src/visctensor_sph.c:126-206
template <int G, int DZ, class LR, class LE, class LVX, class LVY, class LVZ>
__device__ __forceinline__ Taus visctensor(const dmesh& m, int i, int j, int k, LR rho, LE energy, LVX vx, LVY vy,
                                           LVZ vz, const ViscPar& vp) {
  real visc, viscm, visczm = real(0.0), visczmym = real(0.0);
  if (vp.model == 2) {
    const real sq_med = sqrt(m.ymed[j] * m.ymed[j] * m.ymed[j] / (real(1.0) * real(1.0)));
    const real sq_min = sqrt(m.ymin[j] * m.ymin[j] * m.ymin[j] / (real(1.0) * real(1.0)));
    if (vp.iso) {
      visc = vp.alpha * energy(0, 0, 0) * energy(0, 0, 0) * sq_med;
      real s = energy(0, 0, 0) + energy(-1, 0, 0) + energy(0, -1, 0) + energy(-1, -1, 0);
      viscm = vp.alpha * real(0.0625) * s * s * sq_min;
      if (DZ) {
        real szm = energy(0, 0, 0) + energy(0, 0, -1) + energy(-1, 0, 0) + energy(-1, 0, -1);
        visczm = vp.alpha * real(0.0625) * szm * szm * sq_med;
        real szy = (G == SPH) ? energy(0, 0, 0) + energy(0, 0, -1) + energy(0, -1, 0) + energy(0, -1, -1)
                                : energy(0, 0, 0) + energy(0, -1, 0) + energy(0, 0, -1) + energy(0, -1, -1);
        visczmym = vp.alpha * real(0.0625) * szy * szy * sq_min;
      }
    } else {
      const real gg = vp.alpha * vp.gamma * (vp.gamma - real(1.0));
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
  real dv = div_v<DZ>(m, i, j, k, vx, vy, vz);
  dv *= mul_invvol(m, real(2.0) / real(3.0), i, j, k);
  const real r0 = rho(0, 0, 0);
  Taus t;
  t.xx = visc * r0 * (real(2.0) * (vx(1, 0, 0) - vx(0, 0, 0)) / zone_size_x<G>(m, i, j, k) - dv);
  t.xx += visc * r0 * (vy(0, 1, 0) + vy(0, 0, 0)) / m.ymed[j];
  if (G == SPH && DZ) t.xx += visc * r0 * (vz(0, 0, 1) + vz(0, 0, 0)) * m.coszmed[k] / (m.sinzmed[k] * m.ymed[j]);
  t.yy = visc * r0 * (real(2.0) * (vy(0, 1, 0) - vy(0, 0, 0)) / (m.ymin[j + 1] - m.ymin[j]) - dv);
  t.zz = t.xz = t.zy = real(0.0);
  if (DZ) {
    if (G == SPH) {
      t.zz = visc * r0 * (real(2.0) * (vz(0, 0, 1) - vz(0, 0, 0)) / (m.ymed[j] * (m.zmin[k + 1] - m.zmin[k])) - dv);
      t.zz += visc * r0 * (vy(0, 0, 0) + vy(0, 1, 0)) / m.ymed[j];
      t.xz = visczm * real(.25) * (r0 + rho(0, 0, -1) + rho(-1, 0, 0) + rho(-1, 0, -1)) *
             ((vx(0, 0, 0) / m.sinzmed[k] - vx(0, 0, -1) / m.sinzmed[k - 1]) * m.sinzmin[k] /
                  (m.ymed[j] * (m.zmed[k] - m.zmed[k - 1])) +
              ((vz(0, 0, 0) - vz(-1, 0, 0)) * m.invdiffxmed[i] / (m.sinzmin[k] * m.ymed[j])));
    } else {
      t.zz = visc * r0 * (real(2.0) * (vz(0, 0, 1) - vz(0, 0, 0)) / (m.zmin[k + 1] - m.zmin[k]) - dv);
      t.xz = visczm * real(.25) * (r0 + rho(0, 0, -1) + rho(-1, 0, 0) + rho(-1, 0, -1)) *
             ((vx(0, 0, 0) - vx(0, 0, -1)) / (m.zmed[k] - m.zmed[k - 1]) +
              (vz(0, 0, 0) - vz(-1, 0, 0)) * m.invdiffxmed[i] / m.ymed[j]);
    }
  }
  if (G == SPH)
    t.yx = viscm * real(.25) * (r0 + rho(-1, 0, 0) + rho(0, -1, 0) + rho(-1, -1, 0)) *
           ((vy(0, 0, 0) - vy(-1, 0, 0)) * m.invdiffxmed[i] / (m.ymin[j] * m.sinzmed[k]) +
            (vx(0, 0, 0) - vx(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - real(.5) * (vx(0, 0, 0) + vx(0, -1, 0)) / m.ymin[j]);
  else
    t.yx = viscm * real(.25) * (r0 + rho(-1, 0, 0) + rho(0, -1, 0) + rho(-1, -1, 0)) *
           ((vy(0, 0, 0) - vy(-1, 0, 0)) * m.invdiffxmed[i] / m.ymin[j] +
            (vx(0, 0, 0) - vx(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - real(.5) * (vx(0, 0, 0) + vx(0, -1, 0)) / m.ymin[j]);
  if (DZ) {
    if (G == SPH)
      t.zy = visczmym * real(.25) * (r0 + rho(0, -1, 0) + rho(0, 0, -1) + rho(0, -1, -1)) *
             ((vz(0, 0, 0) - vz(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) - real(.5) * (vz(0, 0, 0) + vz(0, -1, 0)) / m.ymin[j] +
              (vy(0, 0, 0) - vy(0, 0, -1)) / (m.ymin[j] * (m.zmed[k] - m.zmed[k - 1])));
    else
      t.zy = visczmym * real(.25) * (r0 + rho(0, -1, 0) + rho(0, 0, -1) + rho(0, -1, -1)) *
             ((vz(0, 0, 0) - vz(0, -1, 0)) / (m.ymed[j] - m.ymed[j - 1]) +
              (vy(0, 0, 0) - vy(0, 0, -1)) / (m.zmed[k] - m.zmed[k - 1]));
  }
  return t;
}

// src/addviscosity_cyl.c:97-131 and src/addviscosity_sph.c:97-143
template <int G, int DZ, class LR, class LXX, class LYY, class LZZ, class LXZ, class LYX, class LZY>
__device__ __forceinline__ void addviscosity(const dmesh& m, int i, int j, int k, real dt, LR rho, LXX txx,
                                             LYY tyy, LZZ tzz, LXZ txz, LYX tyx, LZY tzy, real vx, real vy,
                                             real vz, real& ox, real& oy, real& oz) {
  const real r0 = rho(0, 0, 0), rxm = rho(-1, 0, 0), rym = rho(0, -1, 0);
  const real ymn = m.ymin[j], ymp = m.ymin[j + 1], yd = m.ymed[j], ydm = m.ymed[j - 1];
  vx += real(2.0) * (txx(0, 0, 0) - txx(-1, 0, 0)) * inv_zone_size_xmed<G>(m, i, j, k) / ((r0 + rxm)) * dt;
  vx += real(2.0) * (ymp * ymp * tyx(0, 1, 0) - ymn * ymn * tyx(0, 0, 0)) / ((ymp - ymn) * yd * yd * (rxm + r0)) * dt;
  if (G == SPH) {
    vx += (tyx(0, 1, 0) + tyx(0, 0, 0)) / ((rxm + r0) * yd) * dt;
    if (DZ) {
      vx += real(2.0) * (txz(0, 0, 1) - txz(0, 0, 0)) / (zone_size_z<G>(m, j, k) * (rxm + r0)) * dt;
      vx += real(2.0) * (txz(0, 0, 1) + txz(0, 0, 0)) * m.coszmed[k] / (m.sinzmed[k] * yd * (rxm + r0)) * dt;
    }
    vy += real(2.0) * (yd * yd * tyy(0, 0, 0) - ydm * ydm * tyy(0, -1, 0)) / ((yd - ydm) * (r0 + rym) * ymn * ymn) * dt;
    vy += real(2.0) * (tyx(1, 0, 0) - tyx(0, 0, 0)) / ((m.xmin[i + 1] - m.xmin[i]) * ymn * m.sinzmed[k] * (r0 + rym)) * dt;
    vy -= (txx(0, 0, 0) + txx(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
    if (DZ) {
      const real rzm = rho(0, 0, -1);
      vy -= (tzz(0, 0, 0) + tzz(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
      vy += real(2.0) * (tzy(0, 0, 1) * m.sinzmin[k + 1] - tzy(0, 0, 0) * m.sinzmin[k]) /
            ((m.zmin[k + 1] - m.zmin[k]) * ymn * m.sinzmed[k] * (r0 + rym)) * dt;
      vz += real(2.0) * (tzz(0, 0, 0) * m.sinzmed[k] - tzz(0, 0, -1) * m.sinzmed[k - 1]) /
            ((m.zmed[k] - m.zmed[k - 1]) * yd * m.sinzmin[k] * (r0 + rzm)) * dt;
      vz += real(2.0) * (txz(1, 0, 0) - txz(0, 0, 0)) / ((m.sxi[i] * yd * m.sinzmin[k]) * (r0 + rzm)) * dt;
      vz += real(2.0) * (ymp * ymp * ymp * tzy(0, 1, 0) - ymn * ymn * ymn * tzy(0, 0, 0)) /
            ((ymp - ymn) * yd * yd * yd * (r0 + rzm)) * dt;
      vz -= (txx(0, 0, 0) + txx(0, 0, -1)) * m.coszmin[k] / (yd * m.sinzmin[k] * (r0 + rzm)) * dt;
    }
  } else {
    if (DZ) vx += real(2.0) * (txz(0, 0, 1) - txz(0, 0, 0)) / ((m.zmin[k + 1] - m.zmin[k]) * (rxm + r0)) * dt;
    vy += real(2.0) * (yd * tyy(0, 0, 0) - ydm * tyy(0, -1, 0)) / ((yd - ydm) * (r0 + rym) * ymn) * dt;
    vy += real(2.0) * (tyx(1, 0, 0) - tyx(0, 0, 0)) / ((m.xmin[i + 1] - m.xmin[i]) * ymn * (r0 + rym)) * dt;
    vy -= (txx(0, 0, 0) + txx(0, -1, 0)) / (ymn * (r0 + rym)) * dt;
    if (DZ) {
      const real rzm = rho(0, 0, -1);
      vy += real(2.0) * (tzy(0, 0, 1) - tzy(0, 0, 0)) / ((m.zmin[k + 1] - m.zmin[k]) * (r0 + rym)) * dt;
      vz += real(2.0) * (tzz(0, 0, 0) - tzz(0, 0, -1)) / ((m.zmed[k] - m.zmed[k - 1]) * (r0 + rzm)) * dt;
      vz += real(2.0) * (txz(1, 0, 0) - txz(0, 0, 0)) / (zone_size_x<G>(m, i, j, k) * (r0 + rzm)) * dt;
      vz += real(2.0) * (ymp * tzy(0, 1, 0) - ymn * tzy(0, 0, 0)) / ((ymp - ymn) * yd * (r0 + rzm)) * dt;
    }
  }
  ox = vx;
  oy = vy;
  oz = vz;
}

// ---------------------------------------------------------------- potential
struct PlanetSet {
  int n, indirect;
  real x[kMaxPlanets], y[kMaxPlanets], z[kMaxPlanets], mp[kMaxPlanets], smooth2[kMaxPlanets];
};

// src/potential.c:100-196 (default star, planets, planet indirect term);
// planet mass already tapered, smoothing^2 evaluated on the host.
template <int G>
__device__ __forceinline__ real potential(const dmesh& m, int i, int j, int k, const PlanetSet& ps) {
  real XC, YC, ZC, pot;
  if (G == SPH) {
    XC = m.ymed[j] * m.cosxmed[i] * m.sinzmed[k];
    YC = m.ymed[j] * m.sinxmed[i] * m.sinzmed[k];
    ZC = m.ymed[j] * m.coszmed[k];
    pot = -real(1.0) * real(1.0) / m.ymed[j];
  } else {
    XC = m.ymed[j] * m.cosxmed[i];
    YC = m.ymed[j] * m.sinxmed[i];
    ZC = m.zmed[k];
    pot = -real(1.0) * real(1.0) / sqrt(m.ymed[j] * m.ymed[j] + ZC * ZC);
  }
  for (int n = 0; n < ps.n; n++) {
    const real mp = ps.mp[n];
    const real xp = ps.x[n], yp = ps.y[n], zp = ps.z[n];
    real planetdistance = sqrt(xp * xp + yp * yp + zp * zp);
    real dist = ((XC - xp) * (XC - xp) + (YC - yp) * (YC - yp) + (ZC - zp) * (ZC - zp));
    if (ps.indirect)
      pot += real(1.0) * mp * (XC * xp + YC * yp + ZC * zp) / (planetdistance * planetdistance * planetdistance);
    pot += -real(1.0) * mp / sqrt(dist + ps.smooth2[n]);
  }
  return pot;
}

// The same for C cells i0 + u * di of one ring (j, k), what depends on the
// ring alone (the star term, each planet's distance and the indirect term's
// divisor, shared through rdiv()) evaluated once: every cell's value is
// potential()'s, bit for bit.  Cells with i >= nx are skipped.
template <int G, int C>
__device__ __forceinline__ void potential_cells(const dmesh& m, int i0, int di, int j, int k, const PlanetSet& ps,
                                                real* pot, int l0) {
  real ZC, pot0;
  if (G == SPH) {
    ZC = m.ymed[j] * m.coszmed[k];
    pot0 = -real(1.0) * real(1.0) / m.ymed[j];
  } else {
    ZC = m.zmed[k];
    pot0 = -real(1.0) * real(1.0) / sqrt(m.ymed[j] * m.ymed[j] + ZC * ZC);
  }
  real acc[C], XC[C], YC[C];
#pragma unroll
  for (int u = 0; u < C; u++) {
    const int i = min(i0 + u * di, m.nx - 1);
    acc[u] = pot0;
    if (G == SPH) {
      XC[u] = m.ymed[j] * m.cosxmed[i] * m.sinzmed[k];
      YC[u] = m.ymed[j] * m.sinxmed[i] * m.sinzmed[k];
    } else {
      XC[u] = m.ymed[j] * m.cosxmed[i];
      YC[u] = m.ymed[j] * m.sinxmed[i];
    }
  }
  for (int n = 0; n < ps.n; n++) {
    const real mp = ps.mp[n];
    const real xp = ps.x[n], yp = ps.y[n], zp = ps.z[n];
    const real planetdistance = sqrt(xp * xp + yp * yp + zp * zp);
    const Recip D = recip(planetdistance * planetdistance * planetdistance);
#pragma unroll
    for (int u = 0; u < C; u++) {
      const real dist = ((XC[u] - xp) * (XC[u] - xp) + (YC[u] - yp) * (YC[u] - yp) + (ZC - zp) * (ZC - zp));
      if (ps.indirect) acc[u] += rdiv(real(1.0) * mp * (XC[u] * xp + YC[u] * yp + ZC * zp), D);
      acc[u] += -real(1.0) * mp / sqrt(dist + ps.smooth2[n]);
    }
  }
#pragma unroll
  for (int u = 0; u < C; u++)
    if (i0 + u * di < m.nx) pot[l0 + u * di] = acc[u];
}

// ---------------------------------------------------------------- momenta / new velocities
// src/momenta_x.c:76-84
template <int G>
__device__ __forceinline__ void momenta_x(const dmesh& m, int i, int j, int k, real rho, real vx, real vxp,
                                          real omegaframe, real& mmx, real& mpx) {
  if constexpr (kFix64) {
    // deviations from rho l0: rho vx' r (vx, vxp are deviations), with the
    // FP64 radius
    const double rcyl = rcyl64<G>(m, j, k);
    mmx = (real)((double)vx * rcyl * (double)rho);
    mpx = (real)((double)vxp * rcyl * (double)rho);
    return;
  }
  if (G == SPH) {
    real rcyl = m.ymed[j] * m.sinzmed[k];
    mmx = (vx + rcyl * omegaframe) * rcyl * rho;
    mpx = (vxp + rcyl * omegaframe) * rcyl * rho;
  } else {
    mmx = (vx + m.ymed[j] * omegaframe) * m.ymed[j] * rho;
    mpx = (vxp + m.ymed[j] * omegaframe) * m.ymed[j] * rho;
  }
}

// src/newvel_x.c:78-86
template <int G>
__device__ __forceinline__ real newvel_x(const dmesh& m, int i, int j, int k, real rho, real rhoxm, real mmx,
                                           real mpxm, real omegaframe) {
  const int im = ixm_of(i, m.nx);
  if constexpr (kFix64) {
    // from the momenta's deviations: l0/r - Omega_f r = V0 drops out, and the
    // result is the deviation vx'
    const double r = rcyl64<G>(m, j, k);
    return (real)(((double)mul_vol(m, mmx, i, j, k) + (double)mul_vol(m, mpxm, im, j, k)) /
                  (((double)mul_vol(m, rho, i, j, k) + (double)mul_vol(m, rhoxm, im, j, k)) * r));
  }
  real r = (G == SPH) ? m.ymed[j] * m.sinzmed[k] : m.ymed[j];
  return (mul_vol(m, mmx, i, j, k) + mul_vol(m, mpxm, im, j, k)) /
             ((mul_vol(m, rho, i, j, k) + mul_vol(m, rhoxm, im, j, k)) * r) -
         omegaframe * r;
}

// src/newvel_y.c:67
__device__ __forceinline__ real newvel_y(const dmesh& m, int i, int j, int k, real rho, real rhoym, real mmy,
                                           real mpym) {
  return (mul_vol(m, mmy, i, j, k) + mul_vol(m, mpym, i, j - 1, k)) /
         (mul_vol(m, rho, i, j, k) + mul_vol(m, rhoym, i, j - 1, k));
}

// ---------------------------------------------------------------- PPA (src/fargo_ppa_*.c)
// monotonised-centred slope (src/fargo_ppa_a.c:58-69)
__device__ __forceinline__ real ppa_slope(real qm, real q, real qp) {
  real dqm = (q - qm);
  real dqp = (qp - q);
  if (dqp * dqm <= real(0.0)) return real(0.0);
  real s = real(0.5) * (qp - qm);
  real work = fabs(s);
  if (real(2.0) * fabs(dqm) < work) work = real(2.0) * fabs(dqm);
  if (real(2.0) * fabs(dqp) < work) work = real(2.0) * fabs(dqp);
  return (s < 0) ? -work : work;
}
// interface value q_{i+1/2} (src/fargo_ppa_b.c:52-55)
__device__ __forceinline__ real ppa_face(real q, real qp, real s, real sp) {
  return q + real(0.5) * (qp - q) - real(1.0) / real(6.0) * (sp - s);
}
// Colella-Woodward monotonicity (src/fargo_ppa_c.c:52-63)
__device__ __forceinline__ void ppa_limit(real q, real& qL, real& qR) {
  if ((qR - q) * (q - qL) < real(0.0)) {
    qL = q;
    qR = q;
  }
  real diff = qR - qL;
  real cord = q - real(0.5) * (qL + qR);
  if (real(6.0) * diff * cord > diff * diff) qL = real(3.0) * q - real(2.0) * qR;
  if (-diff * diff > real(6.0) * diff * cord) qR = real(3.0) * q - real(2.0) * qL;
}
// the swept fraction ksi of ppa_flux_value, the same for every quantity of a cell
__device__ __forceinline__ real ppa_ksi(real v, real dt, real zsx) {
  return (v > real(0.0)) ? v * dt / zsx : -v * dt / zsx;
}
__device__ __forceinline__ real ppa_flux_ksi(real v, real ksi, real qm, real qLm, real qRm, real q, real qL, real qR) {
  real qs;
  if (v > real(0.0)) {
    qs = qRm + ksi * (qm - qRm);
    qs += ksi * (real(1.0) - ksi) * (real(2.0) * qm - qRm - qLm);
  } else {
    qs = qL + ksi * (q - qL);
    qs += ksi * (real(1.0) - ksi) * (real(2.0) * q - qR - qL);
  }
  return qs;
}
// parabolic average over the swept zone (src/fargo_ppa_d.c:62-70)
__device__ __forceinline__ real ppa_flux_value(real v, real dt, real zsx, real qm, real qLm, real qRm,
                                                 real q, real qL, real qR) {
  real qs;
  if (v > real(0.0)) {
    real ksi = v * dt / zsx;
    qs = qRm + ksi * (qm - qRm);
    qs += ksi * (real(1.0) - ksi) * (real(2.0) * qm - qRm - qLm);
  } else {
    real ksi = -v * dt / zsx;
    qs = qL + ksi * (q - qL);
    qs += ksi * (real(1.0) - ksi) * (real(2.0) * q - qR - qL);
  }
  return qs;
}

// ---------------------------------------------------------------- conservative updates
// src/update_x.c:69-71: ((a - b)*SurfX*dt*InvVol)
__device__ __forceinline__ real flux_x(const dmesh& m, int i, int j, int k, real dt, real a, real b) {
  return mul_invvol(m, mul_surfx(m, a - b, j, k) * dt, i, j, k);
}
// the same with the cell's InvVol divisor shared by its quantities
__device__ __forceinline__ real flux_x(const dmesh& m, int j, int k, real dt, real a, real b, const Recip& V) {
  return mul_invvol(m, mul_surfx(m, a - b, j, k) * dt, j, V);
}
// src/update_y.c:68-70: (a*SurfY(j) - b*SurfY(j+1))*dt*InvVol
__device__ __forceinline__ real flux_y(const dmesh& m, int i, int j, int k, real dt, real a, real b) {
  return mul_invvol(m, (mul_surfy(m, a, i, j, k) - mul_surfy(m, b, i, j + 1, k)) * dt, i, j, k);
}
__device__ __forceinline__ real flux_z(const dmesh& m, int i, int j, int k, real dt, real a, real b) {
  return mul_invvol(m, (mul_surfz(m, a, i, j, k) - mul_surfz(m, b, i, j, k + 1)) * dt, i, j, k);
}
// The mesh factors of one cell's Y or Z flux difference, loaded once for
// all its quantities: (a SurfA - b SurfB) dt InvVol with Surf = (. * fj) * fk
// * Sxi[i] on either face, in the association of mul_surfy/mul_surfz and
// mul_invvol.
struct FluxMesh {
  real aj, ak, bj, bk, si, iv;
};
__device__ __forceinline__ FluxMesh flux_y_mesh(const dmesh& m, int i, int j, int k) {
  return FluxMesh{m.syj[j], m.syk[k], m.syj[j + 1], m.syk[k], m.sxi[i], m.invvj[j]};
}
__device__ __forceinline__ FluxMesh flux_z_mesh(const dmesh& m, int i, int j, int k) {
  return FluxMesh{m.szj[j], m.szk[k], m.szj[j], m.szk[k + 1], m.sxi[i], m.invvj[j]};
}
// flux_y / flux_z below from those factors, with the InvVol divisor V
__device__ __forceinline__ real flux_yz(const FluxMesh& F, real dt, real a, real b, const Recip& V) {
  return rdiv(((((a * F.aj) * F.ak) * F.si - ((b * F.bj) * F.bk) * F.si) * dt) * F.iv, V);
}
// the same with the cell's InvVol divisor shared by its quantities
__device__ __forceinline__ real flux_y(const dmesh& m, int i, int j, int k, real dt, real a, real b, const Recip& V) {
  return mul_invvol(m, (mul_surfy(m, a, i, j, k) - mul_surfy(m, b, i, j + 1, k)) * dt, j, V);
}
__device__ __forceinline__ real flux_z(const dmesh& m, int i, int j, int k, real dt, real a, real b, const Recip& V) {
  return mul_invvol(m, (mul_surfz(m, a, i, j, k) - mul_surfz(m, b, i, j, k + 1)) * dt, j, V);
}

// ---------------------------------------------------------------- cfl (src/cfl.c:174-253)
struct CflPar {
  double cfl, nu, alpha, gamma;
  int iso, alpha_visc, fargo;
};

// Sum of the squared inverse time scales of one cell (src/cfl.c:185-253),
// evaluated in FP64 in every build.  The reference's maxima over directions
// of c/zone_size are formed as c/min(zone_size): IEEE division is monotone
// in the divisor and rounding is monotone, so max_k fl(c/z_k) = fl(c/min z_k)
// exactly (c >= 0, z_k > 0), with one division instead of three.
// FX fixes the flags at compile time (1 isothermal, constant nu; 2
// isothermal, alpha; 3 adiabatic, constant nu; FARGO in all three; 0: as
// cp holds them), so that no branch separates the cells' loads.
template <int G, int DZ, int FX = 0>
__device__ __forceinline__ double cfl_sumsq(const dmesh& m, int i, int j, int k, int l, const real* cs,
                                            const real* rho, const real* vx, const real* vy, const real* vz,
                                            double vxmed, CflPar cp) {
  if constexpr (FX != 0) {
    cp.fargo = 1;
    cp.iso = FX != 3;
    cp.alpha_visc = FX == 2;
  }
  const int lxp = lxp_of(l, i, m.nx), lyp = l + m.pitch, lzp = l + m.stride;
  double vx0 = vx[l], vx1 = vx[lxp];
  if constexpr (kFix64) {
    vx0 += vbase(m, j, k);
    vx1 += vbase(m, j, k);
  }
  double vxx, vxxp;
  if (!cp.fargo) {
    vxx = vx0;
    vxxp = vx1;
  } else {
    vxx = vx0 - vxmed;
    vxxp = vx1 - vxmed;
  }
  const double c0 = cs[l];
  double soundspeed2, viscosity;
  if (cp.iso) soundspeed2 = c0 * c0;
  else soundspeed2 = cp.gamma * (cp.gamma - 1) * c0 / (double)rho[l];
  if (cp.alpha_visc) {
    const double y = m.ymed64[j];
    const double sq = sqrt(y * y * y / (1.0 * 1.0));
    if (cp.iso) viscosity = cp.alpha * c0 * c0 * sq;
    else viscosity = cp.alpha * cp.gamma * (cp.gamma - 1.0) * c0 / (double)rho[l] * sq;
  } else {
    viscosity = cp.nu;
  }
  const double soundspeed = sqrt(soundspeed2);
  // every quotient below divides by one of the three zone sizes: one
  // reciprocal each, shared (RecipD), with zmin's the one of the smallest
  const double zx = zone_size_x64<G>(m, i, j, k), zy = zone_size_y64(m, j);
  const double zz = DZ ? zone_size_z64<G>(m, j, k) : 1.0;
  const RecipD Zx = recipd(zx), Zy = recipd(zy);
  const RecipD Zz = DZ ? recipd(zz) : RecipD{1.0, 1.0};
  RecipD Zmin = zx < zy ? Zx : Zy;
  if (DZ && zz < Zmin.d) Zmin = Zz;
  const double cfl1 = rdivd(soundspeed, Zmin);
  const double ax = fabs(vxx), axp = fabs(vxxp);
  const double cfl2 = rdivd(ax > axp ? ax : axp, Zx);
  const double vy0 = vy[l], vy1 = vy[lyp];
  const double ay = fabs(vy0), ayp = fabs(vy1);
  const double cfl3 = rdivd(ay > ayp ? ay : ayp, Zy);
  double cfl4 = 0.0, cfl5_c = 0.0;
  if (DZ) {
    const double vz0 = vz[l], vz1 = vz[lzp];
    const double az = fabs(vz0), azp = fabs(vz1);
    cfl4 = rdivd(az > azp ? az : azp, Zz);
    cfl5_c = rdivd(fabs(vz1 - vz0), Zz);
  }
  const double cfl5_a = rdivd(fabs(vx1 - vx0), Zx);
  const double cfl5_b = rdivd(fabs(vy1 - vy0), Zy);
  const double t5 = cfl5_a > cfl5_b ? cfl5_a : cfl5_b;
  const double cfl5 = (t5 > cfl5_c ? t5 : cfl5_c) * 4.0 * 1.41;  // 4 CVNR
  const double m7 = Zmin.r;  // 1.0 / zmin
  const double cfl7 = 4.0 * viscosity * (m7 * m7);
  // the reference sums ten squares; the MHD/strong-shock terms (cfl6, cfl8-10)
  // are exact zeros here and adding +0.0 does not change the sum
  return cfl1 * cfl1 + cfl2 * cfl2 + cfl3 * cfl3 + cfl4 * cfl4 + cfl5 * cfl5 + 0.0 + cfl7 * cfl7;
}

// The cell's time step CFL/sqrt(sum) (the field src/cfl.c:253 writes).
template <int G, int DZ>
__device__ __forceinline__ double cfl_cell(const dmesh& m, int i, int j, int k, int l, const real* cs,
                                           const real* rho, const real* vx, const real* vy, const real* vz,
                                           double vxmed, const CflPar& cp) {
  return cp.cfl / sqrt(cfl_sumsq<G, DZ>(m, i, j, k, l, cs, rho, vx, vy, vz, vxmed, cp));
}

// ---------------------------------------------------------------- Stockholm damping
struct DampPar {
  real Y_inf, Y_sup, Z_inf, Z_sup, y_min, y_max, z_min, z_max, ds;
  double of, of0;  // the frame rate and its initial value, FP64 in every build
  int periodic_z;
};

// src/stockholm.c:129-184: the ramps of ring (j,k) (ramp for rho, vx, vy;
// rampi for vz) and the damping time, then the relaxation of the values
struct Ramps {
  real ramp, rampi, tau;
};
template <int DZ>
__device__ __forceinline__ Ramps stockholm_ramps(const dmesh& m, int j, int k, const DampPar& d) {
  const real ym = m.ymed[j];
  real rampy = real(0.0), rampz = real(0.0), rampzz = real(0.0);
  if (ym > d.Y_sup) rampy = (ym - d.Y_sup) / (d.y_max - d.Y_sup);
  if (ym < d.Y_inf) rampy = (d.Y_inf - ym) / (d.Y_inf - d.y_min);
  rampy *= rampy;
  if (DZ) {
    const real zm = m.zmed[k], zn = m.zmin[k];
    if (zm > d.Z_sup) rampz = (zm - d.Z_sup) / (d.z_max - d.Z_sup);
    if (zm < d.Z_inf) rampz = (d.Z_inf - zm) / (d.Z_inf - d.z_min);
    rampz = rampz * rampz;
    if (zn > d.Z_sup) rampzz = (zn - d.Z_sup) / (d.z_max - d.Z_sup);
    if (zn < d.Z_inf) rampzz = (d.Z_inf - zn) / (d.Z_inf - d.z_min);
    rampzz = rampzz * rampzz;
  }
  if (d.periodic_z) {
    rampz = real(0.0);
    rampzz = real(0.0);
  }
  return Ramps{rampy + rampz, rampy + rampzz, d.ds * sqrt(ym * ym * ym / real(1.0) / real(1.0))};
}
// rho, vx, vy of a cell where R.ramp > 0
template <int G>
__device__ __forceinline__ void stockholm_rxy(const dmesh& m, int j, int k, real dt, const Ramps& R, real& rho,
                                              real& vx, real& vy, real rho0, real vx0, real vy0, const DampPar& d) {
  real taud = R.tau / R.ramp;
  rho = (rho * taud + rho0 * dt) / (dt + taud);
  real vx0_target = vx0;
  real radius = m.ymed[j];
  if (G == SPH) radius *= m.sinzmed[k];
  vx0_target -= (real)((d.of - d.of0) * radius);
  vx = (vx * taud + vx0_target * dt) / (dt + taud);
  vy = (vy * taud + vy0 * dt) / (dt + taud);
}
// vz of a cell where R.rampi > 0
__device__ __forceinline__ real stockholm_z(real dt, const Ramps& R, real vz, real vz0) {
  real taud = R.tau / R.rampi;
  return (vz * taud + vz0 * dt) / (dt + taud);
}
// the same on the arrays (cell l of ring (j,k))
template <int G, int DZ>
__device__ __forceinline__ void stockholm(const dmesh& m, int j, int k, int l, real dt, real* rho, real* vx,
                                          real* vy, real* vz, const real* rho0, const real* vx0,
                                          const real* vy0, const real* vz0, const DampPar& d) {
  const Ramps R = stockholm_ramps<DZ>(m, j, k, d);
  const int l2 = j + k * m.sy;
  if (R.ramp > real(0.0)) {
    real r = rho[l], x = vx[l], y = vy[l];
    stockholm_rxy<G>(m, j, k, dt, R, r, x, y, rho0[l2], vx0[l2], vy0[l2], d);
    rho[l] = r;
    vx[l] = x;
    vy[l] = y;
  }
  if (DZ) {
    if (R.rampi > real(0.0)) vz[l] = stockholm_z(dt, R, vz[l], vz0[l2]);
  }
}

// ---------------------------------------------------------------- boundaries
struct BcPar {
  int p3d;
  int energy;  // Energy ghosts SYMMETRIC (centered |a|a|, std/boundaries.txt:55-57)
  double of;   // the frame rate, FP64 in every build
  real dens[3], vsq[3], vfac[3], zvfac[3];
};

// Y sides (scripts/boundparser.py:323-340 indices; std/boundaries.txt:35-40,
// 62-66, 81-85 formulas; field lists setups/fargo/fargo.bound.0,
// setups/p3disof/p3disof.bound.0)
// One ghost cell of one side: its indices and the values it copies.  The
// loads are separate from the stores so that a kernel can issue the loads of
// every layer before any store (a thread's stores never touch what it reads).
struct GhostCell {
  int lgh, lghs, lnull, gh, act;  // ghost, staggered ghost, null face; ghost/active row or plane
  bool ws;                        // the staggered ghost exists on this layer
  real d0, x0, y0, z0, e0;
};

__device__ __forceinline__ GhostCell ghost_y_load(const dmesh& m, int side, int i, int mm, int k, const real* dens,
                                                  const real* vx, const real* vy, const real* vz, const real* e,
                                                  const BcPar& b) {
  const int pitch = m.pitch, stride = m.stride, ny = m.ny, nghy = m.nghy;
  GhostCell c;
  int lact, lacts;
  if (!side) {
    c.lgh = c.lghs = i + mm * pitch + k * stride;
    lact = i + (2 * nghy - mm - 1) * pitch + k * stride;
    lacts = i + (2 * nghy - mm) * pitch + k * stride;
    c.lnull = i + nghy * pitch + k * stride;
    c.gh = mm;
    c.act = 2 * nghy - mm - 1;
  } else {
    c.lgh = i + (ny + nghy + mm) * pitch + k * stride;
    c.lghs = i + (ny + nghy + 1 + mm) * pitch + k * stride;
    lact = lacts = i + (ny + nghy - 1 - mm) * pitch + k * stride;
    c.lnull = i + (ny + nghy) * pitch + k * stride;
    c.gh = ny + nghy + mm;
    c.act = ny + nghy - 1 - mm;
  }
  c.ws = !side || mm < nghy - 1;
  c.d0 = dens[lact];
  c.x0 = vx[lact];
  c.z0 = b.p3d ? vz[lact] : real(0.0);
  c.y0 = c.ws ? vy[lacts] : real(0.0);
  c.e0 = b.energy ? e[lact] : real(0.0);
  return c;
}

// the density and azimuthal velocity of Y ghost layer mm (row gh, plane k)
// from d0 and x0, the stored values of its active row act
__device__ __forceinline__ void ghost_y_dx(const dmesh& m, const BcPar& b, int mm, int k, int act, int gh, real d0,
                                           real x0, const real* dexp, real& dg, real& xg) {
  if constexpr (kFix64) {
    // the full velocities in FP64, the ghost's deviation from its own V0
    const double of = b.of, xa = (double)x0 + vbase(m, act, k);
    double x;
    if (!b.p3d) {
      dg = d0 * b.dens[mm];
      x = (xa + m.ymed64[act] * of) * (double)b.vsq[mm] - m.ymed64[gh] * of;
    } else {
      dg = d0 * b.dens[mm] * dexp[mm * m.sz + k];
      x = (xa + of * m.ymed64[act] * m.sinzmed64[k]) * (double)b.vsq[mm] * (double)b.vfac[mm] -
          m.ymed64[gh] * of * m.sinzmed64[k];
    }
    xg = (real)(x - vbase(m, gh, k));
  } else {
    const real of = b.of;
    if (!b.p3d) {
      dg = d0 * b.dens[mm];
      xg = (x0 + m.ymed[act] * of) * b.vsq[mm] - m.ymed[gh] * of;
    } else {
      dg = d0 * b.dens[mm] * dexp[mm * m.sz + k];
      xg = (x0 + of * m.ymed[act] * m.sinzmed[k]) * b.vsq[mm] * b.vfac[mm] - m.ymed[gh] * of * m.sinzmed[k];
    }
  }
}

__device__ __forceinline__ void ghost_y_store(const dmesh& m, int mm, int k, const GhostCell& c, real* dens, real* vx,
                                              real* vy, real* vz, real* e, const BcPar& b, const real* dexp) {
  real dg, xg;
  ghost_y_dx(m, b, mm, k, c.act, c.gh, c.d0, c.x0, dexp, dg, xg);
  dens[c.lgh] = dg;
  vx[c.lgh] = xg;
  if (b.p3d) vz[c.lgh] = c.z0;
  if (c.ws) vy[c.lghs] = -c.y0;
  vy[c.lnull] = 0;
  if (b.energy) e[c.lgh] = c.e0;
}

__device__ __forceinline__ void bound_y_cell(const dmesh& m, int side, int i, int mm, int k, real* dens, real* vx,
                                             real* vy, real* vz, real* e, const BcPar& b, const real* dexp) {
  ghost_y_store(m, mm, k, ghost_y_load(m, side, i, mm, k, dens, vx, vy, vz, e, b), dens, vx, vy, vz, e, b, dexp);
}

// Z sides (scripts/boundparser.py:341-358; std/boundaries.txt:31-40, 87-91)
__device__ __forceinline__ GhostCell ghost_z_load(const dmesh& m, int side, int i, int mm, int j, const real* dens,
                                                  const real* vx, const real* vy, const real* vz, const real* e,
                                                  const BcPar& b) {
  const int pitch = m.pitch, stride = m.stride, nz = m.nz, nghz = m.nghz;
  GhostCell c;
  int lact, lacts;
  if (!side) {
    c.lgh = c.lghs = i + j * pitch + mm * stride;
    lact = i + j * pitch + (2 * nghz - mm - 1) * stride;
    lacts = i + j * pitch + (2 * nghz - mm) * stride;
    c.lnull = i + j * pitch + nghz * stride;
    c.gh = mm;
    c.act = 2 * nghz - mm - 1;
  } else {
    c.lgh = i + j * pitch + (nz + nghz + mm) * stride;
    c.lghs = i + j * pitch + (nz + nghz + 1 + mm) * stride;
    lact = lacts = i + j * pitch + (nz + nghz - 1 - mm) * stride;
    c.lnull = i + j * pitch + (nz + nghz) * stride;
    c.gh = nz + nghz + mm;
    c.act = nz + nghz - 1 - mm;
  }
  c.ws = !side || mm < nghz - 1;
  c.d0 = dens[lact];
  c.x0 = vx[lact];
  c.y0 = vy[lact];
  c.z0 = c.ws ? vz[lacts] : real(0.0);
  c.e0 = b.energy ? e[lact] : real(0.0);
  return c;
}

// the azimuthal velocity of Z ghost layer mm (plane gh) from x0, the stored
// velocity of its active plane act
__device__ __forceinline__ real ghost_z_vx(const dmesh& m, const BcPar& b, int mm, int j, int act, int gh, real x0) {
  if constexpr (kFix64) {
    const double of = b.of, xa = (double)x0 + vbase(m, j, act);
    return (real)((xa + of * m.ymed64[j] * m.sinzmed64[act]) * (double)b.zvfac[mm] -
                  m.ymed64[j] * of * m.sinzmed64[gh] - vbase(m, j, gh));
  } else {
    const real of = b.of;
    return (x0 + of * m.ymed[j] * m.sinzmed[act]) * b.zvfac[mm] - m.ymed[j] * of * m.sinzmed[gh];
  }
}

__device__ __forceinline__ void ghost_z_store(const dmesh& m, int mm, int j, const GhostCell& c, real* dens, real* vx,
                                              real* vy, real* vz, real* e, const BcPar& b, const real* zdens) {
  const real dg = c.d0 * zdens[mm * m.sy + j];
  const real xg = ghost_z_vx(m, b, mm, j, c.act, c.gh, c.x0);
  dens[c.lgh] = dg;
  vx[c.lgh] = xg;
  vy[c.lgh] = c.y0;
  if (c.ws) vz[c.lghs] = -c.z0;
  vz[c.lnull] = 0;
  if (b.energy) e[c.lgh] = c.e0;
}

__device__ __forceinline__ void bound_z_cell(const dmesh& m, int side, int i, int mm, int j, real* dens, real* vx,
                                             real* vy, real* vz, real* e, const BcPar& b, const real* zdens) {
  ghost_z_store(m, mm, j, ghost_z_load(m, side, i, mm, j, dens, vx, vy, vz, e, b), dens, vx, vy, vz, e, b, zdens);
}

}  // namespace cell
}  // namespace F3D_NS
