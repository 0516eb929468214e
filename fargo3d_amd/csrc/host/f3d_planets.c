/*
 * f3d_planets.c -- planetary system on the host.
 *
 * Planet motion stays on the host, as in FARGO3D: the RK5 N-body step
 * (src/runge_kutta.c:10-192), the guiding centre of the first planet, the
 * frame-tracking helpers GetPsysInfo / RotatePsys (src/psys.c:362-440) and
 * the planet set-up (src/psys.c:140-330).  The arithmetic follows those files
 * statement by statement so that a run is reproducible bit for bit.
 */
#include "f3d_host.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#define G F3D_G
#define MSTAR F3D_MSTAR
#define R0 F3D_R0

void f3d_psys_add(f3d_psys *sys, const f3d_params *p, double dist, double mass,
                  double accret, int feeldisk, int feelothers) {
  /* src/psys.c:178-214 */
  int i = sys->nb;
  if (i >= F3D_MAX_PLANETS) return;
  if ((p->semimajoraxis > 0.0) && (i == 0)) dist = p->semimajoraxis;
  if (p->orbitalradius > 1e-30) dist *= p->orbitalradius;
  sys->mass[i] = mass;
  if (p->planetmass > 1e-18) sys->mass[0] = p->planetmass;
  sys->x[i] = (double)dist * (1.0 + p->eccentricity);
  sys->y[i] = 0.0;
  sys->z[i] = 0.0;
  sys->vy[i] = (double)sqrt(G * (MSTAR + sys->mass[i]) / dist) *
               sqrt((1.0 - p->eccentricity) / (1.0 + p->eccentricity)) * cos(p->inclination);
  sys->vx[i] = -0.0000000000 * sys->vy[i];
  sys->vz[i] = sys->vy[i] * sin(p->inclination) / cos(p->inclination);
  sys->acc[i] = accret;
  sys->feeldisk[i] = feeldisk;
  sys->feelothers[i] = feelothers;
  sys->nb = i + 1;
}

int f3d_psys_init(f3d_psys *sys, const f3d_params *p, char *err, size_t errlen) {
  memset(sys, 0, sizeof(*sys));
  if (strcmp(p->planetconfig, "NONE") == 0) return 0;
  FILE *in = fopen(p->planetconfig, "r");
  if (!in) {
    if (err) snprintf(err, errlen, "%s cannot be opened", p->planetconfig);
    return -1;
  }
  char s[512], nm[512], t1[512], t2[512];
  while (fgets(s, 510, in)) {
    if (sscanf(s, "%511s ", nm) != 1) continue;
    if (!isalpha((unsigned char)s[0])) continue;
    char *s1 = s + strlen(nm);
    double dist = 0, mass = 0, accret = 0;
    t1[0] = t2[0] = 'y';
    t1[1] = t2[1] = 0;
    sscanf(s1 + strspn(s1, "\t :=>_"), "%lf %lf %lf %511s %511s", &dist, &mass, &accret, t1, t2);
    f3d_psys_add(sys, p, dist, mass, accret, tolower((unsigned char)t1[0]) != 'n',
                 tolower((unsigned char)t2[0]) != 'n');
  }
  fclose(in);
  return 0;
}

/* ---- RK5 (src/runge_kutta.c) ---- */
static void deriv_rk5(const double *q, const double *masses, double *deriv, int n, double dt,
                      const int *feelothers, int indirect) {
  const double *x = q, *y = q + n, *z = q + 2 * n, *vx = q + 3 * n, *vy = q + 4 * n, *vz = q + 5 * n;
  double *dx = deriv, *dy = deriv + n, *dz = deriv + 2 * n;
  double *dvx = deriv + 3 * n, *dvy = deriv + 4 * n, *dvz = deriv + 5 * n;
  double Dist[F3D_MAX_PLANETS];
  for (int i = 0; i < n; i++) Dist[i] = sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
  for (int i = 0; i < n; i++) {
    dx[i] = vx[i];
    dy[i] = vy[i];
    dz[i] = vz[i];
    double coef = -G * MSTAR / Dist[i] / Dist[i] / Dist[i];
    dvx[i] = coef * x[i];
    dvy[i] = coef * y[i];
    dvz[i] = coef * z[i];
    for (int j = 0; j < n; j++) {
      if (indirect) {
        coef = G * masses[j] / Dist[j] / Dist[j] / Dist[j];
        dvx[i] -= coef * x[j];
        dvy[i] -= coef * y[j];
        dvz[i] -= coef * z[j];
      }
      if ((j != i) && feelothers[i]) {
        double dist = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]) +
                      (z[i] - z[j]) * (z[i] - z[j]);
        dist = sqrt(dist);
        coef = G * masses[j] / dist / dist / dist;
        dvx[i] += coef * (x[j] - x[i]);
        dvy[i] += coef * (y[j] - y[i]);
        dvz[i] += coef * (z[j] - z[i]);
      }
    }
  }
  for (int i = 0; i < 6 * n; i++) deriv[i] *= dt;
}

static void translate(const double *qold, double c1, double c2, double c3, double c4, double c5,
                      double *qnew, int n, const double *k1, const double *k2, const double *k3,
                      const double *k4, const double *k5) {
  for (int i = 0; i < 6 * n; i++)
    qnew[i] = qold[i] + c1 * k1[i] + c2 * k2[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
}

void f3d_psys_rk5(f3d_psys *sys, const f3d_params *p, const f3d_clock *c, double dt) {
  int n = sys->nb;
  if (n <= 0) return;
  double q0[6 * F3D_MAX_PLANETS], q1[6 * F3D_MAX_PLANETS], m[F3D_MAX_PLANETS];
  double *k1 = sys->rk[0], *k2 = sys->rk[1], *k3 = sys->rk[2];
  double *k4 = sys->rk[3], *k5 = sys->rk[4], *k6 = sys->rk[5];
  for (int i = 0; i < n; i++) {
    q0[i] = sys->x[i];
    q0[i + n] = sys->y[i];
    q0[i + 2 * n] = sys->z[i];
    q0[i + 3 * n] = sys->vx[i];
    q0[i + 4 * n] = sys->vy[i];
    q0[i + 5 * n] = sys->vz[i];
    m[i] = sys->mass[i];
  }
  const int *fo = sys->feelothers;
  int ind = p->indirectterm;
  deriv_rk5(q0, m, k1, n, dt, fo, ind);
  translate(q0, 0.2, 0.0, 0.0, 0.0, 0.0, q1, n, k1, k2, k3, k4, k5);
  deriv_rk5(q1, m, k2, n, dt, fo, ind);
  translate(q0, 0.075, 0.225, 0.0, 0.0, 0.0, q1, n, k1, k2, k3, k4, k5);
  deriv_rk5(q1, m, k3, n, dt, fo, ind);
  translate(q0, 0.3, -0.9, 1.2, 0.0, 0.0, q1, n, k1, k2, k3, k4, k5);
  deriv_rk5(q1, m, k4, n, dt, fo, ind);
  translate(q0, -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, q1, n, k1, k2, k3, k4, k5);
  deriv_rk5(q1, m, k5, n, dt, fo, ind);
  translate(q0, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
            253.0 / 4096.0, q1, n, k1, k2, k3, k4, k5);
  deriv_rk5(q1, m, k6, n, dt, fo, ind);
  for (int i = 0; i < 6 * n; i++)
    q1[i] = (q0[i] + 37.0 / 378.0 * k1[i] + 250.0 / 621.0 * k3[i] + 125.0 / 594.0 * k4[i] +
             512.0 / 1771.0 * k6[i]);
  for (int i = 1 - (c->PhysicalTime >= p->releasedate); i < n; i++) {
    sys->x[i] = q1[i];
    sys->y[i] = q1[i + n];
    sys->z[i] = q1[i + 2 * n];
    sys->vx[i] = q1[i + 3 * n];
    sys->vy[i] = q1[i + 4 * n];
    sys->vz[i] = q1[i + 5 * n];
  }
  if (c->PhysicalTime < p->releasedate) {
    /* migrating-release branch (src/runge_kutta.c:160-180) */
    double x = sys->x[0], y = sys->y[0];
    double r = sqrt(x * x + y * y);
    double theta = atan2(y, x);
    double rdot = (p->releaseradius - r) / (p->releasedate - c->PhysicalTime);
    double omega = sqrt((1. + sys->mass[0]) / r / r / r);
    double new_r = r + rdot * dt;
    double denom = r - new_r;
    double dtheta;
    if (denom != 0.0)
      dtheta = 2. * dt * r * omega / denom * (sqrt(r / new_r) - 1.);
    else
      dtheta = omega * dt;
    double vx = rdot;
    double vy = new_r * sqrt((1. + sys->mass[0]) / new_r / new_r / new_r);
    sys->x[0] = new_r * cos(dtheta + theta);
    sys->y[0] = new_r * sin(dtheta + theta);
    sys->z[0] = 0.0;
    sys->vx[0] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
    sys->vy[0] = vx * sin(dtheta + theta) + vy * cos(dtheta + theta);
    sys->vz[0] = 0.0;
  }
}

/* ---- the guiding centre of planet 0 (GetPsysInfo, src/psys.c:362-415) ----
 * The frame of a Frame G run follows the guiding centre of the first
 * planet: the point at mean anomaly M on the circle of radius a, azimuth
 * measured from the perihelion, tilted by the inclination.  The elements are
 * those of the two-body orbit about the mass m (star + planet):
 *   h = r x v (specific angular momentum), inclination asin(|h_xy| / |h|);
 *   A = v x h - G m r/|r| (the Laplace-Runge-Lenz vector, pointing to the
 *   perihelion, |A| = G m e);  a = h^2 / (G m (1 - e^2)) (vis-viva);
 *   E from r = a (1 - e cos E), negative while r.v < 0; M = E - e sin E.
 * The reference's conventions for degenerate orbits are kept, because the
 * frame's rotation (and so every velocity of a Frame G run) depends on
 * them: E = 0 for e = 0, E taken as 0 or pi when the acos argument leaves
 * [-1, 1] by rounding, and the perihelion azimuth of a circular orbit taken
 * as the planet's own azimuth.  The arithmetic is evaluated in the
 * reference's order (src/planets.c:117-195), so OMEGAFRAME matches its CPU
 * build bit for bit. */
typedef struct {
  double e, a, M, i, Perihelion_Phi;
} orb_t;

static orb_t guiding_elements(double x, double y, double z, double vx, double vy, double vz, double m) {
  orb_t o;
  double d = sqrt(x * x + y * y + z * z);
  double hx = y * vz - z * vy;
  double hy = z * vx - x * vz;
  double hz = x * vy - y * vx;
  double hhor = sqrt(hx * hx + hy * hy);
  double h = sqrt(hx * hx + hy * hy + hz * hz);
  o.i = asin(hhor / h);
  double Ax = vy * hz - vz * hy - G * m * x / d;
  double Ay = vz * hx - vx * hz - G * m * y / d;
  double Az = vx * hy - vy * hx - G * m * z / d;
  double e = sqrt(Ax * Ax + Ay * Ay + Az * Az) / (G * m);
  o.e = e;
  o.a = h * h / (G * m * (1. - e * e));
  double E = (e != 0.0) ? acos((1.0 - d / o.a) / e) : 0.0;
  if (x * vx + y * vy + z * vz < 0) E = -E;
  if (isnan(E)) E = (d < o.a) ? 0.0 : M_PI;
  o.M = E - e * sin(E);
  o.Perihelion_Phi = (Ax * Ax + Ay * Ay > 0.0) ? atan2(Ay, Ax) : atan2(y, x);
  return o;
}

double f3d_psys_info(f3d_psys *sys, const f3d_params *p, int action) {
  /* src/psys.c:362-415 */
  double x, y, xc, yc;
  double vx, vy;
  double m = sys->mass[0] + MSTAR;
  xc = x = sys->x[0];
  yc = y = sys->y[0];
  vx = sys->vx[0];
  vy = sys->vy[0];
  orb_t o = guiding_elements(sys->x[0], sys->y[0], sys->z[0], sys->vx[0], sys->vy[0], sys->vz[0], m);
  if (p->guidingcenter) {
    xc = o.a * cos(o.M + o.Perihelion_Phi) * cos(o.i);
    yc = o.a * sin(o.M + o.Perihelion_Phi) * cos(o.i);
  }
  if (o.e < 1e-8) {
    xc = x;
    yc = y;
  }
  switch (action) {
    case F3D_PSYS_MARK:
      sys->X_planet = xc;
      sys->Y_planet = yc;
      return 0.;
    case F3D_PSYS_GET: {
      x = xc;
      y = yc;
      double d2 = sqrt(x * x + y * y);
      double d1 = sqrt(sys->X_planet * sys->X_planet + sys->Y_planet * sys->Y_planet);
      double cross = sys->X_planet * y - x * sys->Y_planet;
      sys->X_planet = x;
      sys->Y_planet = y;
      return asin(cross / (d1 * d2));
    }
    case F3D_PSYS_FREQUENCY:
      if (p->guidingcenter)
        return sqrt(G * m / pow(o.a, 3.0));
      else
        return (x * vy - y * vx) / (x * x + y * y);
  }
  return 0.0;
}

void f3d_psys_rotate(f3d_psys *sys, f3d_clock *c, double angle) {
  /* src/psys.c:417-440 */
  double sint = sin(angle), cost = cos(angle);
  c->XAxisRotationAngle += angle;
  if (c->XAxisRotationAngle >= 2.0 * M_PI) c->XAxisRotationAngle -= 2.0 * M_PI;
  if (c->XAxisRotationAngle < 0.0) c->XAxisRotationAngle += 2.0 * M_PI;
  for (int i = 0; i < sys->nb; i++) {
    double xt = sys->x[i], yt = sys->y[i];
    sys->x[i] = xt * cost + yt * sint;
    sys->y[i] = -xt * sint + yt * cost;
    xt = sys->vx[i];
    yt = sys->vy[i];
    sys->vx[i] = xt * cost + yt * sint;
    sys->vy[i] = -xt * sint + yt * cost;
  }
}

double f3d_planet_taper(const f3d_params *p, double PhysicalTime) {
  /* src/potential.c:56-59 */
  if (p->masstaper == 0.0) return 1.0;
  return (PhysicalTime >= p->masstaper ? 1.0 : .5 * (1.0 - cos(M_PI * PhysicalTime / p->masstaper)));
}

double f3d_planet_smoothing2(const f3d_psys *sys, const f3d_params *p, int n, double taper) {
  /* src/potential.c:115-131 */
  double mp = sys->mass[n] * taper;
  double planetdistance = sqrt(sys->x[n] * sys->x[n] + sys->y[n] * sys->y[n] + sys->z[n] * sys->z[n]);
  double rroche = planetdistance * pow((1.0 / 3.0 * mp / MSTAR), 1.0 / 3.0);
  double smoothing;
  if (p->rochesmoothing != 0)
    smoothing = rroche * p->rochesmoothing;
  else
    smoothing = p->aspectratio * pow(planetdistance / R0, p->flaringindex) * planetdistance *
                p->thicknesssmoothing;
  smoothing *= smoothing;
  return smoothing;
}

void f3d_host_advance_planets(f3d_psys *sys, const f3d_params *p, f3d_clock *c, double dt) {
  /* src/potential.c:32-48, after Potential() and AdvanceSystemFromDisk() */
  int subcycling = 5;
  for (int i = 0; i < subcycling; i++) f3d_psys_rk5(sys, p, c, 1.0 / ((double)(subcycling)) * dt);
  if (p->corotating) {
    double omeganew = f3d_psys_info(sys, p, F3D_PSYS_GET) / dt;
    c->Domega = omeganew - c->OMEGAFRAME;
    c->OMEGAFRAME = omeganew;
  }
  f3d_psys_rotate(sys, c, c->OMEGAFRAME * dt);
}
