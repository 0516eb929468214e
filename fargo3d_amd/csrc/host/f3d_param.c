/*
 * f3d_param.c -- .par reader and the .opt-equivalent setup switches.
 *
 * Same file syntax as FARGO3D's ReadVarFile (src/param.c:98-217): one
 * "Name value" per line, separators "\t :=>", '#' starts a comment line,
 * names are case-insensitive, INCLUDE reads another file, booleans are true
 * when the value starts with 'Y'/'y'.  Defaults follow std/stdpar.par.
 */
#include "f3d_host.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { T_INT, T_REAL, T_STR, T_BOOL, T_ENUM };

typedef struct {
  const char *name;
  int type;
  size_t off;
  size_t len; /* strings */
} var_t;

#define OFF(f) offsetof(f3d_params, f)
#define SLEN(f) sizeof(((f3d_params *)0)->f)

static const var_t VARS[] = {
    {"SETUP", T_STR, OFF(setup), SLEN(setup)},
    {"NX", T_INT, OFF(nx), 0},
    {"NY", T_INT, OFF(ny), 0},
    {"NZ", T_INT, OFF(nz), 0},
    {"XMIN", T_REAL, OFF(xmin), 0},
    {"XMAX", T_REAL, OFF(xmax), 0},
    {"YMIN", T_REAL, OFF(ymin), 0},
    {"YMAX", T_REAL, OFF(ymax), 0},
    {"ZMIN", T_REAL, OFF(zmin), 0},
    {"ZMAX", T_REAL, OFF(zmax), 0},
    {"SPACING", T_STR, OFF(spacing), SLEN(spacing)},
    {"ASPECTRATIO", T_REAL, OFF(aspectratio), 0},
    {"SIGMA0", T_REAL, OFF(sigma0), 0},
    {"SIGMASLOPE", T_REAL, OFF(sigmaslope), 0},
    {"FLARINGINDEX", T_REAL, OFF(flaringindex), 0},
    {"NU", T_REAL, OFF(nu), 0},
    {"ALPHA", T_REAL, OFF(alpha), 0},
    {"GAMMA", T_REAL, OFF(gamma), 0},
    {"NOISE", T_REAL, OFF(noise), 0},
    {"CFL", T_REAL, OFF(cfl), 0},
    {"PLANETCONFIG", T_STR, OFF(planetconfig), SLEN(planetconfig)},
    {"THICKNESSSMOOTHING", T_REAL, OFF(thicknesssmoothing), 0},
    {"ROCHESMOOTHING", T_REAL, OFF(rochesmoothing), 0},
    {"ECCENTRICITY", T_REAL, OFF(eccentricity), 0},
    {"INCLINATION", T_REAL, OFF(inclination), 0},
    {"INDIRECTTERM", T_BOOL, OFF(indirectterm), 0},
    {"EXCLUDEHILL", T_BOOL, OFF(excludehill), 0},
    {"MASSTAPER", T_REAL, OFF(masstaper), 0},
    {"RELEASEDATE", T_REAL, OFF(releasedate), 0},
    {"RELEASERADIUS", T_REAL, OFF(releaseradius), 0},
    {"SEMIMAJORAXIS", T_REAL, OFF(semimajoraxis), 0},
    {"ORBITALRADIUS", T_REAL, OFF(orbitalradius), 0},
    {"PLANETMASS", T_REAL, OFF(planetmass), 0},
    {"FRAME", T_STR, OFF(frame), SLEN(frame)},
    {"OMEGAFRAME", T_REAL, OFF(omegaframe), 0},
    {"DAMPINGZONE", T_REAL, OFF(dampingzone), 0},
    {"TAUDAMP", T_REAL, OFF(taudamp), 0},
    {"KILLINGBCCOLATITUDE", T_REAL, OFF(killingbccolatitude), 0},
    {"PERIODICZ", T_BOOL, OFF(periodicz), 0},
    {"PERIODICY", T_BOOL, OFF(periodicy), 0},
    {"DT", T_REAL, OFF(dt_out), 0},
    {"NINTERM", T_INT, OFF(ninterm), 0},
    {"NTOT", T_INT, OFF(ntot), 0},
    {"NSNAP", T_INT, OFF(nsnap), 0},
    {"FUNCARCHFILE", T_STR, OFF(funcarchfile), SLEN(funcarchfile)},
    {"OUTPUTDIR", T_STR, OFF(outputdir), SLEN(outputdir)},
    /* .opt switches, settable at run time */
    {"NFLUIDS", T_INT, OFF(nfluids), 0},
    {"INVSTOKES1", T_REAL, OFF(invstokes[1]), 0},
    {"INVSTOKES2", T_REAL, OFF(invstokes[2]), 0},
    {"INVSTOKES3", T_REAL, OFF(invstokes[3]), 0},
    {"INVSTOKES4", T_REAL, OFF(invstokes[4]), 0},
    {"INVSTOKES5", T_REAL, OFF(invstokes[5]), 0},
    {"INVSTOKES6", T_REAL, OFF(invstokes[6]), 0},
    {"INVSTOKES7", T_REAL, OFF(invstokes[7]), 0},
    {"EPSILON", T_REAL, OFF(epsilon), 0},
    {"DRAGFORCE", T_BOOL, OFF(dragforce), 0},
    {"DUSTDIFFUSION", T_BOOL, OFF(dustdiffusion), 0},
    {"FEEDBACK", T_BOOL, OFF(feedback), 0},
    {"KEPLERIANIC", T_BOOL, OFF(ic_keplerian), 0},
};

/* Variables FARGO3D accepts but the hydro step does not use. */
static const char *IGNORED[] = {"XMA", "XMB", "XMC", "YMA", "YMB", "YMC", "YMY0", "OORTA",
                                "CS", "VERTICALDAMPING", "COORDINATES", "OHMICDIFFUSIONCOEFF",
                                "HALLEFFECTCOEFF", "AMBIPOLARDIFFUSIONCOEFF", "RESONANCE",
                                "BETA", "VTK", "REALTYPE", "FIELD", "CMAP", "PLOTLOG",
                                "COLORBAR", "AUTOCOLOR", "ASPECT", "VMIN", "VMAX", "PLOTLINE",
                                "WRITEDENSITY", "WRITEENERGY", "WRITEBX", "WRITEBY",
                                "WRITEBZ", "WRITEVX", "WRITEVY", "WRITEVZ", "WRITEDIVERGENCE",
                                "WRITEENERGYRAD", "WRITETAU", "DISK", "INVSTOKES1",
                                "INVSTOKES2", "INVSTOKES3", "EPSILON", "END", NULL};

void f3d_params_defaults(f3d_params *p) {
  memset(p, 0, sizeof(*p));
  /* std/stdpar.par */
  strcpy(p->setup, "fargo");
  p->nx = p->ny = p->nz = 1;
  p->xmin = 0.0; p->xmax = 1.0;
  p->ymin = 0.0; p->ymax = 1.0;
  p->zmin = 0.0; p->zmax = 1.0;
  strcpy(p->spacing, "Lin");
  p->sigmaslope = 0.0;
  p->eccentricity = 0.0;
  p->inclination = 0.0;
  p->thicknesssmoothing = 0.0;
  p->rochesmoothing = 0.0;
  p->aspectratio = 0.05;
  p->sigma0 = 1.0;
  p->flaringindex = 0.0;
  p->nu = 0.0;
  strcpy(p->frame, "fixed");
  p->omegaframe = 0.0;
  p->indirectterm = 0;
  p->excludehill = 0;
  p->masstaper = 0.0;
  p->noise = 0.0;
  p->gamma = 1.66666667;
  p->alpha = 0.0;
  strcpy(p->planetconfig, "NONE");
  p->periodicz = 0;
  p->periodicy = 0;
  p->dampingzone = 1.15;
  p->taudamp = 0.3;
  p->killingbccolatitude = -0.2;
  p->cfl = 0.44;
  p->orbitalradius = 0.0;
  p->releasedate = 0.0;
  p->releaseradius = 0.0;
  p->semimajoraxis = 0.0;
  p->planetmass = 0.0;
  strcpy(p->funcarchfile, "std/func_arch.cfg");
  strcpy(p->outputdir, "outputs");
  p->dt_out = 0.314159265359;
  p->ninterm = 1;
  p->ntot = 1;
  /* default switches: the "fargo" setup */
  p->geometry = F3D_CYLINDRICAL;
  p->dimz = 0;
  p->eos = F3D_ISOTHERMAL;
  p->visc = F3D_VISC_NU;
  p->potential = 1;
  p->stockholm = 1;
  p->fargo = 1;
  p->bc = F3D_BC_FARGO2D;
  p->energy_bc = 0;
  p->epsilon = 0.0;
  p->dragforce = p->dustdiffusion = 0;
  p->feedback = 1;
  p->precision = 64;
  p->nfluids = 1;
  p->ic_keplerian = 0;
  p->seed = 0;
  p->mon_2d = p->mon_y_raw = 0;
  p->mon_scalar = F3D_MON_MASS | F3D_MON_MOM_X | F3D_MON_TORQ;
}

/* The switches each setup's .opt file sets (setups/<name>/<name>.opt). */
int f3d_params_setup_options(f3d_params *p, const char *setup) {
  if (strcmp(setup, "fargo") == 0 || strcmp(setup, "fargo_keplerian") == 0) {
    /* setups/fargo/fargo.opt:11-37 */
    p->geometry = F3D_CYLINDRICAL; p->dimz = 0; p->eos = F3D_ISOTHERMAL;
    p->visc = F3D_VISC_NU; p->potential = 1; p->stockholm = 1; p->fargo = 1;
    p->bc = F3D_BC_FARGO2D;
    p->ic_keplerian = (strcmp(setup, "fargo_keplerian") == 0);
    /* setups/fargo/fargo.opt:16-17 */
    p->mon_scalar = F3D_MON_MASS | F3D_MON_MOM_X | F3D_MON_TORQ; p->mon_y_raw = F3D_MON_TORQ; p->mon_2d = 0;
  } else if (strcmp(setup, "fargo_adiabatic") == 0) {
    /* fargo.opt with -DADIABATIC in place of -DISOTHERMAL */
    p->geometry = F3D_CYLINDRICAL; p->dimz = 0; p->eos = F3D_ADIABATIC;
    p->visc = F3D_VISC_NU; p->potential = 1; p->stockholm = 1; p->fargo = 1;
    p->bc = F3D_BC_FARGO2D;
    /* a setup-local .bound entry "Energy: Ymin/Ymax: SYMMETRIC" */
    p->energy_bc = 1;
    p->mon_scalar = F3D_MON_MASS | F3D_MON_MOM_X | F3D_MON_TORQ; p->mon_y_raw = F3D_MON_TORQ; p->mon_2d = 0;
  } else if (strcmp(setup, "fargo_multifluid") == 0) {
    /* setups/fargo_multifluid/fargo_multifluid.opt:13-44: one gas and
       NFLUIDS-1 dust fluids, implicit drag with Stokes numbers, dust diffusion */
    p->geometry = F3D_CYLINDRICAL; p->dimz = 0; p->eos = F3D_ISOTHERMAL;
    p->visc = F3D_VISC_NU; p->potential = 1; p->stockholm = 1; p->fargo = 1;
    p->bc = F3D_BC_FARGO2D;
    p->nfluids = 4;
    p->dragforce = 1;
    p->dustdiffusion = 1;
    /* setups/fargo_multifluid/fargo_multifluid.opt:21-22 */
    p->mon_scalar = F3D_MON_MASS | F3D_MON_MOM_X | F3D_MON_TORQ; p->mon_y_raw = F3D_MON_TORQ; p->mon_2d = 0;
  } else if (strcmp(setup, "p3disof") == 0 || strcmp(setup, "p3disof_alpha") == 0) {
    /* setups/p3disof/p3disof.opt:15-33 (ALPHAVISCOSITY variant on request) */
    p->geometry = F3D_SPHERICAL; p->dimz = 1; p->eos = F3D_ISOTHERMAL;
    p->visc = (strcmp(setup, "p3disof_alpha") == 0) ? F3D_VISC_ALPHA : F3D_VISC_NU;
    p->potential = 1; p->stockholm = 1; p->fargo = 1;
    p->bc = F3D_BC_P3D;
    /* setups/p3disof/p3disof.opt:19-20 */
    p->mon_2d = F3D_MON_MASS | F3D_MON_MOM_X; p->mon_scalar = F3D_MON_TORQ; p->mon_y_raw = 0;
  } else {
    return -1;
  }
  snprintf(p->setup, sizeof(p->setup), "%s", setup);
  return 0;
}

static void upcase(char *s) {
  for (; *s; s++) *s = (char)toupper((unsigned char)*s);
}

int f3d_params_set(f3d_params *p, const char *name_in, const char *value) {
  char name[128];
  snprintf(name, sizeof(name), "%s", name_in);
  upcase(name);
  for (size_t i = 0; i < sizeof(VARS) / sizeof(VARS[0]); i++) {
    if (strcmp(name, VARS[i].name) != 0) continue;
    char *base = (char *)p + VARS[i].off;
    switch (VARS[i].type) {
      case T_INT: *(int *)base = (int)strtod(value, NULL); break;
      case T_REAL: *(double *)base = strtod(value, NULL); break;
      case T_BOOL: *(int *)base = (toupper((unsigned char)value[0]) == 'Y'); break;
      case T_STR: snprintf(base, VARS[i].len, "%s", value); break;
    }
    if (strcmp(name, "SETUP") == 0) f3d_params_setup_options(p, value);
    return 0;
  }
  /* .opt switches spelled as FARGO3D's flags */
  if (strcmp(name, "GEOMETRY") == 0) {
    p->geometry = (toupper((unsigned char)value[0]) == 'S') ? F3D_SPHERICAL : F3D_CYLINDRICAL;
    return 0;
  }
  if (strcmp(name, "EOS") == 0) {
    p->eos = (toupper((unsigned char)value[0]) == 'A') ? F3D_ADIABATIC : F3D_ISOTHERMAL;
    return 0;
  }
  if (strcmp(name, "VISCOSITYMODEL") == 0) {
    char c = (char)toupper((unsigned char)value[0]);
    p->visc = (c == 'A') ? F3D_VISC_ALPHA : (c == 'N' && toupper((unsigned char)value[1]) == 'U') ? F3D_VISC_NU : F3D_VISC_NONE;
    return 0;
  }
  if (strcmp(name, "STOCKHOLM") == 0) { p->stockholm = (toupper((unsigned char)value[0]) == 'Y'); return 0; }
  if (strcmp(name, "FARGOADVECTION") == 0) { p->fargo = (toupper((unsigned char)value[0]) == 'Y'); return 0; }
  if (strcmp(name, "POTENTIALTERM") == 0) { p->potential = (toupper((unsigned char)value[0]) == 'Y'); return 0; }
  if (strcmp(name, "PRECISION") == 0) {
    p->precision = (strstr(value, "32") != NULL) ? 32 : 64;
    return 0;
  }
  if (strcmp(name, "ENERGYBOUNDARY") == 0) { p->energy_bc = (toupper((unsigned char)value[0]) == 'S'); return 0; }
  if (strcmp(name, "SEED") == 0) { p->seed = strtoull(value, NULL, 10); return 0; }
  if (strcmp(name, "CONDINITONDEVICE") == 0) { p->ic_device = (toupper((unsigned char)value[0]) == 'Y'); return 0; }
  for (int i = 0; IGNORED[i]; i++)
    if (strcmp(name, IGNORED[i]) == 0) return 0;
  return -1;
}

int f3d_params_read(f3d_params *p, const char *path, char *err, size_t errlen) {
  FILE *in = fopen(path, "r");
  if (!in) {
    if (err) snprintf(err, errlen, "Unable to read '%s'", path);
    return -1;
  }
  char s[512], name[128], strval[256];
  const char *sep = "\t :=>";
  while (fgets(s, sizeof(s) - 1, in)) {
    if (sscanf(s, "%127s", name) != 1 || name[0] == '#') continue;
    char *s1 = s + strlen(name);
    s1 += strspn(s1, sep);
    if (sscanf(s1, "%255s", strval) != 1) strval[0] = 0;
    char up[128];
    snprintf(up, sizeof(up), "%s", name);
    upcase(up);
    if (strcmp(up, "INCLUDE") == 0) {
      if (f3d_params_read(p, strval, err, errlen) != 0) { fclose(in); return -1; }
      continue;
    }
    if (f3d_params_set(p, name, strval) != 0)
      fprintf(stderr, "Warning: variable %s defined but does not exist in code.\n", up);
  }
  fclose(in);
  return 0;
}

void f3d_params_finalize(f3d_params *p) {
  /* src/param.c:236-240 */
  p->corotating = p->guidingcenter = 0;
  if (p->frame[0] == 'C' || p->frame[0] == 'c') p->corotating = 1;
  if (p->frame[0] == 'G' || p->frame[0] == 'g') {
    p->corotating = 1;
    p->guidingcenter = 1;
  }
  if (!p->dimz) p->nz = 1; /* src/main.c:211-213 */
  if (p->nfluids < 1) p->nfluids = 1;
}
