/*
 * Parameter file, frequency table and descriptor glue of the GetHI host side.
 *
 * Follows read_run_params / param_gethi_new / read_nutable (src/io_gh.c:29-57, 133-296): the same
 * keys, the same "key= value" grammar and the same defaults; errors are returned instead of
 * exit()ing.  Extension keys are documented in gethi_host.h.
 */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gethi_host.h"

static int set_err(gh_params *p, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(p->err, sizeof(p->err), fmt, ap);
  va_end(ap);
  return -1;
}

void gh_params_defaults(gh_params *p)
{
  /* src/io_gh.c:133-186 */
  memset(p, 0, sizeof(*p));
  snprintf(p->prefix_out, sizeof(p->prefix_out), "default");
  snprintf(p->pk_filename, sizeof(p->pk_filename), "default");
  p->omega_m = 0.3;
  p->omega_l = 0.7;
  p->omega_b = 0.05;
  p->hhub = 0.7;
  p->weos = -1.;
  p->n_s = 0.96;
  p->sigma_8 = 0.83;
  p->r_smooth = 2.0;
  p->nu_max = 1050.;
  p->nu_min = 350.;
  p->n_nu = 150;
  p->n_grid = 512;
  p->n_side = 128;
  p->seed = 1234;
  p->do_psources = 0; /* left uninitialised by the reference (src/io_gh.c:133-186) */
  p->nu_filename[0] = '\0';
  p->l_box_override = -1;
  p->bias_model = CRIME_GH_BIAS_REFERENCE;
  p->bias_b0 = 1.0;
  p->bias_b1 = 0.0;
  p->n_subcell = 0;
  p->precision = CRIME_GH_FP32;
  p->rsd = 1;
  p->velocity_mode = CRIME_GH_VEL_STENCIL;
  p->deterministic = 0;
  p->mode_source = CRIME_GH_MODES_PHILOX;
  p->device = 0;
}

static int lower_eq(const char *a, const char *b)
{
  for (; *a && *b; a++, b++)
    if (tolower((unsigned char)*a) != tolower((unsigned char)*b)) return 0;
  return *a == *b;
}

int gh_params_read(gh_params *p, const char *fname)
{
  /* src/io_gh.c:188-252 */
  FILE *fi = fopen(fname, "r");
  if (!fi) return set_err(p, "Couldn't open file %s", fname);
  char s0[512];
  int nlin = 0;
  while (fgets(s0, sizeof(s0), fi) != NULL) {
    char s1[64], s2[256];
    nlin++;
    if (s0[0] == '#' || s0[0] == '\n') continue;
    if (sscanf(s0, "%63s %255s", s1, s2) != 2) {
      fclose(fi);
      return set_err(p, "Error reading file %s, line %d", fname, nlin);
    }
    if (!strcmp(s1, "prefix_out=")) snprintf(p->prefix_out, sizeof(p->prefix_out), "%s", s2);
    else if (!strcmp(s1, "pk_filename=")) snprintf(p->pk_filename, sizeof(p->pk_filename), "%s", s2);
    else if (!strcmp(s1, "omega_M=")) p->omega_m = atof(s2);
    else if (!strcmp(s1, "omega_L=")) p->omega_l = atof(s2);
    else if (!strcmp(s1, "omega_B=")) p->omega_b = atof(s2);
    else if (!strcmp(s1, "h=")) p->hhub = atof(s2);
    else if (!strcmp(s1, "w=")) p->weos = atof(s2);
    else if (!strcmp(s1, "ns=")) p->n_s = atof(s2);
    else if (!strcmp(s1, "sigma_8=")) p->sigma_8 = atof(s2);
    else if (!strcmp(s1, "r_smooth=")) p->r_smooth = atof(s2);
    else if (!strcmp(s1, "frequencies_filename="))
      snprintf(p->nu_filename, sizeof(p->nu_filename), "%s", s2);
    else if (!strcmp(s1, "nu_min=")) p->nu_min = atof(s2);
    else if (!strcmp(s1, "nu_max=")) p->nu_max = atof(s2);
    else if (!strcmp(s1, "n_nu=")) p->n_nu = atoi(s2);
    else if (!strcmp(s1, "n_grid=")) p->n_grid = atoi(s2);
    else if (!strcmp(s1, "n_side=")) p->n_side = atoi(s2);
    else if (!strcmp(s1, "seed=")) p->seed = (uint32_t)atoi(s2);
    else if (!strcmp(s1, "do_psources=")) p->do_psources = atoi(s2);
    /* extension keys */
    else if (!strcmp(s1, "l_box=")) p->l_box_override = atof(s2);
    else if (!strcmp(s1, "bias_model=")) {
      if (lower_eq(s2, "reference") || !strcmp(s2, "0")) p->bias_model = CRIME_GH_BIAS_REFERENCE;
      else if (lower_eq(s2, "const") || !strcmp(s2, "1")) p->bias_model = CRIME_GH_BIAS_CONST;
      else if (lower_eq(s2, "linear") || !strcmp(s2, "2")) p->bias_model = CRIME_GH_BIAS_LINEAR;
      else {
        fclose(fi);
        return set_err(p, "Unknown bias_model %s (reference|const|linear)", s2);
      }
    }
    else if (!strcmp(s1, "bias_b0=")) p->bias_b0 = atof(s2);
    else if (!strcmp(s1, "bias_b1=")) p->bias_b1 = atof(s2);
    else if (!strcmp(s1, "n_subcell=")) p->n_subcell = atoi(s2);
    else if (!strcmp(s1, "precision=")) {
      if (lower_eq(s2, "single") || lower_eq(s2, "fp32")) p->precision = CRIME_GH_FP32;
      else if (lower_eq(s2, "double") || lower_eq(s2, "fp64")) p->precision = CRIME_GH_FP64;
      else {
        fclose(fi);
        return set_err(p, "Unknown precision %s (single|double)", s2);
      }
    }
    else if (!strcmp(s1, "rsd=")) p->rsd = atoi(s2) != 0;
    else if (!strcmp(s1, "deterministic=")) p->deterministic = atoi(s2) != 0;
    else if (!strcmp(s1, "velocity_mode=")) {
      if (lower_eq(s2, "stencil") || !strcmp(s2, "0")) p->velocity_mode = CRIME_GH_VEL_STENCIL;
      else if (lower_eq(s2, "spectral") || !strcmp(s2, "1")) p->velocity_mode = CRIME_GH_VEL_SPECTRAL;
      else {
        fclose(fi);
        return set_err(p, "Unknown velocity_mode %s (stencil|spectral)", s2);
      }
    }
    else if (!strcmp(s1, "modes=")) {
      if (lower_eq(s2, "philox")) p->mode_source = CRIME_GH_MODES_PHILOX;
      else if (lower_eq(s2, "mt19937")) p->mode_source = CRIME_GH_MODES_INJECTED;
      else {
        fclose(fi);
        return set_err(p, "Unknown modes %s (philox|mt19937)", s2);
      }
    }
    else if (!strcmp(s1, "device=")) p->device = atoi(s2);
    else fprintf(stderr, "CRIME: Unknown parameter %s\n", s1);
  }
  fclose(fi);
  return 0;
}

int gh_read_nutable(gh_params *p, const char *fname)
{
  /* src/io_gh.c:29-57: n_nu = #lines - 1, nu0[i] = edge i, nuf[i] = edge i+1, bins must ascend */
  FILE *fi = fopen(fname, "r");
  if (!fi) return set_err(p, "Couldn't open file %s", fname);
  char line[1000];
  int nlines = 0;
  while (fgets(line, sizeof(line), fi) != NULL) nlines++;
  if (nlines < 2) {
    fclose(fi);
    return set_err(p, "Frequency table %s needs at least two edges", fname);
  }
  rewind(fi);
  free(p->nu0);
  free(p->nuf);
  p->n_nu = nlines - 1;
  p->nu0 = (double *)malloc(p->n_nu * sizeof(double));
  p->nuf = (double *)malloc(p->n_nu * sizeof(double));
  if (!p->nu0 || !p->nuf) {
    fclose(fi);
    return set_err(p, "out of memory");
  }
  for (int inu = 0; inu <= p->n_nu; inu++) {
    double nu;
    if (fscanf(fi, "%lf ", &nu) != 1) {
      fclose(fi);
      return set_err(p, "Error reading file %s, line %d", fname, inu + 1);
    }
    if (inu != p->n_nu) p->nu0[inu] = nu;
    if (inu != 0) p->nuf[inu - 1] = nu;
  }
  fclose(fi);
  for (int inu = 0; inu < p->n_nu; inu++)
    if (p->nuf[inu] <= p->nu0[inu]) return set_err(p, "Frequency bins don't make sense");
  p->nu_max = p->nuf[p->n_nu - 1];
  p->nu_min = p->nu0[0];
  p->nu_mode = CRIME_GH_NU_TABLE;
  return 0;
}

void gh_uniform_nutable(gh_params *p)
{
  /* the non-_IRREGULAR_NUTABLE build: edges nu_min + (nu_max-nu_min)*i/n_nu (src/io_gh.c:93-96) */
  free(p->nu0);
  free(p->nuf);
  p->nu0 = (double *)malloc(p->n_nu * sizeof(double));
  p->nuf = (double *)malloc(p->n_nu * sizeof(double));
  for (int i = 0; i < p->n_nu; i++) {
    p->nu0[i] = p->nu_min + (p->nu_max - p->nu_min) * (i + 0.0) / p->n_nu;
    p->nuf[i] = p->nu_min + (p->nu_max - p->nu_min) * (i + 1.0) / p->n_nu;
  }
  p->nu_mode = CRIME_GH_NU_UNIFORM;
}

static int grid_size_ok(int n)
{
  if (n < 8 || n > 2048) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

int gh_params_setup(gh_params *p)
{
  if (!grid_size_ok(p->n_grid))
    return set_err(p, "n_grid=%d is not supported (2^a 3^b between 8 and 2048)", p->n_grid);
  if (p->n_side < 1 || p->n_side > (1L << 13)) return set_err(p, "n_side=%ld out of range", p->n_side);
  if (p->n_subcell < 0 || p->n_subcell > 8) return set_err(p, "n_subcell=%d out of range", p->n_subcell);

  /* src/io_gh.c:255-260 */
  if (p->r_smooth > 0) {
    p->r2_smooth = pow(p->r_smooth, 2);
    p->do_smoothing = 1;
  }
  else {
    p->r2_smooth = p->r_smooth;
    p->do_smoothing = 0;
  }

  if (p->nu_filename[0]) {
    if (gh_read_nutable(p, p->nu_filename)) return -1;
  }
  else {
    if (p->n_nu < 1 || !(p->nu_max > p->nu_min)) return set_err(p, "Frequency bins don't make sense");
    gh_uniform_nutable(p);
  }

  if (gh_cosmo_tables(p)) return -1;
  if (gh_pk_read(p, p->pk_filename)) return -1;
  if (gh_pk_normalise(p)) return -1;

  int nsub = p->n_subcell > 0 ? p->n_subcell * p->n_subcell * p->n_subcell : GH_N_SUBPART;
  free(p->sub_off);
  p->sub_off = (double *)malloc(3 * (size_t)nsub * sizeof(double));
  if (!p->sub_off) return set_err(p, "out of memory");
  p->n_sub = gh_subcell_offsets(p->n_subcell, p->l_box / p->n_grid, p->seed, p->sub_off);

  free(p->unit_modes);
  p->unit_modes = NULL;
  if (p->mode_source == CRIME_GH_MODES_INJECTED) {
    size_t nm = (size_t)p->n_grid * p->n_grid * (p->n_grid / 2 + 1);
    p->unit_modes = (double *)malloc(2 * nm * sizeof(double));
    if (!p->unit_modes) return set_err(p, "out of memory for %zu unit modes", nm);
    gh_unit_modes_mt(p->n_grid, p->seed, p->unit_modes);
  }
  return 0;
}

void gh_params_free(gh_params *p)
{
  free(p->nu0);
  free(p->nuf);
  free(p->logk);
  free(p->pk);
  free(p->sub_off);
  free(p->unit_modes);
  p->nu0 = p->nuf = p->logk = p->pk = p->sub_off = p->unit_modes = NULL;
}

void gh_params_print_summary(const gh_params *p, FILE *f)
{
  /* src/io_gh.c:268-291 */
  double dk = 2 * M_PI / p->l_box;
  double dtheta = sqrt(41253. / (12 * p->n_side * p->n_side));
  double drTmax = p->r_max * dtheta / 57.2957795;
  double drTmin = p->r_min * dtheta / 57.2957795;
  double dnu = (p->nu_max - p->nu_min) / p->n_nu;
  double drL = (p->r_max - p->r_min) / p->n_nu;
  double esz = p->precision == CRIME_GH_FP64 ? 8.0 : 4.0;
  double output_size = 12.0 * p->n_side * p->n_side * p->n_nu * esz / (1024. * 1024 * 1024);
  fprintf(f, "Run parameters: \n");
  fprintf(f, "  %.3lf < nu/MHz < %.3lf\n", p->nu_min, p->nu_max);
  fprintf(f, "  %.3lf < z < %.3lf\n", p->z_min, p->z_max);
  fprintf(f, "  %.3lf < r/(Mpc/h) < %.3lf\n", p->r_min, p->r_max);
  fprintf(f, "  L_box = %.3lf Mpc/h, N_grid = %d \n", p->l_box, p->n_grid);
  fprintf(f, "  Scales resolved: %.3lE < k < %.3lE h/Mpc\n", dk, 0.5 * (p->n_grid - 1) * dk);
  fprintf(f, "  Fourier-space resolution: dk = %.3lE h/Mpc\n", dk);
  fprintf(f, "  Real-space resolution: dx = %.3lE Mpc/h\n", p->l_box / p->n_grid);
  if (p->do_smoothing)
    fprintf(f, "  Density field pre-smoothed on scales: x_s = %.3lE Mpc/h\n", sqrt(p->r2_smooth));
  else
    fprintf(f, "  No extra smoothing\n");
  fprintf(f, "  n_nu = %d, d_nu= %.3lf MHz, drL ~ %.3lf Mpc/h\n", p->n_nu, dnu, drL);
  fprintf(f, "  n_side = %ld, dtheta = %.3lf deg, %.3lf < drT/(Mpc/h) < %.3lf\n", p->n_side, dtheta,
          drTmin, drTmax);
  fprintf(f, "  Estimated output size ~ %.1lf GB \n", output_size);
  fprintf(f, "\n");
}

void gh_params_to_desc(const gh_params *p, crime_gh_desc *d)
{
  memset(d, 0, sizeof(*d));
  d->abi_version = CRIME_GH_ABI_VERSION;
  d->precision = p->precision;
  d->n_grid = p->n_grid;
  d->n_nu = p->n_nu;
  d->n_side = p->n_side;
  d->nu_mode = p->nu_mode;
  d->rsd = p->rsd;
  d->nu0 = p->nu0;
  d->nuf = p->nuf;
  d->nu_min = p->nu_min;
  d->nu_max = p->nu_max;
  d->l_box = p->l_box;
  for (int i = 0; i < 3; i++) d->pos_obs[i] = p->pos_obs[i];
  d->omega_b = p->omega_b;
  d->hhub = p->hhub;
  d->f0h0 = p->fgrowth_0 * p->hubble_0;
  d->r2_smooth = p->r2_smooth;
  d->do_smoothing = p->do_smoothing;
  d->nz = GH_NZ;
  d->z_z2r = p->z_z2r;
  d->r_z2r = p->r_z2r;
  d->z_r2z = p->z_r2z;
  d->r_r2z = p->r_r2z;
  d->d_r2z = p->d_r2z;
  d->v_r2z = p->v_r2z;
  d->idr = p->idr;
  d->logk = p->logk;
  d->pk = p->pk;
  d->numk = p->numk;
  d->logkmin = p->logkmin;
  d->logkmax = p->logkmax;
  d->idlogk = p->idlogk;
  d->n_s = p->n_s;
  d->bias_model = p->bias_model;
  d->bias_b0 = p->bias_b0;
  d->bias_b1 = p->bias_b1;
  d->seed = p->seed;
  d->n_sub = p->n_sub;
  d->sub_off = p->sub_off;
  d->mode_source = p->mode_source;
  d->unit_modes = p->unit_modes;
  d->device = p->device;
  d->velocity_mode = p->velocity_mode;
  d->deterministic = p->deterministic;
}

double gh_fraction_hi(double z)
{
  /* src/user_defined.c:27-30 */
  return 0.008 * pow(1 + z, 0.6);
}

double gh_bias(const gh_params *p, double z)
{
  /* src/user_defined.c:32-35 for REFERENCE; extension models otherwise */
  switch (p->bias_model) {
  case CRIME_GH_BIAS_CONST: return p->bias_b0;
  case CRIME_GH_BIAS_LINEAR: return p->bias_b0 + p->bias_b1 * (1 + z);
  default: return 0.904 + 0.135 * pow(1 + z, 1.696);
  }
}

gh_params *gh_params_new(void)
{
  gh_params *p = (gh_params *)calloc(1, sizeof(gh_params));
  if (p) gh_params_defaults(p);
  return p;
}

void gh_params_delete(gh_params *p)
{
  if (!p) return;
  gh_params_free(p);
  free(p);
}

int gh_params_load(gh_params *p, const char *fname)
{
  if (gh_params_read(p, fname)) return -1;
  return gh_params_setup(p);
}

const char *gh_params_error(const gh_params *p) { return p->err; }

double gh_params_get(const gh_params *p, const char *name)
{
  struct { const char *k; double v; } t[] = {
      {"r_min", p->r_min}, {"r_max", p->r_max}, {"z_min", p->z_min}, {"z_max", p->z_max},
      {"l_box", p->l_box}, {"sigma8_in", p->sigma8_in}, {"fgrowth_0", p->fgrowth_0},
      {"hubble_0", p->hubble_0}, {"idr", p->idr}, {"nu_min", p->nu_min}, {"nu_max", p->nu_max},
      {"n_nu", p->n_nu}, {"n_grid", p->n_grid}, {"n_side", (double)p->n_side}, {"n_sub", p->n_sub},
      {"precision", p->precision}, {"seed", p->seed}, {"velocity_mode", p->velocity_mode},
      {"deterministic", p->deterministic}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (!strcmp(name, t[i].k)) return t[i].v;
  return NAN;
}
