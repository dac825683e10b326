/*
 * FP64 background cosmology, the five radial tables and the linear P(k) of the GetHI host side.
 *
 * Background (src/cosmo_mad.c:160-507): particle horizon, growth factor and growth rate with the
 * same split at a_lim = 0.01 a_eq and the same closed-form matter-era pieces.  The integrals that the
 * reference hands to GSL QNG (relerr 1e-6 / 1e-4, :283, :314) and QAGIL (src/cosmo.c:273) are done
 * by an adaptive Gauss-Kronrod 7/15 rule to a much tighter tolerance; GSL is not vendored, so the
 * last bits of those integrals are not reproduced (they are pinned by closed forms in the tests).
 * Tables (src/cosmo.c:341-394) and lookups (:27-86) follow the reference's geometry and arithmetic
 * exactly: NZ=5001 points, z_i = i*DZ, uniform r_i = i/idr, the same lerp expressions.
 */
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

/* ------------------------------------------------------------------------------------------ */
/* adaptive Gauss-Kronrod 7/15 (nodes and weights of the Kronrod extension of 7-point Gauss)   */
/* ------------------------------------------------------------------------------------------ */
static const double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

typedef double (*gh_fn)(double, const void *);

static void gk15(gh_fn f, const void *par, double a, double b, double *res, double *err)
{
  double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double fc = f(c, par);
  double rk = fc * gk_wk[7], rg = fc * gk_wg[3];
  for (int j = 0; j < 7; j++) {
    double dx = h * gk_x[j];
    double f1 = f(c - dx, par), f2 = f(c + dx, par);
    rk += gk_wk[j] * (f1 + f2);
    if (j % 2 == 1) rg += gk_wg[j / 2] * (f1 + f2);
  }
  *res = rk * h;
  *err = fabs((rk - rg) * h);
}

static double gk_adapt(gh_fn f, const void *par, double a, double b, double tol, double whole, int depth)
{
  double r, e;
  gk15(f, par, a, b, &r, &e);
  if (e <= tol * fabs(whole) || e < 1e-300 || depth >= 40) return r;
  double m = 0.5 * (a + b);
  return gk_adapt(f, par, a, m, tol, whole, depth + 1) + gk_adapt(f, par, m, b, tol, whole, depth + 1);
}

/* integral of f over [a,b] to relative tolerance ~tol */
double gh_integrate(gh_fn f, const void *par, double a, double b, double tol)
{
  if (a == b) return 0;
  double r, e;
  gk15(f, par, a, b, &r, &e);
  if (e <= tol * fabs(r)) return r;
  double m = 0.5 * (a + b);
  return gk_adapt(f, par, a, m, tol, r, 1) + gk_adapt(f, par, m, b, tol, r, 1);
}

#define GH_QUAD_TOL 1e-13

/* ------------------------------------------------------------------------------------------ */
/* background: src/cosmo_mad.c:160-507                                                         */
/* ------------------------------------------------------------------------------------------ */
static double e2a3(const gh_cosmo *c, double a)
{
  /* (H/H0)^2 a^3 = OM + OL a^{-3w} + OK a  (the normalDE branch writes OL a^3) */
  if (c->normal_de) return c->om + c->ol * a * a * a + c->ok * a;
  return c->om + c->ol * pow(a, -3 * c->w0) + c->ok * a;
}

static double na_hm1(double a, const gh_cosmo *c) { return sqrt(a / e2a3(c, a)); }  /* H0/(a H) */

static double na2_hm1(double a, const void *vc)
{
  const gh_cosmo *c = (const gh_cosmo *)vc;
  return 1 / sqrt(a * e2a3(c, a)); /* H0/(a^2 H) */
}

static double na_hm3(double a, const void *vc)
{
  double d = na_hm1(a, (const gh_cosmo *)vc);
  return d * d * d;
}

double gh_cosmo_hubble(const gh_cosmo *c, double a)
{
  /* csm_hubble, src/cosmo_mad.c:321-337 */
  return sqrt(e2a3(c, a) / (a * a * a)) / GH_HMPC;
}

double gh_cosmo_particle_horizon(const gh_cosmo *c, double a)
{
  /* parthor, src/cosmo_mad.c:260-289 */
  double alim = 0.01 * c->a_eq;
  if (a <= alim) return 2 * sqrt(a / c->om) * GH_HMPC;
  double int0 = 2 * sqrt(alim / c->om);
  return (int0 + gh_integrate(na2_hm1, c, alim, a, GH_QUAD_TOL)) * GH_HMPC;
}

double gh_cosmo_chi(const gh_cosmo *c, double a)
{
  /* csm_radial_comoving_distance, src/cosmo_mad.c:349-357 */
  return c->phorizon - gh_cosmo_particle_horizon(c, a);
}

double gh_cosmo_growth(const gh_cosmo *c, double a)
{
  /* gfac, src/cosmo_mad.c:291-319 */
  double alim = 0.01 * c->a_eq;
  double int0 = 0.4 * sqrt(alim * alim * alim * alim * alim / (c->om * c->om * c->om));
  if (a <= alim) return a;
  double integral = gh_integrate(na_hm3, c, alim, a, GH_QUAD_TOL);
  return (int0 + integral) * 2.5 * c->om / (a * na_hm1(a, c));
}

double gh_cosmo_fgrowth(const gh_cosmo *c, double a)
{
  /* csm_f_growth, src/cosmo_mad.c:385-409 */
  double d = gh_cosmo_growth(c, a), coeff, apow;
  if (c->normal_de) {
    coeff = 0;
    apow = a * a * a;
  }
  else {
    coeff = 1 + c->w0;
    apow = pow(a, -3 * c->w0);
  }
  return 0.5 * (5 * c->om * a / d - (3 * c->om + 3 * coeff * c->ol * apow + 2 * c->ok * a)) /
         (c->om + c->ol * apow + c->ok * a);
}

int gh_cosmo_init(gh_cosmo *c, double om, double ol, double ob, double w, double h, char *err,
                  size_t errlen)
{
  /* csm_background_set, src/cosmo_mad.c:411-507 (w_a = 0: cosmo_set never passes another) */
  memset(c, 0, sizeof(*c));
  c->om = om;
  c->ol = ol;
  c->ob = ob;
  c->h = h;
  c->w0 = w;
  c->ok = 1 - om - ol;
  c->normal_de = fabs(w + 1) < 1e-6;
  if (fabs(c->ok) < 1e-6) c->ksign = 0;
  else if (c->ok > 0) c->ksign = -1;
  else c->ksign = 1;
  if (om <= 0) {
    snprintf(err, errlen, "Wrong matter parameter %.3lf", om);
    return -1;
  }
  if (om < ob) {
    snprintf(err, errlen, "Wrong M/B parameter %.3lf > %.3lf", ob, om);
    return -1;
  }
  if (w > -0.333333) {
    snprintf(err, errlen, "DE is too exotic (w=%.3lf)", w);
    return -1;
  }
  /* aeqmin, src/cosmo_mad.c:160-179 */
  double aeq_k = 1, aeq_l = 1;
  if (c->ksign != 0) aeq_k = om / fabs(c->ok);
  if (ol != 0) aeq_l = c->normal_de ? pow(om / ol, 0.333) : pow(om / ol, -1 / (3 * w));
  c->a_eq = aeq_k < aeq_l ? aeq_k : aeq_l;
  c->phorizon = 0;
  c->phorizon = gh_cosmo_particle_horizon(c, 1);
  c->growth0 = gh_cosmo_growth(c, 1);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* tables: src/cosmo.c:27-86, 341-394                                                          */
/* ------------------------------------------------------------------------------------------ */
static double z_of_r_provisional(const gh_params *p, double r)
{
  if (r <= 0) return 0;
  if (r >= p->r_z2r[GH_NZ - 1]) return p->z_z2r[GH_NZ - 1];
  int iz = 0;
  while (r >= p->r_z2r[iz]) iz++;
  return p->z_z2r[iz - 1] +
         (p->z_z2r[iz] - p->z_z2r[iz - 1]) * (r - p->r_z2r[iz - 1]) / (p->r_z2r[iz] - p->r_z2r[iz - 1]);
}

double gh_r_of_z(const gh_params *p, double z)
{
  if (z <= 0) return 0;
  if (z >= p->z_z2r[GH_NZ - 1]) return p->r_z2r[GH_NZ - 1];
  int iz = (int)(z / GH_DZ);
  return p->r_z2r[iz] + (p->r_z2r[iz + 1] - p->r_z2r[iz]) * (z - p->z_z2r[iz]) / GH_DZ;
}

double gh_z_of_r(const gh_params *p, double r)
{
  if (r <= 0) return 0;
  if (r >= p->r_r2z[GH_NZ - 1]) return p->z_r2z[GH_NZ - 1];
  int ir = (int)(r * p->idr);
  return p->z_r2z[ir] + (p->z_r2z[ir + 1] - p->z_r2z[ir]) * (r - p->r_r2z[ir]) * p->idr;
}

double gh_dgrowth_of_r(const gh_params *p, double r)
{
  if (r <= 0) return 1;
  if (r >= p->r_r2z[GH_NZ - 1]) return p->d_r2z[GH_NZ - 1];
  int ir = (int)(r * p->idr);
  return p->d_r2z[ir] + (p->d_r2z[ir + 1] - p->d_r2z[ir]) * (r - p->r_r2z[ir]) * p->idr;
}

double gh_vgrowth_of_r(const gh_params *p, double r)
{
  if (r <= 0) return 1;
  if (r >= p->r_r2z[GH_NZ - 1]) return p->v_r2z[GH_NZ - 1];
  int ir = (int)(r * p->idr);
  return p->v_r2z[ir] + (p->v_r2z[ir + 1] - p->v_r2z[ir]) * (r - p->r_r2z[ir]) * p->idr;
}

int gh_cosmo_tables(gh_params *p)
{
  gh_cosmo c;
  if (gh_cosmo_init(&c, p->omega_m, p->omega_l, p->omega_b, p->weos, p->hhub, p->err, sizeof(p->err)))
    return -1;
  p->fgrowth_0 = gh_cosmo_fgrowth(&c, 1);
  p->hubble_0 = gh_cosmo_hubble(&c, 1);

  p->z_min = GH_NU_21 / p->nu_max - 1;
  p->z_max = GH_NU_21 / p->nu_min - 1;
  p->r_min = gh_cosmo_chi(&c, 1 / (1 + p->z_min));
  p->r_max = gh_cosmo_chi(&c, 1 / (1 + p->z_max));

  if (p->l_box_override > 0) p->l_box = p->l_box_override;
  else p->l_box = 2 * p->r_max * (1 + 2. / p->n_grid);
  p->pos_obs[0] = 0.5 * p->l_box;
  p->pos_obs[1] = 0.5 * p->l_box;
  p->pos_obs[2] = 0.5 * p->l_box;

  for (int ii = 0; ii < GH_NZ; ii++) {
    double z = ((double)ii) * GH_DZ;
    double a = 1 / (1 + z);
    p->z_z2r[ii] = z;
    p->r_z2r[ii] = gh_cosmo_chi(&c, a);
  }
  if (p->z_z2r[GH_NZ - 1] <= p->z_max || p->r_z2r[GH_NZ - 1] <= p->r_max)
    return set_err(p, "Redshift range z_max=%.3lf beyond the z<=5 tables", p->z_max);

  double growth0 = gh_cosmo_growth(&c, 1);
  p->idr = (GH_NZ - 1) / (p->r_z2r[GH_NZ - 1] - p->r_z2r[0]);
  for (int ii = 0; ii < GH_NZ; ii++) {
    double r = ii / p->idr;
    double z = z_of_r_provisional(p, r);
    double a = 1 / (1 + z);
    p->z_r2z[ii] = z;
    p->r_r2z[ii] = r;
    double gz = gh_cosmo_growth(&c, a) / growth0;
    double fz = gh_cosmo_fgrowth(&c, a);
    double hhz = gh_cosmo_hubble(&c, a);
    p->d_r2z[ii] = gz;
    p->v_r2z[ii] = (gz * hhz * fz) / (p->fgrowth_0 * p->hubble_0);
  }
  if (p->z_r2z[GH_NZ - 1] <= p->z_max || p->r_r2z[GH_NZ - 1] <= p->r_max)
    return set_err(p, "Redshift range z_max=%.3lf beyond the z<=5 tables", p->z_max);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* linear power spectrum: src/cosmo.c:127-170, 232-339                                         */
/* ------------------------------------------------------------------------------------------ */
int gh_pk_read(gh_params *p, const char *fname)
{
  FILE *f = fopen(fname, "r");
  if (!f) return set_err(p, "Couldn't open file %s", fname);
  char line[1000];
  int n = 0;
  while (fgets(line, sizeof(line), f) != NULL) n++;
  if (n < 2) {
    fclose(f);
    return set_err(p, "P(k) file %s has fewer than two rows", fname);
  }
  rewind(f);
  free(p->logk);
  free(p->pk);
  p->numk = n;
  p->logk = (double *)malloc(n * sizeof(double));
  p->pk = (double *)malloc(n * sizeof(double));
  for (int i = 0; i < n; i++) {
    double kk, ppk;
    if (fscanf(f, "%lf %lf", &kk, &ppk) != 2) {
      fclose(f);
      return set_err(p, "Error reading file %s, line %d", fname, i + 1);
    }
    p->pk[i] = ppk;
    p->logk[i] = log10(kk);
  }
  fclose(f);
  p->logkmin = p->logk[0];
  p->logkmax = p->logk[n - 1];
  p->idlogk = (n - 1) / (p->logk[n - 1] - p->logk[0]);
  return 0;
}

double gh_pk_linear0(const gh_params *p, double lgk)
{
  /* src/cosmo.c:153-170.  The reference interpolates with pkarr[ik+1] also for ik == numk-1,
   * reading one past the table; here that last half-open bin uses the k^-3 tail instead. */
  int ik = (int)((lgk - p->logkmin) * p->idlogk);
  if (ik < 0) return p->pk[0] * pow(10, p->n_s * (lgk - p->logkmin));
  if (ik < p->numk - 1)
    return p->pk[ik] + (lgk - p->logk[ik]) * (p->pk[ik + 1] - p->pk[ik]) * p->idlogk;
  return p->pk[p->numk - 1] * pow(10, -3 * (lgk - p->logkmax));
}

static double w_tophat(double x)
{
  /* wind(x, TopHat), src/cosmo.c:137-144: Taylor series below 0.1 */
  if (x < 0.1) {
    double x2 = x * x;
    return 1. - x2 / 10. + x2 * x2 / 280. - x2 * x2 * x2 / 15120. + x2 * x2 * x2 * x2 / 1330560.;
  }
  return 3 * (sin(x) - x * cos(x)) / (x * x * x);
}

typedef struct { const gh_params *p; double R; } sig_par;

static double sig_integrand(double lgk, const void *vp)
{
  /* integxiL_NO at r=0 (j0 = 1), src/cosmo.c:212-230: ln10/(2 pi^2) P k^3 W(kR)^2 */
  const sig_par *s = (const sig_par *)vp;
  double kk = pow(10, lgk);
  double w = w_tophat(kk * s->R);
  return 0.1166503235296796 * gh_pk_linear0(s->p, lgk) * kk * kk * kk * w * w;
}

double gh_pk_sigma2_tophat(const gh_params *p, double R)
{
  /* sigL2(R,R,TopHat,TopHat) = QAGIL over (-inf, logkmax] (src/cosmo.c:232-304).  The range is cut
   * at the table's interpolation nodes so every piece is smooth, and the k^(3+ns) tail below the
   * table is integrated over 40 decades. */
  sig_par s = {p, R};
  double sum = gh_integrate(sig_integrand, &s, p->logkmin - 40.0, p->logkmin, 1e-12);
  int nb = p->numk - 1;
  for (int i = 0; i < nb; i++) {
    double a = p->logkmin + i / p->idlogk;
    double b = (i == nb - 1) ? p->logkmax : p->logkmin + (i + 1) / p->idlogk;
    sum += gh_integrate(sig_integrand, &s, a, b, 1e-12);
  }
  return sum;
}

int gh_pk_normalise(gh_params *p)
{
  /* src/cosmo.c:333-338 */
  double s2 = gh_pk_sigma2_tophat(p, 8);
  if (!(s2 > 0)) return set_err(p, "P(k) table has non-positive variance");
  p->sigma8_in = sqrt(s2);
  double norm = p->sigma_8 * p->sigma_8 / s2;
  for (int i = 0; i < p->numk; i++) p->pk[i] *= norm;
  return 0;
}

int gh_write_bbks_pk(const char *fname, double om, double ob, double h, double ns, double kmin,
                     double kmax, int n)
{
  /* BBKS transfer function (Bardeen et al. 1986, eq. G3) with Sugiyama's (1995) shape parameter
   * Gamma = Om h exp(-Ob (1 + sqrt(2h)/Om)); k in h/Mpc, arbitrary amplitude (rescaled to sigma_8) */
  FILE *f = fopen(fname, "w");
  if (!f) return -1;
  double gam = om * h * exp(-ob * (1 + sqrt(2 * h) / om));
  double l0 = log10(kmin), l1 = log10(kmax);
  for (int i = 0; i < n; i++) {
    double lk = l0 + (l1 - l0) * i / (n - 1);
    double k = pow(10, lk);
    double q = k / gam;
    double t = log(1 + 2.34 * q) / (2.34 * q) *
               pow(1 + 3.89 * q + pow(16.1 * q, 2) + pow(5.46 * q, 3) + pow(6.71 * q, 4), -0.25);
    fprintf(f, "%.17e %.17e\n", k, 1e7 * pow(k, ns) * t * t);
  }
  fclose(f);
  return 0;
}
