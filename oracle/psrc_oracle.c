/* ORACLE -- TEST INFRASTRUCTURE ONLY.  FP64 restatement of the reference's point sources:
 *   setup_psources      src/psources.c:99-128    (n(z) and max_Lpdf tables)
 *   n_of_z_psources     src/psources.c:74-90
 *   draw_luminosity     src/psources.c:130-152    (rejection sampling of log10 L)
 *   temp_of_l / spec_ed src/psources.c:54-64, 157-167
 *   get_point_sources   src/grid_tools.c:24-101   (Poisson count per cell; x offset by pos_obs[1], :66)
 *   mk_psources_maps    src/pixelize.c:58-148     (uniform position in the cell, every channel)
 * The random numbers follow the GPU library's documented streams (Philox-4x32-10, key (seed,
 * 0x50535243), counter (cell, cell >> 32, purpose, draw)) so that both sides draw the same sources;
 * the Poisson sampler is inversion below lambda = 10 and PTRS (Hormann 1993) above.  The maps are
 * summed here per source and per channel as the reference sums them; the GPU's per-pixel two-term
 * form is pinned against this sum in the tests. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "crime_gh.h"
#include "gethi_oracle.h"

#define NZ_PS 256
#define NL_PS 256
#define LLOGMIN -5.0
#define LLOGMAX 6.0
#define LX_TOY 2.11
#define ALPHA_TOY -0.633
#define RHO_TOY 3.2E-4
#define SED_NORM 0.114461
#define FLUX2TEMP 3.2548291E-2
#define LUM2FLUX 8.35774E7
#define PS_KEY 0x50535243u

static double ps_u53(uint32_t hi, uint32_t lo)
{
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)(m & ((1ull << 53) - 1)) / 9007199254740992.0;
}

static void ps_uniforms(uint32_t seed, uint64_t cell, uint32_t w2, uint32_t w3, double *u1, double *u2)
{
  uint32_t c[4] = {(uint32_t)cell, (uint32_t)(cell >> 32), w2, w3}, r[4];
  orc_philox4x32_10(c, seed, PS_KEY, r);
  *u1 = ps_u53(r[0], r[1]);
  *u2 = ps_u53(r[2], r[3]);
}

/* l_z_function (src/psources.c:37-52), as written there */
double orc_ps_lfunc(double L, double z)
{
  double rho_l = 2.5 * M_LN10 * RHO_TOY * pow(L / LX_TOY, ALPHA_TOY + 1) * exp(-L / LX_TOY);
  double fz = z < 1.5 ? pow(1 + z, 3.1) : 17.124;
  return fz * rho_l;
}

/* spec_ed (src/psources.c:54-64) */
double orc_ps_sed(double nu)
{
  double nu_ghz = nu * 0.001;
  return SED_NORM * (pow(nu_ghz, -0.1) + 10 * pow(nu_ghz, -0.75));
}

/* temp_of_l (src/psources.c:157-167): mK of a source of luminosity L (1e22 W/Hz/sr) at nu_obs */
double orc_ps_temp(double hhub, double L, double nu_obs, double z, double r, double domega)
{
  double s_nu = LUM2FLUX * 4 * M_PI * L * orc_ps_sed((1 + z) * nu_obs) * hhub * hhub / (r * r * (1 + z));
  return FLUX2TEMP * s_nu / (domega * nu_obs * nu_obs);
}

void orc_ps_tables(double z_max, double *nz, double *maxpdf)
{
  double dlogL = (LLOGMAX - LLOGMIN) / NL_PS, dz = z_max / NZ_PS;
  for (int ii = 0; ii < NZ_PS; ii++) {
    double z = ii * dz, mx = -1;
    nz[ii] = 0;
    for (int jj = 0; jj < NL_PS; jj++) {
      double lf = orc_ps_lfunc(pow(10., LLOGMIN + (jj + 0.5) * dlogL), z);
      if (lf >= mx) mx = lf;
      nz[ii] += dlogL * lf;
    }
    maxpdf[ii] = nz[ii] > 0 ? 1.1 * mx / nz[ii] : 0;
  }
}

/* n_of_z_psources (src/psources.c:74-90), with glob_inv_dz = 1 / glob_dz (:107) */
static double nofz(const double *nz, double dz, double z)
{
  const double inv_dz = 1. / dz;
  int iz = (int)(z * inv_dz);
  if (iz >= NZ_PS || iz < 0) return -1;
  if (iz == NZ_PS - 1) return nz[NZ_PS - 1];
  double zi = iz * dz;
  return nz[iz] + (nz[iz + 1] - nz[iz]) * (z - zi) * inv_dz;
}

uint32_t orc_ps_poisson(double lam, uint32_t seed, uint64_t cell)
{
  if (!(lam > 0)) return 0;
  if (lam < 10.0) {
    double u, unused;
    ps_uniforms(seed, cell, 0, 0, &u, &unused);
    double p = exp(-lam), F = p;
    uint32_t k = 0;
    while (u > F && k < 1000) {
      k++;
      p *= lam / (double)k;
      F += p;
    }
    return k;
  }
  double slam = sqrt(lam), loglam = log(lam);
  double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t j = 0; j < 1024; j++) {
    double U, V;
    ps_uniforms(seed, cell, 0, j, &U, &V);
    U -= 0.5;
    double us = 0.5 - fabs(U);
    double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (uint32_t)k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return (uint32_t)k;
  }
  return (uint32_t)lam;
}

/* draw_luminosity with the library's counters; the acceptance test as the reference writes it */
double orc_ps_luminosity(const double *nz, const double *maxpdf, double z_max, double z, uint32_t seed,
                         uint64_t cell, uint32_t src)
{
  if (z >= z_max) return 0;
  double dz = z_max / NZ_PS;
  int iz = (int)(z * (1. / dz));
  iz = iz < 0 ? 0 : (iz < NZ_PS ? iz : NZ_PS - 1);
  double mp = maxpdf[iz], n = nofz(nz, dz, z), L = 0;
  for (uint32_t att = 0; att < 4096; att++) {
    double u, w;
    ps_uniforms(seed, cell, 4 * src + 2, att, &u, &w);
    double logL = LLOGMIN + (LLOGMAX - LLOGMIN) * w;
    L = pow(10., logL);
    if (u * mp < orc_ps_lfunc(L, z) / n) return L;
  }
  return L;
}

/* Sources of every cell from dens (N^3 Gaussian delta) and dzrsd (N^3 Delta z_RSD), x fastest.
 * maps (n_nu * Npix, zeroed by the caller) get every source's temperature in every channel;
 * A, B (Npix each, zeroed by the caller, may be NULL) get the two per-pixel terms of the GPU's form.
 * Returns the number of sources. */
int64_t orc_psources(const crime_gh_desc *d, const double *dens, const double *dzrsd, double sigma2, double *maps,
                     double *A, double *B)
{
  const int n = d->n_grid;
  const double dx = d->l_box / n, cell_vol = dx * dx * dx;
  const double z_max = CRIME_GH_NU_21 / d->nu_min - 1;
  double nz[NZ_PS], mp[NZ_PS];
  orc_ps_tables(z_max, nz, mp);
  const double dz = z_max / NZ_PS;
  const int64_t npix = 12 * d->n_side * d->n_side;
  const double domega = 4 * M_PI / (double)npix;
  double *nu = malloc(sizeof(double) * (size_t)d->n_nu);
  for (int i = 0; i < d->n_nu; i++)
    nu[i] = d->nu_mode == CRIME_GH_NU_TABLE ? (d->nu0[i] + d->nuf[i]) * 0.5
                                            : d->nu_min + (d->nu_max - d->nu_min) * (i + 0.5) / d->n_nu;
  int64_t total = 0;
  for (int iz = 0; iz < n; iz++)
    for (int iy = 0; iy < n; iy++)
      for (int ix = 0; ix < n; ix++) {
        const uint64_t cell = (uint64_t)ix + (uint64_t)n * ((uint64_t)iy + (uint64_t)n * iz);
        double x0 = (ix + 0.5) * dx - d->pos_obs[1];
        double y0 = (iy + 0.5) * dx - d->pos_obs[1];
        double z0 = (iz + 0.5) * dx - d->pos_obs[2];
        double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
        double redshift = orc_z_of_r(d, r);
        double ndens = nofz(nz, dz, redshift);
        if (!(ndens > 0)) continue;
        double gfb = orc_dgrowth_of_r(d, r);
        double lambda = ndens * cell_vol * exp(gfb * (dens[cell] - 0.5 * gfb * sigma2));
        uint32_t np = orc_ps_poisson(lambda, d->seed, cell);
        total += np;
        double cx0 = dx * ix - d->pos_obs[0], cy0 = dx * iy - d->pos_obs[1], cz0 = dx * iz - d->pos_obs[2];
        for (uint32_t s = 0; s < np; s++) {
          double ux, uy, uz, unused;
          ps_uniforms(d->seed, cell, 4 * s + 1, 0, &ux, &uy);
          ps_uniforms(d->seed, cell, 4 * s + 1, 1, &uz, &unused);
          double px = cx0 + dx * ux, py = cy0 + dx * uy, pz = cz0 + dx * uz;
          double rs = sqrt(px * px + py * py + pz * pz);
          double red_cosmo = orc_z_of_r(d, rs);
          double red_true = red_cosmo + dzrsd[cell];
          int64_t ipix = orc_vec2pix_ring(d->n_side, px, py, pz);
          double L = orc_ps_luminosity(nz, mp, z_max, red_cosmo, d->seed, cell, s);
          if (maps)
            for (int i = 0; i < d->n_nu; i++) {
              double t = orc_ps_temp(d->hhub, L, nu[i], red_true, rs, domega);
              if (t > 0) maps[ipix + npix * i] += t;
            }
          if (A && L > 0) {
            double opz = 1 + red_true, w = L / (rs * rs * opz);
            A[ipix] += w * pow(opz * 1e-3, -0.1);
            B[ipix] += 10.0 * w * pow(opz * 1e-3, -0.75);
          }
        }
      }
  free(nu);
  return total;
}
