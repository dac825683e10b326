/* ORACLE -- TEST INFRASTRUCTURE ONLY.  Direct-sum HEALPix map2alm and angular power spectra in FP64
 * (Legendre functions in x87 long double), restating what the reference obtains from libsharp:
 *
 *   he_map2alm (src/healpix_extra.c:88-130): sharp_make_triangular_alm_info(lmax, lmax, 1) and
 *   sharp_make_weighted_healpix_geom_info(nside, 1, NULL): every pixel weighs 4 pi / Npix, so
 *     a_lm = (4 pi / Npix) sum_p f(p) Y*_lm(theta_p, phi_p),
 *   with Y_lm = lambda_lm(cos theta) e^{i m phi} (orthonormal, Condon-Shortley phase) and a_lm stored
 *   at he_indexlm(l, m, lmax) = l + m lmax - m (m - 1) / 2 (src/healpix_extra.c:36-42), m >= 0.
 *
 * libsharp (its algorithm: Reinecke 2011, A&A 526, A108; Reinecke & Seljebotn 2013) is not vendored in
 * the reference and is not installed here; the pixel centres follow Gorski et al. 2005 (ApJ 622, 759)
 * sec. 4.  Nothing is shared with the GPU code: rings are summed pixel by pixel (no FFT) and the
 * Legendre recursion runs in long double (64-bit mantissa, 15-bit exponent) instead of the GPU's
 * rescaled FP64 recursion. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "gethi_oracle.h"

/* ring ir (1 .. 4 nside - 1): pixel count, first pixel, z = cos theta, sin theta, phase
 * offset in units of pi / nphi (phi_j = pi (2 j + s) / nphi) */
void orc_ring_info(int64_t nside, int64_t ir, int64_t *nphi, int64_t *first, double *z, double *st, int *s)
{
  const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  int64_t i = ir <= 2 * nside ? ir : 4 * nside - ir; /* northern-hemisphere ring number */
  double zn, sn;
  if (i < nside) {
    double t = (double)(i * i) / (3.0 * (double)(nside * nside));
    zn = 1.0 - t;
    sn = sqrt(t * (2.0 - t));
    *nphi = 4 * i;
    *s = 1;
    *first = ir <= 2 * nside ? 2 * i * (i - 1) : npix - 2 * i * (i + 1);
  } else {
    zn = (4.0 / 3.0) - 2.0 * (double)i / (3.0 * (double)nside);
    sn = sqrt((1.0 - zn) * (1.0 + zn));
    *nphi = 4 * nside;
    *s = (int)((i - nside + 1) & 1);
    *first = ncap + (ir - nside) * 4 * nside;
  }
  *z = ir <= 2 * nside ? zn : -zn;
  *st = sn;
}

/* lambda_lm(x) for l = m .. lmax into out[l - m] (Y_lm = lambda_lm e^{i m phi}) */
void orc_legendre(int lmax, int m, double x, double st, double *out)
{
  long double lam = 1.0L / sqrtl(4.0L * (long double)M_PI);
  for (int k = 1; k <= m; k++) lam *= -sqrtl((2.0L * k + 1.0L) / (2.0L * k)) * (long double)st;
  out[0] = (double)lam;
  if (m == lmax) return;
  long double prev = lam, cur = sqrtl(2.0L * m + 3.0L) * (long double)x * lam;
  out[1] = (double)cur;
  for (int l = m + 2; l <= lmax; l++) {
    long double l2 = (long double)l * l, m2 = (long double)m * m, l1 = (long double)(l - 1) * (l - 1);
    long double a = sqrtl((4.0L * l2 - 1.0L) / (l2 - m2));
    long double b = sqrtl((l1 - m2) / (4.0L * l1 - 1.0L));
    long double nxt = a * ((long double)x * cur - b * prev);
    prev = cur;
    cur = nxt;
    out[l - m] = (double)cur;
  }
}

int64_t orc_alm_index(int l, int m, int lmax)
{
  return m > 0 ? (int64_t)l + (int64_t)m * lmax - (int64_t)m * (m - 1) / 2 : l;
}

/* maps: nmaps x Npix (RING); alms: nmaps x nalm complex, interleaved (re, im) */
int orc_map2alm(int64_t nside, int lmax, int nmaps, const double *maps, double *alms)
{
  if (nside < 1 || lmax < 0 || nmaps < 1) return -1;
  const int64_t npix = 12 * nside * nside, nring = 4 * nside - 1;
  const int64_t nalm = (int64_t)(lmax + 1) * (lmax + 2) / 2;
  const double w = 4.0 * M_PI / (double)npix;
  for (int64_t i = 0; i < 2 * nalm * nmaps; i++) alms[i] = 0;
#pragma omp parallel for schedule(dynamic)
  for (int m = 0; m <= lmax; m++) {
    double *lam = malloc(sizeof(double) * (size_t)(lmax - m + 1));
    double *fr = malloc(sizeof(double) * 2 * (size_t)nmaps);
    for (int64_t ir = 1; ir <= nring; ir++) {
      int64_t nphi, first;
      double z, st;
      int s;
      orc_ring_info(nside, ir, &nphi, &first, &z, &st, &s);
      /* F_m = sum_j f_j e^{-i m phi_j}, m phi_j = pi t / nphi, t = m (2 j + s) mod 2 nphi */
      for (int c = 0; c < nmaps; c++) {
        const double *f = maps + (size_t)c * npix + first;
        double re = 0, im = 0;
        for (int64_t j = 0; j < nphi; j++) {
          int64_t t = ((int64_t)m * (2 * j + s)) % (2 * nphi);
          double ang = M_PI * (double)t / (double)nphi;
          re += f[j] * cos(ang);
          im -= f[j] * sin(ang);
        }
        fr[2 * c] = re;
        fr[2 * c + 1] = im;
      }
      orc_legendre(lmax, m, z, st, lam);
      for (int l = m; l <= lmax; l++) {
        int64_t idx = orc_alm_index(l, m, lmax);
        for (int c = 0; c < nmaps; c++) {
          double *a = alms + 2 * ((size_t)c * nalm + idx);
          a[0] += w * lam[l - m] * fr[2 * c];
          a[1] += w * lam[l - m] * fr[2 * c + 1];
        }
      }
    }
    free(lam);
    free(fr);
  }
  return 0;
}

/* he_alm2map (src/healpix_extra.c:44-86, libsharp SHARP_ALM2MAP, unit weights): for a real map,
 *   f(p) = sum_{l, |m| <= l} a_lm Y_lm(p) = sum_l a_l0 lambda_l0 + 2 Re sum_{m>0} F_m e^{i m phi_p},
 *   F_m(ring) = sum_{l >= m} a_lm lambda_lm(z_ring),
 * summed pixel by pixel with the exact phase m phi_j = pi t / nphi, t = m (2 j + s) mod 2 nphi.
 * alms: nmaps x nalm complex (he_indexlm order); maps: nmaps x Npix (RING). */
int orc_alm2map(int64_t nside, int lmax, int nmaps, const double *alms, double *maps)
{
  if (nside < 1 || lmax < 0 || nmaps < 1) return -1;
  const int64_t npix = 12 * nside * nside, nring = 4 * nside - 1;
  const int64_t nalm = (int64_t)(lmax + 1) * (lmax + 2) / 2;
#pragma omp parallel for schedule(dynamic)
  for (int64_t ir = 1; ir <= nring; ir++) {
    int64_t nphi, first;
    double z, st;
    int s;
    orc_ring_info(nside, ir, &nphi, &first, &z, &st, &s);
    double *lam = malloc(sizeof(double) * (size_t)(lmax + 1));
    double *fm = malloc(sizeof(double) * 2 * (size_t)(lmax + 1));
    for (int c = 0; c < nmaps; c++) {
      const double *a = alms + 2 * (size_t)c * nalm;
      for (int m = 0; m <= lmax; m++) {
        orc_legendre(lmax, m, z, st, lam);
        double re = 0, im = 0;
        for (int l = m; l <= lmax; l++) {
          const int64_t idx = orc_alm_index(l, m, lmax);
          re += a[2 * idx] * lam[l - m];
          im += a[2 * idx + 1] * lam[l - m];
        }
        fm[2 * m] = re;
        fm[2 * m + 1] = im;
      }
      double *f = maps + (size_t)c * npix + first;
      for (int64_t j = 0; j < nphi; j++) {
        double v = fm[0];   /* Im a_l0 does not enter a real map */
        for (int m = 1; m <= lmax; m++) {
          int64_t t = ((int64_t)m * (2 * j + s)) % (2 * nphi);
          double ang = M_PI * (double)t / (double)nphi;
          v += 2.0 * (fm[2 * m] * cos(ang) - fm[2 * m + 1] * sin(ang));
        }
        f[j] = v;
      }
    }
    free(lam);
    free(fm);
  }
  return 0;
}

/* he_generate_beam_window (src/healpix_extra.c:389-400): b_l = exp(-l (l + 1) sigma^2 / 2),
 * sigma = FWHM2SIGMA * fwhm_arcmin with FWHM2SIGMA = (pi / 10800) / sqrt(8 ln 2) */
void orc_beam_window(int lmax, double fwhm_arcmin, double *out)
{
  const double sigma = (M_PI / 10800.0) / sqrt(8.0 * log(2.0)) * fwhm_arcmin;
  for (int l = 0; l <= lmax; l++) out[l] = exp(-0.5 * (double)l * (double)(l + 1) * sigma * sigma);
}

/* 53-bit uniform from two Philox words, as the GPU library draws its modes */
static double sht_u53(uint32_t hi, uint32_t lo)
{
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)(m & ((1ull << 53) - 1)) / 9007199254740992.0;
}

/* add_noise (src/main_jt.c:270-311) with rng_delta_gauss (src/common.c:154-164):
 *   maps[c][p] += sqrt(-sigma2[c] ln(1 - u2)) cos(2 pi u1),
 * (u1, u2) from Philox-4x32-10 with counter (p, p >> 32, c, 0) and key (seed, 0x4A4F4E54), each a
 * 53-bit uniform from two 32-bit words as the mode generator draws them. */
void orc_add_noise(int64_t npix, int nmaps, const double *sigma2, uint32_t seed, double *maps)
{
  for (int c = 0; c < nmaps; c++)
    for (int64_t p = 0; p < npix; p++) {
      uint32_t ctr[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), (uint32_t)c, 0u}, r[4];
      orc_philox4x32_10(ctr, seed, 0x4A4F4E54u, r);
      double u1 = sht_u53(r[0], r[1]), u2 = sht_u53(r[2], r[3]);
      maps[(size_t)c * npix + p] += sqrt(-sigma2[c] * log(1.0 - u2)) * cos(2.0 * M_PI * u1);
    }
}

/* C_l^{ab} = (a_l0^a a_l0^b + 2 sum_{m>0} Re(a_lm^a conj a_lm^b)) / (2l + 1), for every pair (a, b):
 * cls[(a * nmaps + b) * (lmax + 1) + l] */
void orc_alm2cl(int lmax, int nmaps, const double *alms, double *cls)
{
  const int64_t nalm = (int64_t)(lmax + 1) * (lmax + 2) / 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int a = 0; a < nmaps; a++)
    for (int b = 0; b < nmaps; b++)
      for (int l = 0; l <= lmax; l++) {
        double acc = 0;
        for (int m = 0; m <= l; m++) {
          const double *x = alms + 2 * ((size_t)a * nalm + orc_alm_index(l, m, lmax));
          const double *y = alms + 2 * ((size_t)b * nalm + orc_alm_index(l, m, lmax));
          double p = x[0] * y[0] + x[1] * y[1];
          acc += m ? 2 * p : p;
        }
        cls[((size_t)a * nmaps + b) * (lmax + 1) + l] = acc / (2 * l + 1);
      }
}
