/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.  A plain FP64 C restatement of the reference GetHI
 * density-to-sky path, used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
 * checker of libcrime_gh.so.  No product code links, loads or calls it.
 *
 * It takes the same crime_gh_desc as the GPU library (same tables, edges, offsets, seed) and follows
 * the reference stage by stage:
 *   orc_fill_modes        src/fourier.c:234-305 (+ src/common.c:154-164 for the draws)
 *   orc_c2r_3d            src/fourier.c:78-99   (FFTW c2r; own FP64 mixed-radix FFT)
 *   orc_fields            src/fourier.c:375-438 (normalisation :395-412, sigma^2 :24-76)
 *   orc_radial_velocity   src/fourier.c:307-373
 *   orc_velocity_spectral the three components v_j(k) = i sin(k_j dx)/dx phi_k (the Fourier symbol
 *                         of the central difference of src/fourier.c:359-366; no reference code)
 *   orc_get_hi            src/grid_tools.c:103-153, src/user_defined.c:27-35
 *   orc_mk_T_maps         src/pixelize.c:150-261 with get_inu :28-55 (the warm-started walk)
 *   orc_vec2pix_ring      chealpix vec2pix_ring (called at src/pixelize.c:222), restated from
 *                         Gorski et al. 2005, ApJ 622:759
 * Parity pins (see tests/): FFT vs brute-force DFT and numpy.fft.irfftn; the mode fill through single
 * injected modes against the closed forms delta = norm c A Re(g e^{ik.x}), phi = delta f0 H0 / k^2 and
 * the smoothing factor e^{-r_s^2 k^2 / 2} (tests/test_oracle_modes.py); sigma^2 vs numpy; v_r vs the
 * closed-form derivative of single Fourier modes; vec2pix vs pix2ang round trips, poles and equal-area
 * counts; channels vs a brute-force scan; maps vs total-mass conservation.  The last bits of GSL's
 * integrators, FFTW's plans and chealpix are not reproducible here (none is vendored): parity with
 * them is unpinned beyond those checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "crime_gh.h"
#include "gethi_oracle.h"

/* ------------------------------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon et al., SC'11) -- the GPU library's mode generator                  */
/* ------------------------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static double u53(uint32_t hi, uint32_t lo)
{
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)(m & ((1ull << 53) - 1)) / 9007199254740992.0;
}

#define ORC_PHILOX_KEY1 0x47484931u

static void unit_mode(const crime_gh_desc *d, size_t idx, double *re, double *im)
{
  if (d->mode_source == CRIME_GH_MODES_INJECTED) {
    *re = d->unit_modes[2 * idx];
    *im = d->unit_modes[2 * idx + 1];
    return;
  }
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0, 0}, r[4];
  orc_philox4x32_10(c, d->seed, ORC_PHILOX_KEY1, r);
  double u1 = u53(r[0], r[1]), u2 = u53(r[2], r[3]);
  double mod = sqrt(-log(1.0 - u2));
  double ph = 2.0 * M_PI * u1;
  *re = mod * cos(ph);
  *im = mod * sin(ph);
}

/* ------------------------------------------------------------------------------------------ */
/* P(k) and tables (src/cosmo.c:40-86, 153-170)                                               */
/* ------------------------------------------------------------------------------------------ */
double orc_pk_linear0(const crime_gh_desc *d, double lgk)
{
  int ik = (int)((lgk - d->logkmin) * d->idlogk);
  if (ik < 0) return d->pk[0] * pow(10, d->n_s * (lgk - d->logkmin));
  if (ik < d->numk - 1) return d->pk[ik] + (lgk - d->logk[ik]) * (d->pk[ik + 1] - d->pk[ik]) * d->idlogk;
  return d->pk[d->numk - 1] * pow(10, -3 * (lgk - d->logkmax));
}

static double tab_r(const crime_gh_desc *d, const double *y, double r, double at0)
{
  if (r <= 0) return at0;
  if (r >= d->r_r2z[CRIME_GH_NZ - 1]) return y[CRIME_GH_NZ - 1];
  int ir = (int)(r * d->idr);
  if (ir > CRIME_GH_NZ - 2) ir = CRIME_GH_NZ - 2; /* r within one ulp of the last node */
  return y[ir] + (y[ir + 1] - y[ir]) * (r - d->r_r2z[ir]) * d->idr;
}

double orc_z_of_r(const crime_gh_desc *d, double r) { return tab_r(d, d->z_r2z, r, 0); }
double orc_dgrowth_of_r(const crime_gh_desc *d, double r) { return tab_r(d, d->d_r2z, r, 1); }
double orc_vgrowth_of_r(const crime_gh_desc *d, double r) { return tab_r(d, d->v_r2z, r, 1); }

double orc_r_of_z(const crime_gh_desc *d, double z)
{
  if (z <= 0) return 0;
  if (z >= d->z_z2r[CRIME_GH_NZ - 1]) return d->r_z2r[CRIME_GH_NZ - 1];
  int iz = (int)(z / CRIME_GH_DZ);
  return d->r_z2r[iz] + (d->r_z2r[iz + 1] - d->r_z2r[iz]) * (z - d->z_z2r[iz]) / CRIME_GH_DZ;
}

/* ------------------------------------------------------------------------------------------ */
/* mode fill: src/fourier.c:234-305                                                           */
/* ------------------------------------------------------------------------------------------ */
void orc_fill_modes(const crime_gh_desc *d, double *dk_out, double *pk_out)
{
  const int n = d->n_grid, nh = n / 2 + 1;
  const double dk = 2 * M_PI / d->l_box, idk3 = 1. / (dk * dk * dk);
  const double factor = d->f0h0;
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < n; ii++) {
    double kz = (2 * ii <= n) ? ii * dk : -(n - ii) * dk;
    for (int jj = 0; jj < n; jj++) {
      double ky = (2 * jj <= n) ? jj * dk : -(n - jj) * dk;
      for (int kk = 0; kk < nh; kk++) {
        double kx = kk * dk;
        size_t idx = (size_t)kk + (size_t)nh * ((size_t)jj + (size_t)n * ii);
        double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 <= 0) {
          dk_out[2 * idx] = dk_out[2 * idx + 1] = 0;
          pk_out[2 * idx] = pk_out[2 * idx + 1] = 0;
          continue;
        }
        double lgk = 0.5 * log10(k2);
        double sigma2 = orc_pk_linear0(d, lgk) * idk3;
        if (d->do_smoothing) sigma2 *= exp(-d->r2_smooth * k2);
        double gr, gi;
        unit_mode(d, idx, &gr, &gi);
        double a = sqrt(sigma2);
        dk_out[2 * idx] = a * gr;
        dk_out[2 * idx + 1] = a * gi;
        pk_out[2 * idx] = dk_out[2 * idx] * factor / k2;
        pk_out[2 * idx + 1] = dk_out[2 * idx + 1] * factor / k2;
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 inverse FFT: recursive mixed radix 2/3 (decimation in time), sign +                    */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double re, im; } orc_c;

/* out[k] = sum_j in[j*stride] exp(+2 pi i j k / n), out contiguous; w holds exp(+2 pi i e / wn) */
static void fft_rec(const orc_c *in, size_t stride, orc_c *out, int n, const orc_c *w, int wn)
{
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = (n % 2 == 0) ? 2 : 3;
  int m = n / r;
  for (int s = 0; s < r; s++) fft_rec(in + s * stride, stride * r, out + s * m, m, w, wn);
  /* X[k + m q] = sum_s W_n^{s (k + m q)} Y_s[k]; the r outputs of one k overwrite its r inputs */
  for (int k = 0; k < m; k++) {
    orc_c y[3];
    for (int s = 0; s < r; s++) y[s] = out[s * m + k];
    for (int q = 0; q < r; q++) {
      int kk = k + m * q;
      double sr = 0, si = 0;
      for (int s = 0; s < r; s++) {
        long e = ((long)s * kk) % n;
        orc_c tw = w[e * (wn / n)];
        sr += y[s].re * tw.re - y[s].im * tw.im;
        si += y[s].re * tw.im + y[s].im * tw.re;
      }
      out[kk] = (orc_c){sr, si};
    }
  }
}

static void fft1d(const orc_c *line, orc_c *out, int n, const orc_c *w) { fft_rec(line, 1, out, n, w, n); }

static int fft_size_ok(int n)
{
  if (n < 1) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

/* inverse 3-D DFT of a full complex N^3 array (index x + N(y + N z)), unnormalised, sign + */
int orc_ifft3d(int n, double *data)
{
  if (!fft_size_ok(n)) return -1;
  orc_c *w = (orc_c *)malloc(sizeof(orc_c) * n);
  for (int k = 0; k < n; k++) w[k] = (orc_c){cos(2 * M_PI * k / n), sin(2 * M_PI * k / n)};
  orc_c *a = (orc_c *)data;
  const size_t nn = (size_t)n, n2 = nn * nn;
  for (int axis = 0; axis < 3; axis++) {
    size_t stride = axis == 0 ? 1 : (axis == 1 ? nn : n2);
#pragma omp parallel
    {
      orc_c *line = (orc_c *)malloc(sizeof(orc_c) * n);
      orc_c *res = (orc_c *)malloc(sizeof(orc_c) * n);
#pragma omp for schedule(static)
      for (long q = 0; q < (long)n2; q++) {
        size_t i1 = (size_t)q % nn, i2 = (size_t)q / nn, base;
        if (axis == 0) base = nn * i1 + n2 * i2;
        else if (axis == 1) base = i1 + n2 * i2;
        else base = i1 + nn * i2;
        for (int e = 0; e < n; e++) line[e] = a[base + e * stride];
        fft1d(line, res, n, w);
        for (int e = 0; e < n; e++) a[base + e * stride] = res[e];
      }
      free(line);
      free(res);
    }
  }
  free(w);
  return 0;
}

/* c2r of a half spectrum (index kx + (N/2+1)(ky + N kz)): Hermitian completion, in which the kx=0
 * and kx=N/2 planes contribute their Hermitian part, then the inverse DFT's real part */
int orc_c2r_3d(int n, const double *half, double *out)
{
  const int nh = n / 2 + 1;
  const size_t nn = (size_t)n, n3 = nn * nn * nn;
  double *full = (double *)malloc(2 * n3 * sizeof(double));
  if (!full) return -2;
#pragma omp parallel for schedule(static)
  for (int kz = 0; kz < n; kz++) {
    for (int ky = 0; ky < n; ky++) {
      int myn = (n - ky) % n, mzn = (n - kz) % n;
      for (int kx = 0; kx < n; kx++) {
        double re, im;
        if (kx > 0 && kx < n - kx) {
          size_t i = (size_t)kx + nh * ((size_t)ky + nn * kz);
          re = half[2 * i];
          im = half[2 * i + 1];
        }
        else if (kx > n / 2) {
          size_t i = (size_t)(n - kx) + nh * ((size_t)myn + nn * mzn);
          re = half[2 * i];
          im = -half[2 * i + 1];
        }
        else {
          size_t i = (size_t)kx + nh * ((size_t)ky + nn * kz);
          size_t j = (size_t)kx + nh * ((size_t)myn + nn * mzn);
          re = 0.5 * (half[2 * i] + half[2 * j]);
          im = 0.5 * (half[2 * i + 1] - half[2 * j + 1]);
        }
        size_t o = (size_t)kx + nn * ((size_t)ky + nn * kz);
        full[2 * o] = re;
        full[2 * o + 1] = im;
      }
    }
  }
  int rc = orc_ifft3d(n, full);
  if (rc == 0)
    for (size_t i = 0; i < n3; i++) out[i] = full[2 * i];
  free(full);
  return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* create_d_and_vr_fields: src/fourier.c:375-438                                              */
/* ------------------------------------------------------------------------------------------ */
void orc_sigma2(int n, const double *dens, double *sigma2, double *mean)
{
  /* compute_sigma_dens, src/fourier.c:24-76 */
  const size_t n3 = (size_t)n * n * n;
  double s1 = 0, s2 = 0;
#pragma omp parallel for reduction(+ : s1, s2) schedule(static)
  for (long i = 0; i < (long)n3; i++) {
    s1 += dens[i];
    s2 += dens[i] * dens[i];
  }
  double m = s1 / n3;
  *mean = m;
  *sigma2 = s2 / n3 - m * m;
}

int orc_fields(const crime_gh_desc *d, double *dens, double *vpot, double *sigma2, double *mean)
{
  const int n = d->n_grid;
  const size_t nh = (size_t)n * n * (n / 2 + 1), n3 = (size_t)n * n * n;
  double *dk = (double *)malloc(2 * nh * sizeof(double));
  double *pk = (double *)malloc(2 * nh * sizeof(double));
  if (!dk || !pk) {
    free(dk);
    free(pk);
    return -2;
  }
  orc_fill_modes(d, dk, pk);
  int rc = orc_c2r_3d(n, dk, dens);
  if (!rc) rc = orc_c2r_3d(n, pk, vpot);
  free(dk);
  free(pk);
  if (rc) return rc;
  double norm = pow(sqrt(2 * M_PI) / d->l_box, 3);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n3; i++) {
    dens[i] *= norm;
    vpot[i] *= norm;
  }
  orc_sigma2(n, dens, sigma2, mean);
  return 0;
}

void orc_radial_velocity(const crime_gh_desc *d, const double *vpot, double *rvel)
{
  /* src/fourier.c:307-373 with the serial halo (slice_left = plane N-1, slice_right = plane 0) */
  const int n = d->n_grid;
  const size_t nn = n, n2 = nn * nn;
  const double dx = d->l_box / n, idx = 1. / dx;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < n; iz++) {
    size_t izh = (iz == n - 1 ? 0 : iz + 1) * n2, izl = (iz == 0 ? n - 1 : iz - 1) * n2, iz0 = iz * n2;
    double z = dx * (iz + 0.5) - d->pos_obs[2];
    for (int iy = 0; iy < n; iy++) {
      size_t iyh = (iy == n - 1 ? 0 : iy + 1) * nn, iyl = (iy == 0 ? n - 1 : iy - 1) * nn, iy0 = iy * nn;
      double y = dx * (iy + 0.5) - d->pos_obs[1];
      for (int ix = 0; ix < n; ix++) {
        size_t ixh = ix == n - 1 ? 0 : ix + 1, ixl = ix == 0 ? n - 1 : ix - 1;
        double x = dx * (ix + 0.5) - d->pos_obs[0];
        double irr = 1. / sqrt(x * x + y * y + z * z);
        double ur0 = x * irr, ur1 = y * irr, ur2 = z * irr;
        double v0 = 0.5 * idx * (vpot[ixh + iy0 + iz0] - vpot[ixl + iy0 + iz0]);
        double v1 = 0.5 * idx * (vpot[ix + iyh + iz0] - vpot[ix + iyl + iz0]);
        double v2 = 0.5 * idx * (vpot[ix + iy0 + izh] - vpot[ix + iy0 + izl]);
        rvel[ix + iy0 + iz0] = v0 * ur0 + v1 * ur1 + v2 * ur2;
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* spectral velocity components (no counterpart in the reference): v_j(k) = i sin(k_j dx)/dx phi_k, */
/* the Fourier symbol of the central difference of src/fourier.c:359-366, zero at Nyquist          */
/* ------------------------------------------------------------------------------------------ */
static double sin_dx(int m, int n, double dx)
{
  /* sin(2 pi m / n) / dx with exact zeros at m = 0 and n/2, odd under m -> n - m */
  m %= n;
  if (m == 0 || 2 * m == n) return 0;
  if (2 * m > n) return -sin_dx(n - m, n, dx);
  return (double)(sinl(2.0L * 3.14159265358979323846264338327950288L * m / n) / (long double)dx);
}

int orc_velocity_spectral(const crime_gh_desc *d, double *vx, double *vy, double *vz)
{
  const int n = d->n_grid, nh = n / 2 + 1;
  const size_t nhalf = (size_t)n * n * nh, n3 = (size_t)n * n * n;
  const double dx = d->l_box / n, norm = pow(sqrt(2 * M_PI) / d->l_box, 3);
  double *dk = (double *)malloc(2 * nhalf * sizeof(double));
  double *pk = (double *)malloc(2 * nhalf * sizeof(double));
  double *vk = (double *)malloc(2 * nhalf * sizeof(double));
  int rc = (!dk || !pk || !vk) ? -2 : 0;
  if (!rc) orc_fill_modes(d, dk, pk);
  double *out[3] = {vx, vy, vz};
  for (int j = 0; j < 3 && !rc; j++) {
#pragma omp parallel for schedule(static)
    for (int ii = 0; ii < n; ii++)
      for (int jj = 0; jj < n; jj++)
        for (int kk = 0; kk < nh; kk++) {
          size_t idx = (size_t)kk + (size_t)nh * ((size_t)jj + (size_t)n * ii);
          double s = sin_dx(j == 0 ? kk : (j == 1 ? jj : ii), n, dx);
          /* i s phi_k */
          vk[2 * idx] = -s * pk[2 * idx + 1];
          vk[2 * idx + 1] = s * pk[2 * idx];
        }
    rc = orc_c2r_3d(n, vk, out[j]);
    if (!rc)
      for (size_t i = 0; i < n3; i++) out[j][i] *= norm;
  }
  free(dk);
  free(pk);
  free(vk);
  return rc;
}

void orc_radial_velocity_spectral(const crime_gh_desc *d, const double *vx, const double *vy, const double *vz,
                                  double *rvel)
{
  /* the projection of src/fourier.c:327-368 with the velocity components given */
  const int n = d->n_grid;
  const size_t nn = n, n2 = nn * nn;
  const double dx = d->l_box / n;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < n; iz++) {
    double z = dx * (iz + 0.5) - d->pos_obs[2];
    for (int iy = 0; iy < n; iy++) {
      double y = dx * (iy + 0.5) - d->pos_obs[1];
      for (int ix = 0; ix < n; ix++) {
        size_t i = ix + nn * iy + n2 * iz;
        double x = dx * (ix + 0.5) - d->pos_obs[0];
        double irr = 1. / sqrt(x * x + y * y + z * z);
        rvel[i] = vx[i] * (x * irr) + vy[i] * (y * irr) + vz[i] * (z * irr);
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* get_HI: src/grid_tools.c:103-153                                                            */
/* ------------------------------------------------------------------------------------------ */
double orc_bias(const crime_gh_desc *d, double z)
{
  if (d->bias_model == CRIME_GH_BIAS_CONST) return d->bias_b0;
  if (d->bias_model == CRIME_GH_BIAS_LINEAR) return d->bias_b0 + d->bias_b1 * (1 + z);
  return 0.904 + 0.135 * pow(1 + z, 1.696);
}

static double fraction_hi(double z) { return 0.008 * pow(1 + z, 0.6); }

void orc_get_hi(const crime_gh_desc *d, const double *dens, const double *rvel, double sigma2, double *mass,
                double *dzrsd)
{
  const int n = d->n_grid;
  const size_t nn = n, n2 = nn * nn;
  const double dx = d->l_box / n, mass_prefac = dx * dx * dx;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < n; iz++) {
    double z = dx * (iz + 0.5) - d->pos_obs[2];
    for (int iy = 0; iy < n; iy++) {
      double y = dx * (iy + 0.5) - d->pos_obs[1];
      for (int ix = 0; ix < n; ix++) {
        size_t index = iz * n2 + iy * nn + ix;
        double x = dx * (ix + 0.5) - d->pos_obs[0];
        double r = sqrt(x * x + y * y + z * z);
        double redshift = orc_z_of_r(d, r);
        double gfd = orc_dgrowth_of_r(d, r) * orc_bias(d, redshift);
        double dz_rsd = d->rsd ? rvel[index] * orc_vgrowth_of_r(d, r) : 0.0;
        double dens_ln = exp(gfd * (dens[index] - 0.5 * gfd * sigma2));
        mass[index] = mass_prefac * fraction_hi(redshift) * dens_ln;
        dzrsd[index] = dz_rsd;
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* mk_T_maps: src/pixelize.c:28-55, 150-261                                                  */
/* ------------------------------------------------------------------------------------------ */
static int64_t imod64(int64_t v, int64_t m)
{
  int64_t r = v % m;
  return r < 0 ? r + m : r;
}

int64_t orc_vec2pix_ring(int64_t nside, double x, double y, double z)
{
  const double twopi = 6.283185307179586476925286766559;
  const double inv_halfpi = 0.636619772367581343075535053490057;
  double vlen = sqrt(x * x + y * y + z * z);
  double cth = z / vlen;
  double za = fabs(cth);
  double phi = atan2(y, x);
  if (phi < 0) {
    phi += twopi;
    if (phi == twopi) phi = 0;
  }
  double tt = phi * inv_halfpi;
  double dn = (double)nside;
  if (za <= 2.0 / 3.0) {
    double temp1 = dn * (0.5 + tt);
    double temp2 = dn * cth * 0.75;
    int64_t jp = (int64_t)(temp1 - temp2);
    int64_t jm = (int64_t)(temp1 + temp2);
    int64_t ir = nside + 1 + jp - jm;
    int64_t kshift = 1 - (ir & 1);
    int64_t ip = imod64((jp + jm - nside + kshift + 1) / 2, 4 * nside);
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  double tp = tt - (double)(int)tt;
  double tmp = za > 0.99 ? dn * (sqrt(x * x + y * y) / vlen) / sqrt((1.0 + za) / 3.0)
                         : dn * sqrt(3.0 * (1.0 - za));
  int64_t jp = (int64_t)(tp * tmp);
  int64_t jm = (int64_t)((1.0 - tp) * tmp);
  int64_t ir = jp + jm + 1;
  int64_t ip = imod64((int64_t)(tt * (double)ir), 4 * ir);
  if (cth > 0) return 2 * ir * (ir - 1) + ip;
  return 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

int orc_get_inu(const crime_gh_desc *d, double nu, int inu_start)
{
  /* get_inu, src/pixelize.c:28-55 (table mode); the uniform (int) cast of :216 otherwise */
  if (d->nu_mode == CRIME_GH_NU_UNIFORM) {
    double inv_dnu = d->n_nu / (d->nu_max - d->nu_min);
    double t = inv_dnu * (nu - d->nu_min);
    if (!(t > -1.0 && t < (double)d->n_nu)) return -1; /* (int)t outside [0,n_nu), cast kept defined */
    return (int)t;
  }
  int inu0 = inu_start < 0 ? 0 : (inu_start >= d->n_nu ? d->n_nu - 1 : inu_start);
  for (;;) {
    if (inu0 == -1 || inu0 == d->n_nu) return inu0;
    if (nu < d->nu0[inu0]) inu0--;
    else if (nu >= d->nuf[inu0]) inu0++;
    else return inu0;
  }
}

void orc_channel_factors(const crime_gh_desc *d, double *prefac)
{
  /* src/pixelize.c:155, 246-256 */
  const double m2t = 90.057156 * d->omega_b * d->hhub * (double)(12 * d->n_side * d->n_side) / (4 * M_PI);
  for (int inu = 0; inu < d->n_nu; inu++) {
    double dnu, nu;
    if (d->nu_mode == CRIME_GH_NU_TABLE) {
      dnu = d->nuf[inu] - d->nu0[inu];
      nu = (d->nuf[inu] + d->nu0[inu]) * 0.5;
    }
    else {
      dnu = (d->nu_max - d->nu_min) / d->n_nu;
      nu = d->nu_min + (inu + 0.5) * dnu;
    }
    double r = orc_r_of_z(d, CRIME_GH_NU_21 / nu - 1);
    prefac[inu] = m2t / (r * r * dnu);
  }
}

/* samples of one cell: channel (-1 if none) and pixel (-1 then) */
static void cell_samples(const crime_gh_desc *d, int ix, int iy, int iz, double dz_rsd, int *inu_state,
                         int32_t *inu_out, int64_t *ipix_out)
{
  const double dx = d->l_box / d->n_grid;
  double x0 = dx * (ix + 0.5) - d->pos_obs[0];
  double y0 = dx * (iy + 0.5) - d->pos_obs[1];
  double z0 = dx * (iz + 0.5) - d->pos_obs[2];
  for (int isub = 0; isub < d->n_sub; isub++) {
    double x = x0 + d->sub_off[3 * isub];
    double y = y0 + d->sub_off[3 * isub + 1];
    double z = z0 + d->sub_off[3 * isub + 2];
    double r = sqrt(x * x + y * y + z * z);
    double redshift = orc_z_of_r(d, r) + dz_rsd;
    double nu = CRIME_GH_NU_21 / (1 + redshift);
    int inu = orc_get_inu(d, nu, *inu_state);
    *inu_state = inu;
    if (inu >= 0 && inu < d->n_nu && r > 0) {
      inu_out[isub] = inu;
      ipix_out[isub] = orc_vec2pix_ring(d->n_side, x, y, z);
    }
    else {
      inu_out[isub] = -1;
      ipix_out[isub] = -1;
    }
  }
}

void orc_sample_indices(const crime_gh_desc *d, const double *dzrsd, int64_t cell0, int64_t ncell, int32_t *inu,
                        int64_t *ipix)
{
  const int n = d->n_grid;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; c++) {
    int64_t cell = cell0 + c;
    int ix = (int)(cell % n), iy = (int)((cell / n) % n), iz = (int)(cell / ((int64_t)n * n));
    int state = 0;
    cell_samples(d, ix, iy, iz, dzrsd[cell], &state, inu + c * d->n_sub, ipix + c * d->n_sub);
  }
}

int orc_mk_T_maps(const crime_gh_desc *d, const double *mass, const double *dzrsd, double *maps)
{
  const int n = d->n_grid;
  const int64_t npix = 12 * d->n_side * d->n_side;
  const size_t nn = n, n2 = nn * nn, nmaps = (size_t)npix * d->n_nu;
  memset(maps, 0, nmaps * sizeof(double));
#pragma omp parallel
  {
    int32_t *inu_s = (int32_t *)malloc(sizeof(int32_t) * d->n_sub);
    int64_t *pix_s = (int64_t *)malloc(sizeof(int64_t) * d->n_sub);
#pragma omp for schedule(static)
    for (int iz = 0; iz < n; iz++) {
      int state = 0; /* the walk restarts at channel 0 on every z-plane (pixelize.c:192) */
      for (int iy = 0; iy < n; iy++) {
        for (int ix = 0; ix < n; ix++) {
          size_t index = ix + iy * nn + iz * n2;
          double mass_sub = mass[index] / d->n_sub;
          cell_samples(d, ix, iy, iz, dzrsd[index], &state, inu_s, pix_s);
          for (int isub = 0; isub < d->n_sub; isub++) {
            if (inu_s[isub] < 0) continue;
            size_t k = (size_t)pix_s[isub] + (size_t)npix * inu_s[isub];
#pragma omp atomic
            maps[k] += mass_sub;
          }
        }
      }
    }
    free(inu_s);
    free(pix_s);
  }
  double *prefac = (double *)malloc(sizeof(double) * d->n_nu);
  orc_channel_factors(d, prefac);
#pragma omp parallel for schedule(static)
  for (int inu = 0; inu < d->n_nu; inu++)
    for (int64_t ip = 0; ip < npix; ip++) maps[ip + npix * inu] *= prefac[inu];
  free(prefac);
  return 0;
}

int orc_run(const crime_gh_desc *d, double *maps, double *sigma2, double *mean)
{
  const int n = d->n_grid;
  const size_t n3 = (size_t)n * n * n;
  double *dens = (double *)malloc(n3 * sizeof(double));
  double *vpot = (double *)malloc(n3 * sizeof(double));
  double *rvel = (double *)malloc(n3 * sizeof(double));
  double *mass = (double *)malloc(n3 * sizeof(double));
  double *dz = (double *)malloc(n3 * sizeof(double));
  int rc = (!dens || !vpot || !rvel || !mass || !dz) ? -2 : 0;
  if (!rc) rc = orc_fields(d, dens, vpot, sigma2, mean);
  if (!rc && d->velocity_mode == CRIME_GH_VEL_SPECTRAL) {
    /* vpot, mass and dz serve as the three components before they are needed */
    rc = orc_velocity_spectral(d, vpot, mass, dz);
    if (!rc) orc_radial_velocity_spectral(d, vpot, mass, dz, rvel);
  }
  else if (!rc)
    orc_radial_velocity(d, vpot, rvel);
  if (!rc) {
    orc_get_hi(d, dens, rvel, *sigma2, mass, dz);
    rc = orc_mk_T_maps(d, mass, dz, maps);
  }
  free(dens);
  free(vpot);
  free(rvel);
  free(mass);
  free(dz);
  return rc;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
