/* ORACLE -- TEST INFRASTRUCTURE ONLY: FP64 restatement of the reference GetHI density-to-sky path
 * (see gethi_oracle.c).  Grids are N^3 doubles, x fastest; half spectra are interleaved (re,im)
 * with index kx + (N/2+1)(ky + N kz); maps are n_nu * 12 Nside^2 doubles, channel-major. */
#ifndef GETHI_ORACLE_H
#define GETHI_ORACLE_H
#include <stdint.h>
#include "crime_gh.h"

void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);
double orc_pk_linear0(const crime_gh_desc *d, double lgk);
double orc_z_of_r(const crime_gh_desc *d, double r);
double orc_dgrowth_of_r(const crime_gh_desc *d, double r);
double orc_vgrowth_of_r(const crime_gh_desc *d, double r);
double orc_r_of_z(const crime_gh_desc *d, double z);
void orc_fill_modes(const crime_gh_desc *d, double *dens_k, double *vpot_k);
int orc_ifft3d(int n, double *data);
int orc_c2r_3d(int n, const double *half, double *out);
void orc_sigma2(int n, const double *dens, double *sigma2, double *mean);
int orc_fields(const crime_gh_desc *d, double *dens, double *vpot, double *sigma2, double *mean);
void orc_radial_velocity(const crime_gh_desc *d, const double *vpot, double *rvel);
int orc_velocity_spectral(const crime_gh_desc *d, double *vx, double *vy, double *vz);
void orc_radial_velocity_spectral(const crime_gh_desc *d, const double *vx, const double *vy, const double *vz,
                                  double *rvel);
double orc_bias(const crime_gh_desc *d, double z);
void orc_get_hi(const crime_gh_desc *d, const double *dens, const double *rvel, double sigma2, double *mass,
                double *dzrsd);
int64_t orc_vec2pix_ring(int64_t nside, double x, double y, double z);
int orc_get_inu(const crime_gh_desc *d, double nu, int inu_start);
void orc_channel_factors(const crime_gh_desc *d, double *prefac);
void orc_sample_indices(const crime_gh_desc *d, const double *dzrsd, int64_t cell0, int64_t ncell, int32_t *inu,
                        int64_t *ipix);
int orc_mk_T_maps(const crime_gh_desc *d, const double *mass, const double *dzrsd, double *maps);
int orc_run(const crime_gh_desc *d, double *maps, double *sigma2, double *mean);
int orc_num_threads(void);

/* sht_oracle.c: direct-sum HEALPix map2alm and angular power spectra */
void orc_ring_info(int64_t nside, int64_t ir, int64_t *nphi, int64_t *first, double *z, double *st, int *s);
void orc_legendre(int lmax, int m, double x, double st, double *out);
int64_t orc_alm_index(int l, int m, int lmax);
int orc_map2alm(int64_t nside, int lmax, int nmaps, const double *maps, double *alms);
void orc_alm2cl(int lmax, int nmaps, const double *alms, double *cls);
int orc_alm2map(int64_t nside, int lmax, int nmaps, const double *alms, double *maps);
double orc_ps_lfunc(double L, double z);
double orc_ps_sed(double nu);
double orc_ps_temp(double hhub, double L, double nu_obs, double z, double r, double domega);
void orc_ps_tables(double z_max, double *nz, double *maxpdf);
uint32_t orc_ps_poisson(double lam, uint32_t seed, uint64_t cell);
double orc_ps_luminosity(const double *nz, const double *maxpdf, double z_max, double z, uint32_t seed,
                         uint64_t cell, uint32_t src);
int64_t orc_psources(const crime_gh_desc *d, const double *dens, const double *dzrsd, double sigma2, double *maps,
                     double *A, double *B);
void orc_beam_window(int lmax, double fwhm_arcmin, double *out);
void orc_add_noise(int64_t npix, int nmaps, const double *sigma2, uint32_t seed, double *maps);
#endif
