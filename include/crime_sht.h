/*
 * crime_sht.h -- HEALPix analysis (map2alm), synthesis (alm2map), angular power spectra and a JoinT-like
 * beam / noise / mask stage on one MI355X, in libcrime_gh.so.
 *
 * Replaces he_map2alm (src/healpix_extra.c:88-130), which calls libsharp with
 * sharp_make_triangular_alm_info(lmax, lmax, 1) and sharp_make_weighted_healpix_geom_info(nside, 1,
 * NULL): every pixel weighs 4 pi / Npix (no quadrature ring weights, no iterations), so
 *     a_lm = (4 pi / Npix) sum_p f(p) Y*_lm(theta_p, phi_p),     0 <= m <= l <= lmax,
 * with orthonormal Y_lm carrying the Condon-Shortley phase.  Host alms follow he_indexlm
 * (src/healpix_extra.c:36-42): a_lm of map t at alms[t][l + m lmax - m (m - 1) / 2], complex as
 * (re, im) doubles -- the reference's fcomplex of its double-precision build.
 *
 * The angular power spectra (no counterpart in the reference's GetHI; the reference's ForGet draws
 * maps from given C_l, src/gsync_qu.c) are
 *     C_l^{ab} = ( a_l0^a a_l0^b + 2 sum_{m>0} Re(a_lm^a conj a_lm^b) ) / (2 l + 1).
 *
 * All arithmetic is FP64.  A plan owns one HIP stream and its device buffers (the Legendre table of
 * the plan's Nside and lmax is kept resident when it fits the plan's budget); plans are not shared
 * across threads.  Status codes are those of crime_gh.h.
 */
#ifndef CRIME_SHT_H
#define CRIME_SHT_H

#include <stddef.h>
#include <stdint.h>

#include "crime_gh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crime_sht crime_sht_t;

/* Plan for RING maps of the given Nside (1..8192) up to lmax (0..4 nside - 1; 3 nside - 1 is the
 * usual choice) and up to max_maps maps per call, on HIP device `device`.  The device buffers hold
 * the ring spectra ((lmax+1) x 2 x 2 nside x max_maps complex doubles), the alms of max_maps maps
 * and the Legendre table or a chunk of it. */
int crime_sht_create(int64_t nside, int lmax, int max_maps, int device, crime_sht_t **out);
void crime_sht_destroy(crime_sht_t *p);
const char *crime_sht_last_error(const crime_sht_t *p); /* NULL plan: the last failed create */

/* he_map2alm (src/healpix_extra.c:88-130) for nmaps <= max_maps host maps maps[t][Npix] (RING)
 * into host alms[t][(lmax+1)(lmax+2)/2] (he_indexlm order, interleaved re, im). */
int crime_sht_map2alm(crime_sht_t *p, int nmaps, const double *maps, double *alms);

/* Same, from maps already in device memory: d_maps[t * Npix + pix], FP64 (maps_fp64 = 1) or FP32
 * (maps_fp64 = 0), e.g. the resident maps of crime_gh_device_maps. */
int crime_sht_map2alm_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, double *alms);

/* Angular power spectra of device maps.  cross = 1: every pair, cls[(a * nmaps + b) * (lmax + 1) + l];
 * cross = 0: autospectra only, cls[a * (lmax + 1) + l].  cls = NULL keeps them in the plan's device
 * buffer (crime_sht_device_cls) without a copy to the host. */
int crime_sht_anafast_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, int cross,
                             double *cls);

/* Cross spectra within each of nsets sets of maps_per_set device maps, in one call: set s has its
 * maps at d_maps[s] (a host array of nsets device pointers, each to maps_per_set * Npix values),
 * nsets * maps_per_set <= max_maps.  cls[((s * maps_per_set + a) * maps_per_set + b) * (lmax + 1) + l];
 * NULL keeps them on the device (crime_sht_device_cls).  For an ensemble of mocks of one Nside: the
 * per-member spectra of many members share the launches, GEMM tiles and LDS of one analysis. */
int crime_sht_anafast_sets_device(crime_sht_t *p, int nsets, const void *const *d_maps, int maps_per_set,
                                  int maps_fp64, double *cls);

/* Device pointer and byte count of the last spectra computed (valid until the next call). */
int crime_sht_device_cls(const crime_sht_t *p, const double **dptr, size_t *bytes);

/* C_l^{ab} for a list of map pairs only: pairs[2 i], pairs[2 i + 1] index the nmaps device maps;
 * cls[i * (lmax + 1) + l]. */
int crime_sht_cls_pairs_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, int npairs,
                               const int32_t *pairs, double *cls);

/* Same as crime_sht_anafast_device, for nmaps <= max_maps host maps maps[t][Npix] (FP64, RING). */
int crime_sht_anafast(crime_sht_t *p, int nmaps, const double *maps, int cross, double *cls);

/* he_alm2map (src/healpix_extra.c:44-86; libsharp SHARP_ALM2MAP with the same alm and geometry
 * info as he_map2alm): for host alms[t][(lmax+1)(lmax+2)/2] (he_indexlm order, interleaved re, im) of
 * nmaps <= max_maps real maps, maps[t][pix] = sum_{l, |m| <= l} a_lm Y_lm(pixel centre), with
 * a_{l,-m} = (-1)^m conj a_lm; Im a_l0 is ignored, as for any real map.  maps: host FP64, RING. */
int crime_sht_alm2map(crime_sht_t *p, int nmaps, const double *alms, double *maps);

/* Same with device alms (same layout) into device maps d_maps[t * Npix + pix], FP64 (maps_fp64 = 1)
 * or FP32 (maps_fp64 = 0). */
int crime_sht_alm2map_device(crime_sht_t *p, int nmaps, const double *d_alms, void *d_maps, int maps_fp64);

/* apply_beam of JoinT (src/main_jt.c:211-246) in place on device maps: map2alm, then
 * he_alter_alm (src/healpix_extra.c:402-421) with the Gaussian window b_l = exp(-l(l+1) sigma_t^2 / 2),
 * sigma_t = FWHM2SIGMA fwhm_arcmin[t] (src/healpix_extra.c:389-400), then alm2map.  fwhm_arcmin is a
 * host array of nmaps values >= 0 (JoinT uses 1030610.55 / (nu_t [MHz] * dish diameter [m])). */
int crime_sht_smooth_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64, const double *fwhm_arcmin);

/* add_noise of JoinT (src/main_jt.c:270-311) on device maps: maps[t][pix] += sqrt(-sigma2[t] ln(1 - u2))
 * cos(2 pi u1), the draw of rng_delta_gauss (src/common.c:154-164), whose variance is sigma2[t] / 2.
 * (u1, u2) are 53-bit uniforms from Philox-4x32-10 with counter (pix, pix >> 32, t, 0) and key
 * (seed, 0x4A4F4E54) -- the reference's per-thread GSL streams depend on its thread count, so no
 * stream can match them; this one depends on nothing but (seed, t, pix).  sigma2: host, nmaps values. */
int crime_sht_add_noise_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64, const double *sigma2,
                               uint32_t seed);

/* apply_mask of JoinT (src/main_jt.c:248-268): maps[t][pix] = 0 where mask[pix] == 0.  mask: host,
 * Npix bytes (nonzero = observed). */
int crime_sht_mask_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64, const uint8_t *mask);

/* Milliseconds of the last call's stages: ring transforms, Legendre table, alm GEMMs, spectra (the beam
 * for crime_sht_smooth_device), total.  A call with both an analysis and a synthesis reports the sum of
 * the two for each stage. */
typedef struct crime_sht_stats {
  double ms_rings, ms_legendre, ms_alm, ms_cls, ms_total;
} crime_sht_stats;
int crime_sht_get_stats(const crime_sht_t *p, crime_sht_stats *st);

#ifdef __cplusplus
}
#endif
#endif
