/*
 * crime_gh.h -- C ABI of the MI355X GetHI density-to-sky library (libcrime_gh.so).
 *
 * The reference GetHI is a monolithic CLI whose hot path is three in-process calls that mutate a
 * ParamGetHI record in place (src/main_gh.c:52-62):
 *     create_d_and_vr_fields(par);   src/fourier.c:375-438
 *     get_HI(par);                   src/grid_tools.c:103-153
 *     mk_T_maps(par);                src/pixelize.c:150-286 (scatter :190-231, mK scaling :246-261)
 * This header is the device-side replacement of that seam.  Everything the three calls read from
 * ParamGetHI (src/common_gh.h:58-129) is handed over once, by value, in crime_gh_desc; the library
 * keeps the fields and maps resident in HBM and returns the maps in the layout of maps_HI after
 * pixelize.c:259 (maps[ipix + N_pix*inu], RING ordering, channel 0 = lowest frequency, mK).
 *
 * Conventions (differences from the reference are deliberate and listed):
 *  - every entry point returns a status code (0 = OK, <0 = error); nothing calls exit().  The
 *    reference exits from report_error (src/common_gh.c:104-120) and my_malloc (src/common.c:31-51).
 *  - one handle owns one HIP stream and its device buffers; handles are not shared across threads.
 *  - the caller owns every host pointer; the descriptor's arrays are copied during crime_gh_create
 *    and may be freed afterwards.
 *  - tables are FP64 and are produced by the host library (gethi_host.h) exactly as cosmo_set
 *    (src/cosmo.c:341-413) lays them out: NZ=5001 entries, uniform-r grid r_i = i/idr.
 */
#ifndef CRIME_GH_H
#define CRIME_GH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRIME_GH_ABI_VERSION 2u   /* 2: velocity_mode appended to crime_gh_desc */
#define CRIME_GH_NZ 5001            /* table length, NZ in src/common_gh.h:34 */
#define CRIME_GH_DZ 0.001           /* z step of the z->r table, DZ in src/common_gh.h:33 */
#define CRIME_GH_NU_21 1420.40575177 /* MHz, NU_21 in src/common.h:38 */

/* status codes */
enum {
  CRIME_GH_OK = 0,
  CRIME_GH_EINVAL = -1,   /* malformed descriptor or argument */
  CRIME_GH_ENOMEM = -2,   /* device allocation failed */
  CRIME_GH_EHIP = -3,     /* a HIP runtime call or kernel launch failed */
  CRIME_GH_ENODEV = -4,   /* no usable gfx950 device */
  CRIME_GH_ESTATE = -5,   /* call out of order (e.g. maps before fields) */
  CRIME_GH_EABI = -6      /* abi_version mismatch */
};

/* arithmetic of the grids, the FFT and the maps.  FP32 mirrors the reference's _SPREC build
 * (Makefile:10-20); FP64 mirrors the double build.  Per-cell geometry, tables, sigma^2 and
 * ang2pix are FP64 in both. */
enum { CRIME_GH_FP32 = 0, CRIME_GH_FP64 = 1 };

/* where the unit-variance Fourier modes come from */
enum {
  CRIME_GH_MODES_PHILOX = 0,   /* generated on the device, counter-based (Philox-4x32-10) */
  CRIME_GH_MODES_INJECTED = 1  /* unit_modes array supplied by the caller (parity runs) */
};

/* frequency binning: table of edges (reference build flag _IRREGULAR_NUTABLE, get_inu
 * src/pixelize.c:28-55) or uniform bins with the reference's (int) truncation (:216) */
enum { CRIME_GH_NU_TABLE = 0, CRIME_GH_NU_UNIFORM = 1 };

/* HI bias model.  REFERENCE = bias_HI(z)=0.904+0.135(1+z)^1.696 (src/user_defined.c:32-35);
 * CONST = b0; LINEAR = b0 + b1 (1+z).  The HI fraction is always fraction_HI (:27-30). */
enum { CRIME_GH_BIAS_REFERENCE = 0, CRIME_GH_BIAS_CONST = 1, CRIME_GH_BIAS_LINEAR = 2 };

/* Velocity field behind the RSD shift.
 *  STENCIL (the reference): one inverse FFT of the velocity potential phi_k = delta_k f0 H0 / k^2
 *    (src/fourier.c:297-298), then v_r = (central difference of phi) . r_hat (:307-373).
 *  SPECTRAL: three inverse FFTs of the velocity components v_j(k) = i sin(k_j dx)/dx phi_k, the
 *    Fourier symbol of that central difference (0 at Nyquist), projected on r_hat per cell.  The
 *    two modes give the same v_r up to rounding; SPECTRAL holds vx, vy, vz as grids (one more
 *    complex N^3 grid in HBM) and the map stage reads them instead of six neighbours of phi. */
enum { CRIME_GH_VEL_STENCIL = 0, CRIME_GH_VEL_SPECTRAL = 1 };

typedef struct crime_gh_desc {
  uint32_t abi_version;       /* CRIME_GH_ABI_VERSION */
  int32_t precision;          /* CRIME_GH_FP32 or CRIME_GH_FP64 */
  int32_t n_grid;             /* N: 2^a 3^b, 8 <= N <= 2048 */
  int32_t n_nu;               /* number of channels */
  int64_t n_side;             /* HEALPix Nside (power of two not required) */
  int32_t nu_mode;            /* CRIME_GH_NU_TABLE or CRIME_GH_NU_UNIFORM */
  int32_t rsd;                /* 1: add the RSD shift (reference), 0: real-space maps */
  const double *nu0;          /* [n_nu] lower channel edges, ascending (TABLE mode) */
  const double *nuf;          /* [n_nu] upper channel edges (TABLE mode) */
  double nu_min, nu_max;      /* MHz */
  double l_box;               /* box side, Mpc/h */
  double pos_obs[3];          /* observer position, Mpc/h */
  double omega_b, hhub;       /* for the mass-to-temperature factor (src/pixelize.c:155) */
  double f0h0;                /* fgrowth_0 * hubble_0 (src/fourier.c:255) */
  double r2_smooth;           /* squared smoothing scale (already squared, src/io_gh.c:255-260) */
  int32_t do_smoothing;
  int32_t nz;                 /* table length, must be CRIME_GH_NZ */
  const double *z_z2r;        /* [nz] z_i = i*DZ                  (z_arr_z2r) */
  const double *r_z2r;        /* [nz] chi(z_i)                     (r_arr_z2r) */
  const double *z_r2z;        /* [nz] z(r_i)                       (z_arr_r2z) */
  const double *r_r2z;        /* [nz] r_i = i/idr                  (r_arr_r2z) */
  const double *d_r2z;        /* [nz] D(r_i)/D(0)                  (growth_d_arr) */
  const double *v_r2z;        /* [nz] D H f/(f0 H0) at r_i         (growth_v_arr) */
  double idr;                 /* glob_idr */
  const double *logk;         /* [numk] log10 k of the P(k) table  (logkarr) */
  const double *pk;           /* [numk] sigma8-normalised P(k)      (pkarr) */
  int32_t numk;
  double logkmin, logkmax, idlogk, n_s;
  int32_t bias_model;         /* CRIME_GH_BIAS_* */
  double bias_b0, bias_b1;
  uint32_t seed;              /* Philox key for CRIME_GH_MODES_PHILOX */
  int32_t n_sub;              /* samples per cell (reference N_SUBPART=10, src/pixelize.c:25) */
  const double *sub_off;      /* [n_sub*3] sample offsets from the cell centre, Mpc/h */
  int32_t mode_source;        /* CRIME_GH_MODES_* */
  const double *unit_modes;   /* [2*N*N*(N/2+1)] interleaved (re,im), index kx+(N/2+1)(ky+N kz);
                                 unit-variance modes of the reference's fill order
                                 (src/fourier.c:260-302), used when mode_source is INJECTED */
  int32_t device;             /* HIP device ordinal */
  int32_t velocity_mode;      /* CRIME_GH_VEL_* (ABI 2) */
  int32_t deterministic;      /* 1: bit-reproducible maps (fixed-point scatter, see below) (ABI 2) */
} crime_gh_desc;

/* per-call timings (hipEvent, ms) and diagnostics */
typedef struct crime_gh_stats {
  double ms_fill_fft_z;       /* mode fill fused with the kz pass */
  double ms_fft_y;
  double ms_fft_x;            /* kx pass + normalisation + FP64 sigma^2 partials */
  double ms_sigma2;           /* final FP64 reduction */
  double ms_maps_clear;
  double ms_pixelize;         /* get_HI + RSD + sub-cell ang2pix + (nu,pix) scatter + mK scale */
  double ms_total;
  double sigma2_gauss;        /* <delta^2> - <delta>^2 (src/fourier.c:24-76) */
  double mean_gauss;
} crime_gh_stats;

typedef struct crime_gh crime_gh_t;

/* Grids exchanged with the parity accessors: N^3 FP64 values, x fastest (index x + N(y + N z)),
 * i.e. the reference's padded arrays with the padding dropped. */
enum {
  CRIME_GH_GRID_DENS = 0,   /* normalised Gaussian delta   (grid_dens after fourier.c:409) */
  CRIME_GH_GRID_VPOT = 1,   /* normalised velocity potential (grid_vpot after :410) */
  CRIME_GH_GRID_RVEL = 2,   /* radial velocity (grid_rvel after fourier.c:368), read only */
  CRIME_GH_GRID_MASS = 3,   /* HI mass per cell (grid_dens after grid_tools.c:145), read only */
  CRIME_GH_GRID_DZRSD = 4,  /* RSD redshift shift (grid_rvel after grid_tools.c:146), read only */
  CRIME_GH_GRID_VX = 5,     /* velocity components of CRIME_GH_VEL_SPECTRAL handles (vel[0..2] of */
  CRIME_GH_GRID_VY = 6,     /* fourier.c:359-366 before the projection); VPOT is not held by them */
  CRIME_GH_GRID_VZ = 7
};

/* point-source pass: timings (hipEvent, ms) and the number of sources drawn */
typedef struct crime_gh_ps_stats {
  double ms_sources;          /* Poisson counts, positions, luminosities, per-pixel sums */
  double ms_maps;             /* the per-channel expansion into the point-source maps */
  int64_t n_sources;
} crime_gh_ps_stats;

/* sizeof(crime_gh_desc) as compiled into the library, for FFI bindings to check their layout. */
size_t crime_gh_desc_size(void);

/* Allocates the device arena, uploads tables; no compute.  Replaces read_run_params' allocation
 * part (init_fftw src/fourier.c:101-199, allocate_maps src/io_gh.c:60-67). */
int crime_gh_create(const crime_gh_desc *desc, crime_gh_t **out);

/* Frees everything; NULL is accepted.  Replaces param_gethi_free (src/io_gh.c:298-324). */
void crime_gh_destroy(crime_gh_t *h);

/* Message of the last failed call on h (h may be NULL for a failed crime_gh_create). */
const char *crime_gh_last_error(const crime_gh_t *h);

/* create_d_and_vr_fields (src/fourier.c:375-438): mode fill, 3-D inverse FFT of delta and of the
 * velocity potential, normalisation by (sqrt(2 pi)/L)^3 and the FP64 sigma^2 of delta.  The radial
 * velocity stencil (:307-373) is evaluated inside crime_gh_mk_T_maps. */
int crime_gh_create_d_and_vr_fields(crime_gh_t *h, crime_gh_stats *st);

/* get_HI (src/grid_tools.c:103-153) fused with mk_T_maps (src/pixelize.c:150-261): per cell the
 * radial velocity, lognormal HI mass and RSD shift; per sample the channel and RING pixel; the
 * scatter-add into zeroed maps, scaled to mK per channel. */
int crime_gh_mk_T_maps(crime_gh_t *h, crime_gh_stats *st);

/* Both stages in order (src/main_gh.c:52-62). */
int crime_gh_run(crime_gh_t *h, crime_gh_stats *st);

/* Deterministic scatter (desc.deterministic = 1; no counterpart in the reference, whose FP32 OpenMP
 * atomics make the maps depend on the thread schedule, src/pixelize.c:223-226).  The map stage first
 * sums the HI mass of the whole box in a fixed order, takes the power of two s with
 * s * (total mass) in [2^61, 2^62), and scatters every sample as the integer round(s * mass / n_sub)
 * with 64-bit integer atomics, which are associative: the sums are the same whatever the order.  The
 * maps are then (sum / s) * M2T/(r^2 dnu) per channel -- the reference's own order, which scales the
 * finished mass maps (pixelize.c:246-261).  No bin can overflow: all bins together hold at most
 * 2^62 + n_samples/2.  Quantisation: s^-1 <= 2^-61 of the total mass per sample.  The maps are
 * bit-identical from run to run, on any launch order (batches included).  Costs one more read of
 * delta, an 8-byte integer map buffer and a conversion pass. */

/* Device-to-host copy of the maps as float (what write_maps writes, src/io_gh.c:69-82) or double. */
int crime_gh_copy_maps(crime_gh_t *h, float *maps_out);
int crime_gh_copy_maps_f64(crime_gh_t *h, double *maps_out);

/* Channels [c0, c0 + nc) of the maps as float into out[nc * N_pix] (no counterpart in the reference;
 * the pieces of write_maps' input, src/io_gh.c:109-131, so that a caller can write finished channels
 * while later ones are copied).  With out from crime_gh_host_alloc the copy runs at the pinned rate. */
int crime_gh_copy_maps_channels(crime_gh_t *h, int c0, int nc, float *out);

/* Page-locked host memory (hipHostMalloc) for the copies above; free with crime_gh_host_free. */
int crime_gh_host_alloc(size_t bytes, void **out);
void crime_gh_host_free(void *p);

/* sigma^2 and mean of the Gaussian field from the last create_d_and_vr_fields. */
int crime_gh_get_sigma2(crime_gh_t *h, double *sigma2, double *mean);

/* Parity accessors: read a grid (computing RVEL/MASS/DZRSD on demand from the resident fields),
 * or overwrite DENS/VPOT (stencil mode) or DENS/VX/VY/VZ (spectral mode), and sigma^2, so the map
 * stage can be fed the oracle's fields. */
int crime_gh_copy_grid(crime_gh_t *h, int which, double *dst);
/* the same for cells [cell0, cell0 + ncell) only (x-fastest order), for grids too large to copy whole */
int crime_gh_copy_grid_range(crime_gh_t *h, int which, int64_t cell0, int64_t ncell, double *dst);
int crime_gh_set_grid(crime_gh_t *h, int which, const double *src);
int crime_gh_set_sigma2(crime_gh_t *h, double sigma2, double mean);

/* Per-sample channel and pixel of cells [cell0, cell0+ncell) in x-fastest order, n_sub samples
 * each: inu[c*n_sub+s] (-1 when the sample falls outside every channel) and ipix[...] (-1 then). */
int crime_gh_sample_indices(crime_gh_t *h, int64_t cell0, int64_t ncell, int32_t *inu,
                            int64_t *ipix);

/* Parity accessors of the map stage at any size (no counterpart in the reference).  Both walk every
 * sample of every cell through the same per-sample decisions as crime_gh_sample_indices -- no cell
 * screen, no LDS aggregation -- and add mass/n_sub * M2T/(r^2 dnu) in FP64:
 *  - crime_gh_channel_totals: per-channel sums over cells [cell0, cell0+ncell) into tot[n_nu], and
 *    the number of binned samples per channel into count[n_nu] (may be NULL);
 *  - crime_gh_reference_maps: the whole maps, FP64, out[n_nu * N_pix] (allocates N_pix n_nu doubles
 *    on the device for the call). */
int crime_gh_channel_totals(crime_gh_t *h, int64_t cell0, int64_t ncell, double *tot, int64_t *count);
int crime_gh_reference_maps(crime_gh_t *h, double *out);

/* Check of the map stage's FP32 screen (no counterpart in the reference).  The map kernels take each
 * sample's channel and pixel from FP32 arithmetic wherever the argument of the decision lies farther
 * than a certified margin from the decision boundary, and from the FP64 path (the arithmetic of
 * crime_gh_sample_indices' reference form) otherwise.  This call runs every sample of the box through
 * both: counts[0] samples, [1] channel decisions the screen left to FP64, [2] pixel decisions it left
 * to FP64, [3] samples whose (channel, pixel, mK factor) differ between the two (must be 0), [4] binned
 * samples.  eps (may be NULL) receives the frequency margin (MHz) and the pixel margin. */
int crime_gh_screen_check(crime_gh_t *h, int64_t *counts, double *eps);

/* Power spectrum of the resident Gaussian field delta (no counterpart in the reference): a forward
 * 3-D FFT of the grid and |delta_k|^2 L^3/N^6 averaged over shells |k| in [i, i+1) * bin_dk (h/Mpc),
 * i < nbins <= 2048.  Outputs per shell: mean |k| (h/Mpc), mean P (Mpc/h)^3 and the number of grid
 * modes (k and -k both counted).  The transform runs in place: the fields are consumed, and
 * crime_gh_create_d_and_vr_fields must run again before crime_gh_mk_T_maps. */
int crime_gh_power_spectrum(crime_gh_t *h, int nbins, double bin_dk, double *k_mean, double *pk, int64_t *nmodes);

/* crime_gh_power_spectrum without any allocation (no counterpart in the reference).  For callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc): the shell sums live in a static device array of the library, one per GPU, 96 KB, set
 * up by crime_gh_create; nbins <= 2048 (larger is CRIME_GH_EINVAL).  Same arguments, results and
 * consumption of the fields as crime_gh_power_spectrum.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (two threads, two handles) race and must be serialised by the caller. */
int crime_gh_power_spectrum_noalloc(crime_gh_t *h, int nbins, double bin_dk, double *k_mean, double *pk,
                                    int64_t *nmodes);

/* A new realisation on the same handle (no counterpart in the reference, which reads the seed once,
 * src/io_gh.c:239): the Philox key of the mode fill becomes `seed`; the next field stage draws the new
 * modes.  Batches capture the seed at creation.  ESTATE for handles with injected modes. */
int crime_gh_set_seed(crime_gh_t *h, uint32_t seed);

/* Kernel schedule (no counterpart in the reference): sched[4] = variants of the fill+kz pass, the ky
 * pass and the kx pass, each in [0, CRIME_GH_N_SCHED), and of the map-stage aggregation table, in
 * [0, 3); 0 is the default.  FFT variants 0-2 differ in tiling with radix-16 Stockham plans; variant 3
 * is variant 0's tiling with radix-8 plans (one pass more, smaller butterflies).  All FFT variants give
 * the same fields up to the rounding of their different butterflies (variants 0-2 bit-identical).  The
 * map variants differ in the size of the LDS aggregation table (2048, 1024, 4096 bins: the atomic order
 * of the FP32 sums changes).  Grids below N = 128 have one FFT variant (their entries read back as 0). */
#define CRIME_GH_N_SCHED 4
int crime_gh_get_schedule(const crime_gh_t *h, int *sched);
int crime_gh_set_schedule(crime_gh_t *h, const int *sched);

/* Autotuner: times every variant of every stage `reps` times on this handle's buffers, keeps the
 * fastest of each (stages in order, each tuned with the winners before it) and sets the schedule.
 * sched_out[4] (optional) receives it, ms_out[4*CRIME_GH_N_SCHED] (optional) the best time of every
 * variant (-1 when the variant does not exist for this grid or stage).  The handle then holds fields and
 * maps, as after crime_gh_run.  Batches created afterwards capture the tuned schedule. */
int crime_gh_autotune(crime_gh_t *h, int reps, int *sched_out, double *ms_out);

/* The autotuner with a schedule cache (a text file, one "<key> s0 s1 s2 s3" line per tuned kind).  The
 * key holds the device architecture, the variants' version and the handle's grid, precision, Nside,
 * channels, samples per cell, velocity, scatter and frequency modes.  A hit sets the cached schedule
 * without timing anything (*hit = 1); a miss runs crime_gh_autotune and appends its result (*hit = 0).
 * hit and sched_out may be NULL. */
int crime_gh_autotune_cached(crime_gh_t *h, const char *path, int reps, int *sched_out, int *hit);

/* Device pointer and size of the resident maps (n_nu*12*Nside^2 elements of the grid precision),
 * for callers that hand the maps to further device work without a host round trip. */
int crime_gh_device_maps(crime_gh_t *h, void **dptr, size_t *bytes);

/* Point sources (do_psources, src/main_gh.c:55-66): setup_psources (src/psources.c:99-128),
 * get_point_sources (src/grid_tools.c:24-101) and mk_psources_maps (src/pixelize.c:58-148) on the
 * resident fields, into point-source maps of their own (maps_PS, written as <prefix>_ps_###.fits by
 * src/io_gh.c:122-127), layout maps_PS[ipix + N_pix * inu], mK, the grid's precision.
 *  - Per cell a Poisson count with mean n(z) dx^3 exp(g (delta - g sigma^2 / 2)), g = D(r) (bias 1),
 *    at the cell centre offset as the reference offsets it (x by pos_obs[1], grid_tools.c:66).
 *  - Per source a uniform position in the cell, its pixel, z + Delta z_RSD of the cell, and a
 *    luminosity from the reference's rejection sampler; its temperature enters every channel.
 *  - The SED is separable, so each source adds two terms to its pixel and the maps are expanded per
 *    channel afterwards; the sum is the reference's up to rounding.
 *  - Random numbers: Philox-4x32-10 keyed by (seed, 0x50535243) and counted by (cell, purpose, draw);
 *    the reference's per-thread GSL streams depend on its OpenMP thread count.
 * Requires crime_gh_create_d_and_vr_fields; may run before or after crime_gh_mk_T_maps. */
int crime_gh_mk_psources_maps(crime_gh_t *h, crime_gh_ps_stats *st);

/* Host copy of the point-source maps as FP64 (n_nu * N_pix values). */
int crime_gh_copy_ps_maps(crime_gh_t *h, double *out);

/* Device pointer and size of the point-source maps (NULL / 0 before the first pass). */
int crime_gh_device_ps_maps(crime_gh_t *h, void **dptr, size_t *bytes);

/* The two per-pixel sums A, B (N_pix FP64 each) of the last pass: T(pix, nu) = c1(nu) A + c2(nu) B. */
int crime_gh_ps_pixel_sums(crime_gh_t *h, double *A, double *B);

/* Batched runs (no counterpart in the reference, which runs one realisation per process).
 * A batch captures every member's field and map stages into ONE hipGraph, members forked over
 * n_streams concurrent streams; crime_gh_batch_run replays it with one launch.  Members may differ in
 * N, Nside, channels, cosmology and seed; they must live on one device and appear once each.  The
 * graph clears the maps with a kernel (no memset nodes).  Each member's sigma^2 stays on the device
 * (crime_gh_get_sigma2). */
typedef struct crime_gh_batch crime_gh_batch_t;
int crime_gh_batch_create(crime_gh_t **hs, int n, int n_streams, crime_gh_batch_t **out);
int crime_gh_batch_run(crime_gh_batch_t *b, double *ms); /* ms: wall time of the launch, may be NULL */
void crime_gh_batch_destroy(crime_gh_batch_t *b);
const char *crime_gh_batch_last_error(const crime_gh_batch_t *b);

/* One-shot convenience: batch_create (4 streams) + batch_run + batch_destroy.  st may be NULL or an
 * array of n; only sigma2_gauss/mean_gauss and ms_total (the whole launch) are filled. */
int crime_gh_run_batch(crime_gh_t **hs, int n, crime_gh_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* CRIME_GH_H */
