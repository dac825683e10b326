/*
 * gethi_host.h -- host C side of the GetHI drop-in (libgethi_host.so): parameter file, frequency
 * table, FP64 cosmology tables, linear P(k), MT19937 streams and the FITS/nu-table writers.
 *
 * It replaces the set-up half of the reference (read_run_params src/io_gh.c:188-296, cosmo_set
 * src/cosmo.c:341-413, the GSL RNG wrappers src/common.c:133-164) and its output half
 * (write_maps src/io_gh.c:109-131, he_write_healpix_map src/healpix_extra.c:132-164) and fills the
 * crime_gh_desc consumed by libcrime_gh.so.  Errors are returned, never exit()ed: every function
 * that can fail returns 0 on success and writes a message into gh_params.err.
 */
#ifndef GETHI_HOST_H
#define GETHI_HOST_H

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include "crime_gh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GH_NZ CRIME_GH_NZ
#define GH_DZ CRIME_GH_DZ
#define GH_NU_21 CRIME_GH_NU_21
#define GH_HMPC 2997.92458 /* c/H0 in Mpc/h (CSM_HMPC, src/cosmo_mad.h) */
#define GH_N_SUBPART 10    /* src/pixelize.c:25 */

typedef struct gh_params {
  /* ---- keys of the reference parameter file (src/io_gh.c:211-251) ---- */
  char prefix_out[256];
  char pk_filename[256];
  double omega_m, omega_l, omega_b, hhub, weos, n_s, sigma_8;
  double r_smooth;          /* as read; r2_smooth holds its square when > 0 */
  char nu_filename[256];    /* frequencies_filename= ; empty -> uniform nu_min/nu_max/n_nu */
  double nu_min, nu_max;
  int n_nu;
  int n_grid;
  long n_side;
  uint32_t seed;
  int do_psources;          /* 1: point-source maps <prefix>_ps_###.fits (crime_gh_mk_psources_maps) */
  /* ---- extension keys (ignored with a warning by the reference, src/io_gh.c:250-251) ---- */
  double l_box_override;    /* l_box=    (<=0: derived, src/cosmo.c:361) */
  int bias_model;           /* bias_model= reference|const|linear */
  double bias_b0, bias_b1;  /* bias_b0= bias_b1= */
  int n_subcell;            /* n_subcell= 0: 10 random shared offsets; k>0: k^3 regular sub-cells */
  int precision;            /* precision= single|double */
  int rsd;                  /* rsd= 1|0 */
  int velocity_mode;        /* velocity_mode= stencil|spectral (CRIME_GH_VEL_*) */
  int deterministic;        /* deterministic= 0|1 (bit-reproducible fixed-point scatter) */
  int mode_source;          /* modes= philox|mt19937  (mt19937: the reference's 1-thread stream,
                               injected into the device) */
  int device;               /* device= */
  /* ---- derived (cosmo_set) ---- */
  int nu_mode;              /* CRIME_GH_NU_TABLE / CRIME_GH_NU_UNIFORM */
  double *nu0, *nuf;        /* [n_nu] channel edges */
  double fgrowth_0, hubble_0, z_min, z_max, r_min, r_max, l_box, pos_obs[3];
  double r2_smooth;
  int do_smoothing;
  double z_z2r[GH_NZ], r_z2r[GH_NZ], z_r2z[GH_NZ], r_r2z[GH_NZ], d_r2z[GH_NZ], v_r2z[GH_NZ];
  double idr;
  int numk;
  double *logk, *pk;
  double logkmin, logkmax, idlogk;
  double sigma8_in;         /* sigma_8 of the table before rescaling (src/cosmo.c:335) */
  int n_sub;
  double *sub_off;          /* [n_sub*3] */
  double *unit_modes;       /* [2*N*N*(N/2+1)] when mode_source is mt19937 */
  char err[512];
} gh_params;

/* defaults of param_gethi_new (src/io_gh.c:133-186) */
void gh_params_defaults(gh_params *p);
/* parse a parameter file with the reference grammar: '#'/empty lines skipped, "key= value" with
 * sscanf("%s %s") == 2 required, unknown keys warned about on stderr (src/io_gh.c:195-252).
 * File names inside the file are used as written (relative to the working directory). */
int gh_params_read(gh_params *p, const char *fname);
/* set-up after parsing: frequency table, cosmology tables, P(k) normalisation, sub-cell offsets
 * and (mt19937 mode) the unit modes.  Mirrors the tail of read_run_params (src/io_gh.c:255-266). */
int gh_params_setup(gh_params *p);
void gh_params_free(gh_params *p);
/* the run summary of src/io_gh.c:268-291 */
void gh_params_print_summary(const gh_params *p, FILE *f);
/* fill a crime_gh_desc that points into p (p must outlive the desc until crime_gh_create) */
void gh_params_to_desc(const gh_params *p, crime_gh_desc *d);

/* opaque-handle helpers for FFI callers (Python ctypes, other languages): heap-allocated params,
 * read + setup in one call, named access to the derived scalars */
gh_params *gh_params_new(void);
void gh_params_delete(gh_params *p);
int gh_params_load(gh_params *p, const char *fname);
const char *gh_params_error(const gh_params *p);
/* "r_min" "r_max" "z_min" "z_max" "l_box" "sigma8_in" "fgrowth_0" "hubble_0" "idr" "nu_min"
 * "nu_max" "n_nu" "n_grid" "n_side" "n_sub" "precision" "seed"; NaN for an unknown name */
double gh_params_get(const gh_params *p, const char *name);

/* ---- frequency table (src/io_gh.c:29-57) ---- */
int gh_read_nutable(gh_params *p, const char *fname);
void gh_uniform_nutable(gh_params *p);

/* ---- cosmology (src/cosmo_mad.c, src/cosmo.c) ---- */
typedef struct gh_cosmo {
  double om, ol, ok, ob, h, w0;
  int ksign, normal_de;
  double a_eq, phorizon, growth0;
} gh_cosmo;
int gh_cosmo_init(gh_cosmo *c, double om, double ol, double ob, double w, double h, char *err,
                  size_t errlen);
double gh_cosmo_hubble(const gh_cosmo *c, double a);            /* h/Mpc */
double gh_cosmo_particle_horizon(const gh_cosmo *c, double a);  /* Mpc/h */
double gh_cosmo_chi(const gh_cosmo *c, double a);               /* radial comoving distance */
double gh_cosmo_growth(const gh_cosmo *c, double a);            /* unnormalised D(a) */
double gh_cosmo_fgrowth(const gh_cosmo *c, double a);           /* dlnD/dlna */
/* the five tables + idr + z/r ranges + l_box + observer (src/cosmo.c:341-394) */
int gh_cosmo_tables(gh_params *p);

/* table lookups, same arithmetic as src/cosmo.c:40-86 */
double gh_r_of_z(const gh_params *p, double z);
double gh_z_of_r(const gh_params *p, double r);
double gh_dgrowth_of_r(const gh_params *p, double r);
double gh_vgrowth_of_r(const gh_params *p, double r);

/* ---- linear power spectrum (src/cosmo.c:153-170, 232-339) ---- */
int gh_pk_read(gh_params *p, const char *fname);
double gh_pk_linear0(const gh_params *p, double lgk);
/* sigma^2 of the linear field in a top-hat of radius R (Mpc/h) with the current table */
double gh_pk_sigma2_tophat(const gh_params *p, double R);
int gh_pk_normalise(gh_params *p);
/* write A k^ns T_BBKS(k/Gamma)^2 with Sugiyama's Gamma on n log-uniform k in [kmin,kmax] */
int gh_write_bbks_pk(const char *fname, double om, double ob, double h, double ns, double kmin,
                     double kmax, int n);

/* ---- HI astrophysics (src/user_defined.c:27-35) + bias extension ---- */
double gh_fraction_hi(double z);
double gh_bias(const gh_params *p, double z);

/* ---- MT19937 with GSL seeding and gsl_rng_uniform (src/common.c:133-164) ---- */
typedef struct gh_mt { uint32_t mt[624]; int mti; } gh_mt;
void gh_mt_seed(gh_mt *s, uint32_t seed);
uint32_t gh_mt_next(gh_mt *s);
double gh_mt_uniform(gh_mt *s);
/* unit modes of the reference's single-thread stream: for every mode with k^2>0 in fill order
 * (src/fourier.c:260-302) phase 2 pi u1, modulus sqrt(-ln(1-u2)) (src/common.c:154-164);
 * zero for k=0.  out has 2*N*N*(N/2+1) doubles. */
void gh_unit_modes_mt(int n_grid, uint32_t seed, double *out);
/* sample offsets: n_subcell==0 -> 10 offsets drawn from MT19937(seed) in x,y,z order
 * (src/pixelize.c:157-164); n_subcell=k>0 -> k^3 regular sub-cell centres.  Returns n_sub. */
int gh_subcell_offsets(int n_subcell, double lcell, uint32_t seed, double *out /* [3*n_sub] */);

/* ---- outputs (src/io_gh.c:69-131, src/healpix_extra.c:132-164) ---- */
/* BINTABLE with one 1E column "T" in mK, PIXTYPE=HEALPIX, ORDERING=RING, NSIDE, COORDSYS=G */
int gh_write_healpix_fits(const char *fname, const float *map, long n_side);
/* <prefix>_nuTable.dat lines "%d %lf %lf %lf %lf" (i+1, nu0, nuf, z0, zf) */
int gh_write_nutable(const gh_params *p, const char *prefix);
/* nu-table + n_nu files <prefix>_%03d.fits from channel-major maps */
int gh_write_maps(const gh_params *p, const float *maps);

#ifdef __cplusplus
}
#endif
#endif /* GETHI_HOST_H */
