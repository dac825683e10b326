/*
 * gethi -- drop-in for the reference GetHI binary (src/main_gh.c:24-80).
 *
 *   gethi <param.ini>
 *
 * Same parameter file, same outputs (<prefix>_001..n_nu.fits and <prefix>_nuTable.dat), same stage
 * order; the hot path (create_d_and_vr_fields, get_HI, mk_T_maps) runs on the GPU through the
 * C ABI of include/crime_gh.h.  Stage times are printed in the reference's timer() format
 * (src/common.c:77-131).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

#include "crime_gh.h"
#include "gethi_host.h"

static double wall_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Map output, pipelined: the main thread copies blocks of channels from the device into a pinned
 * buffer while writer threads turn every copied channel into its <prefix>_###.fits (write_maps,
 * src/io_gh.c:109-131, one file per channel).  Channel i may be written once i < ncopied. */
typedef struct {
  const gh_params *par;
  const float *maps;
  long npix;
  int n_nu;
  int next;          /* next channel to claim */
  int ncopied;       /* channels [0, ncopied) are on the host */
  int failed;
  pthread_mutex_t mu;
  pthread_cond_t cv;
} writer_state;

static void *writer_main(void *arg)
{
  writer_state *w = (writer_state *)arg;
  for (;;) {
    pthread_mutex_lock(&w->mu);
    while (w->next >= w->ncopied && w->next < w->n_nu && !w->failed) pthread_cond_wait(&w->cv, &w->mu);
    if (w->next >= w->n_nu || w->failed) {
      pthread_mutex_unlock(&w->mu);
      return NULL;
    }
    const int inu = w->next++;
    pthread_mutex_unlock(&w->mu);
    char fname[600];
    snprintf(fname, sizeof(fname), "%s_%03d.fits", w->par->prefix_out, inu + 1);
    if (gh_write_healpix_fits(fname, w->maps + (size_t)w->npix * inu, w->par->n_side)) {
      pthread_mutex_lock(&w->mu);
      w->failed = 1;
      pthread_cond_broadcast(&w->cv);
      pthread_mutex_unlock(&w->mu);
      return NULL;
    }
  }
}

/* copy + write; returns 0 or -1 (error printed), fills the copy and total times */
static int write_maps_pipelined(crime_gh_t *h, const gh_params *par, double *ms_copy, double *ms_all)
{
  const long npix = 12 * par->n_side * par->n_side;
  const size_t nmaps = (size_t)par->n_nu * npix;
  double t0 = wall_ms();
  float *maps = NULL;
  if (crime_gh_host_alloc(nmaps * sizeof(float), (void **)&maps)) {
    fprintf(stderr, "CRIME: %s\n", crime_gh_last_error(h));
    return -1;
  }
  if (gh_write_nutable(par, par->prefix_out)) {
    fprintf(stderr, "CRIME: could not write %s_nuTable.dat\n", par->prefix_out);
    crime_gh_host_free(maps);
    return -1;
  }
  writer_state w;
  w.par = par;
  w.maps = maps;
  w.npix = npix;
  w.n_nu = par->n_nu;
  w.next = 0;
  w.ncopied = 0;
  w.failed = 0;
  pthread_mutex_init(&w.mu, NULL);
  pthread_cond_init(&w.cv, NULL);
  long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
  int nthr = (int)(ncpu < 2 ? 1 : (ncpu > 16 ? 16 : ncpu - 1));
  if (nthr > par->n_nu) nthr = par->n_nu;
  pthread_t thr[16];
  int started = 0;
  for (; started < nthr; started++)
    if (pthread_create(&thr[started], NULL, writer_main, &w)) break;
  const int block = 8;
  int rc = 0;
  for (int c0 = 0; c0 < par->n_nu && !rc; c0 += block) {
    const int nc = par->n_nu - c0 < block ? par->n_nu - c0 : block;
    rc = crime_gh_copy_maps_channels(h, c0, nc, maps + (size_t)c0 * npix);
    pthread_mutex_lock(&w.mu);
    if (rc) w.failed = 1;
    else w.ncopied = c0 + nc;
    pthread_cond_broadcast(&w.cv);
    pthread_mutex_unlock(&w.mu);
  }
  *ms_copy = wall_ms() - t0;
  if (started == 0 && !rc) writer_main(&w); /* no thread could start: write here */
  for (int i = 0; i < started; i++) pthread_join(thr[i], NULL);
  pthread_mutex_destroy(&w.mu);
  pthread_cond_destroy(&w.cv);
  crime_gh_host_free(maps);
  *ms_all = wall_ms() - t0;
  if (rc) {
    fprintf(stderr, "CRIME: %s\n", crime_gh_last_error(h));
    return -1;
  }
  if (w.failed) {
    fprintf(stderr, "CRIME: could not write the maps with prefix %s\n", par->prefix_out);
    return -1;
  }
  return 0;
}

int main(int argc, char **argv)
{
  if (argc != 2) {
    fprintf(stderr, "Usage: ./GetHI file_name\n");
    return 0;
  }
  setbuf(stdout, NULL);
  printf("\n|-------------------------------------------------|\n");
  printf("|                      GetHI                      |\n");
  printf("|-------------------------------------------------|\n\n");
  double t_all = wall_ms();

  static gh_params par;
  gh_params_defaults(&par);
  printf("*** Reading run parameters \n");
  if (gh_params_read(&par, argv[1]) || gh_params_setup(&par)) {
    fprintf(stderr, "CRIME: %s\n", par.err);
    return 1;
  }
  printf("  Original sigma8=%lf\n", par.sigma8_in);
  gh_params_print_summary(&par, stdout);
  printf("Seed : %u\n", par.seed);

  crime_gh_desc desc;
  gh_params_to_desc(&par, &desc);
  crime_gh_t *h = NULL;
  int rc = crime_gh_create(&desc, &h);
  if (rc) {
    fprintf(stderr, "CRIME: %s\n", crime_gh_last_error(NULL));
    gh_params_free(&par);
    return 1;
  }

  crime_gh_stats st;
  printf("*** Creating Gaussian density field \n");
  rc = crime_gh_create_d_and_vr_fields(h, &st);
  if (rc) goto fail;
  printf("Creating Fourier-space density and velocity potential, transforming along kz \n");
  printf(">    Relative time ellapsed %.1lf ms\n", st.ms_fill_fft_z);
  printf("Transforming along ky and kx, normalizing \n");
  printf(">    Relative time ellapsed %.1lf ms\n", st.ms_fft_y + st.ms_fft_x);
  printf(" <d>=%.3lE, <d^2>=%.3lE\n", st.mean_gauss, sqrt(st.sigma2_gauss));
  printf(">    Relative time ellapsed %.1lf ms\n\n", st.ms_sigma2);

  printf("*** Gettin' HI, making maps\n");
  rc = crime_gh_mk_T_maps(h, &st);
  if (rc) goto fail;
  printf(">    Relative time ellapsed %.1lf ms\n\n", st.ms_maps_clear + st.ms_pixelize);

  {
    printf("*** Writing files %s_###.fits\n", par.prefix_out);
    double ms_copy = 0, ms_write = 0;
    if (write_maps_pipelined(h, &par, &ms_copy, &ms_write)) {
      crime_gh_destroy(h);
      gh_params_free(&par);
      return 1;
    }
    printf("  maps to the host in %.1lf ms (pinned, blocks of 8 channels), written by then + %.1lf ms\n", ms_copy,
           ms_write - ms_copy);
    printf(">    Relative time ellapsed %.1lf ms\n", ms_write);
  }
  if (par.do_psources) {
    /* src/main_gh.c:55-66 and the <prefix>_ps_###.fits files of src/io_gh.c:122-127 */
    crime_gh_ps_stats ps;
    printf("*** Getting point sources, making source maps\n");
    rc = crime_gh_mk_psources_maps(h, &ps);
    if (rc) goto fail;
    printf("  There will be %lld particles in total \n", (long long)ps.n_sources);
    printf(">    Relative time ellapsed %.1lf ms\n", ps.ms_sources + ps.ms_maps);
    size_t npix = (size_t)12 * par.n_side * par.n_side, nmaps = (size_t)par.n_nu * npix;
    double *ps64 = (double *)malloc(nmaps * sizeof(double));
    float *ps32 = (float *)malloc(nmaps * sizeof(float));
    if (!ps64 || !ps32 || crime_gh_copy_ps_maps(h, ps64)) {
      free(ps64);
      free(ps32);
      fprintf(stderr, "CRIME: could not fetch the point-source maps\n");
      crime_gh_destroy(h);
      gh_params_free(&par);
      return 1;
    }
    for (size_t i = 0; i < nmaps; i++) ps32[i] = (float)ps64[i];
    printf("*** Writing single slice files %s_ps_###.fits\n", par.prefix_out);
    for (int inu = 0; inu < par.n_nu && !rc; inu++) {
      char fname[600];
      snprintf(fname, sizeof(fname), "%s_ps_%03d.fits", par.prefix_out, inu + 1);
      rc = gh_write_healpix_fits(fname, ps32 + npix * inu, par.n_side);
    }
    free(ps64);
    free(ps32);
    if (rc) {
      fprintf(stderr, "CRIME: could not write the point-source maps\n");
      crime_gh_destroy(h);
      gh_params_free(&par);
      return 1;
    }
  }
  printf(">    Total time ellapsed %.1lf ms\n", wall_ms() - t_all);
  printf("\n|-------------------------------------------------|\n\n");
  crime_gh_destroy(h);
  gh_params_free(&par);
  return 0;

fail:
  fprintf(stderr, "CRIME: %s\n", crime_gh_last_error(h));
  crime_gh_destroy(h);
  gh_params_free(&par);
  return 1;
}
