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
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "crime_gh.h"
#include "gethi_host.h"

static double wall_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
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
    size_t nmaps = (size_t)par.n_nu * 12 * par.n_side * par.n_side;
    float *maps = (float *)malloc(nmaps * sizeof(float));
    if (!maps) {
      fprintf(stderr, "CRIME: out of memory for %zu map pixels\n", nmaps);
      crime_gh_destroy(h);
      gh_params_free(&par);
      return 1;
    }
    rc = crime_gh_copy_maps(h, maps);
    if (rc) {
      free(maps);
      goto fail;
    }
    printf("*** Writing files %s_###.fits\n", par.prefix_out);
    double t0 = wall_ms();
    if (gh_write_maps(&par, maps)) {
      fprintf(stderr, "CRIME: could not write the maps with prefix %s\n", par.prefix_out);
      free(maps);
      crime_gh_destroy(h);
      gh_params_free(&par);
      return 1;
    }
    printf(">    Relative time ellapsed %.1lf ms\n", wall_ms() - t0);
    free(maps);
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
