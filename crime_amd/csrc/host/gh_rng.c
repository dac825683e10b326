/*
 * MT19937 streams of the GetHI host side.
 *
 * The reference draws every random number from GSL's mt19937 (init_rng/rng_01, src/common.c:133-146):
 * Matsumoto & Nishimura's generator with the 2002 initialisation (mt[i] = 1812433253 (mt[i-1] ^
 * (mt[i-1] >> 30)) + i), GSL's rule that seed 0 means 4357, and gsl_rng_uniform = x / 2^32.
 * This is an independent implementation of that published algorithm.
 */
#include <math.h>
#include <string.h>

#include "gethi_host.h"

#define MT_N 624
#define MT_M 397

void gh_mt_seed(gh_mt *s, uint32_t seed)
{
  if (seed == 0) seed = 4357;
  s->mt[0] = seed;
  for (int i = 1; i < MT_N; i++)
    s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
  s->mti = MT_N;
}

static void mt_twist(gh_mt *s)
{
  for (int i = 0; i < MT_N; i++) {
    uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % MT_N] & 0x7fffffffu);
    uint32_t v = s->mt[(i + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    s->mt[i] = v;
  }
  s->mti = 0;
}

uint32_t gh_mt_next(gh_mt *s)
{
  if (s->mti >= MT_N) mt_twist(s);
  uint32_t y = s->mt[s->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double gh_mt_uniform(gh_mt *s) { return gh_mt_next(s) / 4294967296.0; }

void gh_unit_modes_mt(int n, uint32_t seed, double *out)
{
  /* fill order of create_density_and_velpot_fourier with one thread (src/fourier.c:253-302):
   * kz slowest, then ky, then kx in [0, N/2]; no draws for k = 0 */
  gh_mt s;
  gh_mt_seed(&s, seed);
  int nh = n / 2 + 1;
  for (int ii = 0; ii < n; ii++) {
    for (int jj = 0; jj < n; jj++) {
      for (int kk = 0; kk < nh; kk++) {
        size_t idx = (size_t)kk + (size_t)nh * ((size_t)jj + (size_t)n * ii);
        if (ii == 0 && jj == 0 && kk == 0) {
          out[2 * idx] = 0;
          out[2 * idx + 1] = 0;
          continue;
        }
        double phase = 2 * M_PI * gh_mt_uniform(&s);
        double u = gh_mt_uniform(&s);
        double mod = sqrt(-log(1 - u));
        out[2 * idx] = mod * cos(phase);
        out[2 * idx + 1] = mod * sin(phase);
      }
    }
  }
}

int gh_subcell_offsets(int n_subcell, double lcell, uint32_t seed, double *out)
{
  if (n_subcell <= 0) {
    /* src/pixelize.c:157-164 */
    gh_mt s;
    gh_mt_seed(&s, seed);
    for (int ii = 0; ii < GH_N_SUBPART; ii++) {
      out[3 * ii + 0] = lcell * (gh_mt_uniform(&s) - 0.5);
      out[3 * ii + 1] = lcell * (gh_mt_uniform(&s) - 0.5);
      out[3 * ii + 2] = lcell * (gh_mt_uniform(&s) - 0.5);
    }
    return GH_N_SUBPART;
  }
  int k = n_subcell, n = 0;
  for (int iz = 0; iz < k; iz++)
    for (int iy = 0; iy < k; iy++)
      for (int ix = 0; ix < k; ix++) {
        out[3 * n + 0] = lcell * ((ix + 0.5) / k - 0.5);
        out[3 * n + 1] = lcell * ((iy + 0.5) / k - 0.5);
        out[3 * n + 2] = lcell * ((iz + 0.5) / k - 0.5);
        n++;
      }
  return n;
}
