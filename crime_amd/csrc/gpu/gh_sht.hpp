// Declarations shared by the HEALPix analysis kernels (k_sht.hip) and their runtime (crime_sht.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace crime {

// one northern RING ring (index r = ir - 1, ir = 1 .. 2 nside) and its southern mirror 4 nside - ir
struct ShtRing {
  int64_t first_n, first_s;  // first pixel of the ring and of its mirror (equal on the equator)
  int nphi, s;               // pixels in the ring; phi_j = pi (2 j + s) / nphi
};

size_t ring_dft_lds_bytes(int n, int lmax);
hipError_t sht_set_lds_limits(size_t bytes);
// compiled Stockham lengths of the equatorial rings (4 Nside for Nside 32 .. 1024)
#define CRIME_FOR_EACH_RING_FFT_L(X) X(128) X(256) X(512) X(1024) X(2048) X(4096)
// 4 Nside when the equatorial rings have a compiled Stockham plan, else 0
int ring_fft_length(int64_t nside);

// convolution lengths of the Bluestein ring transforms (rings of 32 .. 4096 pixels)
#define CRIME_FOR_EACH_BLU_M(X) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

// how a contiguous range of rings is transformed
enum { SHT_RING_SPLIT = 0, SHT_RING_FFT = 1, SHT_RING_BLUESTEIN = 2 };

struct ShtLaunch {
  int kind, len;                 // SHT_RING_*; FFT length (4 Nside) or Bluestein M
  int r0, nr;                    // rings [r0, r0 + nr)
  int nmaps;
  int64_t npix;
  const ShtRing *rings;
  int R, RP, lmax, ldg;
  double w;                      // 4 pi / Npix
  size_t lds;                    // dynamic LDS of the split DFT
  double *G;
  const void *tw;                // exp(+2 pi i m / len), m < len
  const int64_t *woff;           // Bluestein: per-ring offsets of the chirp and filter tables
  const void *wtab;
  const int64_t *boff;
  const void *B;
  int mps;                       // maps per entry of the map table (analysis)
  hipStream_t s;
};
// analysis: map c of the nmaps is mtab[c / mps] + (c % mps) Npix (mtab: device table of map sets)
hipError_t launch_rings(const ShtLaunch &a, const void *const *mtab, int fp64);
// synthesis: the same ring groups, from G into FP32 or FP64 maps
hipError_t launch_syn_rings(const ShtLaunch &a, void *maps, int fp64);
// B of the Bluestein rings [r0, r0 + nr) of one convolution length M
hipError_t launch_blu_setup(int M, const ShtRing *rings, int r0, int nr, const int64_t *woff, const void *wtab,
                            const int64_t *boff, void *B, const void *tw, hipStream_t s);
hipError_t launch_legendre(const double2 *coef, const int64_t *off, int m0, int m1, int lmax, const double *zs,
                           const double *sts, int R, int RP, double *lam, hipStream_t s);
// tiles: (m, parity, row tile, rows of that parity)
// rows (l values of one parity) per alm GEMM tile
#ifndef CRIME_ALM_ROWS
#define CRIME_ALM_ROWS 128
#endif
constexpr int ALM_ROWS = CRIME_ALM_ROWS;
hipError_t launch_alm_gemm(const double *lam, const double *G, const int4 *tiles, int ntiles, const int64_t *off,
                           int m0, int lmax, int RP, int ldg, int ncols, double *alm, int ldo, hipStream_t s);
// cross spectra within each of nsets consecutive sets of mps maps: cls[set][a][b][l]
hipError_t launch_cls_sets(const double *alm, int lmax, int nsets, int mps, int ldo, double *cls, hipStream_t s);
hipError_t launch_cls(const double *alm, int lmax, int nmaps, int ldo, int cross, double *cls, hipStream_t s);
hipError_t launch_cls_pairs(const double *alm, int lmax, int ldo, int npairs, const int *pairs, double *cls,
                            hipStream_t s);
hipError_t launch_syn_gemm(const double *lam, const double *alm, double *G, const int4 *tiles, int ntiles,
                           const int64_t *off, int m0, int lmax, int R, int RP, int ldg, int ncols, int ldo,
                           hipStream_t s);
hipError_t launch_alm_pack(const double *in, int64_t nalm, int nmaps, int ldo, double *alm, hipStream_t s);
hipError_t launch_alm_beam(double *alm, int lmax, int nmaps, int ldo, const double *sig2, hipStream_t s);
// second Philox key word of the JoinT-like noise ("JONT")
constexpr uint32_t JT_NOISE_KEY = 0x4A4F4E54u;
hipError_t launch_noise(void *maps, int fp64, int64_t npix, int nmaps, const double *sig2, uint32_t seed,
                        hipStream_t s);
hipError_t launch_mask(void *maps, int fp64, int64_t npix, int nmaps, const uint8_t *mask, hipStream_t s);
hipError_t launch_alm_unpack(const double *alm, int64_t nalm, int nmaps, int ldo, double *out, hipStream_t s);

}  // namespace crime
