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
hipError_t launch_ring_dft(const void *maps, int fp64, int nmaps, int64_t npix, const ShtRing *rings, int R, int RP,
                           int lmax, int ldg, double w, size_t lds, double *G, hipStream_t s);
hipError_t launch_legendre(const double2 *coef, const int64_t *off, int m0, int m1, int lmax, const double *zs,
                           const double *sts, int R, int RP, double *lam, hipStream_t s);
// tiles: (m, parity, row tile, rows of that parity)
hipError_t launch_alm_gemm(const double *lam, const double *G, const int4 *tiles, int ntiles, const int64_t *off,
                           int m0, int lmax, int RP, int ldg, int ncols, double *alm, int ldo, hipStream_t s);
hipError_t launch_cls(const double *alm, int lmax, int nmaps, int ldo, int cross, double *cls, hipStream_t s);
hipError_t launch_cls_pairs(const double *alm, int lmax, int ldo, int npairs, const int *pairs, double *cls,
                            hipStream_t s);
hipError_t launch_alm_unpack(const double *alm, int64_t nalm, int nmaps, int ldo, double *out, hipStream_t s);

}  // namespace crime
