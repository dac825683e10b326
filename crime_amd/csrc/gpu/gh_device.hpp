// Device-side arithmetic of the GetHI density-to-sky path for gfx950.
//
// Everything that decides an integer (table index, channel, HEALPix pixel) is FP64 with FMA
// contraction disabled, so that it performs the same IEEE operations in the same order as the
// host-side restatement of the reference (src/cosmo.c:52-86, src/pixelize.c:28-55, 190-228).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crime {

template <typename T>
struct __align__(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by +i
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.y, a.x}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }

// Workgroups are dealt to the 8 XCDs round robin (b -> XCD b % 8).  This bijection of [0, nb)
// hands each XCD one contiguous run of logical blocks, so neighbours share that XCD's L2.
__device__ __forceinline__ int xcd_swizzle(int b, int nb)
{
  const int q = nb / 8, rem = nb % 8, x = b % 8;
  return x * q + (x < rem ? x : rem) + b / 8;
}

// ---------------------------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): counter-based, so mode k can be drawn by
// any thread in any order and the mode array does not depend on the launch geometry.
// ---------------------------------------------------------------------------------------------
struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c.v[0];
    uint64_t p1 = (uint64_t)M1 * c.v[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u32x4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    c = o;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53-bit uniform in [0,1) from two 32-bit words
__host__ __device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)(m & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------------------------
// uniform-r table lookups (src/cosmo.c:52-86)
// ---------------------------------------------------------------------------------------------
struct RTables {
  const double *z;    // z_arr_r2z
  const double *r;    // r_arr_r2z
  const double *d;    // growth_d_arr
  const double *v;    // growth_v_arr
  const double *f;    // 0.008 (1 + z)^0.6 at the nodes (the HI fraction of FP32 grids' map stage)
  const double *b;    // 0.904 + 0.135 (1 + z)^1.696 at the nodes (the reference bias, likewise)
  double idr;
  double r_last;      // r_arr_r2z[NZ-1]
  double z_last, d_last, v_last, f_last, b_last;
};

// Branch-free: the node index is clamped into the table for every r (NaN gives node 0), so the three
// loads are always in bounds and are issued together, and the clamped cases are selected afterwards.
// For 0 < r < r_last the node is (int)(r idr), at most NZ - 2 (r within one ulp of the last node).
__device__ __forceinline__ double tab_lerp(const double *y, const double *rr, double idr, double r,
                                           double r_last, double y_last, double y_at0)
{
#pragma clang fp contract(off)
  const int ir = (int)fmin(fmax(r * idr, 0.0), (double)(5001 - 2));
  const double y0 = y[ir], y1 = y[ir + 1], r0 = rr[ir];
  const double v = y0 + (y1 - y0) * (r - r0) * idr;
  return r <= 0 ? y_at0 : (r >= r_last ? y_last : v);
}

__device__ __forceinline__ double z_of_r(const RTables &t, double r)
{
  return tab_lerp(t.z, t.r, t.idr, r, t.r_last, t.z_last, 0.0);
}
__device__ __forceinline__ double dgrowth_of_r(const RTables &t, double r)
{
  return tab_lerp(t.d, t.r, t.idr, r, t.r_last, t.d_last, 1.0);
}
__device__ __forceinline__ double vgrowth_of_r(const RTables &t, double r)
{
  return tab_lerp(t.v, t.r, t.idr, r, t.r_last, t.v_last, 1.0);
}
// the node tables of the HI fraction and of the reference bias, interpolated like z(r): within 1e-8
// (relative) of the formulas at z(r) for tables of the reference's resolution
__device__ __forceinline__ double fhi_of_r(const RTables &t, double r)
{
  return tab_lerp(t.f, t.r, t.idr, r, t.r_last, t.f_last, 0.008);
}
__device__ __forceinline__ double bref_of_r(const RTables &t, double r)
{
  return tab_lerp(t.b, t.r, t.idr, r, t.r_last, t.b_last, 0.904 + 0.135);
}

// ---------------------------------------------------------------------------------------------
// channel of a frequency.  TABLE mode: the unique i with nu0[i] <= nu < nuf[i] (edges[i] = nu0[i],
// edges[n] = nuf[n-1]); get_inu's walk (src/pixelize.c:28-55) converges to this index from any
// start because the bins are contiguous.  UNIFORM mode: (int)(inv_dnu*(nu-nu_min)) (:369).
// Returns -1 when no channel receives the sample.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int channel_of(double nu, int n_nu, int nu_mode, const double *edges,
                                          double nu_min, double inv_dnu)
{
#pragma clang fp contract(off)
  if (nu_mode == 1) {
    // (int)t lands in [0, n_nu) exactly when -1 < t < n_nu; testing t first keeps the cast defined
    double t = inv_dnu * (nu - nu_min);
    if (!(t > -1.0 && t < (double)n_nu)) return -1;
    return (int)t;
  }
  if (!(nu >= edges[0]) || !(nu < edges[n_nu])) return -1;
  int lo = 0, hi = n_nu;  // edges[lo] <= nu < edges[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (nu >= edges[mid]) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Per-bucket channel record of the TABLE mode.  Bucket j covers the frequencies with
// (int)((nu - g_lo) g_inv) = j; c is the channel holding g_lo + j / g_inv, e0..e2 are edges[c..c+2]
// (e2 = e1 for the last channel) and pf0, pf1 the mK factors of channels c and c + 1.  One record
// (one 64-byte line) decides the channel and its factor of almost every sample with a single
// dependent load; frequencies outside [e0, e2) walk the edges from c.
struct __align__(64) ChanRec {
  double e0, e1, e2, pf0, pf1;
  int c, pad0;
  double pad1;
};

// bucket of a frequency: (int)((nu - g_lo) g_inv), clamped into the table (NaN gives bucket 0)
__device__ __forceinline__ int channel_bucket(double nu, int n_guess, double g_lo, double g_inv)
{
#pragma clang fp contract(off)
  const double t = (nu - g_lo) * g_inv;
  return (int)fmin(fmax(t, 0.0), (double)(n_guess - 1));
}

// channel_of for nu_mode 0 given the record r of nu's bucket, and the channel's mK factor; -1 outside
// [nu_lo, nu_hi) = [edges[0], edges[n_nu])
__device__ __forceinline__ int channel_of_rec(double nu, const ChanRec &r, int n_nu, const double *edges,
                                              const double *prefac, double nu_lo, double nu_hi, double &pf)
{
  int inu;
  if (nu >= r.e0 && nu < r.e1) {
    inu = r.c;
    pf = r.pf0;
  }
  else if (nu >= r.e1 && nu < r.e2) {
    inu = r.c + 1;
    pf = r.pf1;
  }
  else {
    inu = r.c;
    while (inu > 0 && nu < edges[inu]) inu--;
    while (inu < n_nu - 1 && nu >= edges[inu + 1]) inu++;
    pf = prefac[inu];
  }
  if (!(nu >= nu_lo) || !(nu < nu_hi)) inu = -1;
  return inu;
}

// ---------------------------------------------------------------------------------------------
// RING pixel of a (non-normalised) direction, Gorski et al. 2005 (ApJ 622:759) section 4; vlen is
// the vector's length sqrt(x*x+y*y+z*z).  Near the poles (|cos theta| > 0.99) sin theta is taken
// from sqrt(x*x+y*y)/vlen to avoid the cancellation in 1-|cos theta|.
// ---------------------------------------------------------------------------------------------
// v mod m for v within a few multiples of m (no integer division)
__device__ __forceinline__ int imod_small(int v, int m)
{
  while (v >= m) v -= m;
  while (v < 0) v += m;
  return v;
}

// Integer work is 32-bit: for Nside <= 8192 every index below is < 12 Nside^2 < 2^31, and the
// truncations are of non-negative values < 5 Nside, where (int) and (int64_t) casts agree.
__device__ __forceinline__ int64_t vec2pix_ring(int64_t nside64, double x, double y, double z, double vlen)
{
#pragma clang fp contract(off)
  const double twopi = 6.283185307179586476925286766559;
  const double inv_halfpi = 0.636619772367581343075535053490057;
  const double twothird = 2.0 / 3.0;
  const int nside = (int)nside64;
  double cth = z / vlen;
  double za = fabs(cth);
  double phi = atan2(y, x);
  double ph = phi;
  if (ph < 0) {
    ph = ph + twopi;
    if (ph == twopi) ph = 0;
  }
  double tt = ph * inv_halfpi;  // [0,4)
  double dn = (double)nside;
  if (za <= twothird) {
    double temp1 = dn * (0.5 + tt);
    double temp2 = dn * cth * 0.75;
    int jp = (int)(temp1 - temp2);  // in [0, 5 Nside]
    int jm = (int)(temp1 + temp2);
    int ir = nside + 1 + jp - jm;
    int kshift = 1 - (ir & 1);
    int ip = (jp + jm - nside + kshift + 1) / 2;  // in [-Nside/2, 5 Nside + 1]
    ip = imod_small(ip, 4 * nside);
    return (int64_t)(nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip);
  }
  double tp = tt - (double)(int)tt;
  double tmp;
  if (za > 0.99) {
    double sth = sqrt(x * x + y * y) / vlen;
    tmp = dn * sth / sqrt((1.0 + za) / 3.0);
  }
  else {
    tmp = dn * sqrt(3.0 * (1.0 - za));
  }
  int jp = (int)(tp * tmp);
  int jm = (int)((1.0 - tp) * tmp);
  int ir = jp + jm + 1;
  int ip = (int)(tt * (double)ir);  // in [0, 4 ir]
  ip = imod_small(ip, 4 * ir);
  if (cth > 0) return (int64_t)(2 * ir * (ir - 1) + ip);
  return (int64_t)(12 * nside * nside - 2 * ir * (ir + 1) + ip);
}
}  // namespace crime
