// Map stage of the GetHI density-to-sky path on gfx950: get_HI (src/grid_tools.c:103-153) fused
// with mk_T_maps (src/pixelize.c:150-261) and the radial-velocity stencil (src/fourier.c:307-373).
//
// One thread per cell reads (delta, phi) of the cell and phi of its six periodic neighbours, forms
// v_r, the lognormal HI mass and the RSD redshift shift in FP64, and scatters its n_sub samples into
// the (channel, RING pixel) maps with hardware FP32/FP64 atomics, already multiplied by the channel's
// mass-to-mK factor (the reference scales the finished maps, pixelize.c:246-261; the product of the
// factor with each sample differs from it only in rounding).  Cells whose whole sample cloud lies
// outside every channel -- decided with a conservative bound on r and the cell's exact RSD shift --
// skip the per-sample work; a sample that the reference would bin is never skipped.
//
// This file is compiled with -ffp-contract=off: positions, radii, redshifts, frequencies and the
// pixel arithmetic are the same IEEE operations, in the same order, as the reference's C.
#include "gh_kernels.hpp"

namespace crime {

__device__ __forceinline__ double bias_of(const MapParams &m, double z)
{
  if (m.bias_model == 1) return m.b0;
  if (m.bias_model == 2) return m.b0 + m.b1 * (1 + z);
  return 0.904 + 0.135 * pow(1 + z, 1.696);  // src/user_defined.c:32-35
}

__device__ __forceinline__ double fraction_hi(double z) { return 0.008 * pow(1 + z, 0.6); }

struct CellState {
  double x0, y0, z0;   // cell centre relative to the observer
  double r;            // |centre|
  double vr;           // radial velocity (fourier.c:368)
  double dz;           // RSD shift (grid_tools.c:142)
  double mass;         // HI mass (grid_tools.c:144)
  double zc;           // z_of_r(r) (grid_tools.c:138)
};

template <typename T>
__device__ __forceinline__ void cell_coords(const MapParams &m, size_t cell, int &ix, int &iy, int &iz)
{
  const size_t n = (size_t)m.n;
  ix = (int)(cell % n);
  size_t q = cell / n;
  iy = (int)(q % n);
  iz = (int)(q / n);
}

// centre, r and v_r (fourier.c:327-368).  STENCIL: phi lives in the imaginary part of F and the
// velocity is its central difference; SPECTRAL: vx is the imaginary part of F and (vy, vz) are G.
template <typename T>
__device__ __forceinline__ void cell_velocity(const MapParams &m, const cx<T> *__restrict__ F, int ix,
                                              int iy, int iz, CellState &s)
{
  const int n = m.n;
  s.x0 = m.dx * (ix + 0.5) - m.obs[0];
  s.y0 = m.dx * (iy + 0.5) - m.obs[1];
  s.z0 = m.dx * (iz + 0.5) - m.obs[2];
  const size_t nn = (size_t)n, n2 = nn * nn;
  const size_t row = nn * iy + n2 * iz;
  double v0, v1, v2;
  if (m.vmode) {
    const cx<T> g = ((const cx<T> *)m.G)[row + ix];
    v0 = (double)F[row + ix].y;
    v1 = (double)g.x;
    v2 = (double)g.y;
  }
  else {
    int ixp = ix == n - 1 ? 0 : ix + 1, ixm = ix == 0 ? n - 1 : ix - 1;
    int iyp = iy == n - 1 ? 0 : iy + 1, iym = iy == 0 ? n - 1 : iy - 1;
    int izp = iz == n - 1 ? 0 : iz + 1, izm = iz == 0 ? n - 1 : iz - 1;
    double px1 = (double)F[row + ixp].y, px0 = (double)F[row + ixm].y;
    double py1 = (double)F[ix + nn * iyp + n2 * iz].y, py0 = (double)F[ix + nn * iym + n2 * iz].y;
    double pz1 = (double)F[ix + nn * iy + n2 * izp].y, pz0 = (double)F[ix + nn * iy + n2 * izm].y;
    double idx = m.inv_dx;
    v0 = 0.5 * idx * (px1 - px0);
    v1 = 0.5 * idx * (py1 - py0);
    v2 = 0.5 * idx * (pz1 - pz0);
  }
  double rr2 = s.x0 * s.x0 + s.y0 * s.y0 + s.z0 * s.z0;
  double irr = 1. / sqrt(rr2);
  s.r = sqrt(rr2);
  s.vr = v0 * (s.x0 * irr) + v1 * (s.y0 * irr) + v2 * (s.z0 * irr);
}

// lognormal HI mass and RSD shift (grid_tools.c:138-146)
// the radial tables at the cell centre, split from the arithmetic so that a kernel can issue their loads
// with the grid's: z, D and V of r
// NODES (the map stage of FP32 grids): the HI fraction and the reference bias from their node tables
// (within 1e-8 relative of the formulas, far below FP32 map precision) instead of two FP64 pow per cell;
// the FP64 grids' map stage and every parity kernel evaluate the formulas.
struct CellTabs {
  double z, d, v, f, b;
};
template <bool NODES = false>
__device__ __forceinline__ void cell_tabs(const MapParams &m, const CellState &s, CellTabs &t)
{
  t.z = z_of_r(m.tab, s.r);
  t.d = dgrowth_of_r(m.tab, s.r);
  t.v = vgrowth_of_r(m.tab, s.r);
  if constexpr (NODES) {
    t.f = fhi_of_r(m.tab, s.r);
    t.b = m.bias_model == 0 ? bref_of_r(m.tab, s.r) : 0.0;
  }
}

template <bool NODES = false>
__device__ __forceinline__ void cell_hi_tabs(const MapParams &m, double delta, double sigma2, const CellTabs &t,
                                             CellState &s)
{
  const double redshift = t.z;
  s.zc = redshift;
  double bias, fhi;
  if constexpr (NODES) {
    bias = m.bias_model == 0 ? t.b : bias_of(m, redshift);
    fhi = t.f;
  }
  else {
    bias = bias_of(m, redshift);
    fhi = fraction_hi(redshift);
  }
  double gfd = t.d * bias;
  s.dz = m.rsd ? s.vr * t.v : 0.0;
  double dens_ln = exp(gfd * (delta - 0.5 * gfd * sigma2));
  s.mass = m.mass_prefac * fhi * dens_ln;
}

template <bool NODES = false>
__device__ __forceinline__ void cell_hi(const MapParams &m, double delta, double sigma2, CellState &s)
{
  CellTabs t;
  cell_tabs<NODES>(m, s, t);
  cell_hi_tabs<NODES>(m, delta, sigma2, t, s);
}

// The sample offsets are the same for every lane and never written by a kernel: read through the
// constant address space they are scalar loads (with a generic pointer the atomics in the loop keep the
// compiler from proving that, and every sample issued two vector loads).
typedef const __attribute__((address_space(4))) double const_double;
typedef const __attribute__((address_space(4))) int const_int;

// Channel and pixel of one sample (pixelize.c:207-222), in two halves so that the per-sample loop can
// interleave two samples: sample_pre computes the position, r, the frequency and loads the channel
// record (table mode); sample_post decides the channel and the pixel and returns the channel or -1,
// with the channel's mK factor in pf.  The pixel is computed only for binned samples: computing it for
// every sample needs more VGPRs than the occupancy allows.
struct SamplePre {
  double x, y, z, r, nu;
  ChanRec rec;  // table mode only
};

// from the sample's position (x, y, z) and the cell's RSD shift dz
template <bool TABLE>
__device__ __forceinline__ void sample_pre_pos(const MapParams &m, double x, double y, double z, double dz,
                                               SamplePre &q)
{
  q.x = x;
  q.y = y;
  q.z = z;
  q.r = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  const double redshift = z_of_r(m.tab, q.r) + dz;
  q.nu = 1420.40575177 / (1 + redshift);
  if constexpr (TABLE) q.rec = m.crec[channel_bucket(q.nu, m.n_guess, m.g_lo, m.g_inv)];
}

template <bool TABLE>
__device__ __forceinline__ void sample_pre(const MapParams &m, const CellState &s, int isub, SamplePre &q)
{
  const const_double *so = (const const_double *)m.sub_off;
  sample_pre_pos<TABLE>(m, s.x0 + so[3 * isub], s.y0 + so[3 * isub + 1], s.z0 + so[3 * isub + 2], s.dz, q);
}

template <bool TABLE>
__device__ __forceinline__ int sample_post(const MapParams &m, const SamplePre &q, int64_t &ipix, double &pf)
{
  int inu;
  if constexpr (TABLE) {
    inu = channel_of_rec(q.nu, q.rec, m.n_nu, m.edges, m.prefac, m.nu_lo, m.nu_hi, pf);
  }
  else {
    inu = channel_of(q.nu, m.n_nu, 1, m.edges, m.nu_min, m.inv_dnu);
    pf = m.prefac[inu < 0 ? 0 : inu];
  }
  if (!(q.r > 0)) inu = -1;  // a sample on the observer has no direction
  ipix = inu >= 0 ? vec2pix_ring(m.nside, q.x, q.y, q.z, q.r) : -1;
  return inu;
}

template <bool TABLE>
__device__ __forceinline__ int sample_bin_t(const MapParams &m, const CellState &s, int isub, int64_t &ipix,
                                            double &pf)
{
  SamplePre q;
  sample_pre<TABLE>(m, s, isub, q);
  return sample_post<TABLE>(m, q, ipix, pf);
}

__device__ __forceinline__ int sample_bin(const MapParams &m, const CellState &s, int isub, int64_t &ipix,
                                          double &pf)
{
  return m.nu_mode == 1 ? sample_bin_t<false>(m, s, isub, ipix, pf) : sample_bin_t<true>(m, s, isub, ipix, pf);
}

// ---------------------------------------------------------------------------------------------
// FP32 screen of the per-sample decisions.  A decision (channel; jp, jm, ip of the RING pixel) is read
// off FP32 arithmetic when its argument lies farther than a margin from the integer or edge that would
// change it; the margins (ScreenParams, set by the host from the error bounds of the FP32 operations)
// exceed the difference between the FP32 argument and the exact one, and the FP64 path's own
// difference from the exact one is far below them, so an accepted decision is the FP64 path's.  Any
// decision that misses its margin is taken by the FP64 path of sample_pre/sample_post/vec2pix_ring.
//
// The channel needs no memory access per sample.  Per cell: z(r) over the sample cloud (|r - r_c| <= h)
// is the parabola through the FP64 table lerp at r_c - h, r_c and r_c + h, whose distance from the lerp
// the host bounds and adds to the margin; and the decided intervals [lo, hi] of 1 + z of the three
// channels around the centre's channel (CellScreen).  Per sample: the parabola, 1 + z and three interval
// tests give the window slot k of the channel.  Margin tests |d - 1/2| <= half imply eps <= d <= 1 - eps
// (the host takes half 2^-24 below 1/2 - eps, more than the rounding of d - 1/2).
// ---------------------------------------------------------------------------------------------
struct SampleS {
  float xf, yf, zf, rf;   // position relative to the observer (the reference's FP64 sums) in FP32, and |p|
  int inu;                // channel, -1 none, -2 undecided
  int k;                  // window slot of the channel (0..2)
};

struct CellScreen {
  float rc, zcd, a, b;                 // z(r_c + t) + dz = zcd + t (a + b t); zcd NaN: no screen for the cell
  float lo0, hi0, lo1, hi1, lo2, hi2;  // decided intervals of 1 + z of channels c0, c0 + 1, c0 + 2 (empty
                                       // outside the channels), disjoint
  int c0;
};

// the FP64 position of sample isub of a cell (the reference's sums; recomputed where the FP64 path needs it)
__device__ __forceinline__ void sample_pos(const MapParams &m, const CellState &s, int isub, double &x, double &y,
                                           double &z)
{
  const const_double *so = (const const_double *)m.sub_off;
  x = s.x0 + so[3 * isub];
  y = s.y0 + so[3 * isub + 1];
  z = s.z0 + so[3 * isub + 2];
}

// decided interval of 1 + z for channel c; empty (+inf, -inf) below channel 0 and above the last
__device__ __forceinline__ float2 chan_window(const MapParams &m, int c)
{
  if (c < 0 || c >= m.n_nu) return make_float2(INFINITY, -INFINITY);
  return m.scr.chwin[c];
}

// true when no sample of any cell of tile (tx, ty, tz) can land in a channel, with every cell's RSD shift
// within dzb: the tile's box of cell centres, widened by h_sub, spans the radii [rlo, rhi], so every
// sample's redshift lies in [z(rlo) - dzb, z(rhi) + dzb] (z(r) is non-decreasing); the same test as
// cell_may_bin on that range.  dzb infinite (or NaN): false.
__device__ __forceinline__ bool tile_out_of_band(const MapParams &m, int tx, int ty, int tz, double dzb)
{
  if (!(dzb < INFINITY)) return false;
  const int t0[3] = {tx * TILE_X, ty * TILE_Y, tz * TILE_Z}, tn[3] = {TILE_X, TILE_Y, TILE_Z};
  double lo2 = 0, hi2 = 0;
  for (int d = 0; d < 3; d++) {
    const double a = m.dx * (t0[d] + 0.5) - m.obs[d], b = m.dx * (t0[d] + tn[d] - 0.5) - m.obs[d];
    const double nearest = (a <= 0 && b >= 0) ? 0.0 : fmin(fabs(a), fabs(b));
    const double farthest = fmax(fabs(a), fabs(b));
    lo2 += nearest * nearest;
    hi2 += farthest * farthest;
  }
  const double zl = z_of_r(m.tab, sqrt(lo2) - m.h_sub) - dzb, zh = z_of_r(m.tab, sqrt(hi2) + m.h_sub) + dzb;
  if (1 + zl > 0 && 1420.40575177 * (1 + 1e-10) < m.nu_lo * (1 + zl)) return true;
  if (1 + zh > 0 && 1420.40575177 * (1 - 1e-10) >= m.nu_hi * (1 + zh)) return true;
  return false;
}

// true when no tile of distance bin b can land a sample in a channel, with every cell's RSD shift within
// dzb: the bin's tiles have their centres at [b dx, (b + 1) dx) from the observer (the last bin: beyond
// b dx), so their cell centres at [b dx - tile_hd, (b + 1) dx + tile_hd]; widened by h_sub, the test of
// tile_out_of_band on that range.  dzb infinite (or NaN): false.
__device__ __forceinline__ bool bin_out_of_band(const MapParams &m, int b, double dzb)
{
  if (!(dzb < INFINITY)) return false;
  const double rlo = b * m.dx - m.tile_hd, rhi = b + 1 >= m.nbins ? INFINITY : (b + 1) * m.dx + m.tile_hd;
  const double zl = z_of_r(m.tab, rlo - m.h_sub) - dzb, zh = z_of_r(m.tab, rhi + m.h_sub) + dzb;
  if (1 + zl > 0 && 1420.40575177 * (1 + 1e-10) < m.nu_lo * (1 + zl)) return true;
  if (1 + zh > 0 && 1420.40575177 * (1 - 1e-10) >= m.nu_hi * (1 + zh)) return true;
  return false;
}

// false when no sample of the cell can land in a channel (conservative); zm, zp as cell_z_range
__device__ __forceinline__ bool cell_may_bin(const MapParams &m, const CellState &s, double zm, double zp)
{
  double zlo = zm + s.dz;
  double zhi = zp + s.dz;
  if (!(1 + zlo > 0) || !(1 + zhi > 0)) return true;
  // nu_top = NU_21 / (1 + zlo) below nu_lo or nu_bot = NU_21 / (1 + zhi) at or above nu_hi, multiplied out;
  // the 1e-10 slack covers every rounding of these and of the samples' frequencies
  if (1420.40575177 * (1 + 1e-10) < m.nu_lo * (1 + zlo)) return false;
  if (1420.40575177 * (1 - 1e-10) >= m.nu_hi * (1 + zhi)) return false;
  return true;
}

// per-cell part of the screen; zm, zp = z_of_r(r_c -+ h_sub) and s.zc = z_of_r(r_c) (TABLE mode)
__device__ __forceinline__ void screen_cell(const MapParams &m, const CellState &s, double zm, double zp,
                                            CellScreen &cs)
{
  const ScreenParams &c = m.scr;
  const bool ok = s.r > c.rc_lo && s.r < c.rc_hi;  // the cloud keeps off r <= r_min and the table's end
  cs.rc = (float)s.r;
  // a cell none of whose samples can reach a channel: z = inf gives nu = 0, no channel for every sample
  // (the FP64 path's decision); otherwise NaN leaves every sample of an unscreened cell to FP64
  cs.zcd = !cell_may_bin(m, s, zm, zp) ? INFINITY : (ok ? (float)(s.zc + s.dz) : __builtin_nanf(""));
  cs.a = (float)((zp - zm) * c.inv_2h);
  cs.b = (float)((zp - 2.0 * s.zc + zm) * c.inv_2h2);
  // the channel of the centre's frequency, from its bucket record; the window is that channel and its
  // neighbours (any window is correct: frequencies outside it are left to FP64)
  const double nu_c = 1420.40575177f * __builtin_amdgcn_rcpf((float)(1 + (s.zc + s.dz)));
  const ChanRec &r = m.crec[channel_bucket(nu_c, m.n_guess, m.g_lo, m.g_inv)];
  cs.c0 = (nu_c < r.e1 ? r.c : r.c + 1) - 1;
  const float2 w0 = chan_window(m, cs.c0), w1 = chan_window(m, cs.c0 + 1), w2 = chan_window(m, cs.c0 + 2);
  cs.lo0 = w0.x;
  cs.hi0 = w0.y;
  cs.lo1 = w1.x;
  cs.hi1 = w1.y;
  cs.lo2 = w2.x;
  cs.hi2 = w2.y;
}

// atan(t) on [0, 1] as t P(t^2), P of degree 7 fitted to atan in absolute error; evaluated with FMAs the
// result is within 1.4e-7 of atan(t) for every float t in [0, 1]
__device__ __forceinline__ float atan01(float t)
{
  const float s = t * t;
  float p = -0.004054564982652664f;
  p = fmaf(p, s, 0.021862955763936043f);
  p = fmaf(p, s, -0.055912334471940994f);
  p = fmaf(p, s, 0.0964219868183136f);
  p = fmaf(p, s, -0.1390863060951233f);
  p = fmaf(p, s, 0.19946566224098206f);
  p = fmaf(p, s, -0.33329859375953674f);
  p = fmaf(p, s, 0.9999993443489075f);
  return p * t;
}

// position, |p| and the channel from FP32 (TABLE mode): q.inu = channel, -1 (no channel) or -2 (left to
// FP64), q.k its window slot
__device__ __forceinline__ void screen_channel_at(const MapParams &m, double x, double y, double z, const CellScreen &cs,
                                                  SampleS &q)
{
  const ScreenParams &c = m.scr;
  q.xf = (float)x;
  q.yf = (float)y;
  q.zf = (float)z;
  // |p|^2 in FP64 (within 4e-16), its root in FP32: |p| within 1.25 ulp
  q.rf = __builtin_amdgcn_sqrtf((float)fma(x, x, fma(y, y, z * z)));
  const float t = q.rf - cs.rc;
  const float zs = fmaf(t, fmaf(t, cs.b, cs.a), cs.zcd);
  // the channel from 1 + z against the windows' bounds on 1 + z (the frequency needs no reciprocal); three
  // interval tests (a select among the struct's fields by k would leave the struct in memory)
  const float g = 1.0f + zs;
  const bool d0 = g >= cs.lo0 && g <= cs.hi0;
  const bool d1 = g >= cs.lo1 && g <= cs.hi1;
  const bool d2 = g >= cs.lo2 && g <= cs.hi2;
  const int k = d1 ? 1 : (d2 ? 2 : 0);
  q.inu = (d0 || d1 || d2) ? cs.c0 + k : ((g < c.out_lo || g > c.out_hi) ? -1 : -2);
  q.k = k;
}

__device__ __forceinline__ void screen_channel(const MapParams &m, const CellState &s, const CellScreen &cs, int isub,
                                               SampleS &q)
{
  double x, y, z;
  sample_pos(m, s, isub, x, y, z);
  screen_channel_at(m, x, y, z, cs, q);
}

// RING pixel from FP32, or -2 when a margin is not met (see vec2pix_ring for the FP64 form); npix < 2^31
__device__ __forceinline__ int screen_pixel(const MapParams &m, const SampleS &q)
{
  // every margin test feeds one flag (no early exits: a wave stays on one path), and the folds of ip into
  // [0, 4 ir) are one conditional add or subtract (|ip| stays within one period either side)
  const ScreenParams &c = m.scr;
  const int nside = (int)m.nside;
  const float N = (float)nside;
  const float cth = q.zf * __builtin_amdgcn_rcpf(q.rf);
  const float za = fabsf(cth);
  const float ax = fabsf(q.xf), ay = fabsf(q.yf);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  // on a coordinate plane atan2's signed-zero conventions decide: FP64
  bool ok = mn > 0.0f && fabsf(za - (2.0f / 3.0f)) >= c.eps_z;
  const float u = atan01(mn * __builtin_amdgcn_rcpf(mx)) * 0.63661977236758134f;  // quarter turns, [0, 0.5]
  const float w = ay > ax ? 1.0f - u : u;  // angle of (|x|, |y|) in quarter turns, [0, 1]
  ok = ok && fabsf(w - 0.5f) <= c.half_ax;
  // tt = phi / (pi/2) in [0, 4) = B + sw; the signs from the sign bits (x, y are nonzero wherever ok holds)
  const bool xneg = (int)__float_as_uint(q.xf) < 0, yneg = (int)__float_as_uint(q.yf) < 0;
  const int B = xneg ? 2 : (yneg ? 4 : 0);
  const float sw = xneg != yneg ? -w : w;
  const float hp = c.half_pix;
  int pix;
  if (za < 2.0f / 3.0f) {
    // jp, jm = floor(N (0.5 + tt) -+ 0.75 N cth) = K0 + floor(h + N sw -+ 0.75 N cth), N (0.5 + B) = K0 + h
    // products of operands below 2^24 (Nside <= 2^13) are 24-bit multiplies
    const int twoK = (int)__umul24((unsigned)nside, (unsigned)(1 + 2 * B));
    const float h = (twoK & 1) ? 0.5f : 0.0f;
    const float nw = N * sw, nc = 0.75f * N * cth;
    const float v1 = (nw - nc) + h, v2 = (nw + nc) + h;
    const float f1 = floorf(v1), f2 = floorf(v2);
    const float d1 = v1 - f1, d2 = v2 - f2;
    ok = ok && fabsf(d1 - 0.5f) <= hp && fabsf(d2 - 0.5f) <= hp;
    const int jp = (twoK >> 1) + (int)f1, jm = (twoK >> 1) + (int)f2;
    const int ir = nside + 1 + jp - jm;
    const int kshift = 1 - (ir & 1);
    // (jp + jm - Nside + kshift + 1) / 2: jp + jm > 2 N (1/2 + tt) - 2 >= N - 2, so the numerator is at least 0
    // and the floor below is C's truncating division; the even offset keeps the shift a floor also for the
    // values of samples that miss a margin (discarded)
    int ip = (int)((unsigned)(jp + jm - nside + kshift + 1 + 8 * nside) >> 1) - 4 * nside;
    ip = ip < 0 ? ip + 4 * nside : (ip >= 4 * nside ? ip - 4 * nside : ip);
    pix = nside * (nside - 1) * 2 + (int)__umul24((unsigned)(ir - 1), (unsigned)(4 * nside)) + ip;
  }
  else {
    // polar caps: tp = tt - (int)tt, jp = floor(tp tmp), jm = floor((1 - tp) tmp), tmp = N sin(theta) sqrt(3/(1+za))
    const float tp = sw > 0.0f ? w : 1.0f - w;
    const float sth = __builtin_amdgcn_sqrtf(fmaf(q.xf, q.xf, q.yf * q.yf)) * __builtin_amdgcn_rcpf(q.rf);
    const float tmp = N * sth * __builtin_amdgcn_rsqf((1.0f + za) * (1.0f / 3.0f));
    const float p1 = tp * tmp, p2 = (1.0f - tp) * tmp;
    const float g1 = floorf(p1), g2 = floorf(p2);
    const float d1 = p1 - g1, d2 = p2 - g2;
    const int ir = (int)g1 + (int)g2 + 1;
    // ip = floor(tt ir) = B ir + floor(sw ir)
    const float p3 = sw * (float)ir;
    const float g3 = floorf(p3);
    const float d3 = p3 - g3;
    ok = ok && fabsf(d1 - 0.5f) <= hp && fabsf(d2 - 0.5f) <= hp && fabsf(d3 - 0.5f) <= hp;
    int ip = (int)__umul24((unsigned)B, (unsigned)ir) + (int)g3;  // in [-ir, 5 ir]
    ip = ip < 0 ? ip + 4 * ir : (ip >= 4 * ir ? ip - 4 * ir : ip);
    const int tri = (int)__umul24((unsigned)ir, (unsigned)(ir - 1));  // ir (ir - 1); written 2 * tri, the
                                                                      // product is no longer 24-bit
    pix = cth > 0.0f ? tri + tri + ip : 12 * nside * nside - (tri + tri) - 4 * ir + ip;
  }
  return ok ? pix : -2;
}

// finish a screened sample in its step: undecided channels take the whole FP64 path, undecided pixels
// the FP64 vec2pix_ring; returns the channel (-1: not binned) with its pixel in ipix
__device__ __forceinline__ int resolve_sample(const MapParams &m, const CellState &s, int isub, SampleS &q,
                                              int64_t &ipix)
{
  if (q.inu == -2) {
    SamplePre p;
    double pf;
    sample_pre<true>(m, s, isub, p);
    q.inu = sample_post<true>(m, p, ipix, pf);
    return q.inu;
  }
  ipix = -1;
  if (q.inu < 0) return -1;
  ipix = screen_pixel(m, q);
  if (ipix == -2) {
    double x, y, z;
    sample_pos(m, s, isub, x, y, z);
    ipix = vec2pix_ring(m.nside, x, y, z, sqrt(x * x + y * y + z * z));
  }
  return q.inu;
}

// z of the cloud's extreme radii (the screen's parabola and cell_may_bin)
__device__ __forceinline__ void cell_z_range(const MapParams &m, const CellState &s, double &zm, double &zp)
{
  zm = z_of_r(m.tab, s.r - m.h_sub);
  zp = z_of_r(m.tab, s.r + m.h_sub);
}

// Per-cell state of sample_decide: the screen's when it is on (TABLE mode)
__device__ __forceinline__ void decide_cell(const MapParams &m, const CellState &s, CellScreen &cs)
{
  if (m.nu_mode == 0 && m.scr.on) {
    double zm, zp;
    cell_z_range(m, s, zm, zp);
    screen_cell(m, s, zm, zp, cs);
  }
}

// the channel and pixel every map kernel uses: screened in TABLE mode when the screen is on; pf is the
// channel's mK factor (prefac[inu], the value the FP64 path reads from the channel record)
__device__ __forceinline__ int sample_decide(const MapParams &m, const CellState &s, const CellScreen &cs, int isub,
                                             int64_t &ipix, double &pf)
{
  if (m.nu_mode == 0 && m.scr.on) {
    SampleS q;
    screen_channel(m, s, cs, isub, q);
    const int inu = resolve_sample(m, s, isub, q, ipix);
    pf = inu >= 0 ? m.prefac[inu] : 0.0;
    return inu;
  }
  return sample_bin(m, s, isub, ipix, pf);
}

// Add val to bin `key` of the workgroup's LDS table (open addressing, linear probing from slot h); a sample
// that finds no slot within the probe budget goes straight to the map in HBM.
template <typename T>
__device__ __forceinline__ void global_add(T *p, T v)
{
  if constexpr (std::is_same<T, unsigned long long>::value) atomicAdd(p, v);
  else unsafeAtomicAdd(p, v);
}

// Home slot of a bin: the 2^SEG keys of one aligned run of map addresses (one 2^SEG-word segment) start at
// a slot given by 32-bit Fibonacci hashing of the segment and follow each other from there, so that the
// flush, where the lanes of a wave read consecutive slots, hands the lanes of one segment to one atomic
// instruction: the memory system merges the lanes of an instruction that fall into one segment (runs of
// 4 lanes per 32-byte segment cost a third of scattered lanes).
template <int SLOTS, int SEG, typename K>
__device__ __forceinline__ unsigned home_slot(K key)
{
  constexpr int SHIFT = 32 - __builtin_ctz(SLOTS);
  const K seg = key >> SEG;
  unsigned folded;
  if constexpr (sizeof(K) == 8) folded = (unsigned)seg ^ (unsigned)((unsigned long long)seg >> 32);
  else folded = (unsigned)seg;
  return (((folded * 0x9E3779B1u) >> SHIFT) + ((unsigned)key & ((1u << SEG) - 1))) & (SLOTS - 1);
}

template <typename T, int SLOTS, typename K>
__device__ __forceinline__ void agg_add(K key, unsigned h, T val, K *keys, T *vals, T *__restrict__ maps,
                                        int probes = 64)
{
  const K EMPTY = (K)~(K)0;
  // the compare-and-swap claims an empty slot and returns the key of an occupied one: one LDS round trip
  // per probe (unrolled by 8 only: a collision is rare, and a longer chain costs scalar registers)
#pragma unroll 8
  for (int p = 0; p < probes; p++) {
    const K old = atomicCAS(&keys[h], EMPTY, key);
    if (old == EMPTY || old == key) {
      atomicAdd(&vals[h], val);
      return;
    }
    h = (h + 1) & (SLOTS - 1);
  }
  global_add(maps + key, val);
}

// agg_add in two halves, so that the first probe's round trip overlaps the next sample's arithmetic: start
// issues the compare-and-swap at the home slot, finish (called later, before any other LDS access of this
// thread to the table) adds the value there or continues probing from the next slot
template <typename K, typename T>
struct PendingAdd {
  K key, old;
  unsigned h;
  T val;
  bool on;
};

template <typename T, int SLOTS, typename K>
__device__ __forceinline__ void agg_start(PendingAdd<K, T> &pa, K key, unsigned h, T val, K *keys)
{
  pa.old = atomicCAS(&keys[h], (K)~(K)0, key);
  pa.key = key;
  pa.h = h;
  pa.val = val;
  pa.on = true;
}

template <typename T, int SLOTS, typename K>
__device__ __forceinline__ void agg_finish(PendingAdd<K, T> &pa, K *keys, T *vals, T *__restrict__ maps)
{
  if (!pa.on) return;
  pa.on = false;
  if (pa.old == (K)~(K)0 || pa.old == pa.key) atomicAdd(&vals[pa.h], pa.val);
  else agg_add<T, SLOTS, K>(pa.key, (pa.h + 1) & (SLOTS - 1), pa.val, keys, vals, maps, 63);
}

// Samples whose FP32 decision missed a margin, queued per workgroup in LDS and resolved by the FP64
// path after the workgroup's last sample, one entry per thread: a tile's few entries (about a dozen at
// 1024^3) then cost one wave's pass through the FP64 path instead of one pass per wave that queued any.
// A wave's step claims its entries with one LDS atomic; a sample that finds the queue full is resolved in
// its step.  dz is the cell's RSD shift, w its mass per sample, inu the channel or -2 when the channel is
// undecided.  n counts the claims, full or not.
constexpr int DEFER_CAP = 128;  // entries per workgroup
struct DeferQueue {
  double x[DEFER_CAP], y[DEFER_CAP], z[DEFER_CAP], dz[DEFER_CAP], w[DEFER_CAP];
  int inu[DEFER_CAP];
  int n;
};

// rank of this lane among the lanes set in mask below it
__device__ __forceinline__ int lane_rank(unsigned long long mask)
{
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// the FP64 decision of a deferred sample: the whole sample_pre/sample_post path when the channel is
// undecided, else the pixel only (the same operations as the unscreened path); the channel's factor
// prefac[inu] is the one its screen record holds
__device__ __forceinline__ int resolve_deferred(const MapParams &m, double x, double y, double z, double dz, int inu,
                                                int64_t &ipix, double &pf)
{
  if (inu == -2) {
    SamplePre p;
    sample_pre_pos<true>(m, x, y, z, dz, p);
    return sample_post<true>(m, p, ipix, pf);
  }
  pf = m.prefac[inu];
  ipix = vec2pix_ring(m.nside, x, y, z, sqrt(x * x + y * y + z * z));
  return inu;
}

// One workgroup = one TILE_X x TILE_Y x TILE_Z block of cells (one cell per thread).  Samples are
// summed per (channel, pixel) bin in LDS and every occupied bin is added to HBM once, so the
// global atomics scale with the bins a tile touches instead of with its samples.  SLOTS (2048, 1024 or
// 4096 bins) trades occupancy against probe length and overflow; SEG groups the bins of 2^SEG adjacent
// map words (home_slot); K is the bin key type (32-bit when n_nu npix < 2^32 - 1, which halves the key
// table).  In TABLE mode the FP32 screen decides the samples and its FP64 fallbacks are queued per workgroup.
// FX (deterministic scatter): every sample adds the integer round(s mass / n_sub) to the 64-bit bin of
// m.fxmaps instead of its FP mK value to maps; k_fx_finalize applies 1/s and the channel factor.
template <typename T, int SLOTS, int SEG, bool FX, typename K>
__device__ __forceinline__ void pixelize_body(const cx<T> *__restrict__ F, T *__restrict__ maps, const MapParams &m)
{
  using V = typename std::conditional<FX, unsigned long long, T>::type;
  __shared__ K keys[SLOTS];
  __shared__ V vals[SLOTS];
  __shared__ DeferQueue dq;
  V *__restrict__ vmaps;
  if constexpr (FX) vmaps = m.fxmaps;
  else vmaps = maps;
  const int n = m.n;
  const int t = threadIdx.x;
  // tiles run in order of distance from the observer, so the workgroups resident at one time cover a
  // thin shell and their atomics fall on a few channels of the maps.  2-D grid: a dispatch holds at
  // most 2^32 work-items per dimension.
  // the workgroups take the entries [first, last) of the order that k_tile_range kept; the rest exit here
  const size_t bid = (size_t)m.tile_range[0] + (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (bid >= (size_t)m.tile_range[1]) return;
  const uint32_t b = m.tile_order[bid];
  const int tx = (int)(b & 1023u), ty = (int)((b >> 10) & 1023u), tz = (int)(b >> 20);
  // whole tiles out of the band, by the grid's bound on the RSD shift (k_tile_cull): before any grid load
  if (m.tile_cull[bid]) return;
  const int ix = tx * TILE_X + (t % TILE_X), iy = ty * TILE_Y + ((t / TILE_X) % TILE_Y),
            iz = tz * TILE_Z + t / (TILE_X * TILE_Y);
  CellState s;
  CellTabs tb;
  double zm = 0, zp = 0, delta = 0;
  bool active = ix < n && iy < n && iz < n;
  if (active) {
    // every load of the cell (the grid's centre and stencil, the radial tables at r and r -+ h) depends only
    // on the tile: issued together, before the workgroup's first barrier
    const size_t cell = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
    delta = (double)F[cell].x;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_tabs<sizeof(T) == 4>(m, s, tb);
    cell_z_range(m, s, zm, zp);
    s.dz = m.rsd ? s.vr * tb.v : 0.0;
    active = cell_may_bin(m, s, zm, zp);
  }
  if (!__syncthreads_or(active)) return;  // no cell of the tile reaches a channel
  for (int i = t; i < SLOTS; i += 256) {
    keys[i] = (K)~(K)0;
    vals[i] = (V)0;
  }
  if (t == 0) dq.n = 0;
  __syncthreads();
  // the value a sample adds: its mK (FP) or its fixed-point mass (FX), from the cell's mass per sample
  auto value = [&](double w, double pf) -> V {
    if constexpr (FX) return (V)__double2ull_rn(w * m.fxs[0]);
    else return (V)(w * pf);
  };
  // npix = 12 Nside^2 < 2^32: the key is one 32 x 32 -> 64-bit multiply-add
  auto add_v = [&](int inu, int64_t ipix, V v) {
    if (inu >= 0 && (uint64_t)ipix < (uint64_t)m.npix) {
      const K key = (K)((unsigned long long)(unsigned)inu * (unsigned)m.npix + (unsigned)ipix);
      agg_add<V, SLOTS, K>(key, home_slot<SLOTS, SEG, K>(key), v, keys, vals, vmaps);
    }
  };
  if (active) {
    cell_hi_tabs<sizeof(T) == 4>(m, delta, m.sig[0], tb, s);
    double mass_sub = s.mass / m.n_sub;
    auto add = [&](int inu, int64_t ipix, double pf) { add_v(inu, ipix, value(mass_sub, pf)); };
    if (m.nu_mode == 0 && m.scr.on) {
      CellScreen cs;
      screen_cell(m, s, zm, zp, cs);
      // the value of a sample in each window channel
      auto win_value = [&](int c) -> V { return value(mass_sub, (c >= 0 && c < m.n_nu) ? m.prefac[c] : 0.0); };
      const V v0 = win_value(cs.c0), v1 = win_value(cs.c0 + 1), v2 = win_value(cs.c0 + 2);
      // the key of window slot k is kb + k npix (kb = c0 npix, modulo 2^bits(K) like every key)
      const K kb = (K)((long long)cs.c0 * m.npix), kn1 = (K)m.npix, kn2 = (K)(2 * m.npix);
      // the current run of samples with one key: its value is summed in registers, its slot claimed when it
      // starts and its sum added when it ends
      PendingAdd<K, V> run;
      run.on = false;
      for (int j = 0; j < m.n_sub; j++) {
        const const_double *o = (const const_double *)m.sub_offp + 3 * j;
        const double x = s.x0 + o[0], y = s.y0 + o[1], z = s.z0 + o[2];  // sample_pos's sums
        SampleS q;
        screen_channel_at(m, x, y, z, cs, q);
        const int p = q.inu >= 0 ? screen_pixel(m, q) : -1;
        const bool d = q.inu == -2 || p == -2;
        if (!d && q.inu >= 0 && (uint64_t)p < (uint64_t)m.npix) {
          const K key = kb + (q.k == 2 ? kn2 : (q.k == 1 ? kn1 : (K)0)) + (K)p;
          const V v = q.k == 2 ? v2 : (q.k == 1 ? v1 : v0);
          if (run.on && key == run.key) {
            run.val += v;
          }
          else {
            agg_finish<V, SLOTS, K>(run, keys, vals, vmaps);
            agg_start<V, SLOTS, K>(run, key, home_slot<SLOTS, SEG, K>(key), v, keys);
          }
        }
        const unsigned long long md = __ballot(d);
        if (md) {
          // the step's entries, claimed by the first lane that has one
          const int lead = __builtin_ctzll(md);
          int base = 0;
          if ((t & 63) == lead) base = atomicAdd(&dq.n, __popcll(md));
          base = __builtin_amdgcn_readlane(base, lead);
          if (d) {
            const int e = base + lane_rank(md);
            if (e < DEFER_CAP) {
              dq.x[e] = x;
              dq.y[e] = y;
              dq.z[e] = z;
              dq.dz[e] = s.dz;
              dq.w[e] = mass_sub;
              dq.inu[e] = q.inu;
            }
            else {  // queue full: the FP64 decision in this step
              int64_t ipix;
              double pf;
              const int c = resolve_deferred(m, x, y, z, s.dz, q.inu, ipix, pf);
              add(c, ipix, pf);
            }
          }
        }
      }
      agg_finish<V, SLOTS, K>(run, keys, vals, vmaps);
    }
    else {
      for (int isub = 0; isub < m.n_sub; isub++) {
        int64_t ipix;
        double pf;
        const int inu = sample_bin_t<false>(m, s, isub, ipix, pf);
        add(inu, ipix, pf);
      }
    }
  }
  __syncthreads();
  // the queued samples, one per thread (the first wave only, for up to 64 entries)
  const int nq = min(dq.n, DEFER_CAP);
  for (int e = t; e < nq; e += 256) {
    int64_t ipix;
    double pf;
    const int inu = resolve_deferred(m, dq.x[e], dq.y[e], dq.z[e], dq.dz[e], dq.inu[e], ipix, pf);
    if (inu >= 0) add_v(inu, ipix, value(dq.w[e], pf));
  }
  __syncthreads();
  for (int i = t; i < SLOTS; i += 256) {
    const K k = keys[i];
    if (k != (K)~(K)0) global_add(vmaps + k, vals[i]);
  }
}

// The range [first, last) of tile_order that the map kernel visits, for the current grid's RSD bound: from
// the first entry of the nearest distance bin that bin_out_of_band keeps to the last entry of the farthest
// (every tile when the bound is off).  One workgroup, before the map kernel.
__global__ __launch_bounds__(256) void k_tile_range(MapParams m)
{
  __shared__ int lo, hi;
  if (threadIdx.x == 0) {
    lo = 0x7fffffff;
    hi = -1;
  }
  __syncthreads();
  const double dzb = m.precull ? m.dz_per_phi * m.sig[2] : INFINITY;
  for (int b = threadIdx.x; b < m.nbins; b += 256)
    if (m.bin_start[b] < m.bin_start[b + 1] && !bin_out_of_band(m, b, dzb)) {
      atomicMin(&lo, b);
      atomicMax(&hi, b);
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool any = lo <= hi;
    m.tile_range[0] = any ? m.bin_start[lo] : 0u;
    m.tile_range[1] = any ? m.bin_start[hi + 1] : 0u;
  }
}

// The out-of-band flag of every entry of tile_order in [first, last), one tile per thread, with the grid's RSD
// bound (0 when the bound is off): the map kernel's workgroups read it instead of each evaluating the test.
__global__ __launch_bounds__(256) void k_tile_cull(MapParams m)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < (size_t)m.tile_range[0] || i >= (size_t)m.tile_range[1]) return;
  const uint32_t b = m.tile_order[i];
  const int tx = (int)(b & 1023u), ty = (int)((b >> 10) & 1023u), tz = (int)(b >> 20);
  m.tile_cull[i] = m.precull && tile_out_of_band(m, tx, ty, tz, m.dz_per_phi * m.sig[2]) ? 1 : 0;
}

template <typename T, int SLOTS, int SEG, bool FX, typename K>
__global__ __launch_bounds__(256) void k_pixelize(const cx<T> *__restrict__ F, T *__restrict__ maps, MapParams m)
{
  pixelize_body<T, SLOTS, SEG, FX, K>(F, maps, m);
}

// the same with the register budget of 7 waves per SIMD (72 VGPRs), for the tables whose LDS footprint
// lets 7 workgroups share a CU (32-bit keys, at most 2048 bins)
template <typename T, int SLOTS, int SEG, bool FX, typename K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_pixelize_w7(
    const cx<T> *__restrict__ F, T *__restrict__ maps, MapParams m)
{
  pixelize_body<T, SLOTS, SEG, FX, K>(F, maps, m);
}

// Deterministic scatter, step 1: the box's HI mass (the same cell_hi as the map kernel) summed in a fixed
// order -- a fixed grid of FX_PARTS blocks, each thread over a fixed stride of cells, then a tree in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_mass_partials(const cx<T> *__restrict__ F, MapParams m)
{
  __shared__ double red[256];
  const size_t ncell = (size_t)m.n * m.n * m.n;
  double acc = 0;
  for (size_t cell = (size_t)blockIdx.x * 256 + threadIdx.x; cell < ncell; cell += (size_t)FX_PARTS * 256) {
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    s.x0 = m.dx * (ix + 0.5) - m.obs[0];
    s.y0 = m.dx * (iy + 0.5) - m.obs[1];
    s.z0 = m.dx * (iz + 0.5) - m.obs[2];
    s.r = sqrt(s.x0 * s.x0 + s.y0 * s.y0 + s.z0 * s.z0);
    s.vr = 0;
    cell_hi<sizeof(T) == 4>(m, (double)F[cell].x, m.sig[0], s);
    acc += s.mass;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) m.mass_part[blockIdx.x] = red[0];
}

// step 2: total -> the power of two s with s total in (2^61, 2^62]; fxs = {s, 1/s}
__global__ __launch_bounds__(256) void k_fx_scale(MapParams m)
{
  __shared__ double red[256];
  double acc = 0;
  for (int i = threadIdx.x; i < FX_PARTS; i += 256) acc += m.mass_part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tot = red[0];
    int e = 0;
    (void)frexp(tot > 0 ? tot : 1.0, &e);  // tot in [2^(e-1), 2^e)
    m.fxs[0] = ldexp(1.0, 62 - e);          // s tot in [2^61, 2^62)
    m.fxs[1] = ldexp(1.0, e - 62);
  }
}

// step 4: maps[c][p] = (fx[c][p] / s) * prefac[c]
template <typename T>
__global__ __launch_bounds__(256) void k_fx_finalize(const unsigned long long *__restrict__ fx, T *__restrict__ maps,
                                                     MapParams m)
{
  const int c = blockIdx.y;
  const double f = m.prefac[c], inv = m.fxs[1];
  const size_t base = (size_t)c * m.npix;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < m.npix; p += (int64_t)gridDim.x * 256)
    maps[base + p] = (T)((double)fx[base + p] * inv * f);
}

// ---------------------------------------------------------------------------------------------
// Point sources.  get_point_sources (grid_tools.c:24-101) draws a Poisson number of sources per cell
// from the Gaussian delta; mk_psources_maps (pixelize.c:58-148) places each uniformly in its cell,
// draws its luminosity from the toy luminosity function by rejection (psources.c:130-152) and adds
//   T(nu) = FLUX2TEMP LUM2FLUX 4 pi L SED((1+z) nu) h^2 / (r^2 (1+z) dOmega nu^2)      (:173-184)
// to every channel of its pixel.  SED = SED_NORM (x^-0.1 + 10 x^-0.75), x = (1+z) nu / 1000, factors
// into a source part and a channel part, so each source adds two numbers to its pixel,
//   A += L/(r^2 (1+z)) ((1+z)/1000)^-0.1,   B += 10 L/(r^2 (1+z)) ((1+z)/1000)^-0.75,
// and T(pix, nu) = K (A nu^-2.1 + B nu^-2.75) afterwards: n_nu times fewer atomics than the
// reference's per-channel adds, the same sum up to rounding.
// Random numbers: Philox-4x32-10 with key (seed, PS_KEY1) and counter (cell, cell >> 32, w2, w3):
// Poisson draws w2 = 0, w3 = j; source s's position w2 = 4 s + 1 (w3 = 0: x, y; 1: z); its luminosity
// w2 = 4 s + 2, w3 = attempt.  The reference's per-thread GSL streams depend on its thread count.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void ps_uniforms(uint32_t seed, uint64_t cell, uint32_t w2, uint32_t w3, double &u1,
                                            double &u2)
{
  u32x4 c;
  c.v[0] = (uint32_t)cell;
  c.v[1] = (uint32_t)(cell >> 32);
  c.v[2] = w2;
  c.v[3] = w3;
  const u32x4 r = philox4x32_10(c, seed, PS_KEY1);
  u1 = u53(r.v[0], r.v[1]);
  u2 = u53(r.v[2], r.v[3]);
}

// Poisson(lambda): inversion by sequential search below 10, PTRS (Hormann 1993, Insurance: Math. and
// Econ. 12, 39) above.  Both loops are bounded; the bounds are never reached in practice.
__device__ uint32_t ps_poisson(double lam, uint32_t seed, uint64_t cell)
{
  if (!(lam > 0)) return 0;
  if (lam < 10.0) {
    double u, unused;
    ps_uniforms(seed, cell, 0, 0, u, unused);
    double p = exp(-lam), F = p;
    uint32_t k = 0;
    while (u > F && k < 1000) {
      k++;
      p *= lam / (double)k;
      F += p;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t j = 0; j < 1024; j++) {
    double U, V;
    ps_uniforms(seed, cell, 0, j, U, V);
    U -= 0.5;
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (uint32_t)k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return (uint32_t)k;
  }
  return (uint32_t)lam;
}

// n_of_z_psources (psources.c:74-90)
__device__ __forceinline__ double ps_nofz(const PsParams &p, double z)
{
  const int iz = (int)(z * p.inv_dz);
  if (iz >= PS_NZ || iz < 0) return -1;
  if (iz == PS_NZ - 1) return p.nz[PS_NZ - 1];
  const double zi = iz * p.dz;
  return p.nz[iz] + (p.nz[iz + 1] - p.nz[iz]) * (z - zi) * p.inv_dz;
}

// draw_luminosity (psources.c:130-152): u, then log10 L uniform in [LLOGMIN, LLOGMAX], accepted when
// u max_pdf < l_z_function(L, z) / n(z); z >= z_max gives L = 0.  10^x = exp(x ln 10) and
// (L/LX)^(alpha+1) = exp((alpha+1) ln(L/LX)) are the same functions written with one exponential each.
// ps_lum_setup: the per-source constants (false when z >= z_max); ps_lum_try: attempt `att`.
__device__ __forceinline__ bool ps_lum_setup(const PsParams &p, double z, double &maxpdf, double &c)
{
  if (z >= p.z_max) return false;
  int iz = (int)(z * p.inv_dz);
  iz = iz < 0 ? 0 : (iz < PS_NZ ? iz : PS_NZ - 1);   // z just below z_max can round to 256
  maxpdf = p.maxpdf[iz];
  const double fz = z < 1.5 ? pow(1.0 + z, 3.1) : 17.124;
  c = fz * 2.5 * M_LN10 * PS_RHO / ps_nofz(p, z);
  return true;
}

__device__ __forceinline__ bool ps_lum_try(uint32_t seed, uint64_t cell, uint32_t src, uint32_t att, double maxpdf,
                                           double c, double &L)
{
  double u, w;
  ps_uniforms(seed, cell, 4 * src + 2, att, u, w);
  const double logL = PS_LLOGMIN + (PS_LLOGMAX - PS_LLOGMIN) * w;
  L = exp(logL * M_LN10);
  const double x = L / PS_LX;
  return u * maxpdf < c * exp((PS_ALPHA + 1.0) * log(x) - x);
}

// Add (a, b) to pixel `key` of the workgroup's LDS table (open addressing, linear probing); a source
// that finds no slot within the probe budget goes straight to A, B in HBM.
template <int SLOTS>
__device__ __forceinline__ void ps_agg_add(unsigned long long key, double a, double b, unsigned long long *keys,
                                           double *va, double *vb, double *__restrict__ A, double *__restrict__ B)
{
  const unsigned long long EMPTY = ~0ull;
  constexpr int SHIFT = 64 - __builtin_ctz(SLOTS);
  unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> SHIFT) & (SLOTS - 1);
  for (int p = 0; p < 64; p++) {
    unsigned long long k = keys[h];
    if (k == EMPTY) {
      unsigned long long old = atomicCAS(&keys[h], EMPTY, key);
      if (old == EMPTY || old == key) {
        atomicAdd(&va[h], a);
        atomicAdd(&vb[h], b);
        return;
      }
    }
    else if (k == key) {
      atomicAdd(&va[h], a);
      atomicAdd(&vb[h], b);
      return;
    }
    h = (h + 1) & (SLOTS - 1);
  }
  unsafeAtomicAdd(A + key, a);
  unsafeAtomicAdd(B + key, b);
}

// One workgroup per map-stage tile (TILE_X x TILE_Y x TILE_Z cells, one per thread, tiles in order of
// distance from the observer).  A thread merges consecutive sources of its cell that share a pixel,
// the workgroup sums per pixel in LDS, and each occupied pixel is added to A, B once.
template <typename T>
__global__ __launch_bounds__(256) void k_psources(const cx<T> *__restrict__ F, MapParams m, PsParams ps)
{
  constexpr int SLOTS = PS_SLOTS;
  __shared__ unsigned long long keys[SLOTS];
  __shared__ double va[SLOTS], vb[SLOTS];
  __shared__ unsigned long long red[256];
  const int n = m.n;
  const int t = threadIdx.x;
  const size_t bid = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (bid >= (size_t)m.ntx * m.nty * m.ntz) return;   // whole workgroup
  const uint32_t tb = m.tile_order[bid];
  const int tx = (int)(tb & 1023u), ty = (int)((tb >> 10) & 1023u), tz = (int)(tb >> 20);
  const int ix = tx * TILE_X + (t % TILE_X), iy = ty * TILE_Y + ((t / TILE_X) % TILE_Y),
            iz = tz * TILE_Z + t / (TILE_X * TILE_Y);
  for (int i = t; i < SLOTS; i += 256) {
    keys[i] = ~0ull;
    va[i] = 0.0;
    vb[i] = 0.0;
  }
  __syncthreads();
  unsigned long long mine = 0;
  if (ix < n && iy < n && iz < n) {
    const size_t cell = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
    // get_point_sources: cell centre, with the reference's x offset by pos_obs[1] (grid_tools.c:66)
    const double xc = (ix + 0.5) * m.dx - m.obs[1];
    const double yc = (iy + 0.5) * m.dx - m.obs[1];
    const double zc = (iz + 0.5) * m.dx - m.obs[2];
    const double rc = sqrt(xc * xc + yc * yc + zc * zc);
    const double zred = z_of_r(m.tab, rc);
    const double ndens = ps_nofz(ps, zred);
    uint32_t np = 0;
    if (ndens > 0) {
      const double gfb = dgrowth_of_r(m.tab, rc);   // bias_psources = 1 (psources.c:66-69)
      const double lam = ndens * ps.cell_vol * exp(gfb * ((double)F[cell].x - 0.5 * gfb * m.sig[0]));
      np = ps_poisson(lam, ps.seed, cell);
    }
    mine = np;
    if (np > 0) {
      // mk_psources_maps: Delta z_RSD of the cell (grid_rvel after get_HI) and the cell's corner
      CellState s;
      cell_velocity<T>(m, F, ix, iy, iz, s);
      const double dz_rsd = m.rsd ? s.vr * vgrowth_of_r(m.tab, s.r) : 0.0;
      const double x0 = m.dx * ix - m.obs[0], y0 = m.dx * iy - m.obs[1], z0 = m.dx * iz - m.obs[2];
      unsigned long long ck = ~0ull;
      double ca = 0.0, cb = 0.0;
      for (uint32_t src = 0; src < np; src++) {
        double ux, uy, uz, unused;
        ps_uniforms(ps.seed, cell, 4 * src + 1, 0, ux, uy);
        ps_uniforms(ps.seed, cell, 4 * src + 1, 1, uz, unused);
        const double px = x0 + m.dx * ux, py = y0 + m.dx * uy, pz = z0 + m.dx * uz;
        const double r = sqrt(px * px + py * py + pz * pz);
        const double red_cosmo = z_of_r(m.tab, r);
        const double opz = 1.0 + red_cosmo + dz_rsd;
        const int64_t ipix = vec2pix_ring(m.nside, px, py, pz, r);
        double maxpdf, c, L = 0;
        if (!ps_lum_setup(ps, red_cosmo, maxpdf, c) || (uint64_t)ipix >= (uint64_t)m.npix) continue;   // L = 0
        for (uint32_t att = 0; att < PS_MAX_TRIES; att++)
          if (ps_lum_try(ps.seed, cell, src, att, maxpdf, c, L)) break;
        const double wgt = L / (r * r * opz);
        const double a = wgt * pow(opz * 1e-3, -0.1), b = 10.0 * wgt * pow(opz * 1e-3, -0.75);
        if ((unsigned long long)ipix == ck) {
          ca += a;
          cb += b;
        }
        else {
          if (ck != ~0ull) ps_agg_add<SLOTS>(ck, ca, cb, keys, va, vb, ps.A, ps.B);
          ck = (unsigned long long)ipix;
          ca = a;
          cb = b;
        }
      }
      if (ck != ~0ull) ps_agg_add<SLOTS>(ck, ca, cb, keys, va, vb, ps.A, ps.B);
    }
  }
  red[t] = mine;
  __syncthreads();
  for (int i = t; i < SLOTS; i += 256) {
    const unsigned long long k = keys[i];
    if (k != ~0ull) {
      unsafeAtomicAdd(ps.A + k, va[i]);
      unsafeAtomicAdd(ps.B + k, vb[i]);
    }
  }
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0 && red[0]) atomicAdd(ps.count, red[0]);
}

// maps_PS[inu][pix] = c1[inu] A[pix] + c2[inu] B[pix]
template <typename T>
__global__ __launch_bounds__(256) void k_ps_expand(const double *__restrict__ A, const double *__restrict__ B,
                                                   const double *__restrict__ c1, const double *__restrict__ c2,
                                                   int n_nu, int64_t npix, T *__restrict__ maps)
{
  const int inu = blockIdx.y;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < npix; p += (int64_t)gridDim.x * 256)
    maps[(size_t)inu * npix + p] = (T)(c1[inu] * A[p] + c2[inu] * B[p]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(T *__restrict__ p, size_t n)
{
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n; i += stride) p[i] = (T)0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_cell_field(const cx<T> *__restrict__ F, MapParams m, size_t cell0,
                                                    size_t ncell, int which, double *__restrict__ out)
{
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (size_t)gridDim.x * 256) {
    const size_t cell = cell0 + c;
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    out[c] = which == 2 ? s.vr : (which == 3 ? s.mass : s.dz);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sample_indices(const cx<T> *__restrict__ F, MapParams m,
                                                        int64_t cell0, int64_t ncell,
                                                        int32_t *__restrict__ inu_out,
                                                        int64_t *__restrict__ ipix_out)
{
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * 256) {
    size_t cell = (size_t)(cell0 + c);
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    CellScreen cs;
    decide_cell(m, s, cs);
    for (int isub = 0; isub < m.n_sub; isub++) {
      int64_t ipix;
      double pf;
      int inu = sample_decide(m, s, cs, isub, ipix, pf);
      inu_out[c * m.n_sub + isub] = inu;
      ipix_out[c * m.n_sub + isub] = ipix;
    }
  }
}

// Parity scatter: every sample of cells [cell0, cell0 + ncell) in FP64, through sample_bin.  ref != NULL:
// whole maps with FP64 global atomics; otherwise per-channel sums in LDS (n_nu <= TOT_MAX_NU) flushed
// to tot/cnt.
constexpr int TOT_MAX_NU = 2048;
template <typename T>
__global__ __launch_bounds__(256) void k_parity_scatter(const cx<T> *__restrict__ F, MapParams m, int64_t cell0,
                                                        int64_t ncell, double *__restrict__ ref,
                                                        double *__restrict__ tot,
                                                        unsigned long long *__restrict__ cnt)
{
  __shared__ double ht[TOT_MAX_NU];
  __shared__ unsigned long long hc[TOT_MAX_NU];
  if (!ref) {
    for (int i = threadIdx.x; i < m.n_nu; i += 256) {
      ht[i] = 0;
      hc[i] = 0;
    }
    __syncthreads();
  }
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * 256) {
    size_t cell = (size_t)(cell0 + c);
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    const double mass_sub = s.mass / m.n_sub;
    for (int isub = 0; isub < m.n_sub; isub++) {
      int64_t ipix;
      double pf;
      const int inu = sample_bin(m, s, isub, ipix, pf);
      if (inu < 0 || (uint64_t)ipix >= (uint64_t)m.npix) continue;
      if (ref) {
        atomicAdd(ref + (size_t)inu * m.npix + ipix, mass_sub * pf);
      }
      else {
        atomicAdd(&ht[inu], mass_sub * pf);
        atomicAdd(&hc[inu], 1ull);
      }
    }
  }
  if (!ref) {
    __syncthreads();
    for (int i = threadIdx.x; i < m.n_nu; i += 256) {
      if (hc[i]) {
        atomicAdd(tot + i, ht[i]);
        atomicAdd(cnt + i, hc[i]);
      }
    }
  }
}

// every sample through sample_bin (FP64) and the screen: counts of samples, decisions left to FP64 and
// mismatches (see launch_screen_check)
template <typename T>
__global__ __launch_bounds__(256) void k_screen_check(const cx<T> *__restrict__ F, MapParams m,
                                                      unsigned long long *__restrict__ counts)
{
  __shared__ unsigned long long red[5];
  if (threadIdx.x < 5) red[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long nsmp = 0, nch = 0, npix = 0, nbad = 0, nbin = 0;
  const size_t ncell = (size_t)m.n * m.n * m.n;
  for (size_t cell = (size_t)blockIdx.x * 256 + threadIdx.x; cell < ncell; cell += (size_t)gridDim.x * 256) {
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    CellScreen cs;
    decide_cell(m, s, cs);
    for (int isub = 0; isub < m.n_sub; isub++) {
      int64_t p64;
      double pf64;
      const int i64 = sample_bin(m, s, isub, p64, pf64);
      SampleS q;
      screen_channel(m, s, cs, isub, q);
      nch += q.inu == -2;
      if (q.inu >= 0 && screen_pixel(m, q) == -2) npix++;
      // a decided channel must be the FP64 one with the same factor as its window value
      const bool chan_bad = q.inu >= 0 && (q.inu != i64 || m.prefac[q.inu] != pf64);
      int64_t pq;
      const int iq = resolve_sample(m, s, isub, q, pq);
      nsmp++;
      nbin += i64 >= 0;
      nbad += chan_bad || (iq != i64) || (i64 >= 0 && pq != p64);
    }
  }
  atomicAdd(&red[0], nsmp);
  atomicAdd(&red[1], nch);
  atomicAdd(&red[2], npix);
  atomicAdd(&red[3], nbad);
  atomicAdd(&red[4], nbin);
  __syncthreads();
  if (threadIdx.x < 5) atomicAdd(counts + threadIdx.x, red[threadIdx.x]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_grid_get(const cx<T> *__restrict__ F, size_t n, int which,
                                                  double *__restrict__ out)
{
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    out[i] = which == 0 ? (double)F[i].x : (double)F[i].y;
}

template <typename T>
__global__ __launch_bounds__(256) void k_grid_set(cx<T> *__restrict__ F, size_t n, int which,
                                                  const double *__restrict__ in)
{
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    if (which == 0) F[i].x = (T)in[i];
    else F[i].y = (T)in[i];
  }
}

// blocks for a grid-stride loop over n items: a dispatch holds at most 2^32 work-items per dimension
static inline unsigned nblk(size_t n)
{
  size_t b = (n + 255) / 256;
  return (unsigned)(b < (1u << 20) ? b : (1u << 20));
}

template <typename T, bool FX>
static void launch_pixelize(const cx<T> *F, T *maps, const MapParams &m, hipStream_t s, const Schedule &sc,
                            size_t ntiles)
{
  if (ntiles == 0) return;
  unsigned gx = (unsigned)(ntiles < 65536 ? ntiles : 65536);
  unsigned gy = (unsigned)((ntiles + gx - 1) / gx);
  // 32-bit bin keys when every key and the empty marker fit
  const bool k32 = (uint64_t)m.n_nu * (uint64_t)m.npix < 0xFFFFFFFFull;
  using K64 = unsigned long long;
  const dim3 g(gx, gy), blk(256);
  // FP32 maps with 32-bit keys and at most 2048 bins fit 7 workgroups per CU in LDS: the 7-wave kernel
  constexpr bool W7 = sizeof(T) == 4 && !FX;
  auto run = [&](auto slots, auto seg, bool w7) {
    constexpr int S = decltype(slots)::value, G = decltype(seg)::value;
    if (k32) {
      if constexpr (W7 && S <= 2048) {
        if (w7) {
          hipLaunchKernelGGL((k_pixelize_w7<T, S, G, FX, unsigned>), g, blk, 0, s, F, maps, m);
          return;
        }
      }
      hipLaunchKernelGGL((k_pixelize<T, S, G, FX, unsigned>), g, blk, 0, s, F, maps, m);
    }
    else {
      hipLaunchKernelGGL((k_pixelize<T, S, G, FX, K64>), g, blk, 0, s, F, maps, m);
    }
  };
  using I2048 = std::integral_constant<int, 2048>;
  using I1024 = std::integral_constant<int, 1024>;
  using S2 = std::integral_constant<int, 2>;
  using S3 = std::integral_constant<int, 3>;
  if (sc.pix == 1) run(I2048{}, S3{}, true);
  else if (sc.pix == 2) run(I2048{}, S2{}, false);
  else if (sc.pix == 3) run(I1024{}, S2{}, true);
  else run(I2048{}, S2{}, true);
}

// ev[0] before clearing, ev[1] before the scatter, ev[2] after the last map kernel
template <typename T>
hipError_t launch_maps(const cx<T> *F, T *maps, size_t nmaps, const MapParams &m, hipStream_t s,
                       hipEvent_t *ev, const Schedule &sc, const RangeReadback *rb)
{
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_tile_range, dim3(1), dim3(256), 0, s, m);
  {
    const size_t nt = (size_t)m.ntx * m.nty * m.ntz;
    hipLaunchKernelGGL(k_tile_cull, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, m);
  }
  if (rb) {
    hipError_t e = hipMemcpyAsync(rb->h_range, m.tile_range, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(rb->ev, s);
    if (e != hipSuccess) return e;
  }
  if (m.fx) {
    if (!m.fxmaps || !m.fxs || !m.mass_part) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(nblk(nmaps)), dim3(256), 0, s, m.fxmaps, nmaps);
  }
  else {
    hipLaunchKernelGGL(k_zero<T>, dim3(nblk(nmaps)), dim3(256), 0, s, maps, nmaps);
  }
  if (ev) (void)hipEventRecord(ev[1], s);
  // the workgroups: the range's size when the host has it (the clear above runs meanwhile), else every tile
  size_t ntiles = (size_t)m.ntx * m.nty * m.ntz;
  auto host_range = [&]() -> hipError_t {
    const hipError_t e = hipEventSynchronize(rb->ev);
    if (e == hipSuccess) ntiles = rb->h_range[1] > rb->h_range[0] ? rb->h_range[1] - rb->h_range[0] : 0;
    return e;
  };
  if (m.fx) {
    hipLaunchKernelGGL(k_mass_partials<T>, dim3(FX_PARTS), dim3(256), 0, s, F, m);
    hipLaunchKernelGGL(k_fx_scale, dim3(1), dim3(256), 0, s, m);
    if (rb) {
      const hipError_t e = host_range();
      if (e != hipSuccess) return e;
    }
    launch_pixelize<T, true>(F, maps, m, s, sc, ntiles);
    int64_t b = (m.npix + 255) / 256;
    hipLaunchKernelGGL(k_fx_finalize<T>, dim3((unsigned)(b < 4096 ? b : 4096), m.n_nu), dim3(256), 0, s,
                       (const unsigned long long *)m.fxmaps, maps, m);
  }
  else {
    if (rb) {
      const hipError_t e = host_range();
      if (e != hipSuccess) return e;
    }
    launch_pixelize<T, false>(F, maps, m, s, sc, ntiles);
  }
  if (ev) (void)hipEventRecord(ev[2], s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_channel_totals(const cx<T> *F, const MapParams &m, int64_t cell0, int64_t ncell, double *tot,
                                 unsigned long long *cnt, hipStream_t s)
{
  if (m.n_nu > TOT_MAX_NU) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(256), 0, s, tot, (size_t)m.n_nu);
  hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(1), dim3(256), 0, s, cnt, (size_t)m.n_nu);
  hipLaunchKernelGGL(k_parity_scatter<T>, dim3(nblk((size_t)ncell)), dim3(256), 0, s, F, m, cell0, ncell,
                     (double *)nullptr, tot, cnt);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_screen_check(const cx<T> *F, const MapParams &m, unsigned long long *counts, hipStream_t s)
{
  hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(1), dim3(256), 0, s, counts, (size_t)5);
  const size_t ncell = (size_t)m.n * m.n * m.n;
  hipLaunchKernelGGL(k_screen_check<T>, dim3(nblk(ncell)), dim3(256), 0, s, F, m, counts);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_reference_maps(const cx<T> *F, const MapParams &m, double *ref, hipStream_t s)
{
  const int64_t ncell = (int64_t)m.n * m.n * m.n;
  hipLaunchKernelGGL(k_parity_scatter<T>, dim3(nblk((size_t)ncell)), dim3(256), 0, s, F, m, (int64_t)0, ncell, ref,
                     (double *)nullptr, (unsigned long long *)nullptr);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cell_field(const cx<T> *F, const MapParams &m, int which, size_t cell0, size_t ncell, double *out,
                             hipStream_t s)
{
  hipLaunchKernelGGL(k_cell_field<T>, dim3(nblk(ncell)), dim3(256), 0, s, F, m, cell0, ncell, which, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sample_indices(const cx<T> *F, const MapParams &m, int64_t cell0, int64_t ncell,
                                 int32_t *inu, int64_t *ipix, hipStream_t s)
{
  hipLaunchKernelGGL(k_sample_indices<T>, dim3(nblk((size_t)ncell)), dim3(256), 0, s, F, m, cell0, ncell,
                     inu, ipix);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_grid_get(const cx<T> *F, size_t ncell, int which, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_grid_get<T>, dim3(nblk(ncell)), dim3(256), 0, s, F, ncell, which, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_grid_set(cx<T> *F, size_t ncell, int which, const double *in, hipStream_t s)
{
  hipLaunchKernelGGL(k_grid_set<T>, dim3(nblk(ncell)), dim3(256), 0, s, F, ncell, which, in);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_psources(const cx<T> *F, const MapParams &m, const PsParams &ps, hipStream_t s)
{
  size_t ntiles = (size_t)m.ntx * m.nty * m.ntz;
  unsigned gx = (unsigned)(ntiles < 65536 ? ntiles : 65536);
  unsigned gy = (unsigned)((ntiles + gx - 1) / gx);
  hipLaunchKernelGGL(k_psources<T>, dim3(gx, gy), dim3(256), 0, s, F, m, ps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ps_expand(const double *A, const double *B, const double *c1, const double *c2, int n_nu,
                            int64_t npix, T *maps, hipStream_t s)
{
  int64_t b = (npix + 255) / 256;
  hipLaunchKernelGGL(k_ps_expand<T>, dim3((unsigned)(b < 4096 ? b : 4096), n_nu), dim3(256), 0, s, A, B, c1, c2,
                     n_nu, npix, maps);
  return hipGetLastError();
}

#define CRIME_INST(T)                                                                                   \
  template hipError_t launch_maps<T>(const cx<T> *, T *, size_t, const MapParams &, hipStream_t,      \
                                     hipEvent_t *, const Schedule &, const RangeReadback *);            \
  template hipError_t launch_cell_field<T>(const cx<T> *, const MapParams &, int, size_t, size_t, double *,   \
                                           hipStream_t);                                                \
  template hipError_t launch_sample_indices<T>(const cx<T> *, const MapParams &, int64_t, int64_t,     \
                                               int32_t *, int64_t *, hipStream_t);                      \
  template hipError_t launch_channel_totals<T>(const cx<T> *, const MapParams &, int64_t, int64_t, double *, \
                                               unsigned long long *, hipStream_t);                      \
  template hipError_t launch_reference_maps<T>(const cx<T> *, const MapParams &, double *, hipStream_t);  \
  template hipError_t launch_screen_check<T>(const cx<T> *, const MapParams &, unsigned long long *, hipStream_t); \
  template hipError_t launch_grid_get<T>(const cx<T> *, size_t, int, double *, hipStream_t);          \
  template hipError_t launch_grid_set<T>(cx<T> *, size_t, int, const double *, hipStream_t);          \
  template hipError_t launch_psources<T>(const cx<T> *, const MapParams &, const PsParams &, hipStream_t); \
  template hipError_t launch_ps_expand<T>(const double *, const double *, const double *, const double *, int, \
                                          int64_t, T *, hipStream_t);
CRIME_INST(float)
CRIME_INST(double)
#undef CRIME_INST

}  // namespace crime
