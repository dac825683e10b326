// HEALPix analysis on gfx950: ring spectra, Legendre table, alm GEMMs (FP64 MFMA) and C_l.
//
// a_lm = w sum_rings lambda_lm(z_r) F_m(r),  F_m(r) = sum_j f(r, j) e^{-i m phi_j},  w = 4 pi / Npix,
// the sum of he_map2alm (src/healpix_extra.c:88-130, libsharp with unit ring weights) split the way
// every fast SHT splits it (Reinecke 2011, A&A 526, A108): an FFT-like pass per ring, then per m a
// real (l x ring) by complex (ring x map) product.  North/south ring pairs are folded first:
// lambda_lm(-z) = (-1)^{l+m} lambda_lm(z), so rows with l - m even use E = F_north + F_south and rows
// with l - m odd use O = F_north - F_south over the 2 Nside northern rings (equator included once).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "gh_fft.hpp"
#include "gh_sht.hpp"

#include <algorithm>
#include <vector>

namespace crime {

// Layout of the ring spectra G: blocks of 16 rings, then m, then parity (E, O), then the 16 rings of
// the block, then [re, im] per map (ldg doubles per ring row).  A GEMM k-step (16 rings at one m and
// parity) is one contiguous 16 x ldg block, and the m-sweep of one ring stays within
// (lmax + 1) * 32 * ldg doubles instead of spanning the whole array.
__device__ __forceinline__ size_t g_block(int rb, int m, int par, int lmax, int ldg)
{
  return (((size_t)rb * (lmax + 1) + m) * 2 + par) * 16 * (size_t)ldg;
}

// ---------------------------------------------------------------------------------------------
// ring spectra: one workgroup per (map, northern ring); the mirror ring is transformed alongside.
// DFT of length n = P Q (P a power of two): S[j1][k1] = sum_j2 f[P j2 + j1] w^{P j2 k1},
// D[k] = sum_j1 w^{j1 k} S[j1][k mod Q], w = e^{-2 pi i / n}; F_m = e^{-i pi m s / n} D[m mod n].
// LDS: twiddles (n complex), f (n real), S (n complex), D north and south (K complex each).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int sht_pick_p(int n, int K)
{
  int best = 1, p = 1;
  double bestc = 1e300;
  while (n % p == 0 && p <= n) {
    int q = n / p;
    double c = 2.0 * n * q + 4.0 * K * p;
    if (c < bestc) { bestc = c; best = p; }
    p *= 2;
  }
  return best;
}

template <typename TM>
__device__ void ring_spectrum(const TM *__restrict__ f_g, int n, int K, int P, const double2 *tw, double *f,
                              double2 *S, double2 *D)
{
  const int Q = n / P;
  for (int j = threadIdx.x; j < n; j += blockDim.x) f[j] = (double)f_g[j];
  __syncthreads();
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    int j1 = idx / Q, k1 = idx - j1 * Q;
    double re = 0, im = 0;
    int t = 0;                   // (P j2 k1) mod n
    const int step = P * k1 % n;
    for (int j2 = 0; j2 < Q; j2++) {
      double v = f[P * j2 + j1];
      double2 w = tw[t];
      re = fma(v, w.x, re);
      im = fma(v, w.y, im);
      t += step;
      if (t >= n) t -= n;
    }
    S[idx] = make_double2(re, im);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    const int k1 = k % Q;
    double re = 0, im = 0;
    int t = 0;                   // (j1 k) mod n
    const int step = k % n;
    for (int j1 = 0; j1 < P; j1++) {
      double2 s = S[j1 * Q + k1];
      double2 w = tw[t];
      re = fma(s.x, w.x, fma(-s.y, w.y, re));
      im = fma(s.x, w.y, fma(s.y, w.x, im));
      t += step;
      if (t >= n) t -= n;
    }
    D[k] = make_double2(re, im);
  }
  __syncthreads();
}

template <typename TM>
__global__ __launch_bounds__(512) void k_ring_dft(const TM *const *__restrict__ mtab, int mps, int64_t npix,
                                                  const ShtRing *__restrict__ rings, int R, int RP, int lmax,
                                                  int ldg, double w, double *__restrict__ G, int r0)
{
  extern __shared__ double2 lds_sht[];
  const int c = blockIdx.x, r = r0 + blockIdx.y;
  const ShtRing ri = rings[r];
  const int n = ri.nphi;
  const int K = n < lmax + 1 ? n : lmax + 1;
  const int P = sht_pick_p(n, K);
  double2 *tw = lds_sht;                         // n
  double2 *S = tw + n;                           // n
  double2 *Dn = S + n;                           // K
  double2 *Ds = Dn + K;                          // K
  double *f = (double *)(Ds + K);                // n
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    double sn, cs;
    sincospi(-2.0 * (double)t / (double)n, &sn, &cs);
    tw[t] = make_double2(cs, sn);
  }
  const TM *mp = mtab[c / mps] + (size_t)(c % mps) * npix;
  ring_spectrum(mp + ri.first_n, n, K, P, tw, f, S, Dn);
  const bool eq = (r == R - 1);
  if (!eq) ring_spectrum(mp + ri.first_s, n, K, P, tw, f, S, Ds);
  for (int m = threadIdx.x; m <= lmax; m += blockDim.x) {
    const int k = m % n;
    double pr = 1, pi = 0;
    if (ri.s) {
      const int t = (int)((int64_t)m % (2 * n));
      sincospi(-(double)t / (double)n, &pi, &pr);
    }
    // the equator has no mirror: E = O = its own spectrum
    double2 a = Dn[k];
    double2 b = eq ? make_double2(0.0, 0.0) : Ds[k];
    double er = a.x + b.x, ei = a.y + b.y;
    double orr = a.x - b.x, oi = a.y - b.y;
    double *ge = G + g_block(r >> 4, m, 0, lmax, ldg) + (size_t)(r & 15) * ldg + 2 * c;
    double *go = ge + 16 * (size_t)ldg;
    ge[0] = w * (er * pr - ei * pi);
    ge[1] = w * (er * pi + ei * pr);
    go[0] = w * (orr * pr - oi * pi);
    go[1] = w * (orr * pi + oi * pr);
  }
}

// ---------------------------------------------------------------------------------------------
// equatorial rings (all 4 Nside long) by the Stockham engine: one line per (map, ring pair) packs
// the north ring and its mirror as z = f_N + i f_S; Z = sum_j z_j e^{+2 pi i jk/L} gives
// A_N = (Z_k + conj Z_{-k})/2, A_S = (Z_k - conj Z_{-k})/(2i), and the forward spectra D = conj(A).
// The equator (the last line of a map) has no mirror: f_S = 0 there.
// ---------------------------------------------------------------------------------------------
constexpr int sht_imax(int a, int b) { return a > b ? a : b; }
constexpr int sht_imin(int a, int b) { return a < b ? a : b; }

// LDS slot of element e of a line: one pad slot per 16 elements, so the radix-16/8 first pass,
// whose lanes write 16 elements apart, spreads over all banks instead of hitting one
__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }
constexpr int lpadded(int L) { return L + L / 16; }

// threads per line: L / RING_TC_DIV, at least 64; largest radix of the ring transforms
#ifndef CRIME_RING_TC_DIV
#define CRIME_RING_TC_DIV 16
#endif
#ifndef CRIME_RING_MAXR
#define CRIME_RING_MAXR 16
#endif

template <int L>
struct RingFftCfg {
  static constexpr int MAXR = CRIME_RING_MAXR;
  static constexpr int TC = plan_tc(L, sht_imax(64, L / CRIME_RING_TC_DIV), MAXR);
  // lines per workgroup: at least 128 threads where 64 KB of LDS allows
  static constexpr int G = sht_imax(1, sht_imin(sht_imax(4, 128 / TC), 65536 / (16 * L)));
  static constexpr int NT = G * TC;
};

template <typename TM, int L>
__global__ __launch_bounds__(RingFftCfg<L>::NT) void k_ring_fft(const TM *const *__restrict__ mtab, int mps,
                                                               int64_t npix,
                                                               const ShtRing *__restrict__ rings, int R, int RP,
                                                               int lmax, int ldg, double w, double *__restrict__ G,
                                                               int r0, int nmaps, int nlines,
                                                               const cx<double> *__restrict__ tw)
{
  using C = RingFftCfg<L>;
  __shared__ cx<double> lds[C::G * lpadded(L)];
  const int g = threadIdx.x / C::TC, t = threadIdx.x % C::TC;
  const int line = xcd_swizzle(blockIdx.x, gridDim.x) * C::G + g;
  const bool live = line < nlines;            // padding lines still join every barrier
  const int c = live ? line % nmaps : 0, r = r0 + (live ? line / nmaps : 0);
  const ShtRing ri = rings[r];
  const bool eq = (r == R - 1);
  const TM *base = mtab[c / mps] + (size_t)(c % mps) * npix;   // map c of set c / mps
  const TM *fn = base + ri.first_n;
  const TM *fs = base + ri.first_s;
  cx<double> *buf = lds + g * lpadded(L);
  auto load = [&](int e) -> cx<double> {
    if (!live) return {0.0, 0.0};
    return {(double)fn[e], eq ? 0.0 : (double)fs[e]};
  };
  auto store = [&](int e, cx<double> v) { buf[lpad(e)] = v; };
  auto slot = [&](int e) -> cx<double> & { return buf[lpad(e)]; };
  Stockham<double, L, C::TC, C::MAXR>::run(t, load, store, slot, tw);
  __syncthreads();
  if (!live) return;
  for (int m = t; m <= lmax; m += C::TC) {
    const cx<double> zk = buf[lpad(m)], zm = buf[lpad((L - m) & (L - 1))];
    // D_N = conj(A_N), D_S = conj(A_S)
    const double dnr = 0.5 * (zk.x + zm.x), dni = -0.5 * (zk.y - zm.y);
    const double dsr = 0.5 * (zk.y + zm.y), dsi = 0.5 * (zk.x - zm.x);
    double pr = 1, pi = 0;
    if (ri.s) {
      const int tt = (int)((int64_t)m % (2 * L));
      sincospi(-(double)tt / (double)L, &pi, &pr);
    }
    const double er = dnr + dsr, ei = dni + dsi, orr = dnr - dsr, oi = dni - dsi;
    double *ge = G + g_block(r >> 4, m, 0, lmax, ldg) + (size_t)(r & 15) * ldg + 2 * c;
    double *go = ge + 16 * (size_t)ldg;
    ge[0] = w * (er * pr - ei * pi);
    ge[1] = w * (er * pi + ei * pr);
    go[0] = w * (orr * pr - oi * pi);
    go[1] = w * (orr * pi + oi * pr);
  }
}

// ---------------------------------------------------------------------------------------------
// Bluestein (1970) for rings of any length n: with w_j = e^{+i pi j^2 / n} (exact: j^2 mod 2n),
//   Z_k = sum_j z_j e^{+2 pi i jk/n} = w_k sum_j (z_j w_j) conj(w_{k-j}),
// a cyclic convolution of length M = 2^a >= 2n - 1 done with the Stockham engine:
//   A = P(a),  conv = conj(P(conj(A B))),  B = P(b) / M,  b_m = conj(w_m) (m < n), b_{M-m} = conj(w_m)
// (P: the engine's sign-+ transform of length M).  B and w are tabulated per ring at plan creation.
// z packs the north ring and its mirror as in k_ring_fft.
// ---------------------------------------------------------------------------------------------
template <int M>
struct BluCfg {
  static constexpr int MAXR = CRIME_RING_MAXR;
  static constexpr int TC = plan_tc(M, sht_imax(64, M / CRIME_RING_TC_DIV), MAXR);
  static constexpr int G = sht_imax(1, sht_imin(sht_imax(4, 128 / TC), 65536 / (16 * M)));
  static constexpr int NT = G * TC;
};

template <int M>
__global__ __launch_bounds__(BluCfg<M>::NT) void k_blu_setup(const ShtRing *__restrict__ rings, int r0, int nr,
                                                            const int64_t *__restrict__ woff,
                                                            const cx<double> *__restrict__ wtab,
                                                            const int64_t *__restrict__ boff,
                                                            cx<double> *__restrict__ B,
                                                            const cx<double> *__restrict__ tw)
{
  using C = BluCfg<M>;
  __shared__ cx<double> lds[C::G * lpadded(M)];
  const int g = threadIdx.x / C::TC, t = threadIdx.x % C::TC;
  const int line = xcd_swizzle(blockIdx.x, gridDim.x) * C::G + g;
  const bool live = line < nr;
  const int r = r0 + (live ? line : 0);
  const int n = rings[r].nphi;
  const cx<double> *w = wtab + woff[r];
  cx<double> *buf = lds + g * lpadded(M);
  auto load = [&](int e) -> cx<double> {
    if (!live) return {0.0, 0.0};
    const int m = e < n ? e : (e > M - n ? M - e : -1);
    if (m < 0) return {0.0, 0.0};
    return {w[m].x, -w[m].y};
  };
  auto store = [&](int e, cx<double> v) {
    if (live) B[boff[r] + e] = {v.x / (double)M, v.y / (double)M};
  };
  auto slot = [&](int e) -> cx<double> & { return buf[lpad(e)]; };
  Stockham<double, M, C::TC, C::MAXR>::run(t, load, store, slot, tw);
}

template <typename TM, int M>
__global__ __launch_bounds__(BluCfg<M>::NT) void k_ring_blu(const TM *const *__restrict__ mtab, int mps,
                                                           int64_t npix,
                                                           const ShtRing *__restrict__ rings, int R, int RP, int lmax,
                                                           int ldg, double wn, double *__restrict__ G, int r0,
                                                           int nmaps, int nlines, const int64_t *__restrict__ woff,
                                                           const cx<double> *__restrict__ wtab,
                                                           const int64_t *__restrict__ boff,
                                                           const cx<double> *__restrict__ B,
                                                           const cx<double> *__restrict__ tw)
{
  using C = BluCfg<M>;
  __shared__ cx<double> lds[C::G * lpadded(M)];
  const int g = threadIdx.x / C::TC, t = threadIdx.x % C::TC;
  const int line = xcd_swizzle(blockIdx.x, gridDim.x) * C::G + g;
  const bool live = line < nlines;
  const int c = live ? line % nmaps : 0, r = r0 + (live ? line / nmaps : 0);
  const ShtRing ri = rings[r];
  const int n = ri.nphi;
  const bool eq = (r == R - 1);
  const TM *base = mtab[c / mps] + (size_t)(c % mps) * npix;   // map c of set c / mps
  const TM *fn = base + ri.first_n;
  const TM *fs = base + ri.first_s;
  const cx<double> *w = wtab + woff[r];
  const cx<double> *Bk = B + boff[r];
  cx<double> *buf = lds + g * lpadded(M);
  auto load1 = [&](int e) -> cx<double> {
    if (!live || e >= n) return {0.0, 0.0};
    const cx<double> z = {(double)fn[e], eq ? 0.0 : (double)fs[e]};
    return cmul(z, w[e]);
  };
  auto store1 = [&](int e, cx<double> v) { buf[lpad(e)] = v; };
  auto slot = [&](int e) -> cx<double> & { return buf[lpad(e)]; };
  Stockham<double, M, C::TC, C::MAXR>::run(t, load1, store1, slot, tw);
  __syncthreads();
  auto load2 = [&](int e) -> cx<double> { return cconj(cmul(buf[lpad(e)], Bk[e])); };
  auto store2 = [&](int e, cx<double> v) {
    if (e < n) buf[lpad(e)] = cmul(w[e], cconj(v));
  };
  Stockham<double, M, C::TC, C::MAXR>::template run<true>(t, load2, store2, slot, tw);
  __syncthreads();
  if (!live) return;
  for (int m = t; m <= lmax; m += C::TC) {
    const int k = m % n, km = k == 0 ? 0 : n - k;
    const cx<double> zk = buf[lpad(k)], zm = buf[lpad(km)];
    const double dnr = 0.5 * (zk.x + zm.x), dni = -0.5 * (zk.y - zm.y);
    const double dsr = 0.5 * (zk.y + zm.y), dsi = 0.5 * (zk.x - zm.x);
    double pr = 1, pi = 0;
    if (ri.s) {
      const int tt = (int)((int64_t)m % (2 * n));
      sincospi(-(double)tt / (double)n, &pi, &pr);
    }
    const double er = dnr + dsr, ei = dni + dsi, orr = dnr - dsr, oi = dni - dsi;
    double *ge = G + g_block(r >> 4, m, 0, lmax, ldg) + (size_t)(r & 15) * ldg + 2 * c;
    double *go = ge + 16 * (size_t)ldg;
    ge[0] = wn * (er * pr - ei * pi);
    ge[1] = wn * (er * pi + ei * pr);
    go[0] = wn * (orr * pr - oi * pi);
    go[1] = wn * (orr * pi + oi * pr);
  }
}

// ---------------------------------------------------------------------------------------------
// Synthesis (alm2map) ring stage.  For a real map f_j = sum_{|m| <= lmax} F_m e^{i m phi_j} with
// F_{-m} = conj F_m and phi_j = pi (2 j + s) / n, so f_j = sum_k D_k e^{+2 pi i jk/n} with the folded,
// phase-shifted spectrum
//   D_k = sum_{m = k mod n} F'_m + sum_{m > 0, -m = k mod n} conj F'_m,   F'_m = F_m e^{+i pi m s / n},
// F_N = E + O and F_S = E - O from the synthesis GEMM (the same 16-ring blocks of G).  The north ring
// and its mirror share one transform: z_k = D^N_k + i D^S_k gives f_N = Re Z, f_S = Im Z, both real
// because D^N and D^S are Hermitian (Im a_l0 is dropped, as a real-map synthesis does).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ cx<double> syn_fold(const double *__restrict__ G, int r, int c, int lmax, int ldg, int n,
                                               int s, bool eq, int k)
{
  const size_t rowoff = (size_t)(r & 15) * ldg + 2 * c;
  double nr = 0, ni = 0, sr = 0, si = 0;
  auto term = [&](int m, bool conj_) {
    const double *ge = G + g_block(r >> 4, m, 0, lmax, ldg) + rowoff;
    const double *go = ge + 16 * (size_t)ldg;
    const double er = ge[0], ei = ge[1], orr = go[0], oi = go[1];
    double fnr = er + orr, fni = ei + oi, fsr = er - orr, fsi = ei - oi;
    if (m == 0) {
      nr += fnr;
      sr += fsr;
      return;
    }
    double pr = 1, pi = 0;
    if (s) {
      const int tt = (int)((int64_t)m % (2 * n));
      sincospi((double)tt / (double)n, &pi, &pr);
    }
    const double anr = fnr * pr - fni * pi, ani = fnr * pi + fni * pr;
    const double asr = fsr * pr - fsi * pi, asi = fsr * pi + fsi * pr;
    nr += anr;
    sr += asr;
    ni += conj_ ? -ani : ani;
    si += conj_ ? -asi : asi;
  };
  for (int m = k; m <= lmax; m += n) term(m, false);
  for (int m = k == 0 ? n : n - k; m <= lmax; m += n) term(m, true);
  if (eq) sr = si = 0.0;
  return {nr - si, ni + sr};
}

template <typename TM, int L>
__global__ __launch_bounds__(RingFftCfg<L>::NT) void k_syn_fft(TM *__restrict__ maps, int64_t npix,
                                                              const ShtRing *__restrict__ rings, int R, int lmax,
                                                              int ldg, const double *__restrict__ G, int r0,
                                                              int nmaps, int nlines, const cx<double> *__restrict__ tw)
{
  using C = RingFftCfg<L>;
  __shared__ cx<double> lds[C::G * lpadded(L)];
  const int g = threadIdx.x / C::TC, t = threadIdx.x % C::TC;
  const int line = xcd_swizzle(blockIdx.x, gridDim.x) * C::G + g;
  const bool live = line < nlines;            // padding lines still join every barrier
  const int c = live ? line % nmaps : 0, r = r0 + (live ? line / nmaps : 0);
  const ShtRing ri = rings[r];
  const bool eq = (r == R - 1);
  TM *fn = maps + (size_t)c * npix + ri.first_n;
  TM *fs = maps + (size_t)c * npix + ri.first_s;
  cx<double> *buf = lds + g * lpadded(L);
  auto load = [&](int e) -> cx<double> {
    if (!live) return {0.0, 0.0};
    return syn_fold(G, r, c, lmax, ldg, L, ri.s, eq, e);
  };
  auto store = [&](int e, cx<double> v) {
    if (!live) return;
    fn[e] = (TM)v.x;
    if (!eq) fs[e] = (TM)v.y;
  };
  auto slot = [&](int e) -> cx<double> & { return buf[lpad(e)]; };
  Stockham<double, L, C::TC, C::MAXR>::run(t, load, store, slot, tw);
}

template <typename TM, int M>
__global__ __launch_bounds__(BluCfg<M>::NT) void k_syn_blu(TM *__restrict__ maps, int64_t npix,
                                                          const ShtRing *__restrict__ rings, int R, int lmax, int ldg,
                                                          const double *__restrict__ G, int r0, int nmaps, int nlines,
                                                          const int64_t *__restrict__ woff,
                                                          const cx<double> *__restrict__ wtab,
                                                          const int64_t *__restrict__ boff,
                                                          const cx<double> *__restrict__ B,
                                                          const cx<double> *__restrict__ tw)
{
  using C = BluCfg<M>;
  __shared__ cx<double> lds[C::G * lpadded(M)];
  const int g = threadIdx.x / C::TC, t = threadIdx.x % C::TC;
  const int line = xcd_swizzle(blockIdx.x, gridDim.x) * C::G + g;
  const bool live = line < nlines;
  const int c = live ? line % nmaps : 0, r = r0 + (live ? line / nmaps : 0);
  const ShtRing ri = rings[r];
  const int n = ri.nphi;
  const bool eq = (r == R - 1);
  TM *fn = maps + (size_t)c * npix + ri.first_n;
  TM *fs = maps + (size_t)c * npix + ri.first_s;
  const cx<double> *w = wtab + woff[r];
  const cx<double> *Bk = B + boff[r];
  cx<double> *buf = lds + g * lpadded(M);
  auto load1 = [&](int e) -> cx<double> {
    if (!live || e >= n) return {0.0, 0.0};
    return cmul(syn_fold(G, r, c, lmax, ldg, n, ri.s, eq, e), w[e]);
  };
  auto store1 = [&](int e, cx<double> v) { buf[lpad(e)] = v; };
  auto slot = [&](int e) -> cx<double> & { return buf[lpad(e)]; };
  Stockham<double, M, C::TC, C::MAXR>::run(t, load1, store1, slot, tw);
  __syncthreads();
  auto load2 = [&](int e) -> cx<double> { return cconj(cmul(buf[lpad(e)], Bk[e])); };
  auto store2 = [&](int e, cx<double> v) {
    if (!live || e >= n) return;
    const cx<double> z = cmul(w[e], cconj(v));
    fn[e] = (TM)z.x;
    if (!eq) fs[e] = (TM)z.y;
  };
  Stockham<double, M, C::TC, C::MAXR>::template run<true>(t, load2, store2, slot, tw);
}

// rings shorter than 32 pixels: one workgroup per (map, ring pair), direct sums with exact phases
template <typename TM>
__global__ __launch_bounds__(64) void k_syn_dft(TM *__restrict__ maps, int64_t npix, const ShtRing *__restrict__ rings,
                                                int R, int lmax, int ldg, const double *__restrict__ G, int r0)
{
  __shared__ cx<double> z[32];
  const int c = blockIdx.x, r = r0 + blockIdx.y;
  const ShtRing ri = rings[r];
  const int n = ri.nphi;
  const bool eq = (r == R - 1);
  for (int k = threadIdx.x; k < n; k += 64) z[k] = syn_fold(G, r, c, lmax, ldg, n, ri.s, eq, k);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 64) {
    double zr = 0, zi = 0;
    for (int k = 0; k < n; k++) {
      double sn, cs;
      sincospi(2.0 * (double)((j * k) % n) / (double)n, &sn, &cs);
      zr += z[k].x * cs - z[k].y * sn;
      zi += z[k].x * sn + z[k].y * cs;
    }
    maps[(size_t)c * npix + ri.first_n + j] = (TM)zr;
    if (!eq) maps[(size_t)c * npix + ri.first_s + j] = (TM)zi;
  }
}

size_t ring_dft_lds_bytes(int n, int lmax)
{
  int K = n < lmax + 1 ? n : lmax + 1;
  return (size_t)(2 * n + 2 * K) * sizeof(double2) + (size_t)n * sizeof(double);
}

// ---------------------------------------------------------------------------------------------
// Legendre rows lambda_lm(z_r), l = m..lmax, for the m of one chunk: thread per (m, ring).
// Start: lambda_mm = (-1)^m 2^{C_m} sin^m(theta), C_m = log2(1/sqrt(4 pi)) + sum_k log2 sqrt((2k+1)/2k);
// lambda_{m+1,m} = sqrt(2m+3) z lambda_mm;
// lambda_lm = a_lm (z lambda_{l-1,m} - b_lm lambda_{l-2,m}), a = sqrt((4l^2-1)/(l^2-m^2)),
// b = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)).  Values below 2^-300 are carried as v 2^{-600 s}, s >= 1
// (rescaled upwards when they grow; Holmes & Featherstone 2002, J. Geodesy 76, 279), and enter the
// table as 0 until s returns to 0.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_legendre(const double2 *__restrict__ coef, const int64_t *__restrict__ off,
                                                  int m0, int lmax, const double *__restrict__ zs,
                                                  const double *__restrict__ sts, int R, int RP,
                                                  double *__restrict__ lam)
{
  const int m = m0 + blockIdx.y;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= RP) return;
  const int L = lmax - m + 1;
  const int64_t o = off[m];
  double *out = lam + (size_t)(o - off[m0]) * RP + r;
  if (r >= R) {
    for (int i = 0; i < L; i++) out[(size_t)i * RP] = 0.0;
    return;
  }
  const double z = zs[r], st = sts[r];
  const double ISC = 0x1p-600;
  double e = coef[o].x + (double)m * log2(st);
  int s = 0;
  if (e < -300.0) s = (int)ceil((-300.0 - e) / 600.0);
  double p = exp2(e + 600.0 * s);
  if (m & 1) p = -p;
  out[0] = s == 0 ? p : 0.0;
  if (L == 1) return;
  double q = coef[o + 1].x * z * p;
  out[RP] = s == 0 ? q : 0.0;
  for (int i = 2; i < L; i++) {
    const double2 ab = coef[o + i];
    double t = ab.x * (z * q - ab.y * p);
    p = q;
    q = t;
    if (s > 0 && fabs(q) > 0x1p300) {
      p *= ISC;
      q *= ISC;
      s--;
    }
    out[(size_t)i * RP] = s == 0 ? q : 0.0;
  }
}

// ---------------------------------------------------------------------------------------------
// alm GEMM: for one (m, parity, 64-row tile) and one 64-column tile,
//   alm[idx(l, m)][col] = sum_r lambda_lm(z_r) G[m][parity][r][col],  l = m + parity + 2 i.
// 256 threads = 4 waves, each a 32 x 32 block of 2 x 2 v_mfma_f64_16x16x4f64 tiles; K in chunks of 16
// rings through LDS.  Operand layout of the 16x16x4 FP64 MFMA (64 lanes, lane = i + 16 k for A[i][k],
// j + 16 k for B[k][j]; D[i][j] in lane j + 16 (i % 4), element i / 4 -- measured on gfx950).
// ---------------------------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64, GK = 16;
static_assert(GK == 16, "a GEMM k-step is one 16-ring block of G");
static_assert(ALM_ROWS == 64 || ALM_ROWS == 128, "alm tiles of 64 or 128 rows");
#ifndef CRIME_SYN_COLS
#define CRIME_SYN_COLS 64
#endif
constexpr int SYN_COLS = CRIME_SYN_COLS;   // columns per synthesis GEMM tile (64 or 128)
static_assert(SYN_COLS == 64 || SYN_COLS == 128, "synthesis tiles of 64 or 128 columns");

// alm[l][col] = sum_r lambda_lm(r) G[r][col] per (m, parity): a workgroup computes ALM_ROWS x GT of
// it, its 4 waves 2 x 2 blocks of (ALM_ROWS / 2) x 32 with 16x16x4 FP64 MFMAs.
__global__ __launch_bounds__(256) void k_alm_gemm(const double *__restrict__ lam, const double *__restrict__ G,
                                                  const int4 *__restrict__ tiles, int ntiles, int ncoltiles,
                                                  int tiles_per_xcd, const int64_t *__restrict__ off, int m0,
                                                  int lmax, int RP, int ldg, int ncols, double *__restrict__ alm,
                                                  int ldo)
{
  constexpr int RW = ALM_ROWS / 2, NI = RW / 16;     // rows per wave, MFMA row blocks per wave
  constexpr int AL = ALM_ROWS * GK / 256;            // A doubles loaded per thread
  // XCD-aware order: consecutive logical tiles (which share Legendre rows) stay on one XCD's L2
  const int b = blockIdx.x;
  const int logical = (b % 8) * tiles_per_xcd + b / 8;
  if (logical >= ntiles * ncoltiles) return;
  const int4 tl = tiles[logical / ncoltiles];
  const int ct = logical % ncoltiles;
  const int m = tl.x, par = tl.y, rt = tl.z, nrows = tl.w;
  __shared__ double As[ALM_ROWS][GK + 1];
  __shared__ double Bs[GK][GT + 2];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int row0 = rt * ALM_ROWS, col0 = ct * GT;
  const double *lrow = lam + (size_t)(off[m] - off[m0]) * RP;   // row (l - m) of this m

  dbl4 acc[NI][2];
  for (int i = 0; i < NI; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (dbl4){0, 0, 0, 0};
  const int ar = tid / (GK / AL), ak = (tid % (GK / AL)) * AL;   // A loads: row ar, AL rings
  const int bk = tid >> 4, bc = (tid & 15) * 4;                    // B loads: ring bk, 4 columns
  // Operands out of range are read from an in-range address and zeroed by a select, so the loop
  // has no divergent branches (they would make the compiler shuttle the accumulators between
  // AGPRs and VGPRs every step).  ncols is even and col0 + bc a multiple of 4, so each double2
  // of B is wholly in range or wholly out.
  const bool arow_ok = row0 + ar < nrows;
  const double *ap = lrow + (size_t)(par + 2 * (arow_ok ? row0 + ar : 0)) * RP + ak;
  const bool b0_ok = col0 + bc < ncols, b1_ok = col0 + bc + 2 < ncols;
  const int bcol0 = b0_ok ? col0 + bc : 0, bcol1 = b1_ok ? col0 + bc + 2 : 0;
  const int wr = (wv >> 1) * RW, wc = (wv & 1) * 32;
  // the next k-step's operands are fetched into registers while this one's MFMAs run
  double2 ra[AL / 2], rb0, rb1;
  auto fetch = [&](int k0) {
    for (int e = 0; e < AL / 2; e++) ra[e] = *(const double2 *)(ap + k0 + 2 * e);
    const double *bp = G + g_block(k0 >> 4, m, par, lmax, ldg) + (size_t)bk * ldg;
    rb0 = *(const double2 *)(bp + bcol0);
    rb1 = *(const double2 *)(bp + bcol1);
  };
  fetch(0);
  for (int k0 = 0; k0 < RP; k0 += GK) {
    for (int e = 0; e < AL / 2; e++) {
      As[ar][ak + 2 * e] = arow_ok ? ra[e].x : 0.0;
      As[ar][ak + 2 * e + 1] = arow_ok ? ra[e].y : 0.0;
    }
    Bs[bk][bc] = b0_ok ? rb0.x : 0.0;
    Bs[bk][bc + 1] = b0_ok ? rb0.y : 0.0;
    Bs[bk][bc + 2] = b1_ok ? rb1.x : 0.0;
    Bs[bk][bc + 3] = b1_ok ? rb1.y : 0.0;
    __syncthreads();
    fetch(k0 + GK < RP ? k0 + GK : k0);   // the last step refetches its own operands: no branch
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      double af[NI], bf[2];
      for (int i = 0; i < NI; i++) af[i] = As[wr + 16 * i + (lane & 15)][kk + (lane >> 4)];
      for (int j = 0; j < 2; j++) bf[j] = Bs[kk + (lane >> 4)][wc + 16 * j + (lane & 15)];
      for (int i = 0; i < NI; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  const int64_t base_idx = (int64_t)m * lmax - (int64_t)m * (m - 1) / 2;   // he_indexlm(l, m) - l
  for (int i = 0; i < NI; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 4; v++) {
        const int row = row0 + wr + 16 * i + (lane >> 4) + 4 * v;
        const int col = col0 + wc + 16 * j + (lane & 15);
        if (row < nrows && col < ncols) {
          const int l = m + par + 2 * row;
          const int64_t idx = m > 0 ? base_idx + l : l;
          alm[(size_t)idx * ldo + col] = acc[i][j][v];
        }
      }
}

// ---------------------------------------------------------------------------------------------
// C_l from alm[idx][2 t + (0 re | 1 im)]: cross spectra of all map pairs, autospectra, listed pairs.
// ---------------------------------------------------------------------------------------------
// Cross spectra on FP64 MFMA: for each l, C_l = X Y^T / (2l + 1) over k = 2m + (0 re | 1 im), with
// X[a][k] = w_m alm_a, Y[b][k] = alm_b, w_0 = 1, w_m = 2, for the pairs within each of nsets sets of
// mps maps (one set: every pair).  One workgroup per (set, 64 x 64 map tile with ta <= tb, l);
// off-diagonal tiles also write the mirror block, so C^{ab} and C^{ba} are equal bit for bit.
// Each XCD takes one contiguous run of logical blocks [xr.start[x], xr.start[x + 1]), so consecutive
// blocks of one XCD take consecutive l of one tile and the [set][a][b][l] stores of neighbouring l meet
// in that XCD's L2.  A block's work grows with l, so the host cuts the runs at equal work, not at equal
// block counts (launch_cls_sets).
struct XcdRuns {
  int start[9];
};

__global__ __launch_bounds__(256) void k_cls_cross(const double *__restrict__ alm, int lmax, int mps, int ldo,
                                                   int ntile, XcdRuns xr, double *__restrict__ cls)
{
  const int x = blockIdx.x & 7;   // blocks are dealt to the 8 XCDs round robin
  const int logical = xr.start[x] + (int)(blockIdx.x >> 3);
  if (logical >= xr.start[x + 1]) return;
  const int nl = lmax + 1;
  const int l = lmax - logical % nl;
  const int npair = ntile * (ntile + 1) / 2;
  const int set = logical / nl / npair;
  int rem = logical / nl % npair, ta = 0;
  while (rem >= ntile - ta) {
    rem -= ntile - ta;
    ta++;
  }
  const int tb = ta + rem;
  const int a0 = ta * GT, b0 = tb * GT;           // within the set
  const int cbase = set * mps;                    // the set's first map in the alm columns
  __shared__ double As[GT][GK + 1];
  __shared__ double Bs[GK][GT + 2];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const int mi = tid >> 5, q = 2 * (tid & 31);      // loads: m chunk row mi, maps q and q + 1
  dbl4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (dbl4){0, 0, 0, 0};
  for (int mc = 0; mc <= l; mc += GK / 2) {
    const int m = mc + mi;
    double2 xa0 = make_double2(0, 0), xa1 = xa0, yb0 = xa0, yb1 = xa0;
    if (m <= l) {
      const int64_t idx = m > 0 ? l + (int64_t)m * lmax - (int64_t)m * (m - 1) / 2 : l;
      const double *row = alm + (size_t)idx * ldo + 2 * cbase;
      const double wgt = m ? 2.0 : 1.0;
      if (a0 + q < mps) xa0 = *(const double2 *)(row + 2 * (a0 + q));
      if (a0 + q + 1 < mps) xa1 = *(const double2 *)(row + 2 * (a0 + q + 1));
      if (b0 + q < mps) yb0 = *(const double2 *)(row + 2 * (b0 + q));
      if (b0 + q + 1 < mps) yb1 = *(const double2 *)(row + 2 * (b0 + q + 1));
      xa0.x *= wgt; xa0.y *= wgt; xa1.x *= wgt; xa1.y *= wgt;
    }
    As[q][2 * mi] = xa0.x;
    As[q][2 * mi + 1] = xa0.y;
    As[q + 1][2 * mi] = xa1.x;
    As[q + 1][2 * mi + 1] = xa1.y;
    Bs[2 * mi][q] = yb0.x;
    Bs[2 * mi + 1][q] = yb0.y;
    Bs[2 * mi][q + 1] = yb1.x;
    Bs[2 * mi + 1][q + 1] = yb1.y;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      double af[2], bf[2];
      for (int i = 0; i < 2; i++) af[i] = As[wr + 16 * i + (lane & 15)][kk + (lane >> 4)];
      for (int j = 0; j < 2; j++) bf[j] = Bs[kk + (lane >> 4)][wc + 16 * j + (lane & 15)];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  const double inv = 1.0 / (2.0 * l + 1.0);
  double *out = cls + (size_t)set * mps * mps * nl;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 4; v++) {
        const int a = a0 + wr + 16 * i + (lane >> 4) + 4 * v;
        const int b = b0 + wc + 16 * j + (lane & 15);
        if (a < mps && b < mps) {
          const double c = acc[i][j][v] * inv;
          out[((size_t)a * mps + b) * nl + l] = c;
          if (ta != tb) out[((size_t)b * mps + a) * nl + l] = c;
        }
      }
}

__global__ __launch_bounds__(256) void k_cls_auto(const double *__restrict__ alm, int lmax, int nmaps, int ldo,
                                                  double *__restrict__ cls)
{
  const int l = blockIdx.y;
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nmaps) return;
  double acc = 0;
  for (int m = 0; m <= l; m++) {
    const int64_t idx = m > 0 ? l + (int64_t)m * lmax - (int64_t)m * (m - 1) / 2 : l;
    const double *row = alm + (size_t)idx * ldo + 2 * a;
    const double p = row[0] * row[0] + row[1] * row[1];
    acc += m ? 2.0 * p : p;
  }
  cls[(size_t)a * (lmax + 1) + l] = acc / (2.0 * l + 1.0);
}

// C_l of listed map pairs: one workgroup per (l, pair), 64 threads over m
__global__ __launch_bounds__(64) void k_cls_pairs(const double *__restrict__ alm, int lmax, int ldo,
                                                  const int *__restrict__ pairs, double *__restrict__ cls)
{
  const int l = blockIdx.x, pr = blockIdx.y;
  const int a = pairs[2 * pr], b = pairs[2 * pr + 1];
  double acc = 0;
  for (int m = threadIdx.x; m <= l; m += 64) {
    const int64_t idx = m > 0 ? l + (int64_t)m * lmax - (int64_t)m * (m - 1) / 2 : l;
    const double *row = alm + (size_t)idx * ldo;
    const double p = row[2 * a] * row[2 * b] + row[2 * a + 1] * row[2 * b + 1];
    acc += m ? 2.0 * p : p;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (threadIdx.x == 0) cls[(size_t)pr * (lmax + 1) + l] = acc / (2.0 * l + 1.0);
}

// alm[idx][2 t + c] -> out[t][idx][c]
__global__ __launch_bounds__(256) void k_alm_unpack(const double *__restrict__ alm, int64_t nalm, int nmaps, int ldo,
                                                    double *__restrict__ out)
{
  const int t = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nalm; i += (int64_t)gridDim.x * 256) {
    const double *src = alm + (size_t)i * ldo + 2 * t;
    double *dst = out + ((size_t)t * nalm + i) * 2;
    dst[0] = src[0];
    dst[1] = src[1];
  }
}

// G[ring][col] = sum_l lambda_lm(ring) alm[l][col] per (m, parity), l = m + parity + 2 row: the
// transpose of the analysis product, into the same 16-ring blocks of G.  A workgroup computes
// ALM_ROWS rings x CT columns, its 4 waves 2 x 2 blocks of (ALM_ROWS / 2) x (CT / 2); K runs over the
// rows of this parity, 16 at a time.  K averages a third of the analysis' (2 Nside rings), so the
// wider tile keeps each workgroup's loop and its A tile busy for longer.  Tiles with no rows
// (m = lmax, odd parity) write zeros.
template <int CT>
__global__ __launch_bounds__(256) void k_syn_gemm(const double *__restrict__ lam, const double *__restrict__ alm,
                                                  double *__restrict__ G, const int4 *__restrict__ tiles, int ntiles,
                                                  int ncoltiles, const int64_t *__restrict__ off,
                                                  int m0, int lmax, int R, int RP, int ldg, int ncols, int ldo)
{
  constexpr int RW = ALM_ROWS / 2, NI = RW / 16;
  constexpr int CW = CT / 2, NJ = CW / 16;
  constexpr int AL = ALM_ROWS / 16;                 // rings loaded per thread
  constexpr int BL = CT / 16;                       // columns loaded per thread
  // XCD-balanced order.  A unit is one (m, parity): its ring tiles x column tiles share the alm rows
  // (B), and the column tiles of one ring tile share the Legendre rows (A).  Units are dealt to the
  // eight XCDs in turn (block b runs on XCD b mod 8), so each XCD's L2 serves whole units and every XCD
  // gets every eighth unit: K (the l rows of the unit) falls with m, and contiguous ranges of units
  // would leave the XCD with the smallest m doing most of the work.
  const int b = blockIdx.x;
  const int nringtiles = (R + ALM_ROWS - 1) / ALM_ROWS, UB = nringtiles * ncoltiles;
  const int j = b >> 3, r = j % UB;
  const int unit = (j / UB) * 8 + (b & 7);
  const int group = unit * nringtiles + r / ncoltiles;
  if (group >= ntiles) return;
  const int4 tl = tiles[group];
  const int ct = r % ncoltiles;
  const int m = tl.x, par = tl.y, rt = tl.z, nrows = tl.w;
  __shared__ double As[GK][ALM_ROWS + 2];   // [k][ring]: A stored as loaded, read down a column
  __shared__ double Bs[GK][CT + 2];         // [k][column]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ring0 = rt * ALM_ROWS, col0 = ct * CT;
  const double *lrow = lam + (size_t)(off[m] - off[m0]) * RP;
  const int64_t base_idx = (int64_t)m * lmax - (int64_t)m * (m - 1) / 2;   // he_indexlm(l, m) - l
  dbl4 acc[NI][NJ];
  for (int i = 0; i < NI; i++)
    for (int j = 0; j < NJ; j++) acc[i][j] = (dbl4){0, 0, 0, 0};
  const int ak = tid >> 4, ar = (tid & 15) * AL;     // A loads: row k, AL consecutive rings
  const int bk = tid >> 4, bc = (tid & 15) * BL;     // B loads: row k, BL consecutive columns
  const bool aring_ok = ring0 + ar < RP;             // AL | RP: a chunk is wholly in range or not
  // ncols is even and bc a multiple of 2: each double2 of B is wholly in range or not
  bool bok[BL / 2];
  int bcol[BL / 2];
  for (int e = 0; e < BL / 2; e++) {
    bok[e] = col0 + bc + 2 * e < ncols;
    bcol[e] = bok[e] ? col0 + bc + 2 * e : 0;
  }
  const int wr = (wv >> 1) * RW, wc = (wv & 1) * CW;
  // operands out of range come from an in-range address and are zeroed by selects (no branches)
  double2 ra[AL / 2], rb[BL / 2];
  bool a_ok = false, b_ok = false;
  auto fetch = [&](int k0) {
    a_ok = aring_ok && k0 + ak < nrows;
    // row 0 of this m always exists (row `par` does not when m = lmax and the parity is odd)
    const double *pa = lrow + (size_t)(a_ok ? par + 2 * (k0 + ak) : 0) * RP + (aring_ok ? ring0 + ar : 0);
    for (int e = 0; e < AL / 2; e++) ra[e] = *(const double2 *)(pa + 2 * e);
    b_ok = k0 + bk < nrows;
    const int l = b_ok ? m + par + 2 * (k0 + bk) : m;
    const double *pb = alm + (size_t)(m > 0 ? base_idx + l : l) * ldo;
    for (int e = 0; e < BL / 2; e++) rb[e] = *(const double2 *)(pb + bcol[e]);
  };
  fetch(0);
  for (int k0 = 0; k0 < nrows; k0 += GK) {
    for (int e = 0; e < AL / 2; e++) {
      As[ak][ar + 2 * e] = a_ok ? ra[e].x : 0.0;
      As[ak][ar + 2 * e + 1] = a_ok ? ra[e].y : 0.0;
    }
    for (int e = 0; e < BL / 2; e++) {
      Bs[bk][bc + 2 * e] = b_ok && bok[e] ? rb[e].x : 0.0;
      Bs[bk][bc + 2 * e + 1] = b_ok && bok[e] ? rb[e].y : 0.0;
    }
    __syncthreads();
    fetch(k0 + GK < nrows ? k0 + GK : k0);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      double af[NI], bf[NJ];
      for (int i = 0; i < NI; i++) af[i] = As[kk + (lane >> 4)][wr + 16 * i + (lane & 15)];
      for (int j = 0; j < NJ; j++) bf[j] = Bs[kk + (lane >> 4)][wc + 16 * j + (lane & 15)];
      for (int i = 0; i < NI; i++)
        for (int j = 0; j < NJ; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int i = 0; i < NI; i++)
    for (int j = 0; j < NJ; j++)
      for (int v = 0; v < 4; v++) {
        const int ring = ring0 + wr + 16 * i + (lane >> 4) + 4 * v;
        const int col = col0 + wc + 16 * j + (lane & 15);
        if (ring < R && col < ncols)
          G[g_block(ring >> 4, m, par, lmax, ldg) + (size_t)(ring & 15) * ldg + col] = acc[i][j][v];
      }
}

// alms of the caller, [map][he_indexlm][re, im], into the plan's alm[idx][2 map + (re | im)]
__global__ __launch_bounds__(256) void k_alm_pack(const double *__restrict__ in, int64_t nalm, int nmaps, int ldo,
                                                  double *__restrict__ alm)
{
  const int t = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nalm; i += (int64_t)gridDim.x * 256) {
    const double *src = in + ((size_t)t * nalm + i) * 2;
    double *dst = alm + (size_t)i * ldo + 2 * t;
    dst[0] = src[0];
    dst[1] = src[1];
  }
}

// he_alter_alm (src/healpix_extra.c:402-421): a_lm *= exp(-l (l + 1) sigma_c^2 / 2).  Block row m;
// threads over (l, map) with the map fastest, so a wave touches consecutive alms of one row.
__global__ __launch_bounds__(256) void k_alm_beam(double *__restrict__ alm, int lmax, int nmaps, int ldo,
                                                  const double *__restrict__ sig2)
{
  const int m = blockIdx.y;
  const int64_t base_idx = m > 0 ? (int64_t)m * lmax - (int64_t)m * (m - 1) / 2 : 0;
  const int64_t n = (int64_t)(lmax - m + 1) * nmaps;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int l = m + (int)(i / nmaps), c = (int)(i % nmaps);
    double *a = alm + (size_t)(base_idx + l) * ldo + 2 * c;
    const double bl = exp(-0.5 * (double)l * (double)(l + 1) * sig2[c]);
    a[0] *= bl;
    a[1] *= bl;
  }
}

// add_noise (src/main_jt.c:270-311), rng_delta_gauss (src/common.c:154-164): += sqrt(-sigma2 ln(1-u2))
// cos(2 pi u1), (u1, u2) from Philox-4x32-10, counter (pixel, pixel >> 32, map, 0), key (seed, 'JONT')
template <typename TM>
__global__ __launch_bounds__(256) void k_noise(TM *__restrict__ maps, int64_t npix, int nmaps,
                                               const double *__restrict__ sig2, uint32_t seed)
{
  const int64_t n = npix * nmaps;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int c = (int)(i / npix);
    const int64_t p = i - (int64_t)c * npix;
    u32x4 ctr;
    ctr.v[0] = (uint32_t)p;
    ctr.v[1] = (uint32_t)((uint64_t)p >> 32);
    ctr.v[2] = (uint32_t)c;
    ctr.v[3] = 0;
    const u32x4 rr = philox4x32_10(ctr, seed, JT_NOISE_KEY);
    const double u1 = u53(rr.v[0], rr.v[1]), u2 = u53(rr.v[2], rr.v[3]);
    const double noise = sqrt(-sig2[c] * log(1.0 - u2)) * cos(2.0 * M_PI * u1);
    maps[i] = (TM)((double)maps[i] + noise);
  }
}

// apply_mask (src/main_jt.c:248-268)
template <typename TM>
__global__ __launch_bounds__(256) void k_mask(TM *__restrict__ maps, int64_t npix, int nmaps,
                                              const uint8_t *__restrict__ mask)
{
  const int64_t n = npix * nmaps;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    if (!mask[i % npix]) maps[i] = (TM)0;
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
hipError_t sht_set_lds_limits(size_t bytes)
{
  hipError_t e = hipFuncSetAttribute((const void *)k_ring_dft<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_ring_dft<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

template <typename TM, int L>
static void launch_ring_fft_t(const ShtLaunch &a, const TM *const *maps)
{
  using C = RingFftCfg<L>;
  const int nlines = a.nmaps * a.nr;
  hipLaunchKernelGGL((k_ring_fft<TM, L>), dim3((nlines + C::G - 1) / C::G), dim3(C::NT), 0, a.s, maps, a.mps, a.npix,
                     a.rings,
                     a.R, a.RP, a.lmax, a.ldg, a.w, a.G, a.r0, a.nmaps, nlines, (const cx<double> *)a.tw);
}

template <typename TM, int M>
static void launch_ring_blu_t(const ShtLaunch &a, const TM *const *maps)
{
  using C = BluCfg<M>;
  const int nlines = a.nmaps * a.nr;
  hipLaunchKernelGGL((k_ring_blu<TM, M>), dim3((nlines + C::G - 1) / C::G), dim3(C::NT), 0, a.s, maps, a.mps, a.npix,
                     a.rings,
                     a.R, a.RP, a.lmax, a.ldg, a.w, a.G, a.r0, a.nmaps, nlines, a.woff, (const cx<double> *)a.wtab,
                     a.boff, (const cx<double> *)a.B, (const cx<double> *)a.tw);
}

template <typename TM, int L>
static void launch_syn_fft_t(const ShtLaunch &a, TM *maps)
{
  using C = RingFftCfg<L>;
  const int nlines = a.nmaps * a.nr;
  hipLaunchKernelGGL((k_syn_fft<TM, L>), dim3((nlines + C::G - 1) / C::G), dim3(C::NT), 0, a.s, maps, a.npix, a.rings,
                     a.R, a.lmax, a.ldg, (const double *)a.G, a.r0, a.nmaps, nlines, (const cx<double> *)a.tw);
}

template <typename TM, int M>
static void launch_syn_blu_t(const ShtLaunch &a, TM *maps)
{
  using C = BluCfg<M>;
  const int nlines = a.nmaps * a.nr;
  hipLaunchKernelGGL((k_syn_blu<TM, M>), dim3((nlines + C::G - 1) / C::G), dim3(C::NT), 0, a.s, maps, a.npix, a.rings,
                     a.R, a.lmax, a.ldg, (const double *)a.G, a.r0, a.nmaps, nlines, a.woff,
                     (const cx<double> *)a.wtab, a.boff, (const cx<double> *)a.B, (const cx<double> *)a.tw);
}

template <typename TM>
static hipError_t launch_syn_t(const ShtLaunch &a, TM *maps)
{
  switch (a.kind) {
  case SHT_RING_SPLIT:
    hipLaunchKernelGGL(k_syn_dft<TM>, dim3(a.nmaps, a.nr), dim3(64), 0, a.s, maps, a.npix, a.rings, a.R, a.lmax, a.ldg,
                       (const double *)a.G, a.r0);
    break;
  case SHT_RING_FFT:
    switch (a.len) {
#define CRIME_FFT(LL) \
  case LL: launch_syn_fft_t<TM, LL>(a, maps); break;
      CRIME_FOR_EACH_RING_FFT_L(CRIME_FFT)
#undef CRIME_FFT
    default: return hipErrorInvalidValue;
    }
    break;
  case SHT_RING_BLUESTEIN:
    switch (a.len) {
#define CRIME_BLU(MM) \
  case MM: launch_syn_blu_t<TM, MM>(a, maps); break;
      CRIME_FOR_EACH_BLU_M(CRIME_BLU)
#undef CRIME_BLU
    default: return hipErrorInvalidValue;
    }
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_syn_rings(const ShtLaunch &a, void *maps, int fp64)
{
  return fp64 ? launch_syn_t<double>(a, (double *)maps) : launch_syn_t<float>(a, (float *)maps);
}

template <typename TM>
static hipError_t launch_rings_t(const ShtLaunch &a, const TM *const *maps)
{
  switch (a.kind) {
  case SHT_RING_SPLIT:
    hipLaunchKernelGGL(k_ring_dft<TM>, dim3(a.nmaps, a.nr), dim3(512), a.lds, a.s, maps, a.mps, a.npix, a.rings, a.R,
                       a.RP,
                       a.lmax, a.ldg, a.w, a.G, a.r0);
    break;
  case SHT_RING_FFT:
    switch (a.len) {
#define CRIME_FFT(LL) \
  case LL: launch_ring_fft_t<TM, LL>(a, maps); break;
      CRIME_FOR_EACH_RING_FFT_L(CRIME_FFT)
#undef CRIME_FFT
    default: return hipErrorInvalidValue;
    }
    break;
  case SHT_RING_BLUESTEIN:
    switch (a.len) {
#define CRIME_BLU(MM) \
  case MM: launch_ring_blu_t<TM, MM>(a, maps); break;
      CRIME_FOR_EACH_BLU_M(CRIME_BLU)
#undef CRIME_BLU
    default: return hipErrorInvalidValue;
    }
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int ring_fft_length(int64_t nside)
{
  switch (4 * nside) {
#define CRIME_FFT(LL) \
  case LL: return LL;
    CRIME_FOR_EACH_RING_FFT_L(CRIME_FFT)
#undef CRIME_FFT
  default: return 0;
  }
}

hipError_t launch_rings(const ShtLaunch &a, const void *const *mtab, int fp64)
{
  return fp64 ? launch_rings_t<double>(a, (const double *const *)mtab)
              : launch_rings_t<float>(a, (const float *const *)mtab);
}

template <int M>
static void launch_blu_setup_t(const ShtRing *rings, int r0, int nr, const int64_t *woff, const void *wtab,
                               const int64_t *boff, void *B, const void *tw, hipStream_t s)
{
  using C = BluCfg<M>;
  hipLaunchKernelGGL((k_blu_setup<M>), dim3((nr + C::G - 1) / C::G), dim3(C::NT), 0, s, rings, r0, nr, woff,
                     (const cx<double> *)wtab, boff, (cx<double> *)B, (const cx<double> *)tw);
}

hipError_t launch_blu_setup(int M, const ShtRing *rings, int r0, int nr, const int64_t *woff, const void *wtab,
                            const int64_t *boff, void *B, const void *tw, hipStream_t s)
{
  switch (M) {
#define CRIME_BLU(MM) \
  case MM: launch_blu_setup_t<MM>(rings, r0, nr, woff, wtab, boff, B, tw, s); break;
    CRIME_FOR_EACH_BLU_M(CRIME_BLU)
#undef CRIME_BLU
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_legendre(const double2 *coef, const int64_t *off, int m0, int m1, int lmax, const double *zs,
                           const double *sts, int R, int RP, double *lam, hipStream_t s)
{
  dim3 grid((RP + 255) / 256, m1 - m0);
  hipLaunchKernelGGL(k_legendre, grid, dim3(256), 0, s, coef, off, m0, lmax, zs, sts, R, RP, lam);
  return hipGetLastError();
}

hipError_t launch_alm_gemm(const double *lam, const double *G, const int4 *tiles, int ntiles, const int64_t *off,
                           int m0, int lmax, int RP, int ldg, int ncols, double *alm, int ldo, hipStream_t s)
{
  const int ncoltiles = (ncols + GT - 1) / GT;
  const long total = (long)ntiles * ncoltiles;
  const int per_xcd = (int)((total + 7) / 8);
  hipLaunchKernelGGL(k_alm_gemm, dim3(8 * per_xcd), dim3(256), 0, s, lam, G, tiles, ntiles, ncoltiles, per_xcd, off,
                     m0, lmax, RP, ldg, ncols, alm, ldo);
  return hipGetLastError();
}

hipError_t launch_cls_sets(const double *alm, int lmax, int nsets, int mps, int ldo, double *cls, hipStream_t s)
{
  const int nt = (mps + GT - 1) / GT;
  const int nl = lmax + 1;
  const int64_t ngroups = (int64_t)nsets * nt * (nt + 1) / 2;   // (set, tile pair) groups of nl blocks
  const int64_t nblk = ngroups * nl;
  if (nblk > 0x7fffffff) return hipErrorInvalidValue;
  // cost of logical block j of a group (l = lmax - j): its m chunks of 8 plus two for loads and stores
  std::vector<int64_t> pre(nl + 1, 0);
  for (int j = 0; j < nl; j++) pre[j + 1] = pre[j] + (lmax - j + 8) / 8 + 2;
  const int64_t wg = pre[nl];
  const double total = (double)wg * (double)ngroups;
  XcdRuns xr;
  xr.start[0] = 0;
  xr.start[8] = (int)nblk;
  for (int x = 1; x < 8; x++) {
    const double target = total * x / 8.0;
    int64_t q = (int64_t)(target / (double)wg);
    if (q > ngroups) q = ngroups;
    const int64_t rem = (int64_t)(target - (double)q * (double)wg);
    const int64_t i = std::lower_bound(pre.begin(), pre.end(), rem) - pre.begin();
    int64_t b = q * nl + (i < nl ? i : nl);
    if (b > nblk) b = nblk;
    if (b < xr.start[x - 1]) b = xr.start[x - 1];
    xr.start[x] = (int)b;
  }
  int per_xcd = 0;
  for (int x = 0; x < 8; x++) per_xcd = std::max(per_xcd, xr.start[x + 1] - xr.start[x]);
  if (per_xcd == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cls_cross, dim3((unsigned)(8 * per_xcd)), dim3(256), 0, s, alm, lmax, mps, ldo, nt, xr, cls);
  return hipGetLastError();
}

hipError_t launch_cls(const double *alm, int lmax, int nmaps, int ldo, int cross, double *cls, hipStream_t s)
{
  if (cross) {
    return launch_cls_sets(alm, lmax, 1, nmaps, ldo, cls, s);
  } else {
    dim3 grid((nmaps + 255) / 256, lmax + 1);
    hipLaunchKernelGGL(k_cls_auto, grid, dim3(256), 0, s, alm, lmax, nmaps, ldo, cls);
  }
  return hipGetLastError();
}

hipError_t launch_cls_pairs(const double *alm, int lmax, int ldo, int npairs, const int *pairs, double *cls,
                            hipStream_t s)
{
  hipLaunchKernelGGL(k_cls_pairs, dim3(lmax + 1, npairs), dim3(64), 0, s, alm, lmax, ldo, pairs, cls);
  return hipGetLastError();
}

hipError_t launch_syn_gemm(const double *lam, const double *alm, double *G, const int4 *tiles, int ntiles,
                           const int64_t *off, int m0, int lmax, int R, int RP, int ldg, int ncols, int ldo,
                           hipStream_t s)
{
  // SYN_COLS-column tiles unless the columns fit one 64-column tile
  const int ct = ncols > SYN_COLS / 2 ? SYN_COLS : 64;
  const int ncoltiles = (ncols + ct - 1) / ct;
  // the tile list holds every ring tile of each (m, parity) unit, consecutively
  const int nringtiles = (R + ALM_ROWS - 1) / ALM_ROWS;
  const int nunits = ntiles / nringtiles;
  if (nunits * nringtiles != ntiles) return hipErrorInvalidValue;
  const int per_xcd = (nunits + 7) / 8 * nringtiles * ncoltiles;
  if (ct == 64)
    hipLaunchKernelGGL(k_syn_gemm<64>, dim3(8 * per_xcd), dim3(256), 0, s, lam, alm, G, tiles, ntiles, ncoltiles,
                       off, m0, lmax, R, RP, ldg, ncols, ldo);
  else
    hipLaunchKernelGGL(k_syn_gemm<SYN_COLS>, dim3(8 * per_xcd), dim3(256), 0, s, lam, alm, G, tiles, ntiles,
                       ncoltiles, off, m0, lmax, R, RP, ldg, ncols, ldo);
  return hipGetLastError();
}

hipError_t launch_alm_pack(const double *in, int64_t nalm, int nmaps, int ldo, double *alm, hipStream_t s)
{
  int64_t nb = (nalm + 255) / 256;
  dim3 grid((unsigned)(nb < 4096 ? nb : 4096), nmaps);
  hipLaunchKernelGGL(k_alm_pack, grid, dim3(256), 0, s, in, nalm, nmaps, ldo, alm);
  return hipGetLastError();
}

hipError_t launch_alm_beam(double *alm, int lmax, int nmaps, int ldo, const double *sig2, hipStream_t s)
{
  const int64_t per_m = (int64_t)(lmax + 1) * nmaps;   // the largest row (m = 0)
  const int64_t bx = (per_m + 255) / 256;
  dim3 grid((unsigned)(bx < 64 ? bx : 64), lmax + 1);
  hipLaunchKernelGGL(k_alm_beam, grid, dim3(256), 0, s, alm, lmax, nmaps, ldo, sig2);
  return hipGetLastError();
}

static unsigned grid_stride_blocks(int64_t n)
{
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? b : 65536);
}

hipError_t launch_noise(void *maps, int fp64, int64_t npix, int nmaps, const double *sig2, uint32_t seed,
                        hipStream_t s)
{
  const dim3 grid(grid_stride_blocks(npix * nmaps));
  if (fp64) hipLaunchKernelGGL(k_noise<double>, grid, dim3(256), 0, s, (double *)maps, npix, nmaps, sig2, seed);
  else hipLaunchKernelGGL(k_noise<float>, grid, dim3(256), 0, s, (float *)maps, npix, nmaps, sig2, seed);
  return hipGetLastError();
}

hipError_t launch_mask(void *maps, int fp64, int64_t npix, int nmaps, const uint8_t *mask, hipStream_t s)
{
  const dim3 grid(grid_stride_blocks(npix * nmaps));
  if (fp64) hipLaunchKernelGGL(k_mask<double>, grid, dim3(256), 0, s, (double *)maps, npix, nmaps, mask);
  else hipLaunchKernelGGL(k_mask<float>, grid, dim3(256), 0, s, (float *)maps, npix, nmaps, mask);
  return hipGetLastError();
}

hipError_t launch_alm_unpack(const double *alm, int64_t nalm, int nmaps, int ldo, double *out, hipStream_t s)
{
  int64_t nb = (nalm + 255) / 256;
  dim3 grid((unsigned)(nb < 4096 ? nb : 4096), nmaps);
  hipLaunchKernelGGL(k_alm_unpack, grid, dim3(256), 0, s, alm, nalm, nmaps, ldo, out);
  return hipGetLastError();
}

}  // namespace crime
