// Runtime behind include/crime_sht.h: plan set-up (ring geometry, recursion coefficients, Legendre
// table, GEMM tile lists) and the map2alm / C_l entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "crime_sht.h"
#include "gh_sht.hpp"

using namespace crime;

namespace {

// Legendre table budget: the whole table stays resident when it fits (Nside 512, lmax 1535: 9.7 GB),
// otherwise it is rebuilt per call in chunks of m of at most this size
constexpr size_t LAMBDA_BUDGET = (size_t)16 << 30;
constexpr size_t LDS_LIMIT = 160 * 1024;
constexpr int BLU_MIN_N = 32;     // shorter rings: the split DFT
constexpr int BLU_MAX_M = 8192;   // largest compiled Bluestein convolution

struct Chunk {
  int m0, m1;        // [m0, m1)
  int tile0, ntile;  // slice of the analysis tile list
  int syn0, nsyn;    // slice of the synthesis tile list
};

}  // namespace

struct crime_sht {
  int64_t nside = 0, npix = 0;
  int lmax = 0, max_maps = 0, device = 0;
  int R = 0, RP = 0, ldg = 0, ldo = 0;
  int64_t nalm = 0;
  size_t lds = 0;
  hipStream_t stream = nullptr;
  // the ring groups write disjoint rings of G: all but every third run on two side streams, forked from
  // and joined back into `stream`, so that the many short Bluestein and split-DFT launches overlap
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  ShtRing *d_rings = nullptr;
  double *d_z = nullptr, *d_st = nullptr;
  double2 *d_coef = nullptr;
  int64_t *d_off = nullptr;
  int4 *d_tiles = nullptr, *d_syn_tiles = nullptr;
  double *d_sig2 = nullptr;     // per-map beam sigma^2 or noise variance
  uint8_t *d_mask = nullptr;    // Npix, allocated by the first mask call
  const void **d_mtab = nullptr;          // device table of map-set base pointers (max_maps entries)
  std::vector<const void *> h_mtab;       // its host image (kept for the asynchronous upload)
  bool syn_ok = true;           // every split-DFT ring is shorter than 32 pixels
  double *d_G = nullptr, *d_alm = nullptr, *d_lam = nullptr;
  double *d_cls = nullptr;
  size_t cls_cap = 0, cls_bytes = 0;
  double *d_unpack = nullptr;
  double *d_hmaps = nullptr;
  // ring transforms: groups of rings with one kernel each, their twiddles and Bluestein tables
  std::vector<ShtLaunch> groups;
  double *d_tw = nullptr;
  int64_t *d_woff = nullptr, *d_boff = nullptr;
  double *d_wtab = nullptr, *d_btab = nullptr;
  std::vector<Chunk> chunks;
  bool resident = false;
  std::vector<hipEvent_t> ev;
  crime_sht_stats st{};
  std::string err;
};

static thread_local std::string g_sht_err;

static int sht_error(crime_sht *p, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (p) p->err = buf;
  else g_sht_err = buf;
  return code;
}

#define SHT_HIP(p, call)                                                                          \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      (void)hipGetLastError(); /* reported here: a later launch check must not see it again */        \
      return sht_error(p, e_ == hipErrorOutOfMemory ? CRIME_GH_ENOMEM : CRIME_GH_EHIP, "%s: %s", #call,     \
                       hipGetErrorString(e_));                                                        \
    }                                                                                                 \
  } while (0)

extern "C" const char *crime_sht_last_error(const crime_sht_t *p) { return p ? p->err.c_str() : g_sht_err.c_str(); }

extern "C" void crime_sht_destroy(crime_sht_t *p)
{
  if (!p) return;
  (void)hipSetDevice(p->device);
  // every stream that may still run ring kernels reading these buffers is drained before the frees
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (int k = 0; k < 2; k++)
    if (p->side[k]) (void)hipStreamSynchronize(p->side[k]);
  void *bufs[] = {p->d_rings, p->d_z, p->d_st, p->d_coef, p->d_off, p->d_tiles, p->d_syn_tiles, p->d_sig2,
                  p->d_mask, (void *)p->d_mtab, p->d_G,
                  p->d_alm, p->d_lam, p->d_cls, p->d_unpack, p->d_hmaps, p->d_tw,
                  p->d_woff, p->d_boff, p->d_wtab, p->d_btab};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; k++) {
    if (p->join[k]) (void)hipEventDestroy(p->join[k]);
    if (p->side[k]) (void)hipStreamDestroy(p->side[k]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// Ring geometry of Gorski et al. 2005 (ApJ 622, 759) eqs. 4-9, northern rings ir = 1 .. 2 nside.
static void ring_geometry(int64_t nside, std::vector<ShtRing> &rings, std::vector<double> &z, std::vector<double> &st)
{
  const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  for (int64_t i = 1; i <= 2 * nside; i++) {
    ShtRing r;
    double zz, ss;
    if (i < nside) {
      double t = (double)(i * i) / (3.0 * (double)(nside * nside));
      zz = 1.0 - t;
      ss = sqrt(t * (2.0 - t));
      r.nphi = (int)(4 * i);
      r.s = 1;
      r.first_n = 2 * i * (i - 1);
      r.first_s = npix - 2 * i * (i + 1);
    } else {
      zz = i == 2 * nside ? 0.0 : (4.0 / 3.0) - 2.0 * (double)i / (3.0 * (double)nside);
      ss = sqrt((1.0 - zz) * (1.0 + zz));
      r.nphi = (int)(4 * nside);
      r.s = (int)((i - nside + 1) & 1);
      r.first_n = ncap + (i - nside) * 4 * nside;
      r.first_s = ncap + (3 * nside - i) * 4 * nside;   // ring 4 nside - i
    }
    rings.push_back(r);
    z.push_back(zz);
    st.push_back(ss);
  }
}

// How a ring of n pixels is transformed: the equatorial rings by a Stockham FFT of length 4 Nside
// where one is compiled, rings of 32 .. 4096 pixels by Bluestein, the rest by the split DFT.
static int ring_kind(int64_t nside, int n, int *len)
{
  const int fft = ring_fft_length(nside);
  if (fft && n == fft) {
    if (len) *len = fft;
    return SHT_RING_FFT;
  }
  if (n >= BLU_MIN_N) {
    int m = 1;
    while (m < 2 * n - 1) m *= 2;
    if (m <= BLU_MAX_M) {
      if (len) *len = m;
      return SHT_RING_BLUESTEIN;
    }
  }
  if (len) *len = 0;
  return SHT_RING_SPLIT;
}

static void push_cis(std::vector<double> &v, long double num, long double den)
{
  const long double a = 3.14159265358979323846264338327950288L * num / den;
  v.push_back((double)cosl(a));
  v.push_back((double)sinl(a));
}

// Groups of consecutive rings that share a kernel, twiddle tables exp(+2 pi i m / len) per length,
// and for the Bluestein rings the chirp w_j = exp(+i pi j^2 / n) (j^2 reduced mod 2n exactly) and
// the transformed filter B, computed here on the device by the Stockham engine.
static hipError_t plan_rings(crime_sht *p, const std::vector<ShtRing> &rings)
{
  const int R = (int)rings.size();
  std::vector<int> kind(R), len(R);
  for (int r = 0; r < R; r++) kind[r] = ring_kind(p->nside, rings[r].nphi, &len[r]);

  std::vector<double> tw;
  std::vector<std::pair<int, size_t>> tw_off;    // (len, offset in complex)
  auto twiddles = [&](int L) -> size_t {
    for (auto &e : tw_off)
      if (e.first == L) return e.second;
    const size_t off = tw.size() / 2;
    for (int m = 0; m < L; m++) push_cis(tw, 2.0L * m, (long double)L);
    tw_off.push_back({L, off});
    return off;
  };
  std::vector<int64_t> woff(R, 0), boff(R, 0);
  std::vector<double> wtab;
  std::vector<std::pair<int, int>> first_of_n;   // (n, ring whose tables every ring of this n shares)
  int64_t btab = 0;
  std::vector<size_t> group_tw;
  for (int r = 0; r < R;) {
    int e = r + 1;
    while (e < R && kind[e] == kind[r] && len[e] == len[r]) e++;
    if (kind[r] == SHT_RING_SPLIT && rings[r].nphi > 32) p->syn_ok = false;
    ShtLaunch g{};
    g.kind = kind[r];
    g.len = len[r];
    g.r0 = r;
    g.nr = e - r;
    p->groups.push_back(g);
    group_tw.push_back(g.kind == SHT_RING_SPLIT ? 0 : twiddles(g.len));
    for (int q = r; q < e && g.kind == SHT_RING_BLUESTEIN; q++) {
      const int n = rings[q].nphi;
      int src = -1;
      for (auto &f : first_of_n)
        if (f.first == n) src = f.second;
      if (src >= 0) {
        woff[q] = woff[src];
        boff[q] = boff[src];
        continue;
      }
      first_of_n.push_back({n, q});
      woff[q] = (int64_t)(wtab.size() / 2);
      for (int64_t j = 0; j < n; j++) push_cis(wtab, (long double)((j * j) % (2 * (int64_t)n)), (long double)n);
      boff[q] = btab;
      btab += g.len;
    }
    r = e;
  }
  hipError_t err;
  if ((err = hipMalloc((void **)&p->d_tw, std::max<size_t>(tw.size(), 2) * sizeof(double))) != hipSuccess) return err;
  if (!tw.empty() &&
      (err = hipMemcpy(p->d_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return err;
  for (size_t i = 0; i < p->groups.size(); i++) p->groups[i].tw = p->d_tw + 2 * group_tw[i];
  if (btab == 0) return hipSuccess;
  if ((err = hipMalloc((void **)&p->d_woff, R * sizeof(int64_t))) != hipSuccess) return err;
  if ((err = hipMalloc((void **)&p->d_boff, R * sizeof(int64_t))) != hipSuccess) return err;
  if ((err = hipMalloc((void **)&p->d_wtab, wtab.size() * sizeof(double))) != hipSuccess) return err;
  if ((err = hipMalloc((void **)&p->d_btab, (size_t)btab * 2 * sizeof(double))) != hipSuccess) return err;
  if ((err = hipMemcpy(p->d_woff, woff.data(), R * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess) return err;
  if ((err = hipMemcpy(p->d_boff, boff.data(), R * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess) return err;
  if ((err = hipMemcpy(p->d_wtab, wtab.data(), wtab.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return err;
  for (ShtLaunch &g : p->groups) {
    if (g.kind != SHT_RING_BLUESTEIN) continue;
    g.woff = p->d_woff;
    g.wtab = p->d_wtab;
    g.boff = p->d_boff;
    g.B = p->d_btab;
    // rings of one length share their tables: fill them once
    const bool shared = rings[g.r0].nphi == rings[g.r0 + g.nr - 1].nphi;
    if ((err = launch_blu_setup(g.len, p->d_rings, g.r0, shared ? 1 : g.nr, p->d_woff, p->d_wtab, p->d_boff,
                                p->d_btab, g.tw, p->stream)) != hipSuccess)
      return err;
  }
  return hipStreamSynchronize(p->stream);
}

static int create_impl(crime_sht *p, int64_t nside, int lmax, int max_maps, int device)
{
  if (nside < 1 || nside > 8192) return sht_error(p, CRIME_GH_EINVAL, "nside %lld outside [1, 8192]", (long long)nside);
  if (lmax < 0 || lmax > 4 * nside - 1)
    return sht_error(p, CRIME_GH_EINVAL, "lmax %d outside [0, 4 nside - 1]", lmax);
  if (max_maps < 1) return sht_error(p, CRIME_GH_EINVAL, "max_maps must be >= 1");
  p->nside = nside;
  p->npix = 12 * nside * nside;
  p->lmax = lmax;
  p->max_maps = max_maps;
  p->device = device;
  p->R = (int)(2 * nside);
  p->RP = (p->R + 15) / 16 * 16;
  p->ldg = 2 * max_maps;
  p->ldo = 2 * max_maps;
  p->nalm = (int64_t)(lmax + 1) * (lmax + 2) / 2;
  p->lds = 0;
  for (int64_t i = 1; i <= 2 * nside; i++) {
    const int n = (int)(4 * (i < nside ? i : nside));
    if (ring_kind(nside, n, nullptr) == SHT_RING_SPLIT) p->lds = std::max(p->lds, ring_dft_lds_bytes(n, lmax));
  }
  if (p->lds > LDS_LIMIT)
    return sht_error(p, CRIME_GH_EINVAL, "nside %lld: a ring spectrum needs %zu bytes of LDS (limit %zu)",
                     (long long)nside, p->lds, LDS_LIMIT);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return sht_error(p, CRIME_GH_ENODEV, "no HIP device");
  if (device < 0 || device >= ndev) return sht_error(p, CRIME_GH_ENODEV, "device %d of %d", device, ndev);
  SHT_HIP(p, hipSetDevice(device));
  SHT_HIP(p, hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  for (int k = 0; k < 2; k++) {
    SHT_HIP(p, hipStreamCreateWithFlags(&p->side[k], hipStreamNonBlocking));
    SHT_HIP(p, hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming));
  }
  SHT_HIP(p, hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
  SHT_HIP(p, sht_set_lds_limits(p->lds));

  std::vector<ShtRing> rings;
  std::vector<double> z, st;
  ring_geometry(nside, rings, z, st);
  SHT_HIP(p, hipMalloc((void **)&p->d_rings, rings.size() * sizeof(ShtRing)));
  SHT_HIP(p, hipMalloc((void **)&p->d_z, z.size() * sizeof(double)));
  SHT_HIP(p, hipMalloc((void **)&p->d_st, st.size() * sizeof(double)));
  SHT_HIP(p, hipMemcpy(p->d_rings, rings.data(), rings.size() * sizeof(ShtRing), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMemcpy(p->d_z, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMemcpy(p->d_st, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));

  SHT_HIP(p, plan_rings(p, rings));

  // recursion coefficients per m: [C_m, sqrt(2m+3), (a_l, b_l) for l = m+2 .. lmax]
  std::vector<int64_t> off(lmax + 2);
  std::vector<double2> coef((size_t)p->nalm);
  long double cm = logl(1.0L / sqrtl(4.0L * (long double)M_PI)) / logl(2.0L);
  off[0] = 0;
  for (int m = 0; m <= lmax; m++) {
    if (m > 0) cm += 0.5L * log2l((2.0L * m + 1.0L) / (2.0L * m));
    off[m + 1] = off[m] + (lmax - m + 1);
    coef[off[m]] = make_double2((double)cm, 0.0);
    if (m < lmax) coef[off[m] + 1] = make_double2((double)sqrtl(2.0L * m + 3.0L), 0.0);
    for (int l = m + 2; l <= lmax; l++) {
      long double l2 = (long double)l * l, m2 = (long double)m * m, q2 = (long double)(l - 1) * (l - 1);
      coef[off[m] + (l - m)] = make_double2((double)sqrtl((4.0L * l2 - 1.0L) / (l2 - m2)),
                                            (double)sqrtl((q2 - m2) / (4.0L * q2 - 1.0L)));
    }
  }
  SHT_HIP(p, hipMalloc((void **)&p->d_coef, coef.size() * sizeof(double2)));
  SHT_HIP(p, hipMalloc((void **)&p->d_off, off.size() * sizeof(int64_t)));
  SHT_HIP(p, hipMemcpy(p->d_coef, coef.data(), coef.size() * sizeof(double2), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMemcpy(p->d_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));

  // m chunks of the Legendre table and the GEMM tiles of each chunk
  const size_t row_bytes = (size_t)p->RP * sizeof(double);
  std::vector<int4> tiles, syn_tiles;
  size_t max_chunk_rows = 0;
  const int nringtiles = (p->R + ALM_ROWS - 1) / ALM_ROWS;
  for (int m = 0; m <= lmax;) {
    Chunk c;
    c.m0 = m;
    c.tile0 = (int)tiles.size();
    c.syn0 = (int)syn_tiles.size();
    size_t rows = 0;
    while (m <= lmax && (rows == 0 || (rows + (size_t)(lmax - m + 1)) * row_bytes <= LAMBDA_BUDGET)) {
      rows += (size_t)(lmax - m + 1);
      const int L = lmax - m + 1;
      for (int par = 0; par < 2; par++) {
        const int nr = (L - par + 1) / 2;
        for (int rt = 0; rt * ALM_ROWS < nr; rt++) tiles.push_back(make_int4(m, par, rt, nr));
        // synthesis: every ring tile, also when this parity has no rows (its spectra are zero)
        for (int rt = 0; rt < nringtiles; rt++) syn_tiles.push_back(make_int4(m, par, rt, nr));
      }
      m++;
    }
    c.m1 = m;
    c.ntile = (int)tiles.size() - c.tile0;
    c.nsyn = (int)syn_tiles.size() - c.syn0;
    if (rows > max_chunk_rows) max_chunk_rows = rows;
    p->chunks.push_back(c);
  }
  SHT_HIP(p, hipMalloc((void **)&p->d_tiles, tiles.size() * sizeof(int4)));
  SHT_HIP(p, hipMemcpy(p->d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMalloc((void **)&p->d_syn_tiles, syn_tiles.size() * sizeof(int4)));
  SHT_HIP(p, hipMemcpy(p->d_syn_tiles, syn_tiles.data(), syn_tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMalloc((void **)&p->d_sig2, (size_t)max_maps * sizeof(double)));
  SHT_HIP(p, hipMalloc((void **)&p->d_mtab, (size_t)max_maps * sizeof(void *)));
  p->h_mtab.resize(max_maps);
  SHT_HIP(p, hipMalloc((void **)&p->d_lam, max_chunk_rows * row_bytes));

  const size_t gsz = (size_t)(lmax + 1) * 2 * p->RP * p->ldg * sizeof(double);
  SHT_HIP(p, hipMalloc((void **)&p->d_G, gsz));
  SHT_HIP(p, hipMemset(p->d_G, 0, gsz));   // rows of the padding rings stay zero
  SHT_HIP(p, hipMalloc((void **)&p->d_alm, (size_t)p->nalm * p->ldo * sizeof(double)));
  p->ev.resize(4 * p->chunks.size() + 8);   // analysis events, then synthesis events
  for (auto &e : p->ev) SHT_HIP(p, hipEventCreate(&e));

  p->resident = p->chunks.size() == 1;
  if (p->resident) {
    const Chunk &c = p->chunks[0];
    SHT_HIP(p, launch_legendre(p->d_coef, p->d_off, c.m0, c.m1, lmax, p->d_z, p->d_st, p->R, p->RP, p->d_lam,
                               p->stream));
    SHT_HIP(p, hipStreamSynchronize(p->stream));
  }
  return 0;
}

extern "C" int crime_sht_create(int64_t nside, int lmax, int max_maps, int device, crime_sht_t **out)
{
  if (!out) return sht_error(nullptr, CRIME_GH_EINVAL, "out is NULL");
  *out = nullptr;
  crime_sht *p = new crime_sht;
  int rc = create_impl(p, nside, lmax, max_maps, device);
  if (rc) {
    g_sht_err = p->err;
    crime_sht_destroy(p);
    return rc;
  }
  *out = p;
  return 0;
}

static float elapsed(hipEvent_t a, hipEvent_t b)
{
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

// ring spectra + alm GEMMs of nmaps device maps into p->d_alm; stage times into p->st
// a ring group's launch record for nmaps maps
static ShtLaunch ring_launch(const crime_sht *p, ShtLaunch g, int nmaps)
{
  g.nmaps = nmaps;
  g.npix = p->npix;
  g.rings = p->d_rings;
  g.R = p->R;
  g.RP = p->RP;
  g.lmax = p->lmax;
  g.ldg = p->ldg;
  g.w = 4.0 * M_PI / (double)p->npix;
  g.lds = p->lds;
  g.G = p->d_G;
  g.s = p->stream;
  return g;
}

// Launch every ring group through `launch` (called with the group and the stream to use): group i on
// the plan's stream when i % 3 == 0, else on side stream i % 3 - 1.  The side streams start after the
// work already queued on the plan's stream, and the plan's stream continues after both.
template <class Launch>
static int run_ring_groups(crime_sht *p, Launch launch)
{
  SHT_HIP(p, hipEventRecord(p->fork, p->stream));
  for (int k = 0; k < 2; k++) SHT_HIP(p, hipStreamWaitEvent(p->side[k], p->fork, 0));
  int rc = 0;
  for (size_t i = 0; i < p->groups.size() && !rc; i++) {
    const hipStream_t s = i % 3 == 0 ? p->stream : p->side[i % 3 - 1];
    rc = launch(p->groups[i], s);
  }
  // join the side streams into the plan's stream also after a failed launch, so that work already
  // queued on them is ordered before anything the caller does next (a destroy included)
  for (int k = 0; k < 2; k++) {
    hipError_t e = hipEventRecord(p->join[k], p->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->stream, p->join[k], 0);
    if (e != hipSuccess && !rc) {
      (void)hipGetLastError();
      rc = sht_error(p, CRIME_GH_EHIP, "ring groups join: %s", hipGetErrorString(e));
    }
  }
  return rc;
}

// ring spectra + alm GEMMs of nsets sets of mps device maps each (set s at d_maps[s]) into p->d_alm,
// columns 2 (s mps + t) + (re | im); stage times into p->st
static int analyse_sets(crime_sht *p, int nsets, const void *const *d_maps, int mps, int fp64)
{
  if (nsets < 1 || mps < 1 || (int64_t)nsets * mps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "%d sets of %d maps: outside [1, %d] maps", nsets, mps, p->max_maps);
  if (!d_maps) return sht_error(p, CRIME_GH_EINVAL, "maps are NULL");
  for (int i = 0; i < nsets; i++) {
    if (!d_maps[i]) return sht_error(p, CRIME_GH_EINVAL, "maps of set %d are NULL", i);
    p->h_mtab[i] = d_maps[i];
  }
  const int nmaps = nsets * mps;
  hipEvent_t *ev = p->ev.data();
  SHT_HIP(p, hipMemcpyAsync(p->d_mtab, p->h_mtab.data(), nsets * sizeof(void *), hipMemcpyHostToDevice, p->stream));
  SHT_HIP(p, hipEventRecord(ev[0], p->stream));
  int rc = run_ring_groups(p, [&](const ShtLaunch &g, hipStream_t s) -> int {
    ShtLaunch a = ring_launch(p, g, nmaps);
    a.mps = mps;
    a.s = s;
    SHT_HIP(p, launch_rings(a, (const void *const *)p->d_mtab, fp64));
    return 0;
  });
  if (rc) return rc;
  SHT_HIP(p, hipEventRecord(ev[1], p->stream));
  for (size_t i = 0; i < p->chunks.size(); i++) {
    const Chunk &c = p->chunks[i];
    if (!p->resident)
      SHT_HIP(p, launch_legendre(p->d_coef, p->d_off, c.m0, c.m1, p->lmax, p->d_z, p->d_st, p->R, p->RP, p->d_lam,
                                 p->stream));
    SHT_HIP(p, hipEventRecord(ev[2 + 2 * i], p->stream));
    SHT_HIP(p, launch_alm_gemm(p->d_lam, p->d_G, p->d_tiles + c.tile0, c.ntile, p->d_off, c.m0, p->lmax, p->RP,
                               p->ldg, 2 * nmaps, p->d_alm, p->ldo, p->stream));
    SHT_HIP(p, hipEventRecord(ev[3 + 2 * i], p->stream));
  }
  return 0;
}

static int analyse(crime_sht *p, int nmaps, const void *d_maps, int fp64)
{
  return analyse_sets(p, 1, &d_maps, nmaps, fp64);
}

static int finish(crime_sht *p, bool with_cls)
{
  hipEvent_t *ev = p->ev.data();
  const size_t nc = p->chunks.size();
  hipEvent_t last = ev[1 + 2 * nc];
  if (with_cls) {
    last = ev[2 + 2 * nc];
    SHT_HIP(p, hipEventRecord(last, p->stream));
  }
  SHT_HIP(p, hipStreamSynchronize(p->stream));
  p->st = crime_sht_stats{};
  p->st.ms_rings = elapsed(ev[0], ev[1]);
  hipEvent_t prev = ev[1];
  for (size_t i = 0; i < nc; i++) {
    p->st.ms_legendre += elapsed(prev, ev[2 + 2 * i]);
    p->st.ms_alm += elapsed(ev[2 + 2 * i], ev[3 + 2 * i]);
    prev = ev[3 + 2 * i];
  }
  if (with_cls) p->st.ms_cls = elapsed(ev[1 + 2 * nc], last);
  p->st.ms_total = elapsed(ev[0], last);
  return 0;
}

extern "C" int crime_sht_map2alm_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, double *alms)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  if (!alms) return sht_error(p, CRIME_GH_EINVAL, "alms are NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  int rc = analyse(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  rc = finish(p, false);
  if (rc) return rc;
  const size_t bytes = (size_t)p->nalm * p->max_maps * 2 * sizeof(double);
  if (!p->d_unpack) SHT_HIP(p, hipMalloc((void **)&p->d_unpack, bytes));
  SHT_HIP(p, launch_alm_unpack(p->d_alm, p->nalm, nmaps, p->ldo, p->d_unpack, p->stream));
  SHT_HIP(p, hipMemcpyAsync(alms, p->d_unpack, (size_t)p->nalm * nmaps * 2 * sizeof(double), hipMemcpyDeviceToHost,
                            p->stream));
  SHT_HIP(p, hipStreamSynchronize(p->stream));
  return 0;
}

// host maps into the plan's staging buffer
static int upload_maps(crime_sht *p, int nmaps, const double *maps)
{
  if (!maps) return sht_error(p, CRIME_GH_EINVAL, "maps are NULL");
  if (nmaps < 1 || nmaps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "nmaps %d outside [1, %d]", nmaps, p->max_maps);
  SHT_HIP(p, hipSetDevice(p->device));
  if (!p->d_hmaps) SHT_HIP(p, hipMalloc((void **)&p->d_hmaps, (size_t)p->npix * p->max_maps * sizeof(double)));
  SHT_HIP(p, hipMemcpyAsync(p->d_hmaps, maps, (size_t)p->npix * nmaps * sizeof(double), hipMemcpyHostToDevice,
                            p->stream));
  return 0;
}

extern "C" int crime_sht_map2alm(crime_sht_t *p, int nmaps, const double *maps, double *alms)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  int rc = upload_maps(p, nmaps, maps);
  if (rc) return rc;
  return crime_sht_map2alm_device(p, nmaps, p->d_hmaps, 1, alms);
}

extern "C" int crime_sht_anafast_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, int cross,
                                        double *cls)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  int rc = analyse(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  const size_t n = (size_t)(cross ? nmaps : 1) * nmaps * (p->lmax + 1);
  if (n > p->cls_cap) {
    if (p->d_cls) SHT_HIP(p, hipFree(p->d_cls));
    p->d_cls = nullptr;
    p->cls_cap = 0;
    SHT_HIP(p, hipMalloc((void **)&p->d_cls, n * sizeof(double)));
    p->cls_cap = n;
  }
  SHT_HIP(p, launch_cls(p->d_alm, p->lmax, nmaps, p->ldo, cross, p->d_cls, p->stream));
  rc = finish(p, true);
  if (rc) return rc;
  p->cls_bytes = n * sizeof(double);
  if (cls) SHT_HIP(p, hipMemcpy(cls, p->d_cls, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int crime_sht_anafast_sets_device(crime_sht_t *p, int nsets, const void *const *d_maps, int maps_per_set,
                                             int maps_fp64, double *cls)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  int rc = analyse_sets(p, nsets, d_maps, maps_per_set, maps_fp64);
  if (rc) return rc;
  const size_t n = (size_t)nsets * maps_per_set * maps_per_set * (p->lmax + 1);
  if (n > p->cls_cap) {
    if (p->d_cls) SHT_HIP(p, hipFree(p->d_cls));
    p->d_cls = nullptr;
    p->cls_cap = 0;
    SHT_HIP(p, hipMalloc((void **)&p->d_cls, n * sizeof(double)));
    p->cls_cap = n;
  }
  SHT_HIP(p, launch_cls_sets(p->d_alm, p->lmax, nsets, maps_per_set, p->ldo, p->d_cls, p->stream));
  rc = finish(p, true);
  if (rc) return rc;
  p->cls_bytes = n * sizeof(double);
  if (cls) SHT_HIP(p, hipMemcpy(cls, p->d_cls, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int crime_sht_device_cls(const crime_sht_t *p, const double **dptr, size_t *bytes)
{
  if (!p || !dptr || !bytes) return CRIME_GH_EINVAL;
  *dptr = p->d_cls;
  *bytes = p->cls_bytes;
  return 0;
}

extern "C" int crime_sht_cls_pairs_device(crime_sht_t *p, int nmaps, const void *d_maps, int maps_fp64, int npairs,
                                          const int32_t *pairs, double *cls)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  if (!pairs || !cls || npairs < 1 || npairs > 65535) return sht_error(p, CRIME_GH_EINVAL, "bad pair list");
  for (int i = 0; i < 2 * npairs; i++)
    if (pairs[i] < 0 || pairs[i] >= nmaps) return sht_error(p, CRIME_GH_EINVAL, "pair index %d outside [0, %d)", pairs[i], nmaps);
  SHT_HIP(p, hipSetDevice(p->device));
  int rc = analyse(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  const size_t n = (size_t)npairs * (p->lmax + 1);
  const size_t need = n + (size_t)npairs;  // spectra, then the pair list (as int32 pairs in doubles' room)
  if (need > p->cls_cap) {
    if (p->d_cls) SHT_HIP(p, hipFree(p->d_cls));
    p->d_cls = nullptr;
    p->cls_cap = 0;
    SHT_HIP(p, hipMalloc((void **)&p->d_cls, need * sizeof(double)));
    p->cls_cap = need;
  }
  int *d_pairs = (int *)(p->d_cls + n);
  SHT_HIP(p, hipMemcpyAsync(d_pairs, pairs, 2 * (size_t)npairs * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
  SHT_HIP(p, launch_cls_pairs(p->d_alm, p->lmax, p->ldo, npairs, d_pairs, p->d_cls, p->stream));
  rc = finish(p, true);
  if (rc) return rc;
  p->cls_bytes = n * sizeof(double);
  SHT_HIP(p, hipMemcpy(cls, p->d_cls, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int crime_sht_anafast(crime_sht_t *p, int nmaps, const double *maps, int cross, double *cls)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  int rc = upload_maps(p, nmaps, maps);
  if (rc) return rc;
  return crime_sht_anafast_device(p, nmaps, p->d_hmaps, 1, cross, cls);
}

// ---------------------------------------------------------------------------------------------
// synthesis: alm2map, the JoinT-like beam, noise and mask
// ---------------------------------------------------------------------------------------------
static hipEvent_t *syn_events(crime_sht *p) { return p->ev.data() + 2 * p->chunks.size() + 4; }

// p->d_alm (nmaps maps) -> d_out (FP32 or FP64 RING maps): per m-chunk Legendre table (unless
// resident) and the transposed GEMM into G, then the ring syntheses
static int synthesize(crime_sht *p, int nmaps, void *d_out, int fp64)
{
  hipEvent_t *ev = syn_events(p);
  SHT_HIP(p, hipEventRecord(ev[0], p->stream));
  for (size_t i = 0; i < p->chunks.size(); i++) {
    const Chunk &c = p->chunks[i];
    if (!p->resident)
      SHT_HIP(p, launch_legendre(p->d_coef, p->d_off, c.m0, c.m1, p->lmax, p->d_z, p->d_st, p->R, p->RP, p->d_lam,
                                 p->stream));
    SHT_HIP(p, hipEventRecord(ev[1 + 2 * i], p->stream));
    SHT_HIP(p, launch_syn_gemm(p->d_lam, p->d_alm, p->d_G, p->d_syn_tiles + c.syn0, c.nsyn, p->d_off, c.m0, p->lmax,
                               p->R, p->RP, p->ldg, 2 * nmaps, p->ldo, p->stream));
    SHT_HIP(p, hipEventRecord(ev[2 + 2 * i], p->stream));
  }
  int rc = run_ring_groups(p, [&](const ShtLaunch &g, hipStream_t s) -> int {
    ShtLaunch a = ring_launch(p, g, nmaps);
    a.s = s;
    SHT_HIP(p, launch_syn_rings(a, d_out, fp64));
    return 0;
  });
  if (rc) return rc;
  SHT_HIP(p, hipEventRecord(ev[1 + 2 * p->chunks.size()], p->stream));
  return 0;
}

// after a synthesis: wait, and add its stage times to p->st (cleared first unless `accumulate`)
static int finish_syn(crime_sht *p, bool accumulate)
{
  SHT_HIP(p, hipStreamSynchronize(p->stream));
  hipEvent_t *ev = syn_events(p);
  const size_t nc = p->chunks.size();
  if (!accumulate) p->st = crime_sht_stats{};
  hipEvent_t prev = ev[0];
  for (size_t i = 0; i < nc; i++) {
    p->st.ms_legendre += elapsed(prev, ev[1 + 2 * i]);
    p->st.ms_alm += elapsed(ev[1 + 2 * i], ev[2 + 2 * i]);
    prev = ev[2 + 2 * i];
  }
  p->st.ms_rings += elapsed(ev[2 * nc], ev[1 + 2 * nc]);
  if (!accumulate) p->st.ms_total = elapsed(ev[0], ev[1 + 2 * nc]);
  return 0;
}

static int syn_check(crime_sht *p, int nmaps)
{
  if (!p->syn_ok) return sht_error(p, CRIME_GH_EINVAL, "nside %lld: no synthesis plan", (long long)p->nside);
  if (nmaps < 1 || nmaps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "nmaps %d outside [1, %d]", nmaps, p->max_maps);
  return 0;
}

extern "C" int crime_sht_alm2map_device(crime_sht_t *p, int nmaps, const double *d_alms, void *d_maps, int maps_fp64)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  int rc = syn_check(p, nmaps);
  if (rc) return rc;
  if (!d_alms || !d_maps) return sht_error(p, CRIME_GH_EINVAL, "alms or maps are NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  SHT_HIP(p, launch_alm_pack(d_alms, p->nalm, nmaps, p->ldo, p->d_alm, p->stream));
  rc = synthesize(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  return finish_syn(p, false);
}

extern "C" int crime_sht_alm2map(crime_sht_t *p, int nmaps, const double *alms, double *maps)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  int rc = syn_check(p, nmaps);
  if (rc) return rc;
  if (!alms || !maps) return sht_error(p, CRIME_GH_EINVAL, "alms or maps are NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  if (!p->d_unpack)
    SHT_HIP(p, hipMalloc((void **)&p->d_unpack, (size_t)p->nalm * p->max_maps * 2 * sizeof(double)));
  if (!p->d_hmaps) SHT_HIP(p, hipMalloc((void **)&p->d_hmaps, (size_t)p->npix * p->max_maps * sizeof(double)));
  SHT_HIP(p, hipMemcpyAsync(p->d_unpack, alms, (size_t)p->nalm * nmaps * 2 * sizeof(double), hipMemcpyHostToDevice,
                            p->stream));
  rc = crime_sht_alm2map_device(p, nmaps, p->d_unpack, p->d_hmaps, 1);
  if (rc) return rc;
  SHT_HIP(p, hipMemcpy(maps, p->d_hmaps, (size_t)p->npix * nmaps * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// FWHM2SIGMA (src/healpix_extra.c:389), used by he_generate_beam_window (:390-400): arcmin -> radians / sqrt(8 ln 2)
static const double SHT_FWHM2SIGMA = 0.00012352884853326381;

extern "C" int crime_sht_smooth_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64,
                                       const double *fwhm_arcmin)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  int rc = syn_check(p, nmaps);
  if (rc) return rc;
  if (!fwhm_arcmin) return sht_error(p, CRIME_GH_EINVAL, "fwhm is NULL");
  std::vector<double> sig2(nmaps);
  for (int c = 0; c < nmaps; c++) {
    if (!(fwhm_arcmin[c] >= 0) || !std::isfinite(fwhm_arcmin[c]))
      return sht_error(p, CRIME_GH_EINVAL, "fwhm[%d] = %g", c, fwhm_arcmin[c]);
    const double sg = SHT_FWHM2SIGMA * fwhm_arcmin[c];
    sig2[c] = sg * sg;
  }
  SHT_HIP(p, hipSetDevice(p->device));
  rc = analyse(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  SHT_HIP(p, hipMemcpyAsync(p->d_sig2, sig2.data(), nmaps * sizeof(double), hipMemcpyHostToDevice, p->stream));
  SHT_HIP(p, launch_alm_beam(p->d_alm, p->lmax, nmaps, p->ldo, p->d_sig2, p->stream));
  rc = synthesize(p, nmaps, d_maps, maps_fp64);
  if (rc) return rc;
  rc = finish(p, false);
  if (rc) return rc;
  const size_t nc = p->chunks.size();
  hipEvent_t *ev = p->ev.data(), *sv = syn_events(p);
  rc = finish_syn(p, true);
  if (rc) return rc;
  p->st.ms_cls = elapsed(ev[1 + 2 * nc], sv[0]);   // the beam
  p->st.ms_total = elapsed(ev[0], sv[1 + 2 * nc]);
  return 0;
}

extern "C" int crime_sht_add_noise_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64,
                                          const double *sigma2, uint32_t seed)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  if (nmaps < 1 || nmaps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "nmaps %d outside [1, %d]", nmaps, p->max_maps);
  if (!d_maps || !sigma2) return sht_error(p, CRIME_GH_EINVAL, "maps or sigma2 are NULL");
  for (int c = 0; c < nmaps; c++)
    if (!(sigma2[c] >= 0) || !std::isfinite(sigma2[c]))
      return sht_error(p, CRIME_GH_EINVAL, "sigma2[%d] = %g", c, sigma2[c]);
  SHT_HIP(p, hipSetDevice(p->device));
  SHT_HIP(p, hipMemcpyAsync(p->d_sig2, sigma2, nmaps * sizeof(double), hipMemcpyHostToDevice, p->stream));
  SHT_HIP(p, launch_noise(d_maps, maps_fp64, p->npix, nmaps, p->d_sig2, seed, p->stream));
  SHT_HIP(p, hipStreamSynchronize(p->stream));
  return 0;
}

extern "C" int crime_sht_mask_device(crime_sht_t *p, int nmaps, void *d_maps, int maps_fp64, const uint8_t *mask)
{
  if (!p) return sht_error(nullptr, CRIME_GH_EINVAL, "plan is NULL");
  if (nmaps < 1 || nmaps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "nmaps %d outside [1, %d]", nmaps, p->max_maps);
  if (!d_maps || !mask) return sht_error(p, CRIME_GH_EINVAL, "maps or mask are NULL");
  SHT_HIP(p, hipSetDevice(p->device));
  if (!p->d_mask) SHT_HIP(p, hipMalloc((void **)&p->d_mask, (size_t)p->npix));
  SHT_HIP(p, hipMemcpyAsync(p->d_mask, mask, (size_t)p->npix, hipMemcpyHostToDevice, p->stream));
  SHT_HIP(p, launch_mask(d_maps, maps_fp64, p->npix, nmaps, p->d_mask, p->stream));
  SHT_HIP(p, hipStreamSynchronize(p->stream));
  return 0;
}

extern "C" int crime_sht_get_stats(const crime_sht_t *p, crime_sht_stats *st)
{
  if (!p || !st) return CRIME_GH_EINVAL;
  *st = p->st;
  return 0;
}
