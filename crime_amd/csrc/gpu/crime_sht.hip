// Runtime behind include/crime_sht.h: plan set-up (ring geometry, recursion coefficients, Legendre
// table, GEMM tile lists) and the map2alm / C_l entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
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
  int tile0, ntile;  // slice of the tile list
};

}  // namespace

struct crime_sht {
  int64_t nside = 0, npix = 0;
  int lmax = 0, max_maps = 0, device = 0;
  int R = 0, RP = 0, ldg = 0, ldo = 0;
  int64_t nalm = 0;
  size_t lds = 0;
  hipStream_t stream = nullptr;
  ShtRing *d_rings = nullptr;
  double *d_z = nullptr, *d_st = nullptr;
  double2 *d_coef = nullptr;
  int64_t *d_off = nullptr;
  int4 *d_tiles = nullptr;
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

#define SHT_HIP(p, call)                                                                              \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return sht_error(p, e_ == hipErrorOutOfMemory ? CRIME_GH_ENOMEM : CRIME_GH_EHIP, "%s: %s", #call, \
                       hipGetErrorString(e_));                                                        \
  } while (0)

extern "C" const char *crime_sht_last_error(const crime_sht_t *p) { return p ? p->err.c_str() : g_sht_err.c_str(); }

extern "C" void crime_sht_destroy(crime_sht_t *p)
{
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  void *bufs[] = {p->d_rings, p->d_z, p->d_st, p->d_coef, p->d_off, p->d_tiles, p->d_G,
                  p->d_alm, p->d_lam, p->d_cls, p->d_unpack, p->d_hmaps, p->d_tw,
                  p->d_woff, p->d_boff, p->d_wtab, p->d_btab};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
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
  std::vector<int4> tiles;
  size_t max_chunk_rows = 0;
  for (int m = 0; m <= lmax;) {
    Chunk c;
    c.m0 = m;
    c.tile0 = (int)tiles.size();
    size_t rows = 0;
    while (m <= lmax && (rows == 0 || (rows + (size_t)(lmax - m + 1)) * row_bytes <= LAMBDA_BUDGET)) {
      rows += (size_t)(lmax - m + 1);
      const int L = lmax - m + 1;
      for (int par = 0; par < 2; par++) {
        const int nr = (L - par + 1) / 2;
        for (int rt = 0; rt * ALM_ROWS < nr; rt++) tiles.push_back(make_int4(m, par, rt, nr));
      }
      m++;
    }
    c.m1 = m;
    c.ntile = (int)tiles.size() - c.tile0;
    if (rows > max_chunk_rows) max_chunk_rows = rows;
    p->chunks.push_back(c);
  }
  SHT_HIP(p, hipMalloc((void **)&p->d_tiles, tiles.size() * sizeof(int4)));
  SHT_HIP(p, hipMemcpy(p->d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  SHT_HIP(p, hipMalloc((void **)&p->d_lam, max_chunk_rows * row_bytes));

  const size_t gsz = (size_t)(lmax + 1) * 2 * p->RP * p->ldg * sizeof(double);
  SHT_HIP(p, hipMalloc((void **)&p->d_G, gsz));
  SHT_HIP(p, hipMemset(p->d_G, 0, gsz));   // rows of the padding rings stay zero
  SHT_HIP(p, hipMalloc((void **)&p->d_alm, (size_t)p->nalm * p->ldo * sizeof(double)));
  p->ev.resize(2 * p->chunks.size() + 4);
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
static int analyse(crime_sht *p, int nmaps, const void *d_maps, int fp64)
{
  if (nmaps < 1 || nmaps > p->max_maps)
    return sht_error(p, CRIME_GH_EINVAL, "nmaps %d outside [1, %d]", nmaps, p->max_maps);
  if (!d_maps) return sht_error(p, CRIME_GH_EINVAL, "maps are NULL");
  hipEvent_t *ev = p->ev.data();
  SHT_HIP(p, hipEventRecord(ev[0], p->stream));
  for (ShtLaunch g : p->groups) {
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
    SHT_HIP(p, launch_rings(g, d_maps, fp64));
  }
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

extern "C" int crime_sht_get_stats(const crime_sht_t *p, crime_sht_stats *st)
{
  if (!p || !st) return CRIME_GH_EINVAL;
  *st = p->st;
  return 0;
}
