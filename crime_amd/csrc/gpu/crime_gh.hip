// C-ABI layer of libcrime_gh.so: handle life cycle, device arena, stage launches and the parity
// accessors declared in include/crime_gh.h.  One handle = one HIP stream + its resident buffers.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "crime_gh.h"
#include "gh_kernels.hpp"

using namespace crime;

struct crime_gh {
  crime_gh_desc d;      // scalars of the descriptor (host array pointers are not kept)
  int prec = 0, n = 0, n_nu = 0;
  int64_t nside = 0, npix = 0;
  hipStream_t stream = nullptr;
  void *F = nullptr;
  void *G = nullptr;           // (vy, vz) of CRIME_GH_VEL_SPECTRAL handles
  double *d_sin = nullptr;     // [N] sin(2 pi m / N) / dx of the spectral velocities
  unsigned long long *d_fx = nullptr;  // deterministic scatter: integer maps [n_nu * npix]
  double *d_fxs = nullptr;     // [2 + FX_PARTS]: s, 1/s, the mass partials
  size_t ncell = 0;
  void *maps = nullptr;
  size_t nmaps = 0;
  double *d_tab = nullptr;     // z_r2z | r_r2z | d_r2z | v_r2z | HI fraction | reference bias (at the nodes)
  double *d_pk = nullptr;      // logk | pk
  double *d_edges = nullptr;   // [n_nu+1]
  double *d_prefac = nullptr;  // [n_nu]
  ChanRec *d_crec = nullptr;   // channel records of the frequency buckets
  float2 *d_chwin = nullptr;   // [n_nu] decided interval of 1 + z of each channel (FP32 screen)
  void *d_amp = nullptr;       // amplitude per |k|^2
  uint32_t *d_order = nullptr; // map-stage tiles by distance from the observer
  uint32_t *d_binstart = nullptr; // [nbins + 1] first tile of each distance bin in d_order
  uint32_t *d_range = nullptr; // [2] the map kernel's range of d_order (k_tile_range)
  uint8_t *d_cull = nullptr;   // [ntiles] out-of-band flags of d_order's entries (k_tile_cull)
  double *d_sub = nullptr;     // [3*n_sub]
  double *d_subp = nullptr;    // [3*n_sub] the offsets in the map stage's visiting order
  double *d_modes = nullptr;   // injected unit modes
  void *d_tw = nullptr;        // twiddles exp(+2 pi i m/N)
  double *d_partials = nullptr;
  double *d_sig = nullptr;     // [sigma2, mean, max |phi| of the grid (infinite when unknown)]
  double *h_sig = nullptr;     // pinned mirror
  uint32_t *h_range = nullptr; // pinned mirror of d_range (the map stage's grid)
  hipEvent_t ev_range = nullptr;
  int nparts = 0;
  FillParams fill{};
  MapParams map{};
  double norm = 0;
  bool fields_done = false;
  double eps_nu = 0;           // frequency margin of the FP32 screen (MHz)
  Schedule sched{};            // kernel variants (crime_gh_autotune / crime_gh_set_schedule)
  hipEvent_t ev[8] = {};
  // point sources (allocated by the first crime_gh_mk_psources_maps)
  double *d_ps = nullptr;      // nz[PS_NZ] | maxpdf[PS_NZ] | c1[n_nu] | c2[n_nu] | A[npix] | B[npix]
  unsigned long long *d_ps_count = nullptr;
  void *ps_maps = nullptr;     // [n_nu * npix], the grid's precision, mK
  std::vector<double> ps_nu;   // channel centres of mk_psources_maps (pixelize.c:62-69)
  std::string err;
};

static thread_local std::string g_create_err;

// ---------------------------------------------------------------------------------------------
// Scratch of crime_gh_power_spectrum_noalloc: a file-scope device array, one copy per GPU, held until
// the process ends.  Layout (doubles): shell sums acc[3 * PK_MAX_BINS] | P[PK_MAX_BINS] |
// k_mean[PK_MAX_BINS] | n_modes as int64[PK_MAX_BINS]: 96 KB.  In host code `pk_scratch` names a
// host-side stand-in, not device memory: its device address is taken once per device with
// hipGetSymbolAddress (scratch_setup, from crime_gh_create) and only that address is used.
// ---------------------------------------------------------------------------------------------
constexpr int PK_SCRATCH_DOUBLES = 6 * PK_MAX_BINS;
static __device__ __align__(256) double pk_scratch[PK_SCRATCH_DOUBLES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[SCRATCH_MAX_DEVICES] = {};

static int set_error(crime_gh *h, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  else g_create_err = buf;
  return code;
}

#define HIP_OR_FAIL(h, call)                                                                          \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      (void)hipGetLastError(); /* reported here: a later launch check must not see it again */        \
      return set_error(h, e_ == hipErrorOutOfMemory ? CRIME_GH_ENOMEM : CRIME_GH_EHIP, "%s: %s", #call,     \
                       hipGetErrorString(e_));                                                        \
    }                                                                                                 \
  } while (0)

static bool grid_supported(int n)
{
  switch (n) {
#define CRIME_CASE(NN) case NN:
    CRIME_FOR_EACH_N(CRIME_CASE)
#undef CRIME_CASE
    return true;
  default: return false;
  }
}

static double r_of_z_host(const crime_gh_desc *d, double z)
{
  // r_of_z, src/cosmo.c:40-50
  if (z <= 0) return 0;
  if (z >= d->z_z2r[CRIME_GH_NZ - 1]) return d->r_z2r[CRIME_GH_NZ - 1];
  int iz = (int)(z / CRIME_GH_DZ);
  return d->r_z2r[iz] + (d->r_z2r[iz + 1] - d->r_z2r[iz]) * (z - d->z_z2r[iz]) / CRIME_GH_DZ;
}

static double pk_host(const crime_gh_desc *d, double lgk)
{
  int ik = (int)((lgk - d->logkmin) * d->idlogk);
  if (ik < 0) return d->pk[0] * pow(10, d->n_s * (lgk - d->logkmin));
  if (ik < d->numk - 1) return d->pk[ik] + (lgk - d->logk[ik]) * (d->pk[ik + 1] - d->pk[ik]) * d->idlogk;
  return d->pk[d->numk - 1] * pow(10, -3 * (lgk - d->logkmax));
}

// rms(delta)/rms(phi) of the grid's modes (rms(delta)/rms(v_j) for spectral velocities), estimated
// with a spherical k-space integral from the fundamental to the radius of the sphere with the grid
// cube's volume.  Only its order of magnitude matters: phi (or v_j) is packed as phi_scale*phi next to
// delta in one complex FP32 element, so that both fields carry rounding errors relative to their own
// size.
static double phi_balance(const crime_gh_desc *d, bool velocity)
{
  double dk = 2 * M_PI / d->l_box;
  double kmax = 0.5 * d->n_grid * dk * cbrt(6 / M_PI);
  const int nk = 4000;
  double l0 = log(dk), l1 = log(kmax), i0 = 0, i2 = 0;
  for (int i = 0; i < nk; i++) {
    double k = exp(l0 + (l1 - l0) * (i + 0.5) / nk);
    double p = pk_host(d, log10(k)) * k * k * k;  // k^2 P(k) dk with dk = k dlnk
    if (d->do_smoothing) p *= exp(-d->r2_smooth * k * k);
    // phi = f0 H0 / k^2 delta; one velocity component ~ f0 H0 / k delta, k_j^2 ~ k^2 / 3 on average
    double psi = velocity ? d->f0h0 / (k * sqrt(3.0)) : d->f0h0 / (k * k);
    i0 += p;
    i2 += p * psi * psi;
  }
  if (!(i0 > 0) || !(i2 > 0)) return 1.0;
  return sqrt(i0 / i2);
}

// float nearest to x in the direction of +inf (up) or -inf (down): a bound that stays on its side
static float f32_up(double x)
{
  float f = (float)x;
  return (double)f < x ? nextafterf(f, INFINITY) : f;
}
static float f32_down(double x)
{
  float f = (float)x;
  return (double)f > x ? nextafterf(f, -INFINITY) : f;
}

// Margins and tables of the FP32 screen (MapParams::scr), from bounds on the FP32 operations (u = 2^-23;
// v_sqrt_f32, v_rcp_f32 and v_rsq_f32 are within 1 ulp, +, *, fma and conversions within 0.5 ulp):
//  - frequency: per cell z(r) is the parabola P through the table lerp L at r_c - h, r_c, r_c + h
//    (h bounds the sample offsets).  On the cloud L is a line plus sum_j kappa_j (r - r_j)_+ over the
//    nodes r_j inside it (kappa_j: the slope jumps, second differences / dr), and P reproduces lines, so
//    L - P = sum_j kappa_j e_j with e_j the parabola's error on one hinge, |e_j| <= h / (3 sqrt 3).  With
//    kappa_j = kb + d_j: the kb part is kb (I[q] - q - P[I[q] - q]) for the lerp I[q] of a parabola q with
//    unit slope jumps, at most 2.25 |kb| dr / 8 (|I[q] - q| <= dr / 8, Lebesgue constant 1.25), and the d_j
//    part at most h / (3 sqrt 3) sum |d_j|.  E_q is the largest such bound over every run of
//    floor(2h / dr) + 1 consecutive nodes of the channels' radii, with kb the run's mid-range.  The FP32
//    evaluation: |p| = sqrt((float)|p|^2_FP64) within 1.25 u of |p| and the rounded r_c within 0.5 u,
//    which moves z by at most 2 A u with A = r dz/dr; the roundings of z_c + dz, a, b, the two fmas and
//    1 + z add at most 2 (1 + z) u (2.5 u more are allowed for: an earlier form took the frequency in
//    FP32).  So NU_21 / (1 + z) in FP32 arithmetic is within |dnu| <= nu (u (2 A / (1 + z) + 4.5) +
//    E_q / (1 + z)) of the exact frequency, taken at the worst table node of the channels' redshift range,
//    doubled; the decided windows bound 1 + z so that NU_21 / (1 + z) keeps eps_nu inside the edges;
//  - pixel: the angle (input rounding 0.32 u, t = min/max 0.48 u, polynomial and its evaluation 1.2 u,
//    scalings 0.75 u) within 2.8 u quarter turns; cos(theta) within 3.3 u; sin(theta) and the cap's
//    square root within 5 u relative; every floor argument of vec2pix is then within 8 N u of the exact
//    one, whose FP64 evaluation is itself within 1e-12 of it: margin 16 N u.  |cos theta| and the
//    quadrant angle w keep 8 u from 2/3 and from 0 or 1.
static int screen_setup(crime_gh *h, const crime_gh_desc *d, const std::vector<double> &edges)
{
  const double u = 1.0 / 8388608.0;
  ScreenParams &c = h->map.scr;
  c = ScreenParams{};
  c.on = d->nu_mode == CRIME_GH_NU_TABLE;
  // bound on the sample offsets, as MapParams::h_sub
  double hmax = 0;
  for (int i = 0; i < d->n_sub; i++) {
    const double *o = d->sub_off + 3 * i;
    hmax = fmax(hmax, sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
  }
  const double hs = hmax * (1 + 1e-9) + 1e-6;
  const double dr = 1 / d->idr;
  // radii of the channels (with slack for the RSD shifts) and the table's derivative bounds over them
  const double nu_lo = edges[0], nu_hi = edges[d->n_nu];
  const double zlo = CRIME_GH_NU_21 / nu_hi - 1 - 0.05, zhi = CRIME_GH_NU_21 / nu_lo - 1 + 0.05;
  const double *z = d->z_r2z;
  int ilo = CRIME_GH_NZ, ihi = 0;
  for (int i = 0; i < CRIME_GH_NZ; i++)
    if (z[i] >= zlo && z[i] <= zhi) {
      ilo = std::min(ilo, i);
      ihi = std::max(ihi, i);
    }
  const int nwin = (int)floor(2 * hs / dr) + 1;  // nodes inside one cloud, at most
  ilo = std::max(1, ilo - nwin - 2);
  ihi = std::min(CRIME_GH_NZ - 2, ihi + nwin + 2);
  double e_q = 0;
  for (int j0 = ilo; j0 + nwin - 1 <= ihi; j0++) {
    double kmin = INFINITY, kmax = -INFINITY;
    for (int j = j0; j < j0 + nwin; j++) {
      const double kap = (z[j + 1] - 2 * z[j] + z[j - 1]) * d->idr;
      kmin = fmin(kmin, kap);
      kmax = fmax(kmax, kap);
    }
    const double kb = 0.5 * (kmax + kmin);
    e_q = fmax(e_q, 2.25 * fabs(kb) * dr / 8 + hs / (3 * sqrt(3.0)) * nwin * 0.5 * (kmax - kmin));
  }
  double worst = 0;
  for (int i = 0; i + 1 < CRIME_GH_NZ; i++) {
    const double z0 = z[i], z1 = z[i + 1];
    if (z1 < zlo || z0 > zhi) continue;
    const double A = (z1 - z0) * d->idr * d->r_r2z[i + 1];
    const double za = z0 > 0 ? z0 : 0, zb = z1 > 0 ? z1 : 0;
    const double nu = CRIME_GH_NU_21 / (1 + za);
    worst = fmax(worst, nu * (u * (2 * A / (1 + zb) + 4.5) + e_q / (1 + za)));
  }
  const double eps_nu = 2 * fmax(worst, nu_hi * 4.5 * u);
  // decided intervals of 1 + z: NU_21 / (1 + z) in [edge + eps, next edge - eps], rounded inwards (the
  // quotients widened by 1e-15 first, more than their FP64 rounding)
  auto g_of = [](double nu) { return CRIME_GH_NU_21 / nu; };
  std::vector<float2> w(d->n_nu);
  for (int k = 0; k < d->n_nu; k++)
    w[k] = make_float2(f32_up(g_of(edges[k + 1] - eps_nu) * (1 + 1e-15)), f32_down(g_of(edges[k] + eps_nu) * (1 - 1e-15)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_chwin, w.size() * sizeof(float2)));
  HIP_OR_FAIL(h, hipMemcpy(h->d_chwin, w.data(), w.size() * sizeof(float2), hipMemcpyHostToDevice));
  c.chwin = h->d_chwin;
  // the cloud [r_c - h, r_c + h] stays inside (r_min, r_last): no clamp of the table lerp inside it, and
  // every sample's |p| is above r_min (radii below it are left to FP64)
  const double r_min = 1e-3 * d->l_box / d->n_grid;
  c.rc_lo = hs * (1 + 1e-9) + 1e-6 + 2 * r_min;
  c.rc_hi = d->r_r2z[CRIME_GH_NZ - 1] - hs * (1 + 1e-9) - 1e-6;
  c.inv_2h = 1 / (2 * hs);
  c.inv_2h2 = 1 / (2 * hs * hs);
  // no channel: NU_21 / (1 + z) above nu_hi + eps or below nu_lo - eps
  c.out_lo = f32_down(g_of(nu_hi + eps_nu) * (1 - 1e-15));
  c.out_hi = nu_lo - eps_nu > 0 ? f32_up(g_of(nu_lo - eps_nu) * (1 + 1e-15)) : INFINITY;
  c.eps_pix = (float)(16 * (double)d->n_side * u);
  c.eps_ax = (float)(8 * u);
  c.eps_z = (float)(8 * u);
  // |fl(d - 1/2)| <= half implies |d - 1/2| <= 1/2 - eps: d - 1/2 rounds by at most 2^-26 and d itself
  // (a difference of floats in [0, 1]) by at most 2^-25
  c.half_pix = f32_down(0.5 - c.eps_pix - 1.0 / 16777216.0);
  c.half_ax = f32_down(0.5 - c.eps_ax - 1.0 / 16777216.0);
  h->eps_nu = eps_nu;
  return 0;
}

static int validate(const crime_gh_desc *d)
{
  if (!d) return set_error(nullptr, CRIME_GH_EINVAL, "descriptor is NULL");
  if (d->abi_version != CRIME_GH_ABI_VERSION)
    return set_error(nullptr, CRIME_GH_EABI, "abi_version %u, library expects %u", d->abi_version,
                     CRIME_GH_ABI_VERSION);
  if (d->precision != CRIME_GH_FP32 && d->precision != CRIME_GH_FP64)
    return set_error(nullptr, CRIME_GH_EINVAL, "precision %d", d->precision);
  if (!grid_supported(d->n_grid))
    return set_error(nullptr, CRIME_GH_EINVAL, "n_grid=%d has no FFT plan (2^a in [8,2048], 3*2^a in [12,1536])",
                     d->n_grid);
  if (d->n_side < 1 || d->n_side > (1 << 13)) return set_error(nullptr, CRIME_GH_EINVAL, "n_side=%ld", (long)d->n_side);
  if (d->n_nu < 1) return set_error(nullptr, CRIME_GH_EINVAL, "n_nu=%d", d->n_nu);
  if (d->nz != CRIME_GH_NZ) return set_error(nullptr, CRIME_GH_EINVAL, "nz=%d, tables must have %d entries", d->nz, CRIME_GH_NZ);
  if (!d->z_z2r || !d->r_z2r || !d->z_r2z || !d->r_r2z || !d->d_r2z || !d->v_r2z)
    return set_error(nullptr, CRIME_GH_EINVAL, "a radial table pointer is NULL");
  if (!(d->idr > 0)) return set_error(nullptr, CRIME_GH_EINVAL, "idr must be positive");
  if (!d->logk || !d->pk || d->numk < 2) return set_error(nullptr, CRIME_GH_EINVAL, "P(k) table missing");
  if (!(d->l_box > 0)) return set_error(nullptr, CRIME_GH_EINVAL, "l_box must be positive");
  if (d->n_sub < 1 || d->n_sub > 4096 || !d->sub_off)
    return set_error(nullptr, CRIME_GH_EINVAL, "n_sub=%d / sub_off", d->n_sub);
  if (d->nu_mode == CRIME_GH_NU_TABLE) {
    if (!d->nu0 || !d->nuf) return set_error(nullptr, CRIME_GH_EINVAL, "channel edges missing");
    for (int i = 0; i < d->n_nu; i++) {
      if (!(d->nuf[i] > d->nu0[i])) return set_error(nullptr, CRIME_GH_EINVAL, "Frequency bins don't make sense");
      if (i > 0 && d->nu0[i] != d->nuf[i - 1])
        return set_error(nullptr, CRIME_GH_EINVAL, "channel %d does not start where channel %d ends", i, i - 1);
    }
  }
  else if (d->nu_mode == CRIME_GH_NU_UNIFORM) {
    if (!(d->nu_max > d->nu_min)) return set_error(nullptr, CRIME_GH_EINVAL, "nu_max must exceed nu_min");
  }
  else {
    return set_error(nullptr, CRIME_GH_EINVAL, "nu_mode %d", d->nu_mode);
  }
  if (d->mode_source != CRIME_GH_MODES_PHILOX && d->mode_source != CRIME_GH_MODES_INJECTED)
    return set_error(nullptr, CRIME_GH_EINVAL, "mode_source %d", d->mode_source);
  if (d->mode_source == CRIME_GH_MODES_INJECTED && !d->unit_modes)
    return set_error(nullptr, CRIME_GH_EINVAL, "unit_modes missing for injected modes");
  if (d->bias_model < 0 || d->bias_model > 2) return set_error(nullptr, CRIME_GH_EINVAL, "bias_model %d", d->bias_model);
  if (d->velocity_mode != CRIME_GH_VEL_STENCIL && d->velocity_mode != CRIME_GH_VEL_SPECTRAL)
    return set_error(nullptr, CRIME_GH_EINVAL, "velocity_mode %d", d->velocity_mode);
  if (d->deterministic != 0 && d->deterministic != 1)
    return set_error(nullptr, CRIME_GH_EINVAL, "deterministic %d", d->deterministic);
  return 0;
}

template <typename T>
static void twiddles(int n, std::vector<cx<T>> &tw)
{
  tw.resize(n);
  for (int m = 0; m < n; m++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    tw[m].x = (T)cosl(a);
    tw[m].y = (T)sinl(a);
  }
}

static int free_all(crime_gh *h)
{
  void *ptrs[] = {h->F, h->G, h->d_sin, h->d_fx, h->d_fxs, h->d_chwin, h->maps, h->d_tab, h->d_pk, h->d_edges, h->d_prefac, h->d_crec, h->d_amp, h->d_order, h->d_binstart, h->d_range, h->d_cull, h->d_sub, h->d_subp,
                  h->d_modes, h->d_tw, h->d_partials, h->d_sig, h->d_ps, h->d_ps_count, h->ps_maps};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (h->h_sig) (void)hipHostFree(h->h_sig);
  if (h->h_range) (void)hipHostFree(h->h_range);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->ev_range) (void)hipEventDestroy(h->ev_range);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  return 0;
}

extern "C" void crime_gh_destroy(crime_gh_t *h)
{
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
}

extern "C" size_t crime_gh_desc_size(void) { return sizeof(crime_gh_desc); }

extern "C" const char *crime_gh_last_error(const crime_gh_t *h)
{
  return h ? h->err.c_str() : g_create_err.c_str();
}

static int create_impl(const crime_gh_desc *d, crime_gh *h)
{
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(h, CRIME_GH_ENODEV, "no HIP device visible");
  if (d->device < 0 || d->device >= ndev) return set_error(h, CRIME_GH_ENODEV, "device %d of %d", d->device, ndev);
  HIP_OR_FAIL(h, hipSetDevice(d->device));
  hipDeviceProp_t prop;
  HIP_OR_FAIL(h, hipGetDeviceProperties(&prop, d->device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(h, CRIME_GH_ENODEV, "device %d is %s; this library holds gfx950 code only", d->device,
                     prop.gcnArchName);

  if (d->device >= SCRATCH_MAX_DEVICES)
    return set_error(h, CRIME_GH_ENODEV, "device %d: at most %d devices are supported", d->device, SCRATCH_MAX_DEVICES);
  {
    // the device is current here: record this device's address of pk_scratch once
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[d->device]) {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(pk_scratch));
      if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return set_error(h, CRIME_GH_EHIP, "hipGetSymbolAddress(pk_scratch): %s", hipGetErrorString(e));
      }
      g_scratch[d->device] = (double *)p;
    }
  }

  h->d = *d;
  h->prec = d->precision;
  h->n = d->n_grid;
  h->n_nu = d->n_nu;
  h->nside = d->n_side;
  h->npix = 12 * d->n_side * d->n_side;
  h->ncell = (size_t)h->n * h->n * h->n;
  h->nmaps = (size_t)h->npix * h->n_nu;
  const size_t esz = h->prec ? sizeof(double) : sizeof(float);

  HIP_OR_FAIL(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (auto &e : h->ev) HIP_OR_FAIL(h, hipEventCreate(&e));
  HIP_OR_FAIL(h, hipEventCreateWithFlags(&h->ev_range, hipEventDisableTiming));
  HIP_OR_FAIL(h, hipHostMalloc((void **)&h->h_range, 2 * sizeof(uint32_t), hipHostMallocDefault));

  HIP_OR_FAIL(h, hipMalloc(&h->F, h->ncell * 2 * esz));
  if (d->velocity_mode == CRIME_GH_VEL_SPECTRAL) {
    HIP_OR_FAIL(h, hipMalloc(&h->G, h->ncell * 2 * esz));
    // sin(k_j dx)/dx at k_j dx = 2 pi m / N: exactly 0 at m = 0 and N/2, odd under m -> N - m
    std::vector<double> sn(h->n, 0.0);
    const double dx = d->l_box / h->n;
    for (int m = 1; 2 * m < h->n; m++) {
      const double v = (double)(sinl(2.0L * 3.14159265358979323846264338327950288L * m / h->n) / (long double)dx);
      sn[m] = v;
      sn[h->n - m] = -v;
    }
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_sin, h->n * sizeof(double)));
    HIP_OR_FAIL(h, hipMemcpy(h->d_sin, sn.data(), h->n * sizeof(double), hipMemcpyHostToDevice));
  }
  HIP_OR_FAIL(h, hipMalloc(&h->maps, h->nmaps * esz));
  if (d->deterministic) {
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_fx, h->nmaps * sizeof(unsigned long long)));
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_fxs, (2 + FX_PARTS) * sizeof(double)));
  }
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_tab, 6 * CRIME_GH_NZ * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_pk, 2 * (size_t)d->numk * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_edges, (d->n_nu + 1) * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_prefac, d->n_nu * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_sub, 3 * (size_t)d->n_sub * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_subp, 3 * (size_t)d->n_sub * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc(&h->d_tw, h->n * 2 * esz));
  h->nparts = fields_partials_count(h->n, h->prec);
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_partials, 3 * (size_t)h->nparts * sizeof(double)));
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_sig, 3 * sizeof(double)));
  HIP_OR_FAIL(h, hipHostMalloc((void **)&h->h_sig, 2 * sizeof(double), hipHostMallocDefault));
  size_t nhalf = (size_t)h->n * h->n * (h->n / 2 + 1);
  if (d->mode_source == CRIME_GH_MODES_INJECTED) {
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_modes, 2 * nhalf * sizeof(double)));
    HIP_OR_FAIL(h, hipMemcpy(h->d_modes, d->unit_modes, 2 * nhalf * sizeof(double), hipMemcpyHostToDevice));
  }

  // tables
  // the HI fraction (src/user_defined.c:27-30) and the reference bias (src/user_defined.c:32-35) at the nodes
  std::vector<double> fhi(CRIME_GH_NZ), bref(CRIME_GH_NZ);
  for (int i = 0; i < CRIME_GH_NZ; i++) {
    fhi[i] = 0.008 * pow(1 + d->z_r2z[i], 0.6);
    bref[i] = 0.904 + 0.135 * pow(1 + d->z_r2z[i], 1.696);
  }
  const double *tabs[6] = {d->z_r2z, d->r_r2z, d->d_r2z, d->v_r2z, fhi.data(), bref.data()};
  for (int i = 0; i < 6; i++)
    HIP_OR_FAIL(h, hipMemcpy(h->d_tab + (size_t)i * CRIME_GH_NZ, tabs[i], CRIME_GH_NZ * sizeof(double),
                             hipMemcpyHostToDevice));
  HIP_OR_FAIL(h, hipMemcpy(h->d_pk, d->logk, d->numk * sizeof(double), hipMemcpyHostToDevice));
  HIP_OR_FAIL(h, hipMemcpy(h->d_pk + d->numk, d->pk, d->numk * sizeof(double), hipMemcpyHostToDevice));
  HIP_OR_FAIL(h, hipMemcpy(h->d_sub, d->sub_off, 3 * (size_t)d->n_sub * sizeof(double), hipMemcpyHostToDevice));
  {
    // the map stage visits the offsets along a nearest-neighbour chain from the first: consecutive samples of
    // a cell then often share a (channel, pixel) bin and are summed in registers before the LDS table
    const int ns = d->n_sub;
    std::vector<int> perm(ns);
    std::vector<char> used(ns, 0);
    perm[0] = 0;
    used[0] = 1;
    for (int j = 1; j < ns; j++) {
      const double *a = d->sub_off + 3 * perm[j - 1];
      int best = -1;
      double bd = INFINITY;
      for (int i = 0; i < ns; i++) {
        if (used[i]) continue;
        const double *b = d->sub_off + 3 * i;
        const double dd = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
        if (dd < bd) {
          bd = dd;
          best = i;
        }
      }
      perm[j] = best;
      used[best] = 1;
    }
    std::vector<double> offp(3 * (size_t)ns);
    for (int j = 0; j < ns; j++)
      for (int k = 0; k < 3; k++) offp[3 * (size_t)j + k] = d->sub_off[3 * (size_t)perm[j] + k];
    HIP_OR_FAIL(h, hipMemcpy(h->d_subp, offp.data(), offp.size() * sizeof(double), hipMemcpyHostToDevice));
  }

  // channel edges and mK factors (pixelize.c:155, 246-256)
  std::vector<double> edges(d->n_nu + 1), prefac(d->n_nu);
  double dnu_u = (d->nu_max - d->nu_min) / d->n_nu;
  for (int i = 0; i < d->n_nu; i++) edges[i] = d->nu_mode == CRIME_GH_NU_TABLE ? d->nu0[i] : d->nu_min + dnu_u * i;
  edges[d->n_nu] = d->nu_mode == CRIME_GH_NU_TABLE ? d->nuf[d->n_nu - 1] : d->nu_max;
  double m2t = 90.057156 * d->omega_b * d->hhub * (double)h->npix / (4 * M_PI);
  for (int i = 0; i < d->n_nu; i++) {
    double dnu, nu;
    if (d->nu_mode == CRIME_GH_NU_TABLE) {
      dnu = d->nuf[i] - d->nu0[i];
      nu = (d->nuf[i] + d->nu0[i]) * 0.5;
    }
    else {
      dnu = (d->nu_max - d->nu_min) / d->n_nu;
      nu = d->nu_min + (i + 0.5) * dnu;
    }
    double r = r_of_z_host(d, CRIME_GH_NU_21 / nu - 1);
    prefac[i] = m2t / (r * r * dnu);
  }
  h->ps_nu.resize(d->n_nu);
  for (int i = 0; i < d->n_nu; i++)
    h->ps_nu[i] = d->nu_mode == CRIME_GH_NU_TABLE ? (d->nu0[i] + d->nuf[i]) * 0.5
                                                  : d->nu_min + (d->nu_max - d->nu_min) * (i + 0.5) / d->n_nu;
  HIP_OR_FAIL(h, hipMemcpy(h->d_edges, edges.data(), edges.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OR_FAIL(h, hipMemcpy(h->d_prefac, prefac.data(), prefac.size() * sizeof(double), hipMemcpyHostToDevice));
  // channel guess table: guess[j] = channel holding g_lo + j/g_inv (table mode)
  int n_guess = 8 * d->n_nu < 64 ? 64 : 8 * d->n_nu;
  double g_lo = edges[0], g_inv = n_guess / (edges[d->n_nu] - edges[0]);
  std::vector<int> guess(n_guess);
  for (int j = 0; j < n_guess; j++) {
    double nu = g_lo + j / g_inv;
    int lo = 0, hi = d->n_nu;
    while (hi - lo > 1) {
      int mid = (lo + hi) / 2;
      if (nu >= edges[mid]) lo = mid;
      else hi = mid;
    }
    guess[j] = lo;
  }
  std::vector<ChanRec> crec(n_guess);
  for (int j = 0; j < n_guess; j++) {
    const int c = guess[j], last = c == d->n_nu - 1;
    crec[j] = ChanRec{};
    crec[j].c = c;
    crec[j].e0 = edges[c];
    crec[j].e1 = edges[c + 1];
    crec[j].e2 = last ? edges[c + 1] : edges[c + 2];  // [e1, e2) empty: no channel above the last
    crec[j].pf0 = prefac[c];
    crec[j].pf1 = last ? 0.0 : prefac[c + 1];
  }
  HIP_OR_FAIL(h, hipMalloc((void **)&h->d_crec, n_guess * sizeof(ChanRec)));
  HIP_OR_FAIL(h, hipMemcpy(h->d_crec, crec.data(), n_guess * sizeof(ChanRec), hipMemcpyHostToDevice));
  {
    const int rc = screen_setup(h, d, edges);
    if (rc) return rc;
  }

  if (h->prec) {
    std::vector<cx<double>> tw;
    twiddles(h->n, tw);
    HIP_OR_FAIL(h, hipMemcpy(h->d_tw, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice));
  }
  else {
    std::vector<cx<float>> tw;
    twiddles(h->n, tw);
    HIP_OR_FAIL(h, hipMemcpy(h->d_tw, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice));
  }

  // fill parameters (fourier.c:245-255)
  FillParams &f = h->fill;
  f.n = h->n;
  f.dk = 2 * M_PI / d->l_box;
  f.idk3 = 1. / (f.dk * f.dk * f.dk);
  f.f0h0 = d->f0h0;
  f.r2_smooth = d->r2_smooth;
  f.do_smoothing = d->do_smoothing;
  f.logk = h->d_pk;
  f.pk = h->d_pk + d->numk;
  f.numk = d->numk;
  f.logkmin = d->logkmin;
  f.logkmax = d->logkmax;
  f.idlogk = d->idlogk;
  f.n_s = d->n_s;
  f.seed = d->seed;
  f.mode_source = d->mode_source;
  f.unit_modes = h->d_modes;
  h->norm = pow(sqrt(2 * M_PI) / d->l_box, 3);  // fourier.c:403
  f.vmode = d->velocity_mode;
  f.comp = 0;
  f.sin_dx = h->d_sin;
  f.phi_scale = phi_balance(d, d->velocity_mode == CRIME_GH_VEL_SPECTRAL);
  f.psi_c = f.phi_scale * d->f0h0 / (f.dk * f.dk);
  HIP_OR_FAIL(h, hipMalloc(&h->d_amp, (size_t)amp_count(h->n) * 2 * esz));
  f.amp = h->d_amp;
  HIP_OR_FAIL(h, h->prec ? launch_amp_table<double>(f, h->d_amp, h->stream)
                         : launch_amp_table<float>(f, h->d_amp, h->stream));
  HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));

  // map parameters
  MapParams &m = h->map;
  m.n = h->n;
  m.dx = d->l_box / h->n;
  for (int i = 0; i < 3; i++) m.obs[i] = d->pos_obs[i];
  m.inv_dx = 1.0 / m.dx;
  m.tab.z = h->d_tab;
  m.tab.r = h->d_tab + CRIME_GH_NZ;
  m.tab.d = h->d_tab + 2 * CRIME_GH_NZ;
  m.tab.v = h->d_tab + 3 * CRIME_GH_NZ;
  m.tab.f = h->d_tab + 4 * CRIME_GH_NZ;
  m.tab.b = h->d_tab + 5 * CRIME_GH_NZ;
  m.tab.idr = d->idr;
  m.tab.r_last = d->r_r2z[CRIME_GH_NZ - 1];
  m.tab.z_last = d->z_r2z[CRIME_GH_NZ - 1];
  m.tab.d_last = d->d_r2z[CRIME_GH_NZ - 1];
  m.tab.v_last = d->v_r2z[CRIME_GH_NZ - 1];
  m.tab.f_last = 0.008 * pow(1 + m.tab.z_last, 0.6);
  m.tab.b_last = 0.904 + 0.135 * pow(1 + m.tab.z_last, 1.696);
  m.bias_model = d->bias_model;
  m.b0 = d->bias_b0;
  m.b1 = d->bias_b1;
  m.mass_prefac = m.dx * m.dx * m.dx;
  m.rsd = d->rsd;
  m.n_sub = d->n_sub;
  m.sub_off = h->d_sub;
  m.sub_offp = h->d_subp;
  double hmax = 0;
  for (int i = 0; i < d->n_sub; i++) {
    const double *o = d->sub_off + 3 * i;
    hmax = fmax(hmax, sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
  }
  m.h_sub = hmax * (1 + 1e-9) + 1e-6;
  m.n_nu = d->n_nu;
  m.nu_mode = d->nu_mode;
  m.edges = h->d_edges;
  m.nu_min = d->nu_min;
  m.inv_dnu = d->n_nu / (d->nu_max - d->nu_min);
  if (d->nu_mode == CRIME_GH_NU_TABLE) {
    m.nu_lo = edges[0];
    m.nu_hi = edges[d->n_nu];
  }
  else {
    // (int) truncation bins frequencies down to nu_min - dnu into channel 0 (pixelize.c:216)
    m.nu_lo = d->nu_min - dnu_u;
    m.nu_hi = d->nu_max;
  }
  m.prefac = h->d_prefac;
  m.crec = h->d_crec;
  m.n_guess = n_guess;
  m.g_lo = g_lo;
  m.g_inv = g_inv;

  // tile order for the map stage: counting sort of the tiles by the distance of their centre from
  // the observer, in bins of one cell
  m.ntx = (h->n + TILE_X - 1) / TILE_X;
  m.nty = (h->n + TILE_Y - 1) / TILE_Y;
  m.ntz = (h->n + TILE_Z - 1) / TILE_Z;
  {
    const size_t ntiles = (size_t)m.ntx * m.nty * m.ntz;
    const double dx = m.dx;
    auto dist_bin = [&](size_t b) {
      size_t tx = b % m.ntx, ty = (b / m.ntx) % m.nty, tz = b / ((size_t)m.ntx * m.nty);
      double cx_ = dx * (tx * TILE_X + 0.5 * TILE_X) - d->pos_obs[0];
      double cy_ = dx * (ty * TILE_Y + 0.5 * TILE_Y) - d->pos_obs[1];
      double cz_ = dx * (tz * TILE_Z + 0.5 * TILE_Z) - d->pos_obs[2];
      return (size_t)(sqrt(cx_ * cx_ + cy_ * cy_ + cz_ * cz_) / dx);
    };
    // distances inside the box are below sqrt(3) N cells; an observer outside it only clamps the order
    const size_t nbins = (size_t)(2.0 * h->n) + 2;
    std::vector<uint32_t> count(nbins + 1, 0), order(ntiles);
    for (size_t b = 0; b < ntiles; b++) count[std::min(dist_bin(b), nbins - 1) + 1]++;
    for (size_t i = 1; i <= nbins; i++) count[i] += count[i - 1];
    const std::vector<uint32_t> bin_start(count);  // count[b]: the first entry of bin b
    // each entry holds the tile's coordinates, tx | ty << 10 | tz << 20 (ntx, nty, ntz <= 512 for n <= 2048)
    for (size_t b = 0; b < ntiles; b++) {
      const uint32_t tx = (uint32_t)(b % m.ntx), ty = (uint32_t)((b / m.ntx) % m.nty),
                     tz = (uint32_t)(b / ((size_t)m.ntx * m.nty));
      order[count[std::min(dist_bin(b), nbins - 1)]++] = tx | ty << 10 | tz << 20;
    }
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_order, ntiles * sizeof(uint32_t)));
    HIP_OR_FAIL(h, hipMemcpy(h->d_order, order.data(), ntiles * sizeof(uint32_t), hipMemcpyHostToDevice));
    m.tile_order = h->d_order;
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_binstart, (nbins + 1) * sizeof(uint32_t)));
    HIP_OR_FAIL(h, hipMemcpy(h->d_binstart, bin_start.data(), (nbins + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    m.bin_start = h->d_binstart;
    m.nbins = (int)nbins;
    // half-diagonal of a tile's box of cell centres, with a slack for the rounding of the bins' distances
    const double hx = 0.5 * (TILE_X - 1), hy = 0.5 * (TILE_Y - 1), hz = 0.5 * (TILE_Z - 1);
    m.tile_hd = dx * sqrt(hx * hx + hy * hy + hz * hz) + 1e-9 * dx * (double)nbins;
    const uint32_t range0[2] = {0, (uint32_t)ntiles};
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_range, sizeof(range0)));
    HIP_OR_FAIL(h, hipMemcpy(h->d_range, range0, sizeof(range0), hipMemcpyHostToDevice));
    m.tile_range = h->d_range;
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_cull, ntiles));
    HIP_OR_FAIL(h, hipMemset(h->d_cull, 0, ntiles));
    m.tile_cull = h->d_cull;
  }
  m.nside = h->nside;
  m.npix = h->npix;
  m.sig = h->d_sig;
  m.vmode = d->velocity_mode;
  m.G = h->G;
  m.fx = d->deterministic ? 1 : 0;
  m.fxmaps = h->d_fx;
  m.fxs = h->d_fxs;
  m.mass_part = h->d_fxs ? h->d_fxs + 2 : nullptr;
  // tile culling before any grid load: |v_r| <= sqrt(3) max|phi| / dx for the central-difference
  // velocities, so |dz| <= that times the largest |V(r)| of the table (with a rounding slack)
  m.precull = d->velocity_mode == CRIME_GH_VEL_STENCIL;
  double vmax = 0;
  for (int i = 0; i < CRIME_GH_NZ; i++) vmax = fmax(vmax, fabs(d->v_r2z[i]));
  m.dz_per_phi = d->rsd ? sqrt(3.0) * m.inv_dx * vmax * (1 + 1e-6) : 0.0;
  double sig0[3] = {0, 0, INFINITY};
  HIP_OR_FAIL(h, hipMemcpy(h->d_sig, sig0, sizeof(sig0), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int crime_gh_create(const crime_gh_desc *desc, crime_gh_t **out)
{
  if (!out) return set_error(nullptr, CRIME_GH_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  crime_gh *h = new crime_gh();
  rc = create_impl(desc, h);
  if (rc) {
    g_create_err = h->err;
    free_all(h);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

static float ev_ms(hipEvent_t a, hipEvent_t b)
{
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1;
  return ms;
}

static int enqueue_fields(crime_gh *h, bool timed, hipStream_t s = nullptr)
{
  hipError_t e;
  if (!s) s = h->stream;
  if (h->prec)
    e = launch_fields<double>(h->n, (cx<double> *)h->F, (cx<double> *)h->G, h->fill, (const cx<double> *)h->d_tw, h->norm,
                              h->d_partials, h->d_sig, s, timed ? h->ev : nullptr, h->sched);
  else
    e = launch_fields<float>(h->n, (cx<float> *)h->F, (cx<float> *)h->G, h->fill, (const cx<float> *)h->d_tw, h->norm,
                             h->d_partials, h->d_sig, s, timed ? h->ev : nullptr, h->sched);
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "field kernels: %s", hipGetErrorString(e));
  return 0;
}

// s: a stream under capture (the workgroups then cover every tile); none: the handle's stream, whose map
// launch is sized by the host from the tile range
static int enqueue_maps(crime_gh *h, bool timed, hipStream_t s = nullptr)
{
  hipError_t e;
  const RangeReadback rb{h->h_range, h->ev_range};
  const RangeReadback *prb = s ? nullptr : &rb;
  if (!s) s = h->stream;
  if (h->prec)
    e = launch_maps<double>((const cx<double> *)h->F, (double *)h->maps, h->nmaps, h->map, s,
                            timed ? h->ev + 5 : nullptr, h->sched, prb);
  else
    e = launch_maps<float>((const cx<float> *)h->F, (float *)h->maps, h->nmaps, h->map, s,
                           timed ? h->ev + 5 : nullptr, h->sched, prb);
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "map kernels: %s", hipGetErrorString(e));
  return 0;
}

static int finish(crime_gh *h, crime_gh_stats *st, bool fields, bool maps)
{
  HIP_OR_FAIL(h, hipMemcpyAsync(h->h_sig, h->d_sig, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));
  if (st) {
    if (fields) {
      st->ms_fill_fft_z = ev_ms(h->ev[0], h->ev[1]);
      st->ms_fft_y = ev_ms(h->ev[1], h->ev[2]);
      st->ms_fft_x = ev_ms(h->ev[2], h->ev[3]);
      st->ms_sigma2 = ev_ms(h->ev[3], h->ev[4]);
    }
    if (maps) {
      st->ms_maps_clear = ev_ms(h->ev[5], h->ev[6]);
      st->ms_pixelize = ev_ms(h->ev[6], h->ev[7]);
    }
    st->ms_total = ev_ms(fields ? h->ev[0] : h->ev[5], maps ? h->ev[7] : h->ev[4]);
    st->sigma2_gauss = h->h_sig[0];
    st->mean_gauss = h->h_sig[1];
  }
  return 0;
}

extern "C" int crime_gh_create_d_and_vr_fields(crime_gh_t *h, crime_gh_stats *st)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  int rc = enqueue_fields(h, true);
  if (rc) return rc;
  rc = finish(h, st, true, false);
  if (rc) return rc;
  h->fields_done = true;
  return 0;
}

extern "C" int crime_gh_mk_T_maps(crime_gh_t *h, crime_gh_stats *st)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "mk_T_maps before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  int rc = enqueue_maps(h, true);
  if (rc) return rc;
  return finish(h, st, false, true);
}

extern "C" int crime_gh_run(crime_gh_t *h, crime_gh_stats *st)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  int rc = enqueue_fields(h, true);
  if (rc) return rc;
  rc = enqueue_maps(h, true);
  if (rc) return rc;
  rc = finish(h, st, true, true);
  if (rc) return rc;
  h->fields_done = true;
  return 0;
}

extern "C" int crime_gh_power_spectrum(crime_gh_t *h, int nbins, double bin_dk, double *k_mean, double *pk,
                                       int64_t *nmodes)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!k_mean || !pk || !nmodes) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (nbins < 1 || nbins > PK_MAX_BINS) return set_error(h, CRIME_GH_EINVAL, "nbins %d not in [1, %d]", nbins, PK_MAX_BINS);
  if (!(bin_dk > 0)) return set_error(h, CRIME_GH_EINVAL, "bin_dk must be > 0");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "power_spectrum before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  const double dk = 2 * M_PI / h->d.l_box, n3 = (double)h->n * h->n * h->n;
  double *acc = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&acc, 3 * (size_t)nbins * sizeof(double)));
  // acc is cleared by a copy (no memset node) before the transform
  std::vector<double> hacc(3 * (size_t)nbins, 0.0);
  hipError_t e = hipMemcpyAsync(acc, hacc.data(), hacc.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
  const double pscale = h->d.l_box * h->d.l_box * h->d.l_box / (n3 * n3), binw = bin_dk / dk;
  if (e == hipSuccess)
    e = h->prec ? launch_power_spectrum<double>(h->n, (cx<double> *)h->F, (const cx<double> *)h->d_tw, pscale, binw,
                                                nbins, acc, h->stream)
                : launch_power_spectrum<float>(h->n, (cx<float> *)h->F, (const cx<float> *)h->d_tw, pscale, binw,
                                               nbins, acc, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hacc.data(), acc, hacc.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(acc);
  h->fields_done = false;  // the grid now holds the transform
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "power_spectrum: %s", hipGetErrorString(e));
  for (int i = 0; i < nbins; i++) {
    double c = hacc[3 * i + 2];
    nmodes[i] = (int64_t)c;
    k_mean[i] = c > 0 ? hacc[3 * i + 1] / c * dk : 0.0;
    pk[i] = c > 0 ? hacc[3 * i] / c : 0.0;
  }
  return 0;
}

// zero n doubles (a kernel, not a memset: no memset node if a caller captures the stream)
__global__ __launch_bounds__(256) void k_scratch_zero(double *__restrict__ p, int n)
{
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

__global__ __launch_bounds__(256) void k_zero_f64(double *__restrict__ p, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0;
}

// shell sums -> P, mean |k| and counts, the same operations as crime_gh_power_spectrum's host loop
__global__ __launch_bounds__(256) void k_pk_finish(const double *__restrict__ acc, int nbins, double dk,
                                                   double *__restrict__ pk, double *__restrict__ k_mean,
                                                   int64_t *__restrict__ nmodes)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nbins) return;
  const double c = acc[3 * i + 2];
  nmodes[i] = (int64_t)c;
  k_mean[i] = c > 0 ? acc[3 * i + 1] / c * dk : 0.0;
  pk[i] = c > 0 ? acc[3 * i] / c : 0.0;
}

extern "C" int crime_gh_power_spectrum_noalloc(crime_gh_t *h, int nbins, double bin_dk, double *k_mean, double *pk,
                                               int64_t *nmodes)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!k_mean || !pk || !nmodes) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (nbins < 1 || nbins > PK_MAX_BINS)
    return set_error(h, CRIME_GH_EINVAL, "nbins %d not in [1, %d] (the size of the static scratch)", nbins, PK_MAX_BINS);
  if (!(bin_dk > 0)) return set_error(h, CRIME_GH_EINVAL, "bin_dk must be > 0");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "power_spectrum before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double *sp;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    sp = g_scratch[h->d.device];
  }
  if (!sp) return set_error(h, CRIME_GH_ESTATE, "no scratch address for device %d", h->d.device);
  double *acc = sp, *d_pk = sp + 3 * PK_MAX_BINS, *d_k = sp + 4 * PK_MAX_BINS;
  int64_t *d_n = (int64_t *)(sp + 5 * PK_MAX_BINS);
  const double dk = 2 * M_PI / h->d.l_box, n3 = (double)h->n * h->n * h->n;
  const double pscale = h->d.l_box * h->d.l_box * h->d.l_box / (n3 * n3), binw = bin_dk / dk;
  hipLaunchKernelGGL(k_scratch_zero, dim3((3 * nbins + 255) / 256), dim3(256), 0, h->stream, acc, 3 * nbins);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = h->prec ? launch_power_spectrum<double>(h->n, (cx<double> *)h->F, (const cx<double> *)h->d_tw, pscale, binw,
                                                nbins, acc, h->stream)
                : launch_power_spectrum<float>(h->n, (cx<float> *)h->F, (const cx<float> *)h->d_tw, pscale, binw,
                                               nbins, acc, h->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pk_finish, dim3((nbins + 255) / 256), dim3(256), 0, h->stream, acc, nbins, dk, d_pk, d_k, d_n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(pk, d_pk, nbins * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(k_mean, d_k, nbins * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(nmodes, d_n, nbins * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  h->fields_done = false;  // the grid now holds the transform
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, CRIME_GH_EHIP, "power_spectrum_noalloc: %s", hipGetErrorString(e));
  }
  return 0;
}

extern "C" int crime_gh_set_seed(crime_gh_t *h, uint32_t seed)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (h->d.mode_source != CRIME_GH_MODES_PHILOX)
    return set_error(h, CRIME_GH_ESTATE, "set_seed: the modes of this handle are injected");
  h->d.seed = seed;
  h->fill.seed = seed;
  h->fields_done = false;
  return 0;
}

extern "C" int crime_gh_get_schedule(const crime_gh_t *h, int *sched)
{
  if (!h || !sched) return CRIME_GH_EINVAL;
  sched[0] = h->sched.fill;
  sched[1] = h->sched.y;
  sched[2] = h->sched.x;
  sched[3] = h->sched.pix;
  return 0;
}

extern "C" int crime_gh_set_schedule(crime_gh_t *h, const int *sched)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!sched) return set_error(h, CRIME_GH_EINVAL, "schedule is NULL");
  for (int i = 0; i < 4; i++)
    if (sched[i] < 0 || sched[i] >= (i < 3 ? N_SCHED : N_SCHED_PIX))
      return set_error(h, CRIME_GH_EINVAL, "schedule[%d] = %d", i, sched[i]);
  const bool tuned = h->n >= TUNE_MIN_N;
  h->sched.fill = tuned ? sched[0] : 0;
  h->sched.y = tuned ? sched[1] : 0;
  h->sched.x = tuned ? sched[2] : 0;
  h->sched.pix = sched[3];
  return 0;
}

// Runs every variant of each stage `reps` times on the handle's own buffers and keeps the fastest
// (minimum over the repetitions).  Stages are tuned in order, each with the winners of the ones
// before it; the field stage is rerun before the map variants.  Leaves fields and maps computed.
extern "C" int crime_gh_autotune(crime_gh_t *h, int reps, int *sched_out, double *ms_out)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (reps < 1) return set_error(h, CRIME_GH_EINVAL, "reps must be >= 1");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double table[4][N_SCHED];
  for (int i = 0; i < 4; i++)
    for (int v = 0; v < N_SCHED; v++) table[i][v] = -1;
  const bool tuned = h->n >= TUNE_MIN_N;
  int *field_slot[3] = {&h->sched.fill, &h->sched.y, &h->sched.x};
  for (int stage = 0; stage < 3; stage++) {
    int best = 0;
    for (int v = 0; v < (tuned ? N_SCHED : 1); v++) {
      *field_slot[stage] = v;
      double t = 1e300;
      for (int r = 0; r < reps; r++) {
        int rc = enqueue_fields(h, true);
        if (rc) return rc;
        HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));
        double ms = ev_ms(h->ev[stage], h->ev[stage + 1]);
        if (ms >= 0 && ms < t) t = ms;
      }
      table[stage][v] = t;
      if (t < table[stage][best]) best = v;
    }
    *field_slot[stage] = best;
  }
  int rc = enqueue_fields(h, true);
  if (rc) return rc;
  rc = finish(h, nullptr, true, false);
  if (rc) return rc;
  h->fields_done = true;
  int best = 0;
  for (int v = 0; v < N_SCHED_PIX; v++) {
    h->sched.pix = v;
    double t = 1e300;
    for (int r = 0; r < reps; r++) {
      rc = enqueue_maps(h, true);
      if (rc) return rc;
      HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));
      double ms = ev_ms(h->ev[6], h->ev[7]);
      if (ms >= 0 && ms < t) t = ms;
    }
    table[3][v] = t;
    if (t < table[3][best]) best = v;
  }
  h->sched.pix = best;
  rc = enqueue_maps(h, false);
  if (rc) return rc;
  HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));
  if (sched_out) crime_gh_get_schedule(h, sched_out);
  if (ms_out)
    for (int i = 0; i < 4; i++)
      for (int v = 0; v < N_SCHED; v++) ms_out[i * N_SCHED + v] = table[i][v];
  return 0;
}

// Key of a handle's kind for the schedule cache: everything the kernel variants' speed depends on
static std::string schedule_key(const crime_gh *h)
{
  hipDeviceProp_t prop;
  const char *arch = hipGetDeviceProperties(&prop, h->d.device) == hipSuccess ? prop.gcnArchName : "gfx950";
  char buf[256];
  snprintf(buf, sizeof(buf), "v%d:%s:N%d:p%d:ns%ld:nu%d:sub%d:vel%d:det%d:mode%d", CRIME_GH_SCHED_VERSION, arch,
           h->n, h->prec, (long)h->nside, h->n_nu, h->d.n_sub, h->d.velocity_mode, h->d.deterministic, h->d.nu_mode);
  std::string k = buf;
  for (char &c : k)
    if (c == ' ') c = '_';
  return k;
}

extern "C" int crime_gh_autotune_cached(crime_gh_t *h, const char *path, int reps, int *sched_out, int *hit)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!path) return set_error(h, CRIME_GH_EINVAL, "cache path is NULL");
  const std::string key = schedule_key(h);
  int sched[4];
  bool found = false;
  if (FILE *f = fopen(path, "r")) {
    char line[512], k[300];
    int a, b, c, d;
    while (fgets(line, sizeof(line), f))
      if (sscanf(line, "%299s %d %d %d %d", k, &a, &b, &c, &d) == 5 && key == k) {
        sched[0] = a;  // the last entry of a key wins
        sched[1] = b;
        sched[2] = c;
        sched[3] = d;
        found = true;
      }
    fclose(f);
  }
  if (hit) *hit = found ? 1 : 0;
  if (found) {
    int rc = crime_gh_set_schedule(h, sched);
    if (rc) return rc;
  }
  else {
    int rc = crime_gh_autotune(h, reps, sched, nullptr);
    if (rc) return rc;
    FILE *f = fopen(path, "a");
    if (!f) return set_error(h, CRIME_GH_EINVAL, "cannot append to the schedule cache %s", path);
    // one short line per append: appends of concurrent tuners do not interleave
    fprintf(f, "%s %d %d %d %d\n", key.c_str(), sched[0], sched[1], sched[2], sched[3]);
    fclose(f);
  }
  if (sched_out) crime_gh_get_schedule(h, sched_out);
  return 0;
}

extern "C" int crime_gh_copy_maps(crime_gh_t *h, float *out)
{
  if (!h || !out) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  if (!h->prec) {
    HIP_OR_FAIL(h, hipMemcpy(out, h->maps, h->nmaps * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<double> tmp(h->nmaps);
  HIP_OR_FAIL(h, hipMemcpy(tmp.data(), h->maps, h->nmaps * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h->nmaps; i++) out[i] = (float)tmp[i];
  return 0;
}

extern "C" int crime_gh_copy_maps_channels(crime_gh_t *h, int c0, int nc, float *out)
{
  if (!h || !out) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (c0 < 0 || nc < 1 || c0 + nc > h->n_nu)
    return set_error(h, CRIME_GH_EINVAL, "channels [%d, %d) outside [0, %d)", c0, c0 + nc, h->n_nu);
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  const size_t n = (size_t)nc * h->npix, off = (size_t)c0 * h->npix;
  if (!h->prec) {
    HIP_OR_FAIL(h, hipMemcpy(out, (const float *)h->maps + off, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<double> tmp(n);
  HIP_OR_FAIL(h, hipMemcpy(tmp.data(), (const double *)h->maps + off, n * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) out[i] = (float)tmp[i];
  return 0;
}

extern "C" int crime_gh_host_alloc(size_t bytes, void **out)
{
  if (!out || !bytes) return set_error(nullptr, CRIME_GH_EINVAL, "host_alloc: NULL or empty");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return set_error(nullptr, CRIME_GH_ENOMEM, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return 0;
}

extern "C" void crime_gh_host_free(void *p)
{
  if (p) (void)hipHostFree(p);
}

extern "C" int crime_gh_copy_maps_f64(crime_gh_t *h, double *out)
{
  if (!h || !out) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  if (h->prec) {
    HIP_OR_FAIL(h, hipMemcpy(out, h->maps, h->nmaps * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<float> tmp(h->nmaps);
  HIP_OR_FAIL(h, hipMemcpy(tmp.data(), h->maps, h->nmaps * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h->nmaps; i++) out[i] = tmp[i];
  return 0;
}

extern "C" int crime_gh_get_sigma2(crime_gh_t *h, double *sigma2, double *mean)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double v[2];
  HIP_OR_FAIL(h, hipMemcpy(v, h->d_sig, sizeof(v), hipMemcpyDeviceToHost));
  if (sigma2) *sigma2 = v[0];
  if (mean) *mean = v[1];
  return 0;
}

extern "C" int crime_gh_set_sigma2(crime_gh_t *h, double sigma2, double mean)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double v[2] = {sigma2, mean};
  HIP_OR_FAIL(h, hipMemcpy(h->d_sig, v, sizeof(v), hipMemcpyHostToDevice));
  return 0;
}

// grid holding a stored field and its half (0 = real, 1 = imaginary); false for a derived field
static bool stored_grid(crime_gh *h, int which, void **grid, int *half)
{
  const bool spec = h->d.velocity_mode == CRIME_GH_VEL_SPECTRAL;
  switch (which) {
  case CRIME_GH_GRID_DENS: *grid = h->F; *half = 0; return true;
  case CRIME_GH_GRID_VPOT: *grid = spec ? nullptr : h->F; *half = 1; return true;
  case CRIME_GH_GRID_VX: *grid = spec ? h->F : nullptr; *half = 1; return true;
  case CRIME_GH_GRID_VY: *grid = h->G; *half = 0; return true;
  case CRIME_GH_GRID_VZ: *grid = h->G; *half = 1; return true;
  default: return false;
  }
}

extern "C" int crime_gh_copy_grid_range(crime_gh_t *h, int which, int64_t cell0, int64_t ncell, double *dst)
{
  if (!h || !dst) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (which < CRIME_GH_GRID_DENS || which > CRIME_GH_GRID_VZ) return set_error(h, CRIME_GH_EINVAL, "grid %d", which);
  if (cell0 < 0 || ncell < 1 || (size_t)(cell0 + ncell) > h->ncell)
    return set_error(h, CRIME_GH_EINVAL, "cell range [%lld,%lld) outside the grid", (long long)cell0,
                     (long long)(cell0 + ncell));
  void *grid = nullptr;
  int half = 0;
  const bool stored = stored_grid(h, which, &grid, &half);
  if (stored && !grid)
    return set_error(h, CRIME_GH_EINVAL, "grid %d is not held in velocity mode %d", which, h->d.velocity_mode);
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double *tmp = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&tmp, (size_t)ncell * sizeof(double)));
  hipError_t e;
  const size_t esz = h->prec ? sizeof(cx<double>) : sizeof(cx<float>);
  if (stored)
    e = h->prec ? launch_grid_get<double>((const cx<double> *)((char *)grid + cell0 * esz), ncell, half, tmp, h->stream)
                : launch_grid_get<float>((const cx<float> *)((char *)grid + cell0 * esz), ncell, half, tmp, h->stream);
  else
    e = h->prec ? launch_cell_field<double>((const cx<double> *)h->F, h->map, which, cell0, ncell, tmp, h->stream)
                : launch_cell_field<float>((const cx<float> *)h->F, h->map, which, cell0, ncell, tmp, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(dst, tmp, (size_t)ncell * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "copy_grid: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int crime_gh_copy_grid(crime_gh_t *h, int which, double *dst)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  return crime_gh_copy_grid_range(h, which, 0, (int64_t)h->ncell, dst);
}

extern "C" int crime_gh_set_grid(crime_gh_t *h, int which, const double *src)
{
  if (!h || !src) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  void *grid = nullptr;
  int half = 0;
  if (which < CRIME_GH_GRID_DENS || which > CRIME_GH_GRID_VZ || !stored_grid(h, which, &grid, &half) || !grid)
    return set_error(h, CRIME_GH_EINVAL, "grid %d cannot be set (DENS and VPOT, or DENS and VX/VY/VZ in spectral mode)",
                     which);
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double *tmp = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&tmp, h->ncell * sizeof(double)));
  hipError_t e = hipMemcpy(tmp, src, h->ncell * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = h->prec ? launch_grid_set<double>((cx<double> *)grid, h->ncell, half, tmp, h->stream)
                : launch_grid_set<float>((cx<float> *)grid, h->ncell, half, tmp, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  // the grid's bound on |phi| no longer holds: the map stage culls by each cell's own shift only
  const double inf = INFINITY;
  if (e == hipSuccess) e = hipMemcpy(h->d_sig + 2, &inf, sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "set_grid: %s", hipGetErrorString(e));
  h->fields_done = true;
  return 0;
}

extern "C" int crime_gh_sample_indices(crime_gh_t *h, int64_t cell0, int64_t ncell, int32_t *inu, int64_t *ipix)
{
  if (!h || !inu || !ipix) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (cell0 < 0 || ncell < 1 || (size_t)(cell0 + ncell) > h->ncell)
    return set_error(h, CRIME_GH_EINVAL, "cell range [%lld,%lld) outside the grid", (long long)cell0,
                     (long long)(cell0 + ncell));
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  size_t ns = (size_t)ncell * h->map.n_sub;
  int32_t *d_inu = nullptr;
  int64_t *d_pix = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&d_inu, ns * sizeof(int32_t)));
  hipError_t e = hipMalloc((void **)&d_pix, ns * sizeof(int64_t));
  if (e == hipSuccess)
    e = h->prec ? launch_sample_indices<double>((const cx<double> *)h->F, h->map, cell0, ncell, d_inu, d_pix, h->stream)
                : launch_sample_indices<float>((const cx<float> *)h->F, h->map, cell0, ncell, d_inu, d_pix, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(inu, d_inu, ns * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ipix, d_pix, ns * sizeof(int64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_inu);
  if (d_pix) (void)hipFree(d_pix);
  if (e != hipSuccess) return set_error(h, CRIME_GH_EHIP, "sample_indices: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int crime_gh_channel_totals(crime_gh_t *h, int64_t cell0, int64_t ncell, double *tot, int64_t *count)
{
  if (!h || !tot) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (cell0 < 0 || ncell < 1 || (size_t)(cell0 + ncell) > h->ncell)
    return set_error(h, CRIME_GH_EINVAL, "cell range [%lld,%lld) outside the grid", (long long)cell0,
                     (long long)(cell0 + ncell));
  if (h->n_nu > 2048) return set_error(h, CRIME_GH_EINVAL, "channel_totals supports n_nu <= 2048");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "channel_totals before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double *d_tot = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&d_tot, (size_t)h->n_nu * (sizeof(double) + sizeof(unsigned long long))));
  unsigned long long *d_cnt = (unsigned long long *)(d_tot + h->n_nu);
  hipError_t e = h->prec ? launch_channel_totals<double>((const cx<double> *)h->F, h->map, cell0, ncell, d_tot, d_cnt,
                                                         h->stream)
                         : launch_channel_totals<float>((const cx<float> *)h->F, h->map, cell0, ncell, d_tot, d_cnt,
                                                        h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(tot, d_tot, h->n_nu * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && count) e = hipMemcpy(count, d_cnt, h->n_nu * sizeof(int64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_tot);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, CRIME_GH_EHIP, "channel_totals: %s", hipGetErrorString(e));
  }
  return 0;
}

extern "C" int crime_gh_screen_check(crime_gh_t *h, int64_t *counts, double *eps)
{
  if (!h || !counts) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "screen_check before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  if (eps) {
    eps[0] = h->eps_nu;
    eps[1] = h->map.scr.eps_pix;
  }
  for (int i = 0; i < 5; i++) counts[i] = 0;
  if (!h->map.scr.on) return 0;
  unsigned long long *d_c = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&d_c, 5 * sizeof(unsigned long long)));
  hipError_t e = h->prec ? launch_screen_check<double>((const cx<double> *)h->F, h->map, d_c, h->stream)
                         : launch_screen_check<float>((const cx<float> *)h->F, h->map, d_c, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(counts, d_c, 5 * sizeof(int64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, CRIME_GH_EHIP, "screen_check: %s", hipGetErrorString(e));
  }
  return 0;
}

extern "C" int crime_gh_reference_maps(crime_gh_t *h, double *out)
{
  if (!h || !out) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "reference_maps before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  double *ref = nullptr;
  HIP_OR_FAIL(h, hipMalloc((void **)&ref, h->nmaps * sizeof(double)));
  hipLaunchKernelGGL(k_zero_f64, dim3(4096), dim3(256), 0, h->stream, ref, h->nmaps);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = h->prec ? launch_reference_maps<double>((const cx<double> *)h->F, h->map, ref, h->stream)
                : launch_reference_maps<float>((const cx<float> *)h->F, h->map, ref, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, ref, h->nmaps * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(ref);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, CRIME_GH_EHIP, "reference_maps: %s", hipGetErrorString(e));
  }
  return 0;
}

extern "C" int crime_gh_device_maps(crime_gh_t *h, void **dptr, size_t *bytes)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (dptr) *dptr = h->maps;
  if (bytes) *bytes = h->nmaps * (h->prec ? sizeof(double) : sizeof(float));
  return 0;
}

// ---------------------------------------------------------------------------------------------
// point sources
// ---------------------------------------------------------------------------------------------
// setup_psources (src/psources.c:99-128): n(z_i) = sum over 256 log10 L midpoints of dlogL
// l_z_function, and max_Lpdf = 1.1 max l_z_function / n(z_i), at z_i = i z_max / 256
static double ps_lfunc_host(double L, double z)
{
  const double rho_l = 2.5 * M_LN10 * PS_RHO * pow(L / PS_LX, PS_ALPHA + 1) * exp(-L / PS_LX);
  const double fz = z < 1.5 ? pow(1 + z, 3.1) : 17.124;
  return fz * rho_l;
}

static void ps_tables(double z_max, double *nz, double *maxpdf)
{
  const double dlogL = (PS_LLOGMAX - PS_LLOGMIN) / PS_NL, dz = z_max / PS_NZ;
  for (int ii = 0; ii < PS_NZ; ii++) {
    const double z = ii * dz;
    double acc = 0, mx = -1;
    for (int jj = 0; jj < PS_NL; jj++) {
      const double lf = ps_lfunc_host(pow(10., PS_LLOGMIN + (jj + 0.5) * dlogL), z);
      if (lf >= mx) mx = lf;
      acc += dlogL * lf;
    }
    nz[ii] = acc;
    maxpdf[ii] = acc > 0 ? 1.1 * mx / acc : 0;
  }
}

extern "C" int crime_gh_mk_psources_maps(crime_gh_t *h, crime_gh_ps_stats *st)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!h->fields_done) return set_error(h, CRIME_GH_ESTATE, "mk_psources_maps before create_d_and_vr_fields");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  const size_t npix = (size_t)h->npix, nps = 2 * PS_NZ + 2 * (size_t)h->n_nu + 2 * npix;
  const size_t esz = h->prec ? sizeof(double) : sizeof(float);
  if (!h->d_ps) {
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_ps, nps * sizeof(double)));
    HIP_OR_FAIL(h, hipMalloc((void **)&h->d_ps_count, sizeof(unsigned long long)));
    HIP_OR_FAIL(h, hipMalloc(&h->ps_maps, (size_t)h->n_nu * npix * esz));
    // tables and channel factors: T(pix, nu) = K (A nu^-2.1 + B nu^-2.75), K = FLUX2TEMP LUM2FLUX
    // 4 pi SED_NORM h^2 / dOmega, nu the channel centre (pixelize.c:62-69)
    std::vector<double> host(2 * PS_NZ + 2 * (size_t)h->n_nu);
    const double z_max = CRIME_GH_NU_21 / h->d.nu_min - 1;   // cosmo.c:357
    ps_tables(z_max, host.data(), host.data() + PS_NZ);
    const double domega = 4 * M_PI / (double)npix;
    const double K = PS_FLUX2TEMP * PS_LUM2FLUX * 4 * M_PI * PS_SED_NORM * h->d.hhub * h->d.hhub / domega;
    for (int i = 0; i < h->n_nu; i++) {
      const double nu = h->ps_nu[i];
      host[2 * PS_NZ + i] = K * pow(nu, -2.1);
      host[2 * PS_NZ + h->n_nu + i] = K * pow(nu, -2.75);
    }
    HIP_OR_FAIL(h, hipMemcpy(h->d_ps, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  PsParams ps{};
  ps.nz = h->d_ps;
  ps.maxpdf = h->d_ps + PS_NZ;
  ps.z_max = CRIME_GH_NU_21 / h->d.nu_min - 1;
  ps.dz = ps.z_max / PS_NZ;
  ps.inv_dz = 1.0 / ps.dz;
  ps.cell_vol = h->map.dx * h->map.dx * h->map.dx;
  ps.seed = h->d.seed;
  ps.A = h->d_ps + 2 * PS_NZ + 2 * (size_t)h->n_nu;
  ps.B = ps.A + npix;
  ps.count = h->d_ps_count;
  const double *c1 = h->d_ps + 2 * PS_NZ, *c2 = c1 + h->n_nu;
  HIP_OR_FAIL(h, hipEventRecord(h->ev[0], h->stream));
  HIP_OR_FAIL(h, hipMemsetAsync(ps.A, 0, 2 * npix * sizeof(double), h->stream));
  HIP_OR_FAIL(h, hipMemsetAsync(h->d_ps_count, 0, sizeof(unsigned long long), h->stream));
  hipError_t e = h->prec ? launch_psources<double>((const cx<double> *)h->F, h->map, ps, h->stream)
                         : launch_psources<float>((const cx<float> *)h->F, h->map, ps, h->stream);
  HIP_OR_FAIL(h, e);
  HIP_OR_FAIL(h, hipEventRecord(h->ev[1], h->stream));
  e = h->prec ? launch_ps_expand<double>(ps.A, ps.B, c1, c2, h->n_nu, h->npix, (double *)h->ps_maps, h->stream)
              : launch_ps_expand<float>(ps.A, ps.B, c1, c2, h->n_nu, h->npix, (float *)h->ps_maps, h->stream);
  HIP_OR_FAIL(h, e);
  HIP_OR_FAIL(h, hipEventRecord(h->ev[2], h->stream));
  HIP_OR_FAIL(h, hipStreamSynchronize(h->stream));
  if (st) {
    unsigned long long nsrc = 0;
    HIP_OR_FAIL(h, hipMemcpy(&nsrc, h->d_ps_count, sizeof(nsrc), hipMemcpyDeviceToHost));
    st->n_sources = (int64_t)nsrc;
    st->ms_sources = ev_ms(h->ev[0], h->ev[1]);
    st->ms_maps = ev_ms(h->ev[1], h->ev[2]);
  }
  return 0;
}

extern "C" int crime_gh_copy_ps_maps(crime_gh_t *h, double *out)
{
  if (!h || !out) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (!h->ps_maps) return set_error(h, CRIME_GH_ESTATE, "no point-source maps yet");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  const size_t n = (size_t)h->n_nu * h->npix;
  if (h->prec) {
    HIP_OR_FAIL(h, hipMemcpy(out, h->ps_maps, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<float> tmp(n);
  HIP_OR_FAIL(h, hipMemcpy(tmp.data(), h->ps_maps, n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) out[i] = tmp[i];
  return 0;
}

extern "C" int crime_gh_ps_pixel_sums(crime_gh_t *h, double *A, double *B)
{
  if (!h || !A || !B) return set_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (!h->d_ps) return set_error(h, CRIME_GH_ESTATE, "no point sources yet");
  HIP_OR_FAIL(h, hipSetDevice(h->d.device));
  const double *dA = h->d_ps + 2 * PS_NZ + 2 * (size_t)h->n_nu;
  HIP_OR_FAIL(h, hipMemcpy(A, dA, (size_t)h->npix * sizeof(double), hipMemcpyDeviceToHost));
  HIP_OR_FAIL(h, hipMemcpy(B, dA + h->npix, (size_t)h->npix * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int crime_gh_device_ps_maps(crime_gh_t *h, void **dptr, size_t *bytes)
{
  if (!h) return set_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (dptr) *dptr = h->ps_maps;
  if (bytes) *bytes = h->ps_maps ? (size_t)h->n_nu * h->npix * (h->prec ? sizeof(double) : sizeof(float)) : 0;
  return 0;
}

// ---------------------------------------------------------------------------------------------
// batches: one captured hipGraph over many handles
// ---------------------------------------------------------------------------------------------
struct crime_gh_batch {
  std::vector<crime_gh *> hs;
  int device = 0;
  hipStream_t origin = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> evs;  // fork + one join per stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
};

static int batch_error(crime_gh_batch *b, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  else g_create_err = buf;
  return code;
}

#define BATCH_HIP(b, call)                                                                           \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) {                                                                         \
      (void)hipGetLastError(); /* reported here: a later launch check must not see it again */      \
      return batch_error(b, CRIME_GH_EHIP, "%s: %s", #call, hipGetErrorString(e_));                 \
    }                                                                                               \
  } while (0)

static void batch_free(crime_gh_batch *b)
{
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->evs) (void)hipEventDestroy(e);
  for (auto s : b->streams) (void)hipStreamDestroy(s);
  if (b->origin) (void)hipStreamDestroy(b->origin);
}

static int batch_capture(crime_gh_batch *b, int n_streams)
{
  BATCH_HIP(b, hipSetDevice(b->device));
  BATCH_HIP(b, hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
  b->streams.resize(n_streams, nullptr);
  b->evs.resize(n_streams + 1, nullptr);
  for (auto &s : b->streams) BATCH_HIP(b, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  for (auto &e : b->evs) BATCH_HIP(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BATCH_HIP(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  hipError_t e = hipEventRecord(b->evs[0], b->origin);
  for (int i = 0; i < n_streams && e == hipSuccess; i++) e = hipStreamWaitEvent(b->streams[i], b->evs[0], 0);
  // members round-robin over the streams, largest grids first so the tail is made of small members
  std::vector<int> order(b->hs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->hs[a]->ncell > b->hs[c]->ncell; });
  for (size_t k = 0; k < order.size() && e == hipSuccess && !rc; k++) {
    crime_gh *h = b->hs[order[k]];
    hipStream_t s = b->streams[k % n_streams];
    rc = enqueue_fields(h, false, s);
    if (!rc) rc = enqueue_maps(h, false, s);
    if (rc) batch_error(b, rc, "member %d: %s", order[k], h->err.c_str());
  }
  for (int i = 0; i < n_streams && e == hipSuccess; i++) {
    e = hipEventRecord(b->evs[i + 1], b->streams[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->evs[i + 1], 0);
  }
  hipError_t e2 = hipStreamEndCapture(b->origin, &b->graph);
  if (rc) return rc;
  if (e != hipSuccess) return batch_error(b, CRIME_GH_EHIP, "capture: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return batch_error(b, CRIME_GH_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
  BATCH_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return 0;
}

extern "C" int crime_gh_batch_create(crime_gh_t **hs, int n, int n_streams, crime_gh_batch_t **out)
{
  if (!out) return batch_error(nullptr, CRIME_GH_EINVAL, "out is NULL");
  *out = nullptr;
  if (!hs || n < 1) return batch_error(nullptr, CRIME_GH_EINVAL, "empty batch");
  if (n_streams < 1 || n_streams > 32) return batch_error(nullptr, CRIME_GH_EINVAL, "n_streams=%d", n_streams);
  for (int i = 0; i < n; i++) {
    if (!hs[i]) return batch_error(nullptr, CRIME_GH_EINVAL, "member %d is NULL", i);
    if (hs[i]->d.device != hs[0]->d.device)
      return batch_error(nullptr, CRIME_GH_EINVAL, "member %d is on device %d, member 0 on %d", i,
                         hs[i]->d.device, hs[0]->d.device);
    for (int j = 0; j < i; j++)
      if (hs[j] == hs[i]) return batch_error(nullptr, CRIME_GH_EINVAL, "member %d repeats member %d", i, j);
  }
  crime_gh_batch *b = new crime_gh_batch();
  b->hs.assign(hs, hs + n);
  b->device = hs[0]->d.device;
  int rc = batch_capture(b, n_streams < n ? n_streams : n);
  if (rc) {
    g_create_err = b->err;
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

extern "C" int crime_gh_batch_run(crime_gh_batch_t *b, double *ms)
{
  if (!b) return batch_error(nullptr, CRIME_GH_EINVAL, "batch is NULL");
  BATCH_HIP(b, hipSetDevice(b->device));
  auto t0 = std::chrono::steady_clock::now();
  BATCH_HIP(b, hipGraphLaunch(b->exec, b->origin));
  BATCH_HIP(b, hipStreamSynchronize(b->origin));
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (auto h : b->hs) h->fields_done = true;
  return 0;
}

extern "C" void crime_gh_batch_destroy(crime_gh_batch_t *b)
{
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  batch_free(b);
  delete b;
}

extern "C" const char *crime_gh_batch_last_error(const crime_gh_batch_t *b)
{
  return b ? b->err.c_str() : g_create_err.c_str();
}

extern "C" int crime_gh_run_batch(crime_gh_t **hs, int n, crime_gh_stats *st)
{
  crime_gh_batch_t *b = nullptr;
  int rc = crime_gh_batch_create(hs, n, 4, &b);
  if (rc) return rc;
  double ms = 0;
  rc = crime_gh_batch_run(b, &ms);
  if (!rc && st) {
    for (int i = 0; i < n; i++) {
      memset(&st[i], 0, sizeof(st[i]));
      st[i].ms_total = ms;
      rc = crime_gh_get_sigma2(hs[i], &st[i].sigma2_gauss, &st[i].mean_gauss);
      if (rc) break;
    }
  }
  if (rc) g_create_err = b->err.empty() ? std::string(crime_gh_last_error(nullptr)) : b->err;
  crime_gh_batch_destroy(b);
  return rc;
}
