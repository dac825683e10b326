// Shared declarations of the gfx950 kernels of the GetHI density-to-sky path: launch shapes,
// kernel parameter blocks and the host-side launchers called by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "gh_fft.hpp"

namespace crime {

// grid sizes with a compiled FFT plan: 2^a (8..2048) and 3*2^a (12..1536)
#define CRIME_FOR_EACH_N(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) \
  X(1536) X(2048)

constexpr uint32_t PHILOX_KEY1 = 0x47484931u;  // second Philox key word ("GHI1")

constexpr int cpow2_div(int n)
{
  int p = 1;
  while (n % (2 * p) == 0) p *= 2;
  return p;
}
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// max elements per thread kept in VGPRs: 64 complex float or 32 complex double
template <typename T>
constexpr int emax() { return sizeof(T) == 4 ? 64 : 32; }

// Schedule variants the autotuner chooses between, per stage (variant 0 is the default).  FFT stages:
// variants 0-2 differ in tiling with radix-16 plans, variant 3 is variant 0's tiling with radix-8 plans
// (one pass more, smaller butterflies).  Map stage: 2048 LDS bins grouped by 4 adjacent map words
// (variant 0) or by 8 (variant 1), 4096 bins (2) and 1024 bins (3).  Grids below TUNE_MIN_N have variant
// 0 only for the FFT stages.
constexpr int N_SCHED = 4;
constexpr int N_SCHED_PIX = 4;
constexpr int CRIME_GH_SCHED_VERSION = 4;  // bump when a variant's meaning changes (schedule cache keys)
constexpr int TUNE_MIN_N = 128;
struct Schedule {
  int fill = 0, y = 0, x = 0, pix = 0;
};

// radix limit and tiling of FFT variant V
constexpr int fft_maxr(int V) { return V == 3 ? 8 : 16; }
constexpr int fft_tiling(int V) { return V == 3 ? 0 : V; }

// strided (kz, ky) passes: TX adjacent columns per workgroup, TC threads per column; variants aim at
// 16, 8 or 4 columns per tile
template <typename T, int N, int V = 0>
struct ColCfg {
  static constexpr int MAXR = fft_maxr(V);
  static constexpr int W = fft_tiling(V);
  static constexpr int TC = plan_tc(N, cmax(64, N / emax<T>()), MAXR);
  static constexpr int CB = N * (int)sizeof(cx<T>);
  static constexpr int TXMAX = cmin(W == 0 ? 16 : (W == 1 ? 8 : 4), cpow2_div(N));
  static constexpr int tx_for(int budget)
  {
    int tx = TXMAX;
    while (tx > 1 && tx * CB > budget) tx /= 2;
    return tx;
  }
  // 16 adjacent complex<float> columns = one 128-byte line per row: every HBM line fetched is used
  // whole (8 columns would fetch each line twice).  Up to 128 KB of LDS: one workgroup of up to
  // 1024 threads per CU.
  static constexpr int TX = cmin(tx_for(131072), 1024 / TC);
  static constexpr int NT = TX * TC;
};

// the fused fill + kz pass only writes (64-byte write requests waste nothing): a tile of at most
// 64 KB keeps two workgroups per CU, so one's mode generation overlaps the other's FFT; the
// variants take 32 KB or 128 KB tiles instead
template <typename T, int N, int V = 0>
struct FillCfg {
  static constexpr int MAXR = fft_maxr(V);
  static constexpr int W = fft_tiling(V);
  static constexpr int TC = ColCfg<T, N, V>::TC;
  static constexpr int TX = cmin(ColCfg<T, N, V>::tx_for(W == 0 ? 65536 : (W == 1 ? 32768 : 131072)), 1024 / TC);
  static constexpr int NT = TX * TC;
};

// contiguous kx pass: TY rows per workgroup, TC threads per row; the variants take half or twice the
// rows (when LDS, threads and N^2 allow)
template <typename T, int N, int V = 0>
struct RowCfg {
  static constexpr int MAXR = fft_maxr(V);
  static constexpr int W = fft_tiling(V);
  static constexpr int TC = plan_tc(N, cmax(64, N / emax<T>()), MAXR);
  static constexpr int RB = N * (int)sizeof(cx<T>);
  static constexpr int ty_for(int budget)
  {
    int ty = cmin(cmax(256 / TC, 1), cpow2_div(N * N));
    while (ty > 1 && ty * RB > budget) ty /= 2;
    return ty;
  }
  static constexpr int TY0 = ty_for(65536) * RB <= 65536 ? ty_for(65536) : ty_for(131072);
  static constexpr bool CAN2 = 2 * TY0 * RB <= 131072 && 2 * TY0 * TC <= 1024 && (N * N) % (2 * TY0) == 0;
  static constexpr int TY = W == 1 ? cmax(TY0 / 2, 1) : (W == 2 && CAN2 ? 2 * TY0 : TY0);
  static constexpr int NT = TY * TC;
};

struct FillParams {
  int n;
  double dk, idk3, f0h0, r2_smooth;
  double phi_scale;  // phi is packed as phi_scale*phi so both halves of F have similar magnitude
  int do_smoothing;
  const double *logk, *pk;
  int numk;
  double logkmin, logkmax, idlogk, n_s;
  uint32_t seed;
  int mode_source;
  const double *unit_modes;
  const void *amp;  // AmpEntry<T>[amp_count(n)] for k^2 = m dk^2, filled by launch_amp_table
  double psi_c;     // phi_scale f0 H0 / dk^2: psi = psi_c / m
  // velocity mode (CRIME_GH_VEL_*).  SPECTRAL: phi_scale scales the velocity components instead of
  // phi, and grid `comp` 0 holds delta + i phi_scale vx, grid 1 phi_scale (vy + i vz), with
  // v_j(k) = i sin_dx[m_j] phi_k and sin_dx[m] = sin(2 pi m / N) / dx (exactly 0 at m = 0, N/2)
  int vmode, comp;
  const double *sin_dx;
};

// amplitude-table entry: A (FP32 grids) or (A, A psi) (FP64 grids)
template <typename T>
using AmpEntry = typename std::conditional<sizeof(T) == 8, cx<double>, float>::type;

// entries of the amplitude table: every m = kx^2 + ky^2 + kz^2 with |k_i| <= N/2
inline int amp_count(int n) { return 3 * (n / 2) * (n / 2) + 1; }

template <typename T>
hipError_t launch_amp_table(const FillParams &f, void *tab, hipStream_t s);

// FP32 screen of the per-sample decisions (channel and RING pixel).  Each decision is taken from FP32
// arithmetic when its argument lies farther than a certified margin from the decision boundary, and
// from the FP64 path (the reference's arithmetic) otherwise; both give the same integers.
struct ScreenParams {
  int on;                  // 0: every decision in FP64
  const float2 *chwin;     // [n_nu] decided interval [lo, hi] of 1 + z for each channel: NU_21 / (1 + z) at
                           // least the frequency margin inside the channel's edges, rounded inwards to float
  double rc_lo, rc_hi;     // cells screened: rc_lo < r_c < rc_hi (the cloud keeps off r = r_min and the table end)
  double inv_2h, inv_2h2;  // 1 / (2 h_sub), 1 / (2 h_sub^2): the per-cell parabola of z(r)
  float out_lo, out_hi;    // 1 + z < out_lo or > out_hi: no channel (nu > nu_hi + eps or nu < nu_lo - eps)
  float eps_pix;           // margin of the pixel floors, in the units of their arguments
  float eps_ax;            // margin of the quadrant / octant boundaries (w near 0 or 1), quarter turns
  float eps_z;             // margin of |cos theta| around 2/3
  float half_pix, half_ax; // |d - 1/2| <= half (FP32) implies eps <= d <= 1 - eps, for the two margins above
};

struct MapParams {
  int n;
  double dx, obs[3];
  double inv_dx;           // 1 / dx (the same double as the division)
  RTables tab;
  int bias_model;
  double b0, b1;
  double mass_prefac;      // dx^3 (grid_tools.c:122)
  int rsd;
  int n_sub;
  const double *sub_off;   // [3*n_sub]
  const double *sub_offp;  // [3*n_sub] the offsets in the map stage's visiting order (a nearest-neighbour chain)
  double h_sub;            // bound on |offset| (with rounding margin)
  int n_nu, nu_mode;
  const double *edges;     // [n_nu+1]
  double nu_min, inv_dnu;
  double nu_lo, nu_hi;     // a sample can be binned only if nu_lo <= nu < nu_hi
  const double *prefac;    // [n_nu] mK factor per channel (pixelize.c:246-256)
  int64_t nside, npix;
  const double *sig;       // device [sigma2, mean, max |phi| of the grid or infinity]
  const ChanRec *crec;     // [n_guess] channel records of the buckets of [g_lo, g_lo + n_guess/g_inv) (table mode)
  int n_guess;
  double g_lo, g_inv;
  int precull;                 // cull whole tiles before any grid load (central-difference velocities)
  double dz_per_phi;           // |dz| <= dz_per_phi * sig[2] for every cell
  const uint32_t *tile_order;  // tiles by increasing distance of their centre from the observer, as
                               // tx | ty << 10 | tz << 20
  const uint32_t *bin_start;   // [nbins + 1] first entry of tile_order in each distance bin (floor(d / dx) of
                               // the tile's centre, the last bin also holding everything beyond); [nbins] = ntiles
  int nbins;
  double tile_hd;              // every cell centre of a tile lies within tile_hd of the tile's centre
  uint32_t *tile_range;        // device [first, last): the entries of tile_order the map kernel visits
  uint8_t *tile_cull;          // device [ntiles]: 1 when tile_order's entry i is out of the band (k_tile_cull;
                               // written for the range only)
  int ntx, nty, ntz;
  int vmode;               // CRIME_GH_VEL_*: SPECTRAL reads vx from F.y and (vy, vz) from G
  const void *G;           // cx<T>[N^3] second grid of SPECTRAL handles (NULL otherwise)
  // deterministic scatter: integer mass sums in fx[n_nu * npix], scale s and 1/s in fxs[0..1],
  // FX_PARTS fixed-order partial sums of the box's mass in mass_part
  int fx;
  unsigned long long *fxmaps;
  double *fxs;
  double *mass_part;
  ScreenParams scr;
};

constexpr int FX_PARTS = 2048;  // blocks of the fixed-order mass reduction

// LDS aggregation of the scatter: one workgroup per TILE_X x TILE_Y x TILE_Z cells, AGG_SLOTS bins.
// TILE_X = 16 cells of complex<float> = one 128-byte line per row of the grid.
constexpr int TILE_X = 16, TILE_Y = 4, TILE_Z = 4;
constexpr int AGG_SLOTS = 2048;

// Point sources (src/grid_tools.c:24-101, src/pixelize.c:58-148, src/psources.c): constants of the
// reference's toy luminosity function and SED (psources.c:24-64, 154-155) and the per-run tables.
constexpr int PS_NZ = 256;                      // NZ_PSOURCES (src/common_gh.h:35)
constexpr double PS_LLOGMIN = -5.0, PS_LLOGMAX = 6.0;
constexpr int PS_NL = 256;                      // NL_PSOURCES
constexpr double PS_LX = 2.11, PS_ALPHA = -0.633, PS_RHO = 3.2e-4;
constexpr double PS_SED_NORM = 0.114461, PS_FLUX2TEMP = 3.2548291e-2, PS_LUM2FLUX = 8.35774e7;
constexpr int PS_MAX_TRIES = 4096;              // cap of the luminosity rejection loop
constexpr uint32_t PS_KEY1 = 0x50535243u;       // second Philox key word ("PSRC")
constexpr int PS_SLOTS = 1024;                  // pixels of a tile's LDS table (24 KB)

struct PsParams {
  const double *nz;          // [PS_NZ] n(z_i) (nz_psources_arr)
  const double *maxpdf;      // [PS_NZ] (max_Lpdf_arr)
  double dz, inv_dz, z_max;  // glob_dz, glob_inv_dz, z_max (psources.c:106-107, cosmo.c:357)
  double cell_vol;
  uint32_t seed;
  double *A, *B;             // [npix] sums of the two SED terms per pixel
  unsigned long long *count; // sources drawn
};

template <typename T>
hipError_t launch_psources(const cx<T> *F, const MapParams &m, const PsParams &ps, hipStream_t s);
template <typename T>
hipError_t launch_ps_expand(const double *A, const double *B, const double *c1, const double *c2, int n_nu,
                            int64_t npix, T *maps, hipStream_t s);

// Field stage.  STENCIL: F = delta + i phi.  SPECTRAL: G = (vy, vz) first, then F = delta + i vx (the
// sigma^2 partials are those of F, whose kx pass runs last).
template <typename T>
hipError_t launch_fields(int n, cx<T> *F, cx<T> *G, const FillParams &f, const cx<T> *tw, double norm,
                         double *partials, double *sig, hipStream_t s, hipEvent_t *ev, const Schedule &sc);

constexpr int PK_MAX_BINS = 2048;
template <typename T>
hipError_t launch_power_spectrum(int n, cx<T> *F, const cx<T> *tw, double pscale, double binw, int nbins,
                                 double *acc, hipStream_t s);
int fields_partials_count(int n, int prec);

// The map stage's tile range to the host: with it, launch_maps waits (on the host, while the maps are
// cleared) for the range k_tile_range found and launches exactly its workgroups; without it (a stream under
// capture) the launch covers every tile and the workgroups past the range exit on their first load.
struct RangeReadback {
  uint32_t *h_range;   // pinned [2]
  hipEvent_t ev;       // recorded after the copy of the range
};
template <typename T>
hipError_t launch_maps(const cx<T> *F, T *maps, size_t nmaps, const MapParams &m, hipStream_t s,
                       hipEvent_t *ev, const Schedule &sc, const RangeReadback *rb = nullptr);
template <typename T>
hipError_t launch_cell_field(const cx<T> *F, const MapParams &m, int which, size_t cell0, size_t ncell, double *out,
                             hipStream_t s);
template <typename T>
hipError_t launch_sample_indices(const cx<T> *F, const MapParams &m, int64_t cell0, int64_t ncell,
                                 int32_t *inu, int64_t *ipix, hipStream_t s);
// FP64 parity scatter of every sample: per-channel sums (tot[n_nu], cnt[n_nu]; cleared here) or whole maps
// (ref[n_nu * npix], cleared by the caller)
template <typename T>
hipError_t launch_channel_totals(const cx<T> *F, const MapParams &m, int64_t cell0, int64_t ncell, double *tot,
                                 unsigned long long *cnt, hipStream_t s);
template <typename T>
hipError_t launch_reference_maps(const cx<T> *F, const MapParams &m, double *ref, hipStream_t s);
// every sample of the box through the screened and the FP64 decisions: counts[0] samples, [1] channel
// decisions left to FP64, [2] pixel decisions left to FP64 (binned samples), [3] samples whose screened
// (channel, pixel) differ from the FP64 ones, [4] binned samples (counts cleared here)
template <typename T>
hipError_t launch_screen_check(const cx<T> *F, const MapParams &m, unsigned long long *counts, hipStream_t s);
template <typename T>
hipError_t launch_grid_get(const cx<T> *F, size_t ncell, int which, double *out, hipStream_t s);
template <typename T>
hipError_t launch_grid_set(cx<T> *F, size_t ncell, int which, const double *in, hipStream_t s);

}  // namespace crime
