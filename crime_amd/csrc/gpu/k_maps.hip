// Map stage of the GetHI density-to-sky path on gfx950: get_HI (src/grid_tools.c:103-153) fused
// with mk_T_maps (src/pixelize.c:303-414) and the radial-velocity stencil (src/fourier.c:307-373).
//
// One thread per cell reads (delta, phi) of the cell and phi of its six periodic neighbours, forms
// v_r, the lognormal HI mass and the RSD redshift shift in FP64, and scatters its n_sub samples into
// the (channel, RING pixel) maps with hardware FP32/FP64 atomics, already multiplied by the channel's
// mass-to-mK factor (the reference scales the finished maps, pixelize.c:399-414; the product of the
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

// centre, r and v_r (fourier.c:327-368); phi lives in the imaginary part of F
template <typename T>
__device__ __forceinline__ void cell_velocity(const MapParams &m, const cx<T> *__restrict__ F, int ix,
                                              int iy, int iz, CellState &s)
{
  const int n = m.n;
  s.x0 = m.dx * (ix + 0.5) - m.obs[0];
  s.y0 = m.dx * (iy + 0.5) - m.obs[1];
  s.z0 = m.dx * (iz + 0.5) - m.obs[2];
  int ixp = ix == n - 1 ? 0 : ix + 1, ixm = ix == 0 ? n - 1 : ix - 1;
  int iyp = iy == n - 1 ? 0 : iy + 1, iym = iy == 0 ? n - 1 : iy - 1;
  int izp = iz == n - 1 ? 0 : iz + 1, izm = iz == 0 ? n - 1 : iz - 1;
  const size_t nn = (size_t)n, n2 = nn * nn;
  const size_t row = nn * iy + n2 * iz;
  double px1 = (double)F[row + ixp].y, px0 = (double)F[row + ixm].y;
  double py1 = (double)F[ix + nn * iyp + n2 * iz].y, py0 = (double)F[ix + nn * iym + n2 * iz].y;
  double pz1 = (double)F[ix + nn * iy + n2 * izp].y, pz0 = (double)F[ix + nn * iy + n2 * izm].y;
  double idx = 1. / m.dx;
  double v0 = 0.5 * idx * (px1 - px0);
  double v1 = 0.5 * idx * (py1 - py0);
  double v2 = 0.5 * idx * (pz1 - pz0);
  double rr2 = s.x0 * s.x0 + s.y0 * s.y0 + s.z0 * s.z0;
  double irr = 1. / sqrt(rr2);
  s.r = sqrt(rr2);
  s.vr = v0 * (s.x0 * irr) + v1 * (s.y0 * irr) + v2 * (s.z0 * irr);
}

// lognormal HI mass and RSD shift (grid_tools.c:138-146)
__device__ __forceinline__ void cell_hi(const MapParams &m, double delta, double sigma2, CellState &s)
{
  double redshift = z_of_r(m.tab, s.r);
  double gfd = dgrowth_of_r(m.tab, s.r) * bias_of(m, redshift);
  s.dz = m.rsd ? s.vr * vgrowth_of_r(m.tab, s.r) : 0.0;
  double dens_ln = exp(gfd * (delta - 0.5 * gfd * sigma2));
  s.mass = m.mass_prefac * fraction_hi(redshift) * dens_ln;
}

// channel and pixel of one sample (pixelize.c:360-375); returns the channel or -1
__device__ __forceinline__ int sample_bin(const MapParams &m, const CellState &s, int isub, int64_t &ipix)
{
  double x = s.x0 + m.sub_off[3 * isub];
  double y = s.y0 + m.sub_off[3 * isub + 1];
  double z = s.z0 + m.sub_off[3 * isub + 2];
  double r = sqrt(x * x + y * y + z * z);
  double redshift = z_of_r(m.tab, r) + s.dz;
  double nu = 1420.40575177 / (1 + redshift);
  int inu = channel_of_guess(nu, m.n_nu, m.nu_mode, m.edges, m.nu_min, m.inv_dnu, m.guess, m.n_guess,
                             m.g_lo, m.g_inv);
  if (!(r > 0)) inu = -1;  // a sample on the observer has no direction
  ipix = inu >= 0 ? vec2pix_ring(m.nside, x, y, z, r) : -1;
  return inu;
}

// false when no sample of the cell can land in a channel (conservative)
__device__ __forceinline__ bool cell_may_bin(const MapParams &m, const CellState &s)
{
  double zlo = z_of_r(m.tab, s.r - m.h_sub) + s.dz;
  double zhi = z_of_r(m.tab, s.r + m.h_sub) + s.dz;
  if (!(1 + zlo > 0) || !(1 + zhi > 0)) return true;
  double nu_top = 1420.40575177 / (1 + zlo);
  double nu_bot = 1420.40575177 / (1 + zhi);
  if (nu_top * (1 + 1e-10) < m.nu_lo) return false;
  if (nu_bot * (1 - 1e-10) >= m.nu_hi) return false;
  return true;
}

// Add val to bin `key` of the workgroup's LDS table (open addressing, linear probing); a sample that
// finds no slot within the probe budget goes straight to the map in HBM.
template <typename T, int SLOTS>
__device__ __forceinline__ void agg_add(unsigned long long key, T val, unsigned long long *keys, T *vals,
                                        T *__restrict__ maps)
{
  const unsigned long long EMPTY = ~0ull;
  constexpr int SHIFT = 64 - __builtin_ctz(SLOTS);  // Fibonacci hashing onto the table
  unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> SHIFT) & (SLOTS - 1);
  for (int p = 0; p < 64; p++) {
    unsigned long long k = keys[h];
    if (k == EMPTY) {
      unsigned long long old = atomicCAS(&keys[h], EMPTY, key);
      if (old == EMPTY || old == key) {
        atomicAdd(&vals[h], val);
        return;
      }
    }
    else if (k == key) {
      atomicAdd(&vals[h], val);
      return;
    }
    h = (h + 1) & (SLOTS - 1);
  }
  unsafeAtomicAdd(maps + key, val);
}

// One workgroup = one TILE_X x TILE_Y x TILE_Z block of cells (one cell per thread).  Samples are
// summed per (channel, pixel) bin in LDS and every occupied bin is added to HBM once, so the
// global atomics scale with the bins a tile touches instead of with its samples.  SLOTS (the
// schedule variant: 2048, 1024 or 4096 bins) trades occupancy against probe length and overflow.
template <typename T, int SLOTS>
__global__ __launch_bounds__(256) void k_pixelize(const cx<T> *__restrict__ F, T *__restrict__ maps,
                                                  MapParams m)
{
  __shared__ unsigned long long keys[SLOTS];
  __shared__ T vals[SLOTS];
  const int n = m.n;
  const int t = threadIdx.x;
  // tiles run in order of distance from the observer, so the workgroups resident at one time cover a
  // thin shell and their atomics fall on a few channels of the maps.  2-D grid: a dispatch holds at
  // most 2^32 work-items per dimension.
  const size_t bid = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (bid >= (size_t)m.ntx * m.nty * m.ntz) return;
  const size_t b = m.tile_order[bid];
  const int tx = (int)(b % m.ntx), ty = (int)((b / m.ntx) % m.nty), tz = (int)(b / ((size_t)m.ntx * m.nty));
  const int ix = tx * TILE_X + (t % TILE_X), iy = ty * TILE_Y + ((t / TILE_X) % TILE_Y),
            iz = tz * TILE_Z + t / (TILE_X * TILE_Y);
  CellState s;
  bool active = ix < n && iy < n && iz < n;
  if (active) {
    cell_velocity<T>(m, F, ix, iy, iz, s);
    s.dz = m.rsd ? s.vr * vgrowth_of_r(m.tab, s.r) : 0.0;
    active = cell_may_bin(m, s);
  }
  if (!__syncthreads_or(active)) return;  // no cell of the tile reaches a channel
  for (int i = t; i < SLOTS; i += 256) {
    keys[i] = ~0ull;
    vals[i] = (T)0;
  }
  __syncthreads();
  if (active) {
    size_t cell = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    double mass_sub = s.mass / m.n_sub;
    for (int isub = 0; isub < m.n_sub; isub++) {
      int64_t ipix;
      int inu = sample_bin(m, s, isub, ipix);
      if (inu >= 0 && (uint64_t)ipix < (uint64_t)m.npix)
        agg_add<T, SLOTS>((unsigned long long)ipix + (unsigned long long)m.npix * inu,
                          (T)(mass_sub * m.prefac[inu]), keys, vals, maps);
    }
  }
  __syncthreads();
  for (int i = t; i < SLOTS; i += 256) {
    unsigned long long k = keys[i];
    if (k != ~0ull) unsafeAtomicAdd(maps + k, vals[i]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(T *__restrict__ p, size_t n)
{
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n; i += stride) p[i] = (T)0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_cell_field(const cx<T> *__restrict__ F, MapParams m, size_t ncell,
                                                    int which, double *__restrict__ out)
{
  for (size_t cell = (size_t)blockIdx.x * 256 + threadIdx.x; cell < ncell; cell += (size_t)gridDim.x * 256) {
    int ix, iy, iz;
    cell_coords<T>(m, cell, ix, iy, iz);
    CellState s;
    cell_velocity<T>(m, F, ix, iy, iz, s);
    cell_hi(m, (double)F[cell].x, m.sig[0], s);
    out[cell] = which == 2 ? s.vr : (which == 3 ? s.mass : s.dz);
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
    for (int isub = 0; isub < m.n_sub; isub++) {
      int64_t ipix;
      int inu = sample_bin(m, s, isub, ipix);
      inu_out[c * m.n_sub + isub] = inu;
      ipix_out[c * m.n_sub + isub] = ipix;
    }
  }
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

template <typename T>
hipError_t launch_maps(const cx<T> *F, T *maps, size_t nmaps, const MapParams &m, hipStream_t s,
                       hipEvent_t *ev, const Schedule &sc)
{
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_zero<T>, dim3(nblk(nmaps)), dim3(256), 0, s, maps, nmaps);
  if (ev) (void)hipEventRecord(ev[1], s);
  size_t ntiles = (size_t)m.ntx * m.nty * m.ntz;
  unsigned gx = (unsigned)(ntiles < 65536 ? ntiles : 65536);
  unsigned gy = (unsigned)((ntiles + gx - 1) / gx);
  if (sc.pix == 1) hipLaunchKernelGGL((k_pixelize<T, AGG_SLOTS / 2>), dim3(gx, gy), dim3(256), 0, s, F, maps, m);
  else if (sc.pix == 2) hipLaunchKernelGGL((k_pixelize<T, AGG_SLOTS * 2>), dim3(gx, gy), dim3(256), 0, s, F, maps, m);
  else hipLaunchKernelGGL((k_pixelize<T, AGG_SLOTS>), dim3(gx, gy), dim3(256), 0, s, F, maps, m);
  if (ev) (void)hipEventRecord(ev[2], s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cell_field(const cx<T> *F, const MapParams &m, int which, double *out, hipStream_t s)
{
  size_t ncell = (size_t)m.n * m.n * m.n;
  hipLaunchKernelGGL(k_cell_field<T>, dim3(nblk(ncell)), dim3(256), 0, s, F, m, ncell, which, out);
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

#define CRIME_INST(T)                                                                                   \
  template hipError_t launch_maps<T>(const cx<T> *, T *, size_t, const MapParams &, hipStream_t,      \
                                     hipEvent_t *, const Schedule &);                                   \
  template hipError_t launch_cell_field<T>(const cx<T> *, const MapParams &, int, double *, hipStream_t); \
  template hipError_t launch_sample_indices<T>(const cx<T> *, const MapParams &, int64_t, int64_t,     \
                                               int32_t *, int64_t *, hipStream_t);                      \
  template hipError_t launch_grid_get<T>(const cx<T> *, size_t, int, double *, hipStream_t);          \
  template hipError_t launch_grid_set<T>(cx<T> *, size_t, int, const double *, hipStream_t);
CRIME_INST(float)
CRIME_INST(double)
#undef CRIME_INST

}  // namespace crime
