// Field stage of the GetHI density-to-sky path on gfx950: create_d_and_vr_fields
// (src/fourier.c:234-305, 375-438) as three HBM passes over one complex N^3 grid.
//
// delta and the velocity potential phi are both real, so the two inverse real FFTs of the
// reference (fourier.c:391-392) are done as ONE complex inverse FFT of
//     F(k) = delta_k + i phi_k,   phi_k = delta_k f0 H0 / k^2            (fourier.c:297-298)
// over the Hermitian completion of the half spectrum: the output is delta in the real part and phi
// in the imaginary part of each cell.  Bytes moved equal those of two r2c-layout transforms, the
// map stage then reads (delta, phi) of a cell in one 8-byte load.  The kx = 0 and kx = N/2 planes,
// which the reference fills without Hermitian symmetry (fourier.c:275-299), enter through their
// Hermitian part (X(k) + X*(-k))/2, which is what a c2r transform uses of them.
//
//   pass 1  k_fill_z : generate F on the fly (P(k) amplitude x unit mode) and transform along kz
//   pass 2  k_fft_y  : transform along ky
//   pass 3  k_fft_x  : transform along kx (contiguous), scale by (sqrt(2 pi)/L)^3 (fourier.c:403),
//                      FP64 partial sums of delta and delta^2 (compute_sigma_dens, fourier.c:24-76)
//   k_sigma_finalize : fixed-order FP64 reduction of the partials -> sigma^2, <delta>, and the largest
//                      |phi| of the grid (bounds the map stage's velocities)
#include "gh_kernels.hpp"

namespace crime {

// ---------------------------------------------------------------------------------------------
// mode fill
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double pk_linear0(const FillParams &f, double lgk)
{
#pragma clang fp contract(off)
  // src/cosmo.c:153-170; the last half-open bin uses the k^-3 tail instead of reading past the table
  int ik = (int)((lgk - f.logkmin) * f.idlogk);
  if (ik < 0) return f.pk[0] * pow(10.0, f.n_s * (lgk - f.logkmin));
  if (ik < f.numk - 1) return f.pk[ik] + (lgk - f.logk[ik]) * (f.pk[ik + 1] - f.pk[ik]) * f.idlogk;
  return f.pk[f.numk - 1] * pow(10.0, -3 * (lgk - f.logkmax));
}

// Philox words of half-spectrum mode idx -> the two uniforms (53-bit) of rng_delta_gauss
__device__ __forceinline__ void mode_uniforms(const FillParams &f, size_t idx, double &u1, double &u2)
{
  u32x4 c;
  c.v[0] = (uint32_t)idx;
  c.v[1] = (uint32_t)(idx >> 32);
  c.v[2] = 0;
  c.v[3] = 0;
  u32x4 r = philox4x32_10(c, f.seed, PHILOX_KEY1);
  u1 = u53(r.v[0], r.v[1]);
  u2 = u53(r.v[2], r.v[3]);
}

// Unit-variance complex mode at half-spectrum position (kx in [0,N/2], ky, kz):
// rng_delta_gauss (src/common.c:154-164), phase 2 pi u1 and modulus sqrt(-ln(1-u2)).  FP64 grids
// evaluate it in FP64; FP32 grids evaluate the logarithm, root and sin/cos in FP32 from the exact
// FP64 value of 1-u2 (relative error ~1e-7 per mode, the precision the mode is stored in).
template <typename T, int N>
__device__ __forceinline__ cx<T> unit_mode(const FillParams &f, int kx, int ky, int kz)
{
  size_t idx = (size_t)kx + (size_t)(N / 2 + 1) * ((size_t)ky + (size_t)N * (size_t)kz);
  if (f.mode_source == 1) return {(T)f.unit_modes[2 * idx], (T)f.unit_modes[2 * idx + 1]};
  double u1, u2;
  mode_uniforms(f, idx, u1, u2);
  if constexpr (sizeof(T) == 8) {
    double mod = sqrt(-log(1.0 - u2));
    double s, co;
    sincos(2.0 * M_PI * u1, &s, &co);
    return {mod * co, mod * s};
  }
  else {
    float mod = sqrtf(-logf((float)(1.0 - u2)));
    float s, co;
    sincospif(2.0f * (float)u1, &s, &co);
    return {mod * co, mod * s};
  }
}

// Amplitude per k^2 = m dk^2: A = sqrt(P(k)/dk^3 [e^{-r_s^2 k^2}]) (fourier.c:292-295) and, for
// FP64 grids, also A psi with the velocity-potential factor psi = phi_scale f0 H0 / k^2
// (fourier.c:298).  FP32 grids keep A alone (4 bytes per entry, 3 MB at N=1024, L2-sized) and form
// psi = psi_c / m in registers.  Built once per handle.
template <typename T>
__global__ __launch_bounds__(256) void k_amp_table(FillParams f, AmpEntry<T> *__restrict__ tab, int mmax)
{
  int m = blockIdx.x * 256 + threadIdx.x;
  if (m > mmax) return;
  double amp = 0, psi = 0;
  if (m > 0) {
    double k2 = (double)m * (f.dk * f.dk);
    double sigma2 = pk_linear0(f, 0.5 * log10(k2)) * f.idk3;
    if (f.do_smoothing) sigma2 *= exp(-f.r2_smooth * k2);
    amp = sqrt(sigma2);
    psi = f.phi_scale * (f.f0h0 / k2);
  }
  if constexpr (sizeof(T) == 8) tab[m] = {amp, amp * psi};
  else tab[m] = (float)amp;
}

// F = delta + i phi_scale phi at full-grid position (kx, ky, kz), kx in [0, N) (N a compile-time
// constant: the mirror indices are masks or constant divisions, not runtime integer divisions)
// (A, A psi) of mode (kx, ky, kz): a function of |k|^2 only, so kz and N - kz share it
template <typename T, int N>
__device__ __forceinline__ cx<T> mode_amp(const FillParams &f, const AmpEntry<T> *__restrict__ amp, int kx, int ky,
                                          int kz)
{
  constexpr int n = N, h = N / 2;
  // |k| components with the reference's +Nyquist convention (fourier.c:264-283); only squares matter
  int ax = kx <= h ? kx : n - kx;
  int ay = ky <= h ? ky : n - ky;
  int az = kz <= h ? kz : n - kz;
  int m2 = ax * ax + ay * ay + az * az;
  cx<T> ab;
  if constexpr (sizeof(T) == 8) ab = amp[m2];
  else {
    // psi = psi_c / m2 through the hardware reciprocal (1 ulp): FP32 fields carry ~1e-7 relative error
    // anyway, and a correctly rounded division costs ten instructions per element
    float a = amp[m2];
    ab = {a, a * ((float)f.psi_c * __builtin_amdgcn_rcpf((float)m2))};
  }
  return ab;
}

// the element from its amplitude pair ab = mode_amp(kx, ky, kz)
template <typename T, int N>
__device__ __forceinline__ cx<T> fill_element_a(const FillParams &f, cx<T> ab, int kx, int ky, int kz)
{
  constexpr int n = N, h = N / 2;
  if ((kx | ky | kz) == 0) return {(T)0, (T)0};
  int myn = (n - ky) % n, mzn = (n - kz) % n;  // -k in y and z
  cx<T> g;
  if (kx > 0 && kx < h) {
    g = unit_mode<T, N>(f, kx, ky, kz);
  }
  else if (kx > h) {
    cx<T> m = unit_mode<T, N>(f, n - kx, myn, mzn);
    g = {m.x, -m.y};
  }
  else {  // kx == 0 or kx == N/2: Hermitian part of the plane
    cx<T> a = unit_mode<T, N>(f, kx, ky, kz), b = unit_mode<T, N>(f, kx, myn, mzn);
    g = {(T)0.5 * (a.x + b.x), (T)0.5 * (a.y - b.y)};
  }
  if (f.vmode == 0) {
    // delta = A g, phi_scale phi = A psi g; F = delta + i phi_scale phi = A g (1 + i psi)
    return {ab.x * g.x - ab.y * g.y, ab.x * g.y + ab.y * g.x};
  }
  // spectral velocities: with A psi = A phi_scale f0 H0 / k^2 and S_j = sin(k_j dx)/dx,
  //   grid 0: delta + i phi_scale vx = A g + i (i S_x A psi g) = (A - S_x A psi) g
  //   grid 1: phi_scale (vy + i vz)  = i S_y A psi g - S_z A psi g = A psi (-S_z + i S_y) g
  // S_j is odd in k_j, so each v_j(k) is Hermitian wherever delta_k is (the kx = 0, N/2 planes included).
  if (f.comp == 0) {
    const T c = ab.x - (T)f.sin_dx[kx] * ab.y;
    return {c * g.x, c * g.y};
  }
  const cx<T> c = {-(T)f.sin_dx[kz] * ab.y, (T)f.sin_dx[ky] * ab.y};
  return {c.x * g.x - c.y * g.y, c.x * g.y + c.y * g.x};
}

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------
template <typename T, int N, int V>
__global__ __launch_bounds__((FillCfg<T, N, V>::NT)) void k_fill_z(cx<T> *__restrict__ F, FillParams f,
                                                                 const cx<T> *__restrict__ tw)
{
  using C = FillCfg<T, N, V>;
  __shared__ cx<T> lds[N * C::TX];
  const int c = threadIdx.x % C::TX, t = threadIdx.x / C::TX;
  // Blocks go to the XCDs round-robin (XCD = linear block index mod 8), and N / TX is a multiple of 8,
  // so without a rotation every row's kx = 0 and kx = N/2 blocks -- whose waves run the one-draw and the
  // two-draw (Hermitian) branches one after the other, three draws in all -- would land on XCD 0.
  // Rotating the column block by the row spreads them over all eight.
  const int xb = (blockIdx.x + blockIdx.y) % gridDim.x;
  const int kx = xb * C::TX + c, ky = blockIdx.y;
  const size_t n2 = (size_t)N * N;
  cx<T> *col = F + (size_t)kx + (size_t)N * ky;
  // the generator is too large to unroll into the register tile: stage the modes in LDS first
  const AmpEntry<T> *amp = (const AmpEntry<T> *)f.amp;
  // kz = e and N - e in one step: one read of their shared amplitude (a scattered gather, a cache line per
  // lane) for two elements
  for (int e = t; e <= N / 2; e += C::TC) {
    const cx<T> ab = mode_amp<T, N>(f, amp, kx, ky, e);
    lds[e * C::TX + c] = fill_element_a<T, N>(f, ab, kx, ky, e);
    if (e > 0 && e < N / 2) lds[(N - e) * C::TX + c] = fill_element_a<T, N>(f, ab, kx, ky, N - e);
  }
  __syncthreads();
  auto load = [&](int e) { return lds[e * C::TX + c]; };
  auto store = [&](int e, cx<T> v) { col[(size_t)e * n2] = v; };
  auto slot = [&](int e) -> cx<T> & { return lds[e * C::TX + c]; };
  Stockham<T, N, C::TC, C::MAXR>::template run<true>(t, load, store, slot, tw);
}

template <typename T, int N, int V>
__global__ __launch_bounds__((ColCfg<T, N, V>::NT)) void k_fft_y(cx<T> *__restrict__ F,
                                                               const cx<T> *__restrict__ tw)
{
  using C = ColCfg<T, N, V>;
  __shared__ cx<T> lds[N * C::TX];
  const int c = threadIdx.x % C::TX, t = threadIdx.x / C::TX;
  // XCD-aware tile order: neighbouring column tiles of one plane run on one XCD, so tiles narrower
  // than a 128-byte line share each line through that XCD's L2 instead of fetching it twice
  const int b = xcd_swizzle((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
  const int kx = (b % (int)gridDim.x) * C::TX + c, kz = b / (int)gridDim.x;
  cx<T> *col = F + (size_t)kx + (size_t)N * N * kz;
  auto load = [&](int e) { return col[(size_t)e * N]; };
  auto store = [&](int e, cx<T> v) { col[(size_t)e * N] = v; };
  auto slot = [&](int e) -> cx<T> & { return lds[e * C::TX + c]; };
  Stockham<T, N, C::TC, C::MAXR>::run(t, load, store, slot, tw);
}

template <typename T, int N, int V>
__global__ __launch_bounds__((RowCfg<T, N, V>::NT)) void k_fft_x(cx<T> *__restrict__ F,
                                                               const cx<T> *__restrict__ tw,
                                                               double norm, double norm_phi,
                                                               double *__restrict__ partials)
{
  using C = RowCfg<T, N, V>;
  __shared__ cx<T> lds[N * C::TY];
  __shared__ double red[2 * C::NT];
  __shared__ unsigned long long imax;
  const int t = threadIdx.x % C::TC, b = threadIdx.x / C::TC;
  const size_t row = (size_t)blockIdx.x * C::TY + b;
  cx<T> *line = F + row * N;
  double s1 = 0, s2 = 0, am = 0;
  if (threadIdx.x == 0) imax = 0;
  __syncthreads();
  auto load = [&](int e) { return line[e]; };
  auto store = [&](int e, cx<T> v) {
    cx<T> o = {(T)((double)v.x * norm), (T)((double)v.y * norm_phi)};
    double d = (double)o.x;
    s1 += d;
    s2 += d * d;
    am = fmax(am, fabs((double)o.y));
    line[e] = o;
  };
  auto slot = [&](int e) -> cx<T> & { return lds[b * N + e]; };
  Stockham<T, N, C::TC, C::MAXR>::run(t, load, store, slot, tw);
  red[threadIdx.x] = s1;
  red[C::NT + threadIdx.x] = s2;
  // largest |imaginary part| of the block: non-negative doubles order as their bit patterns
  atomicMax(&imax, (unsigned long long)__double_as_longlong(am));
  __syncthreads();
  for (int w = C::NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      red[C::NT + threadIdx.x] += red[C::NT + threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = red[0];
    partials[2 * blockIdx.x + 1] = red[C::NT];
    partials[2 * gridDim.x + blockIdx.x] = __longlong_as_double((long long)imax);
  }
}

// ---------------------------------------------------------------------------------------------
// P(k) estimator: forward DFT of the resident grid as conj(IDFT(conj F)) -- kx pass on conj(F),
// then the ky and kz passes -- followed by shell binning of |delta_k|^2.
// ---------------------------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__((RowCfg<T, N>::NT)) void k_fft_x_conj(cx<T> *__restrict__ F,
                                                                 const cx<T> *__restrict__ tw)
{
  using C = RowCfg<T, N>;
  __shared__ cx<T> lds[N * C::TY];
  const int t = threadIdx.x % C::TC, b = threadIdx.x / C::TC;
  cx<T> *line = F + ((size_t)blockIdx.x * C::TY + b) * N;
  auto load = [&](int e) { return cconj(line[e]); };
  auto store = [&](int e, cx<T> v) { line[e] = v; };
  auto slot = [&](int e) -> cx<T> & { return lds[b * N + e]; };
  Stockham<T, N, C::TC>::run(t, load, store, slot, tw);
}

template <typename T, int N>
__global__ __launch_bounds__((ColCfg<T, N>::NT)) void k_fft_z(cx<T> *__restrict__ F, const cx<T> *__restrict__ tw)
{
  using C = ColCfg<T, N>;
  __shared__ cx<T> lds[N * C::TX];
  const int c = threadIdx.x % C::TX, t = threadIdx.x / C::TX;
  const int kx = blockIdx.x * C::TX + c, ky = blockIdx.y;
  const size_t n2 = (size_t)N * N;
  cx<T> *col = F + (size_t)kx + (size_t)N * ky;
  auto load = [&](int e) { return col[(size_t)e * n2]; };
  auto store = [&](int e, cx<T> v) { col[(size_t)e * n2] = v; };
  auto slot = [&](int e) -> cx<T> & { return lds[e * C::TX + c]; };
  Stockham<T, N, C::TC>::run(t, load, store, slot, tw);
}

// G = IDFT(conj F) = conj(DFT F).  delta real => DFT(delta)_k = (DFT F_k + conj(DFT F_{-k}))/2
// = (conj G_k + G_{-k})/2; P_est = |DFT(delta)_k|^2 L^3 / N^6 summed into |k|/dk shells of width binw.
// acc[3*i] += P, acc[3*i+1] += |k|/dk, acc[3*i+2] += 1 (every k of the full grid).
template <typename T>
__global__ __launch_bounds__(256) void k_pk_bin(const cx<T> *__restrict__ G, int n, double pscale, double binw,
                                                int nbins, double *__restrict__ acc)
{
  __shared__ double h[3 * PK_MAX_BINS];
  for (int i = threadIdx.x; i < 3 * nbins; i += 256) h[i] = 0;
  __syncthreads();
  const size_t nn = (size_t)n, n3 = nn * nn * nn;
  const int hf = n / 2;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n3; c += (size_t)gridDim.x * 256) {
    int kx = (int)(c % nn), ky = (int)((c / nn) % nn), kz = (int)(c / (nn * nn));
    int ax = kx <= hf ? kx : n - kx, ay = ky <= hf ? ky : n - ky, az = kz <= hf ? kz : n - kz;
    double kk = sqrt((double)(ax * ax + ay * ay + az * az));
    int bin = (int)(kk / binw);
    if (bin >= nbins) continue;
    size_t cm = (size_t)((n - kx) % n) + nn * ((size_t)((n - ky) % n) + nn * (size_t)((n - kz) % n));
    cx<T> g = G[c], gm = G[cm];
    double re = 0.5 * ((double)g.x + (double)gm.x), im = 0.5 * ((double)gm.y - (double)g.y);
    atomicAdd(&h[3 * bin], (re * re + im * im) * pscale);
    atomicAdd(&h[3 * bin + 1], kk);
    atomicAdd(&h[3 * bin + 2], 1.0);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * nbins; i += 256)
    if (h[i] != 0) atomicAdd(&acc[i], h[i]);
}

template <typename T, int N>
static hipError_t launch_pk_n(cx<T> *F, const cx<T> *tw, double pscale, double binw, int nbins, double *acc,
                              hipStream_t s)
{
  using CC = ColCfg<T, N>;
  using RC = RowCfg<T, N>;
  hipLaunchKernelGGL((k_fft_x_conj<T, N>), dim3((N * N) / RC::TY), dim3(RC::NT), 0, s, F, tw);
  hipLaunchKernelGGL((k_fft_y<T, N, 0>), dim3(N / CC::TX, N), dim3(CC::NT), 0, s, F, tw);
  hipLaunchKernelGGL((k_fft_z<T, N>), dim3(N / CC::TX, N), dim3(CC::NT), 0, s, F, tw);
  size_t n3 = (size_t)N * N * N;
  unsigned nb = (unsigned)((n3 + 255) / 256 < 8192 ? (n3 + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_pk_bin<T>, dim3(nb), dim3(256), 0, s, F, N, pscale, binw, nbins, acc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_power_spectrum(int n, cx<T> *F, const cx<T> *tw, double pscale, double binw, int nbins,
                                 double *acc, hipStream_t s)
{
  switch (n) {
#define CRIME_CASE(NN) \
  case NN: return launch_pk_n<T, NN>(F, tw, pscale, binw, nbins, acc, s);
    CRIME_FOR_EACH_N(CRIME_CASE)
#undef CRIME_CASE
  default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_power_spectrum<float>(int, cx<float> *, const cx<float> *, double, double, int,
                                                 double *, hipStream_t);
template hipError_t launch_power_spectrum<double>(int, cx<double> *, const cx<double> *, double, double, int,
                                                  double *, hipStream_t);

__global__ __launch_bounds__(1024) void k_sigma_finalize(const double *__restrict__ partials, int nparts,
                                                         double ncells, double *__restrict__ out)
{
  __shared__ double r1[1024], r2[1024], r3[1024];
  double a = 0, b = 0, c = 0;
  for (int i = threadIdx.x; i < nparts; i += 1024) {
    a += partials[2 * i];
    b += partials[2 * i + 1];
    c = fmax(c, partials[2 * nparts + i]);
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  r3[threadIdx.x] = c;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
      r3[threadIdx.x] = fmax(r3[threadIdx.x], r3[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double mean = r1[0] / ncells;
    out[0] = r2[0] / ncells - mean * mean;  // sigma2_gauss (fourier.c:74)
    out[1] = mean;
    out[2] = r3[0];  // max |imaginary part| of the grid (the map stage's bound on the velocity)
  }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
template <typename T, int N, int V>
static void launch_fill_v(cx<T> *F, const FillParams &f, const cx<T> *tw, hipStream_t s)
{
  using FC = FillCfg<T, N, V>;
  hipLaunchKernelGGL((k_fill_z<T, N, V>), dim3(N / FC::TX, N), dim3(FC::NT), 0, s, F, f, tw);
}

template <typename T, int N, int V>
static void launch_y_v(cx<T> *F, const cx<T> *tw, hipStream_t s)
{
  using CC = ColCfg<T, N, V>;
  hipLaunchKernelGGL((k_fft_y<T, N, V>), dim3(N / CC::TX, N), dim3(CC::NT), 0, s, F, tw);
}

template <typename T, int N, int V>
static int launch_x_v(cx<T> *F, const cx<T> *tw, double norm, double norm_phi, double *partials, hipStream_t s)
{
  using RC = RowCfg<T, N, V>;
  const int nrowblk = (N * N) / RC::TY;
  hipLaunchKernelGGL((k_fft_x<T, N, V>), dim3(nrowblk), dim3(RC::NT), 0, s, F, tw, norm, norm_phi, partials);
  return nrowblk;
}

template <typename T, int N>
static void fill_sched(cx<T> *F, const FillParams &f, const cx<T> *tw, hipStream_t s, const Schedule &sc)
{
  constexpr bool tuned = N >= TUNE_MIN_N;
  if (tuned && sc.fill == 1) launch_fill_v<T, N, tuned ? 1 : 0>(F, f, tw, s);
  else if (tuned && sc.fill == 2) launch_fill_v<T, N, tuned ? 2 : 0>(F, f, tw, s);
  else if (tuned && sc.fill == 3) launch_fill_v<T, N, tuned ? 3 : 0>(F, f, tw, s);
  else launch_fill_v<T, N, 0>(F, f, tw, s);
}

template <typename T, int N>
static void y_sched(cx<T> *F, const cx<T> *tw, hipStream_t s, const Schedule &sc)
{
  constexpr bool tuned = N >= TUNE_MIN_N;
  if (tuned && sc.y == 1) launch_y_v<T, N, tuned ? 1 : 0>(F, tw, s);
  else if (tuned && sc.y == 2) launch_y_v<T, N, tuned ? 2 : 0>(F, tw, s);
  else if (tuned && sc.y == 3) launch_y_v<T, N, tuned ? 3 : 0>(F, tw, s);
  else launch_y_v<T, N, 0>(F, tw, s);
}

template <typename T, int N>
static int x_sched(cx<T> *F, const cx<T> *tw, double nre, double nim, double *partials, hipStream_t s,
                   const Schedule &sc)
{
  constexpr bool tuned = N >= TUNE_MIN_N;
  if (tuned && sc.x == 1) return launch_x_v<T, N, tuned ? 1 : 0>(F, tw, nre, nim, partials, s);
  if (tuned && sc.x == 2) return launch_x_v<T, N, tuned ? 2 : 0>(F, tw, nre, nim, partials, s);
  if (tuned && sc.x == 3) return launch_x_v<T, N, tuned ? 3 : 0>(F, tw, nre, nim, partials, s);
  return launch_x_v<T, N, 0>(F, tw, nre, nim, partials, s);
}

// Stage timings (ev[0..4]) cover both grids of a spectral handle: the second grid's pass runs right
// before the same pass of F, and its sigma^2 partials are overwritten by F's kx pass.
template <typename T, int N>
static hipError_t launch_fields_n(cx<T> *F, cx<T> *G, const FillParams &f, const cx<T> *tw, double norm,
                                  double *partials, double *sig, hipStream_t s, hipEvent_t *ev,
                                  const Schedule &sc)
{
  const bool two = f.vmode != 0;
  FillParams f1 = f;
  f1.comp = 1;
  if (ev) (void)hipEventRecord(ev[0], s);
  if (two) fill_sched<T, N>(G, f1, tw, s, sc);
  fill_sched<T, N>(F, f, tw, s, sc);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (two) y_sched<T, N>(G, tw, s, sc);
  y_sched<T, N>(F, tw, s, sc);
  if (ev) (void)hipEventRecord(ev[2], s);
  const double nphi = norm / f.phi_scale;
  if (two) x_sched<T, N>(G, tw, nphi, nphi, partials, s, sc);
  const int nrowblk = x_sched<T, N>(F, tw, norm, nphi, partials, s, sc);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_sigma_finalize, dim3(1), dim3(1024), 0, s, partials, nrowblk,
                     (double)N * N * N, sig);
  if (ev) (void)hipEventRecord(ev[4], s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fields(int n, cx<T> *F, cx<T> *G, const FillParams &f, const cx<T> *tw, double norm,
                         double *partials, double *sig, hipStream_t s, hipEvent_t *ev, const Schedule &sc)
{
  if (f.vmode != 0 && (!G || !f.sin_dx)) return hipErrorInvalidValue;
  switch (n) {
#define CRIME_CASE(NN) \
  case NN: return launch_fields_n<T, NN>(F, G, f, tw, norm, partials, sig, s, ev, sc);
    CRIME_FOR_EACH_N(CRIME_CASE)
#undef CRIME_CASE
  default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_fields<float>(int, cx<float> *, cx<float> *, const FillParams &, const cx<float> *,
                                         double, double *, double *, hipStream_t, hipEvent_t *, const Schedule &);
template hipError_t launch_fields<double>(int, cx<double> *, cx<double> *, const FillParams &, const cx<double> *,
                                          double, double *, double *, hipStream_t, hipEvent_t *, const Schedule &);

template <typename T>
hipError_t launch_amp_table(const FillParams &f, void *tab, hipStream_t s)
{
  int mmax = amp_count(f.n) - 1;
  hipLaunchKernelGGL(k_amp_table<T>, dim3((mmax + 256) / 256), dim3(256), 0, s, f, (AmpEntry<T> *)tab, mmax);
  return hipGetLastError();
}
template hipError_t launch_amp_table<float>(const FillParams &, void *, hipStream_t);
template hipError_t launch_amp_table<double>(const FillParams &, void *, hipStream_t);

// the most row blocks any schedule variant launches (variant 1 has the fewest rows per block)
int fields_partials_count(int n, int prec)
{
  switch (n) {
#define CRIME_CASE(NN) \
  case NN: return prec ? (NN * NN) / RowCfg<double, NN, 1>::TY : (NN * NN) / RowCfg<float, NN, 1>::TY;
    CRIME_FOR_EACH_N(CRIME_CASE)
#undef CRIME_CASE
  default: return 0;
  }
}

}  // namespace crime
