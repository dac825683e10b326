// The C_l emulator's network (include/crime_mlp.h): BF16 MFMA forward/backward GEMMs with fused
// epilogues, FP32 edge layers, deterministic split-K weight gradients reduced inside Adam, and one
// training step captured as a HIP graph (step counter, learning rate and minibatch RNG live on the
// device, so the graph replays unchanged).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "crime_mlp.h"
#include "gh_device.hpp"

using namespace crime;

namespace {

constexpr int NIN = CRIME_MLP_IN, WD = CRIME_MLP_WIDTH, NH = CRIME_MLP_HIDDEN;
constexpr int NMID = NH - 1;                          // square WD x WD layers
constexpr int64_t OFF_W0 = 0, OFF_MID = OFF_W0 + (int64_t)NIN * WD + WD;
constexpr int64_t MID_STRIDE = (int64_t)WD * WD + WD; // W_k then b_k
constexpr int64_t OFF_W5 = OFF_MID + NMID * MID_STRIDE;
constexpr int64_t NPARAM = OFF_W5 + WD + 1;
constexpr int KSPLIT = 256;                           // batch rows per split of a weight-gradient GEMM
constexpr int RROWS = 32;                             // batch rows per partial of the column reductions
constexpr uint32_t MLP_KEY1 = 0x4D4C5031u;            // second Philox key word of the minibatch draws

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// per-handle device state read by the captured step
struct DevState {
  int step;        // Adam updates applied since set_params
  int local;       // step index within the current train call (loss slot)
  int total;       // steps of the cosine schedule
  float lr0, b1, b2, eps;
  uint32_t seed;
  int64_t ndata;
};

__device__ __forceinline__ uint16_t f2bf(float x)
{
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// ---------------------------------------------------------------------------------------------
// BF16 GEMM C[M x N] = A[M x K] B[K x N] with FP32 accumulation; TA: A is stored K x M, TB: B is
// stored N x K.  64 x 64 output tile per workgroup (4 waves of 2 x 2 v_mfma_f32_16x16x32_bf16),
// K in chunks of 32 through LDS, split-K over blockIdx.z (rows [z kchunk, (z+1) kchunk) of K).
// MFMA operands: lane = i + 16 (k / 4) holds A[i][k] (element k % 4), lane = j + 16 (k / 4) holds
// B[k][j]; D[i][j] is element i % 4 of lane j + 16 (i / 4).  N and K are multiples of 64 and 32.
// ---------------------------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 32, LP = TK + 8;

template <bool TA, bool TB, class Epi>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
                                              int ldb, int M, int K, int kchunk, Epi epi)
{
  __shared__ __align__(16) uint16_t As[TM * LP];
  __shared__ __align__(16) uint16_t Bs[TN * LP];   // Bs[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  const int kb = blockIdx.z * kchunk;
  const int ke = K < kb + kchunk ? K : kb + kchunk;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  f4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += TK) {
    if (!TA) {
      const int m = tid >> 2, k8 = (tid & 3) * 8;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (m0 + m < M) v = *(const uint4 *)(A + (size_t)(m0 + m) * lda + k0 + k8);
      *(uint4 *)(As + m * LP + k8) = v;
    } else {
      const int k = tid >> 3, m8 = (tid & 7) * 8;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (m0 + m8 < M) v = *(const uint4 *)(A + (size_t)(k0 + k) * lda + m0 + m8);
      const uint16_t *e = (const uint16_t *)&v;
      for (int i = 0; i < 8; i++) As[(m8 + i) * LP + k] = e[i];
    }
    if (!TB) {
      const int k = tid >> 3, n8 = (tid & 7) * 8;
      uint4 v = *(const uint4 *)(B + (size_t)(k0 + k) * ldb + n0 + n8);
      const uint16_t *e = (const uint16_t *)&v;
      for (int i = 0; i < 8; i++) Bs[(n8 + i) * LP + k] = e[i];
    } else {
      const int n = tid >> 2, k8 = (tid & 3) * 8;
      *(uint4 *)(Bs + n * LP + k8) = *(const uint4 *)(B + (size_t)(n0 + n) * ldb + k0 + k8);
    }
    __syncthreads();
    // gfx950's v_mfma_f32_16x16x32_bf16: the whole 32-deep k tile in one instruction per 16 x 16 block.
    // Lane l supplies A[row l % 16][k 8 (l / 16) .. +7] and B[k 8 (l / 16) .. +7][column l % 16].
    {
      bf16x8 af[2], bfr[2];
      for (int i = 0; i < 2; i++)
        af[i] = __builtin_bit_cast(bf16x8, *(const uint4 *)(As + (wr + 16 * i + (lane & 15)) * LP + 8 * (lane >> 4)));
      for (int j = 0; j < 2; j++)
        bfr[j] = __builtin_bit_cast(bf16x8, *(const uint4 *)(Bs + (wc + 16 * j + (lane & 15)) * LP + 8 * (lane >> 4)));
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 4; v++) {
        const int row = m0 + wr + 16 * i + 4 * (lane >> 4) + v;
        const int col = n0 + wc + 16 * j + (lane & 15);
        if (row < M) epi(row, col, acc[i][j][v], (int)blockIdx.z);
      }
}

// forward: z = acc + b; keep z (FP32) for the backward pass, h = bf16(silu(z))
struct EpiFwd {
  const float *bias;
  float *Z;
  uint16_t *H;
  __device__ void operator()(int r, int c, float v, int) const
  {
    const float z = v + bias[c];
    Z[(size_t)r * WD + c] = z;
    H[(size_t)r * WD + c] = f2bf(z * sigm(z));
  }
};

// backward through the layer below: dz_prev = bf16(acc silu'(z_prev)), silu' = s (1 + z (1 - s))
struct EpiBwd {
  const float *Zp;
  uint16_t *dZp;
  __device__ void operator()(int r, int c, float v, int) const
  {
    const float z = Zp[(size_t)r * WD + c], s = sigm(z);
    dZp[(size_t)r * WD + c] = f2bf(v * s * (1.0f + z * (1.0f - s)));
  }
};

// weight gradient: one split's partial sum
struct EpiPart {
  float *part;
  __device__ void operator()(int r, int c, float v, int z) const { part[((size_t)z * WD + r) * WD + c] = v; }
};

// ---------------------------------------------------------------------------------------------
// FP32 edge layers, minibatch gather, reductions, Adam
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_in_fwd(const float *__restrict__ X, const float *__restrict__ p, int n,
                                                float *__restrict__ Z1, uint16_t *__restrict__ H1)
{
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int r = (int)(idx / WD), c = (int)(idx % WD);
  if (r >= n) return;
  float z = p[OFF_W0 + (int64_t)NIN * WD + c];
#pragma unroll
  for (int i = 0; i < NIN; i++) z = fmaf(X[(size_t)r * NIN + i], p[OFF_W0 + (int64_t)i * WD + c], z);
  Z1[(size_t)r * WD + c] = z;
  H1[(size_t)r * WD + c] = f2bf(z * sigm(z));
}

// one wave per row: y = h5 . w5 + b5; when training also dy = 2 w (y - t) / n, the loss and dz5
__global__ __launch_bounds__(256) void k_out(const uint16_t *__restrict__ H5, const float *__restrict__ Z5,
                                             const float *__restrict__ p, const float *__restrict__ T,
                                             const float *__restrict__ Wt, int n,
                                             int train, float *__restrict__ Y, float *__restrict__ dy,
                                             uint16_t *__restrict__ dZ5, const DevState *__restrict__ st,
                                             float *__restrict__ loss_hist)
{
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const float *w5 = p + OFF_W5;
  float s = 0.f;
  for (int c = lane; c < WD; c += 64) s = fmaf(bf2f(H5[(size_t)r * WD + c]), w5[c], s);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const float y = s + w5[WD];
  if (!train) {
    if (lane == 0) Y[r] = y;
    return;
  }
  const float e = y - T[r], wt = Wt[r];
  const float d = 2.0f * wt * e / (float)n;
  if (lane == 0) {
    Y[r] = y;
    dy[r] = d;
    atomicAdd(loss_hist + st->local, wt * e * e / (float)n);
  }
  for (int c = lane; c < WD; c += 64) {
    const float z = Z5[(size_t)r * WD + c], sg = sigm(z);
    dZ5[(size_t)r * WD + c] = f2bf(d * w5[c] * sg * (1.0f + z * (1.0f - sg)));
  }
}

// partials of dw5 = sum_r h5 dy and db5 = sum_r dy over rows [z RROWS, (z+1) RROWS)
__global__ __launch_bounds__(256) void k_out_wgrad(const uint16_t *__restrict__ H5, const float *__restrict__ dy, int n,
                                                   float *__restrict__ part)
{
  const int c = threadIdx.x, z = blockIdx.x;
  const int r0 = z * RROWS, r1 = n < r0 + RROWS ? n : r0 + RROWS;
  float a = 0.f, b = 0.f;
  for (int r = r0; r < r1; r++) {
    const float d = dy[r];
    a = fmaf(bf2f(H5[(size_t)r * WD + c]), d, a);
    b += d;
  }
  part[(size_t)z * (WD + 1) + c] = a;
  if (c == 0) part[(size_t)z * (WD + 1) + WD] = b;
}

// partials of db = sum_r dz (column sums) over one split
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *__restrict__ dZ, int n, float *__restrict__ part)
{
  const int c = threadIdx.x, z = blockIdx.x;
  const int r0 = z * RROWS, r1 = n < r0 + RROWS ? n : r0 + RROWS;
  float a = 0.f;
  for (int r = r0; r < r1; r++) a += bf2f(dZ[(size_t)r * WD + c]);
  part[(size_t)z * WD + c] = a;
}

// partials of dW0 = x^T dz1 and db0 = sum_r dz1: part[z][i][c] (i < 8), part[z][8][c]
__global__ __launch_bounds__(256) void k_in_wgrad(const float *__restrict__ X, const uint16_t *__restrict__ dZ1, int n,
                                                  float *__restrict__ part)
{
  const int c = threadIdx.x, z = blockIdx.x;
  const int r0 = z * RROWS, r1 = n < r0 + RROWS ? n : r0 + RROWS;
  float a[NIN + 1] = {};
  for (int r = r0; r < r1; r++) {
    const float g = bf2f(dZ1[(size_t)r * WD + c]);
#pragma unroll
    for (int i = 0; i < NIN; i++) a[i] = fmaf(X[(size_t)r * NIN + i], g, a[i]);
    a[NIN] += g;
  }
  for (int i = 0; i <= NIN; i++) part[((size_t)z * (NIN + 1) + i) * WD + c] = a[i];
}

__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ Xall, const float *__restrict__ Tall,
                                                const float *__restrict__ Wall, int n, const DevState *__restrict__ st,
                                                float *__restrict__ X, float *__restrict__ T, float *__restrict__ Wt)
{
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  u32x4 c = {{(uint32_t)st->step, (uint32_t)b, 0u, 0u}};
  u32x4 r = philox4x32_10(c, st->seed, MLP_KEY1);
  const uint64_t u = ((uint64_t)r.v[0] << 32) | r.v[1];
  const int64_t i = (int64_t)(u % (uint64_t)st->ndata);
  for (int k = 0; k < NIN; k++) X[(size_t)b * NIN + k] = Xall[(size_t)i * NIN + k];
  T[b] = Tall[i];
  Wt[b] = Wall[i];
}

// Adam (Kingma & Ba 2015) on one contiguous parameter segment; g = sum of nsplit partials
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                              uint16_t *__restrict__ pbf, const float *__restrict__ part, int nsplit,
                                              int64_t count, const DevState *__restrict__ st)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float g = 0.f;
  for (int z = 0; z < nsplit; z++) g += part[(size_t)z * count + i];
  const int s = st->step;
  const float frac = st->total > 0 ? fminf((float)s / (float)st->total, 1.0f) : 0.0f;
  const float lr = st->lr0 * 0.5f * (1.0f + cosf(3.14159265358979f * frac));
  const float b1 = st->b1, b2 = st->b2;
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)(s + 1))), vh = vi / (1.0f - powf(b2, (float)(s + 1)));
  const float pi = p[i] - lr * mh / (sqrtf(vh) + st->eps);
  p[i] = pi;
  if (pbf) pbf[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void k_grad_out(const float *__restrict__ part, int nsplit, int64_t count,
                                                  float *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float g = 0.f;
  for (int z = 0; z < nsplit; z++) g += part[(size_t)z * count + i];
  out[i] = g;
}

__global__ void k_step_done(DevState *st)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->step += 1;
    st->local += 1;
  }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ p, int64_t count, uint16_t *__restrict__ o)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) o[i] = f2bf(p[i]);
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct crime_mlp {
  int max_batch = 0, device = 0;
  hipStream_t stream = nullptr;
  float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t *wbf = nullptr;                  // [NMID][WD][WD]
  float *X = nullptr, *T = nullptr, *Wt = nullptr, *Y = nullptr, *dy = nullptr;
  float *Z = nullptr;                       // [NH][max_batch][WD]
  uint16_t *H = nullptr, *dZ = nullptr;     // [NH][max_batch][WD]
  float *part_in = nullptr, *part_out = nullptr, *part_w = nullptr, *part_b = nullptr;
  DevState *st = nullptr;
  float *loss = nullptr;
  int loss_cap = 0;
  float *Xall = nullptr, *Tall = nullptr, *Wall = nullptr;
  int64_t ndata = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_batch = 0;
  int use_graph = 1;                        // 0: every step enqueued directly (crime_mlp_set_graph)
  std::string err;
};

static thread_local std::string g_mlp_err;

static int mlp_error(crime_mlp *h, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  else g_mlp_err = buf;
  return code;
}

#define MLP_HIP(h, call)                                                                          \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      (void)hipGetLastError(); /* reported here: a later launch check must not see it again */        \
      return mlp_error(h, e_ == hipErrorOutOfMemory ? CRIME_GH_ENOMEM : CRIME_GH_EHIP, "%s: %s", #call,     \
                       hipGetErrorString(e_));                                                        \
    }                                                                                                 \
  } while (0)

extern "C" int64_t crime_mlp_param_count(void) { return NPARAM; }
extern "C" const char *crime_mlp_last_error(const crime_mlp_t *h) { return h ? h->err.c_str() : g_mlp_err.c_str(); }

static void drop_graph(crime_mlp *h)
{
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->exec = nullptr;
  h->graph = nullptr;
  h->graph_batch = 0;
}

extern "C" void crime_mlp_destroy(crime_mlp_t *h)
{
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  void *bufs[] = {h->p, h->m, h->v, h->grad, h->wbf, h->X, h->T, h->Wt, h->Y, h->dy, h->Z, h->H, h->dZ,
                  h->part_in, h->part_out, h->part_w, h->part_b, h->st, h->loss, h->Xall, h->Tall, h->Wall};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static int nsplit_of(int n) { return (n + KSPLIT - 1) / KSPLIT; }
static int nred_of(int n) { return (n + RROWS - 1) / RROWS; }

static int create_impl(crime_mlp *h, int max_batch, int device)
{
  if (max_batch < 64 || max_batch > 65536 || max_batch % 64)
    return mlp_error(h, CRIME_GH_EINVAL, "max_batch %d: a multiple of 64 in [64, 65536]", max_batch);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mlp_error(h, CRIME_GH_ENODEV, "no HIP device");
  if (device < 0 || device >= ndev) return mlp_error(h, CRIME_GH_ENODEV, "device %d of %d", device, ndev);
  h->max_batch = max_batch;
  h->device = device;
  MLP_HIP(h, hipSetDevice(device));
  MLP_HIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const size_t B = (size_t)max_batch, act = (size_t)NH * B * WD;
  const int ns = nsplit_of(max_batch), nr = nred_of(max_batch);
  MLP_HIP(h, hipMalloc((void **)&h->p, NPARAM * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->m, NPARAM * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->v, NPARAM * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->grad, NPARAM * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->wbf, (size_t)NMID * WD * WD * sizeof(uint16_t)));
  MLP_HIP(h, hipMalloc((void **)&h->X, B * NIN * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->T, B * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->Wt, B * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->Y, B * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->dy, B * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->Z, act * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->H, act * sizeof(uint16_t)));
  MLP_HIP(h, hipMalloc((void **)&h->dZ, act * sizeof(uint16_t)));
  MLP_HIP(h, hipMalloc((void **)&h->part_in, (size_t)nr * (NIN + 1) * WD * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->part_out, (size_t)nr * (WD + 1) * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->part_w, (size_t)NMID * ns * WD * WD * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->part_b, (size_t)NMID * nr * WD * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->st, sizeof(DevState)));
  DevState s{};
  MLP_HIP(h, hipMemcpy(h->st, &s, sizeof(s), hipMemcpyHostToDevice));
  std::vector<float> zero(NPARAM, 0.0f);
  MLP_HIP(h, hipMemcpy(h->p, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->m, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->v, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int crime_mlp_create(int max_batch, int device, crime_mlp_t **out)
{
  if (!out) return mlp_error(nullptr, CRIME_GH_EINVAL, "out is NULL");
  *out = nullptr;
  crime_mlp *h = new crime_mlp;
  int rc = create_impl(h, max_batch, device);
  if (rc) {
    g_mlp_err = h->err;
    crime_mlp_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

extern "C" int crime_mlp_set_params(crime_mlp_t *h, const float *params)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!params) return mlp_error(h, CRIME_GH_EINVAL, "params are NULL");
  MLP_HIP(h, hipSetDevice(h->device));
  MLP_HIP(h, hipMemcpy(h->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> zero(NPARAM, 0.0f);
  MLP_HIP(h, hipMemcpy(h->m, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->v, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  DevState s{};
  MLP_HIP(h, hipMemcpy(h->st, &s, sizeof(s), hipMemcpyHostToDevice));
  for (int k = 0; k < NMID; k++)
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((int64_t)WD * WD)), dim3(256), 0, h->stream, h->p + OFF_MID + k * MID_STRIDE,
                       (int64_t)WD * WD, h->wbf + (size_t)k * WD * WD);
  MLP_HIP(h, hipGetLastError());
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int crime_mlp_get_params(crime_mlp_t *h, float *params)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!params) return mlp_error(h, CRIME_GH_EINVAL, "params are NULL");
  MLP_HIP(h, hipSetDevice(h->device));
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  MLP_HIP(h, hipMemcpy(params, h->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// forward over rows [0, n) of h->X
// forward for rows [0, n) of X (the handle's own input buffer unless given) on stream s (its own unless given)
static void enqueue_forward(crime_mlp *h, int n, int train, hipStream_t s = nullptr, const float *X = nullptr)
{
  const size_t B = (size_t)h->max_batch;
  if (!s) s = h->stream;
  if (!X) X = h->X;
  hipLaunchKernelGGL(k_in_fwd, dim3(nblk((int64_t)n * WD)), dim3(256), 0, s, X, h->p, n, h->Z, h->H);
  const dim3 g((n + TM - 1) / TM, WD / TN, 1);
  for (int k = 0; k < NMID; k++) {
    EpiFwd e{h->p + OFF_MID + k * MID_STRIDE + (int64_t)WD * WD, h->Z + (size_t)(k + 1) * B * WD,
             h->H + (size_t)(k + 1) * B * WD};
    hipLaunchKernelGGL((k_gemm<false, false, EpiFwd>), g, dim3(256), 0, s, h->H + (size_t)k * B * WD, WD,
                       h->wbf + (size_t)k * WD * WD, WD, n, WD, WD, e);
  }
  hipLaunchKernelGGL(k_out, dim3((n + 3) / 4), dim3(256), 0, s, h->H + (size_t)(NH - 1) * B * WD,
                     h->Z + (size_t)(NH - 1) * B * WD, h->p, h->T, h->Wt, n, train, h->Y, h->dy,
                     h->dZ + (size_t)(NH - 1) * B * WD, h->st, h->loss);
}

// backward for rows [0, n) (n % 64 == 0) into the split partials
static void enqueue_backward(crime_mlp *h, int n)
{
  const size_t B = (size_t)h->max_batch;
  const int ns = nsplit_of(n), nr = nred_of(n);
  hipLaunchKernelGGL(k_out_wgrad, dim3(nr), dim3(256), 0, h->stream, h->H + (size_t)(NH - 1) * B * WD, h->dy, n,
                     h->part_out);
  for (int k = NMID - 1; k >= 0; k--) {
    const uint16_t *Hin = h->H + (size_t)k * B * WD;         // input of square layer k
    const uint16_t *dZo = h->dZ + (size_t)(k + 1) * B * WD;  // gradient at its output
    EpiPart ep{h->part_w + (size_t)k * nsplit_of(h->max_batch) * WD * WD};
    hipLaunchKernelGGL((k_gemm<true, false, EpiPart>), dim3(WD / TM, WD / TN, ns), dim3(256), 0, h->stream, Hin, WD,
                       dZo, WD, WD, n, KSPLIT, ep);
    hipLaunchKernelGGL(k_colsum, dim3(nr), dim3(256), 0, h->stream, dZo, n,
                       h->part_b + (size_t)k * nred_of(h->max_batch) * WD);
    EpiBwd eb{h->Z + (size_t)k * B * WD, h->dZ + (size_t)k * B * WD};
    hipLaunchKernelGGL((k_gemm<false, true, EpiBwd>), dim3((n + TM - 1) / TM, WD / TN, 1), dim3(256), 0, h->stream,
                       dZo, WD, h->wbf + (size_t)k * WD * WD, WD, n, WD, WD, eb);
  }
  hipLaunchKernelGGL(k_in_wgrad, dim3(nr), dim3(256), 0, h->stream, h->X, h->dZ, n, h->part_in);
}

// segments of the parameter vector and their partials (GEMM splits or row-chunk reductions)
struct Seg {
  int64_t off, count;
  const float *part;
  bool gemm;
  int k;   // square layer whose BF16 copy follows the update, or -1
};

static std::vector<Seg> segments(crime_mlp *h)
{
  const size_t nsm = (size_t)nsplit_of(h->max_batch), nrm = (size_t)nred_of(h->max_batch);
  std::vector<Seg> s;
  s.push_back({OFF_W0, (int64_t)(NIN + 1) * WD, h->part_in, false, -1});
  for (int k = 0; k < NMID; k++) {
    s.push_back({OFF_MID + k * MID_STRIDE, (int64_t)WD * WD, h->part_w + k * nsm * WD * WD, true, k});
    s.push_back({OFF_MID + k * MID_STRIDE + (int64_t)WD * WD, WD, h->part_b + k * nrm * WD, false, -1});
  }
  s.push_back({OFF_W5, WD + 1, h->part_out, false, -1});
  return s;
}

static void enqueue_adam(crime_mlp *h, int n)
{
  for (const Seg &g : segments(h))
    hipLaunchKernelGGL(k_adam, dim3(nblk(g.count)), dim3(256), 0, h->stream, h->p + g.off, h->m + g.off, h->v + g.off,
                       g.k >= 0 ? h->wbf + (size_t)g.k * WD * WD : (uint16_t *)nullptr, g.part,
                       g.gemm ? nsplit_of(n) : nred_of(n), g.count, h->st);
  hipLaunchKernelGGL(k_step_done, dim3(1), dim3(64), 0, h->stream, h->st);
}

static int ensure_loss(crime_mlp *h, int n)
{
  if (n <= h->loss_cap) return 0;
  drop_graph(h);  // the captured step holds the old buffer
  if (h->loss) MLP_HIP(h, hipFree(h->loss));
  h->loss = nullptr;
  h->loss_cap = 0;
  MLP_HIP(h, hipMalloc((void **)&h->loss, (size_t)n * sizeof(float)));
  h->loss_cap = n;
  return 0;
}

extern "C" int crime_mlp_forward(crime_mlp_t *h, int64_t n, const float *x, float *y)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (n < 1 || !x || !y) return mlp_error(h, CRIME_GH_EINVAL, "need n >= 1 and non-NULL x, y");
  MLP_HIP(h, hipSetDevice(h->device));
  int rc = ensure_loss(h, 1);
  if (rc) return rc;
  for (int64_t i0 = 0; i0 < n; i0 += h->max_batch) {
    const int c = (int)(n - i0 < h->max_batch ? n - i0 : h->max_batch);
    MLP_HIP(h, hipMemcpyAsync(h->X, x + i0 * NIN, (size_t)c * NIN * sizeof(float), hipMemcpyHostToDevice, h->stream));
    enqueue_forward(h, c, 0);
    MLP_HIP(h, hipGetLastError());
    MLP_HIP(h, hipMemcpyAsync(y + i0, h->Y, (size_t)c * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  }
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  return 0;
}

// acc[i] (+)= w y[i]: the ensemble mean of crime_mlp_forward_mean
__global__ __launch_bounds__(256) void k_acc_rows(float *__restrict__ acc, const float *__restrict__ y, int n, float w,
                                                  int first)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) acc[i] = first ? w * y[i] : acc[i] + w * y[i];
}

extern "C" int crime_mlp_forward_mean(crime_mlp_t *const *hs, int nh, int64_t n, const float *x, float *y)
{
  if (!hs || nh < 1 || !hs[0]) return mlp_error(nullptr, CRIME_GH_EINVAL, "need nh >= 1 handles");
  crime_mlp *h0 = hs[0];
  if (n < 1 || !x || !y) return mlp_error(h0, CRIME_GH_EINVAL, "need n >= 1 and non-NULL x, y");
  int batch = h0->max_batch;
  for (int k = 0; k < nh; k++) {
    if (!hs[k]) return mlp_error(h0, CRIME_GH_EINVAL, "handle %d is NULL", k);
    if (hs[k]->device != h0->device) return mlp_error(h0, CRIME_GH_EINVAL, "handle %d is on another device", k);
    batch = std::min(batch, hs[k]->max_batch);
  }
  MLP_HIP(h0, hipSetDevice(h0->device));
  for (int k = 0; k < nh; k++) {
    int rc = ensure_loss(hs[k], 1);
    if (rc) return rc;
    MLP_HIP(h0, hipStreamSynchronize(hs[k]->stream));  // no pending work of a handle on its buffers
  }
  // every network reads the first handle's input buffer on the first handle's stream; the mean collects in
  // its target buffer (scratch outside a training step)
  hipStream_t s = h0->stream;
  const float w = 1.0f / (float)nh;
  for (int64_t i0 = 0; i0 < n; i0 += batch) {
    const int c = (int)(n - i0 < batch ? n - i0 : batch);
    MLP_HIP(h0, hipMemcpyAsync(h0->X, x + i0 * NIN, (size_t)c * NIN * sizeof(float), hipMemcpyHostToDevice, s));
    for (int k = 0; k < nh; k++) {
      enqueue_forward(hs[k], c, 0, s, h0->X);
      hipLaunchKernelGGL(k_acc_rows, dim3((c + 255) / 256), dim3(256), 0, s, h0->T, hs[k]->Y, c, w, k == 0);
    }
    MLP_HIP(h0, hipGetLastError());
    MLP_HIP(h0, hipMemcpyAsync(y + i0, h0->T, (size_t)c * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  MLP_HIP(h0, hipStreamSynchronize(s));
  return 0;
}

extern "C" int crime_mlp_loss_grad(crime_mlp_t *h, int n, const float *x, const float *t, const float *w,
                                   double *loss, float *grad)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!x || !t || !loss || !grad) return mlp_error(h, CRIME_GH_EINVAL, "NULL argument");
  if (n < 64 || n > h->max_batch || n % 64)
    return mlp_error(h, CRIME_GH_EINVAL, "n %d: a multiple of 64 in [64, %d]", n, h->max_batch);
  MLP_HIP(h, hipSetDevice(h->device));
  int rc = ensure_loss(h, 1);
  if (rc) return rc;
  DevState s;
  MLP_HIP(h, hipMemcpy(&s, h->st, sizeof(s), hipMemcpyDeviceToHost));
  const int saved_local = s.local;
  s.local = 0;
  MLP_HIP(h, hipMemcpy(h->st, &s, sizeof(s), hipMemcpyHostToDevice));
  const float zero = 0.0f;
  MLP_HIP(h, hipMemcpy(h->loss, &zero, sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->X, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->T, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  if (w) {
    MLP_HIP(h, hipMemcpy(h->Wt, w, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<float> ones((size_t)n, 1.0f);
    MLP_HIP(h, hipMemcpy(h->Wt, ones.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  }
  enqueue_forward(h, n, 1);
  enqueue_backward(h, n);
  for (const Seg &g : segments(h))
    hipLaunchKernelGGL(k_grad_out, dim3(nblk(g.count)), dim3(256), 0, h->stream, g.part,
                       g.gemm ? nsplit_of(n) : nred_of(n), g.count, h->grad + g.off);
  MLP_HIP(h, hipGetLastError());
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  float l = 0;
  MLP_HIP(h, hipMemcpy(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost));
  MLP_HIP(h, hipMemcpy(grad, h->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  s.local = saved_local;
  MLP_HIP(h, hipMemcpy(h->st, &s, sizeof(s), hipMemcpyHostToDevice));
  *loss = l;
  return 0;
}

extern "C" int crime_mlp_set_data(crime_mlp_t *h, int64_t n, const float *x, const float *t, const float *w)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (n < 1 || n > ((int64_t)1 << 40) || !x || !t) return mlp_error(h, CRIME_GH_EINVAL, "bad training set");
  MLP_HIP(h, hipSetDevice(h->device));
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  if (h->Xall) MLP_HIP(h, hipFree(h->Xall));
  if (h->Tall) MLP_HIP(h, hipFree(h->Tall));
  if (h->Wall) MLP_HIP(h, hipFree(h->Wall));
  h->Xall = h->Tall = h->Wall = nullptr;
  h->ndata = 0;
  MLP_HIP(h, hipMalloc((void **)&h->Xall, (size_t)n * NIN * sizeof(float)));
  MLP_HIP(h, hipMalloc((void **)&h->Tall, (size_t)n * sizeof(float)));
  MLP_HIP(h, hipMemcpy(h->Xall, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemcpy(h->Tall, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMalloc((void **)&h->Wall, (size_t)n * sizeof(float)));
  if (w) {
    MLP_HIP(h, hipMemcpy(h->Wall, w, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<float> ones((size_t)n, 1.0f);
    MLP_HIP(h, hipMemcpy(h->Wall, ones.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  }
  h->ndata = n;
  return 0;
}

extern "C" int crime_mlp_set_graph(crime_mlp_t *h, int on)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  h->use_graph = on ? 1 : 0;
  return 0;
}

extern "C" int crime_mlp_train(crime_mlp_t *h, int steps, int batch, float lr0, int total_steps, float beta1,
                               float beta2, float eps, uint32_t seed, float *loss_out)
{
  if (!h) return mlp_error(nullptr, CRIME_GH_EINVAL, "handle is NULL");
  if (!h->ndata) return mlp_error(h, CRIME_GH_ESTATE, "train before set_data");
  if (steps < 1 || batch < 64 || batch > h->max_batch || batch % 64)
    return mlp_error(h, CRIME_GH_EINVAL, "steps %d, batch %d (a multiple of 64 <= %d)", steps, batch, h->max_batch);
  if (!(lr0 > 0) || total_steps < 1 || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
    return mlp_error(h, CRIME_GH_EINVAL, "bad Adam hyper-parameters");
  MLP_HIP(h, hipSetDevice(h->device));
  int rc = ensure_loss(h, steps);
  if (rc) return rc;
  DevState s;
  MLP_HIP(h, hipMemcpy(&s, h->st, sizeof(s), hipMemcpyDeviceToHost));
  s.local = 0;
  s.total = total_steps;
  s.lr0 = lr0;
  s.b1 = beta1;
  s.b2 = beta2;
  s.eps = eps;
  s.seed = seed;
  s.ndata = h->ndata;
  MLP_HIP(h, hipMemcpy(h->st, &s, sizeof(s), hipMemcpyHostToDevice));
  MLP_HIP(h, hipMemset(h->loss, 0, (size_t)steps * sizeof(float)));
  if (!h->use_graph) {
    for (int i = 0; i < steps; i++) {
      hipLaunchKernelGGL(k_gather, dim3(nblk(batch)), dim3(256), 0, h->stream, h->Xall, h->Tall, h->Wall, batch, h->st,
                         h->X, h->T, h->Wt);
      enqueue_forward(h, batch, 1);
      enqueue_backward(h, batch);
      enqueue_adam(h, batch);
      MLP_HIP(h, hipGetLastError());
    }
    MLP_HIP(h, hipStreamSynchronize(h->stream));
    if (loss_out) MLP_HIP(h, hipMemcpy(loss_out, h->loss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
  }
  if (h->exec && h->graph_batch != batch) drop_graph(h);
  if (!h->exec) {
    MLP_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_gather, dim3(nblk(batch)), dim3(256), 0, h->stream, h->Xall, h->Tall, h->Wall, batch, h->st,
                       h->X, h->T, h->Wt);
    enqueue_forward(h, batch, 1);
    enqueue_backward(h, batch);
    enqueue_adam(h, batch);
    hipError_t e = hipStreamEndCapture(h->stream, &h->graph);
    if (e != hipSuccess) {
      h->graph = nullptr;
      return mlp_error(h, CRIME_GH_EHIP, "step capture: %s", hipGetErrorString(e));
    }
    MLP_HIP(h, hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    h->graph_batch = batch;
  }
  for (int i = 0; i < steps; i++) MLP_HIP(h, hipGraphLaunch(h->exec, h->stream));
  MLP_HIP(h, hipStreamSynchronize(h->stream));
  if (loss_out) MLP_HIP(h, hipMemcpy(loss_out, h->loss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}
