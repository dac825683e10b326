// Stockham autosort FFT engine for gfx950: one workgroup transforms a tile of lines of length L.
//
// Layout of the work: a line is split over TC threads; thread t owns E = L/TC elements, held in
// VGPRs as B = E/R radix-R butterflies per pass (butterfly j = t + b*TC).  Pass p reads element
// j + s*L/R, applies the twiddles exp(+2 pi i s (j mod Ns)/(Ns R)) and a radix-R DFT, and writes
// element (j/Ns)*Ns*R + (j mod Ns) + s*Ns (Stockham autosort, so no bit reversal is needed).
// Pass 0 reads through a caller functor (global memory or an on-the-fly generator), intermediate
// passes exchange through LDS, the last pass writes through a caller functor (epilogue).
// Radix plan: as many 16s (or 8s) as possible, then 8/4/2, then 3s.  Inverse transforms (sign +).
#pragma once

#include "gh_device.hpp"

namespace crime {

struct RadixPlan {
  int n;
  int r[10];
};

constexpr RadixPlan make_plan(int L, int maxr = 16)
{
  RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = L, b = 0, a = 0;
  while (m % 3 == 0) {
    m /= 3;
    b++;
  }
  while (m % 2 == 0) {
    m /= 2;
    a++;
  }
  if (maxr == 8) {   // radix 8 at most: twice the threads per line of the radix-16 plan, one pass more
    int n8 = a / 3, rem = a % 3;
    if (rem == 1 && n8 >= 1) n8 -= 1;
    for (int i = 0; i < n8; i++) p.r[p.n++] = 8;
    if (rem == 1) {
      if (a >= 4) {
        p.r[p.n++] = 4;
        p.r[p.n++] = 4;
      }
      else {
        p.r[p.n++] = 2;
      }
    }
    if (rem == 2) p.r[p.n++] = 4;
    for (int i = 0; i < b; i++) p.r[p.n++] = 3;
    return p;
  }
  int n16 = a / 4, rem = a % 4;
  if (rem == 1 && n16 >= 1) n16 -= 1;
  for (int i = 0; i < n16; i++) p.r[p.n++] = 16;
  if (rem == 1) {
    if (a >= 5) {
      p.r[p.n++] = 8;
      p.r[p.n++] = 4;
    }
    else {
      p.r[p.n++] = 2;
    }
  }
  if (rem == 2) p.r[p.n++] = 4;
  if (rem == 3) p.r[p.n++] = 8;
  for (int i = 0; i < b; i++) p.r[p.n++] = 3;
  return p;
}

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

// largest number of threads per line such that every pass has a whole number of butterflies per
// thread, capped at tcmax
constexpr int plan_tc(int L, int tcmax, int maxr = 16)
{
  RadixPlan p = make_plan(L, maxr);
  int g = L / p.r[0];
  for (int i = 1; i < p.n; i++) g = cgcd(g, L / p.r[i]);
  while (g > tcmax) g = (g % 2 == 0) ? g / 2 : g / 3;
  return g;
}

// ---------------------------------------------------------------------------------------------
// small inverse DFTs on registers
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cx<T> &a, cx<T> &b)
{
  cx<T> t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T> &a0, cx<T> &a1, cx<T> &a2, cx<T> &a3)
{
  cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmuli(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename T>
__device__ __forceinline__ void dft3(cx<T> &a0, cx<T> &a1, cx<T> &a2)
{
  const T s3 = (T)0.866025403784438646763723170752936;
  cx<T> t = cadd(a1, a2), u = csub(a1, a2);
  cx<T> m = {a0.x - (T)0.5 * t.x, a0.y - (T)0.5 * t.y};
  a0 = cadd(a0, t);
  cx<T> iu = {-s3 * u.y, s3 * u.x};
  a1 = cadd(m, iu);
  a2 = csub(m, iu);
}

// multiply by exp(+2 pi i k / 16), k in {1,2,3,4,6,9}
template <typename T, int K>
__device__ __forceinline__ cx<T> w16(cx<T> a)
{
  const T c1 = (T)0.923879532511286756128183189396788, s1 = (T)0.382683432365089771728459984030399;
  const T h = (T)0.707106781186547524400844362104849;
  if constexpr (K == 0) return a;
  else if constexpr (K == 1) return {c1 * a.x - s1 * a.y, c1 * a.y + s1 * a.x};
  else if constexpr (K == 2) return {h * (a.x - a.y), h * (a.x + a.y)};
  else if constexpr (K == 3) return {s1 * a.x - c1 * a.y, s1 * a.y + c1 * a.x};
  else if constexpr (K == 4) return cmuli(a);
  else if constexpr (K == 6) return {-h * (a.x + a.y), h * (a.x - a.y)};
  else return {-c1 * a.x + s1 * a.y, -c1 * a.y - s1 * a.x};  // K == 9
}

template <typename T>
__device__ __forceinline__ void dft8(cx<T> *v)
{
  // decimation in time: even/odd radix-4 halves, twiddles exp(+2 pi i m/8)
  dft4(v[0], v[2], v[4], v[6]);
  dft4(v[1], v[3], v[5], v[7]);
  cx<T> o1 = w16<T, 2>(v[3]), o2 = cmuli(v[5]), o3 = w16<T, 6>(v[7]);
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <typename T>
__device__ __forceinline__ void dft16(cx<T> *v)
{
  // n = n1 + 4 n2, m = m2 + 4 m1:  X_{n1}[m2] = DFT4_{n2}(v[n1+4n2]); twiddle W16^{n1 m2};
  // out[m2 + 4 m1] = DFT4_{n1}(X_{n1}[m2])
  dft4(v[0], v[4], v[8], v[12]);
  dft4(v[1], v[5], v[9], v[13]);
  dft4(v[2], v[6], v[10], v[14]);
  dft4(v[3], v[7], v[11], v[15]);
  // X_{n1}[m2] now sits in v[n1 + 4 m2]
  v[5] = w16<T, 1>(v[5]);
  v[9] = w16<T, 2>(v[9]);
  v[13] = w16<T, 3>(v[13]);
  v[6] = w16<T, 2>(v[6]);
  v[10] = w16<T, 4>(v[10]);
  v[14] = w16<T, 6>(v[14]);
  v[7] = w16<T, 3>(v[7]);
  v[11] = w16<T, 6>(v[11]);
  v[15] = w16<T, 9>(v[15]);
  dft4(v[0], v[1], v[2], v[3]);
  dft4(v[4], v[5], v[6], v[7]);
  dft4(v[8], v[9], v[10], v[11]);
  dft4(v[12], v[13], v[14], v[15]);
  // out[m2 + 4 m1] sits in v[m1 + 4 m2]: transpose the 4x4 block
  cx<T> t;
  t = v[1]; v[1] = v[4]; v[4] = t;
  t = v[2]; v[2] = v[8]; v[8] = t;
  t = v[3]; v[3] = v[12]; v[12] = t;
  t = v[6]; v[6] = v[9]; v[9] = t;
  t = v[7]; v[7] = v[13]; v[13] = t;
  t = v[11]; v[11] = v[14]; v[14] = t;
}

template <typename T, int R>
__device__ __forceinline__ void dft(cx<T> *v)
{
  if constexpr (R == 2) dft2(v[0], v[1]);
  else if constexpr (R == 3) dft3(v[0], v[1], v[2]);
  else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8(v);
  else dft16(v);
}

// ---------------------------------------------------------------------------------------------
// the engine
// ---------------------------------------------------------------------------------------------
template <typename T, int L, int TC, int MAXR = 16>
struct Stockham {
  static constexpr RadixPlan plan = make_plan(L, MAXR);
  static constexpr int NP = plan.n;
  static constexpr int E = L / TC;

  template <int P>
  static constexpr int radix() { return plan.r[P]; }
  template <int P>
  static constexpr int ns()
  {
    int n = 1;
    for (int i = 0; i < P; i++) n *= plan.r[i];
    return n;
  }
  template <int P>
  __device__ static __forceinline__ int in_index(int t, int b, int s)
  {
    return t + b * TC + s * (L / radix<P>());
  }
  template <int P>
  __device__ static __forceinline__ int out_index(int t, int b, int s)
  {
    constexpr int R = radix<P>(), Ns = ns<P>();
    int j = t + b * TC;
    return (j / Ns) * Ns * R + (j % Ns) + s * Ns;
  }

  template <int P>
  __device__ static __forceinline__ void compute(cx<T> (&v)[E], int t, const cx<T> *__restrict__ tw)
  {
    constexpr int R = radix<P>(), Ns = ns<P>(), B = E / R, TWS = L / (Ns * R);
#pragma unroll
    for (int b = 0; b < B; ++b) {
      if constexpr (Ns > 1) {
        int k = (t + b * TC) % Ns;
#pragma unroll
        for (int s = 1; s < R; ++s) v[b * R + s] = cmul(v[b * R + s], tw[s * k * TWS]);
      }
      dft<T, R>(&v[b * R]);
    }
  }

  template <int P, class Store, class Lds>
  __device__ static __forceinline__ void step(cx<T> (&v)[E], int t, Store &store, Lds &lds,
                                              const cx<T> *__restrict__ tw)
  {
    compute<P>(v, t, tw);
    constexpr int R = radix<P>(), B = E / R;
    if constexpr (P + 1 < NP) {
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int s = 0; s < R; ++s) lds(out_index<P>(t, b, s)) = v[b * R + s];
      __syncthreads();
      constexpr int R2 = radix<P + 1>(), B2 = E / R2;
#pragma unroll
      for (int b = 0; b < B2; ++b)
#pragma unroll
        for (int s = 0; s < R2; ++s) v[b * R2 + s] = lds(in_index<P + 1>(t, b, s));
      __syncthreads();
      step<P + 1>(v, t, store, lds, tw);
    }
    else {
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int s = 0; s < R; ++s) store(out_index<P>(t, b, s), v[b * R + s]);
    }
  }

  // load(e) -> element e of this thread's line; store(e, value); lds(e) -> LDS slot of element e.
  // LoadFromLds: the inputs were staged in the same LDS tile, so wait before pass 0 overwrites it.
  template <bool LoadFromLds = false, class Load, class Store, class Lds>
  __device__ static __forceinline__ void run(int t, Load &load, Store &store, Lds &lds,
                                             const cx<T> *__restrict__ tw)
  {
    cx<T> v[E];
    constexpr int R = radix<0>(), B = E / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int s = 0; s < R; ++s) v[b * R + s] = load(in_index<0>(t, b, s));
    if constexpr (LoadFromLds) __syncthreads();
    step<0>(v, t, store, lds, tw);
  }
};

}  // namespace crime
