// fft_lds.hpp - workgroup-level Stockham FFT engine for CDNA4 (gfx950).
//
// A workgroup holds a tile of independent complex pencils in LDS: element n of
// pencil c lives at s[n*SN + c*SC] (float2).  Each radix-R stage of the
// autosorting Stockham recursion reads R points j + m*L/R, twiddles them by
// exp(-+2 pi i m k / (NS R)) with k = j mod NS, runs a radix-R DFT in
// registers and writes them to (j/NS)*NS*R + k + m*NS, so the output of the
// last stage is in natural order with no bit reversal.  A thread keeps all of
// its butterflies of a stage in registers between the read phase and the write
// phase, so stages run in place on one LDS buffer with two barriers each.
//
// Twiddles come from a table of exp(-2 pi i t / NTAB) held in LDS (computed in
// double on the host, rounded once); an FFT of length L reads it with stride
// NTAB / L.  CFAST selects which index runs fastest across a wavefront: the
// pencil index (columns of a strided tile) or the point index (contiguous
// rows), whichever keeps a wave's LDS addresses consecutive.
#pragma once
#include <hip/hip_runtime.h>

namespace mrx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// streaming store: the line is not kept in the caches for a later read
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(float2* p, float2 v)
{
  __builtin_nontemporal_store(f32x2{v.x, v.y}, reinterpret_cast<f32x2*>(p));
}
// multiply by (sign * i): INV=false -> -i (forward kernel exp(-i...)), INV=true -> +i
template <bool INV>
__device__ __forceinline__ float2 mul_si(float2 a)
{
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
struct Bfly;

template <bool INV>
struct Bfly<2, INV> {
  __device__ __forceinline__ static void run(float2* v)
  {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Bfly<3, INV> {
  __device__ __forceinline__ static void run(float2* v)
  {
    const float h = 0.866025403784438647f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = cadd(b, c);
    float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    float2 d = cscale(csub(b, c), h);
    float2 e = mul_si<INV>(d);
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, e);
    v[2] = csub(t2, e);
  }
};

template <bool INV>
struct Bfly<4, INV> {
  __device__ __forceinline__ static void run(float2* v)
  {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = mul_si<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <bool INV>
struct Bfly<8, INV> {
  __device__ __forceinline__ static void run(float2* v)
  {
    const float r = 0.707106781186547524f;
    float2 a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = cadd(v[q], v[q + 4]);
      b[q] = csub(v[q], v[q + 4]);
    }
    // b[q] *= w8^q, w8 = exp(-+ i pi/4)
    b[1] = cscale(cadd(b[1], mul_si<INV>(b[1])), r);
    b[2] = mul_si<INV>(b[2]);
    b[3] = cscale(csub(mul_si<INV>(b[3]), b[3]), r);
    Bfly<4, INV>::run(a);
    Bfly<4, INV>::run(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
};

// LDS address of linear element i: with PAD, one float2 of padding every 32
// elements, so the stride-R writes of a Stockham stage over contiguous rows
// land in distinct banks.
template <bool PAD>
__device__ __forceinline__ int lds_addr(int i)
{
  return PAD ? i + (i >> 5) : i;
}

// One Stockham stage over a tile of NCOL pencils of length L.
template <int L, int R, int NS, int NCOL, int NT, bool INV, int SN, int SC, bool CFAST, int TWS, bool PAD = false>
__device__ __forceinline__ void stockham_stage(float2* s, const float2* tw, int tid)
{
  constexpr int NB = NCOL * (L / R);
  constexpr int PT = (NB + NT - 1) / NT;
  float2 v[PT][R];
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    const int b = tid + p * NT;
    if ((NB % NT == 0) || b < NB) {
      const int c = CFAST ? (b % NCOL) : (b / (L / R));
      const int j = CFAST ? (b / NCOL) : (b % (L / R));
#pragma unroll
      for (int m = 0; m < R; ++m) v[p][m] = s[lds_addr<PAD>((j + m * (L / R)) * SN + c * SC)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    const int b = tid + p * NT;
    if ((NB % NT == 0) || b < NB) {
      const int c = CFAST ? (b % NCOL) : (b / (L / R));
      const int j = CFAST ? (b / NCOL) : (b % (L / R));
      const int k = (NS == 1) ? 0 : (j % NS);
      if (NS > 1) {
#pragma unroll
        for (int m = 1; m < R; ++m) {
          float2 w = tw[m * k * (L / (NS * R)) * TWS];
          if (INV) w.y = -w.y;
          v[p][m] = cmul(v[p][m], w);
        }
      }
      Bfly<R, INV>::run(v[p]);
      const int idx = (NS == 1) ? (j * R) : ((j / NS) * NS * R + k);
#pragma unroll
      for (int m = 0; m < R; ++m) s[lds_addr<PAD>((idx + m * NS) * SN + c * SC)] = v[p][m];
    }
  }
  __syncthreads();
}

template <int L, int NS, int NCOL, int NT, bool INV, int SN, int SC, bool CFAST, int TWS, bool PAD, int R,
          int... Rest>
__device__ __forceinline__ void run_stages(float2* s, const float2* tw, int tid)
{
  stockham_stage<L, R, NS, NCOL, NT, INV, SN, SC, CFAST, TWS, PAD>(s, tw, tid);
  if constexpr (sizeof...(Rest) > 0)
    run_stages<L, NS * R, NCOL, NT, INV, SN, SC, CFAST, TWS, PAD, Rest...>(s, tw, tid);
}

// Radix schedules (product = L).  Variant 0 puts radix 8 first, which keeps
// later strides small; variant 1 puts radix 4 first (fewer registers per
// butterfly, one more stage).  Which is faster depends on the tile shape and
// is chosen by measurement.
template <int L, int V>
struct SchedV;
#define MRX_SCHED(L_, V_, ...)                                                                             \
  template <>                                                                                             \
  struct SchedV<L_, V_> {                                                                                 \
    template <int NCOL, int NT, bool INV, int SN, int SC, bool CFAST, int TWS, bool PAD>                  \
    __device__ __forceinline__ static void run(float2* s, const float2* tw, int tid)                      \
    {                                                                                                     \
      run_stages<L_, 1, NCOL, NT, INV, SN, SC, CFAST, TWS, PAD, __VA_ARGS__>(s, tw, tid);                 \
    }                                                                                                     \
  };
MRX_SCHED(16, 0, 4, 4)
MRX_SCHED(24, 0, 8, 3)
MRX_SCHED(32, 0, 8, 4)
MRX_SCHED(48, 0, 8, 2, 3)
MRX_SCHED(64, 0, 8, 8)
MRX_SCHED(96, 0, 8, 4, 3)
MRX_SCHED(128, 0, 8, 4, 4)
MRX_SCHED(192, 0, 8, 8, 3)
MRX_SCHED(256, 0, 8, 8, 4)
MRX_SCHED(384, 0, 8, 4, 4, 3)
MRX_SCHED(512, 0, 8, 8, 8)
MRX_SCHED(768, 0, 8, 8, 4, 3)
MRX_SCHED(1024, 0, 8, 8, 4, 4)
MRX_SCHED(16, 1, 4, 4)
MRX_SCHED(24, 1, 4, 2, 3)
MRX_SCHED(32, 1, 4, 8)
MRX_SCHED(48, 1, 4, 4, 3)
MRX_SCHED(64, 1, 4, 4, 4)
MRX_SCHED(96, 1, 4, 8, 3)
MRX_SCHED(128, 1, 4, 4, 8)
MRX_SCHED(192, 1, 4, 4, 4, 3)
MRX_SCHED(256, 1, 4, 4, 4, 4)
MRX_SCHED(384, 1, 4, 4, 8, 3)
MRX_SCHED(512, 1, 4, 4, 4, 8)
MRX_SCHED(768, 1, 4, 4, 4, 4, 3)
MRX_SCHED(1024, 1, 4, 4, 4, 4, 4)
#undef MRX_SCHED

// FFT of length L over NCOL pencils in LDS; the table has L*TWS entries.
template <int L, int NCOL, int NT, bool INV, int SN, int SC, bool CFAST, int TWS, bool PAD = false, int V = 0>
__device__ __forceinline__ void fft_lds(float2* s, const float2* tw, int tid)
{
  SchedV<L, V>::template run<NCOL, NT, INV, SN, SC, CFAST, TWS, PAD>(s, tw, tid);
}

}  // namespace mrx
