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

// Multiply-adds are fused only within an expression (a * b + c as written),
// never across statements: the fusion the compiler may otherwise choose
// depends on the code a transform is inlined into, and then the same
// transform rounds differently in different kernels.
#pragma clang fp contract(on)

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

// X_k = sum_n x_n w^(n k), w = exp(-+2 pi i / 5): with t1 = x1 + x4, t2 = x2 +
// x3, t3 = x1 - x4, t4 = x2 - x3, X_1,4 = x0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4)
// and X_2,3 = x0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4), c_k = cos(2 pi k / 5),
// s_k = sin(2 pi k / 5) (mul_si carries the sign of the transform).
template <bool INV>
struct Bfly<5, INV> {
  __device__ __forceinline__ static void run(float2* v)
  {
    const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;
    const float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;
    const float2 a = v[0];
    const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const float2 A1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const float2 A2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const float2 B1 = mul_si<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    const float2 B2 = mul_si<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(A1, B1);
    v[4] = csub(A1, B1);
    v[2] = cadd(A2, B2);
    v[3] = csub(A2, B2);
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

// 16 points as 4 x 4: X[k1 + 4 k2] = sum_m1 W^(m1 k1) W4^(m1 k2) sum_m2
// v[m1 + 4 m2] W4^(m2 k1), W = exp(-+2 pi i / 16); in and out in natural order.
template <bool INV>
struct Bfly<16, INV> {
  __device__ __forceinline__ static float2 w16(int j)
  {
    // cos and sin of 2 pi j / 16 for the products m1 k1 of the inner twiddles
    constexpr float c[10] = {1.0f, 0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f, 0.0f,
                             -0.382683432365089772f, -0.707106781186547524f, -0.923879532511286756f, -1.0f,
                             -0.923879532511286756f};
    constexpr float sn[10] = {0.0f, 0.382683432365089772f, 0.707106781186547524f, 0.923879532511286756f, 1.0f,
                              0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f, 0.0f,
                              -0.382683432365089772f};
    return make_float2(c[j], INV ? sn[j] : -sn[j]);
  }
  __device__ __forceinline__ static void run(float2* v)
  {
    float2 a[4][4];
#pragma unroll
    for (int m1 = 0; m1 < 4; ++m1) {
      float2 q[4] = {v[m1], v[m1 + 4], v[m1 + 8], v[m1 + 12]};
      Bfly<4, INV>::run(q);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[m1][k1] = (m1 * k1 == 0) ? q[k1] : cmul(q[k1], w16(m1 * k1));
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Bfly<4, INV>::run(q);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = q[k2];
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

// The same transform with the pencil points in registers, for tiles whose
// thread t owns pencil c = t % NCOL and points n0 + i RPI (n0 = t / NCOL, RPI =
// NT / NCOL, i < L / RPI): the points of the first stage's butterflies j =
// n0 + p RPI, j + m L/R0, are that thread's registers r[p + m PT], and so are
// the outputs of the last stage's butterflies j = n0 + p RPI, j + m L/RL.  The
// first stage therefore runs on the registers and writes its outputs to LDS,
// and the last reads its inputs from LDS and leaves its outputs in the
// registers, in the layout they came in: two LDS passes and four barriers
// fewer than storing the tile, transforming it in LDS and loading it back,
// with the same operations on the same values (the results are identical).
// Needs L/R0 and L/RL to be multiples of RPI; otherwise the tile goes through
// LDS and fft_lds.  No barrier after the last stage: the caller synchronises
// before it next writes s.
template <int L, int NCOL, int NT, bool INV, int SN, int R>
__device__ __forceinline__ void reg_stage_first(float2* r, float2* s, int c, int n0)
{
  constexpr int RPI = NT / NCOL, PT = (L / R) / RPI;
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = r[p + m * PT];
    Bfly<R, INV>::run(v);
    const int j = n0 + p * RPI;
#pragma unroll
    for (int m = 0; m < R; ++m) s[(j * R + m) * SN + c] = v[m];
  }
  __syncthreads();
}

template <int L, int NCOL, int NT, bool INV, int SN, int TWS, int R>
__device__ __forceinline__ void reg_stage_last(float2* r, const float2* s, const float2* tw, int c, int n0)
{
  constexpr int RPI = NT / NCOL, NS = L / R, PT = NS / RPI;
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    const int j = n0 + p * RPI;  // j < NS: twiddle index k = j, outputs j + m NS
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = s[(j + m * NS) * SN + c];
#pragma unroll
    for (int m = 1; m < R; ++m) {
      float2 w = tw[m * j * TWS];
      if (INV) w.y = -w.y;
      v[m] = cmul(v[m], w);
    }
    Bfly<R, INV>::run(v);
#pragma unroll
    for (int m = 0; m < R; ++m) r[p + m * PT] = v[m];
  }
}

template <int L, int NS, int NCOL, int NT, bool INV, int SN, int TWS, int R, int... Rest>
__device__ __forceinline__ void reg_stages_rest(float2* r, float2* s, const float2* tw, int tid, int c, int n0)
{
  if constexpr (sizeof...(Rest) == 0) {
    static_assert(NS * R == L, "radices must multiply to L");
    reg_stage_last<L, NCOL, NT, INV, SN, TWS, R>(r, s, tw, c, n0);
  } else {
    stockham_stage<L, R, NS, NCOL, NT, INV, SN, 1, true, TWS, false>(s, tw, tid);
    reg_stages_rest<L, NS * R, NCOL, NT, INV, SN, TWS, Rest...>(r, s, tw, tid, c, n0);
  }
}

template <int... Rs>
struct LastRadix;
template <int R>
struct LastRadix<R> {
  static constexpr int value = R;
};
template <int R, int R2, int... Rest>
struct LastRadix<R, R2, Rest...> {
  static constexpr int value = LastRadix<R2, Rest...>::value;
};

template <int L, int NCOL, int NT, bool INV, int SN, int TWS, int R0, int... Rest>
__device__ __forceinline__ void reg_fft(float2* r, float2* s, const float2* tw, int tid)
{
  constexpr int RPI = NT / NCOL, EPT = L / RPI;
  static_assert(NT % NCOL == 0 && L % RPI == 0, "tile must split evenly over the threads");
  const int c = tid % NCOL, n0 = tid / NCOL;
  constexpr int RL = LastRadix<R0, Rest...>::value;
  if constexpr (sizeof...(Rest) > 0 && (L / R0) % RPI == 0 && (L / RL) % RPI == 0) {
    reg_stage_first<L, NCOL, NT, INV, SN, R0>(r, s, c, n0);
    reg_stages_rest<L, R0, NCOL, NT, INV, SN, TWS, Rest...>(r, s, tw, tid, c, n0);
  } else {
#pragma unroll
    for (int i = 0; i < EPT; ++i) s[(n0 + i * RPI) * SN + c] = r[i];
    __syncthreads();
    run_stages<L, 1, NCOL, NT, INV, SN, 1, true, TWS, false, R0, Rest...>(s, tw, tid);
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = s[(n0 + i * RPI) * SN + c];
  }
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
    template <int NCOL, int NT, bool INV, int SN, int TWS>                                                \
    __device__ __forceinline__ static void run_reg(float2* r, float2* s, const float2* tw, int tid)       \
    {                                                                                                     \
      reg_fft<L_, NCOL, NT, INV, SN, TWS, __VA_ARGS__>(r, s, tw, tid);                                    \
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
// lengths with a factor 5 (grids 160, 320, 640): the radix-5 stage sits where
// the strided tiles' first and last stages can still run on registers
MRX_SCHED(80, 0, 4, 4, 5)
MRX_SCHED(160, 0, 8, 5, 4)
MRX_SCHED(320, 0, 8, 8, 5)
MRX_SCHED(640, 0, 8, 4, 4, 5)
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
MRX_SCHED(80, 1, 5, 4, 4)
MRX_SCHED(160, 1, 4, 5, 8)
MRX_SCHED(320, 1, 4, 4, 4, 5)
MRX_SCHED(640, 1, 4, 4, 8, 5)
#undef MRX_SCHED

// FFT of length L over NCOL pencils in LDS; the table has L*TWS entries.
template <int L, int NCOL, int NT, bool INV, int SN, int SC, bool CFAST, int TWS, bool PAD = false, int V = 0>
__device__ __forceinline__ void fft_lds(float2* s, const float2* tw, int tid)
{
  SchedV<L, V>::template run<NCOL, NT, INV, SN, SC, CFAST, TWS, PAD>(s, tw, tid);
}

// fft_lds over the pencils of a strided tile whose points are in registers
// (reg_fft's layout), results back in the registers.
template <int L, int NCOL, int NT, bool INV, int SN, int TWS = 1, int V = 0>
__device__ __forceinline__ void fft_tile_reg(float2* r, float2* s, const float2* tw, int tid)
{
  SchedV<L, V>::template run_reg<NCOL, NT, INV, SN, TWS>(r, s, tw, tid);
}

}  // namespace mrx
