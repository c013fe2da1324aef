// reion_kernels.hpp - HIP kernels of the excursion-set reionization path.
//
// Device-internal spectrum layout ("packed"): for an N^3 real field the half
// spectrum is stored as C[x][y][kz], kz in [0, N/2), complex64, so one
// spectrum is exactly N^3 floats with power-of-two rows.  The two real-valued
// z-planes of the transform (kz = 0 and kz = N/2) share the kz = 0 slot as
// (DC + i*Nyquist) while the x and y passes run; in the unfiltered spectra the
// kz = 0 slot holds the true kz = 0 plane A and the Nyquist plane B lives in
// its own N x N array V[ky][kx], so each can be filtered with its own |k|.
//
// The 3-D transform is three passes, one per axis:
//   z  rows are contiguous: a length-N/2 complex FFT of the even/odd packed
//      row plus the real-to-half-spectrum twist (R2C), or the reverse (C2R);
//   y, x  strided pencils: a workgroup loads an N x 16 tile (16 consecutive kz,
//      128-byte row segments), transforms the 16 pencils in LDS and stores
//      the tile back, so every HBM access is a coalesced segment.
// The window W(|k|R) is applied while the x pass loads (computed once per
// mode and reused for every field), and the clamp plus the whole per-cell
// ionisation test run inside the final z pass, so filtered fields never
// travel to HBM in real space.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"

namespace mrx {

constexpr int kTW = 16;        // pencils per strided tile
constexpr int kZE = 2048;      // complex elements per z-pass workgroup
constexpr int kNT = 256;       // threads per workgroup
constexpr int kColNT = 512;    // at most this many threads per strided tile
constexpr float kRelTol = 1e-5f;  // REL_TOL (src/meraxes.h:41)
constexpr float kAbsTol = 1e-8f;  // ABS_TOL (src/meraxes.h:42)

constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }

// Threads of the windowed x pass (k_ix_filter): 256 keep a tile's window
// values in registers up to 512^3; longer pencils need 512 threads, or the
// points and window values of a thread (2 N / 256 + N / 256 registers) spill.
constexpr int ix_max_threads(int N) { return N >= 640 ? 512 : 256; }

// Rows per z-pass workgroup for rows of NH complex: about kZE elements, and a
// multiple of 512 elements, so that every thread of kNT owns whole groups of 4
// cells (NH = 80, 160, 320 take 2560 elements instead of 1280).
constexpr int z_rows(int NH)
{
  int r = (NH >= kZE) ? 1 : pow2_floor(kZE / NH);
  while ((r * NH) % 512 != 0) r *= 2;
  return r;
}

// The fewest rows of NH complex holding a multiple of 128 elements (4-cell
// groups over one 64-thread wave).
constexpr int z_rows_wave(int NH)
{
  int r = 1;
  while ((r * NH) % 128 != 0) r *= 2;
  return r;
}

template <int N>
struct Cfg {
  static_assert(N % 16 == 0, "grid dimension must be a multiple of 16");
  static constexpr int NH = N / 2;
  static constexpr int TW = (NH % kTW == 0) ? kTW : kTW / 2;  // pencils per strided tile
  static constexpr int COL_NT_RAW = TW * N / 8;
  static constexpr int COL_NT = COL_NT_RAW < 64 ? 64 : (COL_NT_RAW > kColNT ? kColNT : COL_NT_RAW);
  static constexpr int IX_NT =
    COL_NT_RAW < 64 ? 64 : (COL_NT_RAW > ix_max_threads(N) ? ix_max_threads(N) : COL_NT_RAW);
  static constexpr int RZ = z_rows(NH);  // rows per z workgroup (divides N*N)
  static constexpr int ZE = RZ * NH;                               // complex elements per z workgroup
  static constexpr int CPT = 2 * ZE / kNT;                          // real cells per thread in the z pass
  static_assert((N * N) % RZ == 0, "z workgroups must tile the rows");
  static_assert(ZE % kNT == 0, "z elements must split evenly over the threads");
  static_assert(CPT % 4 == 0, "z-pass threads own whole groups of 4 cells");
  static_assert((N * TW) % COL_NT == 0, "strided tiles must split evenly over the threads");
};

// ------------------------------------------------------------------ window
// filter() of reionization.c:2110-2170, reproduced operation by operation:
// float k components, float |k| by a correctly rounded sqrtf, float kR, and
// the libm transcendental calls replaced by double evaluations rounded to
// float (the correctly rounded value libm's float functions aim for).  The
// top-hat expression cancels catastrophically for small kR, so there its
// terms must be rounded exactly as libm rounds them; above kR = 4 the
// cancellation is gone and single-precision sinf/cosf are accurate enough.
#pragma clang fp contract(off)
__device__ __forceinline__ float kcomp(int n, int N, float dk)
{
  return (n > N / 2) ? (float)(n - N) * dk : (float)n * dk;
}

__device__ __forceinline__ float tophat_window(float kR)
{
  float s, c;
  if (kR < 4.0f) {
    double sd, cd;
    sincos((double)kR, &sd, &cd);
    s = (float)sd;
    c = (float)cd;
  } else {
    sincosf(kR, &s, &c);
  }
  const double x = (double)kR;
  const float p3 = (float)(x * x * x);
  const float p2 = (float)(x * x);
  const float t = s / p3 - c / p2;
  return (float)(3.0 * (double)t);
}

struct WinArgs {
  float R;        // (float)R as passed to filter()
  float delta_k;  // (float)(2 pi / (float)BoxSize)
  int type;       // MRX_FILTER_*
  int apply;      // 0: the unfiltered last step
};

// |k| of mode (nx, ny, nz) as filter() forms it (reionization.c:2116-2136).
__device__ __forceinline__ float k_magnitude(int nx, int ny, int nz, int N, float dk)
{
  const float kx = kcomp(nx, N, dk);
  const float ky = kcomp(ny, N, dk);
  const float kz = (float)nz * dk;
  return sqrtf(kx * kx + ky * ky + kz * kz);
}

// Sharp-k multiplier (reionization.c:2148-2152): 0 or 1.
__device__ __forceinline__ float sharpk_window(float k_mag, float R)
{
  const float kR = (float)((double)(k_mag * R) * 0.413566994);
  return (kR > 1.0f) ? 0.0f : 1.0f;
}

// Multiplier filter() applies to mode (nx, ny, nz); 1 where it leaves it alone.
__device__ __forceinline__ float window(int nx, int ny, int nz, int N, const WinArgs& w)
{
  if (!w.apply) return 1.0f;
  const float k_mag = k_magnitude(nx, ny, nz, N, w.delta_k);
  float kR = k_mag * w.R;
  switch (w.type) {
    case 0:
      if ((double)kR > 1e-4) return tophat_window(kR);
      return 1.0f;
    case 1:
      return sharpk_window(k_mag, w.R);
    default: {
      // powf((float)M_E, y) = exp(y ln((float)M_E)), evaluated in double and
      // rounded once: the correctly rounded value libm's powf returns
      kR = (float)((double)kR * 0.643);
      const float y = (float)((double)(-kR * kR) / 2.0);
      const double ln_ef = 0.99999996963214000;  // log((double)(float)M_E)
      return (float)exp((double)y * ln_ef);
    }
  }
}
#pragma clang fp contract(on)

// ------------------------------------------------------------- z: R2C pass
// in: padded real rows (N+2 floats); out: packed half spectrum rows (N/2).
// Workgroup blockIdx.x transforms rows [blockIdx.x * RZ, +RZ) of one field.
template <int N>
__device__ __forceinline__ void r2c_z_tile(const float* __restrict__ in, float2* __restrict__ dst,
                                           const float2* __restrict__ twN)
{
  using C = Cfg<N>;
  constexpr int NH = C::NH, RZ = C::RZ, ZE = C::ZE;
  __shared__ float2 s[ZE + ZE / 32];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += kNT) tw[i] = twN[i];
  const long row0 = (long)blockIdx.x * RZ;
  const float2* src = reinterpret_cast<const float2*>(in);
  constexpr int EPZ = ZE / kNT;
  float2 v[EPZ];
#pragma unroll
  for (int i = 0; i < EPZ; ++i) {
    const int e = tid + i * kNT, r = e / NH, n = e % NH;
    v[i] = src[(row0 + r) * (NH + 1) + n];
  }
#pragma unroll
  for (int i = 0; i < EPZ; ++i) s[lds_addr<true>(tid + i * kNT)] = v[i];
  __syncthreads();
  fft_lds<NH, RZ, kNT, false, 1, NH, false, 2, true>(s, tw, tid);
  // twist: X[k] = E + w^k O, X[NH-k] = conj(E - w^k O), E=(Z_k+conj Z_{NH-k})/2, O=(Z_k-conj Z_{NH-k})/(2i)
  constexpr int NP = NH / 2 + 1;
  for (int q = tid; q < RZ * NP; q += kNT) {
    const int r = q / NP, k = q % NP;
    const int rb = r * NH;
    float2* orow = dst + (row0 + r) * NH;
    if (k == 0) {
      const float2 z = s[lds_addr<true>(rb)];
      orow[0] = make_float2(z.x + z.y, z.x - z.y);
    } else {
      const float2 zk = s[lds_addr<true>(rb + k)], zm = cconj(s[lds_addr<true>(rb + NH - k)]);
      const float2 E = cscale(cadd(zk, zm), 0.5f);
      const float2 d = csub(zk, zm);
      const float2 O = make_float2(0.5f * d.y, -0.5f * d.x);
      const float2 wo = cmul(tw[k], O);
      orow[k] = cadd(E, wo);
      if (k != NH - k) orow[NH - k] = cconj(csub(E, wo));
    }
  }
}

// blockIdx.y selects the field (strides in_fstride, out_fstride)
template <int N>
__global__ __launch_bounds__(kNT) void k_r2c_z(const float* __restrict__ in, float2* __restrict__ out,
                                              const float2* __restrict__ twN, long in_fstride, long out_fstride)
{
  r2c_z_tile<N>(in + blockIdx.y * in_fstride, out + blockIdx.y * out_fstride, twN);
}

// blockIdx.y selects the field through a table of input pointers (batches of
// separately allocated boxes); outputs are contiguous
template <int N>
__global__ __launch_bounds__(kNT) void k_r2c_z_tab(const float* const* __restrict__ in_tab, float2* __restrict__ out,
                                                  const float2* __restrict__ twN, long out_fstride)
{
  r2c_z_tile<N>(in_tab[blockIdx.y], out + blockIdx.y * out_fstride, twN);
}

// ---------------------------------------------------- z: C2R load + FFT
// Untwists ROWS packed rows held in LDS and runs the inverse length-N/2 FFT
// on all of them with NT threads; afterwards the tile, read as floats, holds
// ROWS real rows of N.
template <int N, int ROWS, int NT>
__device__ __forceinline__ void c2r_untwist_rows(float2* s, const float2* tw, int tid)
{
  constexpr int NH = N / 2;
  constexpr int NP = NH / 2 + 1;
  for (int q = tid; q < ROWS * NP; q += NT) {
    const int r = q / NP, k = q % NP;
    const int rb = r * NH;
    if (k == 0) {
      const float2 p = s[lds_addr<true>(rb)];  // (X_0, X_{N/2}), both real
      s[lds_addr<true>(rb)] = make_float2(p.x + p.y, p.x - p.y);
    } else {
      const float2 xk = s[lds_addr<true>(rb + k)], xm = cconj(s[lds_addr<true>(rb + NH - k)]);
      const float2 E = cadd(xk, xm);
      const float2 O = cmul(csub(xk, xm), cconj(tw[k]));  // * exp(+2 pi i k / N)
      s[lds_addr<true>(rb + k)] = make_float2(E.x - O.y, E.y + O.x);  // E + i O
      if (k != NH - k) s[lds_addr<true>(rb + NH - k)] = make_float2(E.x + O.y, -E.y + O.x);  // conj(E) + i conj(O)
    }
  }
  __syncthreads();
  fft_lds<NH, ROWS, NT, true, 1, NH, false, 2, true>(s, tw, tid);
}

template <int N>
__device__ __forceinline__ void c2r_untwist_fft(float2* s, const float2* tw, int tid)
{
  c2r_untwist_rows<N, Cfg<N>::RZ, kNT>(s, tw, tid);
}

// Loads RZ packed rows, untwists them and runs the inverse length-N/2 FFT in
// LDS; afterwards real cell e of the tile is lds_real<N>(s, e).  Slots kz >=
// kzlive are read as zero (not loaded).
template <int N>
__device__ __forceinline__ void c2r_rows_lds(float2* s, const float2* tw, const float2* __restrict__ src, long row0,
                                             int tid, int kzlive)
{
  using C = Cfg<N>;
  constexpr int NH = C::NH, ZE = C::ZE;
  constexpr int EPZ = ZE / kNT;
  float2 v[EPZ];
#pragma unroll
  for (int i = 0; i < EPZ; ++i) {
    const int e = tid + i * kNT;
    v[i] = (e % NH < kzlive) ? src[row0 * NH + e] : make_float2(0.0f, 0.0f);
  }
#pragma unroll
  for (int i = 0; i < EPZ; ++i) s[lds_addr<true>(tid + i * kNT)] = v[i];
  __syncthreads();
  c2r_untwist_fft<N>(s, tw, tid);
}

__device__ __forceinline__ float lds_real(const float2* s, int e)
{
  const float2 v = s[lds_addr<true>(e >> 1)];
  return (e & 1) ? v.y : v.x;
}

// aliasing clamps of find_HII_bubbles.c:272-311.  Kinds 0-2 are written as
// compares and selects with fmaxf/fminf's results for every input, NaN
// included (fmaxf(NaN, c) = c, and a compare with NaN is false): a max of a
// value loaded from memory costs a canonicalising max before it, a select
// does not.  Kind 1 is max(v, 0) then v < ABS_TOL -> 0, which is v >= ABS_TOL
// ? v : 0.
__device__ __forceinline__ float clamp_kind(float v, int kind)
{
  switch (kind) {
    case 0: {
      constexpr float lo = -1.0f + kRelTol;
      return (v > lo) ? v : lo;
    }
    case 1:
      return (v >= kAbsTol) ? v : 0.0f;
    case 2:
      return (v >= kAbsTol) ? ((v < 0.999f) ? v : 0.999f) : 0.0f;
    case 3:  // spin-temperature shells: max(v, 0) only (ComputeTs.c:417, :488)
      return fmaxf(v, 0.0f);
    default:
      return v;
  }
}

// Plain C2R z pass: packed rows -> padded real rows, with an optional clamp;
// slots kz >= kzlive of the input are zero.
template <int N>
__global__ __launch_bounds__(kNT) void k_c2r_z(const float2* __restrict__ in, float* __restrict__ out,
                                              const float2* __restrict__ twN, long in_fstride, long out_fstride,
                                              int clamp, int pitch, int kzlive)
{
  using C = Cfg<N>;
  constexpr int RZ = C::RZ, ZE = C::ZE;
  __shared__ float2 s[ZE + ZE / 32];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += kNT) tw[i] = twN[i];
  const long row0 = (long)blockIdx.x * RZ;
  c2r_rows_lds<N>(s, tw, in + blockIdx.y * in_fstride, row0, tid, kzlive);
  float* o = out + blockIdx.y * out_fstride;
  if (pitch == N) {
    // unpadded rows: each thread stores groups of 4 cells (two complex slots
    // of one 32-slot LDS bank row) as one 16-byte write
    constexpr int Q = N / 4;
    static_assert((RZ * Q) % kNT == 0, "4-cell groups must split evenly over the threads");
    float4* o4 = reinterpret_cast<float4*>(o + row0 * N);
#pragma unroll
    for (int i = 0; i < RZ * Q / kNT; ++i) {
      const int q = tid + i * kNT;
      const int c0 = (q / Q) * (N / 2) + 2 * (q % Q);
      const float2 a = s[lds_addr<true>(c0)], b = s[lds_addr<true>(c0 + 1)];
      o4[q] = make_float4(clamp_kind(a.x, clamp), clamp_kind(a.y, clamp), clamp_kind(b.x, clamp),
                          clamp_kind(b.y, clamp));
    }
    return;
  }
  for (int e = tid; e < RZ * pitch; e += kNT) {
    const int r = e / pitch, z = e % pitch;
    o[(row0 + r) * pitch + z] = (z < N) ? clamp_kind(lds_real(s, r * N + z), clamp) : 0.0f;
  }
}

// --------------------------------------------------- strided pencil passes
// Thread map of an N x TW tile: thread t owns column t % TW and rows
// t / TW + i * RPI, i < EPT.  Every load of a tile is issued before the first
// is consumed, so a wave keeps EPT row segments in flight.
template <int N, int TW_, int MAXNT = kColNT>
struct TileMap {
  static constexpr int TW = TW_;
  static constexpr int NT_RAW = TW * N / 8;
  static constexpr int NT = NT_RAW < 64 ? 64 : (NT_RAW > MAXNT ? MAXNT : NT_RAW);
  static constexpr int SP = TW + 2;  // LDS row pitch, breaks the bank pattern
  static constexpr int RPI = NT / TW;
  static constexpr int EPT = N / RPI;
  static constexpr int NKZB = (N / 2) / TW;  // kz blocks per line
  static_assert((N / 2) % TW == 0, "tile width must divide N/2");
  static_assert(NT % TW == 0 && N % RPI == 0, "tile must split evenly over the threads");
};

// One N x TW tile per workgroup: pencils along an axis with element stride
// `rstride`; tile t covers kz block (t mod NKZB) of line (t / NKZB), line
// stride astride; blockIdx.y selects the field (stride fstride).
template <int N, bool INV, int TW = Cfg<N>::TW, int SV = 0, int MAXNT = kColNT>
__global__ __launch_bounds__((TileMap<N, TW, MAXNT>::NT), 4) void k_colpass(float2* __restrict__ data,
                                                                   const float2* __restrict__ twN, long fstride,
                                                                   long astride, long rstride, double div)
{
  using M = TileMap<N, TW, MAXNT>;
  constexpr int NT = M::NT, SP = M::SP, RPI = M::RPI, EPT = M::EPT, NKZB = M::NKZB;
  __shared__ float2 s[N * SP];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  const int c = tid % TW, n0 = tid / TW;
  const int a = blockIdx.x / NKZB, b = blockIdx.x % NKZB;
  const unsigned rs = (unsigned)rstride;
  float2* g = data + blockIdx.y * fstride + (long)a * astride + (long)b * TW + (unsigned)n0 * rs + c;
  float2 r[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) r[i] = g[(unsigned)(i * RPI) * rs];
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  fft_tile_reg<N, TW, NT, INV, SP, 1, SV>(r, s, tw, tid);
  if (div > 0.0) {  // complex float / double, in double (find_HII_bubbles.c:168-182)
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = make_float2((float)((double)r[i].x / div), (float)((double)r[i].y / div));
  } else if (div < 0.0) {  // complex float / float (ComputeTs.c:383-384)
    const float fd = (float)(-div);
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = make_float2(r[i].x / fd, r[i].y / fd);
  }
#pragma unroll
  for (int i = 0; i < EPT; ++i) store_nt(g + (unsigned)(i * RPI) * rs, r[i]);
}

// k_colpass over a subset of its tiles: kz blocks below nkzb only (grid
// nline x nkzb per field, tile t = line t / nkzb, kz block t % nkzb), and
// with lmax >= 0 only the lines ky with min(ky, N - ky) <= lmax (line j is ky
// = j for j <= lmax, else N - (nline - j), nline = min(N, 2 lmax + 1)).  Each
// tile is transformed exactly as by k_colpass; the others are left as they
// are.  The forward passes of a spectrum that a sharp-k window will cut to
// the ball of modes |n_x|, |n_y|, n_z <= nmax need only these tiles.
template <int N, bool INV, int TW = Cfg<N>::TW>
__global__ __launch_bounds__((TileMap<N, TW>::NT), 4) void k_colpass_sub(float2* __restrict__ data,
                                                                   const float2* __restrict__ twN, long fstride,
                                                                   long astride, long rstride, double div, int nkzb,
                                                                   int lmax)
{
  using M = TileMap<N, TW>;
  constexpr int NT = M::NT, SP = M::SP, RPI = M::RPI, EPT = M::EPT;
  __shared__ float2 s[N * SP];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  const int c = tid % TW, n0 = tid / TW;
  const int j = blockIdx.x / nkzb, b = blockIdx.x % nkzb;
  int a = j;
  if (lmax >= 0) {
    const int nline = min(N, 2 * lmax + 1);
    a = (j <= lmax) ? j : N - (nline - j);
  }
  const unsigned rs = (unsigned)rstride;
  float2* g = data + blockIdx.y * fstride + (long)a * astride + (long)b * TW + (unsigned)n0 * rs + c;
  float2 r[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) r[i] = g[(unsigned)(i * RPI) * rs];
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  fft_tile_reg<N, TW, NT, INV, SP>(r, s, tw, tid);
  if (div > 0.0) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = make_float2((float)((double)r[i].x / div), (float)((double)r[i].y / div));
  } else if (div < 0.0) {
    const float fd = (float)(-div);
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = make_float2(r[i].x / fd, r[i].y / fd);
  }
#pragma unroll
  for (int i = 0; i < EPT; ++i) g[(unsigned)(i * RPI) * rs] = r[i];
}

// Window table of one radius: W depends on (|nx|, |ny|, nz) only, because
// filter() builds each k component as n*dk or (n-N)*dk and squares it, so
// modes n and N-n give the same float k^2.  Layout for tile width TW (kz
// blocks of TW): W(a, b, c) for c < N/2 at ((b * NKZB + c / TW) * NH1 + a) *
// TW + c % TW, so the x-pass tile (|ky| = b, kz block) reads one contiguous
// NH1 x TW block; the Nyquist column c = N/2 follows at NH1*(N/2)*NH1 + b*NH1 + a.
//
// One workgroup per contiguous block of the table: (b, kz block) for c < N/2,
// NH1 x TW values, or row b of the Nyquist column, NH1 values; the values are
// stored in table order, so the writes are contiguous.  Every (a, b) is
// evaluated (the mirror (b, a) gives the same float, as k^2 sums kx^2 and
// ky^2 first).
template <int N, int TW>
constexpr unsigned window_table_blocks()
{
  return (unsigned)((N / 2 + 1) * ((N / 2) / TW) + (N / 2 + 1));
}

template <int N, int TW>
__global__ __launch_bounds__(kNT) void k_window_table(float* __restrict__ Wt, WinArgs w)
{
  constexpr int NH = N / 2, NH1 = NH + 1, NKZB = NH / TW;
  const int blk = blockIdx.x;
  if (blk < NH1 * NKZB) {
    const int b = blk / NKZB, kb = blk % NKZB;
    float* out = Wt + (long)blk * NH1 * TW;
    for (int e = threadIdx.x; e < NH1 * TW; e += kNT) out[e] = window(e / TW, b, kb * TW + e % TW, N, w);
  } else {
    const int b = blk - NH1 * NKZB;
    float* out = Wt + (long)NH1 * NH * NH1 + (long)b * NH1;
    for (int a = threadIdx.x; a < NH1; a += kNT) out[a] = window(a, b, NH, N, w);
  }
}

// Window modes of the inverse x pass.
enum { kWinNone = 0, kWinTable = 1, kWinSharpK = 2 };

// Inverse x pass with the window applied on load, for every field of the
// ladder step.  U: unfiltered packed spectra (kz=0 slot = true kz=0 plane);
// V: their Nyquist planes [ky][kx]; T: output packed spectra.  WIN selects the
// window: none (the unfiltered last step), the radius's table Wt, or the
// sharp-k window evaluated in place (wa), whose modes with any |component| >
// nmax are zero: then only the tiles with |ky| <= nmax and kz block < nkzb are
// launched (grid (min(N, 2 nmax + 1)) x nkzb), zero modes are stored as 0
// without being read (filter() assigns 0 to them), and T's kz blocks >= nkzb
// and rows |ky| > nmax are left unwritten for the passes that follow to read
// as zero.  The tile's window values are loaded once for all fields, and the
// next field's tile is loaded while the current one is transformed.
// blockIdx.y selects a box of a batch (strides bstride for U/T, bvstride for V).
template <int N, int WIN, int TW = Cfg<N>::TW, int SV = 0>
__global__ __launch_bounds__((TileMap<N, TW, ix_max_threads(N)>::NT)) void k_ix_filter(const float2* __restrict__ U,
                                                                 const float2* __restrict__ V,
                                                                 float2* __restrict__ T,
                                                                 const float2* __restrict__ twN,
                                                                 const float* __restrict__ Wt, long fstride,
                                                                 long vstride, int nfield, long bstride,
                                                                 long bvstride, WinArgs wa, int nmax, int nkzb)
{
  using M = TileMap<N, TW, ix_max_threads(N)>;  // the window values stay in registers
  constexpr int NT = M::NT, SP = M::SP, RPI = M::RPI, EPT = M::EPT, NKZB = M::NKZB;
  constexpr int NH = N / 2, NH1 = NH + 1;
  constexpr unsigned rs = (unsigned)N * NH;
  __shared__ float2 s[N * SP];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  const int c = tid % TW, n0 = tid / TW;
  int ky, b;
  if constexpr (WIN == kWinSharpK) {
    const int j = blockIdx.x / nkzb, nky = min(N, 2 * nmax + 1);
    b = blockIdx.x % nkzb;
    ky = (j <= nmax) ? j : N - (nky - j);
  } else {
    ky = blockIdx.x / NKZB;
    b = blockIdx.x % NKZB;
  }
  const int aky = ky <= NH ? ky : N - ky;
  U += blockIdx.y * bstride;
  T += blockIdx.y * bstride;
  V += blockIdx.y * bvstride;
  const long base = (long)ky * NH + (long)b * TW + (unsigned)n0 * rs + c;
  float2 r[EPT];
  float w[WIN != kWinNone ? EPT : 1];
  if constexpr (WIN == kWinSharpK) {
    // window first, then only the live modes are read
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int n = n0 + i * RPI;
      const int an = n <= NH ? n : N - n;
      w[i] = (an <= nmax) ? sharpk_window(k_magnitude(an, aky, b * TW + c, N, wa.delta_k), wa.R) : 0.0f;
    }
    const float2* u = U + base;
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = (w[i] != 0.0f) ? u[(unsigned)(i * RPI) * rs] : make_float2(0.0f, 0.0f);
  } else {
    const float2* u = U + base;
#pragma unroll
    for (int i = 0; i < EPT; ++i) r[i] = u[(unsigned)(i * RPI) * rs];
  }
  // this tile's block of the window table (k_window_table's layout)
  if constexpr (WIN == kWinTable) {
    const float* wr = Wt + ((long)aky * NKZB + b) * NH1 * TW + c;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int n = n0 + i * RPI;
      const int an = n <= NH ? n : N - n;
      w[i] = wr[(unsigned)an * TW];
    }
  }
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  // the Nyquist plane's window is nonzero somewhere on this tile row
  const bool nyq = (b == 0) && (WIN != kWinSharpK || (nmax >= NH && aky <= nmax));
  for (int f = 0; f < nfield; ++f) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) s[(n0 + i * RPI) * SP + c] = (WIN == kWinTable) ? cscale(r[i], w[i]) : r[i];
    if (nyq) {
      // the kz=0 slot carries A*W0 + i*(B*WN), B the Nyquist plane
      __syncthreads();
      const float2* vp = V + f * vstride + (long)ky * N;
      for (int n = tid; n < N; n += NT) {
        const int an = n <= NH ? n : N - n;
        float2 nb;
        if constexpr (WIN == kWinSharpK) {
          if (an > nmax || sharpk_window(k_magnitude(an, aky, NH, N, wa.delta_k), wa.R) == 0.0f) continue;
          nb = vp[n];
        } else {
          const float wn = (WIN == kWinTable) ? Wt[(long)NH1 * NH * NH1 + (long)aky * NH1 + an] : 1.0f;
          nb = cscale(vp[n], wn);
        }
        const float2 v = s[n * SP];
        s[n * SP] = make_float2(v.x - nb.y, v.y + nb.x);
      }
    }
    __syncthreads();
    if (f + 1 < nfield) {
      const float2* u = U + (f + 1) * fstride + base;
#pragma unroll
      for (int i = 0; i < EPT; ++i) {
        if constexpr (WIN == kWinSharpK)
          r[i] = (w[i] != 0.0f) ? u[(unsigned)(i * RPI) * rs] : make_float2(0.0f, 0.0f);
        else
          r[i] = u[(unsigned)(i * RPI) * rs];
      }
    }
    fft_lds<N, TW, NT, true, SP, 1, true, 1, false, SV>(s, tw, tid);
    float2* t = T + f * fstride + base;
#pragma unroll
    for (int i = 0; i < EPT; ++i) t[(unsigned)(i * RPI) * rs] = s[(n0 + i * RPI) * SP + c];
    __syncthreads();
  }
}

// Inverse x pass of K consecutive ladder radii at once, for the tabulated
// windows (top-hat, Gaussian): for each field, each radius k < K windows the
// field's tile of the unfiltered spectrum U with table k (Wt + k wstride,
// k_window_table's layout), transforms it and writes it to T + k tstride.  Per
// radius the result is k_ix_filter<N, kWinTable>'s, operation for operation.
//
// One workgroup per (tile, radius, field), grid tiles x K x nfield.  The K
// workgroups of a (tile, field) read the same U tile and the nfield of a
// (tile, radius) the same window block, so all K nfield of a tile are made to
// run side by side on one XCD, whose L2 then serves every read but the first:
// workgroups are dealt to the 8 XCDs round robin (workgroup g to XCD g mod
// 8), so with the tiles a multiple of 8, g = ((j K + k) nfield + f) 8 + x
// gives tile 8 j + x, radius k, field f, and a tile's workgroups are dealt
// consecutively to XCD x.  (A workgroup that looped over the K radii itself
// re-read its tile after a transform and a store of the whole tile: by then
// 64 workgroups per XCD had streamed 8 MiB through its 4 MiB L2, and the
// repeat missed.  One that looped over the fields took 1650 us per
// two-radius launch at 512^3 against 1603 us for one field per workgroup.)
// The outputs go out with streaming stores, which keep them from evicting the
// tiles.
template <int N, int TW = Cfg<N>::TW, int MAXNT = kColNT>
__global__ __launch_bounds__((TileMap<N, TW, MAXNT>::NT)) __attribute__((amdgpu_waves_per_eu(4))) void k_ix_filter_k(
  const float2* __restrict__ U, const float2* __restrict__ V, float2* __restrict__ T,
  const float2* __restrict__ twN, const float* __restrict__ Wt, long wstride, long fstride, long vstride,
  long tstride, int nfield, int K)
{
  using M = TileMap<N, TW, MAXNT>;
  constexpr int NT = M::NT, SP = M::SP, RPI = M::RPI, EPT = M::EPT, NKZB = M::NKZB;
  constexpr int NH = N / 2, NH1 = NH + 1;
  constexpr unsigned rs = (unsigned)N * NH;
  __shared__ float2 s[N * SP];
  __shared__ float2 tw[N];
  constexpr int kTiles = N * NKZB, kXcd = 8;
  const int tid = threadIdx.x;
  const int c = tid % TW, n0 = tid / TW;
  // one workgroup per (tile, radius, field), grid tiles x K x nfield: the K
  // nfield workgroups of a tile (the same window block per radius, the same
  // spectrum tile per field) are dealt to one XCD back to back
  int tile, k, f;
  if constexpr (kTiles % kXcd == 0) {
    const int slot = blockIdx.x / kXcd;  // < kTiles K nfield / 8, so tile < kTiles
    f = slot % nfield;
    k = (slot / nfield) % K;
    tile = (slot / (nfield * K)) * kXcd + blockIdx.x % kXcd;
  } else {
    f = blockIdx.x % nfield;
    k = (blockIdx.x / nfield) % K;
    tile = blockIdx.x / (nfield * K);
  }
  const int ky = tile / NKZB, b = tile % NKZB;
  const int aky = ky <= NH ? ky : N - ky;
  const long base = (long)ky * NH + (long)b * TW + (unsigned)n0 * rs + c;
  const float* wr = Wt + k * wstride + ((long)aky * NKZB + b) * NH1 * TW + c;
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  const bool nyq = (b == 0);  // this tile holds the kz=0 slot, which carries the Nyquist plane
  {
    {
      // the tile and this radius's window, written to LDS as the loads arrive
      // (measured: keeping the windowed tile in registers for a first
      // transform stage there, as k_colpass does, made this kernel 36 % slower;
      // so did loading the next field's tile while this one is transformed:
      // 1643 -> 1900 us per launch at 512^3)
      const float2* u = U + f * fstride + base;
      float2 r[EPT];
#pragma unroll
      for (int i = 0; i < EPT; ++i) r[i] = u[(unsigned)(i * RPI) * rs];
#pragma unroll
      for (int i = 0; i < EPT; ++i) {
        const int n = n0 + i * RPI;
        const int an = n <= NH ? n : N - n;
        s[n * SP + c] = cscale(r[i], wr[(unsigned)an * TW]);
      }
    }
    if (nyq) {
      // the kz=0 slot carries A*W0 + i*(B*WN), B the Nyquist plane
      __syncthreads();
      const float2* vp = V + f * vstride + (long)ky * N;
      const float* wn = Wt + k * wstride + (long)NH1 * NH * NH1 + (long)aky * NH1;
      for (int n = tid; n < N; n += NT) {
        const int an = n <= NH ? n : N - n;
        const float2 nb = cscale(vp[n], wn[an]);
        const float2 v = s[n * SP];
        s[n * SP] = make_float2(v.x - nb.y, v.y + nb.x);
      }
    }
    __syncthreads();
    fft_lds<N, TW, NT, true, SP, 1, true, 1, false, 0>(s, tw, tid);
    float2* t = T + k * tstride + f * fstride + base;
#pragma unroll
    for (int i = 0; i < EPT; ++i) store_nt(t + (unsigned)(i * RPI) * rs, s[(n0 + i * RPI) * SP + c]);
    __syncthreads();
  }
}

// Inverse pass over the live tiles of a sharp-k filtered spectrum left by
// k_ix_filter<kWinSharpK>: kz blocks below nkzb only (grid N x nkzb per
// field, tile t = line t / nkzb, kz block t % nkzb), and rows n with
// min(n, N - n) > nlive read as zero.  Otherwise as k_colpass<N, true>.
template <int N, int TW = Cfg<N>::TW>
__global__ __launch_bounds__((TileMap<N, TW>::NT), 4) void k_colpass_live(float2* __restrict__ data,
                                                                    const float2* __restrict__ twN, long fstride,
                                                                    long astride, long rstride, int nkzb,
                                                                    int nlive)
{
  using M = TileMap<N, TW>;
  constexpr int NT = M::NT, SP = M::SP, RPI = M::RPI, EPT = M::EPT;
  constexpr int NH = N / 2;
  __shared__ float2 s[N * SP];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  const int c = tid % TW, n0 = tid / TW;
  const int a = blockIdx.x / nkzb, b = blockIdx.x % nkzb;
  const unsigned rs = (unsigned)rstride;
  float2* g = data + blockIdx.y * fstride + (long)a * astride + (long)b * TW + (unsigned)n0 * rs + c;
  float2 r[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int n = n0 + i * RPI;
    const int an = n <= NH ? n : N - n;
    r[i] = (an <= nlive) ? g[(unsigned)(i * RPI) * rs] : make_float2(0.0f, 0.0f);
  }
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  fft_tile_reg<N, TW, NT, true, SP>(r, s, tw, tid);
#pragma unroll
  for (int i = 0; i < EPT; ++i) g[(unsigned)(i * RPI) * rs] = r[i];
}

// After the forward x pass the kz=0 slot holds P = A + iB (A, B the kz=0 and
// kz=N/2 planes).  Split it using A(-k) = conj A(k), B(-k) = conj B(k):
// A = (P(k) + conj P(-k))/2, B = (P(k) - conj P(-k))/(2i).
template <int N>
__global__ void k_unpack_dc(float2* __restrict__ U, float2* __restrict__ V, long fstride, long vstride)
{
  constexpr int NH = N / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * N) return;
  const int kx = idx / N, ky = idx % N;
  const int mx = (N - kx) % N, my = (N - ky) % N;
  const int midx = mx * N + my;
  if (idx > midx) return;  // the pair is handled by its canonical member
  float2* u = U + blockIdx.y * fstride;
  float2* v = V + blockIdx.y * vstride;
  const float2 P = u[(long)idx * NH], Q = u[(long)midx * NH];
  const float2 qc = cconj(Q);
  const float2 A = cscale(cadd(P, qc), 0.5f);
  const float2 d = csub(P, qc);
  const float2 B = make_float2(0.5f * d.y, -0.5f * d.x);
  u[(long)idx * NH] = A;
  v[ky * N + kx] = B;
  if (midx != idx) {
    u[(long)midx * NH] = cconj(A);
    v[my * N + mx] = cconj(B);
  }
}

// --------------------------------------------------- layout conversions
// packed spectrum (+ separate Nyquist plane) <-> INDEX_COMPLEX_HERM
template <int N>
__global__ void k_packed_to_herm(const float2* __restrict__ U, const float2* __restrict__ V, float2* __restrict__ H)
{
  constexpr int NH = N / 2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)N * N * (NH + 1)) return;
  const long row = idx / (NH + 1);
  const int kz = (int)(idx % (NH + 1));
  const int x = (int)(row / N), y = (int)(row % N);
  H[idx] = (kz < NH) ? U[row * NH + kz] : V[(long)y * N + x];
}

template <int N>
__global__ void k_herm_to_packed(const float2* __restrict__ H, float2* __restrict__ U)
{
  constexpr int NH = N / 2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)N * N * NH) return;
  const long row = idx / NH;
  const int kz = (int)(idx % NH);
  if (kz == 0) {
    const float2 A = H[row * (NH + 1)], B = H[row * (NH + 1) + NH];
    U[idx] = make_float2(A.x - B.y, A.y + B.x);  // A + iB
  } else {
    U[idx] = H[row * (NH + 1) + kz];
  }
}

// filter() on a reference-layout half spectrum slab, in place.
__global__ void k_filter_herm(float2* __restrict__ H, int local_ix_start, int slab_nx, int N, WinArgs w)
{
  const int NH1 = N / 2 + 1;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)slab_nx * N * NH1) return;
  const int nz = (int)(idx % NH1);
  const long r = idx / NH1;
  const int ny = (int)(r % N), nx = (int)(r / N) + local_ix_start;
  const float wv = window(nx, ny, nz, N, w);
  if (wv != 1.0f) H[idx] = cscale(H[idx], wv);
}

// ------------------------------------------------------------ reductions
template <int NTH>
__device__ __forceinline__ double block_sum(double v, double* red)
{
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int st = NTH / 2; st > 0; st >>= 1) {
    if (tid < st) red[tid] += red[tid + st];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Block sums of NV values at once: a butterfly over each wave's 64 lanes
// (lane i adds lane i^m; addition commutes, so every lane holds the same
// bits), then the NT/64 wave sums in wave order.  Two barriers in all.  On
// return every thread holds the block's sums in v; red holds NV * NT / 64
// doubles.
template <int NT, int NV>
__device__ __forceinline__ void block_sum_n(double (&v)[NV], double* red)
{
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v[k] += __shfl_xor(v[k], m, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wave * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = red[k];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) t += red[w * NV + k];
    v[k] = t;
  }
  __syncthreads();
}

// partials[nb][nv] -> out[nv], fixed order (deterministic); block b reduces
// the b-th such set (one box of a batch)
__global__ __launch_bounds__(kNT) void k_reduce_partials(const double* __restrict__ partials, int nb, int nv,
                                                       double* __restrict__ out)
{
  __shared__ double red[kNT];
  partials += (long)blockIdx.x * nb * nv;
  out += (long)blockIdx.x * nv;
  for (int v = 0; v < nv; ++v) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kNT) acc += partials[(long)i * nv + v];
    acc = block_sum<kNT>(acc, red);
    if (threadIdx.x == 0) out[v] = acc;
  }
}

// First level of a large reduction: workgroup (b, box) sums rows
// [b chunk, (b + 1) chunk) of the box's nb x nv row-major partials (boxes
// contiguous) into row b of the box's block of `out` (gridDim.x rows of nv).
// Every lane reads whole rows; nv <= kRedMaxV.  The partition depends only on
// (nb, gridDim.x), so the sums are the same on every run.
constexpr int kRedMaxV = 8;
constexpr int kRedBlocks = 128;
__global__ __launch_bounds__(kNT) void k_reduce_rows(const double* __restrict__ partials, int nb, int nv, int chunk,
                                                   double* __restrict__ out)
{
  __shared__ double red[kNT];
  partials += (long)blockIdx.y * nb * nv;
  out += ((long)blockIdx.y * gridDim.x + blockIdx.x) * nv;
  const int r0 = blockIdx.x * chunk;
  const int r1 = min(nb, r0 + chunk);
  double acc[kRedMaxV];
#pragma unroll
  for (int v = 0; v < kRedMaxV; ++v) acc[v] = 0.0;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += kNT) {
    const double* row = partials + (long)r * nv;
#pragma unroll
    for (int v = 0; v < kRedMaxV; ++v)
      if (v < nv) acc[v] += row[v];
  }
#pragma unroll
  for (int v = 0; v < kRedMaxV; ++v) {
    if (v >= nv) break;
    const double t = block_sum<kNT>(acc[v], red);
    if (threadIdx.x == 0) out[v] = t;
  }
}

// ------------------------------------------- fused C2R + bubble threshold
struct StepArgs {
  double M_mean;        // RtoM(R) (find_HII_bubbles.c:339)
  double R_cubed;       // R*R*R
  double pixel_volume;  // (BoxSize/N)^3
  double J21c;          // J_21_aux_constant (:317-320)
  double eff;           // ReionEfficiency
  double redshift;
  double bias;          // ReionGammaHaloBias
  double kfc;           // (4/3) pi R^3 / (pixel_volume M_mean): f_coll = stars * kfc / (1 + delta)
  double gamma_pref;    // Gamma_R_prefactor(R) of the recombinations (:325-334)
  int step;             // ladder index (0 = largest R)
  int last;             // unfiltered final step
  int uvb;
  int use_xe;
  int nfield;
  int kzlive;           // spectrum slots kz >= kzlive are zero (sharp-k)
};

// One ladder step of _find_HII_bubbles (find_HII_bubbles.c:255-443) for RZ
// rows: inverse z pass of every field, clamp, then the per-cell test in double
// exactly as the reference writes it.  Per-cell state between steps is one
// byte (0xFF = never crossed, else the step of the latest crossing); xH,
// r_bubble and the J_21 zeros are written once, on the last step, and the
// first crossing writes J_21 / z_at_ionization / J_21_at_ionization directly.
// Threads and 4-cell groups of the fused z pass + threshold: all NF fields of
// RZ rows sit in LDS together and are transformed as NF*RZ pencils, so the
// fields share every barrier of the transform.
// Power-of-two grids take 512 complex elements of rows per field (16 KiB of
// LDS for four fields: measured at 512^3, 512 beats 1024 by 10% and 2048 by
// 38%, more resident workgroups hiding the transform's barriers); the others
// keep the z pass's row blocks.
template <int N>
struct ThrCfg {
  static constexpr int NH = N / 2;
  static constexpr bool P2 = (Cfg<N>::ZE % 1024 == 0) && NH <= 512;
  // Row blocks longer than kZE elements (2560 for a factor 5 in NH) would take
  // 80 KiB of LDS for four fields: those grids run one wave per workgroup over
  // the fewest rows whose elements split into 4-cell groups over 64 threads.
  static constexpr bool WAVE = !P2 && Cfg<N>::ZE > kZE;
  static constexpr int RZ = P2 ? 512 / NH : (WAVE ? z_rows_wave(NH) : Cfg<N>::RZ);
  static constexpr int ZE = RZ * NH;
  static constexpr int NT = WAVE ? 64 : 256;
  static constexpr int G = ZE / (2 * NT);  // groups of 4 cells per thread
  static_assert(ZE % (2 * NT) == 0, "z tile must split into 4-cell groups");
  static_assert((N * N) % RZ == 0, "threshold workgroups must tile the rows");
};

// The ladder step's test of one cell (find_HII_bubbles.c:313-443).  v: the
// cell's clamped filtered fields (deltax, stars, weighted_sfr, then x_e when
// there are four besides N_rec, then N_rec when REC); st: its state byte.  The
// first crossing writes the cell's J_21 (and Gamma_12 with REC); on the last
// step xh, rb and jflag (1 where the cell's J_21 is kept) are set, and zcheck
// says whether z_at_ionization is to be set if still unset (:427-434), with
// J_21_at_ionization = jaux * bias.  The reference sets them at the cell's
// first crossing, to the snapshot's redshift and that crossing's J_21; the
// same values are set here at the last step, where every cell that crossed
// during the snapshot is ionised and J21 still holds its first crossing's
// J_21: for those (jflag = 1) the caller takes jaux from J21, read after this
// call.  Deferring the check keeps a dependent load of z_at_ionization out of
// every ladder step.  Returns the new state byte.
template <int NF, bool REC>
__device__ __forceinline__ uint32_t threshold_cell(const float* v, uint32_t st, long ir, const StepArgs& a,
                                                   float* __restrict__ J21, float* __restrict__ Gamma12,
                                                   const float* __restrict__ Rtab, float& xh, float& rb,
                                                   float& jflag, bool& zcheck, float& jaux)
{
#pragma clang fp contract(off)
  constexpr bool XE = (NF - (REC ? 1 : 0)) == 4;
  constexpr int RI = XE ? 4 : 3;  // field of N_rec
  // the double quantities are formed in the branches that use them: the
  // single-precision test below decides almost every cell without them
  auto density_over_mean = [&]() { return 1.0 + (double)v[0]; };
  auto neutral_fraction = [&]() { return XE ? 1.0 - (double)v[3] : 1.0; };
  // recombinations in the cell (:360-363)
  auto rec = [&]() { return REC ? (double)v[RI] / density_over_mean() : 0.0; };
  // the reference's f_coll (four roundings of its exact value), evaluated
  // where needed: the last step's partial xH and near-threshold cells
  auto f_coll = [&]() {
    return (double)v[1] / (a.M_mean * density_over_mean()) * (4.0 / 3.0) * M_PI * a.R_cubed / a.pixel_volume;
  };
  // the test f_coll*eff > n_f as stars*kfc*eff vs n_f*(1+delta), both
  // positive sides within 1e-14 of their exact values: a relative gap
  // above 1e-12 fixes the reference's outcome, closer cells take its
  // own expression
  bool cross, decided = false;
  if constexpr (!REC) {
    // the same test in single precision first: lf and rf lie within 2e-7
    // (relative) of the exact sides, so a gap above the band's 9.5e-7 gives
    // the double test's outcome; cells inside the band take the double test
    const float kfe = (float)(a.kfc * a.eff);
    const float lf = v[1] * kfe, rf = (XE ? 1.0f - v[3] : 1.0f) * (1.0f + v[0]);
    if (lf > rf * 1.000001f) {
      cross = true;
      decided = true;
    } else if (lf < rf * 0.999999f) {
      cross = false;
      decided = true;
    }
  }
  if (!decided) {
    const double lhs = (double)v[1] * a.kfc * a.eff, rhs = neutral_fraction() * (1. + rec()) * density_over_mean();
    if (lhs > rhs * (1.0 + 1e-12))
      cross = true;
    else if (lhs < rhs * (1.0 - 1e-12))
      cross = false;
    else
      cross = f_coll() * a.eff > neutral_fraction() * (1. + rec());
  }
  // the weighted SFR is clamped here, where it is used (the clamp is
  // idempotent, so a caller may pass it clamped or not)
  auto sfr = [&]() { return clamp_kind(v[2], 1); };
  // J_21 of the cell at this radius (:317-320, :331-337)
  auto j21_aux = [&]() { return a.uvb ? (float)((double)sfr() / a.pixel_volume * a.J21c) : 0.0f; };
  const bool never = (st == 0xFFu);
  zcheck = false;
  jaux = 0.0f;
  uint32_t st_new = st;
  if (cross) {
    if (never) {
      if (a.uvb) J21[ir] = j21_aux();
      if (REC) Gamma12[ir] = (float)(a.gamma_pref * ((double)sfr() / a.pixel_volume));  // :398-404
    }
    st_new = (uint32_t)a.step;
  }
  if (a.last) {
    const bool ion = cross || !never;
    if (ion) {
      xh = 0.0f;
    } else {
      xh = (float)(neutral_fraction() - f_coll() * a.eff);
      if (xh < 0.f)
        xh = 0.f;
      else if (xh > 1.0f)
        xh = 1.f;
      if (xh < kRelTol) jaux = j21_aux();
    }
    zcheck = xh < kRelTol;  // every ionised cell, and the partial ones below REL_TOL
    rb = ion ? Rtab[cross ? a.step : (int)st] : 0.0f;
    jflag = ion ? 1.0f : 0.0f;
  }
  return st_new;
}

// Fields in T: deltax, stars, weighted_sfr, then x_e when XE (NF - REC = 4),
// then N_rec when REC (inhomogeneous recombinations).
template <int N, int NF, bool REC>
__device__ __forceinline__ void c2r_threshold_tile(const float2* __restrict__ T, long fstride,
                                                   const float2* __restrict__ twN,
                                                   const float* __restrict__ deltax_in,
                                                   uint8_t* __restrict__ state, float* __restrict__ xH,
                                                   float* __restrict__ r_bubble, float* __restrict__ J21,
                                                   float* __restrict__ z_in, float* __restrict__ J21_ion,
                                                   float* __restrict__ Gamma12, const float* __restrict__ Rtab,
                                                   double* __restrict__ partials, const StepArgs& a)
{
#pragma clang fp contract(off)
  using TC = ThrCfg<N>;
  constexpr int NH = TC::NH, RZ = TC::RZ, ZE = TC::ZE, NT = TC::NT, G = TC::G;
  constexpr bool XE = (NF - (REC ? 1 : 0)) == 4;
  static_assert(NF - (REC ? 1 : 0) == 3 || NF - (REC ? 1 : 0) == 4, "three or four fields besides N_rec");
  constexpr int TOT = NF * ZE;  // complex elements of every field's rows
  constexpr int EPT = TOT / NT;
  static_assert(TOT % NT == 0, "fields must split evenly over the threads");
  __shared__ float2 s[TOT + TOT / 32];
  __shared__ float2 tw[N];
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * RZ;
  // the cells' ladder state, loaded with the spectra (the barriers of the
  // transform would otherwise leave its latency exposed)
  uint32_t st_pre[G];
#pragma unroll
  for (int j = 0; j < G; ++j)
    st_pre[j] = (a.step == 0) ? 0xFFFFFFFFu
                              : *reinterpret_cast<const uint32_t*>(state + row0 * N + 4 * (tid + j * NT));
  {
    float2 v[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT, f = e / ZE, k = e % ZE;
      v[i] = (k % NH < a.kzlive) ? T[f * fstride + row0 * NH + k] : make_float2(0.0f, 0.0f);
    }
    for (int i = tid; i < N; i += NT) tw[i] = twN[i];
#pragma unroll
    for (int i = 0; i < EPT; ++i) s[lds_addr<true>(tid + i * NT)] = v[i];
  }
  __syncthreads();
  c2r_untwist_rows<N, NF * RZ, NT>(s, tw, tid);
  // thread t owns the cell groups 4 * (t + j * NT) + q, q < 4, j < G: four
  // consecutive cells of one row, so state is one word and the outputs are
  // 16-byte stores
  constexpr int CPT = 4 * G;
  float val[NF][CPT];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    // clamps of :272-311: deltax, x_e, the rest (stars, N_rec); the weighted
    // SFR is clamped in threshold_cell, at its uses
    const int kind = (f == 0) ? 0 : (f == 2 ? -1 : ((XE && f == 3) ? 2 : 1));
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int h = f * ZE + 2 * (tid + j * NT);  // float2 index of the group's first pair
      const float2 p0 = s[lds_addr<true>(h)], p1 = s[lds_addr<true>(h + 1)];
      val[f][4 * j + 0] = clamp_kind(p0.x, kind);
      val[f][4 * j + 1] = clamp_kind(p0.y, kind);
      val[f][4 * j + 2] = clamp_kind(p1.x, kind);
      val[f][4 * j + 3] = clamp_kind(p1.y, kind);
    }
  }
  double s_xH = 0.0, s_xHm = 0.0, s_m = 0.0, s_J = 0.0;
  const long cell0 = row0 * N;
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const int e0 = 4 * (tid + j * NT);
    const long ir0 = cell0 + e0;
    const uint32_t st_old = st_pre[j];
    uint32_t st_new = st_old;
    float xh4[4], rb4[4], j4[4];
    int zcheck4 = 0;
    float jaux4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = 4 * j + q;
      float v[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) v[f] = val[f][i];
      float xh = 0.0f, rb = 0.0f, jf = 0.0f, ja = 0.0f;
      bool zc = false;
      const uint32_t st = (st_old >> (8 * q)) & 0xFFu;
      const uint32_t sn = threshold_cell<NF, REC>(v, st, ir0 + q, a, J21, Gamma12, Rtab, xh, rb, jf, zc, ja);
      st_new = (st_new & ~(0xFFu << (8 * q))) | (sn << (8 * q));
      if (a.last) {
        xh4[q] = xh;
        rb4[q] = rb;
        j4[q] = jf;
      }
      jaux4[q] = ja;
      zcheck4 |= (zc ? 1 : 0) << q;
    }
    if (st_new != st_old || a.step == 0) *reinterpret_cast<uint32_t*>(state + ir0) = st_new;
    if (a.last) {
      *reinterpret_cast<float4*>(xH + ir0) = make_float4(xh4[0], xh4[1], xh4[2], xh4[3]);
      *reinterpret_cast<float4*>(r_bubble + ir0) = make_float4(rb4[0], rb4[1], rb4[2], rb4[3]);
      const int r = e0 / N, z = e0 % N;
      const float* dxp = deltax_in + (row0 + r) * (N + 2) + z;
      const float2 d01 = *reinterpret_cast<const float2*>(dxp);
      const float2 d23 = *reinterpret_cast<const float2*>(dxp + 2);
      const float dx4[4] = {d01.x, d01.y, d23.x, d23.y};
      float4 jv = make_float4(0.f, 0.f, 0.f, 0.f);
      if (a.uvb) jv = *reinterpret_cast<const float4*>(J21 + ir0);
      const float jj[4] = {jv.x, jv.y, jv.z, jv.w};
      float jo[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double dm = 1.0 + (double)dx4[q];
        s_xH += (double)xh4[q];
        s_xHm += (double)xh4[q] * dm;
        s_m += dm;
        jo[q] = (j4[q] != 0.0f) ? jj[q] : 0.0f;
        if (a.uvb) s_J += (double)jo[q];
      }
      if (a.uvb) *reinterpret_cast<float4*>(J21 + ir0) = make_float4(jo[0], jo[1], jo[2], jo[3]);
      if (zcheck4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long ir = ir0 + q;
          if (((zcheck4 >> q) & 1) && z_in[ir] < 0) {
            z_in[ir] = (float)a.redshift;
            if (a.uvb) J21_ion[ir] = ((j4[q] != 0.0f) ? jj[q] : jaux4[q]) * (float)a.bias;
          }
        }
      }
    }
  }
  if (a.last) {
    double* p = partials + (long)blockIdx.x * 4;
    double v0 = block_sum<NT>(s_xH, red);
    double v1 = block_sum<NT>(s_xHm, red);
    double v2 = block_sum<NT>(s_m, red);
    double v3 = block_sum<NT>(s_J, red);
    if (tid == 0) {
      p[0] = v0;
      p[1] = v1;
      p[2] = v2;
      p[3] = v3;
    }
  }
}

// The fused z pass + threshold with the row transforms in registers, for
// rows of NH = 16 A complex (A = 16: 512^3, A = 8: 256^3) and three or four
// fields (no N_rec).  A field's packed row is held by A lanes, lane t holding
// points t + A m, m < 16.  The C2R untwist pairs point k with NH - k, held by
// lane (A - t) mod A (a lane shuffle); the length-NH inverse transform is a
// 16-point DFT of each lane's registers, a twiddle, one exchange through LDS,
// and A-point DFTs, after which lane t holds z[t + A j], j < 16, the real
// cells 2n and 2n + 1 of n = t + A j.  A wave holds four field groups of A
// lanes per row (64 / 4A rows); the fields of a cell meet through LDS: lane
// (g, t) tests the cells of n = t + A j, j in [4g, 4g + 4).  Three barriers
// per workgroup and two LDS round trips, where staging the rows in LDS for a
// Stockham transform takes eight and five.  Per cell the test is
// threshold_cell's, as in c2r_threshold_tile.
// Inverse real transform (C2R) of packed rows of NH = 16 A complex held by A
// lanes each: lane t = lane mod A holds points t + A m of its row in x[m] (m <
// 16), and gets z[j] = point t + A j of the row's complex output (real cells
// 2n, 2n + 1 of n = t + A j).  The untwist pairs point k with NH - k, which
// lane (A - t) mod A holds (a lane shuffle); then a 16-point DFT of the lane's
// registers, the twiddle exp(+2 pi i t k1 / NH), one exchange through the
// row's LDS rows xr (16 x A, pitch A + 1) and A-point DFTs over the lanes.
// The untwist is c2r_untwist_rows's; tw is the N-entry table exp(-2 pi i j /
// N) in LDS.  Every thread of the workgroup calls it (one barrier inside).
template <int A>
__device__ __forceinline__ void row_c2r_reg(const float2 (&x)[16], float2 (&z)[16], const float2* tw, float2* xr,
                                            int lane)
{
  constexpr int XP = A + 1;
  const int t = lane % A;
  const int src_lane = lane - t + (A - t) % A;
  float2 y[16];
  {
    float2 pr[16];  // the partner lane's registers
#pragma unroll
    for (int m = 0; m < 16; ++m) pr[m] = make_float2(__shfl(x[m].x, src_lane), __shfl(x[m].y, src_lane));
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      // x[NH - k]: the partner's register 15 - m, or 16 - m in lane 0
      const float2 xm = cconj(t == 0 ? pr[(16 - m) & 15] : pr[15 - m]);
      const int k = t + A * m;
      if (k == 0) {
        y[m] = make_float2(x[m].x + x[m].y, x[m].x - x[m].y);  // the packed DC and Nyquist values
      } else {
        const float2 E = cadd(x[m], xm);
        const float2 O = cmul(csub(x[m], xm), cconj(tw[k]));
        y[m] = make_float2(E.x - O.y, E.y + O.x);
      }
    }
  }
  Bfly<16, true>::run(y);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) y[k1] = cmul(y[k1], cconj(tw[2 * t * k1]));
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) xr[k1 * XP + t] = y[k1];
  __syncthreads();
  // for k1 = t + A h: X[k1 + 16 k2], k2 < A, is z[t + A j], j = h + (16 / A) k2
#pragma unroll
  for (int h = 0; h < 16 / A; ++h) {
    const int k1 = t + A * h;
    float2 u[A];
#pragma unroll
    for (int tp = 0; tp < A; ++tp) u[tp] = xr[k1 * XP + tp];
    Bfly<A, true>::run(u);
#pragma unroll
    for (int k2 = 0; k2 < A; ++k2) z[h + (16 / A) * k2] = u[k2];
  }
}

// k_c2r_z with the row transforms in registers (row_c2r_reg), for N = 32 A:
// 256 / A rows per 256-thread workgroup; the arguments and the output are
// k_c2r_z's (rows of pitch floats, the padding zeroed).
template <int A>
__global__ __launch_bounds__(256) void k_c2r_z_reg(const float2* __restrict__ in, float* __restrict__ out,
                                                   const float2* __restrict__ twN, long in_fstride, long out_fstride,
                                                   int clamp, int pitch, int kzlive)
{
  constexpr int NH = 16 * A, N = 2 * NH, RWG = 256 / A, XP = A + 1;
  __shared__ float2 tw[N];
  __shared__ float2 xb[RWG][16 * XP];
  const int tid = threadIdx.x, lane = tid & 63, t = lane % A, rl = tid / A;
  const long row = (long)blockIdx.x * RWG + rl;
  for (int i = tid; i < N; i += 256) tw[i] = twN[i];
  float2 x[16];
  {
    const float2* src = in + blockIdx.y * in_fstride + row * NH + t;
#pragma unroll
    for (int m = 0; m < 16; ++m) x[m] = (t + A * m < kzlive) ? src[A * m] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();  // the twiddle table
  float2 z[16];
  row_c2r_reg<A>(x, z, tw, xb[rl], lane);
  float* o = out + blockIdx.y * out_fstride + row * pitch;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    store_nt(reinterpret_cast<float2*>(o + 2 * (t + A * j)),
             make_float2(clamp_kind(z[j].x, clamp), clamp_kind(z[j].y, clamp)));
  if (pitch > N && t == 0) *reinterpret_cast<float2*>(o + N) = make_float2(0.0f, 0.0f);
}

template <int A>
struct RegThr {
  static constexpr int NH = 16 * A, N = 2 * NH;
  static constexpr int RW = 64 / (4 * A);  // rows per wave
  static constexpr int RWG = 4 * RW;        // rows per 256-thread workgroup
  static constexpr int XP = A + 1;          // LDS pitch of an exchange row
  static_assert(A == 8 || A == 16, "rows of 128 or 256 complex");
};

template <int A, int NF>
__global__ __launch_bounds__(256) void k_c2r_threshold_reg(const float2* __restrict__ T, long fstride,
                                                           const float2* __restrict__ twN,
                                                           const float* __restrict__ deltax_in,
                                                           uint8_t* __restrict__ state, float* __restrict__ xH,
                                                           float* __restrict__ r_bubble, float* __restrict__ J21,
                                                           float* __restrict__ z_in, float* __restrict__ J21_ion,
                                                           const float* __restrict__ Rtab,
                                                           double* __restrict__ partials, StepArgs a)
{
#pragma clang fp contract(off)
  using C = RegThr<A>;
  constexpr int NH = C::NH, N = C::N, RW = C::RW, RWG = C::RWG, XP = C::XP;
  static_assert(NF == 3 || NF == 4, "three or four fields");
  constexpr bool XE = NF == 4;
  __shared__ float2 tw[N];
  __shared__ float2 xb[RWG][4][16 * XP];  // [row of the workgroup][field group][16 x A, pitch XP]
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int t = lane % A, g = (lane / A) % 4;
  const int rl = (tid >> 6) * RW + lane / (4 * A);  // row within the workgroup
  const long row = (long)blockIdx.x * RWG + rl;     // x N + y
  for (int i = tid; i < N; i += 256) tw[i] = twN[i];
  // the ladder state of this lane's cells, two bytes at cell 2n, n = t + A (4g + q)
  uint32_t st_pre[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    st_pre[q] = (a.step == 0) ? 0xFFFFu
                              : *reinterpret_cast<const uint16_t*>(state + row * N + 2 * (t + A * (4 * g + q)));
  // field g's row; the fourth group carries zeros when there are three fields
  float2 x[16];
  {
    const float2* src = T + (long)(g < NF ? g : 0) * fstride + row * NH + t;
#pragma unroll
    for (int m = 0; m < 16; ++m)
      x[m] = (g < NF && t + A * m < a.kzlive) ? src[A * m] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();  // the twiddle table
  float2* xr = xb[rl][g];
  float2 z[16];
  row_c2r_reg<A>(x, z, tw, xr, lane);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) xr[j * XP + t] = z[j];
  __syncthreads();
  // the cells n = t + A j, j = 4g + q: all fields of the row from LDS
  double s_xH = 0.0, s_xHm = 0.0, s_m = 0.0, s_J = 0.0;
  const long cell0 = row * N;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = 4 * g + q;
    const int c0 = 2 * (t + A * j);  // first of the two cells
    float v[2][NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      // clamps of :272-311; the weighted SFR's is threshold_cell's
      const int kind = (f == 0) ? 0 : (f == 2 ? -1 : ((XE && f == 3) ? 2 : 1));
      const float2 p = xb[rl][f][j * XP + t];
      v[0][f] = clamp_kind(p.x, kind);
      v[1][f] = clamp_kind(p.y, kind);
    }
    const uint32_t st_old = st_pre[q];
    uint32_t st_new = st_old;
    float xh2[2], rb2[2], j2[2], ja2[2];
    bool zc2[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      xh2[e] = rb2[e] = j2[e] = 0.0f;
      const uint32_t st = (st_old >> (8 * e)) & 0xFFu;
      const uint32_t sn = threshold_cell<NF, false>(v[e], st, cell0 + c0 + e, a, J21, nullptr, Rtab, xh2[e], rb2[e],
                                                    j2[e], zc2[e], ja2[e]);
      st_new = (st_new & ~(0xFFu << (8 * e))) | (sn << (8 * e));
    }
    const long ir0 = cell0 + c0;
    if (st_new != st_old || a.step == 0) *reinterpret_cast<uint16_t*>(state + ir0) = (uint16_t)st_new;
    if (a.last) {
      *reinterpret_cast<float2*>(xH + ir0) = make_float2(xh2[0], xh2[1]);
      *reinterpret_cast<float2*>(r_bubble + ir0) = make_float2(rb2[0], rb2[1]);
      const float2 dx = *reinterpret_cast<const float2*>(deltax_in + row * (N + 2) + c0);
      const float dx2[2] = {dx.x, dx.y};
      float2 jv = make_float2(0.f, 0.f);
      if (a.uvb) jv = *reinterpret_cast<const float2*>(J21 + ir0);
      const float jj[2] = {jv.x, jv.y};
      float jo[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const double dm = 1.0 + (double)dx2[e];
        s_xH += (double)xh2[e];
        s_xHm += (double)xh2[e] * dm;
        s_m += dm;
        jo[e] = (j2[e] != 0.0f) ? jj[e] : 0.0f;
        if (a.uvb) s_J += (double)jo[e];
      }
      if (a.uvb) *reinterpret_cast<float2*>(J21 + ir0) = make_float2(jo[0], jo[1]);
      if (zc2[0] || zc2[1]) {
        const float2 zi = *reinterpret_cast<const float2*>(z_in + ir0);
        const float zi2[2] = {zi.x, zi.y};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          if (zc2[e] && zi2[e] < 0) {
            z_in[ir0 + e] = (float)a.redshift;
            if (a.uvb) J21_ion[ir0 + e] = ((j2[e] != 0.0f) ? jj[e] : ja2[e]) * (float)a.bias;
          }
        }
      }
    }
  }
  if (a.last) {
    double v[4] = {s_xH, s_xHm, s_m, s_J};
    block_sum_n<256, 4>(v, red);
    if (tid == 0) {
      double* p = partials + (long)blockIdx.x * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) p[k] = v[k];
    }
  }
}

template <int N, int NF, bool REC = false>
__global__ __launch_bounds__(ThrCfg<N>::NT) void k_c2r_threshold(const float2* __restrict__ T, long fstride,
                                                      const float2* __restrict__ twN,
                                                      const float* __restrict__ deltax_in,
                                                      uint8_t* __restrict__ state, float* __restrict__ xH,
                                                      float* __restrict__ r_bubble, float* __restrict__ J21,
                                                      float* __restrict__ z_in, float* __restrict__ J21_ion,
                                                      float* __restrict__ Gamma12, const float* __restrict__ Rtab,
                                                      double* __restrict__ partials, StepArgs a)
{
  c2r_threshold_tile<N, NF, REC>(T, fstride, twN, deltax_in, state, xH, r_bubble, J21, z_in, J21_ion, Gamma12, Rtab,
                                 partials, a);
}

// Per-box grids and scalars of a batch (device array, one entry per box).
struct BoxDev {
  const float* deltax;
  float* xH;
  float* r_bubble;
  float* J21;
  float* z_in;
  float* J21_ion;
  double eff;
  double redshift;
  double bias;
};

// The ladder step of every box of a cohort: blockIdx.y is the box.  T holds
// nfield spectra per box (box stride bstride); state and partials are
// contiguous per box; J21c_tab[box * nstep + step] is the box's J_21 constant.
template <int N, int NF>
__global__ __launch_bounds__(ThrCfg<N>::NT) void k_c2r_threshold_batch(const float2* __restrict__ T, long fstride, long bstride,
                                                            const float2* __restrict__ twN,
                                                            const BoxDev* __restrict__ boxes,
                                                            uint8_t* __restrict__ state,
                                                            const float* __restrict__ Rtab,
                                                            const double* __restrict__ J21c_tab, int nstep,
                                                            double* __restrict__ partials, StepArgs a)
{
  const int box = blockIdx.y;
  const BoxDev bx = boxes[box];
  a.eff = bx.eff;
  a.redshift = bx.redshift;
  a.bias = bx.bias;
  a.J21c = J21c_tab[(long)box * nstep + a.step];
  c2r_threshold_tile<N, NF, false>(T + box * bstride, fstride, twN, bx.deltax, state + (long)box * N * N * N, bx.xH,
                                   bx.r_bubble, bx.J21, bx.z_in, bx.J21_ion, nullptr, Rtab,
                                   partials + (long)box * gridDim.x * 4, a);
}

// ------------------------------------------------- brightness temperature
struct TbArgs {
  float const_factor;  // BrightnessTemperature.c:50-51
  double T_rad;        // TCMB (1+z)
  int N;
  int use_ts;
};

__global__ __launch_bounds__(kNT) void k_brightness_temp(const float* __restrict__ xH,
                                                        const float* __restrict__ deltax,
                                                        const float* __restrict__ TS, float* __restrict__ dT,
                                                        double* __restrict__ partials, long ncell, TbArgs a)
{
#pragma clang fp contract(off)
  __shared__ double red[kNT];
  double acc = 0.0;
  // rows of the grid over the workgroups, cells of a row over the threads
  // (no per-cell index division); ncell = N^3
  const long nrow = ncell / a.N;
  for (long r = blockIdx.x; r < nrow; r += gridDim.x)
  for (int z = threadIdx.x; z < a.N; z += kNT) {
    const long i = r * a.N + z;
    const double pixel_deltax = deltax[r * (a.N + 2) + z];
    const double pixel_x_HI = xH[i];
    float v = (float)(a.const_factor * pixel_x_HI * (1.0 + pixel_deltax));
    if (a.use_ts) {
      const double pixel_Ts_factor = (1 - a.T_rad / TS[i]);
      v = (float)((double)v * pixel_Ts_factor);
    }
    dT[i] = v;
    acc += (double)v;
  }
  acc = block_sum<kNT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// ------------------------------------------------ recombinations (N_rec)
#pragma clang fp contract(off)
// Recombination-rate table on the device: knots in ln Gamma12 and the natural
// cubic spline of each redshift row (recombinations.c:56-82).
constexpr int kRRZ = 300, kRRG = 150;
struct RRDev {
  double lnG[kRRG];
  double y[kRRZ][kRRG], b[kRRZ][kRRG], c[kRRZ][kRRG], d[kRRZ][kRRG];
};

// splined_recombination_rate (recombinations.c:30-54); ln Gamma12 above the
// last knot is evaluated on the last interval (gsl_spline_eval would abort)
__device__ __forceinline__ double rr_rate(const RRDev* __restrict__ t, float z_eff, double gamma12_bg)
{
  int z_ct = (int)(z_eff / 0.2f + 0.5);
  double lnGamma = log(gamma12_bg);
  if (z_ct < 0)
    z_ct = 0;
  else if (z_ct >= kRRZ)
    z_ct = kRRZ - 1;
  if (lnGamma < -10.0) return 0;
  if (lnGamma >= (-10.0 + 0.1f * kRRG)) lnGamma = -10.0 + 0.1f * kRRG - 1e-7;
  int lo = 0, hi = kRRG - 1;
  if (lnGamma >= t->lnG[hi]) {
    lo = hi - 1;
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (t->lnG[mid] > lnGamma)
        hi = mid;
      else
        lo = mid;
    }
  }
  const double dx = lnGamma - t->lnG[lo];
  return t->y[z_ct][lo] + dx * (t->b[z_ct][lo] + dx * (t->c[z_ct][lo] + dx * t->d[z_ct][lo]));
}

// N_rec += RR(z_eff, Gamma12) |dt/dz| zstep (1 - xH) for every cell after the
// ladder (find_HII_bubbles.c:472-477); deltax is the input grid.
__global__ __launch_bounds__(kNT) void k_nrec_update(float* __restrict__ N_rec, const float* __restrict__ Gamma12,
                                                    const float* __restrict__ xH,
                                                    const float* __restrict__ deltax_in,
                                                    const RRDev* __restrict__ rr, int N, double redshift,
                                                    float fabs_dtdz, float zstep)
{
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= (long)N * N * N) return;
  const long ip = (i / N) * (N + 2) + i % N;
  const double cell_xH = (double)xH[i];
  const double density_over_mean = 1.0 + (double)deltax_in[ip];
  const float z_eff = (float)((1. + redshift) * pow(density_over_mean, 1.0 / 3.0) - 1);
  const double dNrec = rr_rate(rr, z_eff, (double)Gamma12[i]) * (double)fabs_dtdz * (double)zstep * (1. - cell_xH);
  N_rec[ip] = (float)((double)N_rec[ip] + dNrec);
}
#pragma clang fp contract(on)

// ----------------------------------- brightness temperature with velocities
// Peculiar-velocity branch of ComputeBrightnessTemperatureBox
// (BrightnessTemperature.c:126-510), every operation in the reference's
// precision and order (fp contract off).
#pragma clang fp contract(off)
constexpr int kRsdSteps = 50;  // N_RSD_STEPS (reionization.h:34)

struct TbVelArgs {
  float const_factor;  // :50-51
  float redshift;      // (float)ZZ[snapshot]
  float H_z;           // (float)hubble(redshift)
  int N;
  int pec_vels;
  int trees_id;
  double T_rad;
  double max_v_deriv;  // fabs(MAX_DVDR * H_z)
};

// vel (padded, real cells) to comoving velocity in place (:166-195)
__global__ __launch_bounds__(kNT) void k_vel_convert(float* __restrict__ vel, int N, float redshift, int trees_id)
{
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= (long)N * N * N) return;
  float* v = vel + (i / N) * (N + 2) + i % N;
  float x = *v;
  if (trees_id == 1)
    x = (float)(sqrt(1. / (1. + redshift)) * (double)x / 1000.);
  else if (trees_id == 0 || trees_id == 2)
    x = (float)((1. + redshift) * (double)x);
  x = (float)((double)x * (1000. * 100. / 3.086e24));
  x = (float)((double)x / (1. / (1. + redshift)));
  *v = x;
}

// Packed spectrum U (forward transform / N^3) -> T = U * Hubble_h * i k_z
// (:199-205, velocity_gradient reionization.c:2172-2187).  The kz = 0 plane is
// multiplied by 0; the Nyquist plane's product i k B is purely imaginary after
// the inverse x/y passes (B is Hermitian), and the c2r keeps only the real
// part of the Nyquist coefficient, so it contributes 0: the packed kz = 0 slot
// of T is 0.
__global__ __launch_bounds__(kNT) void k_vel_grad_op(const float2* __restrict__ U, float2* __restrict__ T, int N,
                                                    double hubble_h, float delta_k)
{
  const int NH = N / 2;
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= (long)N * N * NH) return;
  const int kz = (int)(i % NH);
  if (kz == 0) {
    T[i] = make_float2(0.0f, 0.0f);
    return;
  }
  const float2 c = U[i];
  const float re = (float)((double)c.x * hubble_h), im = (float)((double)c.y * hubble_h);
  const float k_z = (float)kz * delta_k;
  T[i] = make_float2(re * 0.0f - im * k_z, re * k_z + im * 0.0f);
}

// delta_T with T_S and the velocity gradient grad (padded), pec_vels 1 or >= 2 (:86-120, :210-287)
__global__ __launch_bounds__(kNT) void k_brightness_temp_vel(const float* __restrict__ xH,
                                                            const float* __restrict__ deltax,
                                                            const float* __restrict__ TS,
                                                            const float* __restrict__ grad, float* __restrict__ dT,
                                                            TbVelArgs a)
{
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= (long)a.N * a.N * a.N) return;
  const long ip = (i / a.N) * (a.N + 2) + i % a.N;
  const double pixel_deltax = deltax[ip], pixel_x_HI = xH[i];
  float v = (float)(a.const_factor * pixel_x_HI * (1.0 + pixel_deltax));
  const float ts = TS[i];
  if (a.pec_vels > 1)
    v = (float)((double)v * ((1. + a.redshift) / (1000. * (double)ts)));
  else
    v = (float)((double)v * (1 - a.T_rad / (double)ts));
  if (a.pec_vels == 1) {
    double dvdx = (double)grad[ip];
    if (fabs(dvdx) > a.max_v_deriv) dvdx = (dvdx < 0) ? -a.max_v_deriv : a.max_v_deriv;
    v = (float)((double)v / (dvdx / (double)a.H_z + 1.0));
  } else {
    const double gc = fabs((double)grad[ip] / (double)a.H_z + 1.0);
    if (gc < 1e-7)
      v = (float)(1000. * ((double)ts - a.T_rad) / (1. + a.redshift));
    else
      v = (float)((1. - exp(-(double)v / gc)) * 1000. * ((double)ts - a.T_rad) / (1. + a.redshift));
  }
  dT[i] = v;
}

struct RsdArgs {
  int N;
  float H_z;
  double cell;    // BoxSize / (float)N  [Mpc]
  double half;    // (subcell_width / 2) / cell
  double frac_w;  // subcell_width / cell
  float x_pos[kRsdSteps];
  float x_off[kRsdSteps];
};

// Line-of-sight re-binning of the sub-cells (:312-496), one thread per (x, y)
// column, cell by cell and sub-cell by sub-cell as the reference does, so every
// bin receives its additions in the reference's order.  rsd_cell places the
// 50 sub-cells of cell k and hands each share to a sink: a float share
// (`bin += delta_T / 50.f`) or a double one (`bin += fraction * delta_T /
// 50.f`, rounded to float after the addition).  Indices the reference's single
// +-N correction would leave outside the column are wrapped periodically.
__device__ __forceinline__ int rsd_wrap(int k, int N) { return ((k % N) + N) % N; }

template <class Sink>
__device__ __forceinline__ void rsd_cell(int k, int N, float d, double d1_low, double d2_low, double d1_high,
                                         double d2_high, const RsdArgs& a, Sink& sink)
{
  const float share = d / (float)kRsdSteps;
  const double w = (double)(float)kRsdSteps;
  for (int s = 0; s < kRsdSteps; ++s) {
    const double xp = a.x_pos[s];
    const double disp = (xp <= 0.5) ? d1_low + ((xp + 0.5) - 0.) * (d2_low - d1_low) / (1. - 0.)
                                    : d1_high + ((xp - 0.5) - 0.) * (d2_high - d1_high) / (1. - 0.);
    const double pos = ((double)a.x_off[s] + disp) / a.cell;
    const double lo = pos - a.half, hi = pos + a.half;
    if (lo >= 0.0 && hi < 1.0) {
      sink.f(k, share);
    } else if (lo < 0.0 && hi < 0.0) {
      const double cd = ceil(fabs(lo)) - 1.;
      const int c = (int)cd;
      if (fabs(hi) > cd) {
        sink.f(rsd_wrap(k < c + 1 ? k - (c + 1) + N : k - (c + 1), N), share);
      } else {
        const double fo = (fabs(lo) - cd) / a.frac_w, fw = 1. - fo;
        sink.d(rsd_wrap(k < c ? k - c + N : k - c, N), fw * (double)d / w);
        sink.d(rsd_wrap(k < c + 1 ? k - (c + 1) + N : k - (c + 1), N), fo * (double)d / w);
      }
    } else if (lo < 0.0 && (hi > 0.0 && hi < 1.0)) {
      const double fw = hi / a.frac_w, fo = 1. - fw;
      sink.d(k == 0 ? N - 1 : k - 1, fo * (double)d / w);
      sink.d(k, fw * (double)d / w);
    } else if ((lo >= 0.0 && lo < 1.0) && hi >= 1.0) {
      const double fo = (hi - 1.) / a.frac_w, fw = 1. - fo;
      sink.d(k, fw * (double)d / w);
      sink.d(k == N - 1 ? 0 : k + 1, fo * (double)d / w);
    } else {
      const double cd = floor(fabs(hi));
      const int c = (int)cd;
      if (lo >= cd) {
        sink.f(rsd_wrap(k > N - 1 - c ? k + c - N : k + c, N), share);
      } else {
        const double fo = (hi - cd) / a.frac_w, fw = 1. - fo;
        sink.d(rsd_wrap(k > N - 1 - (c - 1) ? k + c - 1 - N : k + c - 1, N), fw * (double)d / w);
        sink.d(rsd_wrap(k > N - 1 - c ? k + c - N : k + c, N), fo * (double)d / w);
      }
    }
  }
}

// Bins of the column in global memory, bins[k * N^2 + column] (coalesced across a wave).
struct RsdGlobalSink {
  float* b;
  long NN;
  __device__ __forceinline__ void f(int k, float v)
  {
    float* q = b + (long)k * NN;
    *q = *q + v;
  }
  __device__ __forceinline__ void d(int k, double v)
  {
    float* q = b + (long)k * NN;
    *q = (float)((double)*q + v);
  }
};

__global__ __launch_bounds__(kNT) void k_rsd_los(float* __restrict__ dT, const float* __restrict__ xH,
                                                const float* __restrict__ vel, float* __restrict__ bins, RsdArgs a)
{
  const long col = (long)blockIdx.x * kNT + threadIdx.x;
  const int N = a.N;
  const long NN = (long)N * N;
  if (col >= NN) return;
  const float* dt = dT + col * N;
  const float* xh = xH + col * N;
  const float* vl = vel + col * (N + 2);
  float* o = dT + col * N;
  float* b = bins + col;
  for (int k = 0; k < N; ++k) b[k * NN] = 0.0f;
  RsdGlobalSink g{b, NN};
  for (int k = 0; k < N; ++k) {
    const float d = dt[k];
    if (!(fabs((double)d) >= 1e-7 && (double)xh[k] >= 1e-7)) continue;
    const int km = (k == 0) ? N - 1 : k - 1, kp = (k == N - 1) ? 0 : k + 1;
    rsd_cell(k, N, d, vl[km] / a.H_z, vl[k] / a.H_z, vl[k] / a.H_z, vl[kp] / a.H_z, a, g);
  }
  for (int k = 0; k < N; ++k) o[k] = b[k * NN];
}
#pragma clang fp contract(on)

}  // namespace mrx
