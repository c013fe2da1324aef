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
constexpr float kRelTol = 1e-5f;  // REL_TOL (src/meraxes.h:41)
constexpr float kAbsTol = 1e-8f;  // ABS_TOL (src/meraxes.h:42)

constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }

template <int N>
struct Cfg {
  static_assert(N % 16 == 0, "grid dimension must be a multiple of 16");
  static constexpr int NH = N / 2;
  static constexpr int TW = (NH % kTW == 0) ? kTW : kTW / 2;  // pencils per strided tile
  static constexpr int COL_NT_RAW = TW * N / 8;
  static constexpr int COL_NT = COL_NT_RAW < 64 ? 64 : (COL_NT_RAW > kNT ? kNT : COL_NT_RAW);
  static constexpr int RZ = (NH >= kZE) ? 1 : pow2_floor(kZE / NH);  // rows per z workgroup (divides N*N)
  static constexpr int ZE = RZ * NH;                               // complex elements per z workgroup
  static constexpr int CPT = 2 * ZE / kNT;                          // real cells per thread in the z pass
  static_assert((N * N) % RZ == 0, "z workgroups must tile the rows");
  static_assert((2 * ZE) % kNT == 0, "z cells must split evenly over the threads");
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

// Multiplier filter() applies to mode (nx, ny, nz); 1 where it leaves it alone.
__device__ __forceinline__ float window(int nx, int ny, int nz, int N, const WinArgs& w)
{
  if (!w.apply) return 1.0f;
  const float kx = kcomp(nx, N, w.delta_k);
  const float ky = kcomp(ny, N, w.delta_k);
  const float kz = (float)nz * w.delta_k;
  const float k_mag = sqrtf(kx * kx + ky * ky + kz * kz);
  float kR = k_mag * w.R;
  switch (w.type) {
    case 0:
      if ((double)kR > 1e-4) return tophat_window(kR);
      return 1.0f;
    case 1:
      kR = (float)((double)kR * 0.413566994);
      return (kR > 1.0f) ? 0.0f : 1.0f;
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
template <int N>
__global__ __launch_bounds__(kNT) void k_r2c_z(const float* __restrict__ in, float2* __restrict__ out,
                                              const float2* __restrict__ twN, long in_fstride, long out_fstride)
{
  using C = Cfg<N>;
  constexpr int NH = C::NH, RZ = C::RZ, ZE = C::ZE;
  __shared__ float2 s[ZE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += kNT) tw[i] = twN[i];
  const long row0 = (long)blockIdx.x * RZ;
  const float2* src = reinterpret_cast<const float2*>(in + blockIdx.y * in_fstride);
  float2* dst = out + blockIdx.y * out_fstride;
  for (int e = tid; e < ZE; e += kNT) {
    const int r = e / NH, n = e % NH;
    s[e] = src[(row0 + r) * (NH + 1) + n];
  }
  __syncthreads();
  fft_lds<NH, RZ, kNT, false, 1, NH, false, 2>(s, tw, tid);
  // twist: X[k] = E + w^k O, X[NH-k] = conj(E - w^k O), E=(Z_k+conj Z_{NH-k})/2, O=(Z_k-conj Z_{NH-k})/(2i)
  constexpr int NP = NH / 2 + 1;
  for (int q = tid; q < RZ * NP; q += kNT) {
    const int r = q / NP, k = q % NP;
    const float2* row = s + r * NH;
    float2* orow = dst + (row0 + r) * NH;
    if (k == 0) {
      const float2 z = row[0];
      orow[0] = make_float2(z.x + z.y, z.x - z.y);
    } else {
      const float2 zk = row[k], zm = cconj(row[NH - k]);
      const float2 E = cscale(cadd(zk, zm), 0.5f);
      const float2 d = csub(zk, zm);
      const float2 O = make_float2(0.5f * d.y, -0.5f * d.x);
      const float2 wo = cmul(tw[k], O);
      orow[k] = cadd(E, wo);
      if (k != NH - k) orow[NH - k] = cconj(csub(E, wo));
    }
  }
}

// ---------------------------------------------------- z: C2R load + FFT
// Loads RZ packed rows, untwists them and runs the inverse length-N/2 FFT in
// LDS; afterwards the LDS tile, read as floats, holds RZ real rows of N.
template <int N>
__device__ __forceinline__ void c2r_rows_lds(float2* s, const float2* tw, const float2* __restrict__ src, long row0,
                                             int tid)
{
  using C = Cfg<N>;
  constexpr int NH = C::NH, RZ = C::RZ, ZE = C::ZE;
  for (int e = tid; e < ZE; e += kNT) s[e] = src[row0 * NH + e];
  __syncthreads();
  constexpr int NP = NH / 2 + 1;
  for (int q = tid; q < RZ * NP; q += kNT) {
    const int r = q / NP, k = q % NP;
    float2* row = s + r * NH;
    if (k == 0) {
      const float2 p = row[0];  // (X_0, X_{N/2}), both real
      row[0] = make_float2(p.x + p.y, p.x - p.y);
    } else {
      const float2 xk = row[k], xm = cconj(row[NH - k]);
      const float2 E = cadd(xk, xm);
      const float2 O = cmul(csub(xk, xm), cconj(tw[k]));  // * exp(+2 pi i k / N)
      row[k] = make_float2(E.x - O.y, E.y + O.x);       // E + i O
      if (k != NH - k) row[NH - k] = make_float2(E.x + O.y, -E.y + O.x);  // conj(E) + i conj(O)
    }
  }
  __syncthreads();
  fft_lds<NH, RZ, kNT, true, 1, NH, false, 2>(s, tw, tid);
}

// aliasing clamps of find_HII_bubbles.c:272-311
__device__ __forceinline__ float clamp_kind(float v, int kind)
{
  switch (kind) {
    case 0:
      return fmaxf(v, -1.0f + kRelTol);
    case 1:
      v = fmaxf(v, 0.0f);
      return (v < kAbsTol) ? 0.0f : v;
    case 2:
      v = fmaxf(v, 0.0f);
      v = (v < kAbsTol) ? 0.0f : v;
      return fminf(v, 0.999f);
    case 3:  // spin-temperature shells: max(v, 0) only (ComputeTs.c:417, :488)
      return fmaxf(v, 0.0f);
    default:
      return v;
  }
}

// Plain C2R z pass: packed rows -> padded real rows, with an optional clamp.
template <int N>
__global__ __launch_bounds__(kNT) void k_c2r_z(const float2* __restrict__ in, float* __restrict__ out,
                                              const float2* __restrict__ twN, long in_fstride, long out_fstride,
                                              int clamp, int pitch)
{
  using C = Cfg<N>;
  constexpr int RZ = C::RZ, ZE = C::ZE;
  __shared__ float2 s[ZE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += kNT) tw[i] = twN[i];
  const long row0 = (long)blockIdx.x * RZ;
  c2r_rows_lds<N>(s, tw, in + blockIdx.y * in_fstride, row0, tid);
  const float* sf = reinterpret_cast<const float*>(s);
  float* o = out + blockIdx.y * out_fstride;
  for (int e = tid; e < RZ * pitch; e += kNT) {
    const int r = e / pitch, z = e % pitch;
    o[(row0 + r) * pitch + z] = (z < N) ? clamp_kind(sf[r * N + z], clamp) : 0.0f;
  }
}

// --------------------------------------------------- strided pencil passes
// One N x kTW tile: pencils along an axis with element stride `rstride`;
// tile t covers kz block (t mod nkzb) of line (t / nkzb), line stride astride.
template <int N, bool INV>
__global__ __launch_bounds__(Cfg<N>::COL_NT) void k_colpass(float2* __restrict__ data, const float2* __restrict__ twN,
                                                           long fstride, long astride, long rstride, int nkzb,
                                                           double div)
{
  constexpr int NT = Cfg<N>::COL_NT, kTW = Cfg<N>::TW;
  __shared__ float2 s[N * kTW];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  float2* g = data + blockIdx.y * fstride;
  const int a = blockIdx.x / nkzb, b = blockIdx.x % nkzb;
  const long base = (long)a * astride + (long)b * kTW;
  for (int e = tid; e < N * kTW; e += NT) {
    const int n = e / kTW, c = e % kTW;
    s[e] = g[base + n * rstride + c];
  }
  __syncthreads();
  fft_lds<N, kTW, NT, INV, kTW, 1, true, 1>(s, tw, tid);
  for (int e = tid; e < N * kTW; e += NT) {
    const int n = e / kTW, c = e % kTW;
    float2 v = s[e];
    if (div > 0.0) {  // complex float / double, in double (find_HII_bubbles.c:168-182)
      v = make_float2((float)((double)v.x / div), (float)((double)v.y / div));
    } else if (div < 0.0) {  // complex float / float (ComputeTs.c:383-384)
      const float fd = (float)(-div);
      v = make_float2(v.x / fd, v.y / fd);
    }
    g[base + n * rstride + c] = v;
  }
}

// Inverse x pass with the window applied on load, for every field of the
// ladder step.  U: unfiltered packed spectra (kz=0 slot = true kz=0 plane);
// V: their Nyquist planes [ky][kx]; T: output packed spectra.
template <int N>
__global__ __launch_bounds__(Cfg<N>::COL_NT) void k_ix_filter(const float2* __restrict__ U,
                                                             const float2* __restrict__ V, float2* __restrict__ T,
                                                             const float2* __restrict__ twN, long fstride,
                                                             long vstride, int nfield, WinArgs w)
{
  constexpr int NT = Cfg<N>::COL_NT, NH = N / 2, kTW = Cfg<N>::TW, nkzb = NH / kTW;
  constexpr int EPT = N * kTW / NT;
  __shared__ float2 s[N * kTW];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twN[i];
  const int ky = blockIdx.x / nkzb, b = blockIdx.x % nkzb;
  const long base = (long)ky * NH + (long)b * kTW;
  constexpr long rstride = (long)N * NH;
  float wv[EPT], wn[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT, n = e / kTW, c = e % kTW;
    wv[i] = window(n, ky, b * kTW + c, N, w);
    wn[i] = (b == 0 && c == 0) ? window(n, ky, NH, N, w) : 0.0f;
  }
  for (int f = 0; f < nfield; ++f) {
    const float2* u = U + f * fstride;
    const float2* vp = V + f * vstride + (long)ky * N;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT, n = e / kTW, c = e % kTW;
      float2 v = cscale(u[base + n * rstride + c], wv[i]);
      if (b == 0 && c == 0) {
        const float2 nb = cscale(vp[n], wn[i]);
        v = make_float2(v.x - nb.y, v.y + nb.x);  // A*W0 + i*(B*WN)
      }
      s[e] = v;
    }
    __syncthreads();
    fft_lds<N, kTW, NT, true, kTW, 1, true, 1>(s, tw, tid);
    float2* t = T + f * fstride;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT, n = e / kTW, c = e % kTW;
      t[base + n * rstride + c] = s[e];
    }
    __syncthreads();
  }
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

// partials[nb][nv] -> out[nv], fixed order (deterministic)
__global__ __launch_bounds__(kNT) void k_reduce_partials(const double* __restrict__ partials, int nb, int nv,
                                                       double* __restrict__ out)
{
  __shared__ double red[kNT];
  for (int v = 0; v < nv; ++v) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kNT) acc += partials[(long)i * nv + v];
    acc = block_sum<kNT>(acc, red);
    if (threadIdx.x == 0) out[v] = acc;
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
  int step;             // ladder index (0 = largest R)
  int last;             // unfiltered final step
  int uvb;
  int use_xe;
  int nfield;
};

// One ladder step of _find_HII_bubbles (find_HII_bubbles.c:255-443) for RZ
// rows: inverse z pass of every field, clamp, then the per-cell test in double
// exactly as the reference writes it.  Per-cell state between steps is one
// byte (0xFF = never crossed, else the step of the latest crossing); xH,
// r_bubble and the J_21 zeros are written once, on the last step, and the
// first crossing writes J_21 / z_at_ionization / J_21_at_ionization directly.
template <int N>
__global__ __launch_bounds__(kNT) void k_c2r_threshold(const float2* __restrict__ T, long fstride,
                                                      const float2* __restrict__ twN,
                                                      const float* __restrict__ deltax_in,
                                                      uint8_t* __restrict__ state, float* __restrict__ xH,
                                                      float* __restrict__ r_bubble, float* __restrict__ J21,
                                                      float* __restrict__ z_in, float* __restrict__ J21_ion,
                                                      const float* __restrict__ Rtab, double* __restrict__ partials,
                                                      StepArgs a)
{
#pragma clang fp contract(off)
  using C = Cfg<N>;
  constexpr int RZ = C::RZ, ZE = C::ZE, CPT = C::CPT;
  __shared__ float2 s[ZE];
  __shared__ float2 tw[N];
  __shared__ double red[kNT];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += kNT) tw[i] = twN[i];
  const long row0 = (long)blockIdx.x * RZ;
  const float* sf = reinterpret_cast<const float*>(s);
  float val[4][CPT];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (f < a.nfield) {
      c2r_rows_lds<N>(s, tw, T + f * fstride, row0, tid);
      const int kind = (f == 0) ? 0 : (f == 3 ? 2 : 1);
#pragma unroll
      for (int i = 0; i < CPT; ++i) val[f][i] = clamp_kind(sf[tid + i * kNT], kind);
      __syncthreads();
    } else {
#pragma unroll
      for (int i = 0; i < CPT; ++i) val[f][i] = 0.0f;
    }
  }
  double s_xH = 0.0, s_xHm = 0.0, s_m = 0.0, s_J = 0.0;
  const long cell0 = row0 * N;
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int e = tid + i * kNT;
    const long ir = cell0 + e;
    const double density_over_mean = 1.0 + (double)val[0][i];
    const double f_coll_stars =
      (double)val[1][i] / (a.M_mean * density_over_mean) * (4.0 / 3.0) * M_PI * a.R_cubed / a.pixel_volume;
    const double weighted_sfr_density = (double)val[2][i] / a.pixel_volume;
    const double neutral_fraction = a.use_xe ? 1.0 - (double)val[3][i] : 1.0;
    const float J_21_aux = a.uvb ? (float)(weighted_sfr_density * a.J21c) : 0.0f;
    const bool cross = f_coll_stars * a.eff > neutral_fraction * (1. + 0.0);
    const uint8_t st = (a.step == 0) ? (uint8_t)0xFF : state[ir];
    const bool never = (st == 0xFF);
    bool zcheck = false;
    if (cross) {
      if (never) {
        if (a.uvb) J21[ir] = J_21_aux;
        zcheck = true;
      }
      state[ir] = (uint8_t)a.step;
    } else if (a.step == 0) {
      state[ir] = (uint8_t)0xFF;
    }
    if (a.last) {
      const bool ion = cross || !never;
      float xh;
      if (ion) {
        xh = 0.0f;
      } else {
        xh = (float)(neutral_fraction - f_coll_stars * a.eff);
        if (xh < 0.f)
          xh = 0.f;
        else if (xh > 1.0f)
          xh = 1.f;
        if (xh < kRelTol) zcheck = true;
      }
      xH[ir] = xh;
      r_bubble[ir] = ion ? Rtab[cross ? a.step : st] : 0.0f;
      if (a.uvb && !ion) J21[ir] = 0.0f;
      const int r = e / N, z = e % N;
      const double dm = 1.0 + (double)deltax_in[(row0 + r) * (N + 2) + z];
      s_xH += (double)xh;
      s_xHm += (double)xh * dm;
      s_m += dm;
      if (a.uvb) s_J += (double)(ion ? J21[ir] : 0.0f);
    }
    if (zcheck && z_in[ir] < 0) {
      z_in[ir] = (float)a.redshift;
      if (a.uvb) J21_ion[ir] = J_21_aux * (float)a.bias;
    }
  }
  if (a.last) {
    double* p = partials + (long)blockIdx.x * 4;
    double v0 = block_sum<kNT>(s_xH, red);
    double v1 = block_sum<kNT>(s_xHm, red);
    double v2 = block_sum<kNT>(s_m, red);
    double v3 = block_sum<kNT>(s_J, red);
    if (tid == 0) {
      p[0] = v0;
      p[1] = v1;
      p[2] = v2;
      p[3] = v3;
    }
  }
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
  const long stride = (long)gridDim.x * kNT;
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < ncell; i += stride) {
    const long r = i / a.N;
    const int z = (int)(i % a.N);
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

}  // namespace mrx
