// adjoint.hpp - hand-written adjoint of one snapshot (spin temperature ->
// bubble threshold -> brightness temperature) and of the power spectrum.
//
// What is differentiated, per cell:
//   delta_T = C xH (1 + delta) (1 - T_rad / T_S)          (BrightnessTemperature.c:49-124)
//   xH      = n_f - eff f_coll on the unfiltered last step of the ladder for
//             cells never ionised; 0 for ionised cells - the HII mask, held as
//             a stop-gradient (find_HII_bubbles.c:317-443)
//   x_e, T_K, T_S from one Euler step of evolveInt and get_Ts
//             (ComputeTs.c:743-900), whose 40 shell sums are linear in the
//             filtered SFR shells.
// The per-cell spin-temperature map is run once in forward-mode dual numbers
// with seven directions (delta, x_e_prev, T_K_prev and the four shell sums),
// through the same templated code as the forward kernels, so its primal
// values equal the forward's bit for bit.  The shell sums' adjoints are then
// spread over the shells, and each shell's filter - a real symmetric
// convolution - is its own adjoint, so the adjoint of a shell is the same
// R2C -> window -> C2R chain applied to the shell's gradient field.  The SFR
// working-grid operations are replayed backwards.  The two grid means the
// shell tables depend on (collapse fraction, <x_e>) are constants here.
#pragma once
#include "dual.hpp"
#include "ts_kernels.hpp"

namespace mrx {

struct AdjCellArgs {
  float tb_const;  // const_factor of the brightness temperature
  double T_rad;    // TCMB (1 + z)
  int use_ts;      // delta_T carries the (1 - T_rad/T_S) factor
  double eff;      // ReionEfficiency
  double fc_unit;  // f_coll per unit stars at density 1: (4/3) pi R^3 / V / RtoM(R) on the last step
  int use_xe;      // n_f = 1 - x_e
  int ts_mode;     // 0: no spin temperature, 1: homogeneous branch, 2: inhomogeneous
  float h_xe0, h_T0;
  int N;
};

#pragma clang fp contract(off)

// table bracket of the cell's x_e (ComputeTs.c:759-790); dfr is d(fr)/d(x_e)
struct XeBracket {
  int lo, hi;
  double dxh, fr, dfr;
};

__device__ __forceinline__ XeBracket xe_bracket(const TsDev& t, float xe0)
{
  XeBracket b;
  double xHII_call = xe0;
  b.dfr = 1.0;
  if (xHII_call > t.XH[kNXHII - 1] * 0.999) {
    xHII_call = t.XH[kNXHII - 1] * 0.999;
    b.dfr = 0.0;
  } else if (xHII_call < t.XH[0]) {
    xHII_call = 1.001 * t.XH[0];
    b.dfr = 0.0;
  }
  int lo = kNXHII - 1;
  const float xc_f = (float)xHII_call;
  while (xc_f < t.XH[lo]) lo--;
  b.lo = lo;
  b.hi = lo + 1;
  b.dxh = (double)(t.XH[b.hi] - t.XH[lo]);
  b.fr = xHII_call - t.XH[lo];
  return b;
}

// f = slope * fr + f[lo], as the forward kernel forms it (tab/slope at the
// cell's bracket and shell); deriv = d f / d x_e
__device__ __forceinline__ double interp(double tab, double slope, const XeBracket& b, double* deriv)
{
  double f = slope;
  *deriv = f * b.dfr;
  f *= b.fr;
  f += tab;
  return f;
}

template <int K>
__device__ __forceinline__ Dual<K> dual_sum(double v, double d_xe, int dir)
{
  Dual<K> r(v);
  r.d[1] = d_xe;
  r.d[dir] = 1.0;
  return r;
}

using D7 = Dual<7>;

// One cell of the inhomogeneous spin-temperature step in dual numbers
// (directions: 0 delta, 1 x_e_prev, 2 T_K_prev, 3..6 the raw shell sums
// heat, ion, lya, star-lya); returns T_S and writes the new x_e.
__device__ __forceinline__ D7 ts_cell_dual(const TsDev& t, const double* tabs, float dl, float xe0, float Tk0,
                                           const float* shells, long ncell, long i, D7* xn_out)
{
  const XeBracket br = xe_bracket(t, xe0);
  double H = 0, I = 0, Ly = 0, St = 0, dH = 0, dI = 0, dLy = 0;
  if (!t.no_light) {
    // the forward's sums (ts_shell_sums), so the primal values are its own;
    // the slope parts carry d/d(x_e) through fr
    const ShellSums S = ts_shell_sums(t, tabs, br.lo, shells, ncell, i);
    H = S.hf + br.fr * S.hs;
    I = S.xf + br.fr * S.xs;
    Ly = S.lf + br.fr * S.ls;
    St = S.st;
    dH = S.hs * br.dfr;
    dI = S.xs * br.dfr;
    dLy = S.ls * br.dfr;
  }
  const D7 dlD = D7::var((double)dl, 0);
  const D7 xD = D7::var((double)xe0, 1);
  const D7 TD = D7::var((double)Tk0, 2);
  const D7 n_b = t.NBP3 * rnd_f(1.0 + dlD);  // (1 + dl) is float arithmetic in the forward
  D7 dxheat(0.0), dxion(0.0), dxlya(0.0), dstarlya(0.0);
  if (!t.no_light) {
    dxheat = dual_sum<7>(H, dH, 3) * t.const_zp_prefactor;
    dxion = dual_sum<7>(I, dI, 4) * t.const_zp_prefactor;
    dxlya = dual_sum<7>(Ly, dLy, 5) * (t.const_zp_prefactor * n_b);
    dstarlya = dual_sum<7>(St, 0.0, 6) * ((2.9979e10 / (4. * M_PI)) / (1.6726e-24 / 1.989e33));
  }
  const D7 sink = alpha_A_t(TD) * 2.0 * xD * xD * t.f_H * n_b;
  const D7 dxe_dzp = t.dt_dzp * (dxion - sink);
  // ts_cell_update's products with reciprocals, in its order
  const D7 inv1x = 1.0 / (1.0 + xD);
  D7 dadia(t.dadia0);
  if (fabsf(dl) > 1e-7) dadia += t.dgrowth * dlD / (1.0 + t.growth * dlD);
  dadia *= (2.0 / 3.0) * TD;
  const D7 dspec = -dxe_dzp * TD * inv1x;
  D7 dcomp = (-1.51e-4) * (xD / (1.0 + xD + t.f_He)) * t.comp;
  dcomp *= t.Trad - TD;
  const D7 dxheat_dzp = dxheat * t.dt_dzp * kHeatFac * inv1x;
  const D7 der1 = dxheat_dzp + dcomp + dspec + dadia;
  const D7 der2 = dxlya + dstarlya;
  D7 xn = rnd_f(xD + dxe_dzp * t.dzp);
  if (pval(xn) > 1)
    xn = D7((double)(float)(1 - 1e-7));
  else if (pval(xn) < 0)
    xn = D7(0.0);
  D7 Tn = rnd_f(TD);
  if (pval(Tn) < (double)(float)5e4) Tn = rnd_f(Tn + der1 * t.dzp);
  if (pval(Tn) < 0) Tn = D7((double)(float)t.Tk_floor);
  D7 xa;
  const D7 ts = rnd_f(get_Ts_t<D7>(t, &t.kHH, dlD, Tn, xn, rnd_f(der2), &xa));
  *xn_out = xn;
  return ts;
}

// Per-cell adjoint of one snapshot.  In: the forward's inputs (deltax,
// stars, x_e_prev and T_K before the step), its outputs (xH, T_S, x_e after
// the step), the recomputed shells and dL/d(delta_T).  Out: dL/d(deltax),
// dL/d(stars), dL/d(x_e_prev), dL/d(T_K) (padded where the field is) and the
// adjoints of the four raw shell sums, gS[j * ncell + i], for the shell pass.
__global__ __launch_bounds__(kNT) void k_snapshot_adjoint_cell(
  const TsDev* __restrict__ tg, const float* __restrict__ shells, const float* __restrict__ deltax,
  const float* __restrict__ stars, const float* __restrict__ xe_prev, const float* __restrict__ Tk_prev,
  const float* __restrict__ xH, const float* __restrict__ TS, const float* __restrict__ xe_new,
  const float* __restrict__ gdT, float* __restrict__ g_delta, float* __restrict__ g_stars,
  float* __restrict__ g_xe_prev, float* __restrict__ g_Tk, double* __restrict__ gS, AdjCellArgs a)
{
  extern __shared__ double tabs[];  // the folded shell coefficients when ts_mode == 2 (ts_cell_lds_bytes)
  const TsDev& t = *tg;
  if (a.ts_mode == 2) load_ts_coef(t, tabs);
  __syncthreads();
  const int N = a.N;
  const long ncell = (long)N * N * N;
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= ncell) return;
  const long r = i / N;
  const int z = (int)(i % N);
  const long ip = r * (N + 2) + z;
  const float dl = deltax[ip];
  const double xh = xH[i];
  const double gd = gdT[i];
  const double dom = 1.0 + (double)dl;
  double F = 1.0, gTS = 0.0;
  double ts = 0.0;
  if (a.use_ts) {
    ts = TS[i];
    F = 1.0 - a.T_rad / ts;
    gTS = gd * (double)a.tb_const * xh * dom * a.T_rad / (ts * ts);
  }
  const double gxH = gd * (double)a.tb_const * dom * F;
  double gdl = gd * (double)a.tb_const * xh * F;
  double gst = 0.0, gxe_new = 0.0;
  if (xh > 0.0 && xh < 1.0) {  // partially ionised, never crossed: xH = n_f - eff f_coll
    const float st = stars[ip];
    const double fc_per_star = a.fc_unit / dom;
    if (dl > -1.0f + kRelTol) gdl += gxH * a.eff * (double)st * fc_per_star / dom;
    if (st >= kAbsTol) gst = -gxH * a.eff * fc_per_star;
    if (a.use_xe) {
      const float xe = xe_new[ip];
      if (xe >= kAbsTol && xe < 0.999f) gxe_new = -gxH;
    }
  }
  double gxe0 = 0.0, gT0 = 0.0;
  if (a.ts_mode == 2) {
    D7 xn;
    const D7 tsd = ts_cell_dual(t, tabs, dl, xe_prev[ip], Tk_prev[i], shells, ncell, i, &xn);
    double g[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) g[k] = gTS * tsd.d[k] + gxe_new * xn.d[k];
    gdl += g[0];
    gxe0 = g[1];
    gT0 = g[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) gS[(long)j * ncell + i] = g[3 + j];
  } else if (a.ts_mode == 1) {
    using D1 = Dual<1>;
    D1 xa;
    const D1 tsd = get_Ts_t<D1>(t, &t.kHH, D1::var((double)dl, 0), D1((double)a.h_T0), D1((double)a.h_xe0), D1(0.0),
                                &xa);
    gdl += gTS * tsd.d[0];
  }
  g_delta[ip] = (float)gdl;
  g_stars[ip] = (float)gst;
  if (g_xe_prev) g_xe_prev[ip] = (float)gxe0;
  if (g_Tk) g_Tk[i] = (float)gT0;
  if (z == N - 1) {  // padding of the padded gradients
    g_delta[ip + 1] = g_delta[ip + 2] = 0.0f;
    g_stars[ip + 1] = g_stars[ip + 2] = 0.0f;
    if (g_xe_prev) g_xe_prev[ip + 1] = g_xe_prev[ip + 2] = 0.0f;
  }
}

// Gradient of shell ct, dL/d(shell_ct) on the padded grid: the raw sums'
// adjoints times their coefficients at the cell's x_e, where the forward's
// max(v, 0) clamp passed the value (v > 0; the derivative is taken as 0 at
// v = 0).  `unclamped`, when given, is the padded grid v itself (the
// unfiltered shell 0 is the working grid, whose zeros the transform's
// round-off would otherwise make positive or negative at random).
__global__ __launch_bounds__(kNT) void k_ts_shell_grad(const TsDev* __restrict__ tg, const float* __restrict__ shells,
                                                      const float* __restrict__ xe_prev,
                                                      const double* __restrict__ gS, float* __restrict__ out, int ct,
                                                      int N, const float* __restrict__ unclamped)
{
  // this shell's column of the six x_e tables
  __shared__ double col[6][kNXHII];
  const TsDev& t = *tg;
  if (threadIdx.x < 6 * kNXHII) {
    const int tb = threadIdx.x / kNXHII, x = threadIdx.x % kNXHII;
    const double(*src)[kTsMaxSteps] =
      tb == 0 ? t.fheat : tb == 1 ? t.fion : tb == 2 ? t.flya : tb == 3 ? t.sheat : tb == 4 ? t.sion : t.slya;
    col[tb][x] = src[x][ct];
  }
  __syncthreads();
  const long ncell = (long)N * N * N;
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= ncell) return;
  const long r = i / N;
  const int z = (int)(i % N);
  const long ip = r * (N + 2) + z;
  const float sv = shells[(long)ct * ncell + i];
  double g = 0.0;
  if ((unclamped ? unclamped[ip] : sv) > 0.0f) {
    const XeBracket br = xe_bracket(t, xe_prev[ip]);
    double dd;
    const double fh = interp(col[0][br.lo], col[3][br.lo], br, &dd);
    const double fi = interp(col[1][br.lo], col[4][br.lo], br, &dd);
    const double fl = interp(col[2][br.lo], col[5][br.lo], br, &dd);
    const double dSFR = t.sfr_factor;
    const double w = t.dt_dzpp[ct] * t.dzpp[ct] * t.zpp_pow[ct];
    const double wl = t.P2 * t.zpp1[ct] * t.sum_lyn[ct] * t.dt_dzpp[ct] * t.dzpp[ct];
    g = dSFR * (gS[i] * w * fh + gS[ncell + i] * w * fi + gS[2 * ncell + i] * w * fl + gS[3 * ncell + i] * wl);
  }
  out[ip] = (float)g;
  if (z == N - 1) out[ip + 1] = out[ip + 2] = 0.0f;
}

// The shell tables' dependence on the two grid means they are built from
// (the collapse fraction, m = 0, and <x_e>, m = 1; ComputeTs.c:413-475, which
// enter through tau_X): per cell, sum over shells of v_ct times
// sum_j gS_j (E[m][lo][ct][j] + fr E[m][lo][ct][j + 3]), j = heat, ion, Ly-alpha,
// where E is the derivative of the six x_e tables with respect to mean m
// times the shell's weight (sfr_factor dt_dzpp dzpp zpp_pow), from the host;
// block sums of the nm values to partials[block * 2 + m].
__global__ __launch_bounds__(kNT) void k_ts_means_grad(const TsDev* __restrict__ tg, const float* __restrict__ shells,
                                                      const float* __restrict__ xe_prev,
                                                      const double* __restrict__ gS, const double* __restrict__ E,
                                                      int nm, int N, double* __restrict__ partials)
{
  __shared__ double red[2 * kNT / 64];
  const TsDev& t = *tg;
  const int ns = t.nsteps;
  const long ncell = (long)N * N * N;
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < ncell) {
    const long r = i / N;
    const int z = (int)(i % N);
    const XeBracket br = xe_bracket(t, xe_prev[r * (N + 2) + z]);
    const double g0 = gS[i], g1 = gS[ncell + i], g2 = gS[2 * ncell + i];
    for (int ct = 0; ct < ns; ++ct) {
      const double v = (double)shells[(long)ct * ncell + i];
      if (v == 0.0) continue;
      for (int m = 0; m < nm; ++m) {
        const double* e = E + (((long)m * kNXHII + br.lo) * ns + ct) * 6;
        acc[m] += v * (g0 * (e[0] + br.fr * e[3]) + g1 * (e[1] + br.fr * e[4]) + g2 * (e[2] + br.fr * e[5]));
      }
    }
  }
  block_sum_n<kNT, 2>(acc, red);
  if (threadIdx.x == 0) {
    partials[(long)blockIdx.x * 2] = acc[0];
    partials[(long)blockIdx.x * 2 + 1] = acc[1];
  }
}

// The means' gradients back to the cells: <x_e> is the mean of x_e_prev, so
// every cell's x_e_prev gradient gains gxe (dL/d<x_e> / N^3); the collapse
// fraction is the mean of min(c, 1), c = stars / (RtoM(R) h^2 (1 + delta))
// (4/3) pi R^3 / V (k_ts_fcoll), so the cells with c < 1 gain gcf c / stars
// on stars and -gcf c / (1 + delta) on delta (gcf = dL/d(collapse fraction) /
// N^3).
__global__ __launch_bounds__(kNT) void k_ts_means_apply(const float* __restrict__ deltax,
                                                       const float* __restrict__ stars, float* __restrict__ g_delta,
                                                       float* __restrict__ g_stars, float* __restrict__ g_xe_prev,
                                                       double gxe, double gcf, FcollArgs a)
{
  const long ncell = (long)a.N * a.N * a.N;
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i >= ncell) return;
  const long ip = (i / a.N) * (a.N + 2) + i % a.N;
  if (g_xe_prev) g_xe_prev[ip] += (float)gxe;
  if (gcf != 0.0) {
    const double dom = 1.0 + deltax[ip];
    const double k = 1.0 / (a.M_h2 * dom) * (4.0 / 3.0) * M_PI * a.R3 / a.pixel_volume;
    const double c = (double)stars[ip] * k;
    if (c < 1.0) {
      if (g_stars) g_stars[ip] += (float)(gcf * k);
      if (g_delta) g_delta[ip] -= (float)(gcf * c / dom);
    }
  }
}

// acc[i] += x[i] over n floats
__global__ void k_axpy(float* __restrict__ acc, const float* __restrict__ x, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] += x[i];
}

// Adjoint of one working-grid operation of k_sfr_op (same kind codes).
__global__ void k_sfr_op_adjoint(float* __restrict__ gW, float* __restrict__ ghist, int kind, double wd, float wf,
                                 int N, long half)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kind <= 1) {
    if (i >= half) return;
    gW[i] = (kind == 0) ? 0.0f : (float)((double)gW[i] * wd);
    return;
  }
  const long ncell = (long)N * N * N;
  if (i >= ncell) return;
  const long r = i / N;
  const int z = (int)(i % N);
  const long ip = r * (N + 2) + z;
  const float g = gW[ip];
  if (ghist) ghist[i] += wf * g;  // NULL: the history grid is not differentiated
  if (kind == 2) gW[ip] = 0.0f;
}

// ------------------------------------------------------- power spectrum
// Compute_PS (ComputePowerSpectrum.c:44-255): the field is made
// ((dT / <dT> - 1) V / N^3) <dT> in float, transformed, and every mode of the
// half spectrum adds k^3 |F|^2 / (2 pi^2 V) to its bin.  Bins: [0, dk),
// [dk, f dk), [f dk, f^2 dk) ... while the ceiling stays below dk N.
__global__ void k_ps_prep(const float* __restrict__ dT, float* __restrict__ out, int N, double ave, float volume)
{
  const long ncell = (long)N * N * N;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell) return;
  const long r = i / N;
  const int z = (int)(i % N);
  const long ip = r * (N + 2) + z;
  float v = (float)(((double)dT[i] / ave - 1) * (double)volume / (double)(float)((double)N * N * N));
  v = (float)((double)v * ave);
  out[ip] = v;
  if (z == N - 1) out[ip + 1] = out[ip + 2] = 0.0f;
}

constexpr int kMaxPsBins = 64;

struct PsArgs {
  int N;
  int nbins;
  float delta_k;
  double volume;
  float ceil[kMaxPsBins];  // bin ceilings (the floor of bin b is ceil[b-1], 0 for b = 0)
};

__device__ __forceinline__ int ps_bin(const PsArgs& a, float k_mag)
{
  float fl = 0.0f;
  for (int b = 0; b < a.nbins; ++b) {
    if (k_mag >= fl && k_mag < a.ceil[b]) return b;
    fl = a.ceil[b];
  }
  return -1;
}

__device__ __forceinline__ float ps_kmag(const PsArgs& a, int nx, int ny, int nz)
{
  const int H = a.N / 2;
  const float kx = (float)(nx > H ? nx - a.N : nx) * a.delta_k;
  const float ky = (float)(ny > H ? ny - a.N : ny) * a.delta_k;
  const float kz = (float)nz * a.delta_k;
  return (float)sqrt((double)(kx * kx + ky * ky + kz * kz));
}

// The mode of linear index m of the half spectrum (kz fastest, N/2+1 of
// them): its value from the packed spectrum U / Nyquist plane V.
__device__ __forceinline__ float2 ps_mode(const float2* U, const float2* V, int N, long m, int* nx, int* ny, int* nz)
{
  const int NH1 = N / 2 + 1;
  *nz = (int)(m % NH1);
  const long xy = m / NH1;
  *ny = (int)(xy % N);
  *nx = (int)(xy / N);
  return (*nz < N / 2) ? U[xy * (N / 2) + *nz] : V[(long)(*ny) * N + *nx];
}

// Block partial sums [block][bin][3]: p, k, count.  Each thread bins
// kPsJ modes once, then every bin is summed over the block in a fixed order.
constexpr int kPsJ = 16;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kNT) void k_ps_bin(const float2* __restrict__ U, const float2* __restrict__ V,
                                               double* __restrict__ partials, PsArgs a)
{
  __shared__ double red[kNT / 64][3];
  const long nmodes = (long)a.N * a.N * (a.N / 2 + 1);
  int bin[kPsJ];
  double pc[kPsJ];
  float kc[kPsJ];
#pragma unroll
  for (int j = 0; j < kPsJ; ++j) {
    const long m = ((long)blockIdx.x * kPsJ + j) * kNT + threadIdx.x;
    bin[j] = -1;
    pc[j] = 0.0;
    kc[j] = 0.0f;
    if (m < nmodes) {
      int nx, ny, nz;
      const float2 f = ps_mode(U, V, a.N, m, &nx, &ny, &nz);
      const float k_mag = ps_kmag(a, nx, ny, nz);
      bin[j] = ps_bin(a, k_mag);
      const double amp = sqrt((double)f.x * f.x + (double)f.y * f.y);
      pc[j] = ::pow((double)k_mag, 3) * ::pow(amp, 2.) / (2.0 * M_PI * M_PI * a.volume);
      kc[j] = k_mag;
    }
  }
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  for (int b = 0; b < a.nbins; ++b) {
    double p = 0.0, ks = 0.0, c = 0.0;
#pragma unroll
    for (int j = 0; j < kPsJ; ++j)
      if (bin[j] == b) {
        p += pc[j];
        ks += kc[j];
        c += 1.0;
      }
    p = wave_sum(p);
    ks = wave_sum(ks);
    c = wave_sum(c);
    if (lane == 0) {
      red[wv][0] = p;
      red[wv][1] = ks;
      red[wv][2] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double v = 0.0;
      for (int w = 0; w < kNT / 64; ++w) v += red[w][threadIdx.x];
      partials[((long)blockIdx.x * a.nbins + b) * 3 + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// Adjoint of the binned power: the half spectrum F is replaced in place by
// the spectrum G whose C2R is dL/d(delta_T) - G = s_k F on interior modes and
// 2 s_k F on the kz = 0 and kz = N/2 planes (which C2R counts once), with
// s_k = (V/N^3) k^3 / (2 pi^2 V) lam_b / n_b, lam_b = dL/dP_b.
__global__ void k_ps_adjoint_scale(float2* __restrict__ U, float2* __restrict__ V, const double* __restrict__ lam_n,
                                   PsArgs a, double vol_over_n3)
{
  const long nmodes = (long)a.N * a.N * (a.N / 2 + 1);
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmodes) return;
  int nx, ny, nz;
  (void)ps_mode(U, V, a.N, m, &nx, &ny, &nz);
  const float k_mag = ps_kmag(a, nx, ny, nz);
  const int b = ps_bin(a, k_mag);
  double s = 0.0;
  if (b >= 0) s = vol_over_n3 * ::pow((double)k_mag, 3) / (2.0 * M_PI * M_PI * a.volume) * lam_n[b];
  if (nz == 0 || nz == a.N / 2) s *= 2.0;
  const int NH1 = a.N / 2 + 1;
  const long xy = m / NH1;
  float2* p = (nz < a.N / 2) ? &U[xy * (a.N / 2) + nz] : &V[(long)ny * a.N + nx];
  *p = make_float2((float)((double)p->x * s), (float)((double)p->y * s));
}

#pragma clang fp contract(on)

}  // namespace mrx
