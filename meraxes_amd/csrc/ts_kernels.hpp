// ts_kernels.hpp - HIP kernels of the spin-temperature step (_ComputeTs).
//
// The shell loop reuses the bubble finder's transform kernels: the working SFR
// grid is transformed, windowed at each shell radius and brought back to real
// space, and the clamped shell (float) is stored shell-major, [shell][cell],
// so the per-cell kernel reads one coalesced row per shell.  The reference
// keeps the shells as double (ComputeTs.c:420-433); the double it stores is a
// fixed product of the float shell value, so the per-cell kernel rebuilds that
// double from the float (to within 2 ulp, one folded factor instead of the
// reference's two divisions) and the store costs 4 instead of 8 bytes.
//
// The per-cell kernel is ComputeTs.c:743-900 with evolveInt
// (XRayHeatingFunctions.c:1241-1435) and get_Ts (:1513-1558) in FP64; every
// cell-independent sub-expression that the reference evaluates first (a
// left-to-right prefix) is evaluated once on the host and passed in, and the
// reference's divisions by T_K, T_S, 1 + x_e and snapshot constants are
// products with reciprocals (within two ulp of each quotient; x_e, T_K and
// T_S are rounded to float).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "reion_kernels.hpp"
#include "ts_tables.h"

namespace mrx {

constexpr int kTsMaxSteps = 64;
constexpr int kNXHII = 14;

// a natural cubic spline of log(kappa) against log(T_K), its knots and
// coefficients rounded to float: the rates are evaluated in single precision
struct KappaSpline {
  int n;
  float x[27], y[27], b[27], c[27], d[27];
};

// the LDS copy of the packed shell coefficients (the [lo][ct][6] table, then
// cst): kTsCoefRounds 16-byte loads per lane over kTsCoefLanes lanes cover
// 3 kNXHII kTsMaxSteps + kTsMaxSteps / 2 double2
constexpr int kTsCoefLanes = 256;
constexpr int kTsCoefRounds = (3 * kNXHII * kTsMaxSteps + kTsMaxSteps / 2 + kTsCoefLanes - 1) / kTsCoefLanes;

// Everything the per-cell kernel needs, uploaded once per snapshot.
struct TsDev {
  int nsteps;
  int no_light;
  // shell value -> SMOOTHED_SFR double, the reference's ((f / V) * MT) * L3 / MSUN
  // (ComputeTs.c:420-433) folded into one factor, (1/V) MT L3 / MSUN: a
  // multiply instead of two FP64 divisions per shell and cell, within 2 ulp
  double pixel_volume, MT, L3, sfr_factor;
  // per-shell factors
  double dt_dzpp[kTsMaxSteps], dzpp[kTsMaxSteps], zpp1[kTsMaxSteps], zpp_pow[kTsMaxSteps], sum_lyn[kTsMaxSteps];
  double fheat[kNXHII][kTsMaxSteps], fion[kNXHII][kTsMaxSteps], flya[kNXHII][kTsMaxSteps];
  // slopes of the x_e interpolation, (f[lo+1] - f[lo]) / (XH[lo+1] - XH[lo]):
  // the quotient the reference forms per cell, formed once on the host with
  // the same double operation, so it is bit-identical
  double sheat[kNXHII][kTsMaxSteps], sion[kNXHII][kTsMaxSteps], slya[kNXHII][kTsMaxSteps];
  float XH[kNXHII];
  // snapshot constants
  double dzp, dt_dzp, const_zp_prefactor, growth, dgrowth;
  double NBP3;      // N_b0 * pow(1 + zp_f, 3), 1 + zp_f in float
  double P2;        // pow(1 + zp_f, 2), 1 + zp_f in float
  double dadia0;    // 3 / (1.0 + zp_f)
  double Trad;      // TCMB (1.0 + zp_f)
  double T21_Trad;  // T21 / Trad
  double P3;        // pow(1.0 + zp_f, 3.0)
  double TGP;       // (1.342881e-7 / hubble((float)zp)) * No * pow(1 + z, 3)
  double XA;        // 1.66e11 / (1.0 + zp_f)
  double HH;        // hubble((float)zp) / (HUBBLE h)
  double T4;        // pow(Trad, 4.0)
  double invTrad;   // 1 / Trad
  double comp;      // T4 / (HH hubble_h one_p_z): the Compton term's snapshot factor
  double one_p_z;   // 1.0 + zp_f
  double Tk_floor;  // (float)(TCMB (1 + zp)) with zp double
  double No, N_b0, f_H, f_He, hubble_h;
  float zp_f;
  KappaSpline kHH, keH, kpH;
  // the shell sums folded per (x_e bracket, shell): w_c = sfr_factor dt_dzpp dzpp zpp_pow
  // times {fheat, fion, flya, sheat, sion, slya}[lo][c], so that with the float
  // shell value v_c the cell's heat sum is sum_c v_c cw[lo][c][0] + fr sum_c v_c cw[lo][c][3]
  // (the same terms as the reference's per-shell (slope fr + f) w SFR, summed in
  // another order); cst[c] = sfr_factor P2 (1 + zpp) sum_lyn dt_dzpp dzpp for the
  // stellar Ly-alpha sum
  double cw[kNXHII][kTsMaxSteps][6];
  double cst[kTsMaxSteps];
  // the per-cell kernels' LDS image, 16-byte aligned and padded to whole
  // rounds of the copy (kTsCoefRounds loads of 16 bytes per lane): cw packed
  // for the actual shell count, [lo][ct][6] with stride nsteps (6 kNXHII
  // nsteps doubles), then cst[0 .. nsteps)
  alignas(16) double cwp[2 * kTsCoefRounds * kTsCoefLanes];
};

#pragma clang fp contract(off)

// The scalar helpers below are templates over the arithmetic type R: double
// for the forward kernels, a forward-mode dual number for the adjoint, which
// then recomputes exactly the forward's primal values.  pval() is the primal
// value, used for every branch.
__device__ __forceinline__ double pval(double x) { return x; }

// f(x) with its derivative dfdx carried into R: double returns f; a dual
// number gets f and dfdx times x's derivatives (dual.hpp's chain)
__device__ __forceinline__ double chain(double, double f, double) { return f; }

// The spline at xv (float), and its slope in *dy
__device__ __forceinline__ float spline_eval_f(const KappaSpline& s, float xv, float* dy)
{
  int lo = 0, hi = s.n - 1;
  if (xv <= s.x[0]) {
    lo = 0;
  } else if (xv >= s.x[hi]) {
    lo = hi - 1;
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (s.x[mid] > xv)
        hi = mid;
      else
        lo = mid;
    }
  }
  const float dx = xv - s.x[lo];
  *dy = s.b[lo] + dx * (2.0f * s.c[lo] + 3.0f * dx * s.d[lo]);
  return s.y[lo] + dx * (s.b[lo] + dx * (s.c[lo] + dx * s.d[lo]));
}

// The collision rates in single precision: log T_K, the spline and exp in
// float, each within a few ulp, so a rate is within about 1e-6 of its double
// evaluation (the spline's value, log kappa ~ -25, carries the float's
// absolute error into the exponent).  They enter T_S through x_c only, and
// the per-cell results stay within 1e-6 of the double evaluation, far inside
// the 1e-4 per-cell tolerance; the double evaluation of the three rates took
// 1.9 of the per-cell kernel's 11 ms at 512^3.
//
// kappa_10 (H-H) with its below-table clamp and T^0.381 extrapolation
// (XRayHeatingFunctions.c:1684-1693); log(exp(y_n) (T / exp(x_n))^0.381) is
// written y_n + 0.381 (log T - x_n)
template <class R>
__device__ __forceinline__ R kappa_HH(const KappaSpline& s, R TK)
{
  const double T = pval(TK);
  const float lT = logf((float)T);
  float ans, dans = 0.0f;
  if (lT < s.x[0])
    ans = s.y[0];
  else if (lT > s.x[s.n - 1]) {
    ans = s.y[s.n - 1] + 0.381f * (lT - s.x[s.n - 1]);
    dans = 0.381f;
  } else
    ans = spline_eval_f(s, lT, &dans);
  const double k = (double)expf(ans);
  return chain(TK, k, k * (double)dans / T);
}

// kappa_10_elec / kappa_10_pH with linear-in-log extrapolation (:1754-1766, :1827-1839)
template <class R>
__device__ __forceinline__ R kappa_file(const KappaSpline& s, R TK)
{
  const int n = s.n;
  const double T = pval(TK);
  const float lT = logf((float)T);
  float ans, dans = 0.0f;
  if (lT < s.x[0])
    ans = s.y[0];
  else if (lT > s.x[n - 1]) {
    dans = (s.y[n - 1] - s.y[n - 2]) / (s.x[n - 1] - s.x[n - 2]);
    ans = s.y[n - 1] + dans * (lT - s.x[n - 1]);
  } else
    ans = spline_eval_f(s, lT, &dans);
  const double k = (double)expf(ans);
  return chain(TK, k, k * (double)dans / T);
}

// get_Ts (XRayHeatingFunctions.c:1513-1558) at this snapshot's z, on the
// float inputs widened to R; the caller rounds the result to float.  Its
// divisions by T_K, T_S, T_rad, Sden and 2.85e-15 are products with
// reciprocals formed once (each within two ulp of the quotient, far below the
// float the result is rounded to): three divisions before the fixed point and
// two per iteration, against eleven and four.
// ks: the three collision-rate splines kHH, keH, kpH (t's own, or a copy of
// them in LDS, where the per-lane knot searches are cheaper)
template <class R>
__device__ __forceinline__ R get_Ts_t(const TsDev& t, const KappaSpline* ks, R dl, R TK, R xe, R Jalpha, R* xa_out)
{
  constexpr double kInvA10 = 1.0 / 2.85e-15;
  const double invTrad = t.invTrad;
  const R invTK = 1.0 / TK;
  // xcoll = xcoll_HI + xcoll_elec + xcoll_prot (:1560-1600)
  const R nH = (1.0 - xe) * t.No * t.P3 * (1.0 + dl);
  const R xHI = t.T21_Trad * nH * kappa_HH(ks[0], TK) * kInvA10;
  const R ne = xe * t.N_b0 * t.P3 * (1.0 + dl);
  const R xel = t.T21_Trad * ne * kappa_file(ks[1], TK) * kInvA10;
  const R np = xe * t.No * t.P3 * (1.0 + dl);
  const R xpr = t.T21_Trad * np * kappa_file(ks[2], TK) * kInvA10;
  const R xc = xHI + xel + xpr;
  R TS;
  if (pval(Jalpha) > 1.0e-20) {
    R TSold(0.0), xa(0.0);
    TS = R(t.Trad);
    R invTS = R(invTrad);
    const R tgp = t.TGP * (1.0 + dl) * (1.0 - xe);
    // the terms of the iteration that do not involve TS, evaluated once
    // the cube root in single precision (cbrtf, within 1 ulp): xi is a
    // correction to S of a few per cent, so its float rounding moves T_S by
    // well under 1e-7
    const R xarg = 1.0e-7 * tgp * invTK * invTK;
    const double xa0 = pval(xarg), xif = (double)cbrtf((float)xa0);
    const R xi = chain(xarg, xif, xa0 > 0.0 ? xif / (3.0 * xa0) : 0.0);
    const R S0 = 1.0 - 0.0631789 * invTK + 0.115995 * invTK * invTK;
    const R Sden = 1.0 + 2.98394 * xi + 1.53583 * xi * xi + 3.85289 * xi * xi * xi;
    const R ceTK = 0.405535 * invTK, xcTK = xc * invTK;
    // xa / Tceff as xa * ce (Tceff = 1 / ce); 1 / TS of the next iteration is
    // this one's TSinv
    const R aTK = 0.401403 * invTK, bTK = 0.336463 * invTK * invTK, invSden = 1.0 / Sden;
    const R XJ = t.XA * Jalpha;
    // the reference's test |TS - TSold| / TS > 1e-3 (false for TS <= 0 or NaN)
    while (pval(TS) > 0.0 && fabs(pval(TS) - pval(TSold)) > 1.0e-3 * pval(TS)) {
      TSold = TS;
      const R S = (S0 - aTK * invTS + bTK * invTS) * invSden;
      xa = XJ * S;
      const R ce = invTK + ceTK * (invTS - invTK);
      invTS = (invTrad + xa * ce + xcTK) / (1.0 + xa + xc);
      TS = 1.0 / invTS;
    }
    *xa_out = xa;
  } else {
    const R TSinv = (invTrad + xc * invTK) / (1.0 + xc);
    TS = 1.0 / TSinv;
    *xa_out = R(0.0);
  }
  if (pval(TS) < 0.) TS = fabs(TS);
  return TS;
}

// float in, float out, as the reference calls it
__device__ __forceinline__ float dev_get_Ts(const TsDev& t, const KappaSpline* ks, float delta, float TK_f,
                                           float xe_f, float Jalpha, float* xa_out)
{
  double xa;
  const double TS = get_Ts_t<double>(t, ks, (double)delta, (double)TK_f, (double)xe_f, (double)Jalpha, &xa);
  *xa_out = (float)xa;
  return (float)TS;
}

template <class R>
__device__ __forceinline__ R alpha_A_t(R T)
{
  const R logT = log(T * (1.0 / (double)1.1604505e4));
  // the reference's pow(logT, k) as products (each within a few ulp of pow,
  // which is itself not correctly rounded); the sum keeps its order
  const R p2 = logT * logT, p3 = p2 * logT, p4 = p2 * p2, p5 = p4 * logT, p6 = p3 * p3, p7 = p6 * logT,
          p8 = p4 * p4, p9 = p8 * logT;
  return exp(-28.6130338 - 0.72411256 * logT - 2.02604473e-2 * p2 - 2.38086188e-3 * p3 - 3.21260521e-4 * p4 -
             1.42150291e-5 * p5 + 4.98910892e-6 * p6 + 5.75561414e-7 * p7 - 1.85676704e-8 * p8 -
             3.07113524e-9 * p9);
}

__device__ __forceinline__ double dev_alpha_A(double T) { return alpha_A_t<double>(T); }

// 2 / 3 / k_B of evolveInt's heating term (XRayHeatingFunctions.c:1241-1435)
constexpr double kHeatFac = 2.0 / 3.0 / 1.3806e-16;

// Per-cell spin-temperature update (ComputeTs.c:743-900); 7 block partial sums:
// J_alpha, x_alpha, Xheat, Xion, T_S, T_K, x_e.
// Dynamic LDS: the folded shell coefficients cw, [kNXHII][nsteps][6] doubles,
// the only constants indexed per lane (by the cell's x_e bracket), then the
// stellar Ly-alpha coefficients cst[nsteps] (read in the shell loop, where a
// scalar load would stall every group of shells); everything else is read
// uniformly from the constants in global memory.
constexpr size_t ts_cell_lds_bytes(int nsteps)
{
  return sizeof(double) * ((6 * kNXHII + 1) * (size_t)nsteps + (size_t)(nsteps & 1));
}

// Copies the packed coefficients into LDS (3 kNXHII nsteps double2), 16 bytes
// per lane and load, a lane's loads issued together; blockDim.x == kTsCoefLanes.
__device__ __forceinline__ void load_ts_coef(const TsDev& t, double* tabs)
{
  static_assert(kTsCoefLanes == kNT, "one copy lane per thread");
  const int n2 = 3 * kNXHII * t.nsteps + (t.nsteps + 1) / 2;
  const double2* src = reinterpret_cast<const double2*>(t.cwp);
  double2* dst = reinterpret_cast<double2*>(tabs);
  double2 v[kTsCoefRounds];
#pragma unroll
  for (int u = 0; u < kTsCoefRounds; ++u) v[u] = src[u * kTsCoefLanes + threadIdx.x];  // in the padding
#pragma unroll
  for (int u = 0; u < kTsCoefRounds; ++u) {
    const int k = u * kTsCoefLanes + (int)threadIdx.x;
    if (k < n2) dst[k] = v[u];
  }
}

// Copies the three collision-rate splines (kHH, keH, kpH, adjacent in TsDev)
// into LDS in 4-byte words over the workgroup's lanes.
__device__ __forceinline__ void load_kappa(const TsDev& t, KappaSpline* ks)
{
  constexpr int kWords = 3 * (int)sizeof(KappaSpline) / 4;
  static_assert(offsetof(TsDev, keH) == offsetof(TsDev, kHH) + sizeof(KappaSpline) &&
                  offsetof(TsDev, kpH) == offsetof(TsDev, keH) + sizeof(KappaSpline),
                "the splines are adjacent");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&t.kHH);
  uint32_t* dst = reinterpret_cast<uint32_t*>(ks);
  for (int w = (int)threadIdx.x; w < kWords; w += (int)blockDim.x) dst[w] = src[w];
}

// The cell's seven shell sums at its x_e bracket `lo` (ComputeTs.c:759-790,
// evolveInt's 40-term sums): value and slope parts of heat, ion and Ly-alpha,
// and the stellar Ly-alpha sum.  One fused multiply-add per term.
struct ShellSums {
  double hf, hs, xf, xs, lf, ls, st;
};

// shell ct's terms of the seven sums; row: the coefficients at the cell's
// x_e bracket, cst: the stellar Ly-alpha ones; v: the shell's (clamped) value
__device__ __forceinline__ void shell_term(ShellSums& r, const double2* row, const double* cst, int ct, float vf)
{
  const double v = (double)vf;
  const double2 a = row[3 * ct], b = row[3 * ct + 1], c = row[3 * ct + 2];
  r.hf = fma(v, a.x, r.hf);
  r.xf = fma(v, a.y, r.xf);
  r.lf = fma(v, b.x, r.lf);
  r.hs = fma(v, b.y, r.hs);
  r.xs = fma(v, c.x, r.xs);
  r.ls = fma(v, c.y, r.ls);
  r.st = fma(v, cst[ct], r.st);
}

__device__ __forceinline__ ShellSums ts_shell_sums(const TsDev& t, const double* tabs, int lo, const float* shells,
                                                   long ncell, long i)
{
  ShellSums r = {0, 0, 0, 0, 0, 0, 0};
  const int ns = t.nsteps;
  const double2* row = reinterpret_cast<const double2*>(tabs + (size_t)lo * ns * 6);
  const double* cst = tabs + (size_t)6 * kNXHII * ns;
  // shells in groups of kG, the next group's loads in flight while this one
  // is summed (the loads' latency is the kernel's main wait)
  constexpr int kG = 16;
  float nxt[kG];
#pragma unroll
  for (int u = 0; u < kG; ++u) nxt[u] = u < ns ? shells[(long)u * ncell + i] : 0.0f;
  for (int c0 = 0; c0 < ns; c0 += kG) {
    float cur[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) cur[u] = nxt[u];
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      const int cn = c0 + kG + u;
      nxt[u] = cn < ns ? shells[(long)cn * ncell + i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      const int ct = c0 + u;
      if (ct >= ns) break;
      shell_term(r, row, cst, ct, cur[u]);
    }
  }
  return r;
}

// the table bracket of the cell's x_e (ComputeTs.c:759-790): lo and fr
__device__ __forceinline__ int ts_bracket(const TsDev& t, float xe0, double* fr)
{
  double xHII_call = xe0;
  if (xHII_call > t.XH[kNXHII - 1] * 0.999)
    xHII_call = t.XH[kNXHII - 1] * 0.999;
  else if (xHII_call < t.XH[0])
    xHII_call = 1.001 * t.XH[0];
  int lo = kNXHII - 1;
  const float xc_f = (float)xHII_call;
  while (xc_f < t.XH[lo]) lo--;
  *fr = xHII_call - t.XH[lo];
  return lo;
}

// The update of one cell (ComputeTs.c:743-900) from its shell sums S (unused
// when no_light): the new x_e, T_K and T_S, and the cell's terms of the seven
// block sums (J_alpha, x_alpha, Xheat, Xion, T_S, T_K, x_e).  dl, xe0, T: the
// cell's delta, previous x_e and T_K; fr: its x_e bracket offset.
__device__ __forceinline__ void ts_cell_update(const TsDev& t, const KappaSpline* ks, const ShellSums& S, float dl,
                                               float xe0, double T, double fr, float& xn_out, float& Tn_out,
                                               float& ts_out, double (&sum)[7])
{
  const double x_e = xe0;
  const double n_b = t.NBP3 * (1 + dl);
  double dxheat = 0, dxion = 0, dxlya = 0, dstarlya = 0;
  if (!t.no_light) {
    dxheat = (S.hf + fr * S.hs) * t.const_zp_prefactor;
    dxion = (S.xf + fr * S.xs) * t.const_zp_prefactor;
    dxlya = (S.lf + fr * S.ls) * (t.const_zp_prefactor * n_b);
    dstarlya = S.st * ((2.9979e10 / (4. * M_PI)) / (1.6726e-24 / 1.989e33));
  }
  const double sink = dev_alpha_A(T) * 2.0 * x_e * x_e * t.f_H * n_b;
  const double dxe_dzp = t.dt_dzp * (dxion - sink);
  // the divisions by 1 + x_e, by the snapshot constants and by 1/dl + growth
  // as products (each within two ulp; x_e and T_K are rounded to float)
  const double inv1x = 1.0 / (1.0 + x_e);
  double dadia = t.dadia0;
  if (fabs(dl) > 1e-7) dadia += t.dgrowth * dl / (1.0 + t.growth * dl);
  dadia *= (2.0 / 3.0) * T;
  const double dspec = -dxe_dzp * T * inv1x;
  double dcomp = (-1.51e-4) * (x_e / (1.0 + x_e + t.f_He)) * t.comp;
  dcomp *= t.Trad - T;
  const double dxheat_dzp = dxheat * t.dt_dzp * kHeatFac * inv1x;
  const double der1 = dxheat_dzp + dcomp + dspec + dadia;
  const double der2 = dxlya + dstarlya;
  const double der4 = t.dt_dzp * dxion;
  float xn = (float)((double)xe0 + dxe_dzp * t.dzp);
  if (xn > 1)
    xn = (float)(1 - 1e-7);
  else if (xn < 0)
    xn = 0;
  float Tn = (float)T;
  if (Tn < (float)5e4) Tn = (float)((double)Tn + der1 * t.dzp);
  if (Tn < 0) Tn = (float)t.Tk_floor;
  float xa;
  const float ts = dev_get_Ts(t, ks, dl, Tn, xn, (float)der2, &xa);
  xn_out = xn;
  Tn_out = Tn;
  ts_out = ts;
  sum[0] += der2;
  sum[1] += xa;
  sum[2] += dxheat_dzp;
  sum[3] += der4;
  sum[4] += ts;
  sum[5] += Tn;
  sum[6] += xn;
}

// One cell per thread (a grid-stride loop keeps the snapshot constants live
// across iterations and halves the occupancy).
__global__ __launch_bounds__(kNT) void k_ts_cell(const TsDev* __restrict__ tg, const float* __restrict__ shells,
                                                const float* __restrict__ deltax, float* __restrict__ xe_prev,
                                                float* __restrict__ Tk, float* __restrict__ TS, int N,
                                                double* __restrict__ partials)
{
  extern __shared__ double tabs[];
  __shared__ double red[7 * kNT / 64];
  __shared__ KappaSpline ks[3];
  const TsDev& t = *tg;
  load_ts_coef(t, tabs);
  load_kappa(t, ks);
  __syncthreads();
  const long ncell = (long)N * N * N;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i < ncell) {
    const long r = i / N;
    const int z = (int)(i % N);
    const long ip = r * (N + 2) + z;
    const float dl = deltax[ip];
    const float xe0 = xe_prev[ip];
    const double T = Tk[i];
    double fr;
    const int lo = ts_bracket(t, xe0, &fr);
    ShellSums S = {0, 0, 0, 0, 0, 0, 0};
    if (!t.no_light) S = ts_shell_sums(t, tabs, lo, shells, ncell, i);
    float xn, Tn, ts;
    ts_cell_update(t, ks, S, dl, xe0, T, fr, xn, Tn, ts, v);
    xe_prev[ip] = xn;
    Tk[i] = Tn;
    TS[i] = ts;
  }
  double* p = partials + (long)blockIdx.x * 7;
  block_sum_n<kNT, 7>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) p[k] = v[k];
  }
}

// Homogeneous branch (ComputeTs.c:189-217): RECFAST x_e and T_K everywhere, T_S
// from collisions only; block sums of T_S, T_K, x_e (slots 4..6).
__global__ __launch_bounds__(kNT) void k_ts_homogeneous(const TsDev* __restrict__ tg, const float* __restrict__ deltax,
                                                       float* __restrict__ xe_prev, float* __restrict__ Tk,
                                                       float* __restrict__ TS, int N, float xe0, float T0,
                                                       double* __restrict__ partials)
{
  __shared__ double red[kNT];
  const TsDev& t = *tg;
  const long ncell = (long)N * N * N;
  double sTS = 0, sTK = 0, sxe = 0;
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < ncell; i += (long)gridDim.x * kNT) {
    const long r = i / N;
    const int z = (int)(i % N);
    const long ip = r * (N + 2) + z;
    float xa;
    const float ts = dev_get_Ts(t, &t.kHH, deltax[ip], T0, xe0, 0.0f, &xa);
    xe_prev[ip] = xe0;
    Tk[i] = T0;
    TS[i] = ts;
    sTS += ts;
    sTK += T0;
    sxe += xe0;
  }
  double* p = partials + (long)blockIdx.x * 7;
  const double v4 = block_sum<kNT>(sTS, red), v5 = block_sum<kNT>(sTK, red), v6 = block_sum<kNT>(sxe, red);
  if (threadIdx.x == 0) {
    p[0] = p[1] = p[2] = p[3] = 0.0;
    p[4] = v4;
    p[5] = v5;
    p[6] = v6;
  }
}

// Mean collapse fraction of the first shell and <x_e> (ComputeTs.c:413-468):
// per cell min(stars / (RtoM(R) h^2 (1+delta)) * (4/3) pi R^3 / V, 1).
struct FcollArgs {
  double M_h2;         // RtoM(R) * h * h
  double R3_over_V;    // factor pieces, applied in the reference's order
  double R3;           // pow(R, 3.0)
  double pixel_volume;
  int N;
  int with_xe;
};

__global__ __launch_bounds__(kNT) void k_ts_fcoll(const float* __restrict__ stars, const float* __restrict__ deltax,
                                                 const float* __restrict__ xe_prev, double* __restrict__ partials,
                                                 FcollArgs a)
{
  __shared__ double red[kNT];
  double sf = 0.0, sx = 0.0;
  // rows of the grid over the workgroups, cells of a row over the threads
  // (no per-cell index division)
  const long nrow = (long)a.N * a.N;
  for (long r = blockIdx.x; r < nrow; r += gridDim.x)
  for (int z = threadIdx.x; z < a.N; z += kNT) {
    const long ip = r * (a.N + 2) + z;
    const double dom = 1.0 + deltax[ip];
    double c = stars[ip] / (a.M_h2 * dom) * (4.0 / 3.0) * M_PI * a.R3 / a.pixel_volume;
    if (c > 1.0) c = 1.0;
    sf += c;
    if (a.with_xe) sx += xe_prev[ip];
  }
  const double v0 = block_sum<kNT>(sf, red), v1 = block_sum<kNT>(sx, red);
  if (threadIdx.x == 0) {
    partials[(long)blockIdx.x * 2] = v0;
    partials[(long)blockIdx.x * 2 + 1] = v1;
  }
}

// Working SFR grid operations (ComputeTs.c:296-376, reionization.c:1756-1784).
// kind 0: zero the first `half` floats; 1: scale them by a double weight;
// 2: load hist (read at the unpadded index) * float weight into the padded
// cell; 3: add it.
__global__ void k_sfr_op(float* __restrict__ W, const float* __restrict__ hist, int kind, double wd, float wf, int N,
                         long half)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kind <= 1) {
    if (i >= half) return;
    W[i] = (kind == 0) ? 0.0f : (float)((double)W[i] * wd);
    return;
  }
  const long ncell = (long)N * N * N;
  if (i >= ncell) return;
  const long r = i / N;
  const int z = (int)(i % N);
  const float v = hist[i] * wf;
  const long ip = r * (N + 2) + z;
  W[ip] = (kind == 2) ? v : W[ip] + v;
}

// The operations of one shell (ComputeTs.c:276-377, in plan order) in one
// pass: W[ip] = the operations applied in order to src[ip], every padded
// element written (src may be W, or the input grid the working grid starts
// as).  Each element sees k_sfr_op's float operations in the same order.
struct SfrOps {
  int n;
  int kind[MRX_TS_MAX_OPS];
  int slot[MRX_TS_MAX_OPS];
  double wd[MRX_TS_MAX_OPS];
  float wf[MRX_TS_MAX_OPS];
};

// Cells are taken in pairs (z, z+1), z even: a padded row of N + 2 floats
// holds N/2 + 1 pairs, the last the padding, and `half` (a whole number of
// rows' worth of floats, even) never splits a pair.
__global__ __launch_bounds__(kNT) void k_sfr_ops(float* __restrict__ W, const float* src,
                                                const float* __restrict__ hist, long hstride, SfrOps o, int N,
                                                long half)
{
  const unsigned ph = (unsigned)(N / 2 + 1);
  const unsigned npair = (unsigned)N * (unsigned)N * ph;
  const unsigned p = blockIdx.x * kNT + threadIdx.x;
  if (p >= npair) return;
  const unsigned r = p / ph;
  const unsigned z = 2 * (p - r * ph);
  const bool real = z < (unsigned)N;
  const bool lo = 2 * (long)p < half;
  const long i = (long)r * N + z;
  float2 v = reinterpret_cast<const float2*>(src)[p];
  for (int k = 0; k < o.n; ++k) {
    const int kind = o.kind[k];
    if (kind <= 1) {
      if (lo) v = (kind == 0) ? make_float2(0.0f, 0.0f)
                              : make_float2((float)((double)v.x * o.wd[k]), (float)((double)v.y * o.wd[k]));
    } else if (real) {
      const float2 h = *reinterpret_cast<const float2*>(hist + (long)o.slot[k] * hstride + i);
      const float2 hw = make_float2(h.x * o.wf[k], h.y * o.wf[k]);
      v = (kind == 2) ? hw : make_float2(v.x + hw.x, v.y + hw.y);
    }
  }
  reinterpret_cast<float2*>(W)[p] = v;
}

#pragma clang fp contract(on)

}  // namespace mrx
