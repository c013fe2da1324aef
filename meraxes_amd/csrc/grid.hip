// grid.hip - galaxy gridding on the device: construct_baryon_grids
// (src/core/reionization.c:1353-1615) for one process holding the whole box.
//
// The reference walks the galaxies in order and adds each one's value to a
// float buffer cell, `buffer[ind] += v` with v double, so every cell's value is
// a sequence of float roundings in galaxy order.  Here each galaxy gets its
// cell as a key (pos_to_ngp of its position, misc_tools.c:163-173; dead
// galaxies get the key N^3), the (key, galaxy index) pairs are sorted by key
// with a stable radix sort (rocPRIM), and the first galaxy of every run of
// equal keys sums its run in galaxy order with the reference's roundings and
// writes the cell's final values into the padded grids (the reference's final
// pass, :1530-1600).  The grids are therefore identical to the reference's,
// bit for bit, whatever the number of galaxies per cell.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/mrx_reion.h"
#include "ctx.hpp"

namespace mrx {
namespace {

constexpr int kGT = 256;

struct GridArgs {
  int N;
  int bh;    // Flag_BHFeedback
  int spin;  // Flag_IncludeSpinTemp
  unsigned ncell;  // N^3, also the key of a dead galaxy
  double box_size;
  double bh_limit;  // BlackHoleMassLimitReion
  double a_bh, a_uv;  // ReionAlphaUVBH, ReionAlphaUV
  double ts;  // sfr_timescale
};

#pragma clang fp contract(off)
// pos_to_ngp (misc_tools.c:163-173): nearbyint in the default rounding mode
// (ties to even), an index above N-1 wraps to 0; -1 for an index the
// reference's assert rejects (negative, or not a number).
__device__ __forceinline__ int ngp(float x, double side, int nx)
{
  const double v = rint((double)x / side * (double)nx);
  if (!(v > -1.0)) return -1;
  if (v > (double)(nx - 1)) return 0;
  return (int)v;
}

// Keys and counts: counts[0] dead galaxies, counts[1] galaxies below the black
// hole mass limit (N_BlackHoleMassLimitReion, :1473-1478); *bad = the lowest
// index of a galaxy whose grid index is negative.  Grid-stride, with the
// counts reduced per workgroup (one atomic per workgroup and counter).
__global__ __launch_bounds__(kGT) void k_grid_keys(const float* __restrict__ pos, const int* __restrict__ type,
                                                   const double* __restrict__ bhm, GridArgs a, long n,
                                                   unsigned* __restrict__ keys, unsigned* __restrict__ vals,
                                                   unsigned long long* __restrict__ counts, int* __restrict__ bad)
{
  __shared__ unsigned red[2][kGT / 64];
  unsigned n_dead = 0, n_below = 0;
  for (long g = (long)blockIdx.x * kGT + threadIdx.x; g < n; g += (long)gridDim.x * kGT) {
    unsigned key = a.ncell;
    if (type[g] > 2) {
      ++n_dead;
    } else {
      const int ix = ngp(pos[3 * g], a.box_size, a.N);
      const int iy = ngp(pos[3 * g + 1], a.box_size, a.N);
      const int iz = ngp(pos[3 * g + 2], a.box_size, a.N);
      if (ix < 0 || iy < 0 || iz < 0)
        atomicMin(bad, (int)g);
      else
        key = (unsigned)iz + (unsigned)a.N * ((unsigned)iy + (unsigned)a.N * (unsigned)ix);
      if (a.bh && !(bhm[g] >= a.bh_limit)) ++n_below;
    }
    keys[g] = key;
    vals[g] = (unsigned)g;
  }
  for (int o = 32; o > 0; o >>= 1) {
    n_dead += __shfl_xor(n_dead, o);
    n_below += __shfl_xor(n_below, o);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = n_dead;
    red[1][w] = n_below;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned d = 0, b = 0;
    for (int k = 0; k < kGT / 64; ++k) {
      d += red[0][k];
      b += red[1][k];
    }
    if (d) atomicAdd(&counts[0], (unsigned long long)d);
    if (b) atomicAdd(&counts[1], (unsigned long long)b);
  }
}

// The galaxy values in sorted (cell, galaxy) order, so that each cell's run is
// summed from contiguous memory.
__global__ __launch_bounds__(kGT) void k_grid_gather(const unsigned* __restrict__ sval, long n,
                                                     const double* __restrict__ fesc,
                                                     const double* __restrict__ gross,
                                                     const double* __restrict__ bhm,
                                                     const double* __restrict__ ebhm, GridArgs a,
                                                     double* __restrict__ f_s, double* __restrict__ g_s,
                                                     double* __restrict__ m_s, double* __restrict__ e_s)
{
  const long q = (long)blockIdx.x * kGT + threadIdx.x;
  if (q >= n) return;
  const unsigned g = sval[q];
  f_s[q] = fesc[g];
  if (a.spin) g_s[q] = gross[g];
  if (a.bh) {
    m_s[q] = bhm[g];
    e_s[q] = ebhm[g];
  }
}

// Real cells of one padded grid set to 0 (the new slot 0 of the history ring
// is overwritten at every real cell, its padding kept).
__global__ __launch_bounds__(kGT) void k_zero_real(float* __restrict__ g, int N)
{
  const long i = (long)blockIdx.x * kGT + threadIdx.x;
  if (i >= (long)N * N * N) return;
  const long r = i / N;
  g[r * (N + 2) + (i % N)] = 0.0f;
}

// The first galaxy of each run of equal keys sums the run in galaxy order and
// writes the cell (reionization.c:1436-1515 accumulation, :1530-1600 final
// pass); the values come in sorted order from k_grid_gather.
__global__ __launch_bounds__(kGT) void k_grid_accumulate(const unsigned* __restrict__ skey, long n,
                                                         const double* __restrict__ f_s,
                                                         const double* __restrict__ g_s,
                                                         const double* __restrict__ m_s,
                                                         const double* __restrict__ e_s, GridArgs a,
                                                         float* __restrict__ stars, float* __restrict__ wsfr,
                                                         float* __restrict__ sfr, float* __restrict__ hist0)
{
  const long p = (long)blockIdx.x * kGT + threadIdx.x;
  if (p >= n) return;
  const unsigned key = skey[p];
  if (key >= a.ncell) return;
  if (p > 0 && skey[p - 1] == key) return;
  float st = 0.0f, ws = 0.0f, sf = 0.0f;
  for (long q = p; q < n && skey[q] == key; ++q) {
    const double f = f_s[q];
    st = (float)((double)st + f);  // prop_stellar
    ws = (float)((double)ws + f);  // prop_weighted_sfr
    if (a.bh) {
      if (m_s[q] >= a.bh_limit) {
        const double e = e_s[q];
        st = (float)((double)st + e);
        ws = (float)((double)ws + e * a.a_bh / a.a_uv);
      }
    }
    if (a.spin) sf = (float)((double)sf + g_s[q]);  // prop_sfr
  }
  const unsigned N = (unsigned)a.N;
  const long pad = (long)(key / N) * (N + 2) + key % N;  // INDEX_PADDED of the INDEX_REAL key
  stars[pad] = (st < 0) ? 0.0f : st;
  double v = (double)ws;
  v = (v > 0) ? v / a.ts : 0;
  wsfr[pad] = (float)v;
  if (a.spin) {
    double u = (double)sf;
    u = (u > 0) ? u / a.ts : 0;
    sfr[pad] = (float)u;
    hist0[pad] = (float)u;
  }
}
#pragma clang fp contract(on)

#define G_TRY(call)                                                                                   \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return ctx_fail(ctx, MRX_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                                      \
  } while (0)

// device buffer slots of ctx_buffer used here
enum { S_KEY, S_VAL, S_KEY2, S_VAL2, S_SORT, S_CNT, S_POS, S_TYPE, S_FESC, S_GROSS, S_BHM, S_EBHM, S_STARS,
       S_WSFR, S_SFR, S_HIST, S_FS, S_GS, S_MS, S_ES };

template <class T>
int stage_in(mrx_ctx* ctx, int slot, const T* host, size_t n, const T** dev)
{
  T* d = static_cast<T*>(ctx_buffer(ctx, slot, sizeof(T) * n));
  if (!d) return MRX_ENOMEM;
  if (n) G_TRY(hipMemcpyAsync(d, host, sizeof(T) * n, hipMemcpyHostToDevice, ctx_stream(ctx)));
  *dev = d;
  return MRX_OK;
}

}  // namespace
}  // namespace mrx

using namespace mrx;

extern "C" int mrx_construct_baryon_grids(mrx_ctx* ctx, const mrx_grid_params* p, long n_gal, const float* pos,
                                          const int* type, const double* fesc_weighted_gsm,
                                          const double* gross_stellar_mass, const double* black_hole_mass,
                                          const double* effective_bhm, float* stars, float* weighted_sfr, float* sfr,
                                          float* sfr_histories, mrx_grid_stats* stats)
{
  if (!ctx) return MRX_EINVAL;
  if (!p) return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: null params");
  const int N = p->grid_dim;
  if (N < 2 || (N & 1) || N > 1600)
    return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: grid_dim %d must be even, 2..1600", N);
  if (n_gal < 0 || n_gal >= (1L << 31))
    return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: n_gal %ld outside [0, 2^31)", n_gal);
  if (!(p->box_size > 0)) return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: box_size must be > 0");
  if (n_gal > 0 && (!pos || !type || !fesc_weighted_gsm))
    return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: pos, type and fesc_weighted_gsm are required");
  if (p->include_spin_temp && ((n_gal > 0 && !gross_stellar_mass) || !sfr || !sfr_histories || p->n_store_sfr < 1))
    return ctx_fail(ctx, MRX_EINVAL,
                    "construct_baryon_grids: include_spin_temp needs gross_stellar_mass, sfr, sfr_histories and "
                    "n_store_sfr >= 1");
  if (p->bh_feedback && n_gal > 0 && (!black_hole_mass || !effective_bhm))
    return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: bh_feedback needs black_hole_mass and effective_bhm");
  if (!stars || !weighted_sfr) return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: null output grid");
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ctx_fail(ctx, MRX_EDEVICE, "hipSetDevice failed");
  hipStream_t st = ctx_stream(ctx);
  const bool spin = p->include_spin_temp != 0, bh = p->bh_feedback != 0;
  const size_t np = (size_t)2 * N * N * (N / 2 + 1);
  const size_t ng = (size_t)n_gal;

  // sfr_timescale = ReionSfrTimescale * hubble_time(snapshot) (virial_properties.c:79-93)
  const double zplus1 = p->redshift + 1;
  const double H = p->hubble * sqrt(p->omega_m * zplus1 * zplus1 * zplus1 + p->omega_k * zplus1 * zplus1 +
                                    p->omega_lambda);
  GridArgs a;
  a.N = N;
  a.bh = bh ? 1 : 0;
  a.spin = spin ? 1 : 0;
  a.ncell = (unsigned)N * N * N;
  a.box_size = p->box_size;
  a.bh_limit = p->bh_mass_limit_reion;
  a.a_bh = p->alpha_uv_bh;
  a.a_uv = p->alpha_uv;
  a.ts = p->reion_sfr_timescale * (1.0 / H);

  // inputs where they lie, or staged
  const float* dpos = pos;
  const int* dtype = type;
  const double *dfesc = fesc_weighted_gsm, *dgross = gross_stellar_mass, *dbhm = black_hole_mass,
               *debhm = effective_bhm;
  float *dstars = stars, *dwsfr = weighted_sfr, *dsfr = sfr, *dhist = sfr_histories;
  const bool host = !p->pointers_on_device;
  const size_t nstore = spin ? (size_t)p->n_store_sfr : 0;
  if (host) {
    int rc;
    if ((rc = stage_in(ctx, S_POS, pos, 3 * ng, &dpos)) || (rc = stage_in(ctx, S_TYPE, type, ng, &dtype)) ||
        (rc = stage_in(ctx, S_FESC, fesc_weighted_gsm, ng, &dfesc)))
      return rc;
    if (spin && (rc = stage_in(ctx, S_GROSS, gross_stellar_mass, ng, &dgross))) return rc;
    if (bh && ((rc = stage_in(ctx, S_BHM, black_hole_mass, ng, &dbhm)) ||
               (rc = stage_in(ctx, S_EBHM, effective_bhm, ng, &debhm))))
      return rc;
    dstars = static_cast<float*>(ctx_buffer(ctx, S_STARS, sizeof(float) * np));
    dwsfr = static_cast<float*>(ctx_buffer(ctx, S_WSFR, sizeof(float) * np));
    if (!dstars || !dwsfr) return MRX_ENOMEM;
    if (spin) {
      dsfr = static_cast<float*>(ctx_buffer(ctx, S_SFR, sizeof(float) * np));
      dhist = static_cast<float*>(ctx_buffer(ctx, S_HIST, sizeof(float) * np * nstore));
      if (!dsfr || !dhist) return MRX_ENOMEM;
      G_TRY(hipMemcpyAsync(dhist, sfr_histories, sizeof(float) * np * nstore, hipMemcpyHostToDevice, st));
    }
  } else if ((reinterpret_cast<uintptr_t>(dpos) | reinterpret_cast<uintptr_t>(dfesc)) & 7) {
    return ctx_fail(ctx, MRX_EINVAL, "construct_baryon_grids: device arrays must be naturally aligned");
  }

  // keys, counts and the validity check before any output is touched
  unsigned* keys = static_cast<unsigned*>(ctx_buffer(ctx, S_KEY, sizeof(unsigned) * (ng + 1)));
  unsigned* vals = static_cast<unsigned*>(ctx_buffer(ctx, S_VAL, sizeof(unsigned) * (ng + 1)));
  unsigned* keys2 = static_cast<unsigned*>(ctx_buffer(ctx, S_KEY2, sizeof(unsigned) * (ng + 1)));
  unsigned* vals2 = static_cast<unsigned*>(ctx_buffer(ctx, S_VAL2, sizeof(unsigned) * (ng + 1)));
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx_buffer(ctx, S_CNT, 64));
  if (!keys || !vals || !keys2 || !vals2 || !cnt) return MRX_ENOMEM;
  int* bad = reinterpret_cast<int*>(cnt + 2);
  const unsigned long long init[3] = {0ull, 0ull, 0x7fffffffull};
  G_TRY(hipMemcpyAsync(cnt, init, sizeof init, hipMemcpyHostToDevice, st));
  const unsigned nb = (unsigned)((ng + kGT - 1) / kGT);
  if (ng) {
    k_grid_keys<<<nb < 4096 ? nb : 4096, kGT, 0, st>>>(dpos, dtype, dbhm, a, n_gal, keys, vals, cnt, bad);
    G_TRY(hipGetLastError());
  }
  unsigned long long h[3];
  G_TRY(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, st));
  G_TRY(hipStreamSynchronize(st));
  const int first_bad = (int)(h[2] & 0xffffffffull);
  if (first_bad != 0x7fffffff)
    return ctx_fail(ctx, MRX_EINVAL,
                    "construct_baryon_grids: galaxy %d has a negative grid index (pos_to_ngp asserts, "
                    "misc_tools.c:170)",
                    first_bad);

  // the grids start at zero (:1376-1398); the SFR history ring moves by one slot
  G_TRY(hipMemsetAsync(dstars, 0, sizeof(float) * np, st));
  G_TRY(hipMemsetAsync(dwsfr, 0, sizeof(float) * np, st));
  if (spin) {
    G_TRY(hipMemsetAsync(dsfr, 0, sizeof(float) * np, st));
    for (long s = (long)nstore - 2; s >= 0; --s)
      G_TRY(hipMemcpyAsync(dhist + (size_t)(s + 1) * np, dhist + (size_t)s * np, sizeof(float) * np,
                           hipMemcpyDeviceToDevice, st));
    const long nr = (long)N * N * N;
    k_zero_real<<<(unsigned)((nr + kGT - 1) / kGT), kGT, 0, st>>>(dhist, N);
    G_TRY(hipGetLastError());
  }

  if (ng) {
    // stable sort of (cell, galaxy) by cell; the dead galaxies' key N^3 sorts last
    unsigned bits = 1;
    while (bits < 32 && ((unsigned long long)a.ncell >> bits) != 0) ++bits;
    size_t tmp_bytes = 0;
    G_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, vals, vals2, (unsigned)ng, 0, bits, st));
    void* tmp = ctx_buffer(ctx, S_SORT, tmp_bytes);
    if (!tmp) return MRX_ENOMEM;
    G_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, vals, vals2, (unsigned)ng, 0, bits, st));
    double* f_s = static_cast<double*>(ctx_buffer(ctx, S_FS, sizeof(double) * ng));
    double* g_s = spin ? static_cast<double*>(ctx_buffer(ctx, S_GS, sizeof(double) * ng)) : nullptr;
    double* m_s = bh ? static_cast<double*>(ctx_buffer(ctx, S_MS, sizeof(double) * ng)) : nullptr;
    double* e_s = bh ? static_cast<double*>(ctx_buffer(ctx, S_ES, sizeof(double) * ng)) : nullptr;
    if (!f_s || (spin && !g_s) || (bh && (!m_s || !e_s))) return MRX_ENOMEM;
    k_grid_gather<<<nb, kGT, 0, st>>>(vals2, n_gal, dfesc, dgross, dbhm, debhm, a, f_s, g_s, m_s, e_s);
    G_TRY(hipGetLastError());
    k_grid_accumulate<<<nb, kGT, 0, st>>>(keys2, n_gal, f_s, g_s, m_s, e_s, a, dstars, dwsfr, dsfr, dhist);
    G_TRY(hipGetLastError());
  }

  if (host) {
    G_TRY(hipMemcpyAsync(stars, dstars, sizeof(float) * np, hipMemcpyDeviceToHost, st));
    G_TRY(hipMemcpyAsync(weighted_sfr, dwsfr, sizeof(float) * np, hipMemcpyDeviceToHost, st));
    if (spin) {
      G_TRY(hipMemcpyAsync(sfr, dsfr, sizeof(float) * np, hipMemcpyDeviceToHost, st));
      G_TRY(hipMemcpyAsync(sfr_histories, dhist, sizeof(float) * np * nstore, hipMemcpyDeviceToHost, st));
    }
    G_TRY(hipStreamSynchronize(st));
  }
  if (stats) {
    stats->n_dead = (long)h[0];
    stats->n_gridded = n_gal - (long)h[0];
    stats->n_bh_below_limit = (long)h[1];
    stats->sfr_timescale = a.ts;
  }
  return ctx_finish(ctx);
}
