// mrx_reion.hip - host side of the C ABI declared in include/mrx_reion.h.
//
// One context per device.  It owns the work buffers (packed spectra U/T of up
// to four fields, their Nyquist planes V, the per-cell step state, twiddle
// tables, reduction partials and host-pointer staging), sized for the largest
// grid seen and kept across calls.  Every entry point validates its arguments
// on the host, enqueues on the context stream and, unless the caller installed
// its own stream, synchronises before returning.  Errors come back as MRX_E*
// codes with a message in mrx_last_error(); nothing aborts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mrx_reion.h"
#include "ctx.hpp"
#include "reion_kernels.hpp"
#include "ts_kernels.hpp"
#include "adjoint.hpp"
#include "ts_tables.h"
#include "recomb_tables.h"

using namespace mrx;

namespace {
// constants of src/meraxes.h:21-44 and reionization.h:38
constexpr double kPlanck = 6.6262e-27;
constexpr double kHubble = 3.2407789e-18;
constexpr double kSigmaHI = 6.3e-18;  // SIGMA_HI, reionization.h:11
constexpr double kProtonMass = 1.6726e-24;
constexpr double kTcmb = 2.728;
constexpr double kLFactor = 0.620350491;
constexpr int kMaxSteps = 255;  // the per-cell step index is one byte

// Workgroups of k_brightness_temp (grid-stride): one partial sum each, for
// both brightness-temperature entry points, so their sums are formed alike.
constexpr int kTbBlocks = 2048;

// Device scratch of mrx_brightness_temp_noalloc: the kTbBlocks partial sums of
// k_brightness_temp, then the reduced total.  Its size does not depend on the
// grid.  Every device holds its own copy until the process ends.  Host code
// never uses the array's name (on the host it is only a stand-in symbol): it
// uses the device address resolved once per device by resolve_scratch().
constexpr int kNoallocScratch = kTbBlocks + 8;  // doubles (16 KiB + 64 B)
constexpr int kNoallocMaxDim = 2048;            // largest grid_dim the no-allocation entry point accepts
static __device__ __align__(256) double scratch[kNoallocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Resolves this device's copy of `scratch`, with `device` current; called from
// mrx_init.  Returns the address or nullptr (then only the no-allocation entry
// point is unavailable, and it says so).
double* resolve_scratch(int device)
{
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    double* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

// Radii per inverse x pass of the bubble ladder (1..kLadderRadiiMax); the
// MRX_LADDER_RADII environment variable, read when a context is created,
// overrides it.  Two radii per pass bring the x pass to the y pass's rate at
// 512^3 (more gain nothing); the second radius's spectra cost nfield
// N^2 (N/2) float2 of device memory.
constexpr int kLadderRadii = 2;
constexpr int kLadderRadiiMax = 8;

struct mrx_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  bool user_stream = false;
  std::string err;

  int N = 0;          // grid the buffers are laid out for
  int nfield_cap = 0;  // fields U/T/V hold
  float2* U = nullptr;
  float2* T = nullptr;
  float2* V = nullptr;
  uint8_t* state = nullptr;
  float2* tw = nullptr;
  float* Rtab = nullptr;
  double* partials = nullptr;
  size_t partials_cap = 0;
  double* sums = nullptr;
  // staging for host pointers
  float* stage = nullptr;
  size_t stage_cap = 0;
  // kernel timing (mrx_profile): events recorded around each launch
  struct ProfRec {
    const char* name;
    hipEvent_t a, b;
  };
  bool prof = false;
  bool prof_ts = false;  // launches belong to the spin-temperature shell loop (names get a "ts_" prefix)
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> event_pool;
  // spin temperature: host tables (reloaded when the key changes) and device buffers
  mrx_ts_data* tsd = nullptr;
  std::string ts_key;
  mrx_ts_snap* tss = nullptr;
  TsDev* ts_host = nullptr;
  TsDev* ts_dev = nullptr;
  float* ts_work = nullptr;
  size_t ts_work_cap = 0;
  float* ts_shells = nullptr;
  size_t ts_shells_cap = 0;
  int ts_shells_N = 0, ts_shells_n = 0;  // grid and shell count ts_shells holds (0: none)
  float* wtab = nullptr;  // window tables of the current radii, (N/2+1)^3 floats each
  size_t wtab_cap = 0;
  // the bubble ladder's window tables of every filtered radius, built on a
  // stream of their own (default priority) while the fields are transformed
  float* wtab_all = nullptr;
  size_t wtab_all_cap = 0;
  hipStream_t side = nullptr;
  hipEvent_t side_ev[2] = {nullptr, nullptr};
  // the bubble ladder's inverse x pass covers ladder_radii radii per launch
  // (tabulated windows): their x-pass outputs, ladder_radii x nfield spectra
  int ladder_radii = kLadderRadii;
  float2* TK = nullptr;
  size_t TK_cap = 0;  // float2
  float ts_h_xe0 = 0.f, ts_h_T0 = 0.f;  // RECFAST x_e, T_K of the homogeneous branch
  // adjoint work buffers
  float* adj_work = nullptr;
  size_t adj_work_cap = 0;
  double* adj_gS = nullptr;
  size_t adj_gS_cap = 0;
  double* adj_E = nullptr;  // the shell tables' derivatives with respect to their grid means
  size_t adj_E_cap = 0;
  // recombination-rate tables (host, built once per No and Y_He) and their device copy
  mrx_rr_tables* rr = nullptr;
  RRDev* rr_dev = nullptr;
  // brightness temperature with velocities: gradient grid + line-of-sight bins
  float* tbv = nullptr;
  size_t tbv_cap = 0;
  // buffers of the other translation units (ctx_buffer)
  void* ext[mrx::kCtxSlots] = {};
  size_t ext_cap[mrx::kCtxSlots] = {};
  // this device's copy of the file-scope `scratch` array (mrx_brightness_temp_noalloc)
  double* noalloc_scratch = nullptr;
};

namespace {

int set_err(mrx_ctx* c, int code, const char* fmt, ...)
{
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIP_TRY(call)                                                                                \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return set_err(ctx, MRX_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                     __FILE__, __LINE__);                                                            \
  } while (0)

#define MRX_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != MRX_OK) return rc_; \
  } while (0)

bool supported_n(int N)
{
  switch (N) {
    case 32:
    case 48:
    case 64:
    case 96:
    case 128:
    case 192:
    case 256:
    case 384:
    case 512:
    case 160:
    case 320:
    case 640:
    case 1024:
      return true;
    default:
      return false;
  }
}

template <class T>
void dfree(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

int check_launch(mrx_ctx* ctx, const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, MRX_EDEVICE, "launch of %s failed: %s", what, hipGetErrorString(e));
  return MRX_OK;
}

// Sums the nb x nv row-major partials of each of nbox boxes (boxes
// contiguous) into out[box * nv + v]: in one workgroup per box for short
// arrays, otherwise in two levels (k_reduce_rows, then one workgroup per box)
// through `rows`, kRedBlocks * kRedMaxV * nbox doubles.  The order of the
// additions depends only on (nb, nv), never on timing.
int reduce_partials(mrx_ctx* ctx, hipStream_t st, const double* partials, int nb, int nv, int nbox, double* rows,
                    double* out)
{
  if (nv > kRedMaxV || nb < 32 * kNT || !rows) {
    k_reduce_partials<<<nbox, kNT, 0, st>>>(partials, nb, nv, out);
    return check_launch(ctx, "k_reduce_partials");
  }
  const int G = std::min(kRedBlocks, (nb + kNT - 1) / kNT);
  const int chunk = (nb + G - 1) / G;
  k_reduce_rows<<<dim3(G, nbox), kNT, 0, st>>>(partials, nb, nv, chunk, rows);
  MRX_TRY(check_launch(ctx, "k_reduce_rows"));
  k_reduce_partials<<<nbox, kNT, 0, st>>>(rows, G, nv, out);
  return check_launch(ctx, "k_reduce_partials");
}

// Workgroups of kNT threads of `kernel` (with `lds` bytes of dynamic LDS) that
// are resident on the context's device at once: the grid of a grid-stride
// kernel that fills the device in one wave.
template <class K>
int resident_blocks(mrx_ctx* ctx, K kernel, size_t lds)
{
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kNT, lds) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1) {
    (void)hipGetLastError();
    cus = 256;
  }
  return per_cu * cus;
}

int finish(mrx_ctx* ctx)
{
  if (!ctx->user_stream) HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

// Work buffers for an N^3 grid with `nfield` fields.
int ensure_buffers(mrx_ctx* ctx, int N, int nfield)
{
  if (ctx->N == N && ctx->nfield_cap >= nfield) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  dfree(ctx->U);
  dfree(ctx->T);
  dfree(ctx->V);
  dfree(ctx->state);
  dfree(ctx->tw);
  ctx->N = 0;
  ctx->nfield_cap = 0;
  const size_t nc = (size_t)N * N * (N / 2);
  if (hipMalloc(&ctx->U, sizeof(float2) * nc * nfield) != hipSuccess ||
      hipMalloc(&ctx->T, sizeof(float2) * nc * nfield) != hipSuccess ||
      hipMalloc(&ctx->V, sizeof(float2) * (size_t)N * N * nfield) != hipSuccess ||
      hipMalloc(&ctx->state, (size_t)N * N * N) != hipSuccess || hipMalloc(&ctx->tw, sizeof(float2) * N) != hipSuccess) {
    (void)hipGetLastError();
    dfree(ctx->U);
    dfree(ctx->T);
    dfree(ctx->V);
    dfree(ctx->state);
    dfree(ctx->tw);
    return set_err(ctx, MRX_ENOMEM, "device allocation for a %d^3 grid with %d fields failed", N, nfield);
  }
  std::vector<float2> h(N);
  for (int t = 0; t < N; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)N;
    h[t] = make_float2((float)cos(a), (float)sin(a));
  }
  HIP_TRY(hipMemcpy(ctx->tw, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  ctx->N = N;
  ctx->nfield_cap = nfield;
  return MRX_OK;
}

int ensure_partials(mrx_ctx* ctx, size_t n)
{
  if (ctx->partials_cap >= n) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  dfree(ctx->partials);
  ctx->partials_cap = 0;
  HIP_TRY(hipMalloc(&ctx->partials, sizeof(double) * n));
  ctx->partials_cap = n;
  return MRX_OK;
}

// The ladder's multi-radius spectra (float2 elements); a failed allocation is
// not an error (the ladder then runs one radius per x pass).
int ensure_tk(mrx_ctx* ctx, size_t n)
{
  if (ctx->TK_cap >= n) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (ctx->TK) (void)hipFree(ctx->TK);
  ctx->TK = nullptr;
  ctx->TK_cap = 0;
  if (hipMalloc(&ctx->TK, sizeof(float2) * n) != hipSuccess) {
    (void)hipGetLastError();
    ctx->TK = nullptr;
    return MRX_ENOMEM;
  }
  ctx->TK_cap = n;
  return MRX_OK;
}

int ensure_stage(mrx_ctx* ctx, size_t nfloat)
{
  if (ctx->stage_cap >= nfloat) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  dfree(ctx->stage);
  ctx->stage_cap = 0;
  HIP_TRY(hipMalloc(&ctx->stage, sizeof(float) * nfloat));
  ctx->stage_cap = nfloat;
  return MRX_OK;
}

hipEvent_t take_event(mrx_ctx* ctx)
{
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one kernel launch with events on the launch stream when profiling.
struct ProfScope {
  mrx_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr;
  ProfScope(mrx_ctx* c, const char* n) : ctx(c), name(n)
  {
    if (ctx->prof) {
      a = take_event(ctx);
      if (a) (void)hipEventRecord(a, ctx->stream);
    }
  }
  ~ProfScope()
  {
    if (ctx->prof && a) {
      hipEvent_t b = take_event(ctx);
      if (b) {
        (void)hipEventRecord(b, ctx->stream);
        ctx->prof_recs.push_back({name, a, b});
      }
    }
  }
};
#define PROF(name) ProfScope prof_scope_##__LINE__(ctx, ctx->prof_ts ? "ts_" name : name)

int ensure_buf(mrx_ctx* ctx, float*& p, size_t& cap, size_t nfloat)
{
  if (cap >= nfloat) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  dfree(p);
  cap = 0;
  if (hipMalloc(&p, sizeof(float) * nfloat) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, MRX_ENOMEM, "device allocation of %zu floats failed", nfloat);
  }
  cap = nfloat;
  return MRX_OK;
}

template <class T>
int ensure_dev(mrx_ctx* ctx, T*& p, size_t& cap, size_t n)
{
  if (cap >= n) return MRX_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  dfree(p);
  cap = 0;
  if (hipMalloc(&p, sizeof(T) * n) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, MRX_ENOMEM, "device allocation of %zu bytes failed", sizeof(T) * n);
  }
  cap = n;
  return MRX_OK;
}

size_t npad(int N) { return (size_t)N * N * (2 * (N / 2 + 1)); }
size_t nreal(int N) { return (size_t)N * N * N; }

WinArgs win_args(double box_size, float R, int type, int apply)
{
  WinArgs w;
  w.R = R;
  const float bs = (float)box_size;
  w.delta_k = (float)(2.0 * M_PI / bs);
  w.type = type;
  w.apply = apply;
  return w;
}

// Largest |n| component of a mode the sharp-k window of w can keep, capped at
// N/2.  A mode with a component m > nmax has 0.413566994 |k| R > 1 + 1e-5
// before rounding; the float evaluation of the window (relative error below
// 1e-6) then still exceeds 1, so filter() zeroes it.
int sharpk_nmax(const WinArgs& w, int N)
{
  const double kc = 1.0 / (0.413566994 * (double)w.R * (double)w.delta_k);
  const double m = std::floor(kc * (1.0 + 1e-5)) + 1.0;
  return (m >= (double)(N / 2)) ? N / 2 : (int)m;
}

// ----------------------------------------------------- templated pipelines
template <int N>
struct Pipe {
  static constexpr int NH = N / 2;
  static constexpr long nc = (long)N * N * NH;
  // the y passes (tile = (x, kz block), rows along y with stride N/2): at
  // 512^3 tiles 32 kz wide (256-byte row segments), one 1024-thread
  // workgroup per CU (793 us per four-field inverse pass against 847 us for
  // 16-wide tiles and two 512-thread workgroups per CU; the windowed x pass
  // measured the other way, 2063 against 1607 us per two-radius launch)
  static constexpr int TWY = (N == 512) ? 32 : Cfg<N>::TW;
  static constexpr int NTY = (N == 512) ? 1024 : kColNT;
  // (the forward y pass, one field per launch, keeps the 16-wide tiles: 199
  // against 205 us at 512^3)
  template <bool INV>
  static void colpass_y(mrx_ctx* ctx, float2* data, int nfield, double div)
  {
    const dim3 gy(N * (NH / TWY), nfield);
    k_colpass<N, INV, TWY, 0, NTY><<<gy, TileMap<N, TWY, NTY>::NT, 0, ctx->stream>>>(data, ctx->tw, nc, (long)N * NH,
                                                                                   NH, div);
  }

  // padded real fields -> unfiltered packed spectra U (scaled by 1/N^3 when
  // `scale`), Nyquist planes split into V.
  static int forward(mrx_ctx* ctx, const float* in, long in_fstride, int nfield, bool scale, float2* U, float2* V,
                     bool float_scale = false)
  {
    const dim3 gz(N * N / Cfg<N>::RZ, nfield);
    {
    PROF("r2c_z");
    k_r2c_z<N><<<gz, kNT, 0, ctx->stream>>>(in, U, ctx->tw, in_fstride, nc);
    }
    MRX_TRY(check_launch(ctx, "k_r2c_z"));
    const dim3 gc(N * (NH / Cfg<N>::TW), nfield);
    // y pass: tile = (x, kz block); rows along y with stride NH
    {
    PROF("colpass_y_fwd");
    k_colpass<N, false><<<gc, Cfg<N>::COL_NT, 0, ctx->stream>>>(U, ctx->tw, nc, (long)N * NH, NH, 0.0);
    }
    MRX_TRY(check_launch(ctx, "k_colpass y"));
    // x pass: tile = (y, kz block); rows along x with stride N*NH
    double div = scale ? std::pow((double)N, 3) : 0.0;
    if (scale && float_scale) div = -div;
    {
    PROF("colpass_x_fwd");
    k_colpass<N, false><<<gc, Cfg<N>::COL_NT, 0, ctx->stream>>>(U, ctx->tw, nc, NH, (long)N * NH, div);
    }
    MRX_TRY(check_launch(ctx, "k_colpass x"));
    const dim3 gu((N * N + 255) / 256, nfield);
    PROF("unpack_dc");
    k_unpack_dc<N><<<gu, 256, 0, ctx->stream>>>(U, V, nc, (long)N * N);
    return check_launch(ctx, "k_unpack_dc");
  }

  // forward() for one field whose spectrum a sharp-k window will cut to
  // |n_x|, |n_y|, n_z <= nmax (nmax < N/2): the y pass over the kz blocks
  // that hold kz <= nmax, the x pass over those blocks of the lines |ky| <=
  // nmax; the other tiles of U (and the Nyquist plane, which the window
  // zeroes) are left stale and must not be read.  The tiles it computes are
  // forward()'s, bit for bit.
  static int forward_sharpk(mrx_ctx* ctx, const float* in, float2* U, float2* V, int nmax)
  {
    constexpr int TW = Cfg<N>::TW;
    const int nkzb = std::min(NH / TW, nmax / TW + 1);
    {
      PROF("r2c_z");
      k_r2c_z<N><<<dim3(N * N / Cfg<N>::RZ, 1), kNT, 0, ctx->stream>>>(in, U, ctx->tw, 0, nc);
    }
    MRX_TRY(check_launch(ctx, "k_r2c_z"));
    {
      PROF("colpass_y_fwd");
      k_colpass_sub<N, false><<<dim3(N * nkzb, 1), Cfg<N>::COL_NT, 0, ctx->stream>>>(U, ctx->tw, nc, (long)N * NH,
                                                                                     NH, 0.0, nkzb, -1);
    }
    MRX_TRY(check_launch(ctx, "k_colpass_sub y"));
    const int nline = std::min(N, 2 * nmax + 1);
    {
      PROF("colpass_x_fwd");
      k_colpass_sub<N, false><<<dim3(nline * nkzb, 1), Cfg<N>::COL_NT, 0, ctx->stream>>>(
        U, ctx->tw, nc, NH, (long)N * NH, -std::pow((double)N, 3), nkzb, nmax);
    }
    MRX_TRY(check_launch(ctx, "k_colpass_sub x"));
    PROF("unpack_dc");
    k_unpack_dc<N><<<dim3((N * N + 255) / 256, 1), 256, 0, ctx->stream>>>(U, V, nc, (long)N * N);
    return check_launch(ctx, "k_unpack_dc");
  }

  // U (+V) filtered at R -> T, then inverse y pass; T then holds the packed
  // spectra of the filtered fields with x and y already transformed, in kz
  // slots below kzlive (the slots above are zero and left unwritten: a
  // sharp-k window keeps only a ball of modes, so only the tiles that hold
  // part of it are read, transformed and written).
  static int inverse_xy(mrx_ctx* ctx, const float2* U, const float2* V, float2* T, int nfield, const WinArgs& w,
                        int& kzlive)
  {
    kzlive = NH;
    if (w.apply && w.type == MRX_FILTER_SHARPK) {
      constexpr int TW = Cfg<N>::TW;
      const int nmax = sharpk_nmax(w, N);
      const int nkzb = std::min(NH / TW, nmax / TW + 1);
      const int nky = std::min(N, 2 * nmax + 1);
      {
        PROF("ix_filter");
        k_ix_filter<N, kWinSharpK><<<nky * nkzb, Cfg<N>::IX_NT, 0, ctx->stream>>>(
          U, V, T, ctx->tw, nullptr, nc, (long)N * N, nfield, 0, 0, w, nmax, nkzb);
      }
      MRX_TRY(check_launch(ctx, "k_ix_filter (sharp-k)"));
      {
        PROF("colpass_y_inv");
        k_colpass_live<N><<<dim3(N * nkzb, nfield), Cfg<N>::COL_NT, 0, ctx->stream>>>(T, ctx->tw, nc, (long)N * NH,
                                                                                    NH, nkzb, nmax);
      }
      MRX_TRY(check_launch(ctx, "k_colpass_live y inv"));
      kzlive = nkzb * TW;
      return MRX_OK;
    }
    const float* Wt = nullptr;
    if (w.apply) {
      const long nt = (long)(NH + 1) * (NH + 1) * (NH + 1);
      MRX_TRY(ensure_buf(ctx, ctx->wtab, ctx->wtab_cap, (size_t)nt));
      {
        PROF("window_table");
        k_window_table<N, Cfg<N>::TW><<<window_table_blocks<N, Cfg<N>::TW>(), kNT, 0, ctx->stream>>>(ctx->wtab, w);
      }
      MRX_TRY(check_launch(ctx, "k_window_table"));
      Wt = ctx->wtab;
    }
    const dim3 gc(N * (NH / Cfg<N>::TW), 1);
    {
    PROF("ix_filter");
    if (Wt)
      k_ix_filter<N, kWinTable><<<gc, Cfg<N>::IX_NT, 0, ctx->stream>>>(U, V, T, ctx->tw, Wt, nc, (long)N * N, nfield,
                                                                       0, 0, w, N, NH / Cfg<N>::TW);
    else
      k_ix_filter<N, kWinNone><<<gc, Cfg<N>::IX_NT, 0, ctx->stream>>>(U, V, T, ctx->tw, Wt, nc, (long)N * N, nfield,
                                                                      0, 0, w, N, NH / Cfg<N>::TW);
    }
    MRX_TRY(check_launch(ctx, "k_ix_filter"));
    PROF("colpass_y_inv");
    colpass_y<true>(ctx, T, nfield, 0.0);
    return check_launch(ctx, "k_colpass y inv");
  }

  // The inverse x and y passes of K ladder radii with tabulated windows ws[k]
  // (not sharp-k): one x pass reads U once and writes the K spectra TK + k
  // tstride, then each radius's y pass runs in place.  Per radius the spectra
  // are inverse_xy's.
  // wpre: the K radii's window tables already built (window_tables), or null
  static int inverse_xy_k(mrx_ctx* ctx, const float2* U, const float2* V, float2* TK, long tstride, int nfield,
                          const WinArgs* ws, int K, const float* wpre = nullptr)
  {
    constexpr int TW = Cfg<N>::TW;
    const long nt = (long)(NH + 1) * (NH + 1) * (NH + 1);
    const float* wt = wpre;
    if (!wt) {
      MRX_TRY(ensure_buf(ctx, ctx->wtab, ctx->wtab_cap, (size_t)nt * K));
      MRX_TRY(window_tables(ctx, ctx->wtab, ws, K));
      wt = ctx->wtab;
    }
    {
      PROF("ix_filter_k");
      k_ix_filter_k<N><<<N * (NH / TW) * K * nfield, Cfg<N>::COL_NT, 0, ctx->stream>>>(
        U, V, TK, ctx->tw, wt, nt, nc, (long)N * N, tstride, nfield, K);
    }
    MRX_TRY(check_launch(ctx, "k_ix_filter_k"));
    for (int k = 0; k < K; ++k) {
      PROF("colpass_y_inv");
      colpass_y<true>(ctx, TK + k * tstride, nfield, 0.0);
    }
    return check_launch(ctx, "k_colpass y inv");
  }

  // the window tables of K radii into Wt (k_window_table's layout, one table
  // per (N/2+1)^3 floats), on ctx->stream
  static int window_tables(mrx_ctx* ctx, float* Wt, const WinArgs* ws, int K)
  {
    constexpr int TW = Cfg<N>::TW;
    const long nt = (long)(NH + 1) * (NH + 1) * (NH + 1);
    for (int k = 0; k < K; ++k) {
      PROF("window_table");
      k_window_table<N, TW><<<window_table_blocks<N, TW>(), kNT, 0, ctx->stream>>>(Wt + k * nt, ws[k]);
    }
    return check_launch(ctx, "k_window_table");
  }

  static int c2r_z(mrx_ctx* ctx, const float2* T, float* out, int nfield, int clamp, int pitch = N + 2,
                   int kzlive = NH)
  {
    PROF("c2r_z");
    if constexpr (N == 256 || N == 512) {
      // the row transforms in registers
      constexpr int A = N / 32;
      k_c2r_z_reg<A><<<dim3(N * N / (256 / A), nfield), 256, 0, ctx->stream>>>(T, out, ctx->tw, nc,
                                                                              (long)N * N * pitch, clamp, pitch, kzlive);
      return check_launch(ctx, "k_c2r_z_reg");
    }
    const dim3 gz(N * N / Cfg<N>::RZ, nfield);
    k_c2r_z<N><<<gz, kNT, 0, ctx->stream>>>(T, out, ctx->tw, nc, (long)N * N * pitch, clamp, pitch, kzlive);
    return check_launch(ctx, "k_c2r_z");
  }

  // the ladder's last pass; fields deltax, stars, sfr (+ x_e) (+ N_rec when rec)
  // 256^3 and 512^3 without recombinations: the row transforms in registers
  // (k_c2r_threshold_reg); otherwise staged in LDS (k_c2r_threshold)
  static constexpr bool kRegThr = (N == 256 || N == 512);
  static int threshold(mrx_ctx* ctx, const float2* Tsrc, const float* deltax, float* xH, float* rb, float* J21,
                       float* zin, float* J21i, float* Gamma12, bool rec, const StepArgs& a)
  {
    const int nb = nblocks_threshold(rec);
    PROF("c2r_threshold");
    if constexpr (kRegThr) {
      if (!rec) {
        constexpr int A = N / 32;
        if (a.nfield == 4)
          k_c2r_threshold_reg<A, 4><<<nb, 256, 0, ctx->stream>>>(Tsrc, nc, ctx->tw, deltax, ctx->state, xH, rb, J21,
                                                                zin, J21i, ctx->Rtab, ctx->partials, a);
        else
          k_c2r_threshold_reg<A, 3><<<nb, 256, 0, ctx->stream>>>(Tsrc, nc, ctx->tw, deltax, ctx->state, xH, rb, J21,
                                                                zin, J21i, ctx->Rtab, ctx->partials, a);
        return check_launch(ctx, "k_c2r_threshold_reg");
      }
    }
    constexpr int NT = ThrCfg<N>::NT;
#define MRX_THR(NF_, REC_)                                                                                          \
  k_c2r_threshold<N, NF_, REC_><<<nb, NT, 0, ctx->stream>>>(Tsrc, nc, ctx->tw, deltax, ctx->state, xH, rb, J21,   \
                                                           zin, J21i, Gamma12, ctx->Rtab, ctx->partials, a)
    if (!rec && a.nfield == 4)
      MRX_THR(4, false);
    else if (!rec)
      MRX_THR(3, false);
    else if (a.nfield == 5)
      MRX_THR(5, true);
    else
      MRX_THR(4, true);
#undef MRX_THR
    return check_launch(ctx, "k_c2r_threshold");
  }

  static int nblocks_threshold(bool rec)
  {
    if constexpr (kRegThr) {
      if (!rec) return N * N / RegThr<N / 32>::RWG;
    }
    return N * N / ThrCfg<N>::RZ;
  }

  static int pack_herm(mrx_ctx* ctx, const float2* H, float2* U)
  {
    const long n = nc;
    k_herm_to_packed<N><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(H, U);
    return check_launch(ctx, "k_herm_to_packed");
  }
  static int unpack_herm(mrx_ctx* ctx, const float2* U, const float2* V, float2* H)
  {
    const long n = (long)N * N * (NH + 1);
    k_packed_to_herm<N><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(U, V, H);
    return check_launch(ctx, "k_packed_to_herm");
  }
  // inverse x then y pass on packed spectra in place (no window)
  static int inv_xy_plain(mrx_ctx* ctx, float2* T, int nfield)
  {
    const dim3 gc(N * (NH / Cfg<N>::TW), nfield);
    k_colpass<N, true><<<gc, Cfg<N>::COL_NT, 0, ctx->stream>>>(T, ctx->tw, nc, NH, (long)N * NH, 0.0);
    MRX_TRY(check_launch(ctx, "k_colpass x inv"));
    colpass_y<true>(ctx, T, nfield, 0.0);
    return check_launch(ctx, "k_colpass y inv");
  }
};

template <template <int> class F, class... A>
int dispatch_n(int N, A&&... args)
{
  switch (N) {
    case 32: return F<32>::run(args...);
    case 48: return F<48>::run(args...);
    case 64: return F<64>::run(args...);
    case 96: return F<96>::run(args...);
    case 128: return F<128>::run(args...);
    case 192: return F<192>::run(args...);
    case 256: return F<256>::run(args...);
    case 384: return F<384>::run(args...);
    case 512: return F<512>::run(args...);
    case 160: return F<160>::run(args...);
    case 320: return F<320>::run(args...);
    case 640: return F<640>::run(args...);
    case 1024: return F<1024>::run(args...);
    default: return MRX_EUNSUPPORTED;
  }
}

double RtoM(const mrx_hii_params* p, double R)
{
  // find_HII_bubbles.c:26-45
  if (p->rtom_filter_type == 0) return (4.0 / 3.0) * M_PI * pow(R, 3) * (p->omega_m * p->rho_crit);
  return pow(2 * M_PI, 1.5) * p->omega_m * p->rho_crit * pow(R, 3);
}

int ladder(const mrx_hii_params* p, std::vector<double>& R)
{
  // find_HII_bubbles.c:97-100, :184-207, :445
  const double box_size = p->box_size;
  const int N = p->grid_dim;
  double clf = kLFactor;
  if ((box_size / (double)N) < 1.0) clf = 1.0;
  double r = fmin(p->r_bubble_max, kLFactor * box_size);
  R.clear();
  bool last = false;
  while (!last) {
    if (((r / p->delta_r_factor) <= (clf * box_size / (double)N)) || ((r / p->delta_r_factor) <= p->r_bubble_min)) {
      last = true;
      r = clf * box_size / (double)N;
    }
    R.push_back(r);
    if (R.size() > 100000) return MRX_EINVAL;
    r /= p->delta_r_factor;
  }
  return (int)R.size();
}

int validate_hii(mrx_ctx* ctx, const mrx_hii_params* p)
{
  if (!p) return set_err(ctx, MRX_EINVAL, "params is NULL");
  if (!supported_n(p->grid_dim))
    return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported (32,48,64,96,128,192,256,384,512)", p->grid_dim);
  if (p->filter_type < 0 || p->filter_type > 2)
    return set_err(ctx, MRX_EFILTER, "ReionFilterType type %d is undefined", p->filter_type);
  if (p->rtom_filter_type < 0 || p->rtom_filter_type > 1)
    return set_err(ctx, MRX_EINVAL, "Unrecognised RtoM filter (%d)", p->rtom_filter_type);
  if (!(p->box_size > 0) || !(p->delta_r_factor > 1.0) || !(p->r_bubble_max > 0))
    return set_err(ctx, MRX_EINVAL, "box_size, r_bubble_max must be > 0 and delta_r_factor > 1");
  return MRX_OK;
}

// The recombination-rate table for the cosmology of p, on the device.
int rr_tables(mrx_ctx* ctx, const mrx_hii_params* p, const RRDev** out)
{
  const double No = mrx_rr_No(p->hubble_h, p->omega_m, p->baryon_frac, p->y_he);
  if (!ctx->rr || ctx->rr->No != No || ctx->rr->y_he != p->y_he) {
    if (!ctx->rr) ctx->rr = new mrx_rr_tables();
    if (mrx_rr_init(ctx->rr, No, p->y_he) != 0) {
      delete ctx->rr;
      ctx->rr = nullptr;
      return set_err(ctx, MRX_EINVAL, "recombination-rate tables failed");
    }
    if (!ctx->rr_dev) HIP_TRY(hipMalloc(&ctx->rr_dev, sizeof(RRDev)));
    static_assert(sizeof(RRDev) == sizeof(double) * (kRRG + 4 * kRRZ * kRRG), "RRDev layout");
    static_assert(kRRZ == MRX_RR_Z_NPTS && kRRG == MRX_RR_LNG_NPTS, "table sizes");
    std::vector<double> h(kRRG + 4 * (size_t)kRRZ * kRRG);
    memcpy(h.data(), ctx->rr->lnG, sizeof(double) * kRRG);
    const size_t row = (size_t)kRRZ * kRRG;
    memcpy(h.data() + kRRG, ctx->rr->y, sizeof(double) * row);
    memcpy(h.data() + kRRG + row, ctx->rr->b, sizeof(double) * row);
    memcpy(h.data() + kRRG + 2 * row, ctx->rr->c, sizeof(double) * row);
    memcpy(h.data() + kRRG + 3 * row, ctx->rr->d, sizeof(double) * row);
    HIP_TRY(hipMemcpy(ctx->rr_dev, h.data(), sizeof(RRDev), hipMemcpyHostToDevice));
  }
  *out = ctx->rr_dev;
  return MRX_OK;
}

struct HiiRun {
  template <int N>
  struct At {
    // in: deltax, stars, weighted_sfr, x_e, N_rec; out: xH, r_bubble, J_21,
    // z_at_ionization, J_21_at_ionization, Gamma12 (the last in and out only
    // with recombinations)
    static int run(mrx_ctx* ctx, const mrx_hii_params* p, const float* const* in, float* const* out,
                   mrx_hii_globals* g)
    {
      using P = Pipe<N>;
      const bool rec = p->include_recombinations != 0;
      const int nfield = 3 + (p->include_spin_temp ? 1 : 0) + (rec ? 1 : 0);
      MRX_TRY(ensure_buffers(ctx, N, nfield > 4 ? nfield : 4));
      std::vector<double> R;
      const int S = ladder(p, R);
      if (S < 1) return set_err(ctx, MRX_EINVAL, "radius ladder is empty");
      if (S > kMaxSteps) return set_err(ctx, MRX_EUNSUPPORTED, "radius ladder has %d steps (max %d)", S, kMaxSteps);
      MRX_TRY(ensure_partials(ctx, (size_t)P::nblocks_threshold(rec) * 4 + 8 + kRedBlocks * kRedMaxV));
      if (!ctx->Rtab) HIP_TRY(hipMalloc(&ctx->Rtab, sizeof(float) * 256));
      if (!ctx->sums) HIP_TRY(hipMalloc(&ctx->sums, sizeof(double) * 8));
      std::vector<float> Rf(S);
      for (int s = 0; s < S; ++s) Rf[s] = (float)R[s];
      HIP_TRY(hipMemcpyAsync(ctx->Rtab, Rf.data(), sizeof(float) * S, hipMemcpyHostToDevice, ctx->stream));

      // inputs: deltax, stars, weighted_sfr, x_e are padded and laid out at a
      // fixed stride when staged; device pointers are used where they lie.
      const long fs = (long)npad(N);
      const float* deltax = in[0];
      const float* src[5] = {in[0], in[1], in[2], p->include_spin_temp ? in[3] : in[4], in[4]};
      const double box_size = p->box_size;
      // the tabulated windows (top-hat, Gaussian) of every filtered radius
      // are built on the side stream while the fields are transformed: the
      // table kernel is bound by its FP64 sincos, the transforms by memory
      // (up to 4 GiB of tables; beyond that each x pass builds its own)
      const long nt = (long)(N / 2 + 1) * (N / 2 + 1) * (N / 2 + 1);
      const float* wpre = nullptr;
      if (p->filter_type != MRX_FILTER_SHARPK && ctx->ladder_radii >= 1 && S > 1 &&
          (size_t)nt * (S - 1) * sizeof(float) <= ((size_t)4 << 30) &&
          ensure_buf(ctx, ctx->wtab_all, ctx->wtab_all_cap, (size_t)nt * (S - 1)) == MRX_OK) {
        if (!ctx->side) HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        for (auto& e : ctx->side_ev)
          if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        std::vector<WinArgs> ws(S - 1);
        for (int k = 0; k < S - 1; ++k) ws[k] = win_args(box_size, (float)R[k], p->filter_type, 1);
        HIP_TRY(hipEventRecord(ctx->side_ev[0], ctx->stream));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->side_ev[0], 0));
        hipStream_t const main_stream = ctx->stream;
        ctx->stream = ctx->side;
        const int rc = P::window_tables(ctx, ctx->wtab_all, ws.data(), S - 1);
        ctx->stream = main_stream;
        const hipError_t er = (rc == MRX_OK) ? hipEventRecord(ctx->side_ev[1], ctx->side) : hipSuccess;
        if (rc != MRX_OK || er != hipSuccess) {
          // the tables already launched finish before the buffer can be reused
          (void)hipStreamSynchronize(ctx->side);
          if (rc != MRX_OK) return rc;
          return set_err(ctx, MRX_EDEVICE, "hipEventRecord (window tables) failed: %s", hipGetErrorString(er));
        }
        wpre = ctx->wtab_all;
      }
      // the main stream waits for the tables on every exit from here, so no
      // later call can reuse their buffer while they are being written
      struct WaitSide {
        mrx_ctx* c;
        bool on;
        ~WaitSide()
        {
          if (on) (void)hipStreamWaitEvent(c->stream, c->side_ev[1], 0);
        }
      } wait_side{ctx, wpre != nullptr};
      for (int f = 0; f < nfield; ++f) {
        // forward transform of each field into its U slot
        MRX_TRY(P::forward(ctx, src[f], 0, 1, true, ctx->U + (long)f * P::nc, ctx->V + (long)f * N * N));
      }
      if (wpre) {
        HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->side_ev[1], 0));
        wait_side.on = false;
      }
      (void)fs;
      const double pixel_volume = pow(box_size / (double)N, 3);
      StepArgs a;
      a.pixel_volume = pixel_volume;
      a.eff = p->efficiency;
      a.redshift = p->redshift;
      a.bias = p->gamma_halo_bias;
      a.uvb = p->uvb_flag;
      a.use_xe = p->include_spin_temp;
      a.nfield = nfield;
      const double z = p->redshift;
      // the filtered radii go through the x pass ladder_radii at a time
      // (tabulated windows; KL = 0: sharp-k, one radius per pass over its live
      // tiles) when the spectra of that many radii fit, else one at a time
      int KL = (p->filter_type == MRX_FILTER_SHARPK) ? 0 : ctx->ladder_radii;
      const long tstride = (long)nfield * P::nc;
      if (KL > 1 && ensure_tk(ctx, (size_t)KL * tstride) != MRX_OK) KL = 1;
      for (int s = 0; s < S;) {
        const bool last = (s == S - 1);
        int K = 1;
        const float2* Tsrc = ctx->T;
        if (!last && KL >= 1) {
          K = std::min(KL, S - 1 - s);
          WinArgs ws[kLadderRadiiMax];
          for (int k = 0; k < K; ++k) ws[k] = win_args(box_size, (float)R[s + k], p->filter_type, 1);
          float2* TK = KL > 1 ? ctx->TK : ctx->T;
          MRX_TRY(P::inverse_xy_k(ctx, ctx->U, ctx->V, TK, tstride, nfield, ws, K, wpre ? wpre + s * nt : nullptr));
          a.kzlive = N / 2;
        }
        for (int k = 0; k < K; ++k, ++s) {
          if (KL >= 1 && !last) {
            Tsrc = (KL > 1 ? ctx->TK : ctx->T) + k * tstride;
          } else {
            const WinArgs w = win_args(box_size, (float)R[s], p->filter_type, last ? 0 : 1);
            MRX_TRY(P::inverse_xy(ctx, ctx->U, ctx->V, ctx->T, nfield, w, a.kzlive));
          }
          a.J21c = (1.0 + z) * (1.0 + z) / (4.0 * M_PI) * p->alpha_uv * kPlanck * 1e21 * R[s] * p->unit_length_cm *
                   p->nion_phot_per_bary / kProtonMass * p->unit_mass_g / pow(p->unit_length_cm, 3) / p->unit_time_s;
          a.M_mean = RtoM(p, R[s]);
          a.R_cubed = R[s] * R[s] * R[s];
          a.kfc = (4.0 / 3.0) * M_PI * a.R_cubed / a.pixel_volume / a.M_mean;
          if (rec) {
            // Gamma_R_prefactor (find_HII_bubbles.c:325-334): R h^-1 to R, pixel volume (Mpc/h)^3 to cm^3
            double gp = (1.0 + z) * (1.0 + z) * R[s] * (p->unit_length_cm / p->hubble_h) * kSigmaHI * p->alpha_uv /
                        (p->alpha_uv + 2.75) / 1.0e-12;
            gp *= (p->unit_mass_g / p->unit_time_s) * pow(p->unit_length_cm / p->hubble_h, -3.) *
                  p->nion_phot_per_bary / kProtonMass;
            a.gamma_pref = gp;
          } else {
            a.gamma_pref = 0.0;
          }
          a.step = s;
          a.last = last ? 1 : 0;
          MRX_TRY(P::threshold(ctx, Tsrc, deltax, out[0], out[1], out[2], out[3], out[4], out[5], rec, a));
        }
      }
      if (rec) {
        // N_rec += dNrec after the ladder (:472-477)
        const RRDev* rr;
        MRX_TRY(rr_tables(ctx, p, &rr));
        const float zstep = (float)(p->prev_redshift - p->redshift);
        const float fabs_dtdz =
          (float)fabs(mrx_cosmo_dtdz(p->omega_lambda, p->omega_m, p->hubble_h, (float)p->redshift) / p->hubble_h);
        const long nr = (long)N * N * N;
        {
          PROF("nrec_update");
          k_nrec_update<<<(unsigned)((nr + kNT - 1) / kNT), kNT, 0, ctx->stream>>>(
            const_cast<float*>(in[4]), out[5], out[0], deltax, rr, N, p->redshift, fabs_dtdz, zstep);
        }
        MRX_TRY(check_launch(ctx, "k_nrec_update"));
      }
      const int nb = P::nblocks_threshold(rec);
      {
        PROF("reduce_partials");
        MRX_TRY(reduce_partials(ctx, ctx->stream, ctx->partials, nb, 4, 1, ctx->partials + (size_t)nb * 4 + 8,
                                ctx->sums));
      }
      double h[4];
      HIP_TRY(hipMemcpyAsync(h, ctx->sums, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      const double total_n_cells = pow((double)N, 3);
      if (g) {
        g->volume_weighted_global_xH = h[0] / total_n_cells;
        g->mass_weighted_global_xH = h[1] / h[2];
        g->volume_weighted_global_J_21 = p->uvb_flag ? h[3] / total_n_cells : 0.0;
        g->n_radii = S;
      }
      return MRX_OK;
    }
  };
};

struct FilteredField {
  template <int N>
  struct At {
    static int run(mrx_ctx* ctx, double box_size, int filter_type, const float* in, double R, int apply, int clamp,
                   float* out)
    {
      using P = Pipe<N>;
      MRX_TRY(ensure_buffers(ctx, N, 1));
      MRX_TRY(P::forward(ctx, in, 0, 1, true, ctx->U, ctx->V));
      const WinArgs w = win_args(box_size, (float)R, filter_type, apply);
      int kzlive;
      MRX_TRY(P::inverse_xy(ctx, ctx->U, ctx->V, ctx->T, 1, w, kzlive));
      return P::c2r_z(ctx, ctx->T, out, 1, clamp, N + 2, kzlive);
    }
  };
};

struct R2C {
  template <int N>
  struct At {
    static int run(mrx_ctx* ctx, const float* in, float2* out)
    {
      using P = Pipe<N>;
      MRX_TRY(ensure_buffers(ctx, N, 1));
      MRX_TRY(P::forward(ctx, in, 0, 1, false, ctx->U, ctx->V));
      return P::unpack_herm(ctx, ctx->U, ctx->V, out);
    }
  };
};

struct C2R {
  template <int N>
  struct At {
    static int run(mrx_ctx* ctx, const float2* in, float* out)
    {
      using P = Pipe<N>;
      MRX_TRY(ensure_buffers(ctx, N, 1));
      MRX_TRY(P::pack_herm(ctx, in, ctx->T));
      MRX_TRY(P::inv_xy_plain(ctx, ctx->T, 1));
      return P::c2r_z(ctx, ctx->T, out, 1, -1);
    }
  };
};


// ------------------------------------------------------------ spin temperature
void fill_kappa(KappaSpline& k, const mrx_spline& s)
{
  k.n = s.n;
  for (int i = 0; i < s.n && i < 27; ++i) {  // rounded to float: the rates are evaluated in float
    k.x[i] = (float)s.x[i];
    k.y[i] = (float)s.y[i];
    k.b[i] = (float)s.b[i];
    k.c[i] = (float)s.c[i];
    k.d[i] = (float)s.d[i];
  }
}

// Snapshot constants of the per-cell kernel, each the prefix the reference
// evaluates first (see TsDev).
void fill_ts_dev(TsDev& t, const mrx_ts_data* d, const mrx_ts_snap* s, const mrx_ts_params* p)
{
  memset(&t, 0, sizeof t);
  const int N = p->grid_dim;
  const double box_size = p->box_size / p->hubble_h;
  t.nsteps = s->nsteps;
  t.no_light = s->no_light;
  t.pixel_volume = pow(box_size / (double)N, 3);
  t.MT = (p->unit_mass_g / p->unit_time_s);
  t.L3 = pow(p->unit_length_cm, -3.);
  t.sfr_factor = (1.0 / t.pixel_volume) * t.MT * t.L3 / 1.989e33;
  for (int c = 0; c < s->nsteps; ++c) {
    t.dt_dzpp[c] = s->dt_dzpp[c];
    t.dzpp[c] = s->dzpp[c];
    t.zpp1[c] = 1 + s->zpp[c];
    t.zpp_pow[c] = s->zpp_pow[c];
    t.sum_lyn[c] = s->sum_lyn[c];
    for (int x = 0; x < kNXHII; ++x) {
      t.fheat[x][c] = s->freq_heat[x][c];
      t.fion[x][c] = s->freq_ion[x][c];
      t.flya[x][c] = s->freq_lya[x][c];
    }
    for (int x = 0; x + 1 < kNXHII; ++x) {
      const double dxh = (double)(d->xi_XHII[x + 1] - d->xi_XHII[x]);  // float difference, as the reference
      t.sheat[x][c] = (s->freq_heat[x + 1][c] - s->freq_heat[x][c]) / dxh;
      t.sion[x][c] = (s->freq_ion[x + 1][c] - s->freq_ion[x][c]) / dxh;
      t.slya[x][c] = (s->freq_lya[x + 1][c] - s->freq_lya[x][c]) / dxh;
    }
  }
  for (int x = 0; x < kNXHII; ++x) t.XH[x] = d->xi_XHII[x];
  const float zp = (float)s->zp;
  // the folded shell coefficients of k_ts_cell (see TsDev::cw)
  const double P2 = pow(1 + zp, 2);
  for (int c = 0; c < s->nsteps; ++c) {
    const double w = t.sfr_factor * t.dt_dzpp[c] * t.dzpp[c] * t.zpp_pow[c];
    for (int x = 0; x < kNXHII; ++x) {
      t.cw[x][c][0] = w * t.fheat[x][c];
      t.cw[x][c][1] = w * t.fion[x][c];
      t.cw[x][c][2] = w * t.flya[x][c];
      t.cw[x][c][3] = w * t.sheat[x][c];
      t.cw[x][c][4] = w * t.sion[x][c];
      t.cw[x][c][5] = w * t.slya[x][c];
    }
    t.cst[c] = t.sfr_factor * P2 * t.zpp1[c] * t.sum_lyn[c] * t.dt_dzpp[c] * t.dzpp[c];
  }
  for (int x = 0; x < kNXHII; ++x)
    for (int c = 0; c < s->nsteps; ++c)
      for (int k = 0; k < 6; ++k) t.cwp[((size_t)x * s->nsteps + c) * 6 + k] = t.cw[x][c][k];
  for (int c = 0; c < s->nsteps; ++c) t.cwp[(size_t)6 * kNXHII * s->nsteps + c] = t.cst[c];
  t.zp_f = zp;
  t.dzp = s->dzp;
  t.dt_dzp = s->dt_dzp;
  t.const_zp_prefactor = s->const_zp_prefactor;
  t.growth = s->growth_factor_zp;
  t.dgrowth = s->dgrowth_factor_dzp;
  t.NBP3 = d->N_b0 * pow(1 + zp, 3);
  t.P2 = pow(1 + zp, 2);
  t.dadia0 = 3 / (1.0 + zp);
  t.Trad = 2.728 * (1.0 + zp);
  t.T21_Trad = 0.0628 / t.Trad;
  t.P3 = pow(1.0 + zp, 3.0);
  const double H = mrx_ts_hubble(d, zp);
  t.TGP = 1.342881e-7 / H * d->No * pow(1 + (double)zp, 3);
  t.XA = 1.66e11 / (1.0 + zp);
  t.HH = H / (3.2407789e-18 * p->hubble_h);
  t.T4 = pow(t.Trad, 4.0);
  t.one_p_z = 1.0 + zp;
  t.invTrad = 1.0 / t.Trad;
  t.comp = t.T4 / (t.HH * p->hubble_h * t.one_p_z);
  t.Tk_floor = (float)(2.728 * (1 + s->zp));
  t.No = d->No;
  t.N_b0 = d->N_b0;
  t.f_H = d->f_H;
  t.f_He = d->f_He;
  t.hubble_h = p->hubble_h;
  fill_kappa(t.kHH, d->kappa_HH);
  fill_kappa(t.keH, d->kappa_eH);
  fill_kappa(t.kpH, d->kappa_pH);
}

std::string ts_key(const mrx_ts_params* p)
{
  char b[1400];
  snprintf(b, sizeof b, "%s|%.17g|%.17g|%.17g|%.17g|%.17g|%.17g|%.17g|%.17g|%.17g", p->tables_dir, p->hubble_h,
           p->omega_m, p->omega_lambda, p->omega_r, p->omega_k, p->w_lambda, p->baryon_frac, p->y_he,
           p->nion_phot_per_bary);
  return std::string(b) + "|" + std::to_string(p->reion_efficiency);
}

double ts_RtoM(const mrx_ts_params* p, double R)
{
  if (p->rtom_filter_type == 0) return (4.0 / 3.0) * M_PI * pow(R, 3) * (p->omega_m * p->rho_crit);
  return pow(2 * M_PI, 1.5) * p->omega_m * p->rho_crit * pow(R, 3);
}

struct TsRun {
  template <int N>
  struct At {
    // grids are device pointers here: deltax, stars, sfr, hist (slots), xe_prev, Tk, TS, xe
    // prepare_only: the shells and the snapshot constants (ctx->ts_dev) only,
    // for the adjoint, which then runs its own per-cell kernel
    static int run(mrx_ctx* ctx, const mrx_ts_params* p, const float* deltax, const float* stars, const float* sfr,
                   const float* hist, float* xe_prev, float* Tk, float* TS, mrx_ts_globals* g,
                   bool prepare_only = false)
    {
      using P = Pipe<N>;
      struct TsTag {  // kernel timings of this step are named "ts_..."
        mrx_ctx* c;
        explicit TsTag(mrx_ctx* x) : c(x) { c->prof_ts = true; }
        ~TsTag() { c->prof_ts = false; }
      } tag(ctx);
      mrx_ts_data* d = ctx->tsd;
      mrx_ts_snap* s = ctx->tss;
      const long ncell = (long)N * N * N;
      const size_t np = npad(N);
      const double h = p->hubble_h;
      const double box_size = p->box_size / h;
      const double pixel_volume = pow(box_size / (double)N, 3);
      const int nbc = (int)((ncell + kNT - 1) / kNT);
      // workgroups of the cell kernels, whose 7 partial sums are reduced at the end
      // (one cell per thread for k_ts_cell; the homogeneous kernel is grid-stride)
      int nred = nbc;
      const int nred_h = (int)std::min<long>(nbc, resident_blocks(ctx, k_ts_homogeneous, 0));
      MRX_TRY(ensure_partials(ctx, (size_t)nbc * 7 + 2 * 1024 + 16 + kRedBlocks * kRedMaxV));
      if (!ctx->sums) HIP_TRY(hipMalloc(&ctx->sums, sizeof(double) * 8));
      if (!ctx->ts_dev) HIP_TRY(hipMalloc(&ctx->ts_dev, sizeof(TsDev)));
      double collapse_fraction = 0.0, x_e_ave = 0.0;
      auto fcoll = [&](double R, int with_xe) -> int {
        FcollArgs fa;
        fa.M_h2 = ts_RtoM(p, R) * h * h;
        fa.R3 = pow(R, 3.0);
        fa.R3_over_V = 0.0;
        fa.pixel_volume = pixel_volume;
        fa.N = N;
        fa.with_xe = with_xe;
        {
          PROF("fcoll");
          k_ts_fcoll<<<1024, kNT, 0, ctx->stream>>>(stars, deltax, xe_prev, ctx->partials + (size_t)nbc * 7, fa);
        }
        MRX_TRY(check_launch(ctx, "k_ts_fcoll"));
        k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(ctx->partials + (size_t)nbc * 7, 1024, 2, ctx->sums);
        MRX_TRY(check_launch(ctx, "k_reduce_partials"));
        double hsum[2];
        HIP_TRY(hipMemcpyAsync(hsum, ctx->sums, sizeof hsum, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        collapse_fraction = hsum[0] / pow((double)N, 3);
        x_e_ave = hsum[1] / pow((double)N, 3);
        return MRX_OK;
      };
      if (s->homogeneous) {
        const float zp = (float)s->zp;
        const float xe0 = (float)mrx_ts_xion_RECFAST(d, zp);
        const float T0 = (float)mrx_ts_T_RECFAST(d, zp);
        fill_ts_dev(*ctx->ts_host, d, s, p);
        HIP_TRY(hipMemcpyAsync(ctx->ts_dev, ctx->ts_host, sizeof(TsDev), hipMemcpyHostToDevice, ctx->stream));
        ctx->ts_h_xe0 = xe0;
        ctx->ts_h_T0 = T0;
        if (prepare_only) return MRX_OK;
        {
          PROF("homogeneous");
          k_ts_homogeneous<<<nred_h, kNT, 0, ctx->stream>>>(ctx->ts_dev, deltax, xe_prev, Tk, TS, N, xe0, T0,
                                                           ctx->partials);
        }
        nred = nred_h;
        MRX_TRY(check_launch(ctx, "k_ts_homogeneous"));
        ctx->ts_shells_N = ctx->ts_shells_n = 0;  // no shells on the homogeneous branch
        MRX_TRY(fcoll(kLFactor * box_size / (float)N, 0));
        x_e_ave = 0.0;
      } else {
        MRX_TRY(ensure_buffers(ctx, N, 1));
        MRX_TRY(ensure_buf(ctx, ctx->ts_work, ctx->ts_work_cap, np));
        ctx->ts_shells_N = ctx->ts_shells_n = 0;
        MRX_TRY(ensure_buf(ctx, ctx->ts_shells, ctx->ts_shells_cap, (size_t)s->nsteps * ncell));
        // f_coll and <x_e> of the first shell (ComputeTs.c:413-475) read the
        // input grids only: taken first, so the host builds the tables while
        // the device filters the shells
        if (!p->fixed_globals) MRX_TRY(fcoll(s->R[0], 1));
        if (p->fixed_globals) {
          collapse_fraction = p->fixed_collapse_fraction;
          x_e_ave = p->fixed_x_e_ave;
        }
        s->collapse_fraction = collapse_fraction;
        s->x_e_ave = x_e_ave;
        // the host tables are built on a thread of their own while this one
        // queues the shells' kernels (the build reads the plan's radii and
        // edges and writes only the tables; the shell loop reads the plan)
        int tab_rc = MRX_OK;
        std::thread tab;
        try {
          tab = std::thread([&tab_rc, s, d, p] { tab_rc = mrx_ts_build_tables(s, d, p); });
        } catch (...) {  // no thread to be had: build them here, before the shells
          tab_rc = mrx_ts_build_tables(s, d, p);
        }
        struct Join {
          std::thread& t;
          ~Join()
          {
            if (t.joinable()) t.join();
          }
        } join_tab{tab};
        const long half = (long)N * N * (N / 2 + 1);
        // the working grid starts as the input sfr grid: it is read in place
        // until the first shell with operations writes ts_work from it
        const float* wgrid = sfr;
        for (int ct = 0; ct < s->nsteps; ++ct) {
          // a shell's operations in plan order, MRX_TS_MAX_OPS per launch (each
          // launch after the first reads the working grid the previous wrote)
          for (int k0 = 0; k0 < s->n_ops[ct]; k0 += MRX_TS_MAX_OPS) {
            SfrOps o;
            o.n = std::min(MRX_TS_MAX_OPS, s->n_ops[ct] - k0);
            for (int k = 0; k < o.n; ++k) {
              const mrx_sfr_op& op = s->ops[ct][k0 + k];
              o.kind[k] = op.kind;
              o.slot[k] = op.slot;
              o.wd[k] = op.weight;
              o.wf[k] = (float)op.weight;
            }
            PROF("sfr_op");
            k_sfr_ops<<<(unsigned)((np / 2 + kNT - 1) / kNT), kNT, 0, ctx->stream>>>(ctx->ts_work, wgrid, hist, (long)np,
                                                                              o, N, half);
            MRX_TRY(check_launch(ctx, "k_sfr_ops"));
            wgrid = ctx->ts_work;
          }
          // the working grid only changes when operations were applied: the
          // forward transform of an unchanged grid is reused.  With sharp-k
          // shells, the shells that reuse it have radii >= this one's, so
          // their balls of modes lie inside this shell's: the forward passes
          // need only its tiles (ct = 0 is unfiltered and takes them all).
          const WinArgs w = win_args(p->box_size, (float)s->R[ct], p->heating_filter_type, ct > 0 ? 1 : 0);
          if (ct == 0 || s->n_ops[ct] > 0) {
            const int nmax = (w.apply && w.type == MRX_FILTER_SHARPK) ? sharpk_nmax(w, N) : N / 2;
            if (nmax < N / 2)
              MRX_TRY(P::forward_sharpk(ctx, wgrid, ctx->U, ctx->V, nmax));
            else
              MRX_TRY(P::forward(ctx, wgrid, 0, 1, true, ctx->U, ctx->V, true));
          }
          int kzlive;
          MRX_TRY(P::inverse_xy(ctx, ctx->U, ctx->V, ctx->T, 1, w, kzlive));
          MRX_TRY(P::c2r_z(ctx, ctx->T, ctx->ts_shells + (size_t)ct * ncell, 1, 3, N, kzlive));
        }
        ctx->ts_shells_N = N;
        ctx->ts_shells_n = s->nsteps;
        if (tab.joinable()) tab.join();
        if (tab_rc != MRX_OK) return set_err(ctx, tab_rc, "spin-temperature tables failed");
        fill_ts_dev(*ctx->ts_host, d, s, p);
        HIP_TRY(hipMemcpyAsync(ctx->ts_dev, ctx->ts_host, sizeof(TsDev), hipMemcpyHostToDevice, ctx->stream));
        if (prepare_only) return MRX_OK;
        {
          PROF("cell");
          k_ts_cell<<<nred, kNT, ts_cell_lds_bytes(s->nsteps), ctx->stream>>>(ctx->ts_dev, ctx->ts_shells, deltax,
                                                                               xe_prev, Tk, TS, N, ctx->partials);
        }
        MRX_TRY(check_launch(ctx, "k_ts_cell"));
      }
      MRX_TRY(reduce_partials(ctx, ctx->stream, ctx->partials, nred, 7, 1,
                              ctx->partials + (size_t)nbc * 7 + 2 * 1024 + 16, ctx->sums));
      double hs[7];
      HIP_TRY(hipMemcpyAsync(hs, ctx->sums, sizeof hs, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      const double n3 = pow((double)N, 3);
      if (g) {
        g->volume_ave_J_alpha = hs[0] / n3;
        g->volume_ave_xalpha = hs[1] / n3;
        g->volume_ave_Xheat = hs[2] / n3;
        g->volume_ave_Xion = hs[3] / n3;
        g->volume_ave_TS = hs[4] / n3;
        g->volume_ave_TK = hs[5] / n3;
        g->volume_ave_xe = hs[6] / n3;
        if (p->reproduce_quirks) {
          // ComputeTs.c:957 reads x_e_box_prev at the stale index of the previous loop: the last cell
          float last = 0.f;
          const size_t ip_last = ((size_t)N * N - 1) * (N + 2) + (N - 1);
          HIP_TRY(hipMemcpy(&last, xe_prev + ip_last, sizeof(float), hipMemcpyDeviceToHost));
          g->volume_ave_xe = (double)last * n3 / n3;
        }
        g->collapse_fraction = collapse_fraction;
        g->x_e_ave = x_e_ave;
        g->no_light = s->homogeneous ? 0 : s->no_light;
        g->homogeneous = s->homogeneous;
      }
      return MRX_OK;
    }
  };
};

template <class K>
struct Disp {
  template <int N>
  struct F {
    template <class... A>
    static int run(A&&... args)
    {
      return K::template At<N>::run(args...);
    }
  };
};

template <class K, class... A>
int dispatch(int N, A&&... args)
{
  return dispatch_n<Disp<K>::template F>(N, args...);
}

// Brightness-temperature constants exactly as mrx_brightness_temp forms them
// (BrightnessTemperature.c:46-52, :80).
TbArgs tb_args(const mrx_tb_params* p)
{
  TbArgs a;
  const float redshift = (float)(float)p->redshift;
  const float Hubble_h = (float)(float)p->hubble_h;
  const float OmegaM = (float)(float)p->omega_m;
  const float OmegaB = (float)(OmegaM * p->baryon_frac);
  a.const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                           sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  a.T_rad = kTcmb * (1. + redshift);
  a.N = p->grid_dim;
  a.use_ts = p->include_spin_temp;
  return a;
}

// Bin ceilings of Compute_PS (ComputePowerSpectrum.c:18-41, :160-190), in float.
int ps_setup(int N, double box_size_mpc, float k_factor, PsArgs& a)
{
  memset(&a, 0, sizeof a);
  if (!(k_factor > 1.0f)) k_factor = 1.35f;
  const float delta_k = (float)(2. * M_PI / box_size_mpc);
  const float k_max = delta_k * (float)N;
  float k_ceil = delta_k;
  int nb = 0;
  while (k_ceil < k_max) {
    if (nb >= kMaxPsBins) return -1;
    a.ceil[nb++] = k_ceil;
    k_ceil *= k_factor;
  }
  a.N = N;
  a.nbins = nb;
  a.delta_k = delta_k;
  a.volume = (double)powf((float)box_size_mpc, 3);
  return nb;
}

__global__ void k_sum_f32(const float* __restrict__ x, long n, double* __restrict__ partials)
{
  __shared__ double red[kNT];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) acc += x[i];
  acc = block_sum<kNT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

struct TbVel {
  template <int N>
  struct At {
    // delta_T with peculiar velocities; device grids; vel converted in place
    static int run(mrx_ctx* ctx, const mrx_tb_vel_params* p, const float* xH, const float* deltax, const float* TS,
                   float* vel, float* dT, double* ave)
    {
      using P = Pipe<N>;
      constexpr int NH = N / 2;
      const long nr = (long)N * N * N;
      const size_t np = npad(N);
      MRX_TRY(ensure_buffers(ctx, N, 1));
      MRX_TRY(ensure_buf(ctx, ctx->tbv, ctx->tbv_cap, np + (p->pec_vels == 3 ? (size_t)nr : 0)));
      float* grad = ctx->tbv;
      const TbArgs base = tb_args(&p->tb);
      TbVelArgs a;
      a.const_factor = base.const_factor;
      a.redshift = (float)(float)p->tb.redshift;
      a.N = N;
      a.pec_vels = p->pec_vels;
      a.trees_id = p->trees_id;
      a.T_rad = base.T_rad;
      // hubble(z), XRayHeatingFunctions.c:394-404 (z a float there)
      const float z = a.redshift;
      a.H_z = (float)(float)(kHubble * p->tb.hubble_h *
                             sqrt(p->tb.omega_m * std::pow((double)(1 + z), 3.0) +
                                  p->omega_r * std::pow((double)(1 + z), 4.0) + p->omega_lambda));
      a.max_v_deriv = fabs(0.2f * a.H_z);
      const unsigned nb = (unsigned)((nr + kNT - 1) / kNT);
      {
        PROF("vel_convert");
        k_vel_convert<<<nb, kNT, 0, ctx->stream>>>(vel, N, a.redshift, a.trees_id);
      }
      MRX_TRY(check_launch(ctx, "k_vel_convert"));
      MRX_TRY(P::forward(ctx, vel, 0, 1, true, ctx->U, ctx->V));
      const float box_f = (float)p->box_size;
      const float delta_k = (float)(2.0 * M_PI / box_f);
      {
        PROF("vel_grad_op");
        k_vel_grad_op<<<(unsigned)((P::nc + kNT - 1) / kNT), kNT, 0, ctx->stream>>>(ctx->U, ctx->T, N,
                                                                                  p->tb.hubble_h, delta_k);
      }
      MRX_TRY(check_launch(ctx, "k_vel_grad_op"));
      MRX_TRY(P::inv_xy_plain(ctx, ctx->T, 1));
      MRX_TRY(P::c2r_z(ctx, ctx->T, grad, 1, -1, N + 2, NH));
      {
        PROF("brightness_temp_vel");
        k_brightness_temp_vel<<<nb, kNT, 0, ctx->stream>>>(xH, deltax, TS, grad, dT, a);
      }
      MRX_TRY(check_launch(ctx, "k_brightness_temp_vel"));
      if (p->pec_vels == 3) {
        RsdArgs r;
        r.N = N;
        r.H_z = a.H_z;
        const double box_mpc = p->box_size / p->tb.hubble_h;  // BrightnessTemperature.c:44
        const double sub_w = (box_mpc / ((double)N)) / (double)kRsdSteps;
        r.cell = box_mpc / (float)N;
        r.half = (sub_w / 2.) / r.cell;
        r.frac_w = sub_w / r.cell;
        for (int s = 0; s < kRsdSteps; ++s) {
          r.x_off[s] = (float)(sub_w * (float)s + sub_w / 2.);
          r.x_pos[s] = (float)(r.x_off[s] / (box_mpc / (float)N));
        }
        PROF("rsd_los");
        k_rsd_los<<<(unsigned)(((long)N * N + kNT - 1) / kNT), kNT, 0, ctx->stream>>>(dT, xH, vel, grad + np, r);
        MRX_TRY(check_launch(ctx, "k_rsd_los"));
      }
      MRX_TRY(ensure_partials(ctx, 1024 + 8));
      if (!ctx->sums) HIP_TRY(hipMalloc(&ctx->sums, sizeof(double) * 8));
      k_sum_f32<<<1024, kNT, 0, ctx->stream>>>(dT, nr, ctx->partials);
      MRX_TRY(check_launch(ctx, "k_sum_f32"));
      k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(ctx->partials, 1024, 1, ctx->sums);
      MRX_TRY(check_launch(ctx, "k_reduce_partials"));
      double h;
      HIP_TRY(hipMemcpyAsync(&h, ctx->sums, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      if (ave) *ave = h / pow((double)N, 3);
      return MRX_OK;
    }
  };
};

struct PowerSpec {
  template <int N>
  struct At {
    // forward (grad_power == NULL): bins into k/power/error/counts; adjoint:
    // dL/d(delta_T) into grad_out from dL/dP_b
    static int run(mrx_ctx* ctx, double box_size_mpc, float k_factor, const float* dT, int max_bins, double* k_out,
                   double* p_out, double* e_out, long* c_out, const double* grad_power, float* grad_out)
    {
      using P = Pipe<N>;
      PsArgs a;
      const int nb = ps_setup(N, box_size_mpc, k_factor, a);
      if (nb < 1) return set_err(ctx, MRX_EINVAL, "power spectrum: bad binning (%d bins)", nb);
      const long ncell = (long)N * N * N;
      const size_t np = npad(N);
      const long nmodes = (long)N * N * (N / 2 + 1);
      const int nblk = (int)((nmodes + (long)kNT * kPsJ - 1) / ((long)kNT * kPsJ));
      MRX_TRY(ensure_buffers(ctx, N, 1));
      MRX_TRY(ensure_buf(ctx, ctx->adj_work, ctx->adj_work_cap, 2 * np));
      // partials: block sums, then the bin totals, then dL/dP_b / n_b
      const size_t pblk = std::max((size_t)nblk * nb * 3, (size_t)1024);
      MRX_TRY(ensure_partials(ctx, pblk + 4 * kMaxPsBins));
      if (!ctx->sums) HIP_TRY(hipMalloc(&ctx->sums, sizeof(double) * 8));
      double* sums_big = ctx->partials + pblk;
      double* dlam = sums_big + 3 * kMaxPsBins;
      // <delta_T>
      k_sum_f32<<<1024, kNT, 0, ctx->stream>>>(dT, ncell, ctx->partials);
      MRX_TRY(check_launch(ctx, "k_sum_f32"));
      k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(ctx->partials, 1024, 1, ctx->sums);
      MRX_TRY(check_launch(ctx, "k_reduce_partials"));
      double tot = 0.0;
      HIP_TRY(hipMemcpyAsync(&tot, ctx->sums, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      const double ave = tot / pow((double)N, 3);
      float* work = ctx->adj_work;
      k_ps_prep<<<(unsigned)((ncell + 255) / 256), 256, 0, ctx->stream>>>(dT, work, N, ave, (float)a.volume);
      MRX_TRY(check_launch(ctx, "k_ps_prep"));
      MRX_TRY(P::forward(ctx, work, 0, 1, false, ctx->U, ctx->V));
      k_ps_bin<<<nblk, kNT, 0, ctx->stream>>>(ctx->U, ctx->V, ctx->partials, a);
      MRX_TRY(check_launch(ctx, "k_ps_bin"));
      k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(ctx->partials, nblk, nb * 3, sums_big);
      MRX_TRY(check_launch(ctx, "k_reduce_partials"));
      std::vector<double> h(3 * (size_t)nb);
      HIP_TRY(hipMemcpyAsync(h.data(), sums_big, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      if (!grad_power) {
        for (int b = 0; b < nb && b < max_bins; ++b) {
          const double cnt = h[3 * b + 2];
          if (k_out) k_out[b] = h[3 * b + 1] / cnt;
          if (p_out) p_out[b] = h[3 * b] / cnt;
          if (e_out) e_out[b] = (h[3 * b] / cnt) / sqrt(cnt);
          if (c_out) c_out[b] = (long)cnt;
        }
        return nb;
      }
      std::vector<double> lam(nb);
      for (int b = 0; b < nb; ++b) lam[b] = (h[3 * b + 2] > 0) ? grad_power[b] / h[3 * b + 2] : 0.0;
      HIP_TRY(hipMemcpyAsync(dlam, lam.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
      k_ps_adjoint_scale<<<(unsigned)((nmodes + 255) / 256), 256, 0, ctx->stream>>>(
        ctx->U, ctx->V, dlam, a, a.volume / pow((double)N, 3));
      MRX_TRY(check_launch(ctx, "k_ps_adjoint_scale"));
      const WinArgs w = win_args(box_size_mpc, 1.0f, 0, 0);
      int kzlive;
      MRX_TRY(P::inverse_xy(ctx, ctx->U, ctx->V, ctx->T, 1, w, kzlive));
      MRX_TRY(P::c2r_z(ctx, ctx->T, grad_out, 1, -1, N, kzlive));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      return nb;
    }
  };
};

struct SnapAdj {
  template <int N>
  struct At {
    // The gradient through the shell tables' two grid means, the collapse
    // fraction and <x_e> (ComputeTs.c:413-475), which the tables take through
    // tau_X: the derivatives of the six x_e tables with respect to each mean
    // by central differences of the host table build (relative step 1e-4),
    // contracted on the device with every cell's shell values and shell-sum
    // adjoints (k_ts_means_grad), then spread back over the cells the means
    // are taken from (k_ts_means_apply).  <x_e> only when its gradient
    // (g_xe_prev) is wanted.
    static int means_tables(mrx_ctx* ctx, const mrx_ts_params* tp, const mrx_ts_snap* s, int nm,
                            std::vector<double>& E)
    {
      const int ns = s->nsteps;
      E.assign((size_t)nm * kNXHII * ns * 6, 0.0);
      std::unique_ptr<mrx_ts_snap> sc(new mrx_ts_snap);
      std::unique_ptr<TsDev> tplus(new TsDev), tminus(new TsDev);
      const TsDev& t0 = *ctx->ts_host;  // the forward's constants
      // the table build records the collapse fraction in the caller's history
      double* fc_slot = tp->stored_fcoll ? tp->stored_fcoll + tp->snapshot : nullptr;
      const double fc_saved = fc_slot ? *fc_slot : 0.0;
      int rc = MRX_OK;
      for (int m = 0; m < nm && rc == MRX_OK; ++m) {
        const double x0 = (m == 0) ? s->collapse_fraction : s->x_e_ave;
        const double h = 1e-4 * std::max(std::fabs(x0), 1e-8);
        for (int sg = 0; sg < 2 && rc == MRX_OK; ++sg) {
          *sc = *s;
          (m == 0 ? sc->collapse_fraction : sc->x_e_ave) = x0 + (sg ? -h : h);
          rc = mrx_ts_build_tables(sc.get(), ctx->tsd, tp);
          if (rc == MRX_OK) fill_ts_dev(sg ? *tminus : *tplus, ctx->tsd, sc.get(), tp);
        }
        if (rc != MRX_OK) break;
        for (int lo = 0; lo < kNXHII; ++lo)
          for (int ct = 0; ct < ns; ++ct) {
            const double w = t0.sfr_factor * t0.dt_dzpp[ct] * t0.dzpp[ct] * t0.zpp_pow[ct];
            const double d[6] = {tplus->fheat[lo][ct] - tminus->fheat[lo][ct], tplus->fion[lo][ct] - tminus->fion[lo][ct],
                                 tplus->flya[lo][ct] - tminus->flya[lo][ct], tplus->sheat[lo][ct] - tminus->sheat[lo][ct],
                                 tplus->sion[lo][ct] - tminus->sion[lo][ct], tplus->slya[lo][ct] - tminus->slya[lo][ct]};
            for (int j = 0; j < 6; ++j) E[(((size_t)m * kNXHII + lo) * ns + ct) * 6 + j] = w * d[j] / (2.0 * h);
          }
      }
      if (fc_slot) *fc_slot = fc_saved;
      return rc;
    }
    static int means_term(mrx_ctx* ctx, const mrx_ts_params* tp, const mrx_ts_snap* s, int nm,
                          const std::vector<double>& E, const float* deltax, const float* stars, const float* xe_prev,
                          float* g_delta, float* g_stars, float* g_xe_prev)
    {
      const long ncell = (long)N * N * N;
      MRX_TRY(ensure_dev(ctx, ctx->adj_E, ctx->adj_E_cap, E.size()));
      HIP_TRY(hipMemcpyAsync(ctx->adj_E, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice, ctx->stream));
      const unsigned nbc = (unsigned)((ncell + kNT - 1) / kNT);
      {
        PROF("adjoint_means");
        k_ts_means_grad<<<nbc, kNT, 0, ctx->stream>>>(ctx->ts_dev, ctx->ts_shells, xe_prev, ctx->adj_gS, ctx->adj_E, nm,
                                                      N, ctx->partials);
      }
      MRX_TRY(check_launch(ctx, "k_ts_means_grad"));
      MRX_TRY(reduce_partials(ctx, ctx->stream, ctx->partials, (int)nbc, 2, 1, ctx->partials + (size_t)nbc * 2 + 16,
                              ctx->sums));
      double g[2];
      HIP_TRY(hipMemcpyAsync(g, ctx->sums, sizeof g, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      // both means are over the N^3 cells
      const double n3 = std::pow((double)N, 3);
      const double h = tp->hubble_h;
      const double R0 = s->R[0];
      FcollArgs fa;
      fa.M_h2 = ts_RtoM(tp, R0) * h * h;
      fa.R3 = std::pow(R0, 3.0);
      fa.R3_over_V = 0.0;
      fa.pixel_volume = std::pow(tp->box_size / h / (double)N, 3);
      fa.N = N;
      fa.with_xe = 1;
      k_ts_means_apply<<<nbc, kNT, 0, ctx->stream>>>(deltax, stars, g_delta, g_stars, nm > 1 ? g_xe_prev : nullptr,
                                                     nm > 1 ? g[1] / n3 : 0.0, g[0] / n3, fa);
      return check_launch(ctx, "k_ts_means_apply");
    }

    static int run(mrx_ctx* ctx, const mrx_ts_params* tp, const mrx_hii_params* hp, const mrx_tb_params* bp,
                   const float* deltax, const float* stars, const float* sfr, const float* hist,
                   const float* xe_prev, const float* Tk, const float* xH, const float* TS, const float* xe_box,
                   const float* gdT, float* g_delta, float* g_stars, float* g_sfr, float* g_hist, float* g_xe_prev,
                   float* g_Tk)
    {
      using P = Pipe<N>;
      const long ncell = (long)N * N * N;
      const size_t np = npad(N);
      AdjCellArgs a;
      memset(&a, 0, sizeof a);
      a.N = N;
      const TbArgs tb = tb_args(bp);
      a.tb_const = tb.const_factor;
      a.T_rad = tb.T_rad;
      a.use_ts = tb.use_ts;
      std::vector<double> R;
      const int S = ladder(hp, R);
      if (S < 1) return set_err(ctx, MRX_EINVAL, "radius ladder is empty");
      const double Rl = R[S - 1];
      const double pixel_volume = pow(hp->box_size / (double)N, 3);
      a.eff = hp->efficiency;
      a.fc_unit = (4.0 / 3.0) * M_PI * (Rl * Rl * Rl) / pixel_volume / RtoM(hp, Rl);
      a.use_xe = hp->include_spin_temp;
      const mrx_ts_snap* s = nullptr;
      if (tp) {
        MRX_TRY(TsRun::At<N>::run(ctx, tp, deltax, stars, sfr, hist, const_cast<float*>(xe_prev), nullptr, nullptr,
                                  nullptr, true));
        s = ctx->tss;
        a.ts_mode = s->homogeneous ? 1 : 2;
        a.h_xe0 = ctx->ts_h_xe0;
        a.h_T0 = ctx->ts_h_T0;
      }
      const bool shells = (a.ts_mode == 2) && !s->no_light;
      if (a.ts_mode == 2) MRX_TRY(ensure_dev(ctx, ctx->adj_gS, ctx->adj_gS_cap, 4 * (size_t)ncell));
      // the tables' derivatives with respect to their grid means are host
      // work: built on a thread of their own while the device runs the
      // adjoint, and used after the shells' adjoints
      const bool means = shells && !tp->fixed_globals;
      const int nm = g_xe_prev ? 2 : 1;
      std::vector<double> E;
      int E_rc = MRX_OK;
      std::thread E_thread;
      if (means) {
        try {
          E_thread = std::thread([&E_rc, &E, ctx, tp, s, nm] { E_rc = means_tables(ctx, tp, s, nm, E); });
        } catch (...) {
          E_rc = means_tables(ctx, tp, s, nm, E);
        }
      }
      struct JoinE {
        std::thread& t;
        ~JoinE()
        {
          if (t.joinable()) t.join();
        }
      } join_E{E_thread};
      auto finish_means = [&]() -> int {
        if (!means) return MRX_OK;
        if (E_thread.joinable()) E_thread.join();
        if (E_rc != MRX_OK) return set_err(ctx, E_rc, "spin-temperature tables (grid-mean derivatives) failed");
        return means_term(ctx, tp, s, nm, E, deltax, stars, xe_prev, g_delta, g_stars, g_xe_prev);
      };
      const unsigned nbc = (unsigned)((ncell + kNT - 1) / kNT);
      {
        PROF("adjoint_cell");
        k_snapshot_adjoint_cell<<<nbc, kNT, a.ts_mode == 2 ? ts_cell_lds_bytes(s->nsteps) : 0, ctx->stream>>>(
          ctx->ts_dev, ctx->ts_shells, deltax, stars, xe_prev, Tk, xH, TS, xe_box, gdT, g_delta, g_stars,
          a.ts_mode == 2 ? g_xe_prev : nullptr, a.ts_mode == 2 ? g_Tk : nullptr, ctx->adj_gS, a);
      }
      MRX_TRY(check_launch(ctx, "k_snapshot_adjoint_cell"));
      if (a.ts_mode != 2) {
        if (g_xe_prev) HIP_TRY(hipMemsetAsync(g_xe_prev, 0, sizeof(float) * np, ctx->stream));
        if (g_Tk) HIP_TRY(hipMemsetAsync(g_Tk, 0, sizeof(float) * ncell, ctx->stream));
      }
      const int nslot = (tp && s && !s->homogeneous) ? tp->n_history : 0;
      if (g_sfr) HIP_TRY(hipMemsetAsync(g_sfr, 0, sizeof(float) * np, ctx->stream));
      if (g_hist && tp) HIP_TRY(hipMemsetAsync(g_hist, 0, sizeof(float) * np * (size_t)tp->n_history, ctx->stream));
      if (!shells || !g_sfr) return finish_means();
      MRX_TRY(ensure_buf(ctx, ctx->adj_work, ctx->adj_work_cap, 2 * np));
      float* gshell = ctx->adj_work;
      float* tmp = ctx->adj_work + np;
      const long half = (long)N * N * (N / 2 + 1);
      for (int ct = s->nsteps - 1; ct >= 0; --ct) {
        const float* unclamped = nullptr;
        if (ct == 0) {  // shell 0 is the working grid after its operations: rebuild it
          HIP_TRY(hipMemcpyAsync(tmp, sfr, sizeof(float) * np, hipMemcpyDeviceToDevice, ctx->stream));
          for (int k = 0; k < s->n_ops[0]; ++k) {
            const mrx_sfr_op& o = s->ops[0][k];
            const long n = (o.kind <= 1) ? half : ncell;
            const float* hs = (o.kind >= 2) ? hist + (size_t)o.slot * np : nullptr;
            k_sfr_op<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(tmp, hs, o.kind, o.weight,
                                                                          (float)o.weight, N, half);
            MRX_TRY(check_launch(ctx, "k_sfr_op"));
          }
          unclamped = tmp;
        }
        {
          PROF("adjoint_shell_grad");
          k_ts_shell_grad<<<nbc, kNT, 0, ctx->stream>>>(ctx->ts_dev, ctx->ts_shells, xe_prev, ctx->adj_gS, gshell, ct,
                                                        N, unclamped);
        }
        MRX_TRY(check_launch(ctx, "k_ts_shell_grad"));
        const float* add = gshell;
        if (ct > 0) {  // shell 0 is the unfiltered grid
          MRX_TRY(P::forward(ctx, gshell, 0, 1, true, ctx->U, ctx->V, true));
          const WinArgs w = win_args(tp->box_size, (float)s->R[ct], tp->heating_filter_type, 1);
          int kzlive;
          MRX_TRY(P::inverse_xy(ctx, ctx->U, ctx->V, ctx->T, 1, w, kzlive));
          MRX_TRY(P::c2r_z(ctx, ctx->T, tmp, 1, -1, N + 2, kzlive));
          add = tmp;
        }
        k_axpy<<<(unsigned)((np + 255) / 256), 256, 0, ctx->stream>>>(g_sfr, add, (long)np);
        MRX_TRY(check_launch(ctx, "k_axpy"));
        for (int k = s->n_ops[ct] - 1; k >= 0; --k) {
          const mrx_sfr_op& o = s->ops[ct][k];
          const long n = (o.kind <= 1) ? half : ncell;
          float* gh = (o.kind >= 2 && g_hist && o.slot < nslot) ? g_hist + (size_t)o.slot * np : nullptr;
          k_sfr_op_adjoint<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(g_sfr, gh, o.kind, o.weight,
                                                                                 (float)o.weight, N, half);
          MRX_TRY(check_launch(ctx, "k_sfr_op_adjoint"));
        }
      }
      return finish_means();
    }
  };
};

#include "batch.hpp"

}  // namespace

struct mrx_hii_batch {
  mrx_ctx* ctx = nullptr;
  int nbox = 0;
  int flags = 0;
  std::vector<Cohort> cohorts;
  std::vector<int> box_cohort, box_slot;  // cohort and position of each box
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t fork = nullptr;
  size_t graph_nodes = 0;
};

namespace {

void destroy_batch(mrx_hii_batch* b)
{
  if (!b) return;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto& c : b->cohorts) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    free_cohort(c);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
}

// Enqueues every cohort on its own stream, forked from and joined back into
// the context stream (captured into the graph when that stream is capturing).
int enqueue_batch(mrx_ctx* ctx, mrx_hii_batch* b)
{
  HIP_TRY(hipEventRecord(b->fork, ctx->stream));
  for (auto& c : b->cohorts) {
    HIP_TRY(hipStreamWaitEvent(c.stream, b->fork, 0));
    CohortRun op{ctx, &c};
    MRX_TRY(cohort_dispatch(c.N, op));
    HIP_TRY(hipEventRecord(c.done, c.stream));
    HIP_TRY(hipStreamWaitEvent(ctx->stream, c.done, 0));
  }
  return MRX_OK;
}

int capture_batch(mrx_ctx* ctx, mrx_hii_batch* b)
{
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batch(ctx, b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != MRX_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(ctx, MRX_EDEVICE, "batch: graph capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  HIP_TRY(hipGraphGetNodes(g, nullptr, &b->graph_nodes));
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return MRX_OK;
}

}  // namespace

// =============================================================== C ABI
// ------------------------------------------- context access for other units
namespace mrx {
hipStream_t ctx_stream(mrx_ctx* c) { return c->stream; }
int ctx_device(mrx_ctx* c) { return c->device; }
int ctx_fail(mrx_ctx* c, int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}
void* ctx_buffer(mrx_ctx* c, int slot, size_t bytes)
{
  if (slot < 0 || slot >= kCtxSlots) {
    ctx_fail(c, MRX_EINVAL, "ctx_buffer: slot %d out of range", slot);
    return nullptr;
  }
  if (c->ext[slot] && c->ext_cap[slot] >= bytes) return c->ext[slot];
  if (c->ext[slot]) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->ext[slot]);
    c->ext[slot] = nullptr;
    c->ext_cap[slot] = 0;
  }
  const size_t nb = bytes ? bytes : 1;
  if (hipMalloc(&c->ext[slot], nb) != hipSuccess) {
    (void)hipGetLastError();
    c->ext[slot] = nullptr;
    ctx_fail(c, MRX_ENOMEM, "device allocation of %zu bytes failed", nb);
    return nullptr;
  }
  c->ext_cap[slot] = nb;
  return c->ext[slot];
}
int ctx_finish(mrx_ctx* c)
{
  if (!c->user_stream) {
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return ctx_fail(c, MRX_EDEVICE, "stream synchronisation failed: %s", hipGetErrorString(e));
  }
  return MRX_OK;
}
}  // namespace mrx

extern "C" {

int mrx_init(int device, mrx_ctx** out)
{
  if (!out) return MRX_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MRX_EDEVICE;
  if (device < 0 || device >= ndev) return MRX_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return MRX_EDEVICE;
  mrx_ctx* ctx = new mrx_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return MRX_EDEVICE;
  }
  ctx->stream = ctx->own_stream;
  ctx->noalloc_scratch = resolve_scratch(device);
  if (const char* e = getenv("MRX_LADDER_RADII")) {
    const int k = atoi(e);
    if (k >= 1 && k <= kLadderRadiiMax) ctx->ladder_radii = k;
  }
  *out = ctx;
  return MRX_OK;
}

int mrx_device_count(void)
{
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

int mrx_free(mrx_ctx* ctx)
{
  if (!ctx) return MRX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  dfree(ctx->U);
  dfree(ctx->T);
  dfree(ctx->V);
  dfree(ctx->state);
  dfree(ctx->tw);
  dfree(ctx->Rtab);
  dfree(ctx->partials);
  dfree(ctx->sums);
  dfree(ctx->stage);
  dfree(ctx->ts_dev);
  dfree(ctx->ts_work);
  dfree(ctx->ts_shells);
  dfree(ctx->wtab);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  for (auto& e : ctx->side_ev)
    if (e) (void)hipEventDestroy(e);
  dfree(ctx->wtab_all);
  if (ctx->TK) (void)hipFree(ctx->TK);
  dfree(ctx->adj_work);
  dfree(ctx->adj_gS);
  dfree(ctx->adj_E);
  dfree(ctx->tbv);
  dfree(ctx->rr_dev);
  delete ctx->rr;
  for (auto& e : ctx->ext) dfree(e);
  delete ctx->tsd;
  delete ctx->tss;
  delete ctx->ts_host;
  for (auto& r : ctx->prof_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return MRX_OK;
}

const char* mrx_last_error(const mrx_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int mrx_synchronize(mrx_ctx* ctx)
{
  if (!ctx) return MRX_EINVAL;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_set_stream(mrx_ctx* ctx, void* s)
{
  if (!ctx) return MRX_EINVAL;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  ctx->user_stream = (s != nullptr);
  return MRX_OK;
}

int mrx_radius_ladder(const mrx_hii_params* p, double* R_out, int max_n)
{
  if (!p || p->grid_dim <= 0 || !(p->delta_r_factor > 1.0)) return MRX_EINVAL;
  std::vector<double> R;
  const int n = ladder(p, R);
  if (n < 0) return n;
  for (int i = 0; i < n && i < max_n && R_out; ++i) R_out[i] = R[i];
  return n;
}

int mrx_find_hii_bubbles_rec(mrx_ctx* ctx, const mrx_hii_params* p, const float* deltax, const float* stars,
                             const float* weighted_sfr, const float* x_e, float* N_rec, float* xH, float* r_bubble,
                             float* J_21, float* z_at_ionization, float* J_21_at_ionization, float* Gamma12,
                             mrx_hii_globals* out)
{
  if (!ctx) return MRX_EINVAL;
  MRX_TRY(validate_hii(ctx, p));
  if (!deltax || !stars || !weighted_sfr || !xH || !r_bubble || !z_at_ionization)
    return set_err(ctx, MRX_EINVAL, "a required grid pointer is NULL");
  if (p->uvb_flag && (!J_21 || !J_21_at_ionization))
    return set_err(ctx, MRX_EINVAL, "uvb_flag needs J_21 and J_21_at_ionization");
  if (p->include_spin_temp && !x_e) return set_err(ctx, MRX_EINVAL, "include_spin_temp needs x_e");
  const bool rec = p->include_recombinations != 0;
  if (rec && (!N_rec || !Gamma12)) return set_err(ctx, MRX_EINVAL, "include_recombinations needs N_rec and Gamma12");
  if (rec && !(p->hubble_h > 0 && p->y_he > 0 && p->y_he < 1 && p->baryon_frac > 0 && p->omega_lambda > 0))
    return set_err(ctx, MRX_EINVAL, "include_recombinations needs hubble_h, y_he, baryon_frac and omega_lambda");
  HIP_TRY(hipSetDevice(ctx->device));
  const int N = p->grid_dim;
  const size_t np = npad(N), nr = nreal(N);
  const float* in[5] = {deltax, stars, weighted_sfr, x_e, rec ? N_rec : nullptr};
  float* outp[6] = {xH, r_bubble, J_21, z_at_ionization, J_21_at_ionization, rec ? Gamma12 : nullptr};
  if (p->pointers_on_device) {
    // the kernels move four cells per 16-byte access
    for (const float* q : {deltax, stars, weighted_sfr, x_e, (const float*)xH, (const float*)r_bubble,
                           (const float*)J_21, (const float*)z_at_ionization, (const float*)J_21_at_ionization,
                           (const float*)in[4], (const float*)outp[5]})
      if (((uintptr_t)q & 15u) != 0) return set_err(ctx, MRX_EINVAL, "device grids must be 16-byte aligned");
    MRX_TRY(dispatch<HiiRun>(N, ctx, p, in, outp, out));
    return finish(ctx);
  }
  // stage: 5 padded inputs + 6 unpadded outputs
  MRX_TRY(ensure_stage(ctx, 5 * np + 6 * nr));
  float* d_in[5];
  float* d_out[6];
  for (int f = 0; f < 5; ++f) d_in[f] = ctx->stage + f * np;
  for (int o = 0; o < 6; ++o) d_out[o] = ctx->stage + 5 * np + o * nr;
  for (int f = 0; f < 5; ++f)
    if (in[f]) HIP_TRY(hipMemcpyAsync(d_in[f], in[f], sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_out[3], z_at_ionization, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
  if (p->uvb_flag)
    HIP_TRY(hipMemcpyAsync(d_out[4], J_21_at_ionization, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
  if (rec) HIP_TRY(hipMemcpyAsync(d_out[5], Gamma12, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
  const float* din_c[5] = {d_in[0], d_in[1], d_in[2], in[3] ? d_in[3] : nullptr, rec ? d_in[4] : nullptr};
  float* dout_c[6] = {d_out[0], d_out[1], d_out[2], d_out[3], d_out[4], rec ? d_out[5] : nullptr};
  MRX_TRY(dispatch<HiiRun>(N, ctx, p, din_c, dout_c, out));
  for (int o = 0; o < 6; ++o) {
    if ((o == 2 || o == 4) && !p->uvb_flag) continue;
    if (o == 5 && !rec) continue;
    HIP_TRY(hipMemcpyAsync(outp[o], d_out[o], sizeof(float) * nr, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (rec) HIP_TRY(hipMemcpyAsync(N_rec, d_in[4], sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_find_hii_bubbles(mrx_ctx* ctx, const mrx_hii_params* p, const float* deltax, const float* stars,
                         const float* weighted_sfr, const float* x_e, float* xH, float* r_bubble, float* J_21,
                         float* z_at_ionization, float* J_21_at_ionization, mrx_hii_globals* out)
{
  if (!ctx) return MRX_EINVAL;
  if (p && p->include_recombinations)
    return set_err(ctx, MRX_EINVAL, "include_recombinations: call mrx_find_hii_bubbles_rec with N_rec and Gamma12");
  return mrx_find_hii_bubbles_rec(ctx, p, deltax, stars, weighted_sfr, x_e, nullptr, xH, r_bubble, J_21,
                                  z_at_ionization, J_21_at_ionization, nullptr, out);
}

int mrx_filter(mrx_ctx* ctx, float* box, int local_ix_start, int slab_nx, int grid_dim, float R, int filter_type,
               double box_size, int on_device)
{
  if (!ctx) return MRX_EINVAL;
  if (filter_type < 0 || filter_type > 2)
    return set_err(ctx, MRX_EFILTER, "ReionFilterType type %d is undefined", filter_type);
  if (!box || grid_dim < 2 || (grid_dim & 1) || slab_nx < 0 || local_ix_start < 0 ||
      local_ix_start + slab_nx > grid_dim || !(box_size > 0))
    return set_err(ctx, MRX_EINVAL, "bad filter arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  const long n = (long)slab_nx * grid_dim * (grid_dim / 2 + 1);
  if (n == 0) return MRX_OK;
  float2* d = (float2*)box;
  if (!on_device) {
    MRX_TRY(ensure_stage(ctx, 2 * (size_t)n));
    d = (float2*)ctx->stage;
    HIP_TRY(hipMemcpyAsync(d, box, sizeof(float2) * n, hipMemcpyHostToDevice, ctx->stream));
  }
  const WinArgs w = win_args(box_size, R, filter_type, 1);
  k_filter_herm<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(d, local_ix_start, slab_nx, grid_dim, w);
  MRX_TRY(check_launch(ctx, "k_filter_herm"));
  if (!on_device) {
    HIP_TRY(hipMemcpyAsync(box, d, sizeof(float2) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MRX_OK;
  }
  return finish(ctx);
}

int mrx_fft_r2c_3d(mrx_ctx* ctx, int N, const float* in_padded, float* out_herm, int on_device)
{
  if (!ctx) return MRX_EINVAL;
  if (!supported_n(N)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", N);
  if (!in_padded || !out_herm) return set_err(ctx, MRX_EINVAL, "NULL grid");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t np = npad(N);
  if (on_device) {
    MRX_TRY(dispatch<R2C>(N, ctx, in_padded, (float2*)out_herm));
    return finish(ctx);
  }
  MRX_TRY(ensure_stage(ctx, 2 * np));
  float* din = ctx->stage;
  float* dout = ctx->stage + np;
  HIP_TRY(hipMemcpyAsync(din, in_padded, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  MRX_TRY(dispatch<R2C>(N, ctx, (const float*)din, (float2*)dout));
  HIP_TRY(hipMemcpyAsync(out_herm, dout, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_fft_c2r_3d(mrx_ctx* ctx, int N, const float* in_herm, float* out_padded, int on_device)
{
  if (!ctx) return MRX_EINVAL;
  if (!supported_n(N)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", N);
  if (!in_herm || !out_padded) return set_err(ctx, MRX_EINVAL, "NULL grid");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t np = npad(N);
  if (on_device) {
    MRX_TRY(dispatch<C2R>(N, ctx, (const float2*)in_herm, out_padded));
    return finish(ctx);
  }
  MRX_TRY(ensure_stage(ctx, 2 * np));
  float* din = ctx->stage;
  float* dout = ctx->stage + np;
  HIP_TRY(hipMemcpyAsync(din, in_herm, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  MRX_TRY(dispatch<C2R>(N, ctx, (const float2*)din, dout));
  HIP_TRY(hipMemcpyAsync(out_padded, dout, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_filtered_field(mrx_ctx* ctx, int N, double box_size, int filter_type, const float* field_padded, double R,
                       int apply_filter, int clamp_kind, float* out_padded, int on_device)
{
  if (!ctx) return MRX_EINVAL;
  if (!supported_n(N)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", N);
  if (filter_type < 0 || filter_type > 2)
    return set_err(ctx, MRX_EFILTER, "ReionFilterType type %d is undefined", filter_type);
  if (!field_padded || !out_padded || clamp_kind < -1 || clamp_kind > 2 || !(box_size > 0))
    return set_err(ctx, MRX_EINVAL, "bad filtered_field arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t np = npad(N);
  if (on_device) {
    MRX_TRY(dispatch<FilteredField>(N, ctx, box_size, filter_type, field_padded, R, apply_filter, clamp_kind,
                                    out_padded));
    return finish(ctx);
  }
  MRX_TRY(ensure_stage(ctx, 2 * np));
  float* din = ctx->stage;
  float* dout = ctx->stage + np;
  HIP_TRY(hipMemcpyAsync(din, field_padded, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  MRX_TRY(dispatch<FilteredField>(N, ctx, box_size, filter_type, (const float*)din, R, apply_filter, clamp_kind,
                                  dout));
  HIP_TRY(hipMemcpyAsync(out_padded, dout, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}


int mrx_compute_ts(mrx_ctx* ctx, const mrx_ts_params* p, const float* deltax, const float* stars, const float* sfr,
                   const float* sfr_histories, float* x_e_box, float* x_e_box_prev, float* Tk_box, float* TS_box,
                   mrx_ts_globals* out)
{
  if (!ctx) return MRX_EINVAL;
  if (!p) return set_err(ctx, MRX_EINVAL, "params is NULL");
  if (!supported_n(p->grid_dim)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", p->grid_dim);
  if (p->num_filter_steps < 1 || p->num_filter_steps > kTsMaxSteps)
    return set_err(ctx, MRX_EUNSUPPORTED, "TsNumFilterSteps %d outside 1..%d", p->num_filter_steps, kTsMaxSteps);
  if (p->heating_filter_type < 0 || p->heating_filter_type > 2)
    return set_err(ctx, MRX_EFILTER, "TsHeatingFilterType %d is undefined", p->heating_filter_type);
  if (!deltax || !stars || !sfr || !x_e_box || !x_e_box_prev || !Tk_box || !TS_box || !p->redshifts ||
      !p->stored_fcoll || !p->tables_dir || p->snapshot < 0)
    return set_err(ctx, MRX_EINVAL, "a required pointer is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  const std::string key = ts_key(p);
  if (!ctx->tsd || ctx->ts_key != key) {
    if (!ctx->tsd) ctx->tsd = new mrx_ts_data();
    if (mrx_ts_data_init(ctx->tsd, p) != MRX_OK) {
      ctx->ts_key.clear();
      return set_err(ctx, MRX_EINVAL, "cannot read the 21cmFAST tables under %s", p->tables_dir);
    }
    ctx->ts_key = key;
  }
  if (!ctx->tss) ctx->tss = new mrx_ts_snap();
  if (!ctx->ts_host) ctx->ts_host = new TsDev();
  if (const int prc = mrx_ts_plan(ctx->tss, ctx->tsd, p); prc != MRX_OK)
    return prc == MRX_EUNSUPPORTED
             ? set_err(ctx, prc, "a spin-temperature shell needs more than %d SFR-history operations",
                       MRX_TS_MAX_SHELL_OPS)
             : set_err(ctx, MRX_EINVAL, "bad spin-temperature plan");
  const mrx_ts_snap* s = ctx->tss;
  if (!s->homogeneous && s->max_slot >= 0 && (!sfr_histories || s->max_slot >= p->n_history))
    return set_err(ctx, MRX_EINVAL, "the shells need SFR history slot %d, %d held", s->max_slot, p->n_history);
  const int N = p->grid_dim;
  const size_t np = npad(N), nr = nreal(N);
  if (p->pointers_on_device) {
    MRX_TRY(dispatch<TsRun>(N, ctx, p, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box, TS_box, out));
    HIP_TRY(hipMemcpyAsync(x_e_box, x_e_box_prev, sizeof(float) * np, hipMemcpyDeviceToDevice, ctx->stream));
    return finish(ctx);
  }
  const int nslot = s->homogeneous ? 0 : (s->max_slot + 1);
  MRX_TRY(ensure_stage(ctx, (4 + (size_t)nslot) * np + 2 * nr));
  float* st = ctx->stage;
  float *d_dx = st, *d_st = st + np, *d_sfr = st + 2 * np, *d_xe = st + 3 * np;
  float *d_Tk = st + 4 * np, *d_TS = st + 4 * np + nr, *d_hist = st + 4 * np + 2 * nr;
  HIP_TRY(hipMemcpyAsync(d_dx, deltax, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_st, stars, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_sfr, sfr, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_xe, x_e_box_prev, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_Tk, Tk_box, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
  if (nslot > 0)
    HIP_TRY(hipMemcpyAsync(d_hist, sfr_histories, sizeof(float) * np * nslot, hipMemcpyHostToDevice, ctx->stream));
  MRX_TRY(dispatch<TsRun>(N, ctx, p, (const float*)d_dx, (const float*)d_st, (const float*)d_sfr,
                          (const float*)d_hist, d_xe, d_Tk, d_TS, out));
  HIP_TRY(hipMemcpyAsync(x_e_box_prev, d_xe, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(x_e_box, d_xe, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(Tk_box, d_Tk, sizeof(float) * nr, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(TS_box, d_TS, sizeof(float) * nr, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_brightness_temp(mrx_ctx* ctx, const mrx_tb_params* p, const float* xH, const float* deltax,
                        const float* TS_box, float* delta_T, double* ave_Tb)
{
  if (!ctx) return MRX_EINVAL;
  if (!p || p->grid_dim < 2 || (p->grid_dim & 1) || !xH || !deltax || !delta_T)
    return set_err(ctx, MRX_EINVAL, "bad brightness_temp arguments");
  if (p->include_spin_temp && !TS_box) return set_err(ctx, MRX_EINVAL, "include_spin_temp needs TS_box");
  HIP_TRY(hipSetDevice(ctx->device));
  const int N = p->grid_dim;
  const size_t np = npad(N), nr = nreal(N);
  // BrightnessTemperature.c:46-52, :80
  TbArgs a;
  const float redshift = (float)(float)p->redshift;
  const float Hubble_h = (float)(float)p->hubble_h;
  const float OmegaM = (float)(float)p->omega_m;
  const float OmegaB = (float)(OmegaM * p->baryon_frac);
  a.const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                           sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  a.T_rad = kTcmb * (1. + redshift);
  a.N = N;
  a.use_ts = p->include_spin_temp;
  const float *dxH = xH, *ddx = deltax, *dTS = TS_box;
  float* ddT = delta_T;
  if (!p->pointers_on_device) {
    MRX_TRY(ensure_stage(ctx, np + 3 * nr));
    float* s = ctx->stage;
    HIP_TRY(hipMemcpyAsync(s, deltax, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s + np, xH, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
    if (p->include_spin_temp)
      HIP_TRY(hipMemcpyAsync(s + np + nr, TS_box, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
    ddx = s;
    dxH = s + np;
    dTS = p->include_spin_temp ? s + np + nr : nullptr;
    ddT = s + np + 2 * nr;
  }
  const int nb = kTbBlocks;
  MRX_TRY(ensure_partials(ctx, nb + 8));
  if (!ctx->sums) HIP_TRY(hipMalloc(&ctx->sums, sizeof(double) * 8));
  {
  PROF("brightness_temp");
  k_brightness_temp<<<nb, kNT, 0, ctx->stream>>>(dxH, ddx, dTS, ddT, ctx->partials, (long)nr, a);
  }
  MRX_TRY(check_launch(ctx, "k_brightness_temp"));
  k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(ctx->partials, nb, 1, ctx->sums);
  MRX_TRY(check_launch(ctx, "k_reduce_partials"));
  double sum = 0;
  HIP_TRY(hipMemcpyAsync(&sum, ctx->sums, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (!p->pointers_on_device)
    HIP_TRY(hipMemcpyAsync(delta_T, ddT, sizeof(float) * nr, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (ave_Tb) *ave_Tb = sum / pow((double)N, 3);
  return MRX_OK;
}

int mrx_brightness_temp_noalloc(mrx_ctx* ctx, const mrx_tb_params* p, const float* xH, const float* deltax,
                                const float* TS_box, float* delta_T, double* ave_Tb)
{
  if (!ctx) return MRX_EINVAL;
  if (!p || !xH || !deltax || !delta_T) return set_err(ctx, MRX_EINVAL, "bad brightness_temp_noalloc arguments");
  if (!p->pointers_on_device)
    return set_err(ctx, MRX_EINVAL, "brightness_temp_noalloc takes device pointers only (host grids need staging)");
  if (p->grid_dim < 2 || (p->grid_dim & 1))
    return set_err(ctx, MRX_EINVAL, "brightness_temp_noalloc: grid_dim %d must be even and >= 2", p->grid_dim);
  if (p->grid_dim > kNoallocMaxDim)
    return set_err(ctx, MRX_EUNSUPPORTED, "brightness_temp_noalloc: grid_dim %d is above its maximum %d",
                   p->grid_dim, kNoallocMaxDim);
  if (p->include_spin_temp && !TS_box) return set_err(ctx, MRX_EINVAL, "include_spin_temp needs TS_box");
  double* sp = ctx->noalloc_scratch;
  if (!sp) return set_err(ctx, MRX_EDEVICE, "brightness_temp_noalloc: the device scratch address was not resolved");
  HIP_TRY(hipSetDevice(ctx->device));
  const int N = p->grid_dim;
  const TbArgs a = tb_args(p);
  // every partial is written by its workgroup before the reduction reads it, and
  // the total by the reduction: nothing in `scratch` needs clearing first
  k_brightness_temp<<<kTbBlocks, kNT, 0, ctx->stream>>>(xH, deltax, p->include_spin_temp ? TS_box : nullptr,
                                                        delta_T, sp, (long)nreal(N), a);
  MRX_TRY(check_launch(ctx, "k_brightness_temp"));
  k_reduce_partials<<<1, kNT, 0, ctx->stream>>>(sp, kTbBlocks, 1, sp + kTbBlocks);
  MRX_TRY(check_launch(ctx, "k_reduce_partials"));
  double sum = 0;
  HIP_TRY(hipMemcpyAsync(&sum, sp + kTbBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (ave_Tb) *ave_Tb = sum / pow((double)N, 3);
  return MRX_OK;
}

int mrx_brightness_temp_vel(mrx_ctx* ctx, const mrx_tb_vel_params* p, const float* xH, const float* deltax,
                            const float* TS_box, float* vel, float* delta_T, double* ave_Tb)
{
  if (!ctx) return MRX_EINVAL;
  if (!p) return set_err(ctx, MRX_EINVAL, "brightness_temp_vel: null params");
  // velocities act only with the spin temperature (BrightnessTemperature.c:159)
  if (!p->tb.include_spin_temp || p->pec_vels <= 0)
    return mrx_brightness_temp(ctx, &p->tb, xH, deltax, TS_box, delta_T, ave_Tb);
  const int N = p->tb.grid_dim;
  if (p->pec_vels > 3) return set_err(ctx, MRX_EINVAL, "pec_vels %d: Flag_IncludePecVelsFor21cm is 0..3", p->pec_vels);
  if (!xH || !deltax || !TS_box || !vel || !delta_T) return set_err(ctx, MRX_EINVAL, "brightness_temp_vel: null grid");
  if (!supported_n(N)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not built", N);
  if (!(p->box_size > 0)) return set_err(ctx, MRX_EINVAL, "brightness_temp_vel: box_size must be > 0");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t np = npad(N), nr = nreal(N);
  const float *dxH = xH, *ddx = deltax, *dTS = TS_box;
  float *dvel = vel, *ddT = delta_T;
  if (!p->tb.pointers_on_device) {
    MRX_TRY(ensure_stage(ctx, 2 * np + 3 * nr));
    float* s = ctx->stage;
    HIP_TRY(hipMemcpyAsync(s, deltax, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s + np, vel, sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s + 2 * np, xH, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s + 2 * np + nr, TS_box, sizeof(float) * nr, hipMemcpyHostToDevice, ctx->stream));
    ddx = s;
    dvel = s + np;
    dxH = s + 2 * np;
    dTS = s + 2 * np + nr;
    ddT = s + 2 * np + 2 * nr;
  }
  double ave = 0.0;
  MRX_TRY(dispatch<TbVel>(N, ctx, p, dxH, ddx, dTS, dvel, ddT, &ave));
  if (!p->tb.pointers_on_device) {
    HIP_TRY(hipMemcpyAsync(vel, dvel, sizeof(float) * np, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(delta_T, ddT, sizeof(float) * nr, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (ave_Tb) *ave_Tb = ave;
  return finish(ctx);
}

int mrx_profile(mrx_ctx* ctx, int enable)
{
  if (!ctx) return MRX_EINVAL;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->prof_recs) {
    ctx->event_pool.push_back(r.a);
    ctx->event_pool.push_back(r.b);
  }
  ctx->prof_recs.clear();
  ctx->prof = enable != 0;
  return MRX_OK;
}

int mrx_profile_read(mrx_ctx* ctx, int max_kinds, char* names, int name_len, double* ms_total, long* counts)
{
  if (!ctx || max_kinds < 0 || (max_kinds > 0 && (!names || !ms_total || !counts || name_len < 2)))
    return MRX_EINVAL;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::vector<std::string> kinds;
  std::vector<double> tot;
  std::vector<long> cnt;
  for (auto& r : ctx->prof_recs) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    size_t k = 0;
    while (k < kinds.size() && kinds[k] != r.name) ++k;
    if (k == kinds.size()) {
      kinds.push_back(r.name);
      tot.push_back(0.0);
      cnt.push_back(0);
    }
    tot[k] += ms;
    cnt[k] += 1;
    ctx->event_pool.push_back(r.a);
    ctx->event_pool.push_back(r.b);
  }
  ctx->prof_recs.clear();
  const int n = (int)kinds.size();
  for (int k = 0; k < n && k < max_kinds; ++k) {
    snprintf(names + (size_t)k * name_len, (size_t)name_len, "%s", kinds[k].c_str());
    ms_total[k] = tot[k];
    counts[k] = cnt[k];
  }
  return n;
}

int mrx_power_spectrum(mrx_ctx* ctx, int grid_dim, double box_size_mpc, float k_factor, const float* delta_T,
                       int max_bins, double* k, double* power, double* error, long* counts)
{
  if (!ctx) return MRX_EINVAL;
  if (!supported_n(grid_dim)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", grid_dim);
  if (!delta_T || !(box_size_mpc > 0) || max_bins < 0) return set_err(ctx, MRX_EINVAL, "bad power spectrum arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  return dispatch<PowerSpec>(grid_dim, ctx, box_size_mpc, k_factor, delta_T, max_bins, k, power, error, counts,
                             (const double*)nullptr, (float*)nullptr);
}

int mrx_power_spectrum_adjoint(mrx_ctx* ctx, int grid_dim, double box_size_mpc, float k_factor, const float* delta_T,
                               const double* grad_power, float* grad_delta_T)
{
  if (!ctx) return MRX_EINVAL;
  if (!supported_n(grid_dim)) return set_err(ctx, MRX_EUNSUPPORTED, "grid_dim %d not supported", grid_dim);
  if (!delta_T || !grad_power || !grad_delta_T || !(box_size_mpc > 0))
    return set_err(ctx, MRX_EINVAL, "bad power spectrum adjoint arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  const int rc = dispatch<PowerSpec>(grid_dim, ctx, box_size_mpc, k_factor, delta_T, 0, (double*)nullptr,
                                     (double*)nullptr, (double*)nullptr, (long*)nullptr, grad_power, grad_delta_T);
  return rc < 0 ? rc : MRX_OK;
}

int mrx_snapshot_adjoint(mrx_ctx* ctx, const mrx_ts_params* ts, const mrx_hii_params* hii, const mrx_tb_params* tb,
                         const float* deltax, const float* stars, const float* sfr, const float* sfr_histories,
                         const float* x_e_box_prev, const float* Tk_box, const float* xH, const float* TS_box,
                         const float* x_e_box, const float* grad_delta_T, float* grad_deltax, float* grad_stars,
                         float* grad_sfr, float* grad_sfr_histories, float* grad_x_e_box_prev, float* grad_Tk_box)
{
  if (!ctx) return MRX_EINVAL;
  MRX_TRY(validate_hii(ctx, hii));
  if (!tb || tb->grid_dim != hii->grid_dim) return set_err(ctx, MRX_EINVAL, "adjoint: tb params missing or mismatched");
  if (!hii->pointers_on_device || !tb->pointers_on_device || (ts && !ts->pointers_on_device))
    return set_err(ctx, MRX_EINVAL, "adjoint: grids must be device pointers");
  if (!deltax || !stars || !xH || !grad_delta_T || !grad_deltax || !grad_stars)
    return set_err(ctx, MRX_EINVAL, "adjoint: a required grid pointer is NULL");
  if (tb->include_spin_temp && !TS_box) return set_err(ctx, MRX_EINVAL, "adjoint: include_spin_temp needs TS_box");
  if (hii->include_spin_temp && !x_e_box) return set_err(ctx, MRX_EINVAL, "adjoint: include_spin_temp needs x_e_box");
  if (ts) {
    if (ts->grid_dim != hii->grid_dim) return set_err(ctx, MRX_EINVAL, "adjoint: ts grid_dim mismatch");
    if (!sfr || !x_e_box_prev || !Tk_box || !TS_box || !ts->redshifts || !ts->stored_fcoll || !ts->tables_dir)
      return set_err(ctx, MRX_EINVAL, "adjoint: spin-temperature inputs missing");
    if (ts->num_filter_steps < 1 || ts->num_filter_steps > kTsMaxSteps)
      return set_err(ctx, MRX_EUNSUPPORTED, "TsNumFilterSteps %d outside 1..%d", ts->num_filter_steps, kTsMaxSteps);
    if (ts->heating_filter_type < 0 || ts->heating_filter_type > 2)
      return set_err(ctx, MRX_EFILTER, "TsHeatingFilterType %d is undefined", ts->heating_filter_type);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  if (ts) {
    const std::string key = ts_key(ts);
    if (!ctx->tsd || ctx->ts_key != key) {
      if (!ctx->tsd) ctx->tsd = new mrx_ts_data();
      if (mrx_ts_data_init(ctx->tsd, ts) != MRX_OK) {
        ctx->ts_key.clear();
        return set_err(ctx, MRX_EINVAL, "cannot read the 21cmFAST tables under %s", ts->tables_dir);
      }
      ctx->ts_key = key;
    }
    if (!ctx->tss) ctx->tss = new mrx_ts_snap();
    if (!ctx->ts_host) ctx->ts_host = new TsDev();
    if (const int prc = mrx_ts_plan(ctx->tss, ctx->tsd, ts); prc != MRX_OK)
      return prc == MRX_EUNSUPPORTED
               ? set_err(ctx, prc, "a spin-temperature shell needs more than %d SFR-history operations",
                         MRX_TS_MAX_SHELL_OPS)
               : set_err(ctx, MRX_EINVAL, "bad spin-temperature plan");
    const mrx_ts_snap* s = ctx->tss;
    if (!s->homogeneous && s->max_slot >= 0 && (!sfr_histories || s->max_slot >= ts->n_history))
      return set_err(ctx, MRX_EINVAL, "the shells need SFR history slot %d, %d held", s->max_slot, ts->n_history);
  }
  MRX_TRY(dispatch<SnapAdj>(hii->grid_dim, ctx, ts, hii, tb, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box,
                            xH, TS_box, x_e_box, grad_delta_T, grad_deltax, grad_stars, grad_sfr, grad_sfr_histories,
                            grad_x_e_box_prev, grad_Tk_box));
  return finish(ctx);
}

int mrx_hii_batch_create(mrx_ctx* ctx, const mrx_hii_batch_box* boxes, int n_boxes, int flags,
                         mrx_hii_batch** out)
{
  if (!ctx || !out) return MRX_EINVAL;
  *out = nullptr;
  if (!boxes || n_boxes <= 0) return set_err(ctx, MRX_EINVAL, "batch: no boxes");
  for (int i = 0; i < n_boxes; ++i) {
    const mrx_hii_batch_box& b = boxes[i];
    MRX_TRY(validate_hii(ctx, &b.params));
    if (b.params.include_recombinations)
      return set_err(ctx, MRX_EUNSUPPORTED, "batch: box %d: recombinations are not batched (mrx_find_hii_bubbles_rec)", i);
    if (!b.params.pointers_on_device)
      return set_err(ctx, MRX_EINVAL, "batch: box %d: grids must be device pointers", i);
    if (!b.deltax || !b.stars || !b.weighted_sfr || !b.xH || !b.r_bubble || !b.z_at_ionization)
      return set_err(ctx, MRX_EINVAL, "batch: box %d: a required grid pointer is NULL", i);
    if (b.params.uvb_flag && (!b.J_21 || !b.J_21_at_ionization))
      return set_err(ctx, MRX_EINVAL, "batch: box %d: uvb_flag needs J_21 and J_21_at_ionization", i);
    if (b.params.include_spin_temp && !b.x_e)
      return set_err(ctx, MRX_EINVAL, "batch: box %d: include_spin_temp needs x_e", i);
    for (const float* q : {b.deltax, b.stars, b.weighted_sfr, b.x_e, (const float*)b.xH, (const float*)b.r_bubble,
                           (const float*)b.J_21, (const float*)b.z_at_ionization,
                           (const float*)b.J_21_at_ionization})
      if (((uintptr_t)q & 15u) != 0)
        return set_err(ctx, MRX_EINVAL, "batch: box %d: device grids must be 16-byte aligned", i);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  mrx_hii_batch* b = new mrx_hii_batch();
  b->ctx = ctx;
  b->nbox = n_boxes;
  b->flags = flags;
  b->box_cohort.assign(n_boxes, -1);
  b->box_slot.assign(n_boxes, -1);
  for (int i = 0; i < n_boxes; ++i) {
    int k = 0;
    while (k < (int)b->cohorts.size() && !same_cohort(b->cohorts[k].p0, boxes[i].params)) ++k;
    if (k == (int)b->cohorts.size()) {
      b->cohorts.emplace_back();
      b->cohorts[k].p0 = boxes[i].params;
    }
    b->box_cohort[i] = k;
    b->box_slot[i] = (int)b->cohorts[k].boxes.size();
    b->cohorts[k].boxes.push_back(i);
  }
  int rc = MRX_OK;
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    rc = set_err(ctx, MRX_EDEVICE, "batch: event creation failed");
  for (auto& c : b->cohorts) {
    if (rc != MRX_OK) break;
    c.N = c.p0.grid_dim;
    c.F = c.p0.include_spin_temp ? 4 : 3;
    c.nbox = (int)c.boxes.size();
    c.S = ladder(&c.p0, c.R);
    if (c.S < 1) rc = set_err(ctx, MRX_EINVAL, "batch: radius ladder is empty");
    else if (c.S > kMaxSteps)
      rc = set_err(ctx, MRX_EUNSUPPORTED, "batch: radius ladder has %d steps (max %d)", c.S, kMaxSteps);
    if (rc != MRX_OK) break;
    CohortSetup op{ctx, &c, boxes};
    rc = cohort_dispatch(c.N, op);
    if (rc == MRX_OK && !(flags & MRX_BATCH_NO_TUNE)) {
      CohortTune t{ctx, &c};
      rc = cohort_dispatch(c.N, t);
    }
  }
  if (rc == MRX_OK && !(flags & MRX_BATCH_NO_GRAPH)) rc = capture_batch(ctx, b);
  if (rc == MRX_OK) {
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = set_err(ctx, MRX_EDEVICE, "batch: %s", hipGetErrorString(e));
  }
  if (rc != MRX_OK) {
    destroy_batch(b);
    return rc;
  }
  *out = b;
  return MRX_OK;
}

int mrx_hii_batch_run(mrx_hii_batch* b, mrx_hii_globals* out)
{
  if (!b) return MRX_EINVAL;
  mrx_ctx* ctx = b->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  if (b->exec)
    HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
  else
    MRX_TRY(enqueue_batch(ctx, b));
  if (!out) return finish(ctx);
  std::vector<std::vector<double>> sums(b->cohorts.size());
  for (size_t k = 0; k < b->cohorts.size(); ++k) {
    sums[k].resize(4 * (size_t)b->cohorts[k].nbox);
    HIP_TRY(hipMemcpyAsync(sums[k].data(), b->cohorts[k].sums, sizeof(double) * sums[k].size(),
                           hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < b->nbox; ++i) {
    const Cohort& c = b->cohorts[b->box_cohort[i]];
    const double* h = sums[b->box_cohort[i]].data() + 4 * (size_t)b->box_slot[i];
    const double total_n_cells = pow((double)c.N, 3);
    out[i].volume_weighted_global_xH = h[0] / total_n_cells;
    out[i].mass_weighted_global_xH = h[1] / h[2];
    out[i].volume_weighted_global_J_21 = c.p0.uvb_flag ? h[3] / total_n_cells : 0.0;
    out[i].n_radii = c.S;
  }
  return MRX_OK;
}

int mrx_hii_batch_describe(const mrx_hii_batch* b, char* buf, int len)
{
  if (!b || !buf || len <= 0) return MRX_EINVAL;
  std::string s = "{\"graph\": ";
  s += b->exec ? "true" : "false";
  s += ", \"graph_nodes\": " + std::to_string(b->graph_nodes) + ", \"cohorts\": [";
  for (size_t k = 0; k < b->cohorts.size(); ++k) {
    const Cohort& c = b->cohorts[k];
    CohortShape sh{c.variant, 0, 0};
    (void)cohort_dispatch(c.N, sh);
    char tmp[256];
    snprintf(tmp, sizeof tmp, "%s{\"grid_dim\": %d, \"boxes\": %d, \"fields\": %d, \"radii\": %d, \"variant\": %d, "
             "\"tile_width\": %d, \"radix_schedule\": %d, \"tune_ms\": [", k ? ", " : "", c.N, c.nbox, c.F, c.S,
             c.variant, sh.tw, sh.sv);
    s += tmp;
    for (size_t v = 0; v < c.tune_ms.size(); ++v) {
      snprintf(tmp, sizeof tmp, "%s%.4f", v ? ", " : "", c.tune_ms[v]);
      s += tmp;
    }
    s += "]}";
  }
  s += "]}";
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return (int)s.size();
}

int mrx_hii_batch_free(mrx_hii_batch* b)
{
  destroy_batch(b);
  return MRX_OK;
}

// Size of each ABI struct, so bindings can check their mirrors.
int mrx_ts_last_shells(mrx_ctx* ctx, int grid_dim, int n_shells, float* out, int pointers_on_device)
{
  if (!ctx) return MRX_EINVAL;
  if (!out) return set_err(ctx, MRX_EINVAL, "ts_last_shells: null output");
  if (ctx->ts_shells_n == 0 || ctx->ts_shells_N != grid_dim || ctx->ts_shells_n != n_shells)
    return set_err(ctx, MRX_EINVAL, "ts_last_shells: the context holds %d shells of %d^3 (asked %d of %d^3)",
                   ctx->ts_shells_n, ctx->ts_shells_N, n_shells, grid_dim);
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t bytes = sizeof(float) * (size_t)n_shells * nreal(grid_dim);
  HIP_TRY(hipMemcpyAsync(out, ctx->ts_shells, bytes, pointers_on_device ? hipMemcpyDeviceToDevice
                                                                         : hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MRX_OK;
}

int mrx_ts_snapshot_tables(const mrx_ts_params* p, double collapse_fraction, double x_e_ave, mrx_ts_tables* out)
{
  if (!p || !out || !p->tables_dir) return MRX_EINVAL;
  if (p->num_filter_steps < 1 || p->num_filter_steps > MRX_TS_TABLES_MAX_STEPS) return MRX_EUNSUPPORTED;
  std::unique_ptr<mrx_ts_data> d(new mrx_ts_data());
  std::unique_ptr<mrx_ts_snap> s(new mrx_ts_snap());
  if (mrx_ts_data_init(d.get(), p) != MRX_OK) return MRX_EINVAL;
  const int rc = mrx_ts_plan(s.get(), d.get(), p);
  if (rc != MRX_OK) return rc;
  int nop = 0;
  for (int c = 0; c < s->nsteps; ++c) nop += s->n_ops[c];
  if (nop > MRX_TS_TABLES_MAX_OPS) return MRX_EUNSUPPORTED;
  memset(out, 0, sizeof *out);
  out->nsteps = s->nsteps;
  out->homogeneous = s->homogeneous;
  out->zp = s->zp;
  out->dzp = s->dzp;
  out->dt_dzp = s->dt_dzp;
  if (s->homogeneous) return MRX_OK;
  s->collapse_fraction = collapse_fraction;
  s->x_e_ave = x_e_ave;
  MRX_TRY(mrx_ts_build_tables(s.get(), d.get(), p));
  out->no_light = s->no_light;
  out->dt_dzp = s->dt_dzp;
  out->const_zp_prefactor = s->const_zp_prefactor;
  out->growth_factor_zp = s->growth_factor_zp;
  out->dgrowth_factor_dzp = s->dgrowth_factor_dzp;
  for (int c = 0; c < s->nsteps; ++c) {
    out->R[c] = s->R[c];
    out->zpp_edge[c] = s->zpp_edge[c];
    out->zpp[c] = s->zpp[c];
    out->dzpp[c] = s->dzpp[c];
    out->dt_dzpp[c] = s->dt_dzpp[c];
    out->zpp_pow[c] = s->zpp_pow[c];
    out->sum_lyn[c] = s->sum_lyn[c];
    out->nu_lower[c] = s->nu_lower[c];
    for (int x = 0; x < MRX_X_INT_NXHII; ++x) {
      out->freq_heat[x][c] = s->freq_heat[x][c];
      out->freq_ion[x][c] = s->freq_ion[x][c];
      out->freq_lya[x][c] = s->freq_lya[x][c];
    }
    for (int k = 0; k < s->n_ops[c]; ++k) {
      const int i = out->n_ops++;
      out->op_shell[i] = c;
      out->op_kind[i] = s->ops[c][k].kind;
      out->op_slot[i] = s->ops[c][k].slot;
      out->op_weight[i] = s->ops[c][k].weight;
    }
  }
  return MRX_OK;
}

size_t mrx_abi_sizeof(int which)
{
  switch (which) {
    case 0: return sizeof(mrx_hii_params);
    case 1: return sizeof(mrx_hii_globals);
    case 2: return sizeof(mrx_tb_params);
    case 3: return sizeof(mrx_ts_params);
    case 4: return sizeof(mrx_ts_globals);
    case 5: return sizeof(mrx_hii_batch_box);
    case 6: return sizeof(mrx_grid_params);
    case 7: return sizeof(mrx_grid_stats);
    case 8: return sizeof(mrx_tb_vel_params);
    case 9: return sizeof(mrx_ts_tables);
    default: return 0;
  }
}

}  // extern "C"
