// batch.hpp - many boxes through the bubble finder in one graph launch.
//
// Boxes that share the grid, the box size, the radius ladder, the filter, the
// flags and the mass-radius constants form a cohort; every kernel of a cohort
// covers all of its boxes at once (blockIdx.y is the box, or box*F + field), so
// even 64^3 boxes fill the chip.  Per-box scalars that only enter the
// threshold (efficiency, redshift, halo bias and the J_21 constant of each
// step) come from device tables.  Cohorts run on their own streams; the whole
// computation, every radius of every box, is captured into one HIP graph at
// plan time and replayed by a single hipGraphLaunch.
//
// Each cohort's strided passes use a (tile width, radix schedule) variant
// chosen by timing the candidates on the cohort itself when the plan is made.
// Included by mrx_reion.hip inside its anonymous namespace.
#pragma once

// ---------------------------------------------------------- tile variants
// Variant v of grid N: tile width TWd, TWd/2 or 2*TWd (v / 2), radix
// schedule v % 2, where TWd is the default width.  Only the grids of the
// tuned set carry more than the default variant.
template <int N>
constexpr bool kTunedN = (N == 64 || N == 128 || N == 192 || N == 256);
constexpr int kNumVariants = 6;

template <int N, int V>
struct VariantOf {
  static constexpr int TWd = Cfg<N>::TW;
  static constexpr int TW = (V / 2 == 0) ? TWd : (V / 2 == 1 ? TWd / 2 : TWd * 2);
  static constexpr int SV = V % 2;
  static constexpr bool ok = (V == 0) || (kTunedN<N> && TW >= 4 && (N / 2) % TW == 0 &&
                                          (long)N * (TW + 2) * 8 + (long)N * 8 <= 96 * 1024);
};

// Calls fn.template go<TW, SV>() for runtime variant v; MRX_EINVAL if the
// variant does not exist for N.
template <int N, class Fn>
int with_variant(int v, Fn& fn)
{
  switch (v) {
#define MRX_VAR_CASE(V_)                                                                  \
  case V_:                                                                                \
    if constexpr (VariantOf<N, V_>::ok)                                                   \
      return fn.template go<VariantOf<N, V_>::TW, VariantOf<N, V_>::SV>();                \
    else                                                                                  \
      return MRX_EINVAL;
    MRX_VAR_CASE(0)
    MRX_VAR_CASE(1)
    MRX_VAR_CASE(2)
    MRX_VAR_CASE(3)
    MRX_VAR_CASE(4)
    MRX_VAR_CASE(5)
#undef MRX_VAR_CASE
    default:
      return MRX_EINVAL;
  }
}

template <int N>
bool variant_exists(int v)
{
  switch (v) {
    case 0: return VariantOf<N, 0>::ok;
    case 1: return VariantOf<N, 1>::ok;
    case 2: return VariantOf<N, 2>::ok;
    case 3: return VariantOf<N, 3>::ok;
    case 4: return VariantOf<N, 4>::ok;
    case 5: return VariantOf<N, 5>::ok;
    default: return false;
  }
}

template <int N>
void variant_shape(int v, int& tw, int& sv)
{
  tw = (v / 2 == 0) ? Cfg<N>::TW : (v / 2 == 1 ? Cfg<N>::TW / 2 : Cfg<N>::TW * 2);
  sv = v % 2;
}

// ------------------------------------------------------------------ cohort
struct Cohort {
  int N = 0, F = 0, nbox = 0, S = 0;
  int variant = 0;
  std::vector<int> boxes;  // indices into the batch
  mrx_hii_params p0;       // parameters shared by the cohort (box 0's)
  std::vector<double> R;
  float2* U = nullptr;
  float2* T = nullptr;
  float2* V = nullptr;
  uint8_t* state = nullptr;
  float2* tw = nullptr;
  float* Rtab = nullptr;
  float* wtab = nullptr;
  double* partials = nullptr;
  double* sums = nullptr;
  BoxDev* boxes_dev = nullptr;
  double* J21c = nullptr;
  const float** in_tab = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<double> tune_ms;  // measured time of each variant (-1: not a candidate)
};

void free_cohort(Cohort& c)
{
  dfree(c.U);
  dfree(c.T);
  dfree(c.V);
  dfree(c.state);
  dfree(c.tw);
  dfree(c.Rtab);
  dfree(c.wtab);
  dfree(c.partials);
  dfree(c.sums);
  dfree(c.boxes_dev);
  dfree(c.J21c);
  if (c.in_tab) (void)hipFree((void*)c.in_tab);
  c.in_tab = nullptr;
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  if (c.done) (void)hipEventDestroy(c.done);
  c.done = nullptr;
}

// J_21 constant of find_HII_bubbles.c:317-320 at radius R, evaluated exactly
// as the single-box path does.
double j21_const(const mrx_hii_params* p, double R)
{
  const double z = p->redshift;
  return (1.0 + z) * (1.0 + z) / (4.0 * M_PI) * p->alpha_uv * kPlanck * 1e21 * R * p->unit_length_cm *
         p->nion_phot_per_bary / kProtonMass * p->unit_mass_g / pow(p->unit_length_cm, 3) / p->unit_time_s;
}

bool same_cohort(const mrx_hii_params& a, const mrx_hii_params& b)
{
  return a.grid_dim == b.grid_dim && a.filter_type == b.filter_type && a.rtom_filter_type == b.rtom_filter_type &&
         a.uvb_flag == b.uvb_flag && a.include_spin_temp == b.include_spin_temp && a.box_size == b.box_size &&
         a.r_bubble_max == b.r_bubble_max && a.r_bubble_min == b.r_bubble_min &&
         a.delta_r_factor == b.delta_r_factor && a.omega_m == b.omega_m && a.rho_crit == b.rho_crit;
}

template <int N>
struct BatchPipe {
  static constexpr int NH = N / 2;
  static constexpr long nc = (long)N * N * NH;
  static int nb_threshold() { return N * N / ThrCfg<N>::RZ; }

  struct Fwd {
    mrx_ctx* ctx;
    Cohort* c;
    template <int TW, int SV>
    int go()
    {
      using M = TileMap<N, TW>;
      const int nf = c->nbox * c->F;
      const dim3 gc(N * M::NKZB, nf);
      k_colpass<N, false, TW, SV><<<gc, M::NT, 0, c->stream>>>(c->U, c->tw, nc, (long)N * NH, NH, 0.0);
      MRX_TRY(check_launch(ctx, "k_colpass y (batch)"));
      k_colpass<N, false, TW, SV><<<gc, M::NT, 0, c->stream>>>(c->U, c->tw, nc, NH, (long)N * NH, std::pow((double)N, 3));
      return check_launch(ctx, "k_colpass x (batch)");
    }
  };

  struct InvXY {
    mrx_ctx* ctx;
    Cohort* c;
    const WinArgs* w;  // NULL: the unfiltered last step
    template <int TW, int SV>
    int go()
    {
      using M = TileMap<N, TW>;
      using MX = TileMap<N, TW, kNT>;
      const bool win = (w != nullptr);
      if (win) {  // the table layout follows the tile width
        const long nt = (long)(NH + 1) * (NH + 1) * (NH + 1);
        k_window_table<N, TW><<<window_table_blocks<N, TW>(), kNT, 0, c->stream>>>(c->wtab, *w);
        MRX_TRY(check_launch(ctx, "k_window_table (batch)"));
      }
      const dim3 gx(N * M::NKZB, c->nbox);
      const long bs = (long)c->F * nc, bvs = (long)c->F * N * N;
      if (win)
        k_ix_filter<N, kWinTable, TW, SV><<<gx, MX::NT, 0, c->stream>>>(c->U, c->V, c->T, c->tw, c->wtab, nc,
                                                                       (long)N * N, c->F, bs, bvs, *w, N, M::NKZB);
      else
        k_ix_filter<N, kWinNone, TW, SV><<<gx, MX::NT, 0, c->stream>>>(c->U, c->V, c->T, c->tw, c->wtab, nc,
                                                                      (long)N * N, c->F, bs, bvs, WinArgs{}, N,
                                                                      M::NKZB);
      MRX_TRY(check_launch(ctx, "k_ix_filter (batch)"));
      const dim3 gy(N * M::NKZB, c->nbox * c->F);
      k_colpass<N, true, TW, SV><<<gy, M::NT, 0, c->stream>>>(c->T, c->tw, nc, (long)N * NH, NH, 0.0);
      return check_launch(ctx, "k_colpass y inv (batch)");
    }
  };

  static int forward(mrx_ctx* ctx, Cohort& c)
  {
    const int nf = c.nbox * c.F;
    k_r2c_z_tab<N><<<dim3(N * N / Cfg<N>::RZ, nf), kNT, 0, c.stream>>>(c.in_tab, c.U, c.tw, nc);
    MRX_TRY(check_launch(ctx, "k_r2c_z_tab"));
    Fwd f{ctx, &c};
    MRX_TRY(with_variant<N>(c.variant, f));
    k_unpack_dc<N><<<dim3((N * N + 255) / 256, nf), 256, 0, c.stream>>>(c.U, c.V, nc, (long)N * N);
    return check_launch(ctx, "k_unpack_dc (batch)");
  }

  static int inverse_xy(mrx_ctx* ctx, Cohort& c, const WinArgs* w, int variant)
  {
    InvXY f{ctx, &c, w};
    return with_variant<N>(variant, f);
  }

  // the whole ladder of every box of the cohort, enqueued on c.stream
  static int ladder_all(mrx_ctx* ctx, Cohort& c)
  {
    MRX_TRY(forward(ctx, c));
    const mrx_hii_params* p = &c.p0;
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.pixel_volume = pow(p->box_size / (double)N, 3);
    a.uvb = p->uvb_flag;
    a.use_xe = p->include_spin_temp;
    a.nfield = c.F;
    a.kzlive = N / 2;
    const int nb = nb_threshold();
    for (int s = 0; s < c.S; ++s) {
      const bool last = (s == c.S - 1);
      const WinArgs w = win_args(p->box_size, (float)c.R[s], p->filter_type, 1);
      MRX_TRY(inverse_xy(ctx, c, last ? nullptr : &w, c.variant));
      a.M_mean = RtoM(p, c.R[s]);
      a.R_cubed = c.R[s] * c.R[s] * c.R[s];
      a.kfc = (4.0 / 3.0) * M_PI * a.R_cubed / a.pixel_volume / a.M_mean;
      a.step = s;
      a.last = last ? 1 : 0;
      constexpr int NT = ThrCfg<N>::NT;
      if (c.F == 4)
        k_c2r_threshold_batch<N, 4><<<dim3(nb, c.nbox), NT, 0, c.stream>>>(
          c.T, nc, (long)c.F * nc, c.tw, c.boxes_dev, c.state, c.Rtab, c.J21c, c.S, c.partials, a);
      else
        k_c2r_threshold_batch<N, 3><<<dim3(nb, c.nbox), NT, 0, c.stream>>>(
          c.T, nc, (long)c.F * nc, c.tw, c.boxes_dev, c.state, c.Rtab, c.J21c, c.S, c.partials, a);
      MRX_TRY(check_launch(ctx, "k_c2r_threshold_batch"));
    }
    // the same reduction as the single-box path, so each box's globals are identical to it
    return reduce_partials(ctx, c.stream, c.partials, nb, 4, c.nbox, c.partials + (size_t)nb * 4 * c.nbox, c.sums);
  }

  // Allocates the cohort's buffers and uploads its tables.
  static int setup(mrx_ctx* ctx, Cohort& c, const mrx_hii_batch_box* bx)
  {
    const size_t nf = (size_t)c.nbox * c.F;
    auto alloc = [&](void** p, size_t bytes) -> int {
      if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(ctx, MRX_ENOMEM, "batch: device allocation of %zu bytes failed", bytes);
      }
      return MRX_OK;
    };
    MRX_TRY(alloc((void**)&c.U, sizeof(float2) * nc * nf));
    MRX_TRY(alloc((void**)&c.T, sizeof(float2) * nc * nf));
    MRX_TRY(alloc((void**)&c.V, sizeof(float2) * (size_t)N * N * nf));
    MRX_TRY(alloc((void**)&c.state, (size_t)N * N * N * c.nbox));
    MRX_TRY(alloc((void**)&c.tw, sizeof(float2) * N));
    MRX_TRY(alloc((void**)&c.Rtab, sizeof(float) * c.S));
    MRX_TRY(alloc((void**)&c.wtab, sizeof(float) * (size_t)(NH + 1) * (NH + 1) * (NH + 1)));
    MRX_TRY(alloc((void**)&c.partials,
                  sizeof(double) * ((size_t)nb_threshold() * 4 + (size_t)kRedBlocks * kRedMaxV) * c.nbox));
    MRX_TRY(alloc((void**)&c.sums, sizeof(double) * 4 * c.nbox));
    MRX_TRY(alloc((void**)&c.boxes_dev, sizeof(BoxDev) * c.nbox));
    MRX_TRY(alloc((void**)&c.J21c, sizeof(double) * c.nbox * c.S));
    MRX_TRY(alloc((void**)&c.in_tab, sizeof(float*) * nf));
    std::vector<float2> h(N);
    for (int t = 0; t < N; ++t) {
      const double an = -2.0 * M_PI * (double)t / (double)N;
      h[t] = make_float2((float)cos(an), (float)sin(an));
    }
    HIP_TRY(hipMemcpy(c.tw, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
    std::vector<float> Rf(c.S);
    for (int s = 0; s < c.S; ++s) Rf[s] = (float)c.R[s];
    HIP_TRY(hipMemcpy(c.Rtab, Rf.data(), sizeof(float) * c.S, hipMemcpyHostToDevice));
    std::vector<BoxDev> bd(c.nbox);
    std::vector<double> j(c.nbox * c.S);
    std::vector<const float*> in(nf);
    for (int i = 0; i < c.nbox; ++i) {
      const mrx_hii_batch_box& b = bx[c.boxes[i]];
      bd[i].deltax = b.deltax;
      bd[i].xH = b.xH;
      bd[i].r_bubble = b.r_bubble;
      bd[i].J21 = b.J_21;
      bd[i].z_in = b.z_at_ionization;
      bd[i].J21_ion = b.J_21_at_ionization;
      bd[i].eff = b.params.efficiency;
      bd[i].redshift = b.params.redshift;
      bd[i].bias = b.params.gamma_halo_bias;
      for (int s = 0; s < c.S; ++s) j[(size_t)i * c.S + s] = j21_const(&b.params, c.R[s]);
      const float* src[4] = {b.deltax, b.stars, b.weighted_sfr, b.x_e};
      for (int f = 0; f < c.F; ++f) in[(size_t)i * c.F + f] = src[f];
    }
    HIP_TRY(hipMemcpy(c.boxes_dev, bd.data(), sizeof(BoxDev) * c.nbox, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c.J21c, j.data(), sizeof(double) * j.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((void*)c.in_tab, in.data(), sizeof(float*) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
    return MRX_OK;
  }

  // Times every variant's window table + filtered inverse x+y passes over the whole cohort
  // (the radius of step 0, real spectra from one forward transform) and keeps
  // the fastest.
  static int tune(mrx_ctx* ctx, Cohort& c)
  {
    c.tune_ms.assign(kNumVariants, -1.0);
    int ncand = 0;
    for (int v = 0; v < kNumVariants; ++v) ncand += variant_exists<N>(v) ? 1 : 0;
    if (ncand <= 1) return MRX_OK;
    c.variant = 0;
    MRX_TRY(forward(ctx, c));
    const WinArgs w0 = win_args(c.p0.box_size, (float)c.R[0], c.p0.filter_type, 1);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e300;
    int bestv = 0;
    for (int v = 0; v < kNumVariants; ++v) {
      if (!variant_exists<N>(v)) continue;
      MRX_TRY(inverse_xy(ctx, c, &w0, v));  // warm-up
      HIP_TRY(hipEventRecord(e0, c.stream));
      for (int r = 0; r < 3; ++r) MRX_TRY(inverse_xy(ctx, c, &w0, v));
      HIP_TRY(hipEventRecord(e1, c.stream));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      c.tune_ms[v] = ms / 3.0;
      if (ms < best) {
        best = ms;
        bestv = v;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.variant = bestv;
    return MRX_OK;
  }
};

template <class Op>
int cohort_dispatch(int N, Op& op)
{
  switch (N) {
    case 32: return op.template go<32>();
    case 48: return op.template go<48>();
    case 64: return op.template go<64>();
    case 96: return op.template go<96>();
    case 128: return op.template go<128>();
    case 192: return op.template go<192>();
    case 256: return op.template go<256>();
    case 384: return op.template go<384>();
    case 512: return op.template go<512>();
    default: return MRX_EUNSUPPORTED;
  }
}

struct CohortSetup {
  mrx_ctx* ctx;
  Cohort* c;
  const mrx_hii_batch_box* bx;
  template <int N>
  int go() { return BatchPipe<N>::setup(ctx, *c, bx); }
};
struct CohortTune {
  mrx_ctx* ctx;
  Cohort* c;
  template <int N>
  int go() { return BatchPipe<N>::tune(ctx, *c); }
};
struct CohortRun {
  mrx_ctx* ctx;
  Cohort* c;
  template <int N>
  int go() { return BatchPipe<N>::ladder_all(ctx, *c); }
};
struct CohortShape {
  int v, tw, sv;
  template <int N>
  int go()
  {
    variant_shape<N>(v, tw, sv);
    return MRX_OK;
  }
};
