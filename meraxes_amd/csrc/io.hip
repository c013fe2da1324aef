// io.hip - the reionization grid files: gen_grids_fname, save_reion_input_grids
// and save_reion_output_grids (src/core/reionization.c:1617-1646, :1648-1754,
// :1786-2050) for grids held in HBM.
//
// The reference de-pads each padded grid (and converts sfr / weighted_sfr to
// M_sun/yr) in a host loop into an unpadded buffer, then every rank writes its
// x-slab of each N^3 float dataset (chunked {1, N, N}) collectively.  Here one
// process writes the whole box: a kernel de-pads and converts on the device
// (k_unpad), the unpadded grid goes to pinned host memory in x-slabs through
// two staging buffers, and the host hands slab s to HDF5 while the copy of
// slab s+1 runs.  Unpadded grids already in host memory go to HDF5 directly.
#include <hip/hip_runtime.h>

#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/mrx_io.h"
#include "ctx.hpp"

#define MRX_IO_TRY(x)                \
  do {                               \
    const int rc_ = (x);             \
    if (rc_ != MRX_OK) return rc_;   \
  } while (0)

namespace mrx {
namespace {

constexpr int kIT = 256;
constexpr double kSecPerYear = 3.155e7;  // SEC_PER_YEAR (src/meraxes.h:33)
constexpr double kSolarMass = 1.989e33;  // SOLAR_MASS (src/meraxes.h:22)
constexpr size_t kSlabBytes = 16u << 20; // one staging buffer (two are used)

// device buffer slots of ctx_buffer used here (the gridding uses 0-19)
enum { S_IO_PAD = 20, S_IO_OUT = 21 };

#pragma clang fp contract(off)
// dst[i, j, k] = src at the padded index (INDEX_PADDED -> INDEX_REAL), times
// UnitMass_in_g / UnitTime_in_s * SEC_PER_YEAR / SOLAR_MASS in the reference's
// order when `scale` (:1691-1693): float * double, then left to right in double,
// then rounded to float.
__global__ __launch_bounds__(kIT) void k_unpad(const float* __restrict__ src, float* __restrict__ dst, int N,
                                               int scale, double unit_mass_g, double unit_time_s)
{
  const long n = (long)N * N * N;
  for (long i = (long)blockIdx.x * kIT + threadIdx.x; i < n; i += (long)gridDim.x * kIT) {
    const float v = src[(i / N) * (N + 2) + i % N];
    dst[i] = scale ? (float)((double)v * unit_mass_g / unit_time_s * kSecPerYear / kSolarMass) : v;
  }
}
#pragma clang fp contract(on)

// Restores HDF5's automatic error printing on scope exit.
struct QuietH5 {
  H5E_auto2_t func = nullptr;
  void* data = nullptr;
  QuietH5()
  {
    H5Eget_auto2(H5E_DEFAULT, &func, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

// Closes an HDF5 handle on scope exit.
struct H5Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Id()
  {
    if (id >= 0 && close) close(id);
  }
  H5Id(const H5Id&) = delete;
  H5Id& operator=(const H5Id&) = delete;
  operator hid_t() const { return id; }
};

enum Kind { UNPADDED, PADDED, PADDED_MSUN_PER_YR };

class Writer {
 public:
  Writer(mrx_ctx* ctx, const mrx_save_params* p, hid_t file) : ctx_(ctx), p_(p), file_(file), N_(p->grid_dim) {}
  ~Writer()
  {
    // a failed write may leave a slab copy in flight into the pinned buffers
    if (pin_[0]) (void)hipStreamSynchronize(ctx_stream(ctx_));
    for (int b = 0; b < 2; ++b) {
      if (ev_[b]) (void)hipEventDestroy(ev_[b]);
      if (pin_[b]) (void)hipHostFree(pin_[b]);
    }
  }

  // One N^3 float dataset chunked {1, N, N} (write_grid_float, :1617-1638, with
  // the creation properties of :1670-1683).
  int grid(const char* name, const float* src, Kind kind)
  {
    const hsize_t n = (hsize_t)N_;
    const hsize_t dims[3] = {n, n, n};
    H5Id fspace(H5Screate_simple(3, dims, nullptr), H5Sclose);
    H5Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    const hsize_t chunk[3] = {1, n, n};
    if (fspace < 0 || dcpl < 0 || H5Pset_chunk(dcpl, 3, chunk) < 0) return io_fail("dataspace of %s", name);
    H5Id dset(H5Dcreate2(file_, name, H5T_NATIVE_FLOAT, fspace, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
    if (dset < 0) return io_fail("creating dataset %s", name);
    const long nr = (long)N_ * N_ * N_;
    if (!p_->pointers_on_device && kind == UNPADDED) {
      // host grid in the file's layout: HDF5 reads it in place
      if (H5Dwrite(dset, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, src) < 0)
        return io_fail("writing dataset %s", name);
      return MRX_OK;
    }
    hipStream_t st = ctx_stream(ctx_);
    const float* dev = src;
    if (kind != UNPADDED) {
      const long np = (long)N_ * N_ * (N_ + 2);
      if (!p_->pointers_on_device) {
        float* d = static_cast<float*>(ctx_buffer(ctx_, S_IO_PAD, sizeof(float) * np));
        if (!d) return MRX_ENOMEM;
        if (hipMemcpyAsync(d, src, sizeof(float) * np, hipMemcpyHostToDevice, st) != hipSuccess)
          return dev_fail("staging %s", name);
        src = d;
      }
      float* out = static_cast<float*>(ctx_buffer(ctx_, S_IO_OUT, sizeof(float) * nr));
      if (!out) return MRX_ENOMEM;
      const long nb = (nr + kIT - 1) / kIT;
      k_unpad<<<(unsigned)(nb < 8192 ? nb : 8192), kIT, 0, st>>>(src, out, N_, kind == PADDED_MSUN_PER_YR ? 1 : 0,
                                                                 p_->unit_mass_g, p_->unit_time_s);
      if (hipGetLastError() != hipSuccess) return dev_fail("k_unpad for %s", name);
      dev = out;
    }
    MRX_IO_TRY(staging());
    // x-slabs of `planes` planes through the two pinned buffers
    const long plane = (long)N_ * N_;
    const int nslab = (N_ + planes_ - 1) / planes_;
    auto enqueue = [&](int s) -> int {
      const int x0 = s * planes_, px = (N_ - x0 < planes_) ? N_ - x0 : planes_;
      if (hipMemcpyAsync(pin_[s & 1], dev + (long)x0 * plane, sizeof(float) * plane * px, hipMemcpyDeviceToHost,
                         st) != hipSuccess ||
          hipEventRecord(ev_[s & 1], st) != hipSuccess)
        return dev_fail("copying %s", name);
      return MRX_OK;
    };
    auto write = [&](int s) -> int {
      const int x0 = s * planes_, px = (N_ - x0 < planes_) ? N_ - x0 : planes_;
      if (hipEventSynchronize(ev_[s & 1]) != hipSuccess) return dev_fail("copying %s", name);
      const hsize_t start[3] = {(hsize_t)x0, 0, 0}, count[3] = {(hsize_t)px, n, n};
      H5Id mspace(H5Screate_simple(3, count, nullptr), H5Sclose);
      if (mspace < 0 || H5Sselect_hyperslab(fspace, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0)
        return io_fail("selecting planes %d-%d of %s", x0, x0 + px - 1, name);
      if (H5Dwrite(dset, H5T_NATIVE_FLOAT, mspace, fspace, H5P_DEFAULT, pin_[s & 1]) < 0)
        return io_fail("writing planes %d-%d of %s", x0, x0 + px - 1, name);
      return MRX_OK;
    };
    MRX_IO_TRY(enqueue(0));
    for (int s = 1; s < nslab; ++s) {
      // buffer s & 1 held slab s - 2, which the previous iteration wrote
      MRX_IO_TRY(enqueue(s));
      MRX_IO_TRY(write(s - 1));
    }
    return write(nslab - 1);
  }

  // A {len} float dataset with default (contiguous) creation properties from
  // host memory (:1966-2040).
  int vector(const char* name, const float* host, int len)
  {
    const hsize_t dims[1] = {(hsize_t)len};
    H5Id space(H5Screate_simple(1, dims, nullptr), H5Sclose);
    if (space < 0) return io_fail("dataspace of %s", name);
    H5Id dset(H5Dcreate2(file_, name, H5T_NATIVE_FLOAT, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (dset < 0) return io_fail("creating dataset %s", name);
    if (H5Dwrite(dset, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, host) < 0)
      return io_fail("writing dataset %s", name);
    return MRX_OK;
  }

  int attr(const char* dset, const char* name, double v)
  {
    if (H5LTset_attribute_double(file_, dset, name, &v, 1) < 0) return io_fail("attribute %s of %s", name, dset);
    return MRX_OK;
  }

  int io_fail(const char* fmt, ...)
  {
    char msg[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof msg, fmt, ap);
    va_end(ap);
    return ctx_fail(ctx_, MRX_EIO, "HDF5: %s", msg);
  }

  int dev_fail(const char* what, const char* name)
  {
    char msg[192];
    snprintf(msg, sizeof msg, what, name);
    return ctx_fail(ctx_, MRX_EDEVICE, "%s: %s", msg, hipGetErrorString(hipGetLastError()));
  }

 private:
  int staging()
  {
    if (pin_[0]) return MRX_OK;
    const size_t plane = sizeof(float) * (size_t)N_ * N_;
    planes_ = (int)(kSlabBytes / plane);
    if (planes_ < 1) planes_ = 1;
    if (planes_ > N_) planes_ = N_;
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc((void**)&pin_[b], plane * planes_, hipHostMallocDefault) != hipSuccess) {
        pin_[b] = nullptr;
        return ctx_fail(ctx_, MRX_ENOMEM, "pinned staging of %zu bytes", plane * planes_);
      }
      if (hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming) != hipSuccess)
        return ctx_fail(ctx_, MRX_EDEVICE, "hipEventCreate failed");
    }
    return MRX_OK;
  }

  mrx_ctx* ctx_;
  const mrx_save_params* p_;
  hid_t file_;
  int N_;
  int planes_ = 1;
  float* pin_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
};

int check_params(mrx_ctx* ctx, const mrx_save_params* p, const char* path)
{
  if (!ctx) return MRX_EINVAL;
  if (!p || !path) return ctx_fail(ctx, MRX_EINVAL, "null params or path");
  if (p->grid_dim < 2 || p->grid_dim > 4096) return ctx_fail(ctx, MRX_EINVAL, "grid_dim %d", p->grid_dim);
  return MRX_OK;
}

}  // namespace
}  // namespace mrx

using namespace mrx;

extern "C" int mrx_grids_fname(char* buf, size_t len, const char* output_dir, const char* file_name_galaxies,
                               int snapshot, int relative)
{
  if (!buf || !file_name_galaxies || (!relative && !output_dir)) return MRX_EINVAL;
  const int n = relative ? snprintf(buf, len, "%s_grids_%d.hdf5", file_name_galaxies, snapshot)
                         : snprintf(buf, len, "%s/%s_grids_%d.hdf5", output_dir, file_name_galaxies, snapshot);
  if (n < 0 || (size_t)n >= len) return MRX_EINVAL;
  return n;
}

extern "C" int mrx_save_reion_input_grids(mrx_ctx* ctx, const mrx_save_params* p, const char* path,
                                          const float* deltax, const float* stars, const float* weighted_sfr,
                                          const float* sfr)
{
  MRX_IO_TRY(check_params(ctx, p, path));
  if (!deltax || !stars || !weighted_sfr || (p->include_spin_temp && !sfr))
    return ctx_fail(ctx, MRX_EINVAL, "save_reion_input_grids: a grid its flags need is null");
  QuietH5 quiet;
  H5Id file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file < 0) return ctx_fail(ctx, MRX_EIO, "HDF5: cannot create %s", path);
  {
    Writer w(ctx, p, file);
    if (p->include_spin_temp) MRX_IO_TRY(w.grid("sfr", sfr, PADDED_MSUN_PER_YR));
    MRX_IO_TRY(w.grid("deltax", deltax, PADDED));
    MRX_IO_TRY(w.grid("stars", stars, PADDED));
    MRX_IO_TRY(w.grid("weighted_sfr", weighted_sfr, PADDED_MSUN_PER_YR));
  }
  if (H5Fflush(file, H5F_SCOPE_LOCAL) < 0) return ctx_fail(ctx, MRX_EIO, "HDF5: flushing %s", path);
  return MRX_OK;
}

extern "C" int mrx_save_reion_output_grids(mrx_ctx* ctx, const mrx_save_params* p, const char* path,
                                           const mrx_output_grids* g)
{
  MRX_IO_TRY(check_params(ctx, p, path));
  if (!g || !g->xH || !g->z_at_ionization || !g->r_bubble ||
      (p->uvb_flag && (!g->J_21 || !g->J_21_at_ionization || !g->Mvir_crit)) ||
      (p->include_spin_temp && (!g->TS_box || !g->Tk_box || !g->x_e_box_prev)) ||
      (p->compute_21cm_bright_temp && !g->delta_T) ||
      (p->compute_ps && (p->ps_length < 1 || !g->PS_k || !g->PS_data || !g->PS_error)))
    return ctx_fail(ctx, MRX_EINVAL, "save_reion_output_grids: a grid its flags need is null");
  QuietH5 quiet;
  H5Id file(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
  if (file < 0) return ctx_fail(ctx, MRX_EIO, "HDF5: cannot open %s for writing", path);
  {
    Writer w(ctx, p, file);
    MRX_IO_TRY(w.grid("xH", g->xH, UNPADDED));
    MRX_IO_TRY(w.grid("z_at_ionization", g->z_at_ionization, UNPADDED));
    MRX_IO_TRY(w.grid("r_bubble", g->r_bubble, UNPADDED));
    if (p->uvb_flag) {
      MRX_IO_TRY(w.grid("J_21", g->J_21, UNPADDED));
      MRX_IO_TRY(w.attr("J_21", "volume_weighted_global_J_21", g->volume_weighted_global_J_21));
      MRX_IO_TRY(w.grid("J_21_at_ionization", g->J_21_at_ionization, UNPADDED));
      MRX_IO_TRY(w.grid("Mvir_crit", g->Mvir_crit, UNPADDED));
    }
    if (p->include_spin_temp) {
      MRX_IO_TRY(w.grid("TS_box", g->TS_box, UNPADDED));
      MRX_IO_TRY(w.grid("Tk_box", g->Tk_box, UNPADDED));
      MRX_IO_TRY(w.grid("x_e_box", g->x_e_box_prev, PADDED));
    }
    if (p->compute_21cm_bright_temp) MRX_IO_TRY(w.grid("delta_T", g->delta_T, UNPADDED));
    MRX_IO_TRY(w.attr("xH", "volume_weighted_global_xH", g->volume_weighted_global_xH));
    MRX_IO_TRY(w.attr("xH", "mass_weighted_global_xH", g->mass_weighted_global_xH));
    if (p->include_spin_temp) {
      MRX_IO_TRY(w.attr("TS_box", "volume_ave_TS", g->volume_ave_TS));
      MRX_IO_TRY(w.attr("Tk_box", "volume_ave_TK", g->volume_ave_TK));
      MRX_IO_TRY(w.attr("x_e_box", "volume_ave_xe", g->volume_ave_xe));
      MRX_IO_TRY(w.attr("TS_box", "volume_ave_J_alpha", g->volume_ave_J_alpha));
      MRX_IO_TRY(w.attr("TS_box", "volume_ave_xalpha", g->volume_ave_xalpha));
      MRX_IO_TRY(w.attr("TS_box", "volume_ave_Xheat", g->volume_ave_Xheat));
      MRX_IO_TRY(w.attr("TS_box", "volume_ave_Xion", g->volume_ave_Xion));
    }
    if (p->compute_21cm_bright_temp) MRX_IO_TRY(w.attr("delta_T", "volume_ave_Tb", g->volume_ave_Tb));
    if (p->compute_ps) {
      MRX_IO_TRY(w.vector("k_bins", g->PS_k, p->ps_length));
      MRX_IO_TRY(w.vector("PS_data", g->PS_data, p->ps_length));
      MRX_IO_TRY(w.vector("PS_error", g->PS_error, p->ps_length));
    }
  }
  if (H5Fflush(file, H5F_SCOPE_LOCAL) < 0) return ctx_fail(ctx, MRX_EIO, "HDF5: flushing %s", path);
  return MRX_OK;
}

extern "C" size_t mrx_io_abi_sizeof(int which)
{
  switch (which) {
    case 0:
      return sizeof(mrx_save_params);
    case 1:
      return sizeof(mrx_output_grids);
    default:
      return 0;
  }
}
