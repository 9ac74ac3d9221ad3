// C-ABI of the MCV step on MI355X (include/mcv_hip.h): handle lifetime,
// host-side preparation of the static mesh data, launch sequences for the
// RK3_TVD and RK4 schemes, read-back and FP64 integrals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mcv_hip.h"
#include "mcv_handle.hpp"
#include "mcv_kernels_fp64.hpp"

using namespace mcv;

namespace mcv {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

size_t footprint_bytes(const mcv_config *c) {
  const size_t H = c->dof - 1, npv = (size_t)c->n * H + 1, nh = npv + 2 * H;
  const size_t np = 6 * nh * nh, nband = band_values(c->dof, (int)npv), pad = 256;
  const size_t ncell = 6 * (size_t)c->n * c->n;
  const size_t diag_blocks =
      std::min<size_t>((ncell + kDiagThreads - 1) / kDiagThreads, kDiagMaxBlocks);
  auto d = [&](size_t count) { return count * sizeof(double) + pad; };
  auto f = [&](size_t count) { return count * sizeof(float) + pad; };
  // static mesh data (panel-invariant metric terms counted on every panel)
  size_t b = 5 * d(np) + d(np) + d(np) + d(4 * np) + d(H * npv) +
             sizeof(int32_t) * H * npv + pad;
  b += d(4 * np) + d(2 * np);  // packed metric terms
  b += d(4 * 6 * 4 * (H + 1) * npv) + d(4 * 6 * 4 * H * npv) + d(4 * 6 * 4 * npv);  // band tables
  const bool fp32c = c->precision == MCV_FP32C;
  b += (fp32c ? 3 : 9) * d(np);                       // FP64 levels
  if (!fp32c) b += 3 * d(np);                          // RK4 accumulator
  b += d(nband) + 2 * d(np) + d(4 * diag_blocks) + d(2) + 6 * d(np);
  if (fp32c)  // join, phi_ref, phi_ref*sqrtG, FP32 {reference gradient, Coriolis, 0},
              // levels, compensation, accumulator, bands
    b += 3 * d(np) + d(np) + d(np) + 4 * f(np) + 3 * f(3 * np) + f(3 * np) + f(3 * np) +
         2 * f(nband);
  return b;
}

StageArgs stage_args64(const mcv_handle *h, int in, int out, double c1) {
  StageArgs a{};
  a.in_pu = h->lpu[in]; a.in_v = h->lvv[in];
  a.b_pu = h->lpu[0]; a.b_v = h->lvv[0];
  a.acc_pu = reinterpret_cast<D2 *>(h->acc[0]); a.acc_v = h->acc[2];
  a.o_pu = h->lpu[out]; a.o_v = h->lvv[out];
  a.band_raw = h->band_raw;
  a.c1 = c1;
  return a;
}

BoundaryArgs boundary_args64(const mcv_handle *h, int level) {
  return BoundaryArgs{h->lpu[level], h->lvv[level], h->band_raw};
}

StageArgs32 stage_args32(const mcv_handle *h, int in, int out, double c1) {
  StageArgs32 a{};
  a.in = h->l32[in];
  a.b = h->l32[0];
  a.c = h->c32;
  a.acc = h->a32;
  a.o = h->l32[out];
  if (out == 0) a.oc = h->c32;  // the step's last stage writes Y_n+1 with its compensation
  a.band_raw = h->band32;
  a.band_raw_c = h->band32c;
  a.c1 = c1;
  return a;
}

BoundaryArgs32 boundary_args32(const mcv_handle *h, int level) {
  const bool base = level == 0;  // Y_n carries its compensation words
  BoundaryArgs32 b{};
  b.lev = h->l32[level];
  if (base) b.comp = h->c32;
  b.band_raw = h->band32;
  b.band_raw_c = h->band32c;
  return b;
}

}  // namespace mcv

namespace {

// Scratch of the allocation-free entry point mcv_hip_tendency_noalloc: one
// copy per device (the code object's own static memory, no hipMalloc), held
// until the process ends.  Host code never passes `scratch` itself to a
// kernel or a copy (in host code the name is only a stand-in for the device
// symbol): it uses the device address that hipGetSymbolAddress returned for
// the current device, kept in g_scratch.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];
std::mutex g_scratch_mu;
char *g_scratch[kMaxDevices] = {};

// Looks up this device's scratch address once per device (the device must be
// current); called from mcv_hip_init.
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(MCV_EINVAL, "device %d is beyond the %d devices the library supports", device,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return MCV_OK;
  void *p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = static_cast<char *>(p);
  return MCV_OK;
}

char *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

int dev_alloc(mcv_handle *h, double **out, size_t count) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(double) + 256);
  if (e != hipSuccess)
    return fail(MCV_ENOMEM, "hipMalloc(%zu doubles) failed: %s", count, hipGetErrorString(e));
  h->allocs.push_back(p);
  h->device_bytes += count * sizeof(double) + 256;
  *out = static_cast<double *>(p);
  return MCV_OK;
}

int upload(mcv_handle *h, double **dst, const double *src, size_t count) {
  int rc = dev_alloc(h, dst, count);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(*dst, src, count * sizeof(double), hipMemcpyHostToDevice));
  return MCV_OK;
}

bool bits_equal(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// True when the six panels of a (nh*nh*6) array hold identical bits.
bool panel_invariant(const double *a, long long per, int width = 1) {
  for (int p = 1; p < 6; ++p)
    if (std::memcmp(a, a + p * per * width, sizeof(double) * per * width) != 0) return false;
  return true;
}

void destroy(mcv_handle *h) {
  if (!h) return;
  if (h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (void *p : h->allocs) (void)hipFree(p);
  delete h;
}

int validate(const mcv_config *c) {
  if (!c) return fail(MCV_EINVAL, "cfg is NULL");
  if (c->abi_version != MCV_ABI_VERSION)
    return fail(MCV_EINVAL, "abi_version %d, library is %d", c->abi_version, MCV_ABI_VERSION);
  if (c->dof != 3 && c->dof != 4) return fail(MCV_EINVAL, "dof must be 3 or 4, got %d", c->dof);
  // every per-point array the stage kernels index stays below 2^32 bytes
  if (c->n < 2 || c->n > 2048) return fail(MCV_EINVAL, "n must be in [2, 2048], got %d", c->n);
  if (c->scheme != MCV_RK4 && c->scheme != MCV_RK3_TVD)
    return fail(MCV_EINVAL, "Unknown integral scheme %d, select MCV_RK3_TVD or MCV_RK4",
                c->scheme);
  if (c->precision != MCV_FP64 && c->precision != MCV_FP32C)
    return fail(MCV_EINVAL, "precision %d: select MCV_FP64 or MCV_FP32C", c->precision);
  if (!(c->dt > 0.0) || !std::isfinite(c->dt)) return fail(MCV_EINVAL, "dt must be > 0");
  if (!(c->dx > 0.0) || !std::isfinite(c->dx)) return fail(MCV_EINVAL, "dx must be > 0");
  return MCV_OK;
}

template <int DOF, int CX, int CY>
int launch_stage_shape(mcv_handle *h, int kind, const StageArgs &a) {
  const int tx = (h->cfg.n + CX - 1) / CX, ty = (h->cfg.n + CY - 1) / CY;
  dim3 grid(tx * ty * 6), block(kStageThreads);
  switch (kind) {
#define MCV_CASE(K)                                                                   \
  case K:                                                                             \
    hipLaunchKernelGGL((stage_kernel<DOF, K, CX, CY>), grid, block, 0, h->stream, h->mesh, \
                       a, tx);                                                        \
    break;
    MCV_CASE(kTend)
    MCV_CASE(kRK4S1)
    MCV_CASE(kRK4S2)
    MCV_CASE(kRK4S3)
    MCV_CASE(kRK4S4)
    MCV_CASE(kRK3S1)
    MCV_CASE(kRK3S2)
    MCV_CASE(kRK3S3)
#undef MCV_CASE
    default:
      return fail(MCV_EINVAL, "bad stage kind %d", kind);
  }
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

template <int DOF>
int launch_stage(mcv_handle *h, int kind, const StageArgs &a) {
  using T = TileSet<DOF>;
  switch (h->tile) {
    case 0: return launch_stage_shape<DOF, T::cx[0], T::cy[0]>(h, kind, a);
    case 1: return launch_stage_shape<DOF, T::cx[1], T::cy[1]>(h, kind, a);
    case 2: return launch_stage_shape<DOF, T::cx[2], T::cy[2]>(h, kind, a);
    default: return launch_stage_shape<DOF, T::cx[3], T::cy[3]>(h, kind, a);
  }
}

int stage(mcv_handle *h, int kind, const StageArgs &a) {
  return h->cfg.dof == 3 ? launch_stage<3>(h, kind, a) : launch_stage<4>(h, kind, a);
}

int boundary(mcv_handle *h, int level) {
  const BoundaryArgs b = boundary_args64(h, level);
  const int threads = boundary_items(h->cfg.dof, h->npv);
  dim3 grid((threads + 255) / 256), block(256);
  if (h->cfg.dof == 3)
    hipLaunchKernelGGL(boundary_kernel<3>, grid, block, 0, h->stream, h->mesh, b);
  else
    hipLaunchKernelGGL(boundary_kernel<4>, grid, block, 0, h->stream, h->mesh, b);
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

// ---------------------------------------------------------------------------
// FP32C launches
// ---------------------------------------------------------------------------
template <int DOF, int CX, int CY>
int launch_stage32_shape(mcv_handle *h, int kind, const StageArgs32 &a) {
  const int tx = (h->cfg.n + CX - 1) / CX, ty = (h->cfg.n + CY - 1) / CY;
  dim3 grid(tx * ty * 6), block(256);
  switch (kind) {
#define MCV_CASE(K)                                                                       \
  case K:                                                                                 \
    hipLaunchKernelGGL((stage_kernel_f32<DOF, K, CX, CY>), grid, block, 0, h->stream,      \
                       h->mesh32, a, tx);                                                 \
    break;
    MCV_CASE(kRK4S1)
    MCV_CASE(kRK4S2)
    MCV_CASE(kRK4S3)
    MCV_CASE(kRK4S4)
    MCV_CASE(kRK3S1)
    MCV_CASE(kRK3S2)
    MCV_CASE(kRK3S3)
#undef MCV_CASE
    default:
      return fail(MCV_EINVAL, "bad FP32C stage kind %d", kind);
  }
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

template <int DOF>
int launch_stage32(mcv_handle *h, int kind, const StageArgs32 &a) {
  using T = TileSet32<DOF>;
  switch (h->tile) {
    case 0: return launch_stage32_shape<DOF, T::cx[0], T::cy[0]>(h, kind, a);
    case 1: return launch_stage32_shape<DOF, T::cx[1], T::cy[1]>(h, kind, a);
    case 2: return launch_stage32_shape<DOF, T::cx[2], T::cy[2]>(h, kind, a);
    default: return launch_stage32_shape<DOF, T::cx[3], T::cy[3]>(h, kind, a);
  }
}

int stage32(mcv_handle *h, int kind, const StageArgs32 &a) {
  return h->cfg.dof == 3 ? launch_stage32<3>(h, kind, a) : launch_stage32<4>(h, kind, a);
}

int boundary32(mcv_handle *h, int level) {
  const BoundaryArgs32 b = boundary_args32(h, level);
  const int threads = boundary_items(h->cfg.dof, h->npv);
  dim3 grid((threads + 255) / 256), block(256);
  if (h->cfg.dof == 3)
    hipLaunchKernelGGL(boundary_kernel_f32<3>, grid, block, 0, h->stream, h->mesh, b);
  else
    hipLaunchKernelGGL(boundary_kernel_f32<4>, grid, block, 0, h->stream, h->mesh, b);
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

// One RK stage: stage kernel from level `in` (base Yn = level 0) into level
// `out`, then the boundary pass on `out`.
int rk_stage(mcv_handle *h, int kind, int in, int out, double c1) {
  StageArgs a = h->fp32c ? StageArgs{} : stage_args64(h, in, out, c1);
  if (h->first_pending) { a.cu0 = h->cu0; a.cv0 = h->cv0; }
  if (h->timing) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
  int rc = h->fp32c ? stage32(h, kind, stage_args32(h, in, out, c1)) : stage(h, kind, a);
  if (rc) return rc;
  h->first_pending = false;
  if (h->timing) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
  rc = h->fp32c ? boundary32(h, out) : boundary(h, out);
  if (rc) return rc;
  if (h->timing) {
    // events bracket each kernel on the handle's stream; wait so they can be reused
    HIP_TRY(hipEventRecord(h->ev[2], h->stream));
    HIP_TRY(hipEventSynchronize(h->ev[2]));
    float a_ms = 0.f, b_ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&a_ms, h->ev[0], h->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b_ms, h->ev[1], h->ev[2]));
    h->stage_ms += a_ms;
    h->boundary_ms += b_ms;
  }
  return MCV_OK;
}

// time_scheme_mod.F90:13-69.  Levels: 0 = Yn (updated in place by the last
// stage), 1 and 2 = intermediate stage states.
int one_step(mcv_handle *h) {
  const double dt = h->cfg.dt;
  int rc;
  if (h->cfg.scheme == MCV_RK4) {
    if ((rc = rk_stage(h, kRK4S1, 0, 1, 0.5 * dt))) return rc;
    if ((rc = rk_stage(h, kRK4S2, 1, 2, 0.5 * dt))) return rc;
    if ((rc = rk_stage(h, kRK4S3, 2, 1, dt))) return rc;
    if ((rc = rk_stage(h, kRK4S4, 1, 0, dt))) return rc;
  } else {
    if ((rc = rk_stage(h, kRK3S1, 0, 1, dt))) return rc;
    if ((rc = rk_stage(h, kRK3S2, 1, 2, 0.25 * dt))) return rc;
    if ((rc = rk_stage(h, kRK3S3, 2, 0, 2.0 / 3.0 * dt))) return rc;
  }
  return MCV_OK;
}

// FP64 (phi, u, v) of Y_n: the level itself, or for FP32C the reconstruction
// phi_ref + hi + lo, hi + lo.
// The FP64 level is unpacked into pack[0] (phi), pack[2] (u) and pack[3] (v),
// the slots the read-back kernels write those fields to.
int materialize(mcv_handle *h, const double **out) {
  if (!h->fp32c) {
    hipLaunchKernelGGL(unpack_pairs_kernel, dim3(1024), dim3(256), 0, h->stream, h->np,
                       h->lpu[0], h->pack[0], h->pack[2]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h->pack[3], h->lvv[0], h->np * sizeof(double),
                           hipMemcpyDeviceToDevice, h->stream));
    out[0] = h->pack[0];
    out[1] = h->pack[2];
    out[2] = h->pack[3];
    return MCV_OK;
  }
  hipLaunchKernelGGL(join_state_kernel, dim3(1024), dim3(256), 0, h->stream, h->np, h->phi_ref,
                     h->l32[0], h->c32, h->join[0], h->join[1], h->join[2]);
  HIP_TRY(hipGetLastError());
  for (int f = 0; f < 3; ++f) out[f] = h->join[f];
  return MCV_OK;
}

int run_diag(mcv_handle *h, double *mass, double *energy) {
  const double *cu0 = h->first_pending ? h->cu0 : nullptr;
  const double *cv0 = h->first_pending ? h->cv0 : nullptr;
  const double *y[3];
  int rc = materialize(h, y);
  if (rc) return rc;
  dim3 grid(h->diag_blocks), block(kDiagThreads);
  if (h->cfg.dof == 3)
    hipLaunchKernelGGL(diag_partial_kernel<3>, grid, block, 0, h->stream, h->mesh, y[0], y[1],
                       y[2], cu0, cv0, h->partials);
  else
    hipLaunchKernelGGL(diag_partial_kernel<4>, grid, block, 0, h->stream, h->mesh, y[0], y[1],
                       y[2], cu0, cv0, h->partials);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(256), 0, h->stream, h->partials,
                     h->diag_blocks, h->diag_out);
  HIP_TRY(hipGetLastError());
  double r[2];
  HIP_TRY(hipMemcpyAsync(r, h->diag_out, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (mass) *mass = r[0];
  if (energy) *energy = r[1];
  if (!std::isfinite(r[0]) || !std::isfinite(r[1]))
    return fail(MCV_ENONFINITE, "non-finite integrals: mass %g energy %g", r[0], r[1]);
  return MCV_OK;
}

template <typename T>
int dev_alloc_t(mcv_handle *h, T **out, size_t count) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
  if (e != hipSuccess)
    return fail(MCV_ENOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                hipGetErrorString(e));
  h->allocs.push_back(p);
  h->device_bytes += count * sizeof(T) + 256;
  *out = static_cast<T *>(p);
  return MCV_OK;
}

// FP32C set-up from the FP64 initial state in level 0.  One FP64 boundary
// pass on a copy makes it continuous across panel edges with an interpolated
// halo; that projected geopotential is the reference phi_ref (kept in level
// 0).  The static FP64 gradient of phi_ref + phi_s follows, then the FP32
// hi/lo split of the initial state itself (phi - phi_ref, u, v), so the first
// stage sees the handed-over halo and edges as the FP64 path does.
int setup32(mcv_handle *h) {
  const long long np = h->np;
  const size_t nband = band_values(h->cfg.dof, h->npv);
  const int threads = 6 * 4 * h->cfg.dof * h->npv;
  int rc;
  for (int f = 0; f < 3; ++f)
    if ((rc = dev_alloc(h, &h->join[f], np))) return rc;
  hipLaunchKernelGGL(unpack_pairs_kernel, dim3(1024), dim3(256), 0, h->stream, np, h->lpu[0],
                     h->join[0], h->join[1]);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(h->join[2], h->lvv[0], np * sizeof(double), hipMemcpyDeviceToDevice,
                         h->stream));
  hipLaunchKernelGGL(band_copy_kernel, dim3((threads + 255) / 256), dim3(256), 0, h->stream,
                     h->mesh, h->lpu[0], h->lvv[0], h->band_raw);
  HIP_TRY(hipGetLastError());
  if ((rc = boundary(h, 0))) return rc;
  if ((rc = dev_alloc(h, &h->phi_ref, np))) return rc;
  hipLaunchKernelGGL(unpack_pairs_kernel, dim3(1024), dim3(256), 0, h->stream, np, h->lpu[0],
                     h->phi_ref, nullptr);
  HIP_TRY(hipGetLastError());
  h->first_pending = false;  // the FP32C stages form the contravariant winds themselves
  F4 *fdE;
  if ((rc = dev_alloc_t(h, &fdE, np))) return rc;
  if (h->cfg.dof == 3)
    hipLaunchKernelGGL(ref_gradient_kernel<3>, dim3(2048), dim3(256), 0, h->stream, h->mesh,
                       h->phi_ref, fdE);
  else
    hipLaunchKernelGGL(ref_gradient_kernel<4>, dim3(2048), dim3(256), 0, h->stream, h->mesh,
                       h->phi_ref, fdE);
  HIP_TRY(hipGetLastError());
  for (int l = 0; l < 3; ++l)
    if ((rc = dev_alloc_t(h, &h->l32[l], np))) return rc;
  if ((rc = dev_alloc_t(h, &h->c32, np)) || (rc = dev_alloc_t(h, &h->a32, np))) return rc;
  if ((rc = dev_alloc_t(h, &h->band32, nband / 3)) ||
      (rc = dev_alloc_t(h, &h->band32c, nband / 3)))
    return rc;
  double *pgref;
  if ((rc = dev_alloc_t(h, &pgref, np))) return rc;
  hipLaunchKernelGGL(pgref_kernel, dim3(1024), dim3(256), 0, h->stream, h->mesh, h->phi_ref,
                     pgref);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(split_state_kernel, dim3(1024), dim3(256), 0, h->stream, np, h->join[0],
                     h->phi_ref, h->join[1], h->join[2], h->l32[0], h->c32);
  HIP_TRY(hipGetLastError());
  for (int l = 1; l < 3; ++l)
    HIP_TRY(hipMemcpyAsync(h->l32[l], h->l32[0], np * sizeof(F3), hipMemcpyDeviceToDevice,
                           h->stream));
  MeshF32 &m = h->mesh32;
  const MeshDev &d = h->mesh;
  m.dof = d.dof;
  m.n = d.n;
  m.npv = d.npv;
  m.nh = d.nh;
  m.rdx = 1.0 / d.dx;
  m.r2dx = 1.0 / (2.0 * d.dx);
  m.rdxx = 1.0 / (d.dx * d.dx);
  m.dxf = d.dx;
  m.metric_pstride = d.metric_pstride;
  m.sg = d.sqrtG;
  m.rsg = d.rsqrtG;
  m.ig11 = d.ig11;
  m.ig12 = d.ig12;
  m.ig22 = d.ig22;
  m.mpack = d.mpack;
  m.spack = d.spack;
  m.pgref = pgref;
  m.fdE = fdE;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MCV_OK;
}

int init_impl(const mcv_config *cfg, const mcv_mesh *mesh, const mcv_state *st,
              mcv_handle *h) {
  int rc = validate(cfg);
  if (rc) return rc;
  if (!mesh || !mesh->sqrtG || !mesh->matrixG || !mesh->matrixIG || !mesh->matrixA ||
      !mesh->matrixIA || !mesh->f || !mesh->phi_s || !mesh->ghost_coef || !mesh->ghost_iref)
    return fail(MCV_EINVAL, "mesh or one of its arrays is NULL");
  if (!st || !st->phi || !st->phiG || !st->u || !st->v || !st->contraU || !st->contraV)
    return fail(MCV_EINVAL, "initial state or one of its arrays is NULL");
  h->cfg = *cfg;
  const int H = cfg->dof - 1;
  h->H = H;
  h->npv = cfg->n * H + 1;
  h->nh = h->npv + 2 * H;
  const long long per = (long long)h->nh * h->nh;
  h->np = 6 * per;
  const long long np = h->np;

  // ghost tables
  for (long long k = 0; k < (long long)H * h->npv; ++k) {
    const int c = mesh->ghost_iref[k];
    if (c < 1 || c > cfg->n)
      return fail(MCV_EINVAL, "ghost_iref[%lld] = %d outside [1, %d]", k, c, cfg->n);
    if (!std::isfinite(mesh->ghost_coef[k]))
      return fail(MCV_EINVAL, "ghost_coef[%lld] is not finite", k);
  }
  // split IG into its three distinct components; require IG21 == IG12
  std::vector<double> ig11(np), ig12(np), ig22(np), p2sp(4 * np), sp2p(4 * np);
  for (long long g = 0; g < np; ++g) {
    const double *IG = mesh->matrixIG + 4 * g;
    if (!bits_equal(IG[1], IG[2]))
      return fail(MCV_EINVAL, "matrixIG is not symmetric at point %lld", g);
    ig11[g] = IG[0];
    ig12[g] = IG[2];
    ig22[g] = IG[3];
    // (A*IG) and (G*IA), formed first as projection_mod.F90:71,90 do
    const double *A = mesh->matrixA + 4 * g, *G = mesh->matrixG + 4 * g,
                 *IA = mesh->matrixIA + 4 * g;
    double *P = p2sp.data() + 4 * g, *S = sp2p.data() + 4 * g;
    P[0] = A[0] * IG[0] + A[2] * IG[1];
    P[1] = A[1] * IG[0] + A[3] * IG[1];
    P[2] = A[0] * IG[2] + A[2] * IG[3];
    P[3] = A[1] * IG[2] + A[3] * IG[3];
    S[0] = G[0] * IA[0] + G[2] * IA[1];
    S[1] = G[1] * IA[0] + G[3] * IA[1];
    S[2] = G[0] * IA[2] + G[2] * IA[3];
    S[3] = G[1] * IA[2] + G[3] * IA[3];
  }
  // phiG must be phi*sqrtG (the device keeps phi only)
  for (int p = 0; p < 6; ++p)
    for (int j = 0; j < h->nh; ++j)
      for (int i = 0; i < h->nh; ++i) {
        const bool corner = (i < H || i >= H + h->npv) && (j < H || j >= H + h->npv);
        if (corner) continue;
        const long long g = p * per + (long long)j * h->nh + i;
        if (!bits_equal(st->phiG[g], st->phi[g] * mesh->sqrtG[g]))
          return fail(MCV_EINVAL, "initial phiG != phi*sqrtG at point %lld", g);
      }

  HIP_TRY(hipSetDevice(cfg->device));
  if ((rc = scratch_setup(cfg->device))) return rc;
  HIP_TRY(hipStreamCreate(&h->stream));
  for (hipEvent_t &e : h->ev) HIP_TRY(hipEventCreate(&e));

  MeshDev &m = h->mesh;
  m.dof = cfg->dof;
  m.n = cfg->n;
  m.npv = h->npv;
  m.nh = h->nh;
  m.dx = cfg->dx;
  m.dv.dx = make_recip(cfg->dx);
  m.dv.dx2 = make_recip(2.0 * cfg->dx);
  m.dv.dxx = make_recip(cfg->dx * cfg->dx);
  std::vector<double> rsg(np);
  for (long long g = 0; g < np; ++g) rsg[g] = 1.0 / mesh->sqrtG[g];
  const bool metric_inv = panel_invariant(mesh->sqrtG, per) && panel_invariant(ig11.data(), per) &&
                          panel_invariant(ig12.data(), per) && panel_invariant(ig22.data(), per);
  const long long mcount = metric_inv ? per : np;
  m.metric_pstride = metric_inv ? 0 : per;
  const bool phis_inv = panel_invariant(mesh->phi_s, per);
  m.phis_pstride = phis_inv ? 0 : per;
  double *d;
  if ((rc = upload(h, &d, mesh->sqrtG, mcount))) return rc;
  m.sqrtG = d;
  if ((rc = upload(h, &d, rsg.data(), mcount))) return rc;
  m.rsqrtG = d;
  if ((rc = upload(h, &d, ig11.data(), mcount))) return rc;
  m.ig11 = d;
  if ((rc = upload(h, &d, ig12.data(), mcount))) return rc;
  m.ig12 = d;
  if ((rc = upload(h, &d, ig22.data(), mcount))) return rc;
  m.ig22 = d;
  {
    std::vector<double> mp(4 * mcount), sp(2 * mcount);
    for (long long g = 0; g < mcount; ++g) {
      mp[4 * g] = mesh->sqrtG[g];
      mp[4 * g + 1] = ig11[g];
      mp[4 * g + 2] = ig12[g];
      mp[4 * g + 3] = ig22[g];
      sp[2 * g] = mesh->sqrtG[g];
      sp[2 * g + 1] = rsg[g];
    }
    if ((rc = upload(h, &d, mp.data(), mp.size()))) return rc;
    m.mpack = reinterpret_cast<const D2 *>(d);
    if ((rc = upload(h, &d, sp.data(), sp.size()))) return rc;
    m.spack = reinterpret_cast<const D2 *>(d);
  }
  m.phis_zero = 1;
  for (long long g = 0; g < np && m.phis_zero; ++g) {
    uint64_t bits;
    std::memcpy(&bits, &mesh->phi_s[g], sizeof bits);
    if (bits != 0) m.phis_zero = 0;
  }
  if ((rc = upload(h, &d, mesh->f, np))) return rc;
  m.f = d;
  if ((rc = upload(h, &d, mesh->phi_s, phis_inv ? per : np))) return rc;
  m.phi_s = d;
  if ((rc = upload(h, &d, p2sp.data(), 4 * np))) return rc;
  m.p2sp = d;
  {
    // line-major matrix tables of the boundary pass (band, ghost and edge points)
    const int npv = h->npv;
    std::vector<double> pb(4 * (size_t)6 * 4 * (H + 1) * npv), sg(4 * (size_t)6 * 4 * H * npv),
        se(4 * (size_t)6 * 4 * npv);
    for (int p = 1; p <= 6; ++p)
      for (int side = 0; side < 4; ++side)
        for (int al = 1; al <= npv; ++al) {
          for (int dd = 1; dd <= H + 1; ++dd) {
            int i, j;
            band_ij(npv, side, dd, al, i, j);
            std::memcpy(&pb[4 * (size_t)band_point(npv, H, p, side, dd, al)],
                        &p2sp[4 * pidx(h->nh, H, p, i, j)], 4 * sizeof(double));
            if (dd == 1)
              std::memcpy(&se[4 * (size_t)edge_point(npv, p, side, al)],
                          &sp2p[4 * pidx(h->nh, H, p, i, j)], 4 * sizeof(double));
          }
          for (int k = 1; k <= H; ++k) {
            int hi, hj;
            if (side == 0) { hi = 1 - k; hj = al; }
            else if (side == 1) { hi = npv + k; hj = al; }
            else if (side == 2) { hi = al; hj = 1 - k; }
            else { hi = al; hj = npv + k; }
            std::memcpy(&sg[4 * (size_t)ghost_point(npv, H, p, side, k, al)],
                        &sp2p[4 * pidx(h->nh, H, p, hi, hj)], 4 * sizeof(double));
          }
        }
    if ((rc = upload(h, &d, pb.data(), pb.size()))) return rc;
    m.p2sp_band = d;
    if ((rc = upload(h, &d, sg.data(), sg.size()))) return rc;
    m.sp2p_ghost = d;
    if ((rc = upload(h, &d, se.data(), se.size()))) return rc;
    m.sp2p_edge = d;
  }
  if ((rc = upload(h, &d, mesh->ghost_coef, (size_t)H * h->npv))) return rc;
  m.gcoef = d;
  {
    void *pi = nullptr;
    HIP_TRY(hipMalloc(&pi, sizeof(int32_t) * H * h->npv + 256));
    h->allocs.push_back(pi);
    h->device_bytes += sizeof(int32_t) * H * h->npv + 256;
    HIP_TRY(hipMemcpy(pi, mesh->ghost_iref, sizeof(int32_t) * H * h->npv,
                      hipMemcpyHostToDevice));
    m.giref = static_cast<const int32_t *>(pi);
  }
  // state levels (FP32C keeps only level 0 in FP64: the reference state)
  h->fp32c = cfg->precision == MCV_FP32C;
  const int nlev = h->fp32c ? 1 : 3;
  for (int l = 0; l < nlev; ++l) {
    double *pu;
    if ((rc = dev_alloc(h, &pu, 2 * np))) return rc;
    h->lpu[l] = reinterpret_cast<D2 *>(pu);
    if ((rc = dev_alloc(h, &h->lvv[l], np))) return rc;
  }
  if (!h->fp32c) {  // accumulator: {phiG, u} pairs and v
    if ((rc = dev_alloc(h, &h->acc[0], 2 * np))) return rc;
    if ((rc = dev_alloc(h, &h->acc[2], np))) return rc;
  }
  if ((rc = dev_alloc(h, &h->band_raw, band_values(cfg->dof, h->npv)))) return rc;
  {
    std::vector<D2> pu(np);
    for (long long g = 0; g < np; ++g) pu[g] = D2{st->phi[g], st->u[g]};
    HIP_TRY(hipMemcpy(h->lpu[0], pu.data(), np * sizeof(D2), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(h->lvv[0], st->v, np * sizeof(double), hipMemcpyHostToDevice));
  // halo corners of the scratch levels are never read, but keep them defined
  for (int l = 1; l < nlev; ++l) {
    HIP_TRY(hipMemcpy(h->lpu[l], h->lpu[0], np * sizeof(D2), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(h->lvv[l], h->lvv[0], np * sizeof(double), hipMemcpyDeviceToDevice));
  }
  if ((rc = upload(h, &h->cu0, st->contraU, np))) return rc;
  if ((rc = upload(h, &h->cv0, st->contraV, np))) return rc;
  h->first_pending = true;
  if (h->fp32c && (rc = setup32(h))) return rc;
  // integrals
  const long long ncell = 6LL * cfg->n * cfg->n;
  h->diag_blocks =
      (int)std::min<long long>((ncell + kDiagThreads - 1) / kDiagThreads, kDiagMaxBlocks);
  if ((rc = dev_alloc(h, &h->partials, 4 * (size_t)h->diag_blocks))) return rc;
  if ((rc = dev_alloc(h, &h->diag_out, 2))) return rc;
  for (int f = 0; f < 6; ++f)
    if ((rc = dev_alloc(h, &h->pack[f], np))) return rc;
  h->steps = 0;
  HIP_TRY(hipDeviceSynchronize());
  return MCV_OK;
}

// History fields on the device (lazily allocated scratch, kept by the handle).
int history_buffers(mcv_handle *h, bool points) {
  int rc;
  for (int f = 0; f < 8 && points; ++f)
    if (!h->hist[f] && (rc = dev_alloc(h, &h->hist[f], h->np))) return rc;
  const size_t nc = 6 * (size_t)(h->cfg.n + 2) * (h->cfg.n + 2);
  for (int f = 0; f < 3; ++f)
    if (!h->histc[f] && (rc = dev_alloc(h, &h->histc[f], nc))) return rc;
  return MCV_OK;
}

int launch_cell_averages(mcv_handle *h, const double *a, const double *b, const double *c) {
  const long long nc = 6LL * (h->cfg.n + 2) * (h->cfg.n + 2);
  const dim3 grid((unsigned)std::min<long long>((nc + 255) / 256, 4096)), block(256);
  if (h->cfg.dof == 3)
    hipLaunchKernelGGL(cell_average_kernel<3>, grid, block, 0, h->stream, h->mesh, h->cfg.n, a, b,
                       c, h->histc[0], h->histc[1], h->histc[2]);
  else
    hipLaunchKernelGGL(cell_average_kernel<4>, grid, block, 0, h->stream, h->mesh, h->cfg.n, a, b,
                       c, h->histc[0], h->histc[1], h->histc[2]);
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

// The output block of MCV_SW.F90:62-70 on the current state: hist[0..7] and
// histc[0..2] hold the record's fields.
int history_fields_dev(mcv_handle *h) {
  if (h->steps == 0)
    return fail(MCV_ESTATE,
                "no time step has run: the first history record is written from the host's "
                "initial arrays (MCV_SW.F90:50)");
  int rc = history_buffers(h, true);
  if (rc) return rc;
  const double *y[3];
  if ((rc = materialize(h, y))) return rc;
  hipLaunchKernelGGL(history_point_kernel, dim3(2048), dim3(256), 0, h->stream, h->mesh, y[0],
                     y[1], y[2], h->hist[0], h->hist[1], h->hist[2], h->hist[3], h->hist[4],
                     h->hist[5], h->hist[6], h->hist[7]);
  HIP_TRY(hipGetLastError());
  return launch_cell_averages(h, h->hist[0], h->hist[1], h->hist[4]);
}

int download_history(mcv_handle *h, const mcv_history_fields *out) {
  double *dst[11] = {out->u, out->v, out->contraU, out->contraV, out->phi, out->phi_t,
                     out->zonal, out->meridional, out->uC, out->vC, out->phiC};
  const size_t nc = 6 * (size_t)(h->cfg.n + 2) * (h->cfg.n + 2);
  for (int f = 0; f < 11; ++f)
    if (dst[f])
      HIP_TRY(hipMemcpyAsync(dst[f], f < 8 ? h->hist[f] : h->histc[f - 8],
                             (f < 8 ? h->np : nc) * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MCV_OK;
}

}  // namespace

extern "C" {

int mcv_hip_init(const mcv_config *cfg, const mcv_mesh *mesh, const mcv_state *initial,
                 mcv_handle **out) {
  g_last_error.clear();
  if (!out) return fail(MCV_EINVAL, "out is NULL");
  *out = nullptr;
  mcv_handle *h = new (std::nothrow) mcv_handle();
  if (!h) return fail(MCV_ENOMEM, "host allocation failed");
  int rc = init_impl(cfg, mesh, initial, h);
  if (rc) {
    std::string msg = g_last_error;
    destroy(h);
    g_last_error = msg;
    return rc;
  }
  *out = h;
  return MCV_OK;
}

int mcv_hip_step(mcv_handle *h, int32_t nsteps) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (nsteps < 0) return fail(MCV_EINVAL, "nsteps must be >= 0, got %d", nsteps);
  HIP_TRY(hipSetDevice(h->cfg.device));
  for (int32_t s = 0; s < nsteps; ++s) {
    int rc = one_step(h);
    if (rc) return rc;
    ++h->steps;
  }
  return MCV_OK;
}

int mcv_hip_sync(mcv_handle *h, mcv_state *out, double *mass, double *energy) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  HIP_TRY(hipSetDevice(h->cfg.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (out) {
    const double *cu0 = h->first_pending ? h->cu0 : nullptr;
    const double *cv0 = h->first_pending ? h->cv0 : nullptr;
    const double *y[3];
    int rc = materialize(h, y);
    if (rc) return rc;
    hipLaunchKernelGGL(pack_kernel, dim3(1024), dim3(256), 0, h->stream, h->mesh, y[0], y[1],
                       y[2], cu0, cv0, h->pack[0], h->pack[1], h->pack[2], h->pack[3],
                       h->pack[4], h->pack[5]);
    HIP_TRY(hipGetLastError());
    double *dst[6] = {out->phi, out->phiG, out->u, out->v, out->contraU, out->contraV};
    for (int f = 0; f < 6; ++f)
      if (dst[f])
        HIP_TRY(hipMemcpyAsync(dst[f], h->pack[f], h->np * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  if (mass || energy) return run_diag(h, mass, energy);
  return MCV_OK;
}

int mcv_hip_sphere_winds(mcv_handle *h, double *zonal, double *meridional) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!zonal || !meridional) return fail(MCV_EINVAL, "output array is NULL");
  HIP_TRY(hipSetDevice(h->cfg.device));
  const double *y[3];
  int rc = materialize(h, y);
  if (rc) return rc;
  hipLaunchKernelGGL(sphere_winds_kernel, dim3(1024), dim3(256), 0, h->stream, h->mesh, y[1],
                     y[2], h->pack[0], h->pack[1]);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(zonal, h->pack[0], h->np * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipMemcpyAsync(meridional, h->pack[1], h->np * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MCV_OK;
}

int mcv_hip_cell_averages(mcv_handle *h, double *uC, double *vC, double *phiC) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!uC || !vC || !phiC) return fail(MCV_EINVAL, "output array is NULL");
  HIP_TRY(hipSetDevice(h->cfg.device));
  int rc = history_buffers(h, false);
  if (rc) return rc;
  const double *y[3];
  if ((rc = materialize(h, y))) return rc;
  if ((rc = launch_cell_averages(h, y[1], y[2], y[0]))) return rc;
  mcv_history_fields f{};
  f.uC = uC;
  f.vC = vC;
  f.phiC = phiC;
  return download_history(h, &f);
}

int mcv_hip_history_fields(mcv_handle *h, const mcv_history_fields *out) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!out) return fail(MCV_EINVAL, "out is NULL");
  HIP_TRY(hipSetDevice(h->cfg.device));
  int rc = history_fields_dev(h);
  if (rc) return rc;
  return download_history(h, out);
}

int mcv_hip_history_write(mcv_handle *h, const char *path, int32_t time_slot) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!path) return fail(MCV_EINVAL, "path is NULL");
  HIP_TRY(hipSetDevice(h->cfg.device));
  int rc = history_fields_dev(h);
  if (rc) return rc;
  const size_t nc = 6 * (size_t)(h->cfg.n + 2) * (h->cfg.n + 2);
  std::vector<double> buf;
  try {
    buf.resize(8 * (size_t)h->np + 3 * nc);
  } catch (const std::bad_alloc &) {
    return fail(MCV_ENOMEM, "host staging for the history record failed");
  }
  mcv_history_fields f{};
  double **slot[11] = {&f.u, &f.v, &f.contraU, &f.contraV, &f.phi, &f.phi_t,
                       &f.zonal, &f.meridional, &f.uC, &f.vC, &f.phiC};
  size_t off = 0;
  for (int k = 0; k < 11; ++k) {
    *slot[k] = buf.data() + off;
    off += k < 8 ? (size_t)h->np : nc;
  }
  if ((rc = download_history(h, &f))) return rc;
  return history_write_checked(path, time_slot, &f, h->np, (long long)nc);
}

int mcv_hip_tendency(mcv_handle *h, const mcv_state *in, double *t_phiG, double *t_u,
                     double *t_v) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!in || !in->phi || !in->u || !in->v || !in->contraU || !in->contraV || !t_phiG || !t_u ||
      !t_v)
    return fail(MCV_EINVAL, "state or output array is NULL");
  if (h->fp32c)
    return fail(MCV_EINVAL, "the tendency seam evaluates the FP64 operator; this handle is FP32C");
  HIP_TRY(hipSetDevice(h->cfg.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const long long np = h->np, nt = 6LL * h->npv * h->npv;
  const size_t b = np * sizeof(double);
  // scratch: the input as a level ({phi, u} pairs in a temporary, v and the
  // contravariant winds in the pack buffers), the tendencies in temporaries
  std::vector<D2> pu(np);
  for (long long g = 0; g < np; ++g) pu[g] = D2{in->phi[g], in->u[g]};
  double *tin, *tpu, *tv;
  int rc;
  if ((rc = dev_alloc(h, &tin, 2 * np)) || (rc = dev_alloc(h, &tpu, 2 * nt)) ||
      (rc = dev_alloc(h, &tv, nt)))
    return rc;
  HIP_TRY(hipMemcpy(tin, pu.data(), np * sizeof(D2), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->pack[3], in->v, b, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->pack[4], in->contraU, b, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->pack[5], in->contraV, b, hipMemcpyHostToDevice));
  StageArgs a{};
  a.in_pu = reinterpret_cast<const D2 *>(tin); a.in_v = h->pack[3];
  a.cu0 = h->pack[4]; a.cv0 = h->pack[5];
  a.o_pu = reinterpret_cast<D2 *>(tpu); a.o_v = tv;
  if ((rc = stage(h, kTend, a))) return rc;
  std::vector<D2> tk(nt);
  HIP_TRY(hipMemcpyAsync(tk.data(), tpu, nt * sizeof(D2), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipMemcpyAsync(t_v, tv, nt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (long long g = 0; g < nt; ++g) {
    t_phiG[g] = tk[g].a;
    t_u[g] = tk[g].b;
  }
  const long long sizes[3] = {nt, 2 * nt, 2 * np};  // freed in reverse order of allocation
  for (long long c : sizes) {
    (void)hipFree(h->allocs.back());
    h->allocs.pop_back();
    h->device_bytes -= (size_t)c * sizeof(double) + 256;
  }
  return MCV_OK;
}

int mcv_hip_tendency_noalloc(mcv_handle *h, const mcv_state *in, double *t_phiG, double *t_u,
                             double *t_v) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (!in || !in->phi || !in->u || !in->v || !in->contraU || !in->contraV || !t_phiG || !t_u ||
      !t_v)
    return fail(MCV_EINVAL, "state or output array is NULL");
  if (h->fp32c)
    return fail(MCV_EINVAL, "the tendency seam evaluates the FP64 operator; this handle is FP32C");
  const long long np = h->np, nt = 6LL * h->npv * h->npv;
  // scratch: the input as {phi, u} pairs, the {phiG, u} tendency pairs and
  // the v tendency, each 256-byte aligned
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_in = 0, o_pu = up(np * sizeof(D2)), o_v = o_pu + up(nt * sizeof(D2));
  const size_t need = o_v + nt * sizeof(double);
  if (need > kScratchBytes)
    return fail(MCV_EINVAL,
                "mcv_hip_tendency_noalloc: this mesh needs %zu bytes of scratch, the library "
                "holds %zu (at most about %lld points per field); use mcv_hip_tendency",
                need, kScratchBytes, (long long)(kScratchBytes / 40));
  HIP_TRY(hipSetDevice(h->cfg.device));
  char *const sp = scratch_of(h->cfg.device);
  if (!sp) return fail(MCV_ESTATE, "no scratch address for device %d", h->cfg.device);
  HIP_TRY(hipStreamSynchronize(h->stream));
  // the input through the handle's own read-back buffers (allocated at init)
  const size_t b = np * sizeof(double);
  HIP_TRY(hipMemcpyAsync(h->pack[0], in->phi, b, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->pack[2], in->u, b, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->pack[3], in->v, b, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->pack[4], in->contraU, b, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->pack[5], in->contraV, b, hipMemcpyHostToDevice, h->stream));
  D2 *const s_in = reinterpret_cast<D2 *>(sp + o_in);
  D2 *const s_pu = reinterpret_cast<D2 *>(sp + o_pu);
  double *const s_v = reinterpret_cast<double *>(sp + o_v);
  hipLaunchKernelGGL(pack_pairs_kernel, dim3(1024), dim3(256), 0, h->stream, np, h->pack[0],
                     h->pack[2], s_in);
  HIP_TRY(hipGetLastError());
  StageArgs a{};
  a.in_pu = s_in; a.in_v = h->pack[3];
  a.cu0 = h->pack[4]; a.cv0 = h->pack[5];
  a.o_pu = s_pu; a.o_v = s_v;
  int rc = stage(h, kTend, a);
  if (rc) return rc;
  hipLaunchKernelGGL(unpack_pairs_kernel, dim3(1024), dim3(256), 0, h->stream, nt, s_pu,
                     h->pack[0], h->pack[1]);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(t_phiG, h->pack[0], nt * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipMemcpyAsync(t_u, h->pack[1], nt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipMemcpyAsync(t_v, s_v, nt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MCV_OK;
}

int mcv_hip_step_timed(mcv_handle *h, int32_t nsteps, double *stage_ms, double *boundary_ms) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (nsteps < 0) return fail(MCV_EINVAL, "nsteps must be >= 0, got %d", nsteps);
  HIP_TRY(hipSetDevice(h->cfg.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->timing = true;
  h->stage_ms = h->boundary_ms = 0.0;
  int rc = MCV_OK;
  for (int32_t s = 0; s < nsteps && rc == MCV_OK; ++s) {
    rc = one_step(h);
    if (rc == MCV_OK) ++h->steps;
  }
  h->timing = false;
  if (rc) return rc;
  if (stage_ms) *stage_ms = h->stage_ms;
  if (boundary_ms) *boundary_ms = h->boundary_ms;
  return MCV_OK;
}

int mcv_hip_autotune(mcv_handle *h, int32_t reps, int32_t *tile, double *ms_per_stage) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (reps < 1) return fail(MCV_EINVAL, "reps must be >= 1");
  HIP_TRY(hipSetDevice(h->cfg.device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  // A representative stage (reads Yn and the accumulator, writes S1 and the
  // accumulator): S1, the accumulator and the band are rewritten by
  // the first stage of every step, so timing never changes the trajectory.
  StageArgs a{};
  a.in_pu = h->lpu[0]; a.in_v = h->lvv[0];
  a.b_pu = h->lpu[0]; a.b_v = h->lvv[0];
  a.acc_pu = reinterpret_cast<D2 *>(h->acc[0]); a.acc_v = h->acc[2];
  a.o_pu = h->lpu[1]; a.o_v = h->lvv[1];
  a.band_raw = h->band_raw;
  if (h->first_pending) { a.cu0 = h->cu0; a.cv0 = h->cv0; }
  a.c1 = 0.5 * h->cfg.dt;
  const int kind = h->cfg.scheme == MCV_RK4 ? kRK4S2 : kRK3S2;
  const StageArgs32 a32 = h->fp32c ? stage_args32(h, 0, 1, 0.5 * h->cfg.dt) : StageArgs32{};
  auto run = [&]() { return h->fp32c ? stage32(h, kind, a32) : stage(h, kind, a); };
  const int keep = h->tile;
  int best = keep;
  double best_ms = 1e300;
  for (int v = 0; v < kTileVariants; ++v) {
    h->tile = v;
    int rc = run();  // warm-up (code object load, caches)
    if (rc) { h->tile = keep; return rc; }
    HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    for (int r = 0; r < reps; ++r)
      if ((rc = run())) { h->tile = keep; return rc; }
    HIP_TRY(hipEventRecord(h->ev[1], h->stream));
    HIP_TRY(hipEventSynchronize(h->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    if (ms_per_stage) ms_per_stage[v] = ms / reps;
    if (ms < best_ms) { best_ms = ms; best = v; }
  }
  h->tile = best;
  if (tile) *tile = best;
  return MCV_OK;
}

int mcv_hip_set_tile(mcv_handle *h, int32_t tile) {
  if (!h) return fail(MCV_EINVAL, "handle is NULL");
  if (tile < 0 || tile >= kTileVariants)
    return fail(MCV_EINVAL, "tile must be in [0, %d), got %d", kTileVariants, tile);
  h->tile = tile;
  return MCV_OK;
}

int mcv_hip_steps_done(const mcv_handle *h, int64_t *steps) {
  if (!h || !steps) return fail(MCV_EINVAL, "NULL argument");
  *steps = h->steps;
  return MCV_OK;
}

int mcv_hip_free(mcv_handle *h) {
  destroy(h);
  return MCV_OK;
}

int mcv_hip_footprint(const mcv_config *cfg, int64_t *bytes) {
  if (!bytes) return fail(MCV_EINVAL, "bytes is NULL");
  int rc = validate(cfg);
  if (rc) return rc;
  *bytes = (int64_t)footprint_bytes(cfg);
  return MCV_OK;
}

int mcv_hip_device_bytes(const mcv_handle *h, int64_t *bytes) {
  if (!h || !bytes) return fail(MCV_EINVAL, "NULL argument");
  *bytes = (int64_t)h->device_bytes;
  return MCV_OK;
}

const char *mcv_hip_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
