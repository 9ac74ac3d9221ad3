// Internal: the per-solve handle and helpers shared by the single-solve and
// batched entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mcv_hip.h"
#include "mcv_device.hpp"
#include "mcv_fp32c.hpp"

namespace mcv {

// Records the message returned by mcv_hip_last_error() and returns code.
int fail(int code, const char *fmt, ...);
// mcv_history_write that also checks the file's per-record sizes (mcv_netcdf.cpp)
int history_write_checked(const char *path, int32_t time_slot, const mcv_history_fields *rec,
                          long long npoint, long long ncell);

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::mcv::fail(MCV_EDEVICE, "%s failed: %s (%s:%d)", #expr,                   \
                         hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)

// Tile shapes (cells x cells per workgroup) the autotuner chooses from, FP64
// and FP32C.  Phase 2 of the stage kernel visits H*H*CX*CY interior slots plus
// H*(CX+CY)+1 far-end points; 6x4 (MCV4: 247) and 8x7 (MCV3: 255) fill 256
// threads in one pass, the larger shapes re-read less halo per cell.
constexpr int kTileVariants = 4;
template <int DOF> struct TileSet;
template <> struct TileSet<3> {
  static constexpr int cx[kTileVariants] = {8, 16, 10, 12};
  static constexpr int cy[kTileVariants] = {7, 4, 5, 6};
};
template <> struct TileSet<4> {
  static constexpr int cx[kTileVariants] = {6, 8, 5, 6};
  static constexpr int cy[kTileVariants] = {4, 4, 5, 6};
};

template <int DOF> struct TileSet32;
template <> struct TileSet32<3> {
  static constexpr int cx[kTileVariants] = {8, 16, 10, 16};
  static constexpr int cy[kTileVariants] = {7, 8, 5, 4};
};
template <> struct TileSet32<4> {
  static constexpr int cx[kTileVariants] = {6, 8, 8, 10};
  static constexpr int cy[kTileVariants] = {4, 4, 8, 6};
};

// Upper bound of the device bytes one solve of this configuration allocates
// (exact when the metric terms differ between panels).
size_t footprint_bytes(const mcv_config *c);

// Stage and boundary arguments of a solve's levels (level 0 is Y_n).
StageArgs stage_args64(const mcv_handle *h, int in, int out, double c1);
StageArgs32 stage_args32(const mcv_handle *h, int in, int out, double c1);
BoundaryArgs boundary_args64(const mcv_handle *h, int level);
BoundaryArgs32 boundary_args32(const mcv_handle *h, int level);

}  // namespace mcv

// One solve on one device.
struct mcv_handle {
  mcv_config cfg{};
  int H = 0, npv = 0, nh = 0;
  long long np = 0;  // points per field
  hipStream_t stream = nullptr;
  mcv::MeshDev mesh{};
  std::vector<void *> allocs;  // every device allocation, freed together
  size_t device_bytes = 0;     // their sizes summed
  mcv::D2 *lpu[3] = {};        // levels (Yn, S1, S2): {phi, u} pairs
  double *lvv[3] = {};         //   and v
  double *phi_ref = nullptr;   // FP32C: the projected initial geopotential (FP64)
  double *acc[3] = {};         // RK4 accumulator: acc[0] {phiG, u} pairs, acc[2] v
  double *band_raw = nullptr;  // raw band lines written by the stage kernel
  double *cu0 = nullptr, *cv0 = nullptr;  // host contra winds until the first stage
  bool first_pending = true;
  double *partials = nullptr;
  int diag_blocks = 0;
  double *diag_out = nullptr;
  double *pack[6] = {};
  // history fields (allocated by the first mcv_hip_history_fields call):
  // u, v, contraU, contraV, phi, phi_t, zonal, meridional; uC, vC, phiC
  double *hist[8] = {};
  double *histc[3] = {};
  int64_t steps = 0;
  int tile = 0;  // index into the tile shapes of the solve's order
  // per-kernel timing (mcv_hip_step_timed)
  bool timing = false;
  hipEvent_t ev[3] = {};
  double stage_ms = 0.0, boundary_ms = 0.0;
  // FP32C form (precision == MCV_FP32C): FP32 levels of (phi - phi_ref, u, v),
  // compensation words of Y_n, FP32 accumulator and band copies; phi_ref is
  // the projected initial geopotential (FP64).
  bool fp32c = false;
  // levels, compensation words of Y_n and accumulator: {phi', u, v} records
  mcv::F3 *l32[3] = {};
  mcv::F3 *c32 = nullptr;
  mcv::F3 *a32 = nullptr;
  mcv::F3 *band32 = nullptr, *band32c = nullptr;  // FP32 band records and residues
  double *join[3] = {};  // FP64 reconstruction of Y_n for read-back and integrals
  mcv::MeshF32 mesh32{};
};
