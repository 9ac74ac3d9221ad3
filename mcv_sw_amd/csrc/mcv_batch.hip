// Batched and continuously batched MCV solves (include/mcv_hip.h, mcv_batch_*).
//
// A batch holds up to `capacity` solves (slots) of any order (MCV3/MCV4),
// resolution, precision (FP64/FP32C), RK scheme (RK3_TVD/RK4) and time step.
// Each stage (three launches) advances every listed solve by ONE RK stage:
//
//   batch_stage_kernel     one workgroup per item of the concatenated
//                          (solve, panel, tile) work list of all listed
//                          solves, each at its own stage of its own scheme
//                          (the stage index lives on the device, so solves
//                          admitted at different times interleave freely)
//   batch_boundary_kernel  one workgroup per unit of 256 boundary items of
//                          every solve's freshly written level
//   batch_advance_kernel   one workgroup; advances every solve's stage index
//                          and step counters
//
// All per-solve data the kernels need sits in device tables (solve
// descriptors, stage state, work prefix sums), so one captured hipGraph of
// kGraphStages stages serves every batch composition whose work list fits its
// stage grid: admitting or retiring a solve rewrites the tables between
// replays, and the graph is re-captured only when the list outgrows the grid
// or the set of solve kinds changes.
// Solves whose step budget is used up are skipped by the kernels until the
// host retires them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mcv_hip.h"
#include "mcv_handle.hpp"
#include "mcv_kernels_fp64.hpp"

namespace mcv {

constexpr int kMaxStages = 4;
constexpr int kBndChunk = 256;   // boundary items per work unit
constexpr int kGraphStages = 12; // stages per graph (a multiple of 3 and 4)
constexpr int kXcds = 8;

// Static description of one solve, written once at admission.
struct BatchSolve {
  int dof, prec, nstages, nbitems;
  int kind[kMaxStages];   // stage kinds of its scheme, in order
  MeshDev mesh;
  MeshF32 mesh32;
  StageArgs a64[kMaxStages];
  StageArgs32 a32[kMaxStages];
  BoundaryArgs b64[kMaxStages];
  BoundaryArgs32 b32[kMaxStages];
  const double *cu0, *cv0;  // handed-over contravariant winds (FP64, first stage)
};

// Mutable progress of one solve, advanced on the device.
struct BatchProgress {
  int phase;           // next stage index
  int first;           // no stage has run yet
  long long steps_left;
  long long steps_done;
};

// Bodies of the batched stage kernel: (precision, order) pairs, and the RK
// schemes its solves use.  A kernel instantiation compiles the bodies of its
// mask only, so that a batch of one kind of solve runs with that body's
// registers and LDS; with one scheme in the mask the compiler also drops the
// other scheme's stage kinds (the RK4 accumulator paths of an RK3 batch).
enum : int {
  kBody64x3 = 1, kBody64x4 = 2, kBody32x3 = 4, kBody32x4 = 8,
  kSchemeRK4 = 16, kSchemeRK3 = 32, kSchemes = kSchemeRK4 | kSchemeRK3
};
// body sets (one body, then the FP64 pair, the FP32C pair, all four), each
// with RK3 alone, RK4 alone and both
constexpr int kBodyMasks[] = {
    1 | kSchemeRK3,  1 | kSchemeRK4,  1 | kSchemes,  2 | kSchemeRK3,  2 | kSchemeRK4,
    2 | kSchemes,    4 | kSchemeRK3,  4 | kSchemeRK4,  4 | kSchemes,  8 | kSchemeRK3,
    8 | kSchemeRK4,  8 | kSchemes,    3 | kSchemeRK3,  3 | kSchemeRK4,  3 | kSchemes,
    12 | kSchemeRK3, 12 | kSchemeRK4, 12 | kSchemes,   15 | kSchemeRK3, 15 | kSchemeRK4,
    15 | kSchemes};
constexpr int kNumMasks = sizeof(kBodyMasks) / sizeof(kBodyMasks[0]);

inline int body_of(int dof, int prec, int nstages) {
  const int body = prec == MCV_FP64 ? (dof == 3 ? kBody64x3 : kBody64x4)
                                    : (dof == 3 ? kBody32x3 : kBody32x4);
  return body | (nstages == 4 ? kSchemeRK4 : kSchemeRK3);
}

// The first mask (fewest bodies, then fewest schemes) that holds every bit of
// `need`.
inline int mask_index(int need) {
  for (int i = 0; i < kNumMasks; ++i)
    if ((kBodyMasks[i] & need) == need) return i;
  return kNumMasks - 1;
}

// LDS of the batched stage kernel for tile variant V and body mask M: the
// largest of its bodies.
template <int V, int M>
struct BatchSmem {
  static constexpr size_t a =
      (M & kBody64x3) ? TileLayout<double, 3, TileSet<3>::cx[V], TileSet<3>::cy[V]>::kBytes : 0;
  static constexpr size_t b =
      (M & kBody64x4) ? TileLayout<double, 4, TileSet<4>::cx[V], TileSet<4>::cy[V]>::kBytes : 0;
  static constexpr size_t c =
      (M & kBody32x3)
          ? TileLayout<FP32CStencil, 3, TileSet32<3>::cx[V], TileSet32<3>::cy[V]>::kBytes
          : 0;
  static constexpr size_t d =
      (M & kBody32x4)
          ? TileLayout<FP32CStencil, 4, TileSet32<4>::cx[V], TileSet32<4>::cy[V]>::kBytes
          : 0;
  static constexpr size_t ab = a > b ? a : b, cd = c > d ? c : d;
  static constexpr size_t kBytes = ab > cd ? ab : cd;
};

// Tiles per panel edge of a solve for tile variant V.
template <int V>
__host__ __device__ inline int tiles_x_of(int dof, int prec, int n) {
  int cx;
  if (prec == MCV_FP64) cx = dof == 3 ? TileSet<3>::cx[V] : TileSet<4>::cx[V];
  else cx = dof == 3 ? TileSet32<3>::cx[V] : TileSet32<4>::cx[V];
  return (n + cx - 1) / cx;
}
template <int V>
__host__ __device__ inline int tiles_y_of(int dof, int prec, int n) {
  int cy;
  if (prec == MCV_FP64) cy = dof == 3 ? TileSet<3>::cy[V] : TileSet<4>::cy[V];
  else cy = dof == 3 ? TileSet32<3>::cy[V] : TileSet32<4>::cy[V];
  return (n + cy - 1) / cy;
}

// First list entry k whose range [prefix[k], prefix[k+1]) holds w.
__device__ __forceinline__ int find_entry(const int *prefix, int cnt, int w) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Body of the stage kernel.  Work item w of the concatenated list is (solve,
// panel, tile); the work is cut into one contiguous chunk per XCD (blocks are
// dealt round-robin to the 8 XCDs), so the six panel tiles that share metric
// data, and neighbouring tiles, go through the same L2.  The host launches at
// least 8 * ceil(items / 8) workgroups, so each takes at most one item and the
// hardware's dynamic dispatch balances the tiles; a smaller grid would walk
// the chunk in strides.
template <int V, int M>
__device__ __forceinline__ void batch_stage_items(
    const BatchSolve *__restrict__ solves, const BatchProgress *__restrict__ prog,
    const int *__restrict__ list, const int *__restrict__ wprefix, const int *__restrict__ meta,
    char *smem) {
  const int cnt = meta[0];
  if (cnt == 0) return;
  const int total = wprefix[cnt];
  const int nslot = gridDim.x / kXcds;
  const int xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
  if (slot >= nslot) return;
  const int lo = (int)((long long)total * xcd / kXcds);
  const int hi = (int)((long long)total * (xcd + 1) / kXcds);
  int k = -1, kend = 0;
  for (int w = lo + slot; w < hi; w += nslot) {
    if (w >= kend) {
      k = find_entry(wprefix, cnt, w);
      kend = wprefix[k + 1];
    }
    const int sl = list[k];
    const BatchProgress st = prog[sl];
    if (st.steps_left <= 0) continue;
    const BatchSolve &S = solves[sl];
    const int local = w - wprefix[k];
    const int p = local % 6 + 1, tile = local / 6;
    const int ph = st.phase, kind = S.kind[ph];
    // the host launches a mask holding the scheme of every solve in the batch
    if constexpr ((M & kSchemeRK4) == 0) __builtin_assume(kind >= kRK3S1 && kind <= kRK3S3);
    if constexpr ((M & kSchemeRK3) == 0) __builtin_assume(kind >= kRK4S1 && kind <= kRK4S4);
    const int n = S.mesh.n;
    // the host launches a mask holding every body in the batch
    if (S.prec == MCV_FP64) {
      if constexpr ((M & (kBody64x3 | kBody64x4)) != 0) {
        StageArgs a = S.a64[ph];
        if (st.first) { a.cu0 = S.cu0; a.cv0 = S.cv0; }
        const MeshDev &m = S.mesh;  // read where used (scalar loads), not copied
        if (S.dof == 3) {
          if constexpr ((M & kBody64x3) != 0) {
            const int tx = tiles_x_of<V>(3, MCV_FP64, n);
            stage_tile<3, TileSet<3>::cx[V], TileSet<3>::cy[V]>(m, a, kind, p, tile % tx,
                                                                 tile / tx, smem);
          }
        } else {
          if constexpr ((M & kBody64x4) != 0) {
            const int tx = tiles_x_of<V>(4, MCV_FP64, n);
            stage_tile<4, TileSet<4>::cx[V], TileSet<4>::cy[V]>(m, a, kind, p, tile % tx,
                                                                 tile / tx, smem);
          }
        }
      }
    } else {
      if constexpr ((M & (kBody32x3 | kBody32x4)) != 0) {
        const StageArgs32 &a = S.a32[ph];
        const MeshF32 &m = S.mesh32;
        if (S.dof == 3) {
          if constexpr ((M & kBody32x3) != 0) {
            const int tx = tiles_x_of<V>(3, MCV_FP32C, n);
            stage_tile_f32<3, TileSet32<3>::cx[V], TileSet32<3>::cy[V]>(m, a, kind, p, tile % tx,
                                                                        tile / tx, smem);
          }
        } else {
          if constexpr ((M & kBody32x4) != 0) {
            const int tx = tiles_x_of<V>(4, MCV_FP32C, n);
            stage_tile_f32<4, TileSet32<4>::cx[V], TileSet32<4>::cy[V]>(m, a, kind, p, tile % tx,
                                                                        tile / tx, smem);
          }
        }
      }
    }
    __syncthreads();  // LDS is reused by the next work item
  }
}

// FP32C bodies with one RK scheme at the tile shapes whose LDS allows 8
// workgroups per CU (8x7 and 10x5 cells for MCV3, 6x4 for MCV4) fit in 64
// VGPRs; there the kernel is compiled for 8 waves per SIMD, the SGPRs beyond
// that budget spilling to VGPR lanes (no scratch memory).  Elsewhere the
// compiler's own choice stands.
constexpr bool eight_waves(int V, int M) {
  const int bodies = M & (kBody64x3 | kBody64x4 | kBody32x3 | kBody32x4);
  if ((bodies & (kBody64x3 | kBody64x4)) != 0 || (M & kSchemes) == kSchemes) return false;
  return V == 0 || (V == 2 && bodies == kBody32x3);
}

template <int V, int M>
__global__ __launch_bounds__(kStageThreads) void batch_stage_kernel(
    const BatchSolve *__restrict__ solves, const BatchProgress *__restrict__ prog,
    const int *__restrict__ list, const int *__restrict__ wprefix, const int *__restrict__ meta) {
  __shared__ __align__(16) char smem[BatchSmem<V, M>::kBytes];
  batch_stage_items<V, M>(solves, prog, list, wprefix, meta, smem);
}

template <int V, int M>
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
batch_stage_kernel_8w(const BatchSolve *__restrict__ solves, const BatchProgress *__restrict__ prog,
                      const int *__restrict__ list, const int *__restrict__ wprefix,
                      const int *__restrict__ meta) {
  __shared__ __align__(16) char smem[BatchSmem<V, M>::kBytes];
  batch_stage_items<V, M>(solves, prog, list, wprefix, meta, smem);
}

// Boundary kernel: the boundary pass of every listed solve's freshly written
// level, in work units of kBndChunk items.  The host launches at least one
// workgroup per unit; a smaller grid would walk the units in strides.
__global__ __launch_bounds__(256) void batch_boundary_kernel(
    const BatchSolve *__restrict__ solves, const BatchProgress *__restrict__ prog,
    const int *__restrict__ list, const int *__restrict__ bprefix, const int *__restrict__ meta) {
  const int cnt = meta[0];
  const int total = cnt ? bprefix[cnt] : 0;
  int k = -1, kend = 0;
  for (int w = blockIdx.x; w < total; w += gridDim.x) {
    if (w >= kend) {
      k = find_entry(bprefix, cnt, w);
      kend = bprefix[k + 1];
    }
    const int sl = list[k];
    const BatchProgress st = prog[sl];
    if (st.steps_left <= 0) continue;
    const BatchSolve &S = solves[sl];
    const int ph = st.phase;
    const int base = (w - bprefix[k]) * kBndChunk;
    const int lim = min(kBndChunk, S.nbitems - base);
    if (S.prec == MCV_FP64) {
      const BoundaryArgs b = S.b64[ph];
      for (int j = threadIdx.x; j < lim; j += blockDim.x) {
        if (S.dof == 3) boundary_item<3>(S.mesh, b, base + j);
        else boundary_item<4>(S.mesh, b, base + j);
      }
    } else {
      const BoundaryArgs32 b = S.b32[ph];
      for (int j = threadIdx.x; j < lim; j += blockDim.x) {
        if (S.dof == 3) boundary_item_f32<3>(S.mesh, b, base + j);
        else boundary_item_f32<4>(S.mesh, b, base + j);
      }
    }
  }
}

// Advances every listed solve by one stage (stage index, then step counters);
// one workgroup, after the boundary kernel of the stage.
__global__ __launch_bounds__(256) void batch_advance_kernel(const BatchSolve *__restrict__ solves,
                                                            BatchProgress *__restrict__ prog,
                                                            const int *__restrict__ list,
                                                            const int *__restrict__ meta) {
  const int cnt = meta[0];
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const int sl = list[e];
    BatchProgress st = prog[sl];
    if (st.steps_left <= 0) continue;
    st.first = 0;
    if (++st.phase == solves[sl].nstages) {
      st.phase = 0;
      --st.steps_left;
      ++st.steps_done;
    }
    prog[sl] = st;
  }
}

}  // namespace mcv

using namespace mcv;

struct mcv_batch {
  int device = 0;
  int capacity = 0;
  hipStream_t stream = nullptr;
  std::vector<mcv_handle *> slot;       // owned solves; nullptr = free
  std::vector<BatchSolve> solve;        // host copies of the descriptors
  std::vector<BatchProgress> progress;  // host view, refreshed by poll
  BatchSolve *d_solve = nullptr;
  BatchProgress *d_prog = nullptr;
  int *d_list = nullptr, *d_wprefix = nullptr, *d_bprefix = nullptr, *d_meta = nullptr;
  std::vector<void *> allocs;
  int variant = 0;
  bool tables_dirty = true;
  int stage_items = 0;  // length of the stage work list
  int stage_grid = 0;   // workgroups of the stage launch (>= 8 * ceil(items / 8))
  int bnd_units = 0;    // boundary work units
  int bnd_grid = 0;     // workgroups of the boundary launch (>= units)
  int mask = kNumMasks - 1;  // index into kBodyMasks of the launched stage kernel
  bool use_graph = true;
  hipGraphExec_t exec = nullptr;
  int exec_variant = -1, exec_mask = -1, exec_grid = 0, exec_bnd_grid = 0;
  hipEvent_t ev[3] = {};
  int64_t stages_run = 0;
};

namespace {

template <typename T>
int balloc(mcv_batch *b, T **out, size_t count) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
  if (e != hipSuccess)
    return fail(MCV_ENOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                hipGetErrorString(e));
  b->allocs.push_back(p);
  *out = static_cast<T *>(p);
  return MCV_OK;
}

void drop_graph(mcv_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->exec_variant = -1;
  b->exec_mask = -1;
  b->exec_grid = 0;
  b->exec_bnd_grid = 0;
}

void destroy_batch(mcv_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (mcv_handle *h : b->slot)
    if (h) mcv_hip_free(h);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

BatchSolve describe(const mcv_handle *h) {
  BatchSolve s{};
  s.dof = h->cfg.dof;
  s.prec = h->cfg.precision;
  s.nbitems = boundary_items(h->cfg.dof, h->npv);
  s.mesh = h->mesh;
  s.mesh32 = h->mesh32;
  const double dt = h->cfg.dt;
  struct St { int kind, in, out; double c1; };
  // time_scheme_mod.F90:13-69 with the level roles of the single-solve path
  const St rk4[4] = {{kRK4S1, 0, 1, 0.5 * dt}, {kRK4S2, 1, 2, 0.5 * dt}, {kRK4S3, 2, 1, dt},
                     {kRK4S4, 1, 0, dt}};
  const St rk3[3] = {{kRK3S1, 0, 1, dt}, {kRK3S2, 1, 2, 0.25 * dt},
                     {kRK3S3, 2, 0, 2.0 / 3.0 * dt}};
  const St *sched = h->cfg.scheme == MCV_RK4 ? rk4 : rk3;
  s.nstages = h->cfg.scheme == MCV_RK4 ? 4 : 3;
  for (int i = 0; i < s.nstages; ++i) {
    s.kind[i] = sched[i].kind;
    if (h->fp32c) {
      s.a32[i] = stage_args32(h, sched[i].in, sched[i].out, sched[i].c1);
      s.b32[i] = boundary_args32(h, sched[i].out);
    } else {
      s.a64[i] = stage_args64(h, sched[i].in, sched[i].out, sched[i].c1);
      s.b64[i] = boundary_args64(h, sched[i].out);
    }
  }
  s.cu0 = h->cu0;
  s.cv0 = h->cv0;
  return s;
}

template <int V>
int work_of(const BatchSolve &s) {
  return 6 * tiles_x_of<V>(s.dof, s.prec, s.mesh.n) * tiles_y_of<V>(s.dof, s.prec, s.mesh.n);
}

int stage_work(int v, const BatchSolve &s) {
  switch (v) {
    case 0: return work_of<0>(s);
    case 1: return work_of<1>(s);
    case 2: return work_of<2>(s);
    default: return work_of<3>(s);
  }
}

// Rebuilds and uploads the list of occupied slots with its work prefix sums.
int upload_tables(mcv_batch *b) {
  std::vector<int> list, wp(1, 0), bp(1, 0);
  int need = 0;
  for (int s = 0; s < b->capacity; ++s) {
    if (!b->slot[s]) continue;
    need |= body_of(b->solve[s].dof, b->solve[s].prec, b->solve[s].nstages);
    list.push_back(s);
    wp.push_back(wp.back() + stage_work(b->variant, b->solve[s]));
    bp.push_back(bp.back() + (b->solve[s].nbitems + kBndChunk - 1) / kBndChunk);
  }
  const int meta[1] = {(int)list.size()};
  b->stage_items = wp.back();
  b->bnd_units = bp.back();
  if (!list.empty()) {
    HIP_TRY(hipMemcpyAsync(b->d_list, list.data(), list.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_wprefix, wp.data(), wp.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_bprefix, bp.data(), bp.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
  }
  HIP_TRY(hipMemcpyAsync(b->d_meta, meta, sizeof(meta), hipMemcpyHostToDevice, b->stream));
  // the host vectors go out of scope: finish the copies first
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->mask = mask_index(need);
  // a graph keeps its grids while the work fits in them and fills at least
  // half of them
  auto keep = [&](int have, int fit) {
    return b->exec && have >= fit && have <= 2 * fit ? have : fit;
  };
  b->stage_grid = keep(b->exec_grid, std::max(kXcds, (b->stage_items + kXcds - 1) / kXcds * kXcds));
  b->bnd_grid = keep(b->exec_bnd_grid, std::max(1, b->bnd_units));
  b->tables_dirty = false;
  return MCV_OK;
}

template <int V, int I>
void launch_stage_vm(mcv_batch *b) {
  if constexpr (eight_waves(V, kBodyMasks[I])) {
    hipLaunchKernelGGL((batch_stage_kernel_8w<V, kBodyMasks[I]>), dim3(b->stage_grid),
                       dim3(kStageThreads), 0, b->stream, b->d_solve, b->d_prog, b->d_list,
                       b->d_wprefix, b->d_meta);
  } else {
    hipLaunchKernelGGL((batch_stage_kernel<V, kBodyMasks[I]>), dim3(b->stage_grid),
                       dim3(kStageThreads), 0, b->stream, b->d_solve, b->d_prog, b->d_list,
                       b->d_wprefix, b->d_meta);
  }
}

template <int V, size_t... I>
void launch_stage_masks(mcv_batch *b, std::index_sequence<I...>) {
  using Fn = void (*)(mcv_batch *);
  static constexpr Fn fns[] = {&launch_stage_vm<V, (int)I>...};
  fns[b->mask](b);
}

template <int V>
void launch_stage_v(mcv_batch *b) {
  launch_stage_masks<V>(b, std::make_index_sequence<kNumMasks>{});
}

int launch_stage(mcv_batch *b, int v) {
  switch (v) {
    case 0: launch_stage_v<0>(b); break;
    case 1: launch_stage_v<1>(b); break;
    case 2: launch_stage_v<2>(b); break;
    default: launch_stage_v<3>(b); break;
  }
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

int launch_boundary(mcv_batch *b) {
  hipLaunchKernelGGL(batch_boundary_kernel, dim3(b->bnd_grid), dim3(256), 0, b->stream,
                     b->d_solve, b->d_prog, b->d_list, b->d_bprefix, b->d_meta);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(256), 0, b->stream, b->d_solve,
                     b->d_prog, b->d_list, b->d_meta);
  HIP_TRY(hipGetLastError());
  return MCV_OK;
}

int launch_pair(mcv_batch *b) {
  int rc = launch_stage(b, b->variant);
  return rc ? rc : launch_boundary(b);
}

int capture(mcv_batch *b) {
  drop_graph(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = MCV_OK;
  for (int i = 0; i < kGraphStages && rc == MCV_OK; ++i) rc = launch_pair(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(MCV_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  // the batch takes the executable graph only once it exists
  hipGraphExec_t ex = nullptr;
  const hipError_t e2 = hipGraphInstantiateWithFlags(&ex, g, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess)
    return fail(MCV_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e2));
  b->exec = ex;
  b->exec_variant = b->variant;
  b->exec_mask = b->mask;
  b->exec_grid = b->stage_grid;
  b->exec_bnd_grid = b->bnd_grid;
  return MCV_OK;
}

int check_slot(const mcv_batch *b, int32_t s) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (s < 0 || s >= b->capacity || !b->slot[s])
    return fail(MCV_EINVAL, "slot %d is not occupied", s);
  return MCV_OK;
}

// Copies the device progress into the host view and the slots' handles.
int refresh(mcv_batch *b) {
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(b->progress.data(), b->d_prog, sizeof(BatchProgress) * b->capacity,
                    hipMemcpyDeviceToHost));
  for (int s = 0; s < b->capacity; ++s) {
    mcv_handle *h = b->slot[s];
    if (!h) continue;
    h->steps = b->progress[s].steps_done;
    if (!b->progress[s].first) h->first_pending = false;
  }
  return MCV_OK;
}

}  // namespace

extern "C" {

int mcv_batch_create(int32_t device, int32_t capacity, mcv_batch **out) {
  if (!out) return fail(MCV_EINVAL, "out is NULL");
  *out = nullptr;
  if (capacity < 1 || capacity > 65536)
    return fail(MCV_EINVAL, "capacity must be in [1, 65536], got %d", capacity);
  mcv_batch *b = new (std::nothrow) mcv_batch();
  if (!b) return fail(MCV_ENOMEM, "host allocation failed");
  auto init = [&]() -> int {
    b->device = device;
    b->capacity = capacity;
    b->slot.assign(capacity, nullptr);
    b->solve.assign(capacity, BatchSolve{});
    b->progress.assign(capacity, BatchProgress{0, 0, 0, 0});
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreate(&b->stream));
    for (hipEvent_t &e : b->ev) HIP_TRY(hipEventCreate(&e));
    int rc;
    if ((rc = balloc(b, &b->d_solve, capacity)) || (rc = balloc(b, &b->d_prog, capacity)) ||
        (rc = balloc(b, &b->d_list, capacity)) || (rc = balloc(b, &b->d_wprefix, capacity + 1)) ||
        (rc = balloc(b, &b->d_bprefix, capacity + 1)) || (rc = balloc(b, &b->d_meta, 4)))
      return rc;
    HIP_TRY(hipMemcpy(b->d_prog, b->progress.data(), sizeof(BatchProgress) * capacity,
                      hipMemcpyHostToDevice));
    const int zero[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpy(b->d_meta, zero, sizeof(zero), hipMemcpyHostToDevice));
    return MCV_OK;
  };
  int rc = init();
  if (rc) {
    std::string msg = mcv_hip_last_error();
    destroy_batch(b);
    fail(rc, "%s", msg.c_str());
    return rc;
  }
  *out = b;
  return MCV_OK;
}

int mcv_batch_add(mcv_batch *b, const mcv_config *cfg, const mcv_mesh *mesh,
                  const mcv_state *initial, int64_t nsteps, int32_t *slot) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (!slot) return fail(MCV_EINVAL, "slot is NULL");
  if (nsteps < 0) return fail(MCV_EINVAL, "nsteps must be >= 0");
  if (cfg && cfg->device != b->device)
    return fail(MCV_EINVAL, "solve on device %d, batch on device %d", cfg->device, b->device);
  int s = 0;
  while (s < b->capacity && b->slot[s]) ++s;
  if (s == b->capacity) return fail(MCV_ESTATE, "batch is full (%d slots)", b->capacity);
  mcv_handle *h = nullptr;
  int rc = mcv_hip_init(cfg, mesh, initial, &h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(b->device));
  b->solve[s] = describe(h);
  BatchProgress p{0, h->first_pending ? 1 : 0, nsteps, 0};
  b->progress[s] = p;
  // ordered after earlier work on the batch stream; a graph replay never
  // runs concurrently with these copies
  HIP_TRY(hipMemcpyAsync(b->d_solve + s, &b->solve[s], sizeof(BatchSolve), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_prog + s, &b->progress[s], sizeof(BatchProgress),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->slot[s] = h;
  b->tables_dirty = true;
  *slot = s;
  return MCV_OK;
}

int mcv_batch_remove(mcv_batch *b, int32_t s) {
  int rc = check_slot(b, s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  mcv_hip_free(b->slot[s]);
  b->slot[s] = nullptr;
  b->progress[s] = BatchProgress{0, 0, 0, 0};
  HIP_TRY(hipMemcpy(b->d_prog + s, &b->progress[s], sizeof(BatchProgress), hipMemcpyHostToDevice));
  b->tables_dirty = true;
  return MCV_OK;
}

int mcv_batch_run(mcv_batch *b, int32_t nstages) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (nstages < 0) return fail(MCV_EINVAL, "nstages must be >= 0");
  HIP_TRY(hipSetDevice(b->device));
  int rc;
  if (b->tables_dirty && (rc = upload_tables(b))) return rc;
  int left = nstages;
  if (b->use_graph && left >= kGraphStages) {
    if (!b->exec || b->exec_variant != b->variant || b->exec_mask != b->mask ||
        b->exec_grid != b->stage_grid || b->exec_bnd_grid != b->bnd_grid) {
      if ((rc = capture(b))) return rc;
    }
    for (; left >= kGraphStages; left -= kGraphStages)
      HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  }
  for (; left > 0; --left)
    if ((rc = launch_pair(b))) return rc;
  b->stages_run += nstages;
  return MCV_OK;
}

int mcv_batch_poll(mcv_batch *b, int64_t *steps_left, int64_t *steps_done) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  HIP_TRY(hipSetDevice(b->device));
  int rc = refresh(b);
  if (rc) return rc;
  for (int s = 0; s < b->capacity; ++s) {
    const bool occ = b->slot[s] != nullptr;
    if (steps_left) steps_left[s] = occ ? b->progress[s].steps_left : -1;
    if (steps_done) steps_done[s] = occ ? b->progress[s].steps_done : -1;
  }
  return MCV_OK;
}

int mcv_batch_sync(mcv_batch *b, int32_t s, mcv_state *out, double *mass, double *energy) {
  int rc = check_slot(b, s);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(b->device));
  if ((rc = refresh(b))) return rc;
  if (b->progress[s].phase != 0)
    return fail(MCV_ESTATE, "slot %d is inside a time step (stage %d)", s, b->progress[s].phase);
  return mcv_hip_sync(b->slot[s], out, mass, energy);
}

int mcv_batch_handle(mcv_batch *b, int32_t s, mcv_handle **h) {
  int rc = check_slot(b, s);
  if (rc) return rc;
  if (!h) return fail(MCV_EINVAL, "h is NULL");
  if ((rc = refresh(b))) return rc;
  *h = b->slot[s];
  return MCV_OK;
}

int mcv_batch_autotune(mcv_batch *b, int32_t reps, int32_t *tile, double *ms_per_stage) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (reps < 1) return fail(MCV_EINVAL, "reps must be >= 1");
  HIP_TRY(hipSetDevice(b->device));
  int rc = refresh(b);
  if (rc) return rc;
  // The first stage of every scheme reads Y_n and writes only scratch levels,
  // the accumulator (set, not accumulated) and the band, so repeating
  // it leaves the trajectory unchanged; it is only valid at a step boundary.
  for (int s = 0; s < b->capacity; ++s)
    if (b->slot[s] && b->progress[s].phase != 0)
      return fail(MCV_ESTATE, "slot %d is inside a time step; autotune between steps", s);
  const int keep = b->variant;
  int best = keep;
  float best_ms = 1e30f;
  for (int v = 0; v < kTileVariants; ++v) {
    b->variant = v;
    if ((rc = upload_tables(b))) break;
    if ((rc = launch_stage(b, v))) break;
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    for (int r = 0; r < reps && rc == MCV_OK; ++r) rc = launch_stage(b, v);
    if (rc) break;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    if (ms_per_stage) ms_per_stage[v] = ms / reps;
    if (ms < best_ms) { best_ms = ms; best = v; }
  }
  b->variant = rc ? keep : best;
  b->tables_dirty = true;
  if (rc) return rc;
  if (tile) *tile = best;
  return upload_tables(b);
}

int mcv_batch_set_tile(mcv_batch *b, int32_t tile) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (tile < 0 || tile >= kTileVariants)
    return fail(MCV_EINVAL, "tile must be in [0, %d), got %d", kTileVariants, tile);
  if (b->variant != tile) {
    b->variant = tile;
    b->tables_dirty = true;
  }
  return MCV_OK;
}

int mcv_batch_set_graphs(mcv_batch *b, int32_t on) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  b->use_graph = on != 0;
  return MCV_OK;
}

int mcv_batch_run_timed(mcv_batch *b, int32_t nstages, double *stage_ms, double *boundary_ms) {
  if (!b) return fail(MCV_EINVAL, "batch is NULL");
  if (nstages < 0) return fail(MCV_EINVAL, "nstages must be >= 0");
  HIP_TRY(hipSetDevice(b->device));
  int rc;
  if (b->tables_dirty && (rc = upload_tables(b))) return rc;
  double sm = 0.0, bm = 0.0;
  for (int i = 0; i < nstages; ++i) {
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if ((rc = launch_stage(b, b->variant))) return rc;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    if ((rc = launch_boundary(b))) return rc;
    HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[2]));
    float a = 0.f, c = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
    HIP_TRY(hipEventElapsedTime(&c, b->ev[1], b->ev[2]));
    sm += a;
    bm += c;
  }
  b->stages_run += nstages;
  if (stage_ms) *stage_ms = sm;
  if (boundary_ms) *boundary_ms = bm;
  return MCV_OK;
}

int mcv_batch_free(mcv_batch *b) {
  destroy_batch(b);
  return MCV_OK;
}

}  // extern "C"
