// TEST-ONLY CPU shim of the few HIP runtime names the product sources use.
//
// tests/emu builds mcv_sw_amd/csrc/mcv_hip.hip with g++ against this header
// into tests/emu/_build/libmcv_hip_emu.so, which exports the same C-ABI.  The
// CPU suite uses it to check the kernels' indexing (under AddressSanitizer)
// and their bitwise agreement with the oracle without a GPU.  The product
// package never loads it.
//
// Emulation model: a launch runs its blocks one after another, each with a
// single thread (blockDim = 1).  Every kernel in the product loops over its
// work with a block-strided loop, so one thread covers the whole block and
// __syncthreads() has nothing to wait for.
//
// With MCV_EMU_THREADS defined, a launch instead runs each block with its
// real blockDim: one host thread per GPU thread (a persistent team),
// __syncthreads() a barrier among them, __shared__ arrays shared by the team,
// blocks one after another.
// Built with ThreadSanitizer (tests/emu/build_emu.py, tsan_driver), this
// reports any two threads of a block that touch the same LDS or global word
// without a barrier between them (a missing __syncthreads, an LDS race).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#ifdef MCV_EMU_THREADS
#include <barrier>
#include <condition_variable>
#include <mutex>
#include <thread>
#endif

#define __global__
#define __device__
#define __host__
#define __constant__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __restrict__
// a compiler assumption in the product is a checked condition here
#define __builtin_assume(cond)                                                      \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      std::fprintf(stderr, "assumption failed: %s (%s:%d)\n", #cond, __FILE__,       \
                   __LINE__);                                                       \
      std::abort();                                                                 \
    }                                                                               \
  } while (0)

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
inline thread_local dim3 blockIdx(0), threadIdx(0), blockDim(1), gridDim(1);
#ifdef MCV_EMU_THREADS
inline thread_local std::barrier<> *g_emu_block_barrier = nullptr;
inline void __syncthreads() { g_emu_block_barrier->arrive_and_wait(); }
#else
inline void __syncthreads() {}
#endif
inline void __threadfence() {}
// the 24-bit multiply: low 32 bits of the product of the low 24 bits
inline unsigned __umul24(unsigned a, unsigned b) { return (a & 0xffffffu) * (b & 0xffffffu); }
inline unsigned atomicAdd(unsigned *p, unsigned v) {
  const unsigned o = *p;
  *p = o + v;
  return o;
}
inline int atomicAdd(int *p, int v) {
  const int o = *p;
  *p = o + v;
  return o;
}
inline int atomicExch(int *p, int v) {
  const int o = *p;
  *p = v;
  return o;
}
using std::max;
using std::min;

typedef int hipError_t;
typedef void *hipStream_t;
enum { hipSuccess = 0 };
enum hipMemcpyKind {
  hipMemcpyHostToDevice,
  hipMemcpyDeviceToHost,
  hipMemcpyDeviceToDevice,
  hipMemcpyHostToHost
};
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = (void *)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) {
  *p = std::calloc(1, n);
  return *p ? hipSuccess : 2;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind,
                                 hipStream_t = nullptr) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t = nullptr) {
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
  std::memset(d, v, n);
  return hipSuccess;
}
// Cache hints have no meaning here.
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
// A device symbol's address is the host object's own address here.
#define HIP_SYMBOL(x) reinterpret_cast<const void *>(&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
struct double2 {
  double x, y;
};
typedef void *hipEvent_t;
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}

// Device properties of the emulated device.
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount };
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
  *v = 4;
  return hipSuccess;
}
template <typename K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, K, int, size_t) {
  *n = 2;
  return hipSuccess;
}

// Stream capture records launches as closures; a graph launch replays them.
struct EmuGraph {
  std::vector<std::function<void()>> ops;
};
typedef EmuGraph *hipGraph_t;
typedef EmuGraph *hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal };
inline thread_local EmuGraph *g_emu_capture = nullptr;
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  g_emu_capture = new EmuGraph();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  *g = g_emu_capture;
  g_emu_capture = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiateWithFlags(hipGraphExec_t *e, hipGraph_t g,
                                               unsigned long long) {
  *e = new EmuGraph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &f : e->ops) f();
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}

#ifdef MCV_EMU_THREADS
// A persistent team of host threads, one per GPU thread of a block (created
// again when a launch asks for another block size).  run() hands every member
// the same job and returns when all have finished it; the mutex and condition
// variables order the launches with the team's work.
struct EmuTeam {
  std::vector<std::thread> members;
  unsigned size = 0;
  std::mutex mu;
  std::condition_variable start, done;
  std::function<void(unsigned)> job;
  unsigned long generation = 0;
  unsigned pending = 0;
  bool stop = false;
  void member(unsigned t) {
    unsigned long seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu);
      start.wait(lk, [&] { return stop || generation != seen; });
      if (stop) return;
      seen = generation;
      std::function<void(unsigned)> j = job;
      lk.unlock();
      j(t);
      lk.lock();
      if (--pending == 0) done.notify_one();
    }
  }
  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    start.notify_all();
    for (auto &m : members) m.join();
    members.clear();
    stop = false;
    size = 0;
  }
  void run(unsigned n, std::function<void(unsigned)> f) {
    if (size != n) {
      shutdown();
      size = n;
      for (unsigned t = 0; t < n; ++t) members.emplace_back([this, t] { member(t); });
    }
    std::unique_lock<std::mutex> lk(mu);
    job = std::move(f);
    pending = n;
    ++generation;
    start.notify_all();
    done.wait(lk, [&] { return pending == 0; });
  }
  ~EmuTeam() { shutdown(); }
};
inline EmuTeam g_emu_team;

// Each block is run by the team, one member per GPU thread; the members meet
// at a barrier after every block.
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)          \
  do {                                                                       \
    const dim3 g_ = dim3(grid), b_ = dim3(block);                            \
    auto op_ = [=]() {                                                       \
      std::barrier<> bar_(b_.x);                                             \
      g_emu_team.run(b_.x, [&](unsigned t_) {                                \
        g_emu_block_barrier = &bar_;                                         \
        gridDim = g_;                                                        \
        blockDim = dim3(b_.x);                                               \
        threadIdx = dim3(t_);                                                \
        for (unsigned bx_ = 0; bx_ < g_.x; ++bx_) {                          \
          blockIdx = dim3(bx_);                                              \
          kernel(__VA_ARGS__);                                               \
          bar_.arrive_and_wait();                                            \
        }                                                                    \
      });                                                                    \
    };                                                                       \
    if (g_emu_capture)                                                       \
      g_emu_capture->ops.push_back(op_);                                     \
    else                                                                     \
      op_();                                                                 \
  } while (0)
#else
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)          \
  do {                                                                       \
    const dim3 g_ = dim3(grid);                                              \
    auto op_ = [=]() {                                                       \
      gridDim = g_;                                                          \
      blockDim = dim3(1);                                                    \
      threadIdx = dim3(0);                                                   \
      for (unsigned bx_ = 0; bx_ < gridDim.x; ++bx_) {                       \
        blockIdx = dim3(bx_);                                                \
        kernel(__VA_ARGS__);                                                 \
      }                                                                      \
    };                                                                       \
    if (g_emu_capture)                                                       \
      g_emu_capture->ops.push_back(op_);                                     \
    else                                                                     \
      op_();                                                                 \
  } while (0)
#endif
