// Host/launch helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "srmhd_common.h"

namespace srmhd {

constexpr int BLK = 256;

inline int blocks_for(int64_t n) { return (int)((n + BLK - 1) / BLK); }

inline int hip_status(hipError_t e) { return e == hipSuccess ? SRMHD_OK : SRMHD_EHIP + (int)e; }

inline int launch_status() { return hip_status(hipGetLastError()); }

// A second stream per device for work that can run beside the sweeps (the
// FP64 form's y sweep beside its x sweep; the mixed form's FP64 fix-up of the
// x faces, a few latency-bound waves, beside the y sweep; the adjoint's y
// sweep beside its x sweep), forked and joined with events.  Created at
// first use, at the default priority; one per device, shared by every
// translation unit of the library (inline variables).
struct AuxStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool failed = false;   // creation failed once: this device stays on one stream
};
constexpr int AUX_MAX_DEVICES = 64;
inline AuxStream g_aux[AUX_MAX_DEVICES];
inline std::mutex g_aux_mu;   // held while a stage enqueues its fork/join

// The current device's auxiliary stream, created at first use (nullptr when
// unavailable: the caller then stays on its own stream).  Caller holds
// g_aux_mu.
inline AuxStream* aux_stream() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= AUX_MAX_DEVICES) return nullptr;
  AuxStream& a = g_aux[dev];
  if (!a.s && !a.failed) {
    hipStream_t st = nullptr;
    hipEvent_t f = nullptr, j = nullptr;
    const bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&f, hipEventDisableTiming) == hipSuccess &&
                    hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      if (j) (void)hipEventDestroy(j);
      if (f) (void)hipEventDestroy(f);
      if (st) (void)hipStreamDestroy(st);
      (void)hipGetLastError();   // the failure is handled here, not by the next launch check
      a.failed = true;
      return nullptr;
    }
    a.fork = f;
    a.join = j;
    a.s = st;
  }
  return a.s ? &a : nullptr;
}

__device__ inline unsigned long long wave_or(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}
__device__ inline unsigned long long wave_max(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
__device__ inline unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ inline bool lane0() { return (threadIdx.x & 63) == 0; }

// The workgroup's index in an order that gives each of the 8 XCDs one
// contiguous run of the grid: the dispatcher deals workgroup b to XCD b % 8,
// so XCD x holds b = x, x + 8, ...; they are renumbered to the x-th run of
// the grid (runs of nb / 8, the first nb % 8 one longer).  A bijection on
// [0, nb) for any nb, so correctness does not depend on the dealing order;
// the gain does: neighbouring workgroups, whose stencils share rows, then
// share an L2.
__device__ inline int64_t xcd_block() {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  const int64_t x = b & 7, q = b >> 3, per = nb >> 3, rem = nb & 7;
  return x * per + (x < rem ? x : rem) + q;
}

// Appends idx to the work list q for the lanes with want set: one atomic per
// wave on the count *n.  Every lane of the wave that is still running calls
// it.  The order of the entries is unspecified.
__device__ inline void queue_push(int64_t* q, unsigned long long* n, int64_t idx, bool want) {
  const unsigned long long m = __ballot(want);
  if (!m) return;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __builtin_ctzll(m);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(n, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (want) q[base + __popcll(m & ((1ull << lane) - 1ull))] = idx;
}

__device__ inline unsigned long long dbits(double x) { return __double_as_longlong(x); }

__device__ inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

struct Carver {
  char* base;
  int64_t off = 0;
  template <typename T>
  T* take(int64_t count) {
    T* p = (T*)(base + off);
    off += align256(count * (int64_t)sizeof(T));
    return p;
  }
};

inline bool eos_ok(const srmhd_eos* e) { return e && e->gamma > 1.0; }

// live kernel timing (srmhd_timing_enable): HIP event pairs around the
// launches of one kernel class, defined in srmhd_kernels.hip
void timing_begin(int cls, hipStream_t s);
void timing_end(int cls, hipStream_t s);

}  // namespace srmhd
