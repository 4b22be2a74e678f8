// Shared device helpers and host error plumbing for libharmony_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "harmony_hip.h"

namespace hz {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // threads per workgroup for the tiled kernels
constexpr int kTileRows = 64;  // rows per LDS tile in the row/centroid kernels

// ---------------------------------------------------------------- errors

void set_error(const char* fmt, ...);

#define HZ_RET_IF(cond, code, ...)      \
  do {                                  \
    if (cond) {                         \
      ::hz::set_error(__VA_ARGS__);     \
      return code;                      \
    }                                   \
  } while (0)

#define HZ_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      ::hz::set_error("%s failed: %s", #call, hipGetErrorString(e_));     \
      return HZ_ERR_HIP;                                                  \
    }                                                                     \
  } while (0)

#define HZ_LAUNCHED(name)                                                 \
  do {                                                                    \
    hipError_t e_ = hipGetLastError();                                    \
    if (e_ != hipSuccess) {                                               \
      ::hz::set_error("kernel %s launch failed: %s", name,                \
                      hipGetErrorString(e_));                             \
      return HZ_ERR_LAUNCH;                                               \
    }                                                                     \
  } while (0)

#define HZ_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != HZ_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------- math

__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double dpow(double x, double y) { return pow(x, y); }

// Butterfly sum over groups of `width` adjacent lanes (width power of two <= 64).
template <typename V>
__device__ __forceinline__ V group_sum(V v, int width) {
  for (int m = width >> 1; m > 0; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Sum over each group of 16 lanes, result in every lane of the group, with
// DPP lane permutations (quad [1,0,3,2], quad [2,3,0,1], half-row mirror, row
// mirror): each step adds a lane's value to its partner's, so every lane ends
// with the same sum.
// (update_dpp with old = 0 lets each step fold into one v_add_f32_dpp; every
// lane reads a lane of its own row, so the bound control never applies)
__device__ __forceinline__ float sum16(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false));
  return v;
}

template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
  return group_sum(v, kWave);
}

// Sum of one value per thread over the workgroup; result valid in every thread.
// `red` needs blockDim.x / kWave entries.  Deterministic (fixed tree).
template <typename V>
__device__ V block_sum(V v, V* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  V s = V(0);
  const int nw = blockDim.x / kWave;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// Fixed-point accumulation: integer atomics make the sum independent of
// arrival order, so every reduction on the parity path is deterministic.
__device__ __forceinline__ long long to_fixed(double v, double scale) {
  return (long long)llrint(v * scale);
}
__device__ __forceinline__ void fixed_add(long long* dst, double v, double scale) {
  long long q = to_fixed(v, scale);
  if (q != 0)
    atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)q);
}
__device__ __forceinline__ double from_fixed(long long q, double inv_scale) {
  return (double)q * inv_scale;
}

// Fixed-point add that waits until the device-scope atomic has been performed
// (its old value is consumed), so a later relaxed arrival counter orders after
// it without a device-wide release fence.
__device__ __forceinline__ unsigned long long fixed_add_wait(long long* dst, double v, double scale) {
  long long q = to_fixed(v, scale);
  if (q == 0) return 0ull;
  return __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)q,
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ long long load_agent(const long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(long long* p, long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Number of bits needed for a non-negative value.
static inline int bit_length(unsigned long long v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

static inline long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }
__host__ __device__ inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace (256-byte aligned slices).
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* b) : base(static_cast<char*>(b)) {}
  template <typename P>
  P* take(size_t count) {
    off = align_up(off, 256);
    P* p = base ? reinterpret_cast<P*>(base + off) : nullptr;
    off += count * sizeof(P);
    return p;
  }
};

// Fill kernel (used instead of memset so the same code can be captured in a graph).
template <typename P>
__global__ void fill_kernel(P* __restrict__ p, long long n, P v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename P>
static int fill(P* p, long long n, P v, hipStream_t s) {
  if (n <= 0) return HZ_OK;
  int grid = (int)std::min<long long>(ceil_div(n, 256), 2048);
  hipLaunchKernelGGL(fill_kernel<P>, dim3(grid), dim3(256), 0, s, p, n, v);
  HZ_LAUNCHED("fill");
  return HZ_OK;
}

int device_lds_limit();
int device_cu_count();

}  // namespace hz
