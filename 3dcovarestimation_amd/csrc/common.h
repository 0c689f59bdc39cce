// Shared helpers of the cov3d HIP library: error reporting, complex types,
// launch helpers.  All device code targets gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>

#include "../../include/cov3d_hip.h"

namespace cov3d {

// Thread-local last-error message (returned by cov3d_last_error).
void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <typename T>
struct Cplx {
  T x, y;
};

// Workgroup index dealt XCD-contiguously: the hardware hands consecutive workgroups to the 8 XCDs
// in turn, so workgroup b of n runs on XCD x = b mod 8 as that XCD's k-th (k = b / 8); XCD x holds
// q + (x < r) of them (q = n / 8, r = n mod 8), and renumbered as start(x) + k, start(x) =
// x q + min(x, r), each XCD takes one contiguous range of logical indices (a bijection of
// [0, n)), so neighbouring logical workgroups, which share cache lines, share that XCD's L2.
__device__ __forceinline__ int64_t xcd_contiguous(int64_t b, int64_t n) {
  const int64_t q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// One 8-byte (FP64: 16-byte) store of a complex value.  Cplx<T> is only T-aligned, and a
// member-wise store of one may be emitted as two separate stores (two partial writes of one word).
template <typename T>
__device__ __forceinline__ void store_cplx(Cplx<T>* p, T x, T y) {
  if constexpr (sizeof(T) == 4)
    *reinterpret_cast<float2*>(p) = make_float2((float)x, (float)y);
  else
    *reinterpret_cast<double2*>(p) = make_double2((double)x, (double)y);
}

template <typename T>
__device__ __forceinline__ Cplx<T> cadd(Cplx<T> a, Cplx<T> b) {
  return {a.x + b.x, a.y + b.y};
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace cov3d

#define COV3D_CHECK_ARG(cond, ...)       \
  do {                                   \
    if (!(cond)) {                       \
      ::cov3d::set_error(__VA_ARGS__);   \
      return COV3D_ERR_ARG;              \
    }                                    \
  } while (0)

#define COV3D_HIP(call)                                                     \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::cov3d::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                               \
      return COV3D_ERR_HIP;                                                 \
    }                                                                       \
  } while (0)

// Checks the launch status of the kernel just queued.
#define COV3D_LAUNCH_CHECK() COV3D_HIP(hipGetLastError())
