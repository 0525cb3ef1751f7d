// Shared definitions for the gfx950 kernels of the periodic GPW engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hqc_amd.h"

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define HQC_CHECK_LAUNCH()                                   \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) return HQC_E_HIP;                  \
  } while (0)

// Tuned launch parameter `key` of the calling host thread, or dflt when unset
// (tune.hip; keys HQC_TUNE_* in hqc_amd.h).
int hqc_tuned(int key, int dflt);

// Thread-block sizes used throughout (one 64-lane wavefront = 64 threads).
#define HQC_WAVE 64

static inline int hqc_div_up(long a, long b) { return (int)((a + b - 1) / b); }

// Wavefront-wide sum of a double (all 64 lanes receive the result).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide maximum, as block_sum (result valid in thread 0).
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    t = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) t = fmax(t, red[i]);
  }
  __syncthreads();
  return t;
}

// Block-wide sum (blockDim.x a multiple of 64, at most 1024).  Result valid in
// thread 0.  `red` must hold blockDim.x/64 doubles of shared memory.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  }
  __syncthreads();
  return t;
}
