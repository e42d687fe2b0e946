// Wave64 / workgroup helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace rskx {

constexpr int kWave = 64;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// (value, index) maximum; ties go to the smaller index (LAPACK's i*amax
// picks the first maximum, MATLAB's max the first in column-major order)
__device__ inline void argmax_merge(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__device__ inline void wave_argmax(double& v, int& i) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    double v2 = __shfl_xor(v, o, kWave);
    int i2 = __shfl_xor(i, o, kWave);
    argmax_merge(v, i, v2, i2);
  }
}

// Workgroup reductions; every thread gets the result.  `red` must hold
// 2 * (blockDim / 64) doubles of shared memory.
__device__ inline double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < nw; ++w) s += red[w];
  __syncthreads();
  return s;
}

__device__ inline void block_argmax(double& v, int& i, double* red) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  wave_argmax(v, i);
  int* redi = reinterpret_cast<int*>(red + nw);
  __syncthreads();
  if (lane == 0) { red[wid] = v; redi[wid] = i; }
  __syncthreads();
  v = red[0];
  i = redi[0];
  for (int w = 1; w < nw; ++w) argmax_merge(v, i, red[w], redi[w]);
  __syncthreads();
}

template <typename S> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr double v = FLT_EPSILON; };

}  // namespace rskx
