// Wave64 / workgroup helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace rskx {

constexpr int kWave = 64;

// DPP lane moves of a 64-bit value (two 32-bit halves).  Lanes whose source
// is out of range read 0 (row_bcast into row 0).
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const uint64_t u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
template <int CTRL>
__device__ inline int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
__device__ inline double read_lane(double v, int l) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Wave64 sum with every lane getting the total: butterflies inside each
// quad (quad_perm 1032, 2301), rotations inside each 16-lane row (row_ror
// 4, 8), then row_bcast 15 / 31 accumulate the four row sums into lane 63,
// which is read back as a uniform value.  DPP moves stay in the VALU;
// shuffles would go through the LDS crossbar twice per double.
__device__ inline double wave_sum(double v) {
  v += dpp_move<0xb1>(v);
  v += dpp_move<0x4e>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  v += dpp_move<0x142>(v);
  v += dpp_move<0x143>(v);
  return read_lane(v, 63);
}

// (value, index) maximum; ties go to the smaller index (LAPACK's i*amax
// picks the first maximum, MATLAB's max the first in column-major order)
__device__ inline void argmax_merge(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

// Same DPP pattern as wave_sum.  Lanes reading out of range get (0, 0),
// which only happens in row 0 of the broadcasts, whose result is unused.
__device__ inline void wave_argmax(double& v, int& i) {
#define RSKX_AMX(C) { const double v2 = dpp_move<C>(v); const int i2 = dpp_move<C>(i); argmax_merge(v, i, v2, i2); }
  RSKX_AMX(0xb1) RSKX_AMX(0x4e) RSKX_AMX(0x124) RSKX_AMX(0x128) RSKX_AMX(0x142) RSKX_AMX(0x143)
#undef RSKX_AMX
  v = read_lane(v, 63);
  i = __builtin_amdgcn_readlane(i, 63);
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
