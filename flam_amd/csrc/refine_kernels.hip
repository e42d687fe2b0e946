// Orthogonalization step of the flexible GMRES refinement (the Krylov
// solve with an rskelf preconditioner of rskelf/test/ie_square3.m:97):
// classical Gram-Schmidt applied twice to w against the basis V, for all
// right-hand sides at once, with the new vector's norm.
//
// Layout: right-hand side r of the basis is V + r * ldr, its vector i at
// + i * ldv (N contiguous doubles); w + r * ldw likewise.  Each pass reads
// the nv basis vectors and w once to form the projections (k_cgs_dots),
// then once more to subtract them (k_cgs_update), which also sums the
// squares of the updated w; partial sums go through a per-block buffer,
// never through atomics, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "devutil.hpp"

namespace rskx {

constexpr int kCgsMaxV = 32;        // basis vectors per pass
constexpr int kCgsThreads = 256;

// partial[(r * nblk + b) * nv + i] = sum over block b's rows of V[r,i,:] w[r,:]
__global__ void __launch_bounds__(kCgsThreads) k_cgs_dots(const double* __restrict__ V, int64_t ldv,
                                                          int64_t ldr, int nv,
                                                          const double* __restrict__ w,
                                                          int64_t ldw, int64_t N,
                                                          double* __restrict__ partial) {
  __shared__ double red[kCgsThreads / 64][kCgsMaxV];
  const int r = blockIdx.y, nblk = gridDim.x, b = blockIdx.x;
  const int64_t chunk = (N + nblk - 1) / nblk;
  const int64_t n0 = b * chunk, n1 = n0 + chunk < N ? n0 + chunk : N;
  const double* Vr = V + r * ldr;
  const double* wr = w + r * ldw;
  double acc[kCgsMaxV];
#pragma unroll
  for (int i = 0; i < kCgsMaxV; ++i) acc[i] = 0;
  for (int64_t n = n0 + threadIdx.x; n < n1; n += kCgsThreads) {
    const double wv = wr[n];
#pragma unroll
    for (int i = 0; i < kCgsMaxV; ++i)
      if (i < nv) acc[i] += Vr[i * ldv + n] * wv;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kCgsMaxV; ++i) {
    if (i >= nv) break;
    const double s = wave_sum(acc[i]);
    if (lane == 0) red[wid][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0;
    for (int q = 0; q < kCgsThreads / 64; ++q) s += red[q][threadIdx.x];
    partial[((int64_t)r * nblk + b) * nv + threadIdx.x] = s;
  }
}

// h = the projections summed over blocks; w[r,:] -= sum_i V[r,i,:] h[i];
// normpart[r * nblk + b] = block b's sum of squares of the updated w;
// block 0 adds h into hout[r * (nv + 1) + i] (zeroed by the first pass)
__global__ void __launch_bounds__(kCgsThreads) k_cgs_update(const double* __restrict__ V, int64_t ldv,
                                                            int64_t ldr, int nv, double* w,
                                                            int64_t ldw, int64_t N,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ normpart,
                                                            double* __restrict__ hout, int first) {
  __shared__ double h[kCgsMaxV];
  __shared__ double red[kCgsThreads / 64];
  const int r = blockIdx.y, nblk = gridDim.x, b = blockIdx.x;
  if (threadIdx.x < nv) {
    double s = 0;
    for (int q = 0; q < nblk; ++q) s += partial[((int64_t)r * nblk + q) * nv + threadIdx.x];
    h[threadIdx.x] = s;
    if (b == 0) {
      double* ho = hout + (int64_t)r * (nv + 1) + threadIdx.x;
      *ho = first ? s : *ho + s;
    }
  }
  __syncthreads();
  const int64_t chunk = (N + nblk - 1) / nblk;
  const int64_t n0 = b * chunk, n1 = n0 + chunk < N ? n0 + chunk : N;
  const double* Vr = V + r * ldr;
  double* wr = w + r * ldw;
  double sq = 0;
  for (int64_t n = n0 + threadIdx.x; n < n1; n += kCgsThreads) {
    double v = wr[n];
#pragma unroll
    for (int i = 0; i < kCgsMaxV; ++i)
      if (i < nv) v -= Vr[i * ldv + n] * h[i];
    wr[n] = v;
    sq += v * v;
  }
  const double s = wave_sum(sq);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < kCgsThreads / 64; ++q) t += red[q];
    normpart[(int64_t)r * nblk + b] = t;
  }
}

// hout[r * (nv + 1) + nv] = || w[r,:] || from the last update's block sums
__global__ void k_cgs_norm(const double* __restrict__ normpart, int nblk, int nv,
                           double* __restrict__ hout) {
  const int r = blockIdx.x;
  __shared__ double red[kCgsThreads / 64];
  double s = 0;
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) s += normpart[(int64_t)r * nblk + q];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) t += red[q];
    hout[(int64_t)r * (nv + 1) + nv] = sqrt(t);
  }
}

int cgs_max_vectors() { return kCgsMaxV; }
int cgs_blocks(int64_t N) {
  const int64_t b = (N + 8191) / 8192;
  return (int)(b < 1 ? 1 : b > 256 ? 256 : b);
}

void launch_cgs(const double* V, int64_t ldv, int64_t ldr, int nv, double* w, int64_t ldw,
                int64_t N, int nrhs, double* partial, double* normpart, double* hout,
                hipStream_t st) {
  const int nblk = cgs_blocks(N);
  const dim3 grid(nblk, nrhs);
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(k_cgs_dots, grid, dim3(kCgsThreads), 0, st, V, ldv, ldr, nv, w, ldw, N,
                       partial);
    hipLaunchKernelGGL(k_cgs_update, grid, dim3(kCgsThreads), 0, st, V, ldv, ldr, nv, w, ldw, N,
                       partial, normpart, hout, pass == 0 ? 1 : 0);
  }
  hipLaunchKernelGGL(k_cgs_norm, dim3(nrhs), dim3(kCgsThreads), 0, st, normpart, nblk, nv, hout);
}

}  // namespace rskx
