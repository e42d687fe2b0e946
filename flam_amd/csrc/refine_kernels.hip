// Orthogonalization step of the flexible GMRES refinement (the Krylov
// solve with an rskelf preconditioner of rskelf/test/ie_square3.m:97):
// classical Gram-Schmidt applied twice to w against the basis V, for all
// right-hand sides at once, with the new vector's norm.
//
// Layout: right-hand side r of the basis is V + r * ldr, its vector i at
// + i * ldv (N contiguous doubles); w + r * ldw likewise.  Each pass reads
// the nv basis vectors and w once to form the projections (k_cgs_dots),
// then once more to subtract them (k_cgs_update, after k_cgs_reduce sums
// the blocks' projections), which also sums the squares of the updated w;
// partial sums go through per-block buffers, never through atomics, so
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "devutil.hpp"

namespace rskx {

constexpr int kCgsMaxV = 32;        // basis vectors per pass
constexpr int kCgsThreads = 256;

// partial[(r * nblk + b) * NV + i] = sum over block b's rows of V[r,i,:] w[r,:]
// (NV, the number of basis vectors, is a compile-time constant so the
// accumulators stay in registers)
template <int NV>
__global__ void __launch_bounds__(kCgsThreads) k_cgs_dots(const double* __restrict__ V, int64_t ldv,
                                                          int64_t ldr, const double* __restrict__ w,
                                                          int64_t ldw, int64_t N,
                                                          double* __restrict__ partial) {
  constexpr int nv = NV;
  __shared__ double red[kCgsThreads / 64][NV];
  const int r = blockIdx.y, nblk = gridDim.x, b = blockIdx.x;
  const int64_t chunk = (N + nblk - 1) / nblk;
  const int64_t n0 = b * chunk, n1 = n0 + chunk < N ? n0 + chunk : N;
  const double* Vr = V + r * ldr;
  const double* wr = w + r * ldw;
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0;
#pragma unroll 2
  for (int64_t n = n0 + threadIdx.x; n < n1; n += kCgsThreads) {
    const double wv = wr[n];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] += Vr[i * ldv + n] * wv;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
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

// h[r * nv + i] = the projections summed over the nblk blocks (one
// workgroup per right-hand side, one wave per projection); hout's entries
// take this pass's projections (first pass) or add them (second)
__global__ void __launch_bounds__(kCgsThreads) k_cgs_reduce(const double* __restrict__ partial,
                                                            int nblk, int nv, double* __restrict__ h,
                                                            double* __restrict__ hout, int first) {
  const int r = blockIdx.x, lane = threadIdx.x & 63;
  for (int i = threadIdx.x >> 6; i < nv; i += kCgsThreads / 64) {
    double s = 0;
    for (int q = lane; q < nblk; q += 64) s += partial[((int64_t)r * nblk + q) * nv + i];
    s = wave_sum(s);
    if (lane == 0) {
      h[(int64_t)r * nv + i] = s;
      double* ho = hout + (int64_t)r * (nv + 1) + i;
      *ho = first ? s : *ho + s;
    }
  }
}

// w[r,:] -= sum_i V[r,i,:] h[r,i]; normpart[r * nblk + b] = block b's sum
// of squares of the updated w
template <int NV>
__global__ void __launch_bounds__(kCgsThreads) k_cgs_update(const double* __restrict__ V, int64_t ldv,
                                                            int64_t ldr, double* w, int64_t ldw,
                                                            int64_t N, const double* __restrict__ hg,
                                                            double* __restrict__ normpart) {
  constexpr int nv = NV;
  __shared__ double h[NV];
  __shared__ double red[kCgsThreads / 64];
  const int r = blockIdx.y, nblk = gridDim.x, b = blockIdx.x;
  if (threadIdx.x < nv) h[threadIdx.x] = hg[(int64_t)r * nv + threadIdx.x];
  __syncthreads();
  const int64_t chunk = (N + nblk - 1) / nblk;
  const int64_t n0 = b * chunk, n1 = n0 + chunk < N ? n0 + chunk : N;
  const double* Vr = V + r * ldr;
  double* wr = w + r * ldw;
  double sq = 0;
#pragma unroll 2
  for (int64_t n = n0 + threadIdx.x; n < n1; n += kCgsThreads) {
    double v = wr[n];
#pragma unroll
    for (int i = 0; i < NV; ++i) v -= Vr[i * ldv + n] * h[i];
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

// launch the dots (hg == nullptr) or the update kernel for nv basis
// vectors: the instantiation with NV == nv, found from kCgsMaxV down
template <int NV>
struct CgsLaunch {
  static void run(int nv, dim3 grid, const double* V, int64_t ldv, int64_t ldr, double* w,
                  int64_t ldw, int64_t N, double* partial, const double* hg, double* normpart,
                  hipStream_t st) {
    if (nv != NV) return CgsLaunch<NV - 1>::run(nv, grid, V, ldv, ldr, w, ldw, N, partial, hg,
                                               normpart, st);
    if (!hg)
      hipLaunchKernelGGL(k_cgs_dots<NV>, grid, dim3(kCgsThreads), 0, st, V, ldv, ldr, w, ldw, N,
                         partial);
    else
      hipLaunchKernelGGL(k_cgs_update<NV>, grid, dim3(kCgsThreads), 0, st, V, ldv, ldr, w, ldw, N,
                         hg, normpart);
  }
};
template <>
struct CgsLaunch<0> {
  static void run(int, dim3, const double*, int64_t, int64_t, double*, int64_t, int64_t, double*,
                  const double*, double*, hipStream_t) {}
};

int cgs_max_vectors() { return kCgsMaxV; }
int cgs_blocks(int64_t N) {
  const int64_t b = (N + 2047) / 2048;           // 2048 rows per block: ~4 blocks per CU per vector
  return (int)(b < 1 ? 1 : b > 1024 ? 1024 : b);
}

// scratch: partial (nrhs nblk nv), normpart (nrhs nblk), h (nrhs nv)
void launch_cgs(const double* V, int64_t ldv, int64_t ldr, int nv, double* w, int64_t ldw,
                int64_t N, int nrhs, double* partial, double* normpart, double* h, double* hout,
                hipStream_t st) {
  const int nblk = cgs_blocks(N);
  const dim3 grid(nblk, nrhs);
  for (int pass = 0; pass < 2; ++pass) {
    CgsLaunch<kCgsMaxV>::run(nv, grid, V, ldv, ldr, w, ldw, N, partial, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_cgs_reduce, dim3(nrhs), dim3(kCgsThreads), 0, st, partial, nblk, nv, h,
                       hout, pass == 0 ? 1 : 0);
    CgsLaunch<kCgsMaxV>::run(nv, grid, V, ldv, ldr, w, ldw, N, nullptr, h, normpart, st);
  }
  hipLaunchKernelGGL(k_cgs_norm, dim3(nrhs), dim3(kCgsThreads), 0, st, normpart, nblk, nv, hout);
}

}  // namespace rskx
