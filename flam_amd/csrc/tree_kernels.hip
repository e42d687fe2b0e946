// Device side of the tree build (core/hypoct.m:74-133).
//
// Points live in a structure of arrays of coordinates plus their original
// indices, kept in "current order": every box owns a contiguous range.  At
// each level the ranges of the boxes being split are cut into chunks; one
// workgroup per chunk computes each point's child code (bit q set iff
// dimension q is divided and x_q > ctr_q, strictly) and either counts the
// codes or scatters the points to their child's range, keeping the order
// of the points inside each child (hypoct.m:92,108 keep xi's order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "devutil.hpp"

namespace rskx {

constexpr int kMaxCodes = 8;

template <bool WRITE>
__global__ void __launch_bounds__(256) k_tree_split(const TreeChunk* __restrict__ chunks, int d,
                                                    int64_t N, double* xs,
                                                    int* ord, double* txs, int* tord,
                                                    int* __restrict__ counts) {
  const TreeChunk C = chunks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  const int nw = blockDim.x / kWave;
  const int ncode = 1 << d;
  const int divmask = C.divmask;
  __shared__ int wcnt[kMaxCodes][4];
  __shared__ int64_t run[kMaxCodes];
  if (tid < ncode) run[tid] = WRITE ? C.off[tid] : 0;
  __syncthreads();
  for (int b0 = 0; b0 < C.len; b0 += blockDim.x) {
    const int i = b0 + tid;
    const bool act = i < C.len;
    const int64_t p = C.start + i;
    int code = -1;
    if (act) {
      code = 0;
      for (int q = 0; q < d; ++q)
        if (((divmask >> q) & 1) && xs[q * N + p] > C.ctr[q]) code |= 1 << q;
    }
    int rank = 0;
    for (int c = 0; c < ncode; ++c) {
      uint64_t m = __ballot(code == c);
      if (code == c) rank = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wcnt[c][wid] = __popcll(m);
    }
    __syncthreads();
    if (act && WRITE) {
      int64_t dst = run[code] + rank;
      for (int w = 0; w < wid; ++w) dst += wcnt[code][w];
      for (int q = 0; q < d; ++q) txs[q * N + dst] = xs[q * N + p];
      tord[dst] = ord[p];
    }
    __syncthreads();
    if (tid < ncode) {
      int tot = 0;
      for (int w = 0; w < nw; ++w) tot += wcnt[tid][w];
      run[tid] += tot;
    }
    __syncthreads();
  }
  if (!WRITE && tid < ncode) counts[blockIdx.x * kMaxCodes + tid] = (int)run[tid];
}

// copy the scattered ranges back into place
__global__ void k_tree_copyback(const TreeChunk* __restrict__ chunks, int d, int64_t N, double* xs,
                                int* ord, const double* __restrict__ txs,
                                const int* __restrict__ tord) {
  const TreeChunk C = chunks[blockIdx.x];
  for (int i = threadIdx.x; i < C.len; i += blockDim.x) {
    const int64_t p = C.start + i;
    for (int q = 0; q < d; ++q) xs[q * N + p] = txs[q * N + p];
    ord[p] = tord[p];
  }
}

// column-major d x N points into per-dimension arrays; identity order;
// per-block min/max of each problem's points (blockIdx.y: problem, whose
// points are [poff[j], poff[j+1])) for the root extents (hypoct.m:65)
__global__ void k_tree_init(const double* __restrict__ x, int d, int64_t N,
                            const int64_t* __restrict__ poff, double* xs, int* ord,
                            double* __restrict__ part) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int64_t p0 = poff[blockIdx.y], p1 = poff[blockIdx.y + 1];
  for (int64_t p = p0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < p1;
       p += (int64_t)gridDim.x * blockDim.x) {
    for (int q = 0; q < d; ++q) {
      double v = x[d * p + q];
      xs[q * N + p] = v;
      lo[q] = fmin(lo[q], v);
      hi[q] = fmax(hi[q], v);
    }
    ord[p] = (int)p;
  }
  __shared__ double slo[3][256], shi[3][256];
  for (int q = 0; q < d; ++q) { slo[q][threadIdx.x] = lo[q]; shi[q][threadIdx.x] = hi[q]; }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < d; ++q) {
        slo[q][threadIdx.x] = fmin(slo[q][threadIdx.x], slo[q][threadIdx.x + s]);
        shi[q][threadIdx.x] = fmax(shi[q][threadIdx.x], shi[q][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6;
    for (int q = 0; q < d; ++q) {
      o[q] = slo[q][0];
      o[3 + q] = shi[q][0];
    }
  }
}

// a[i] = i
__global__ void k_iota(int* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int)i;
}

void launch_iota(int* a, int64_t n, hipStream_t st) {
  const int grid = (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096);
  hipLaunchKernelGGL(k_iota, dim3(grid), dim3(256), 0, st, a, n);
}

void launch_tree_init(const double* x, int d, int64_t N, const int64_t* poff, int nprob,
                      double* xs, int* ord, double* part, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_tree_init, dim3(nblk, nprob), dim3(256), 0, st, x, d, N, poff, xs, ord,
                     part);
}

void launch_tree_split(bool write, const TreeChunk* chunks, int nchunk, int d, int64_t N,
                       double* xs, int* ord, double* txs, int* tord, int* counts,
                       hipStream_t st) {
  if (write) {
    hipLaunchKernelGGL(k_tree_split<true>, dim3(nchunk), dim3(256), 0, st, chunks, d, N,
                       xs, ord, txs, tord, counts);
    hipLaunchKernelGGL(k_tree_copyback, dim3(nchunk), dim3(256), 0, st, chunks, d, N, xs, ord,
                       txs, tord);
  } else {
    hipLaunchKernelGGL(k_tree_split<false>, dim3(nchunk), dim3(256), 0, st, chunks, d, N,
                       xs, ord, txs, tord, counts);
  }
}

}  // namespace rskx
