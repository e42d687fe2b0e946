// Batched per-level solve / apply sweeps (rskelf/sv/*.m, rskelf/mv/*.m).
//
// The factors of one tree level touch disjoint rows of X, so each level is
// one launch with a workgroup per factor: X(sk,:) and X(rd,:) are staged in
// shared memory, the three steps of the factor are applied there and the
// rows are written back.  X is in tree order and in the factor's precision.
#include <algorithm>

#include "common.hpp"
#include "devutil.hpp"
#include "rskx.h"

namespace rskx {

// Y (ny x cc) += sgn * op(A) Z,  op(A)(i,l) = A[i + l lda] or A[l + i lda]
template <typename S>
__device__ void upd(S* Y, int ny, const S* Z, int nz, const S* __restrict__ A, int lda, bool trA,
                    S sgn, int cc) {
  for (int i = threadIdx.x; i < ny; i += blockDim.x) {
    S acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (trA) {
      const S* a = A + (int64_t)i * lda;
      for (int l = 0; l < nz; ++l) {
        S al = a[l];
        for (int c = 0; c < cc; ++c) acc[c] += al * Z[l + c * nz];
      }
    } else {
      for (int l = 0; l < nz; ++l) {
        S al = A[i + (int64_t)l * lda];
        for (int c = 0; c < cc; ++c) acc[c] += al * Z[l + c * nz];
      }
    }
    for (int c = 0; c < cc; ++c) Y[i + c * ny] += sgn * acc[c];
  }
  __syncthreads();
}

// in-place triangular solve with the r x r triangle of A (leading dim r)
//   lower/trans false: L y = b ; lower/trans true: L' y = b
//   upper/trans false: U y = b ; upper/trans true: U' y = b
template <typename S>
__device__ void trsolve(S* Y, int r, const S* __restrict__ A, bool lower, bool unit, bool tr, int cc) {
  const bool fwd = (lower && !tr) || (!lower && tr);
  for (int t = 0; t < r; ++t) {
    const int j = fwd ? t : r - 1 - t;
    if (!unit) {
      for (int c = threadIdx.x; c < cc; c += blockDim.x) Y[j + c * r] /= A[j + (int64_t)j * r];
      __syncthreads();
    }
    // rows still to be solved: below j (fwd) or above j (bwd)
    const int lo = fwd ? j + 1 : 0, hi = fwd ? r : j;
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      S a = tr ? A[j + (int64_t)i * r] : A[i + (int64_t)j * r];
      for (int c = 0; c < cc; ++c) Y[i + c * r] -= a * Y[j + c * r];
    }
    __syncthreads();
  }
}

// Y = op(tri(A)) Y using tmp (r x cc)
template <typename S>
__device__ void trmul(S* Y, S* tmp, int r, const S* __restrict__ A, bool lower, bool unit, bool tr,
                      int cc) {
  const bool low = lower != tr;   // effective lower-triangular operator
  for (int i = threadIdx.x; i < r; i += blockDim.x) {
    S acc[8];
    for (int c = 0; c < cc; ++c) acc[c] = unit ? Y[i + c * r] : (S)0;
    int l0 = low ? 0 : (unit ? i + 1 : i);
    int l1 = low ? (unit ? i : i + 1) : r;
    for (int l = l0; l < l1; ++l) {
      S a = tr ? A[l + (int64_t)i * r] : A[i + (int64_t)l * r];
      for (int c = 0; c < cc; ++c) acc[c] += a * Y[l + c * r];
    }
    for (int c = 0; c < cc; ++c) tmp[i + c * r] = acc[c];
  }
  __syncthreads();
  for (int c = 0; c < cc; ++c)
    for (int i = threadIdx.x; i < r; i += blockDim.x) Y[i + c * r] = tmp[i + c * r];
  __syncthreads();
}

// gather (tmp(i) = Y(p(i))) or scatter (tmp(p(i)) = Y(i)), then copy back
template <typename S>
__device__ void permrows(S* Y, S* tmp, int r, const int* __restrict__ p, bool gather, int cc) {
  for (int i = threadIdx.x; i < r; i += blockDim.x) {
    int pi = p[i];
    for (int c = 0; c < cc; ++c) {
      if (gather) tmp[i + c * r] = Y[pi + c * r];
      else tmp[pi + c * r] = Y[i + c * r];
    }
  }
  __syncthreads();
  for (int c = 0; c < cc; ++c)
    for (int i = threadIdx.x; i < r; i += blockDim.x) Y[i + c * r] = tmp[i + c * r];
  __syncthreads();
}

// PROG: 0 sv 'n'/'n', 1 sv 'n'/'c', 2 sv 'p', 3 mv 'n'/'n', 4 mv 'n'/'c', 5 mv 'p'
template <typename S, int PROG, bool UP>
__global__ void __launch_bounds__(256) k_sweep(const FacRec* __restrict__ recs,
                                               const S* __restrict__ fval,
                                               const int* __restrict__ fidx, S* X, int64_t ldx,
                                               int nrhs, int cchunk) {
  extern __shared__ char smem_raw[];
  S* sm = reinterpret_cast<S*>(smem_raw);
  const FacRec f = recs[blockIdx.x];
  const int k = f.k, r = f.r;
  const int* sk = fidx + f.sk;
  const int* rd = fidx + f.rd;
  const int* piv = fidx + f.piv;
  const S* T = fval + f.T;
  const S* LU = fval + f.LU;
  const S* E = fval + f.E;
  const S* G = fval + f.G;
  const S one = 1, mone = -1;
  for (int c0 = 0; c0 < nrhs; c0 += cchunk) {
    const int cc = nrhs - c0 < cchunk ? nrhs - c0 : cchunk;
    S* xs = sm;
    S* xr = xs + k * cc;
    S* tmp = xr + r * cc;
    for (int c = 0; c < cc; ++c) {
      for (int a = threadIdx.x; a < k; a += blockDim.x) xs[a + c * k] = X[sk[a] + (c0 + c) * ldx];
      for (int i = threadIdx.x; i < r; i += blockDim.x) xr[i + c * r] = X[rd[i] + (c0 + c) * ldx];
    }
    __syncthreads();
    if (PROG == 0) {            // sv_nn.m
      if (UP) {
        upd(xr, r, xs, k, T, k, true, mone, cc);
        permrows(xr, tmp, r, piv, true, cc);
        trsolve(xr, r, LU, true, true, false, cc);
        upd(xs, k, xr, r, E, k, false, mone, cc);
      } else {
        upd(xr, r, xs, k, G, r, false, mone, cc);
        trsolve(xr, r, LU, false, false, false, cc);
        upd(xs, k, xr, r, T, k, false, mone, cc);
      }
    } else if (PROG == 1) {     // sv_nc.m
      if (UP) {
        upd(xr, r, xs, k, T, k, true, mone, cc);
        trsolve(xr, r, LU, false, false, true, cc);
        upd(xs, k, xr, r, G, r, true, mone, cc);
      } else {
        upd(xr, r, xs, k, E, k, true, mone, cc);
        trsolve(xr, r, LU, true, true, true, cc);
        permrows(xr, tmp, r, piv, false, cc);
        upd(xs, k, xr, r, T, k, false, mone, cc);
      }
    } else if (PROG == 2) {     // sv_p.m
      if (UP) {
        upd(xr, r, xs, k, T, k, true, mone, cc);
        trsolve(xr, r, LU, true, false, false, cc);
        upd(xs, k, xr, r, E, k, false, mone, cc);
      } else {
        upd(xr, r, xs, k, E, k, true, mone, cc);
        trsolve(xr, r, LU, true, false, true, cc);
        upd(xs, k, xr, r, T, k, false, mone, cc);
      }
    } else if (PROG == 3) {     // mv_nn.m
      if (UP) {
        upd(xs, k, xr, r, T, k, false, one, cc);
        trmul(xr, tmp, r, LU, false, false, false, cc);
        upd(xr, r, xs, k, G, r, false, one, cc);
      } else {
        upd(xs, k, xr, r, E, k, false, one, cc);
        trmul(xr, tmp, r, LU, true, true, false, cc);
        permrows(xr, tmp, r, piv, false, cc);
        upd(xr, r, xs, k, T, k, true, one, cc);
      }
    } else if (PROG == 4) {     // mv_nc.m
      if (UP) {
        upd(xs, k, xr, r, T, k, false, one, cc);
        permrows(xr, tmp, r, piv, true, cc);
        trmul(xr, tmp, r, LU, true, true, true, cc);
        upd(xr, r, xs, k, E, k, true, one, cc);
      } else {
        upd(xs, k, xr, r, G, r, true, one, cc);
        trmul(xr, tmp, r, LU, false, false, true, cc);
        upd(xr, r, xs, k, T, k, true, one, cc);
      }
    } else {                    // mv_p.m
      if (UP) {
        upd(xs, k, xr, r, T, k, false, one, cc);
        trmul(xr, tmp, r, LU, true, false, true, cc);
        upd(xr, r, xs, k, E, k, true, one, cc);
      } else {
        upd(xs, k, xr, r, E, k, false, one, cc);
        trmul(xr, tmp, r, LU, true, false, false, cc);
        upd(xr, r, xs, k, T, k, true, one, cc);
      }
    }
    for (int c = 0; c < cc; ++c) {
      for (int a = threadIdx.x; a < k; a += blockDim.x) X[sk[a] + (c0 + c) * ldx] = xs[a + c * k];
      for (int i = threadIdx.x; i < r; i += blockDim.x) X[rd[i] + (c0 + c) * ldx] = xr[i + c * r];
    }
    __syncthreads();
  }
}

template <typename S, int PROG>
static void launch_prog(bool up, const FacRec* recs, int nrec, const S* fval, const int* fidx, S* X,
                        int64_t ldx, int nrhs, int cchunk, size_t shmem, hipStream_t st) {
  if (shmem > 65536) {
    (void)hipFuncSetAttribute((const void*)k_sweep<S, PROG, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    (void)hipFuncSetAttribute((const void*)k_sweep<S, PROG, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  }
  if (up)
    hipLaunchKernelGGL((k_sweep<S, PROG, true>), dim3(nrec), dim3(256), shmem, st, recs, fval, fidx,
                       X, ldx, nrhs, cchunk);
  else
    hipLaunchKernelGGL((k_sweep<S, PROG, false>), dim3(nrec), dim3(256), shmem, st, recs, fval,
                       fidx, X, ldx, nrhs, cchunk);
}

template <typename S>
void launch_sweep(int prog, bool up, const FacRec* recs, int nrec, const S* fval, const int* fidx,
                  S* X, int64_t ldx, int nrhs, int cchunk, size_t shmem, hipStream_t st) {
  switch (prog) {
    case 0: launch_prog<S, 0>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
    case 1: launch_prog<S, 1>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
    case 2: launch_prog<S, 2>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
    case 3: launch_prog<S, 3>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
    case 4: launch_prog<S, 4>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
    default: launch_prog<S, 5>(up, recs, nrec, fval, fidx, X, ldx, nrhs, cchunk, shmem, st); break;
  }
}

// ------------------------------------------------ X to / from tree order --
template <typename U, typename S>
__global__ void k_to_tree(const U* __restrict__ X, int64_t ldx, S* __restrict__ Y, int64_t N,
                          int nrhs, const int* __restrict__ order) {
  int64_t tot = N * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t p = e % N, c = e / N;
    Y[e] = (S)X[order[p] + c * ldx];
  }
}

template <typename U, typename S>
__global__ void k_from_tree(const S* __restrict__ Y, U* __restrict__ X, int64_t ldx, int64_t N,
                            int nrhs, const int* __restrict__ order) {
  int64_t tot = N * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t p = e % N, c = e / N;
    X[order[p] + c * ldx] = (U)Y[e];
  }
}

template <typename U, typename S>
void launch_perm(bool to, void* X, int64_t ldx, S* Y, int64_t N, int nrhs, const int* order,
                 hipStream_t st) {
  int64_t tot = N * nrhs;
  int grid = (int)((tot + 255) / 256);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  if (to)
    hipLaunchKernelGGL((k_to_tree<U, S>), dim3(grid), dim3(256), 0, st, (const U*)X, ldx, Y, N,
                       nrhs, order);
  else
    hipLaunchKernelGGL((k_from_tree<U, S>), dim3(grid), dim3(256), 0, st, (const S*)Y, (U*)X, ldx,
                       N, nrhs, order);
}

// ------------------------------------------- FP64 on-the-fly A X product --
constexpr double kTwoPiM = 6.283185307179586476925286766559;

__global__ void __launch_bounds__(256) k_matvec(const double* __restrict__ x, int64_t N,
                                                const double* __restrict__ nu,
                                                const double* __restrict__ w,
                                                const double* __restrict__ diag, double scale,
                                                const double* __restrict__ X, int64_t ldx,
                                                double* __restrict__ Y, int64_t ldy, int nrhs) {
  __shared__ double sx[256], sy[256], snx[256], sny[256], sw[256];
  __shared__ double sX[8][256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool act = i < N;
  const double xi = act ? x[2 * i] : 0, yi = act ? x[2 * i + 1] : 0;
  for (int c0 = 0; c0 < nrhs; c0 += 8) {
    const int cc = nrhs - c0 < 8 ? nrhs - c0 : 8;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j0 = 0; j0 < N; j0 += 256) {
      __syncthreads();
      int64_t j = j0 + threadIdx.x;
      if (j < N) {
        sx[threadIdx.x] = x[2 * j];
        sy[threadIdx.x] = x[2 * j + 1];
        snx[threadIdx.x] = nu[2 * j];
        sny[threadIdx.x] = nu[2 * j + 1];
        sw[threadIdx.x] = w[j];
        for (int c = 0; c < cc; ++c) sX[c][threadIdx.x] = X[j + (c0 + c) * ldx];
      }
      __syncthreads();
      const int jn = N - j0 < 256 ? (int)(N - j0) : 256;
      for (int t = 0; t < jn; ++t) {
        int64_t jj = j0 + t;
        double a;
        if (jj == i) a = diag[jj];
        else {
          double dx = xi - sx[t], dy = yi - sy[t];
          a = scale * sw[t] * (dx * snx[t] + dy * sny[t]) / (kTwoPiM * (dx * dx + dy * dy));
        }
        for (int c = 0; c < cc; ++c) acc[c] += a * sX[c][t];
      }
    }
    if (act)
      for (int c = 0; c < cc; ++c) Y[i + (c0 + c) * ldy] = acc[c];
  }
}

void launch_matvec(const double* x, int64_t N, const double* nu, const double* w,
                   const double* diag, double scale, const double* X, int64_t ldx, double* Y,
                   int64_t ldy, int nrhs, hipStream_t st) {
  int grid = (int)((N + 255) / 256);
  hipLaunchKernelGGL(k_matvec, dim3(grid), dim3(256), 0, st, x, N, nu, w, diag, scale, X, ldx, Y,
                     ldy, nrhs);
}

// points, normals, weights and diagonal into tree-order structure of arrays
__global__ void k_gather_points(const double* __restrict__ x, const double* __restrict__ nu,
                                const double* __restrict__ w, const double* __restrict__ diag,
                                const int* __restrict__ order, int64_t N, int d,
                                double* __restrict__ pts) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t q = order[p];
    for (int i = 0; i < d; ++i) {
      pts[i * N + p] = x[d * q + i];
      pts[(d + i) * N + p] = nu[d * q + i];
    }
    pts[2 * d * N + p] = w[q];
    pts[(2 * d + 1) * N + p] = diag[q];
  }
}

void launch_gather_points(const double* x, const double* nu, const double* w, const double* diag,
                          const int* order, int64_t N, int d, double* pts, hipStream_t st) {
  int grid = (int)std::min<int64_t>((N + 255) / 256, 16384);
  hipLaunchKernelGGL(k_gather_points, dim3(grid < 1 ? 1 : grid), dim3(256), 0, st, x, nu, w, diag,
                     order, N, d, pts);
}

#define RSKX_SWEEP_INST(S)                                                                    \
  template void launch_sweep<S>(int, bool, const FacRec*, int, const S*, const int*, S*,      \
                                int64_t, int, int, size_t, hipStream_t);                      \
  template void launch_perm<double, S>(bool, void*, int64_t, S*, int64_t, int, const int*,    \
                                       hipStream_t);                                          \
  template void launch_perm<float, S>(bool, void*, int64_t, S*, int64_t, int, const int*,     \
                                      hipStream_t);
RSKX_SWEEP_INST(double)
RSKX_SWEEP_INST(float)

}  // namespace rskx
