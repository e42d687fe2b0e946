// Boxes too large for one workgroup: the few top-level boxes of 3D problems
// (n in the thousands), each processed with the whole GPU.
//
// ID (tol >= 1e-7, where the Gram route is accurate): K = [A(nbr,slf);
// A(slf,nbr)' (two-sided); Kpxy] (rskelf.m:201-212) assembled in FP64, one
// workgroup per column; G = K'K by rocBLAS dsyrk; then the diagonally
// pivoted Cholesky of G, stopped as id.m:105-111 stops the QR, computed
// left-looking so G is never rewritten: step s picks the largest remaining
// diagonal (one workgroup), then every column's R entry is formed as
// (G(pc, j) - R(0:s, pc)' R(0:s, j)) / R(s, pc) by a grid-wide pass.  R's
// leading rows by pivot position go where k_gram_id leaves them, and the
// same tail (k_id<S, 2>: T = R11 \ R12 and the Gu-Eisenstat refinement)
// finishes the ID.
//
// Elimination (symm 'n', 's' (real) and 'p'): K = A(slf,slf) + M in the
// order [rd; sk] assembled densely, the two sparsification products, LU
// with partial pivoting (rocSOLVER getrf) or Cholesky (potrf), E and G by
// triangular solves and the skeleton Schur update by a GEMM (rocBLAS),
// exactly rskelf.m:218-246; the factor is then written in the layout of the
// other elimination kernels (k_pack adds the transposed copies and the
// inverted diagonal blocks).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdio>
#include <mutex>

#include "common.hpp"
#include "devutil.hpp"
#include "entries.hpp"
#include "rskx.h"

namespace rskx {

// ------------------------------------------------------------- the ID --
// column j of K (mb rows, leading dimension mb; rows past the box's m are 0)
__global__ void __launch_bounds__(256) k_big_assemble(int b, const LvlBox* __restrict__ boxes,
                                                      const int* __restrict__ pool,
                                                      const int* __restrict__ nbrar,
                                                      const int* __restrict__ mnbr,
                                                      int* __restrict__ mrows, double* __restrict__ K,
                                                      int mb, PointsDev P, LevelParams L) {
  const LvlBox B = boxes[b];
  const int j = blockIdx.x;
  const int n = B.n, mn = mnbr[b];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  if (j == 0 && threadIdx.x == 0) mrows[b] = m;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  double xj, yj, nxj, nyj, wj;
  point_fields(P, slf[j], xj, yj, nxj, nyj, wj);
  const bool sl3 = P.kind == RSKX_KERNEL_LAPLACE3D_SL;
  const double c0 = P.scale / kTwoPi;
  double* Kj = K + (int64_t)j * mb;
  for (int i = threadIdx.x; i < mb; i += blockDim.x) {
    double v = 0;
    if (i < mn) {                                           // A(nbr, slf)
      double x, y, f2, f3, w;
      point_fields(P, nbr[i], x, y, f2, f3, w);
      if (sl3) {
        v = sl3_entry(P, x, y, f2, xj, yj, nxj, wj);
      } else {
        const double dx = x - xj, dy = y - yj;
        v = c0 * wj * (dx * nxj + dy * nyj) / (dx * dx + dy * dy);
      }
    } else if (i < mt) {                                    // A(slf, nbr)'
      double x, y, f2, f3, w;
      point_fields(P, nbr[i - mn], x, y, f2, f3, w);
      if (sl3) {
        v = sl3_entry(P, xj, yj, nxj, x, y, f2, w);
      } else {
        const double dx = xj - x, dy = yj - y;
        v = c0 * w * (dx * f2 + dy * f3) / (dx * dx + dy * dy);
      }
    } else if (i < m) {                                     // proxy rows
      const int q = i - mt;
      double px, py, pz;
      proxy_point(L, B, q, px, py, pz);
      if (L.pkind == RSKX_PROXY_POINTS_SL) {
        v = proxy_sl3(L, B, q, px, py, pz, xj, yj, nxj, wj);
      } else {
        const double dx = px - xj, dy = py - yj;
        const double r2 = dx * dx + dy * dy;
        v = L.pkind == RSKX_PROXY_RING_SL ? -0.5 * log(r2) / kTwoPi * B.pw
                                          : wj * (dx * nxj + dy * nyj) / (kTwoPi * r2);
      }
    }
    Kj[i] = v;
  }
}

// state of the pivoted Cholesky on the device
struct BigPiv {
  int k;           // rows of R formed so far
  int stop;        // set when the ID's rank is reached
  int pc;          // pivot column of the row being formed
  int pad;
  double nrm;      // its pivot (the remaining column norm)
  double tol_abs;  // tol times the first pivot (id.m:105-111)
};

__global__ void k_big_init(BigPiv* st, const double* __restrict__ G, double* __restrict__ dg,
                           int* __restrict__ p2c, int n) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    dg[j] = G[j + (int64_t)j * n];
    p2c[j] = j;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->k = 0;
    st->stop = 0;
  }
}

// pivot of step s = st->k: the largest remaining diagonal over positions
// s.. (ties to the smaller position), with the stopping tests
__global__ void __launch_bounds__(1024) k_big_pivot(BigPiv* st, const double* __restrict__ dg,
                                                     int* __restrict__ p2c, int n, int smax,
                                                     int kmax, double tol) {
  __shared__ double red[2 * 16];
  if (st->stop) return;
  const int s = st->k;
  if (s >= smax || s >= kmax) {
    if (threadIdx.x == 0) st->stop = 1;
    return;
  }
  double bv = -1.0;
  int bi = n;
  for (int pos = s + threadIdx.x; pos < n; pos += blockDim.x) argmax_merge(bv, bi, dg[p2c[pos]], pos);
  block_argmax(bv, bi, red);
  if (threadIdx.x == 0) {
    if (bi >= n) { st->stop = 1; return; }
    const double nrm = sqrt(fmax(bv, 0.0));
    if (s == 0) st->tol_abs = nrm * tol;
    if (!(nrm > st->tol_abs)) { st->stop = 1; return; }
    const int pc = p2c[bi];
    p2c[bi] = p2c[s];
    p2c[s] = pc;
    st->pc = pc;
    st->nrm = nrm;
    st->k = s + 1;
  }
}

// row s = k - 1 of R for every column: (G(pc, j) - sum_{t<s} R(t, pc) R(t, j)) / nrm,
// R row-major (row t at Rrow + t n); G holds its lower triangle
__global__ void __launch_bounds__(256) k_big_row(const BigPiv* st, const double* __restrict__ G,
                                                 int n, double* __restrict__ Rrow,
                                                 double* __restrict__ dg, int last_k) {
  if (st->k != last_k + 1) return;           // this step formed no row
  const int s = st->k - 1, pc = st->pc;
  const double inv = 1.0 / st->nrm;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double g = j <= pc ? G[pc + (int64_t)j * n] : G[j + (int64_t)pc * n];
  for (int t = 0; t < s; ++t) g -= Rrow[(int64_t)t * n + pc] * Rrow[(int64_t)t * n + j];
  const double r = g * inv;
  Rrow[(int64_t)s * n + j] = r;
  dg[j] -= r * r;
}

// R's leading k rows by pivot position into the box's third T slice (as
// k_gram_id leaves them), the pivot order and the rank
__global__ void k_big_finish(int b, const LvlBox* __restrict__ boxes, const BigPiv* st,
                             const double* __restrict__ Rrow, const int* __restrict__ p2c,
                             double* __restrict__ Tar, int* __restrict__ permar,
                             int* __restrict__ kout) {
  const LvlBox B = boxes[b];
  const int n = B.n, ldt = B.ldt, k = st->k;
  double* Rp = Tar + B.T_off + 2 * (int64_t)ldt * n;
  int* perm = permar + B.perm_off;
  for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += gridDim.x * blockDim.x) {
    const int c = p2c[pos];
    perm[pos] = c;
    for (int i = 0; i < k; ++i) Rp[i + (int64_t)pos * ldt] = Rrow[(int64_t)i * n + c];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) kout[b] = k;
}

// ------------------------------------------------------ the elimination --
// K = A(slf, slf) + M in the order [rd; sk] (n x n, leading dimension n),
// one workgroup per column; with mo_only, only the sk x sk block and only
// M there (the new skeleton block starts from M(sk, sk), rskelf.m:246)
template <typename S>
__global__ void __launch_bounds__(256) k_big_kassemble(const LvlBox* __restrict__ boxes, int b,
                                                       const BoxOut* __restrict__ outs,
                                                       const Seg* __restrict__ csegs,
                                                       const int* __restrict__ pool,
                                                       const int* __restrict__ permar,
                                                       const S* __restrict__ Mprev,
                                                       S* __restrict__ Z, PointsDev P,
                                                       int mo_only) {
  __shared__ ChildMap cm;
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int n = B.n, k = O.k, r = O.r;
  if (threadIdx.x == 0) {
    const int nc = B.ncs > 8 ? 8 : B.ncs;
    cm.nc = nc;
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      const Seg g = csegs[B.cs0 + c];
      cm.start[c] = acc;
      cm.moff[c] = g.moff;
      acc += g.len;
    }
    cm.start[nc] = acc;
  }
  __syncthreads();
  const bool leafbox = cm.nc == 1 && cm.moff[0] < 0;
  const int* slf = pool + B.slf_off;
  const int* q = permar + B.perm_off;           // q[0..k) skeletons, q[k..n) redundant
  const int j = mo_only ? r + blockIdx.x : blockIdx.x;
  const int sj = j < r ? q[k + j] : q[j - r];
  double xj, yj, f2j, f3j, wj;
  point_fields(P, slf[sj], xj, yj, f2j, f3j, wj);
  const double dgj = P.diag[slf[sj]];
  S* Zj = Z + (int64_t)j * n;
  for (int i = (mo_only ? r : 0) + threadIdx.x; i < n; i += blockDim.x) {
    const int si = i < r ? q[k + i] : q[i - r];
    double v = 0;
    if (!mo_only) {
      if (P.kind == RSKX_KERNEL_CALLBACK) {
        v = P.blk[B.A_off + si + (int64_t)sj * n];
      } else if (si == sj) {
        v = dgj;
      } else {
        double x, y, f2, f3, w;
        point_fields(P, slf[si], x, y, f2, f3, w);
        if (P.kind == RSKX_KERNEL_LAPLACE3D_SL) {
          v = sl3_entry(P, x, y, f2, xj, yj, f2j, wj);
        } else {
          const double dx = x - xj, dy = y - yj;
          v = P.scale * wj * (dx * f2j + dy * f3j) / (kTwoPi * (dx * dx + dy * dy));
        }
      }
    }
    S z = (S)v;
    if (!leafbox) z += mval(cm, Mprev, si, sj);
    Zj[i] = z;
  }
}

// T (k x r, FP64, leading dimension ldt) in the factor's precision,
// leading dimension k
template <typename S>
__global__ void k_big_tconv(const LvlBox* __restrict__ boxes, int b, const BoxOut* __restrict__ outs,
                            const double* __restrict__ Tar, S* __restrict__ Ts) {
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int k = O.k, r = O.r;
  const double* Tg = Tar + B.T_off;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)k * r;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(e % k), c = (int)(e / k);
    Ts[e] = (S)Tg[a + (int64_t)c * B.ldt];
  }
}

// the factor from the dense result (Z, leading dimension n): L and U (or
// L), E, G, T, the new skeleton block, the pivot vector (from LAPACK's
// interchanges), the point lists
template <typename S>
__global__ void k_big_out(const LvlBox* __restrict__ boxes, int b, const BoxOut* __restrict__ outs,
                          int* pool, const int* __restrict__ permar, const S* __restrict__ Z,
                          const S* __restrict__ Ts, const int* __restrict__ ipiv,
                          S* __restrict__ fval, int* __restrict__ fidx, S* __restrict__ Mcur,
                          int lu) {
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int n = B.n, k = O.k, r = O.r;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  S* fLU = fval + O.LU;
  S* fE = fval + O.E;
  S* fT = fval + O.T;
  for (int64_t e = tid; e < (int64_t)r * r; e += nt) {
    const int i = (int)(e % r), j = (int)(e / r);
    fLU[e] = (lu || i >= j) ? Z[i + (int64_t)j * n] : (S)0;
  }
  for (int64_t e = tid; e < (int64_t)k * r; e += nt) {
    const int a = (int)(e % k), j = (int)(e / k);
    fE[e] = Z[r + a + (int64_t)j * n];
    fT[e] = Ts[e];
  }
  if (lu) {
    S* fG = fval + O.G;
    for (int64_t e = tid; e < (int64_t)r * k; e += nt) {
      const int i = (int)(e % r), a = (int)(e / r);
      fG[e] = Z[i + (int64_t)(r + a) * n];
    }
  }
  S* Mo = Mcur + O.M_off;
  for (int64_t e = tid; e < (int64_t)k * k; e += nt) {
    const int a = (int)(e % k), c = (int)(e / k);
    Mo[e] = Z[r + a + (int64_t)(r + c) * n];
  }
  const int* slf = pool + B.slf_off;
  const int* q = permar + B.perm_off;
  int* fsk = fidx + O.sk;
  int* frd = fidx + O.rd;
  for (int64_t a = tid; a < k; a += nt) {
    const int pt = slf[q[a]];
    fsk[a] = pt;
    pool[O.nl_off + a] = pt;
  }
  for (int64_t i = tid; i < r; i += nt) frd[i] = slf[q[k + i]];
}

// LAPACK's interchanges (1-based ipiv) as the permutation vector p with
// L U = K(rd(p), rd), [L,U,p] = lu(...,'vector'); one thread walks them
__global__ void k_big_piv(const BoxOut* __restrict__ outs, int b, const int* __restrict__ ipiv,
                          int* __restrict__ fidx) {
  const BoxOut O = outs[b];
  const int r = O.r;
  int* piv = fidx + O.piv;
  for (int i = threadIdx.x; i < r; i += blockDim.x) piv[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < r; ++i) {
      const int t = ipiv[i] - 1;
      if (t != i) { const int a = piv[i]; piv[i] = piv[t]; piv[t] = a; }
    }
}

// inverted 32 x 32 diagonal blocks of L (unit for LU) and of U, one wave
// per block: lane c forms column c of the block's inverse by substitution
// from the factor just written
template <typename S>
__global__ void __launch_bounds__(64) k_big_inv(const BoxOut* __restrict__ outs, int b,
                                                S* __restrict__ fval, int lu) {
  const BoxOut O = outs[b];
  const int r = O.r, nbk = (r + kDiagBlock - 1) / kDiagBlock;
  const int bb = blockIdx.x;
  const bool upper = bb >= nbk;
  if (upper && !lu) return;
  const int blk = upper ? bb - nbk : bb;
  const int j0 = blk * kDiagBlock, bs = r - j0 < kDiagBlock ? r - j0 : kDiagBlock;
  const S* Lb = fval + O.LU + j0 + (int64_t)j0 * r;
  const int c = threadIdx.x;
  S x[kDiagBlock];
#pragma unroll
  for (int t = 0; t < kDiagBlock; ++t) x[t] = (S)0;
  if (!upper) {
#pragma unroll
    for (int t = 0; t < kDiagBlock; ++t) {
      if (t < bs) {
        S v = (t == c) ? (S)1 : (S)0;
#pragma unroll
        for (int u = 0; u < t; ++u) v -= Lb[t + (int64_t)u * r] * x[u];
        x[t] = lu ? v : v / Lb[t + (int64_t)t * r];
      }
    }
  } else {
#pragma unroll
    for (int tt = kDiagBlock - 1; tt >= 0; --tt) {
      if (tt < bs) {
        S v = (tt == c) ? (S)1 : (S)0;
#pragma unroll
        for (int u = tt + 1; u < kDiagBlock; ++u)
          if (u < bs) v -= Lb[tt + (int64_t)u * r] * x[u];
        x[tt] = v / Lb[tt + (int64_t)tt * r];
      }
    }
  }
  if (c < kDiagBlock) {
    S* dst = fval + (upper ? O.Ui : O.Li) + (int64_t)blk * kDiagBlock * kDiagBlock;
#pragma unroll
    for (int t = 0; t < kDiagBlock; ++t) dst[t + c * kDiagBlock] = (t < bs && c < bs) ? x[t] : (S)0;
  }
}

// a zero pivot (getrf) or a non-positive one (potrf) raises the level's
// failure flag as the other elimination kernels do
__global__ void k_big_info(const int* __restrict__ info, int lu, int* __restrict__ fail) {
  if (*info > 0) atomicOr(fail, lu ? 2 : 1);
}

// ------------------------------------------------------------ host side --
static rocblas_handle big_handle(hipStream_t st) {
  static std::mutex mu;
  static rocblas_handle h[64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  rocblas_handle& hd = h[dev & 63];
  if (!hd && rocblas_create_handle(&hd) != rocblas_status_success) hd = nullptr;
  if (hd) rocblas_set_stream(hd, st);
  return hd;
}

#define RSKX_BLAS(call)                                                                   \
  do {                                                                                    \
    const rocblas_status s_ = (call);                                                     \
    if (s_ != rocblas_status_success) {                                                   \
      set_error(std::string("FLAM:rskx:blas: ") + #call + ": " + rocblas_status_to_string(s_)); \
      return RSKX_ERR_HIP;                                                                \
    }                                                                                     \
  } while (0)

// device scratch of the big-box path: sizes in elements
size_t big_id_scratch(int n, int mb) {
  return (size_t)mb * n + (size_t)n * n + (size_t)n * n + 2 * (size_t)n + 64;   // K, G, R rows, dg, p2c, state
}

int big_id(int b, int n, int mb, int kmax, int smax, double tol, const LvlBox* boxes,
           const int* pool, const int* nbrar, const int* mnbr, int* mrows, double* Tar,
           int* permar, int* kout, const PointsDev& P, const LevelParams& L, double* scratch,
           hipStream_t st) {
  double* K = scratch;
  double* G = K + (size_t)mb * n;
  double* Rrow = G + (size_t)n * n;
  double* dg = Rrow + (size_t)n * n;
  int* p2c = reinterpret_cast<int*>(dg + n);
  BigPiv* state = reinterpret_cast<BigPiv*>(dg + 2 * (size_t)n);
  rocblas_handle h = big_handle(st);
  if (!h) return (set_error("FLAM:rskx:blas: rocblas_create_handle failed"), RSKX_ERR_HIP);
  hipLaunchKernelGGL(k_big_assemble, dim3(n), dim3(256), 0, st, b, boxes, pool, nbrar, mnbr, mrows,
                     K, mb, P, L);
  const double one = 1.0, zero = 0.0;
  RSKX_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, n, mb, &one, K, mb,
                          &zero, G, n));
  hipLaunchKernelGGL(k_big_init, dim3((n + 255) / 256), dim3(256), 0, st, state, G, dg, p2c, n);
  int steps = smax < kmax ? smax : kmax;
  if (steps > n) steps = n;
  int done = 0, hstop = 0, hk = 0;
  // steps in chunks, the host reading the stop flag after each chunk
  while (done < steps && !hstop) {
    const int chunk = steps - done < 32 ? steps - done : 32;
    for (int c = 0; c < chunk; ++c) {
      hipLaunchKernelGGL(k_big_pivot, dim3(1), dim3(1024), 0, st, state, dg, p2c, n, smax, kmax, tol);
      hipLaunchKernelGGL(k_big_row, dim3((n + 255) / 256), dim3(256), 0, st, state, G, n, Rrow, dg,
                         done + c);
    }
    done += chunk;
    RSKX_HIP(hipMemcpyAsync(&hstop, &state->stop, sizeof(int), hipMemcpyDeviceToHost, st));
    RSKX_HIP(hipMemcpyAsync(&hk, &state->k, sizeof(int), hipMemcpyDeviceToHost, st));
    RSKX_HIP(hipStreamSynchronize(st));
  }
  hipLaunchKernelGGL(k_big_finish, dim3((n + 255) / 256), dim3(256), 0, st, b, boxes, state, Rrow,
                     p2c, Tar, permar, kout);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

size_t big_elim_scratch(int n, int k, int r, size_t es) {
  return ((size_t)n * n + (size_t)k * r + 64) * es + (size_t)(r + 16) * sizeof(int);
}

template <typename S> struct Blas;
template <> struct Blas<float> {
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                             int n, int k, const float* al, const float* A, int lda, const float* B,
                             int ldb, const float* be, float* C, int ldc) {
    return rocblas_sgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static rocblas_status trsm(rocblas_handle h, rocblas_side sd, rocblas_fill up, rocblas_operation ta,
                             rocblas_diagonal dg, int m, int n, const float* al, const float* A,
                             int lda, float* B, int ldb) {
    return rocblas_strsm(h, sd, up, ta, dg, m, n, al, A, lda, B, ldb);
  }
  static rocblas_status getrf(rocblas_handle h, int m, int n, float* A, int lda, int* ipiv, int* info) {
    return rocsolver_sgetrf(h, m, n, A, lda, ipiv, info);
  }
  static rocblas_status potrf(rocblas_handle h, int n, float* A, int lda, int* info) {
    return rocsolver_spotrf(h, rocblas_fill_lower, n, A, lda, info);
  }
  static rocblas_status laswp(rocblas_handle h, int n, float* A, int lda, int k1, int k2, const int* ipiv) {
    return rocsolver_slaswp(h, n, A, lda, k1, k2, ipiv, 1);
  }
};
template <> struct Blas<double> {
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                             int n, int k, const double* al, const double* A, int lda, const double* B,
                             int ldb, const double* be, double* C, int ldc) {
    return rocblas_dgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static rocblas_status trsm(rocblas_handle h, rocblas_side sd, rocblas_fill up, rocblas_operation ta,
                             rocblas_diagonal dg, int m, int n, const double* al, const double* A,
                             int lda, double* B, int ldb) {
    return rocblas_dtrsm(h, sd, up, ta, dg, m, n, al, A, lda, B, ldb);
  }
  static rocblas_status getrf(rocblas_handle h, int m, int n, double* A, int lda, int* ipiv, int* info) {
    return rocsolver_dgetrf(h, m, n, A, lda, ipiv, info);
  }
  static rocblas_status potrf(rocblas_handle h, int n, double* A, int lda, int* info) {
    return rocsolver_dpotrf(h, rocblas_fill_lower, n, A, lda, info);
  }
  static rocblas_status laswp(rocblas_handle h, int n, double* A, int lda, int k1, int k2, const int* ipiv) {
    return rocsolver_dlaswp(h, n, A, lda, k1, k2, ipiv, 1);
  }
};

// one box's elimination (rskelf.m:218-246); lu: 'n' / 's' (real), else 'p'
template <typename S>
int big_elim(bool lu, int b, int n, int k, int r, const LvlBox* boxes, const BoxOut* outs,
             const Seg* csegs, int* pool, const int* permar, const double* Tar, const S* Mprev,
             S* Mcur, S* fval, int* fidx, const PointsDev& P, int* fail, void* scratch,
             hipStream_t st) {
  S* Z = reinterpret_cast<S*>(scratch);
  S* Ts = Z + (size_t)n * n;
  int* ipiv = reinterpret_cast<int*>(Ts + (size_t)k * r + 64);
  int* info = ipiv + r;
  rocblas_handle h = big_handle(st);
  if (!h) return (set_error("FLAM:rskx:blas: rocblas_create_handle failed"), RSKX_ERR_HIP);
  hipLaunchKernelGGL(k_big_kassemble<S>, dim3(n), dim3(256), 0, st, boxes, b, outs, csegs, pool,
                     permar, Mprev, Z, P, 0);
  const S one = 1, mone = -1;
  const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
  if (k > 0) {
    hipLaunchKernelGGL(k_big_tconv<S>, dim3(((int64_t)k * r + 255) / 256 < 65536 ? ((int64_t)k * r + 255) / 256 : 65536),
                       dim3(256), 0, st, boxes, b, outs, Tar, Ts);
    // K(rd,:) -= T' K(sk,:), then K(:,rd) -= K(:,sk) T
    RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, n, k, &mone, Ts, k, Z + r, n, &one, Z, n));
    RSKX_BLAS(Blas<S>::gemm(h, N_, N_, n, r, k, &mone, Z + (size_t)r * n, n, Ts, k, &one, Z, n));
    // the new skeleton block starts from M(sk, sk)
    hipLaunchKernelGGL(k_big_kassemble<S>, dim3(k), dim3(256), 0, st, boxes, b, outs, csegs, pool,
                       permar, Mprev, Z, P, 1);
  }
  if (lu) {
    RSKX_BLAS(Blas<S>::getrf(h, r, r, Z, n, ipiv, info));
    if (k > 0) {
      // E = K(sk,rd) / U, G = L \ K(rd(p), sk), M(sk,sk) -= E G
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_right, rocblas_fill_upper, N_, rocblas_diagonal_non_unit,
                              k, r, &one, Z, n, Z + r, n));
      RSKX_BLAS(Blas<S>::laswp(h, k, Z + (size_t)r * n, n, 1, r, ipiv));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_lower, N_, rocblas_diagonal_unit,
                              r, k, &one, Z, n, Z + (size_t)r * n, n));
      RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, k, r, &mone, Z + r, n, Z + (size_t)r * n, n, &one,
                              Z + r + (size_t)r * n, n));
    }
  } else {
    RSKX_BLAS(Blas<S>::potrf(h, r, Z, n, info));
    if (k > 0) {
      // E = K(sk,rd) / L', M(sk,sk) -= E E'
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_right, rocblas_fill_lower, T_, rocblas_diagonal_non_unit,
                              k, r, &one, Z, n, Z + r, n));
      RSKX_BLAS(Blas<S>::gemm(h, N_, T_, k, k, r, &mone, Z + r, n, Z + r, n, &one,
                              Z + r + (size_t)r * n, n));
    }
  }
  hipLaunchKernelGGL(k_big_info, dim3(1), dim3(1), 0, st, info, lu ? 1 : 0, fail);
  const int64_t tot = (int64_t)r * r + 2 * (int64_t)k * r;
  int grid = (int)((tot + 255) / 256);
  grid = grid < 1 ? 1 : grid > 65536 ? 65536 : grid;
  hipLaunchKernelGGL(k_big_out<S>, dim3(grid), dim3(256), 0, st, boxes, b, outs, pool, permar, Z,
                     Ts, ipiv, fval, fidx, Mcur, lu ? 1 : 0);
  if (lu) hipLaunchKernelGGL(k_big_piv, dim3(1), dim3(256), 0, st, outs, b, ipiv, fidx);
  const int nbk = (r + kDiagBlock - 1) / kDiagBlock;
  hipLaunchKernelGGL(k_big_inv<S>, dim3(2 * nbk), dim3(64), 0, st, outs, b, fval, lu ? 1 : 0);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

template int big_elim<float>(bool, int, int, int, int, const LvlBox*, const BoxOut*, const Seg*,
                             int*, const int*, const double*, const float*, float*, float*, int*,
                             const PointsDev&, int*, void*, hipStream_t);
template int big_elim<double>(bool, int, int, int, int, const LvlBox*, const BoxOut*, const Seg*,
                              int*, const int*, const double*, const double*, double*, double*,
                              int*, const PointsDev&, int*, void*, hipStream_t);

}  // namespace rskx
