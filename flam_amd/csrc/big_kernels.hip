// Boxes too large for one workgroup's registers or shared memory.
//
// ID (Gram route, tol >= 1e-7): boxes beyond the register-resident Gram
// kernels, batched over a level.  K (rskelf.m:201-212: [A(nbr,slf);
// A(slf,nbr)' (two-sided); Kpxy], stored as S) comes from k_assemble; then
//   G = K'K in FP64 on 16 x 16 x 4 FP64 MFMA tiles (k_gw_gram): products of
//     S values are exact in FP64, so G is the Gram matrix of the stored K;
//   the diagonally pivoted Cholesky of G, blocked by 32 pivots: a panel
//     kernel (one workgroup per box) picks each pivot and forms its row of
//     R from G as updated through the previous panels minus the panel's
//     earlier rows, then a trailing update G -= R_p' R_p (k_gw_update);
//     stopped as id.m:105-111 stops the QR;
//   k_gw_finish leaves R's leading rows by pivot position in the box's
//     third T slice, the pivot order and the rank, as k_gram_id does.
// T = R11 \ R12 (id.m:114) then comes from the Gram tail (k_id<S, 2>), or,
// for the largest ranks, from rocBLAS dtrsm with k_id<S, 3> refining.
//
// Elimination (symm 'n', 's' (real) and 'p') of the largest boxes, one at a
// time with the whole GPU: K = A(slf,slf) + M in the order [rd; sk]
// assembled densely, the two sparsification products, LU with partial
// pivoting (rocSOLVER getrf) or Cholesky (potrf), E and G by triangular
// solves and the skeleton Schur update by a GEMM (rocBLAS), exactly
// rskelf.m:218-246; the factor is then written in the layout of the other
// elimination kernels (k_pack adds the transposed copies).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdio>
#include <mutex>

#include "common.hpp"
#include "devutil.hpp"
#include "entries.hpp"
#include "mfma.hpp"
#include "rskx.h"

namespace rskx {

// ------------------------------------------------- the ID of wide boxes --
// per-box state of the blocked pivoted Cholesky, followed in the box's
// scratch by dg (n doubles: remaining diagonal of G) and p2c (n ints:
// position -> column)
struct GwState {
  int k;           // rows of R formed
  int stop;        // set once the rank is decided
  int pad0, pad1;
  double tol_abs;  // tol times the first pivot (id.m:105-111)
  double pad2;
};
constexpr int kGwPanel = 32;   // pivots per panel

// lower 64 x 64 tile t of an nt x nt tile triangle: t = ti (ti + 1) / 2 + tj
__device__ inline void gw_tile(int t, int& ti, int& tj) {
  ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > t) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  tj = t - ti * (ti + 1) / 2;
}

// G = K'K: workgroup (tile, box), four waves of 32 x 32.  MFMA D(j, i):
// lane l holds i = i0 + (l & 15) and rows j = j0 + (l >> 4) + 4 v, so the
// stores of G(i, j) (column-major, leading dimension n) are coalesced.
// Per 16 rows of K the workgroup stages the tile's 64 i and 64 j columns
// in LDS once (one 4-element vector load per thread and half: the host
// aligns K and its leading dimension to 4 elements), double-buffered with
// the next 16 rows loaded while these are multiplied; lane l then reads
// rows 4 (l >> 4) + [0, 4) of its columns as one LDS vector and MFMA c
// multiplies their c-th entries (both operands pair the same row).  Each
// wave reading its own columns from L1 instead had the four waves load
// every column twice and held the kernel at the L1 bandwidth.
template <typename S> struct Vec4;
template <> struct Vec4<float> { typedef float4 T; };
template <> struct Vec4<double> { typedef double4 T; };

constexpr int kGwLd = 20;          // LDS row pitch of a staged column (16 rows + 4 pad)

template <typename S>
__global__ void __launch_bounds__(256) k_gw_gram(const int* __restrict__ ids,
                                                 const LvlBox* __restrict__ boxes,
                                                 const int* __restrict__ mrows,
                                                 const S* __restrict__ Kar,
                                                 double* __restrict__ Tar) {
  typedef typename Vec4<S>::T V4;
  __shared__ __align__(32) S stg[2][128 * kGwLd];
  const int b = ids[blockIdx.y];
  const LvlBox B = boxes[b];
  const int n = B.n, nt = (n + 63) >> 6;
  if ((int)blockIdx.x >= nt * (nt + 1) / 2) return;
  int ti, tj;
  gw_tile(blockIdx.x, ti, tj);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int m = mrows[b], ld = B.mcap;
  const S* K = Kar + B.K_off;
  // staging: thread tid moves column q = (tid + 256 h) >> 2, rows 4 ((tid + 256 h) & 3) + [0, 4);
  // staged columns 0..63 are the tile's i columns, 64..127 its j columns
  // (columns past n read column n - 1; their results are never stored)
  const S* src[2];
  int dst[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + 256 * h, q = e >> 2, part = e & 3;
    const int col = min((q < 64 ? 64 * ti + q : 64 * tj + q - 64), n - 1);
    src[h] = K + (int64_t)col * ld + 4 * part;
    dst[h] = q * kGwLd + 4 * part;
  }
  auto fetch = [&](int r0, V4* v) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int part = ((tid + 256 * h) & 3) * 4;
      if (r0 + 16 <= m) {
        v[h] = *reinterpret_cast<const V4*>(src[h] + r0);
      } else {                             // the last rows, masked
        const S* p = src[h] + r0;
        v[h].x = r0 + part + 0 < m ? p[0] : (S)0;
        v[h].y = r0 + part + 1 < m ? p[1] : (S)0;
        v[h].z = r0 + part + 2 < m ? p[2] : (S)0;
        v[h].w = r0 + part + 3 < m ? p[3] : (S)0;
      }
    }
  };
  auto stash = [&](int sb, const V4* v) {
#pragma unroll
    for (int h = 0; h < 2; ++h) *reinterpret_cast<V4*>(&stg[sb][dst[h]]) = v[h];
  };
  // this lane's operand columns in the staged block
  int qc[4];
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    qc[x] = 32 * (wid >> 1) + 16 * x + (lane & 15);
    qc[2 + x] = 64 + 32 * (wid & 1) + 16 * x + (lane & 15);
  }
  const int rl = 4 * (lane >> 4);
  mf_d4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = mf_d4{0, 0, 0, 0};
  V4 nx[2];
  if (m > 0) {
    fetch(0, nx);
    stash(0, nx);
  }
  __syncthreads();
  int sb = 0;
  for (int r0 = 0; r0 < m; r0 += 16) {
    const bool more = r0 + 16 < m;
    if (more) fetch(r0 + 16, nx);          // in flight while this block is multiplied
    V4 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const V4*>(&stg[sb][qc[q] * kGwLd + rl]);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int x = 0; x < 2; ++x)
          acc[y][x] = Mf<double>::mac((double)a[2 + y][c], (double)a[x][c], acc[y][x]);
    if (more) stash(sb ^ 1, nx);
    __syncthreads();
    sb ^= 1;
  }
  const int i0 = 64 * ti + 32 * (wid >> 1), j0 = 64 * tj + 32 * (wid & 1);
  double* G = Tar + B.T_off;
#pragma unroll
  for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int i = i0 + 16 * x + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = j0 + 16 * y + Mf<double>::row(lane, v);
        if (i < n && j < n) G[i + (int64_t)j * n] = acc[y][x][v];
      }
    }
}

// Panel p of the pivoted Cholesky: pivots 32 p .. 32 p + 31 of one box per
// workgroup.  G holds the lower triangle updated through panel p - 1; R is
// row-major (row t at Rrow + t n) in the second T slice.
template <bool SH>
__global__ void __launch_bounds__(1024) k_gw_panel(const int* __restrict__ ids,
                                                   const LvlBox* __restrict__ boxes,
                                                   const int* __restrict__ mrows,
                                                   double* __restrict__ Tar, double* __restrict__ wsp,
                                                   int64_t wstride, IdParams prm, int p) {
  extern __shared__ double sh[];
  __shared__ double red[2 * 16];
  const int w = blockIdx.x, b = ids[w];
  const LvlBox B = boxes[b];
  const int n = B.n, m = mrows[b], tid = threadIdx.x, nt = blockDim.x;
  GwState* st = reinterpret_cast<GwState*>(wsp + w * wstride);
  double* dgg = wsp + w * wstride + sizeof(GwState) / sizeof(double);
  int* p2cg = reinterpret_cast<int*>(dgg + n);
  if (p > 0 && st->stop) return;
  const double* G = Tar + B.T_off;
  double* Rrow = Tar + B.T_off + (int64_t)n * n;
  // dg and p2c in shared memory, or (boxes too wide for it) in place
  double* dg = SH ? sh : dgg;
  int* p2c = SH ? reinterpret_cast<int*>(sh + n) : p2cg;
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  const int lim = min(min(m, n), kmax);
  double tol_abs = 0;
  if (p == 0) {
    for (int j = tid; j < n; j += nt) { dg[j] = G[j + (int64_t)j * n]; p2c[j] = j; }
  } else {
    if (SH)
      for (int j = tid; j < n; j += nt) { dg[j] = dgg[j]; p2c[j] = p2cg[j]; }
    tol_abs = st->tol_abs;
  }
  __syncthreads();
  const int s0 = kGwPanel * p, s1 = min(s0 + kGwPanel, lim);
  int k = s0 < lim ? s0 : lim, stop = s1 >= lim ? 1 : 0;
  for (int s = s0; s < s1; ++s) {
    double bv = -1.0;
    int bi = n;
    for (int pos = s + tid; pos < n; pos += nt) argmax_merge(bv, bi, dg[p2c[pos]], pos);
    block_argmax(bv, bi, red);
    if (bi >= n) { stop = 1; break; }                 // no finite candidate
    const int pc = p2c[bi];
    const double nrm = sqrt(fmax(bv, 0.0));
    if (s == 0) tol_abs = nrm * prm.tol;
    if (!(nrm > tol_abs)) { stop = 1; break; }
    k = s + 1;
    __syncthreads();                                   // p2c[bi] read by every thread
    if (tid == 0) { p2c[bi] = p2c[s]; p2c[s] = pc; }
    const double rinv = 1.0 / nrm;
    const double* Rpc = Rrow + pc;
    for (int j = tid; j < n; j += nt) {
      double g = j <= pc ? G[pc + (int64_t)j * n] : G[j + (int64_t)pc * n];
      for (int t = s0; t < s; ++t) g -= Rpc[(int64_t)t * n] * Rrow[(int64_t)t * n + j];
      const double r = g * rinv;
      Rrow[(int64_t)s * n + j] = r;
      dg[j] -= r * r;
    }
    __syncthreads();
  }
  __syncthreads();
  if (SH)
    for (int j = tid; j < n; j += nt) { dgg[j] = dg[j]; p2cg[j] = p2c[j]; }
  if (tid == 0) {
    st->k = k;
    st->stop = stop;
    st->tol_abs = tol_abs;
  }
}

// G -= R_p' R_p over the lower 64 x 64 tiles after a complete panel p
// (nothing once the rank is decided: no later panel reads G)
__global__ void __launch_bounds__(256) k_gw_update(const int* __restrict__ ids,
                                                   const LvlBox* __restrict__ boxes,
                                                   double* __restrict__ Tar,
                                                   const double* __restrict__ wsp, int64_t wstride,
                                                   int p) {
  const int w = blockIdx.y;
  const GwState* st = reinterpret_cast<const GwState*>(wsp + w * wstride);
  if (st->stop) return;
  const LvlBox B = boxes[ids[w]];
  const int n = B.n, nt = (n + 63) >> 6;
  if ((int)blockIdx.x >= nt * (nt + 1) / 2) return;
  int ti, tj;
  gw_tile(blockIdx.x, ti, tj);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int i0 = 64 * ti + 32 * (wid >> 1), j0 = 64 * tj + 32 * (wid & 1);
  double* G = Tar + B.T_off;
  const double* R = Tar + B.T_off + (int64_t)n * n + (int64_t)kGwPanel * p * n;
  const int rl = 4 * (lane >> 4);
  mf_d4 acc[2][2];
#pragma unroll
  for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int i = i0 + 16 * x + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = j0 + 16 * y + Mf<double>::row(lane, v);
        acc[y][x][v] = (i < n && j < n) ? G[i + (int64_t)j * n] : 0.0;
      }
    }
  int ci[2], cj[2];
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    ci[x] = min(i0 + 16 * x + (lane & 15), n - 1);
    cj[x] = min(j0 + 16 * x + (lane & 15), n - 1);
  }
#pragma unroll
  for (int r0 = 0; r0 < kGwPanel; r0 += 16) {
    double ai[2][4], aj[2][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double* Rt = R + (int64_t)(r0 + rl + c) * n;
#pragma unroll
      for (int x = 0; x < 2; ++x) { ai[x][c] = Rt[ci[x]]; aj[x][c] = -Rt[cj[x]]; }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int x = 0; x < 2; ++x) acc[y][x] = Mf<double>::mac(aj[y][c], ai[x][c], acc[y][x]);
  }
#pragma unroll
  for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int i = i0 + 16 * x + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = j0 + 16 * y + Mf<double>::row(lane, v);
        if (i < n && j < n) G[i + (int64_t)j * n] = acc[y][x][v];
      }
    }
}

// R's leading k rows by pivot position into the third T slice (leading
// dimension n), the pivot order and the rank; one wave per position
__global__ void __launch_bounds__(256) k_gw_finish(const int* __restrict__ ids,
                                                   const LvlBox* __restrict__ boxes,
                                                   double* __restrict__ Tar,
                                                   const double* __restrict__ wsp, int64_t wstride,
                                                   int* __restrict__ permar, int* __restrict__ kout) {
  const int w = blockIdx.y, b = ids[w];
  const LvlBox B = boxes[b];
  const int n = B.n;
  const int pos = 4 * blockIdx.x + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const GwState* st = reinterpret_cast<const GwState*>(wsp + w * wstride);
  const int* p2c = reinterpret_cast<const int*>(wsp + w * wstride + sizeof(GwState) / sizeof(double) + n);
  const int k = st->k;
  if (blockIdx.x == 0 && threadIdx.x == 0) kout[b] = k;
  if (pos >= n) return;
  const int c = p2c[pos];
  if (lane == 0) permar[B.perm_off + pos] = c;
  const double* Rrow = Tar + B.T_off + (int64_t)n * n;
  double* Rp = Tar + B.T_off + 2 * (int64_t)n * n + (int64_t)pos * n;
  for (int i = lane; i < k; i += 64) Rp[i] = Rrow[(int64_t)i * n + c];
}

// T = R11 \ R12 (k x r, leading dimension n) starts as a copy of R12
__global__ void k_gw_tcopy(const LvlBox* __restrict__ boxes, int b, int k, double* __restrict__ Tar) {
  const LvlBox B = boxes[b];
  const int n = B.n, r = n - k;
  double* T = Tar + B.T_off;
  const double* X = Tar + B.T_off + 2 * (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)k * r;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % k), c = (int)(e / k);
    T[i + (int64_t)c * n] = X[i + (int64_t)(k + c) * n];
  }
}

// ------------------------------------------------------ the elimination --
// A batch of boxes padded to common sizes rM >= r and kM >= k: box w's
// matrix Z (nM x nM, nM = rM + kM, column-major) holds K = A(slf,slf) + M
// in the order [rd, rd padding, sk, sk padding]; the rd padding is an
// identity block (LU and Cholesky of [K(rd,rd) 0; 0 I] leave the real
// block's factors, and no padded row is ever a pivot candidate with a
// nonzero entry), everything else in the padding is zero, so every
// product and solve below runs on the whole batch with uniform sizes.

// padded index -> position in q (rd first, then sk), or -1 for padding
__device__ inline int bigb_pos(int i, int r, int k, int rM) {
  if (i < rM) return i < r ? k + i : -1;
  const int a = i - rM;
  return a < k ? a : -1;
}

// column blockIdx.x of box blockIdx.y's Z; with mo_only only the sk x sk
// block and only M there (the new skeleton block starts from M(sk, sk),
// rskelf.m:246)
template <typename S>
__global__ void __launch_bounds__(256) k_bigb_assemble(const int* __restrict__ ids,
                                                       const LvlBox* __restrict__ boxes,
                                                       const BoxOut* __restrict__ outs,
                                                       const Seg* __restrict__ csegs,
                                                       const int* __restrict__ pool,
                                                       const int* __restrict__ permar,
                                                       const S* __restrict__ Mprev,
                                                       S* __restrict__ Z, PointsDev P, int rM,
                                                       int kM, int mo_only) {
  __shared__ ChildMap cm;
  const int b = ids[blockIdx.y];
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int n = B.n, k = O.k, r = O.r, nM = rM + kM;
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
  const int j = mo_only ? rM + blockIdx.x : blockIdx.x;
  S* Zj = Z + (int64_t)blockIdx.y * nM * nM + (int64_t)j * nM;
  const int pj = bigb_pos(j, r, k, rM);
  const int i0 = mo_only ? rM : 0;
  if (pj < 0) {                                 // padded column
    for (int i = i0 + threadIdx.x; i < nM; i += blockDim.x) Zj[i] = (S)(i == j && j < rM ? 1 : 0);
    return;
  }
  const int sj = q[pj];
  double xj, yj, f2j, f3j, wj;
  point_fields(P, slf[sj], xj, yj, f2j, f3j, wj);
  const double dgj = P.diag[slf[sj]];
  (void)n;
  for (int i = i0 + threadIdx.x; i < nM; i += blockDim.x) {
    const int pi = bigb_pos(i, r, k, rM);
    if (pi < 0) { Zj[i] = (S)0; continue; }
    const int si = q[pi];
    double v = 0;
    if (!mo_only) {
      if (P.kind == RSKX_KERNEL_CALLBACK) {
        v = P.blk[B.A_off + si + (int64_t)sj * B.n];
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

// T of box blockIdx.y (k x r, FP64, leading dimension ldt) in the factor's
// precision, padded to kM x rM with zeros
template <typename S>
__global__ void k_bigb_tconv(const int* __restrict__ ids, const LvlBox* __restrict__ boxes,
                             const BoxOut* __restrict__ outs, const double* __restrict__ Tar,
                             S* __restrict__ Ts, int rM, int kM) {
  const int b = ids[blockIdx.y];
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int k = O.k, r = O.r;
  const double* Tg = Tar + B.T_off;
  S* Tw = Ts + (int64_t)blockIdx.y * kM * rM;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)kM * rM;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(e % kM), c = (int)(e / kM);
    Tw[e] = (a < k && c < r) ? (S)Tg[a + (int64_t)c * B.ldt] : (S)0;
  }
}

// LAPACK's row interchanges of the LU applied to Krs (rows 0..rM, columns
// rM..nM): G = L \ K(rd(p), sk); one thread per column
template <typename S>
__global__ void k_bigb_laswp(S* __restrict__ Z, const int* __restrict__ ipiv, int rM, int kM) {
  const int nM = rM + kM, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kM) return;
  S* col = Z + (int64_t)blockIdx.y * nM * nM + (int64_t)(rM + c) * nM;
  const int* pv = ipiv + (int64_t)blockIdx.y * rM;
  for (int i = 0; i < rM; ++i) {
    const int t = pv[i] - 1;
    if (t != i) { const S a = col[i]; col[i] = col[t]; col[t] = a; }
  }
}

// the factor from the batch's dense results: L and U (or L), E, G, T, the
// new skeleton block, the point lists
template <typename S>
__global__ void k_bigb_out(const int* __restrict__ ids, const LvlBox* __restrict__ boxes,
                           const BoxOut* __restrict__ outs, int* pool,
                           const int* __restrict__ permar, const S* __restrict__ Z,
                           const S* __restrict__ Ts, S* __restrict__ fval, int* __restrict__ fidx,
                           S* __restrict__ Mcur, int lu, int rM, int kM) {
  const int b = ids[blockIdx.y];
  const LvlBox B = boxes[b];
  const BoxOut O = outs[b];
  const int k = O.k, r = O.r, nM = rM + kM;
  const S* Zw = Z + (int64_t)blockIdx.y * nM * nM;
  const S* Tw = Ts + (int64_t)blockIdx.y * kM * rM;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  S* fLU = fval + O.LU;
  S* fE = fval + O.E;
  S* fT = fval + O.T;
  for (int64_t e = tid; e < (int64_t)r * r; e += nt) {
    const int i = (int)(e % r), j = (int)(e / r);
    fLU[e] = (lu || i >= j) ? Zw[i + (int64_t)j * nM] : (S)0;
  }
  for (int64_t e = tid; e < (int64_t)k * r; e += nt) {
    const int a = (int)(e % k), j = (int)(e / k);
    fE[e] = Zw[rM + a + (int64_t)j * nM];
    fT[e] = Tw[a + (int64_t)j * kM];
  }
  if (lu) {
    S* fG = fval + O.G;
    for (int64_t e = tid; e < (int64_t)r * k; e += nt) {
      const int i = (int)(e % r), a = (int)(e / r);
      fG[e] = Zw[i + (int64_t)(rM + a) * nM];
    }
  }
  S* Mo = Mcur + O.M_off;
  for (int64_t e = tid; e < (int64_t)k * k; e += nt) {
    const int a = (int)(e % k), c = (int)(e / k);
    Mo[e] = Zw[rM + a + (int64_t)(rM + c) * nM];
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
__global__ void k_bigb_piv(const int* __restrict__ ids, const BoxOut* __restrict__ outs,
                           const int* __restrict__ ipiv, int* __restrict__ fidx, int rM) {
  const BoxOut O = outs[ids[blockIdx.x]];
  const int r = O.r;
  const int* pv = ipiv + (int64_t)blockIdx.x * rM;
  int* piv = fidx + O.piv;
  for (int i = threadIdx.x; i < r; i += blockDim.x) piv[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < r; ++i) {
      const int t = pv[i] - 1;
      if (t != i && t < r) { const int a = piv[i]; piv[i] = piv[t]; piv[t] = a; }
    }
}

// inverted 32 x 32 diagonal blocks of L (unit for LU) and of U, one wave
// per block: lane c forms column c of the block's inverse by substitution
// from the factor just written
template <typename S>
__global__ void __launch_bounds__(64) k_bigb_inv(const int* __restrict__ ids,
                                                 const BoxOut* __restrict__ outs,
                                                 S* __restrict__ fval, int lu) {
  const BoxOut O = outs[ids[blockIdx.y]];
  const int r = O.r, nbk = (r + kDiagBlock - 1) / kDiagBlock;
  const int bb = blockIdx.x;
  if (bb >= 2 * nbk) return;
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
__global__ void k_bigb_info(const int* __restrict__ info, int nbx, int lu, int* __restrict__ fail) {
  for (int w = threadIdx.x; w < nbx; w += blockDim.x)
    if (info[w] > 0) atomicOr(fail, lu ? 2 : 1);
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

// scratch of the wide-box ID per box (doubles): state, dg, p2c
int64_t gram_wide_stride(int nmax) { return (int64_t)(sizeof(GwState) / sizeof(double)) + nmax + (nmax + 1) / 2 + 2; }

int gram_wide_rounds(int nmax) { return (nmax + kGwPanel - 1) / kGwPanel; }

// G = K'K for the nw boxes of ids (device list), K assembled
template <typename S>
void gram_wide_gram(const int* ids, int nw, int nmax, const LvlBox* boxes, const int* mrows,
                    const S* Kar, double* Tar, hipStream_t st) {
  const int nt = (nmax + 63) / 64, ntile = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_gw_gram<S>, dim3(ntile, nw), dim3(256), 0, st, ids, boxes, mrows, Kar, Tar);
}

// the blocked pivoted Cholesky of every box's G and R by position; no host
// synchronization
int gram_wide_chol(const int* ids, int nw, int nmax, const LvlBox* boxes, const int* mrows,
                   double* Tar, int* permar, int* kout, const IdParams& prm, double* wsp,
                   hipStream_t st) {
  const int nt = (nmax + 63) / 64, ntile = nt * (nt + 1) / 2;
  const int64_t wstride = gram_wide_stride(nmax);
  const size_t shm = (size_t)nmax * sizeof(double) + (size_t)nmax * sizeof(int) + 16;
  const bool sh = shm <= 160 * 1024;
  if (sh) ensure_dyn_lds((const void*)k_gw_panel<true>, shm);
  const int rounds = gram_wide_rounds(prm.kmax > 0 ? std::min(nmax, prm.kmax) : nmax);
  for (int p = 0; p < rounds; ++p) {
    if (sh)
      hipLaunchKernelGGL(k_gw_panel<true>, dim3(nw), dim3(1024), shm, st, ids, boxes, mrows, Tar,
                         wsp, wstride, prm, p);
    else
      hipLaunchKernelGGL(k_gw_panel<false>, dim3(nw), dim3(1024), 0, st, ids, boxes, mrows, Tar,
                         wsp, wstride, prm, p);
    if (p + 1 < rounds)
      hipLaunchKernelGGL(k_gw_update, dim3(ntile, nw), dim3(256), 0, st, ids, boxes, Tar, wsp,
                         wstride, p);
  }
  hipLaunchKernelGGL(k_gw_finish, dim3((nmax + 3) / 4, nw), dim3(256), 0, st, ids, boxes, Tar, wsp,
                     wstride, permar, kout);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

// T = R11 \ R12 and inv(R11) for one wide box of rank k with the whole GPU
// (FP64)
int gram_wide_tsolve(const LvlBox* boxes, int b, const LvlBox& hbox, int k, double* Tar,
                     hipStream_t st) {
  const int n = hbox.n, r = n - k;
  if (k <= 0 || r <= 0) return RSKX_OK;
  rocblas_handle h = big_handle(st);
  if (!h) return (set_error("FLAM:rskx:blas: rocblas_create_handle failed"), RSKX_ERR_HIP);
  const int64_t e = (int64_t)k * r;
  const int grid = (int)std::min<int64_t>((e + 255) / 256, 65536);
  hipLaunchKernelGGL(k_gw_tcopy, dim3(grid), dim3(256), 0, st, boxes, b, k, Tar);
  double* T = Tar + hbox.T_off;
  const double* X = Tar + hbox.T_off + 2 * (int64_t)n * n;
  const double one = 1.0;
  RSKX_BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, k, r, &one, X, n, T, n));
  // inv(R11) into the second slice for the refinement's inverse row norms
  double* Ri = Tar + hbox.T_off + (int64_t)n * n;
  RSKX_BLAS(rocblas_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, k, X, n, Ri, n));
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

template void gram_wide_gram<float>(const int*, int, int, const LvlBox*, const int*, const float*,
                                    double*, hipStream_t);
template void gram_wide_gram<double>(const int*, int, int, const LvlBox*, const int*,
                                     const double*, double*, hipStream_t);


template <typename S> struct Blas;
template <> struct Blas<float> {
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                             int n, int k, const float* al, const float* A, int lda, const float* B,
                             int ldb, const float* be, float* C, int ldc) {
    return rocblas_sgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                                int n, int k, const float* al, const float* A, int lda, int64_t sa,
                                const float* B, int ldb, int64_t sb, const float* be, float* C,
                                int ldc, int64_t sc, int cnt) {
    return rocblas_sgemm_strided_batched(h, ta, tb, m, n, k, al, A, lda, sa, B, ldb, sb, be, C, ldc,
                                         sc, cnt);
  }
  static rocblas_status trsm(rocblas_handle h, rocblas_side sd, rocblas_fill up, rocblas_operation ta,
                             rocblas_diagonal dg, int m, int n, const float* al, const float* A,
                             int lda, float* B, int ldb) {
    return rocblas_strsm(h, sd, up, ta, dg, m, n, al, A, lda, B, ldb);
  }
  static rocblas_status trsm_sb(rocblas_handle h, rocblas_side sd, rocblas_fill up,
                                rocblas_operation ta, rocblas_diagonal dg, int m, int n,
                                const float* al, const float* A, int lda, int64_t sa, float* B,
                                int ldb, int64_t sb, int cnt) {
    return rocblas_strsm_strided_batched(h, sd, up, ta, dg, m, n, al, A, lda, sa, B, ldb, sb, cnt);
  }
  static rocblas_status getrf_sb(rocblas_handle h, int n, float* A, int lda, int64_t sa, int* ipiv,
                                 int64_t sp, int* info, int cnt) {
    return rocsolver_sgetrf_strided_batched(h, n, n, A, lda, sa, ipiv, sp, info, cnt);
  }
  static rocblas_status potrf_sb(rocblas_handle h, int n, float* A, int lda, int64_t sa, int* info,
                                 int cnt) {
    return rocsolver_spotrf_strided_batched(h, rocblas_fill_lower, n, A, lda, sa, info, cnt);
  }
};
template <> struct Blas<double> {
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                             int n, int k, const double* al, const double* A, int lda,
                             const double* B, int ldb, const double* be, double* C, int ldc) {
    return rocblas_dgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                                int n, int k, const double* al, const double* A, int lda,
                                int64_t sa, const double* B, int ldb, int64_t sb, const double* be,
                                double* C, int ldc, int64_t sc, int cnt) {
    return rocblas_dgemm_strided_batched(h, ta, tb, m, n, k, al, A, lda, sa, B, ldb, sb, be, C, ldc,
                                         sc, cnt);
  }
  static rocblas_status trsm(rocblas_handle h, rocblas_side sd, rocblas_fill up, rocblas_operation ta,
                             rocblas_diagonal dg, int m, int n, const double* al, const double* A,
                             int lda, double* B, int ldb) {
    return rocblas_dtrsm(h, sd, up, ta, dg, m, n, al, A, lda, B, ldb);
  }
  static rocblas_status trsm_sb(rocblas_handle h, rocblas_side sd, rocblas_fill up,
                                rocblas_operation ta, rocblas_diagonal dg, int m, int n,
                                const double* al, const double* A, int lda, int64_t sa, double* B,
                                int ldb, int64_t sb, int cnt) {
    return rocblas_dtrsm_strided_batched(h, sd, up, ta, dg, m, n, al, A, lda, sa, B, ldb, sb, cnt);
  }
  static rocblas_status getrf_sb(rocblas_handle h, int n, double* A, int lda, int64_t sa, int* ipiv,
                                 int64_t sp, int* info, int cnt) {
    return rocsolver_dgetrf_strided_batched(h, n, n, A, lda, sa, ipiv, sp, info, cnt);
  }
  static rocblas_status potrf_sb(rocblas_handle h, int n, double* A, int lda, int64_t sa, int* info,
                                 int cnt) {
    return rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, n, A, lda, sa, info, cnt);
  }
};

// scratch of one batch (bytes): Z, padded T, pivots, info
size_t big_elim_scratch(int nbx, int rM, int kM, size_t es) {
  const size_t nM = (size_t)rM + kM;
  return ((size_t)nbx * (nM * nM + (size_t)kM * rM) + 64) * es +
         ((size_t)nbx * (rM + 1) + 64) * sizeof(int);
}

// the elimination of a batch of boxes (rskelf.m:218-246) with the whole
// GPU; lu: 'n' / 's' (real), else 'p'.  ids: the batch's box indices on the
// device; rM, kM: the largest r and k in the batch
template <typename S>
int big_elim(bool lu, const int* ids, int nbx, int rM, int kM, const LvlBox* boxes,
             const BoxOut* outs, const Seg* csegs, int* pool, const int* permar, const double* Tar,
             const S* Mprev, S* Mcur, S* fval, int* fidx, const PointsDev& P, int* fail,
             void* scratch, hipStream_t st) {
  const int nM = rM + kM;
  const int64_t sZ = (int64_t)nM * nM, sT = (int64_t)kM * rM;
  S* Z = reinterpret_cast<S*>(scratch);
  S* Ts = Z + (size_t)nbx * sZ;
  int* ipiv = reinterpret_cast<int*>(Ts + (size_t)nbx * sT + 64);
  int* info = ipiv + (size_t)nbx * rM;
  rocblas_handle h = big_handle(st);
  if (!h) return (set_error("FLAM:rskx:blas: rocblas_create_handle failed"), RSKX_ERR_HIP);
  hipLaunchKernelGGL(k_bigb_assemble<S>, dim3(nM, nbx), dim3(256), 0, st, ids, boxes, outs, csegs,
                     pool, permar, Mprev, Z, P, rM, kM, 0);
  const S one = 1, mone = -1;
  const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
  auto grid1 = [](int64_t e) { return (int)std::max<int64_t>(1, std::min<int64_t>((e + 255) / 256, 4096)); };
  if (kM > 0) {
    hipLaunchKernelGGL(k_bigb_tconv<S>, dim3(grid1(sT), nbx), dim3(256), 0, st, ids, boxes, outs, Tar,
                       Ts, rM, kM);
    // K(rd,:) -= T' K(sk,:), then K(:,rd) -= K(:,sk) T
    RSKX_BLAS(Blas<S>::gemm_sb(h, T_, N_, rM, nM, kM, &mone, Ts, kM, sT, Z + rM, nM, sZ, &one, Z, nM,
                               sZ, nbx));
    RSKX_BLAS(Blas<S>::gemm_sb(h, N_, N_, nM, rM, kM, &mone, Z + (size_t)rM * nM, nM, sZ, Ts, kM, sT,
                               &one, Z, nM, sZ, nbx));
    // the new skeleton block starts from M(sk, sk)
    hipLaunchKernelGGL(k_bigb_assemble<S>, dim3(kM, nbx), dim3(256), 0, st, ids, boxes, outs, csegs,
                       pool, permar, Mprev, Z, P, rM, kM, 1);
  }
  if (lu) {
    RSKX_BLAS(Blas<S>::getrf_sb(h, rM, Z, nM, sZ, ipiv, rM, info, nbx));
    if (kM > 0) {
      // E = K(sk,rd) / U, G = L \ K(rd(p), sk), M(sk,sk) -= E G
      RSKX_BLAS(Blas<S>::trsm_sb(h, rocblas_side_right, rocblas_fill_upper, N_,
                                 rocblas_diagonal_non_unit, kM, rM, &one, Z, nM, sZ, Z + rM, nM, sZ,
                                 nbx));
      hipLaunchKernelGGL(k_bigb_laswp<S>, dim3((kM + 255) / 256, nbx), dim3(256), 0, st, Z, ipiv, rM,
                         kM);
      RSKX_BLAS(Blas<S>::trsm_sb(h, rocblas_side_left, rocblas_fill_lower, N_, rocblas_diagonal_unit,
                                 rM, kM, &one, Z, nM, sZ, Z + (size_t)rM * nM, nM, sZ, nbx));
      RSKX_BLAS(Blas<S>::gemm_sb(h, N_, N_, kM, kM, rM, &mone, Z + rM, nM, sZ, Z + (size_t)rM * nM, nM,
                                 sZ, &one, Z + rM + (size_t)rM * nM, nM, sZ, nbx));
    }
  } else {
    RSKX_BLAS(Blas<S>::potrf_sb(h, rM, Z, nM, sZ, info, nbx));
    if (kM > 0) {
      // E = K(sk,rd) / L', M(sk,sk) -= E E'
      RSKX_BLAS(Blas<S>::trsm_sb(h, rocblas_side_right, rocblas_fill_lower, T_,
                                 rocblas_diagonal_non_unit, kM, rM, &one, Z, nM, sZ, Z + rM, nM, sZ,
                                 nbx));
      RSKX_BLAS(Blas<S>::gemm_sb(h, N_, T_, kM, kM, rM, &mone, Z + rM, nM, sZ, Z + rM, nM, sZ, &one,
                                 Z + rM + (size_t)rM * nM, nM, sZ, nbx));
    }
  }
  hipLaunchKernelGGL(k_bigb_info, dim3(1), dim3(256), 0, st, info, nbx, lu ? 1 : 0, fail);
  hipLaunchKernelGGL(k_bigb_out<S>, dim3(grid1((int64_t)rM * rM + 2 * (int64_t)kM * rM), nbx),
                     dim3(256), 0, st, ids, boxes, outs, pool, permar, Z, Ts, fval, fidx, Mcur,
                     lu ? 1 : 0, rM, kM);
  if (lu) hipLaunchKernelGGL(k_bigb_piv, dim3(nbx), dim3(256), 0, st, ids, outs, ipiv, fidx, rM);
  const int nbk = (rM + kDiagBlock - 1) / kDiagBlock;
  hipLaunchKernelGGL(k_bigb_inv<S>, dim3(2 * nbk, nbx), dim3(64), 0, st, ids, outs, fval, lu ? 1 : 0);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}


// ------------------------------------------------ the sweeps' largest --
// X(rows, :) (tree order, leading dimension ldx) <-> a dense column-major
// buffer (leading dimension nrow); perm (optional) maps buffer row i to
// rows[perm[i]] (gather) or back (scatter)
template <typename S>
__global__ void k_big_rows(S* __restrict__ X, int64_t ldx, const int* __restrict__ rows, int nrow,
                           int ncol, S* __restrict__ buf, int to_buf) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)nrow * ncol;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nrow), c = (int)(e / nrow);
    S* x = X + rows[i] + (int64_t)c * ldx;
    if (to_buf) buf[e] = *x;
    else *x = buf[e];
  }
}

// xr = xr(piv) (gather) or xr(piv) = xr (scatter), through tmp
template <typename S>
__global__ void k_big_perm(S* __restrict__ xr, S* __restrict__ tmp, const int* __restrict__ piv,
                           int r, int ncol, int gather) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)r * ncol;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % r);
    const int64_t c = e / r;
    if (gather) tmp[e] = xr[piv[i] + c * r];
    else tmp[piv[i] + c * r] = xr[e];
  }
}

size_t big_solve_scratch(int k, int r, int ncol) {
  return (size_t)(k + 2 * (size_t)r) * ncol;
}

// One factor's part of a solve sweep with library calls (the operations of
// k_solve for programs 0 'n', 1 transposed 'n', 2 'p'): sv_nn.m, sv_nc.m,
// sv_p.m.  The factor's rows of X are gathered into dense blocks, updated
// by GEMMs and triangular solves, and scattered back.
template <typename S>
int big_solve(int prog, bool up, const FacRec& f, const S* fval, const int* fidx, S* X,
              int64_t ldx, int ncol, S* scratch, hipStream_t st) {
  const int k = f.k, r = f.r;
  rocblas_handle h = big_handle(st);
  if (!h) return (set_error("FLAM:rskx:blas: rocblas_create_handle failed"), RSKX_ERR_HIP);
  S* xs = scratch;
  S* xr = xs + (size_t)k * ncol;
  S* tmp = xr + (size_t)r * ncol;
  const int* sk = fidx + f.sk;
  const int* rd = fidx + f.rd;
  const int* piv = fidx + f.piv;
  auto grid = [](int64_t e) { return (int)std::max<int64_t>(1, std::min<int64_t>((e + 255) / 256, 65536)); };
  if (k > 0)
    hipLaunchKernelGGL(k_big_rows<S>, dim3(grid((int64_t)k * ncol)), dim3(256), 0, st, X, ldx, sk, k,
                       ncol, xs, 1);
  hipLaunchKernelGGL(k_big_rows<S>, dim3(grid((int64_t)r * ncol)), dim3(256), 0, st, X, ldx, rd, r,
                     ncol, xr, 1);
  const S one = 1, mone = -1;
  const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
  const S* T = fval + f.T;     // k x r
  const S* LU = fval + f.LU;   // r x r
  const S* E = fval + f.E;     // k x r
  const S* G = fval + f.G;     // r x k
  auto perm = [&](int gather) {
    hipLaunchKernelGGL(k_big_perm<S>, dim3(grid((int64_t)r * ncol)), dim3(256), 0, st, xr, tmp, piv,
                       r, ncol, gather);
    RSKX_HIP(hipMemcpyAsync(xr, tmp, (size_t)r * ncol * sizeof(S), hipMemcpyDeviceToDevice, st));
    return RSKX_OK;
  };
  if (prog == 0) {
    if (up) {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, ncol, k, &mone, T, k, xs, k, &one, xr, r));
      if (const int rc = perm(1)) return rc;
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_lower, N_, rocblas_diagonal_unit,
                              r, ncol, &one, LU, r, xr, r));
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, ncol, r, &mone, E, k, xr, r, &one, xs, k));
    } else {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, r, ncol, k, &mone, G, r, xs, k, &one, xr, r));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_upper, N_,
                              rocblas_diagonal_non_unit, r, ncol, &one, LU, r, xr, r));
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, ncol, r, &mone, T, k, xr, r, &one, xs, k));
    }
  } else if (prog == 1) {
    if (up) {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, ncol, k, &mone, T, k, xs, k, &one, xr, r));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_upper, T_,
                              rocblas_diagonal_non_unit, r, ncol, &one, LU, r, xr, r));
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, k, ncol, r, &mone, G, r, xr, r, &one, xs, k));
    } else {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, ncol, k, &mone, E, k, xs, k, &one, xr, r));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_lower, T_, rocblas_diagonal_unit,
                              r, ncol, &one, LU, r, xr, r));
      if (const int rc = perm(0)) return rc;
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, ncol, r, &mone, T, k, xr, r, &one, xs, k));
    }
  } else {
    if (up) {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, ncol, k, &mone, T, k, xs, k, &one, xr, r));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_lower, N_,
                              rocblas_diagonal_non_unit, r, ncol, &one, LU, r, xr, r));
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, ncol, r, &mone, E, k, xr, r, &one, xs, k));
    } else {
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, T_, N_, r, ncol, k, &mone, E, k, xs, k, &one, xr, r));
      RSKX_BLAS(Blas<S>::trsm(h, rocblas_side_left, rocblas_fill_lower, T_,
                              rocblas_diagonal_non_unit, r, ncol, &one, LU, r, xr, r));
      if (k > 0) RSKX_BLAS(Blas<S>::gemm(h, N_, N_, k, ncol, r, &mone, T, k, xr, r, &one, xs, k));
    }
  }
  if (k > 0)
    hipLaunchKernelGGL(k_big_rows<S>, dim3(grid((int64_t)k * ncol)), dim3(256), 0, st, X, ldx, sk, k,
                       ncol, xs, 0);
  hipLaunchKernelGGL(k_big_rows<S>, dim3(grid((int64_t)r * ncol)), dim3(256), 0, st, X, ldx, rd, r,
                     ncol, xr, 0);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

template int big_solve<float>(int, bool, const FacRec&, const float*, const int*, float*, int64_t,
                              int, float*, hipStream_t);
template int big_solve<double>(int, bool, const FacRec&, const double*, const int*, double*,
                               int64_t, int, double*, hipStream_t);

template int big_elim<float>(bool, const int*, int, int, int, const LvlBox*, const BoxOut*,
                             const Seg*, int*, const int*, const double*, const float*, float*,
                             float*, int*, const PointsDev&, int*, void*, hipStream_t);
template int big_elim<double>(bool, const int*, int, int, int, const LvlBox*, const BoxOut*,
                              const Seg*, int*, const int*, const double*, const double*, double*,
                              double*, int*, const PointsDev&, int*, void*, hipStream_t);

}  // namespace rskx
