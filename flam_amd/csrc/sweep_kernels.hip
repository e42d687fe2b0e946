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

// Right-hand sides are processed in blocks of kRhs columns (the host pads
// the work copy of X with zero columns), so every loop over them is
// unrolled at compile time.
constexpr int kRhs = 8;

// Y (ny x kRhs) += sgn * op(A) Z, op(A)(i,l) = A[i + l lda] or A[l + i lda].
// When there are fewer rows than threads, each row's sum over l is split
// between P threads and combined through `red` (P * ny * kRhs elements).
template <typename S>
__device__ void upd(S* Y, int ny, const S* Z, int nz, const S* __restrict__ A, int lda, bool trA,
                    S sgn, S* red) {
  const int nt = blockDim.x, tid = threadIdx.x;
  if (ny == 0 || nz == 0) return;
  int P = nt / ny;
  P = P < 1 ? 1 : (P > 16 ? 16 : P);
  if (P == 1) {
    for (int i = tid; i < ny; i += nt) {
      S acc[kRhs] = {};
      for (int l = 0; l < nz; ++l) {
        const S a = trA ? A[l + (int64_t)i * lda] : A[i + (int64_t)l * lda];
#pragma unroll
        for (int c = 0; c < kRhs; ++c) acc[c] += a * Z[l + c * nz];
      }
#pragma unroll
      for (int c = 0; c < kRhs; ++c) Y[i + c * ny] += sgn * acc[c];
    }
    __syncthreads();
    return;
  }
  if (tid < ny * P) {
    const int i = tid % ny, part = tid / ny;
    S acc[kRhs] = {};
    for (int l = part; l < nz; l += P) {
      const S a = trA ? A[l + (int64_t)i * lda] : A[i + (int64_t)l * lda];
#pragma unroll
      for (int c = 0; c < kRhs; ++c) acc[c] += a * Z[l + c * nz];
    }
#pragma unroll
    for (int c = 0; c < kRhs; ++c) red[(part * kRhs + c) * ny + i] = acc[c];
  }
  __syncthreads();
  for (int o = tid; o < ny * kRhs; o += nt) {
    S sum = 0;
    for (int part = 0; part < P; ++part) sum += red[part * kRhs * ny + o];
    Y[o] += sgn * sum;
  }
  __syncthreads();
}

// In-place blocked triangular solve with the r x r triangle of A (leading
// dimension r) and its inverted 32 x 32 diagonal blocks Ainv:
//   lower/trans false: L y = b ; lower/trans true: L' y = b
//   upper/trans false: U y = b ; upper/trans true: U' y = b
// Per block: y_b = inv(A_bb) y_b as a 32 x kRhs product (one element per
// thread), then the block's contribution is subtracted from the rows still
// to be solved.
template <typename S>
__device__ void trsolve(S* Y, int r, const S* __restrict__ A, const S* __restrict__ Ainv,
                        bool lower, bool tr) {
  const bool fwd = (lower && !tr) || (!lower && tr);
  auto coef = [&](int i, int t) -> S { return tr ? A[t + (int64_t)i * r] : A[i + (int64_t)t * r]; };
  const int nbk = (r + kDiagBlock - 1) / kDiagBlock;
  const int tid = threadIdx.x;
  const int bi = tid & (kDiagBlock - 1), bc = tid / kDiagBlock;   // block row, rhs column
  for (int bk = 0; bk < nbk; ++bk) {
    const int b = fwd ? bk : nbk - 1 - bk;
    const int j0 = b * kDiagBlock, bs = r - j0 < kDiagBlock ? r - j0 : kDiagBlock;
    const S* Ib = Ainv + (int64_t)b * kDiagBlock * kDiagBlock;
    S acc = 0;
    const bool act = bc < kRhs && bi < bs;
    if (act) {
      S iv[kDiagBlock];
#pragma unroll
      for (int u = 0; u < kDiagBlock; ++u)
        iv[u] = u < bs ? (tr ? Ib[u + bi * kDiagBlock] : Ib[bi + u * kDiagBlock]) : (S)0;
#pragma unroll
      for (int u = 0; u < kDiagBlock; ++u)
        if (u < bs) acc += iv[u] * Y[j0 + u + bc * r];
    }
    __syncthreads();
    if (act) Y[j0 + bi + bc * r] = acc;
    __syncthreads();
    const int lo = fwd ? j0 + bs : 0, hi = fwd ? r : j0;
    for (int i = lo + tid; i < hi; i += blockDim.x) {
      S at[kDiagBlock];
#pragma unroll
      for (int t = 0; t < kDiagBlock; ++t) at[t] = t < bs ? coef(i, j0 + t) : (S)0;
      S a2[kRhs] = {};
#pragma unroll
      for (int t = 0; t < kDiagBlock; ++t)
        if (t < bs) {
#pragma unroll
          for (int c = 0; c < kRhs; ++c) a2[c] += at[t] * Y[j0 + t + c * r];
        }
#pragma unroll
      for (int c = 0; c < kRhs; ++c) Y[i + c * r] -= a2[c];
    }
    __syncthreads();
  }
}

// Y = op(tri(A)) Y, the sum over l split between threads as in upd
template <typename S>
__device__ void trmul(S* Y, S* tmp, int r, const S* __restrict__ A, bool lower, bool unit, bool tr,
                      S* red) {
  const bool low = lower != tr;   // effective lower-triangular operator
  const int nt = blockDim.x, tid = threadIdx.x;
  int P = nt / (r > 0 ? r : 1);
  P = P < 1 ? 1 : (P > 16 ? 16 : P);
  for (int t0 = tid; t0 < r * P; t0 += nt) {
    const int i = t0 % r, part = t0 / r;
    const int l0 = low ? 0 : (unit ? i + 1 : i);
    const int l1 = low ? (unit ? i : i + 1) : r;
    S acc[kRhs] = {};
    for (int l = l0 + part; l < l1; l += P) {
      const S a = tr ? A[l + (int64_t)i * r] : A[i + (int64_t)l * r];
#pragma unroll
      for (int c = 0; c < kRhs; ++c) acc[c] += a * Y[l + c * r];
    }
    if (unit && part == 0) {
#pragma unroll
      for (int c = 0; c < kRhs; ++c) acc[c] += Y[i + c * r];
    }
    // one part per row (r >= threads): straight into tmp; else partials
    S* dst = P == 1 ? tmp : red + part * kRhs * r;
#pragma unroll
    for (int c = 0; c < kRhs; ++c) dst[c * r + i] = acc[c];
  }
  __syncthreads();
  if (P > 1) {
    for (int o = tid; o < r * kRhs; o += nt) {
      S sum = 0;
      for (int part = 0; part < P; ++part) sum += red[part * kRhs * r + o];
      tmp[o] = sum;
    }
    __syncthreads();
  }
  for (int o = tid; o < r * kRhs; o += nt) Y[o] = tmp[o];
  __syncthreads();
}

// gather (tmp(i) = Y(p(i))) or scatter (tmp(p(i)) = Y(i)), then copy back
template <typename S>
__device__ void permrows(S* Y, S* tmp, int r, const int* __restrict__ p, bool gather) {
  for (int i = threadIdx.x; i < r; i += blockDim.x) {
    const int pi = p[i];
#pragma unroll
    for (int c = 0; c < kRhs; ++c) {
      if (gather) tmp[i + c * r] = Y[pi + c * r];
      else tmp[pi + c * r] = Y[i + c * r];
    }
  }
  __syncthreads();
  for (int o = threadIdx.x; o < r * kRhs; o += blockDim.x) Y[o] = tmp[o];
  __syncthreads();
}

// Copy n contiguous elements global -> shared with eight independent
// loads in flight per thread.
template <typename S>
__device__ void bulk_copy(S* __restrict__ dst, const S* __restrict__ src, int n) {
  const int nt = blockDim.x;
  int e = threadIdx.x;
  for (; e + 7 * nt < n; e += 8 * nt) {
    S v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[e + u * nt];
#pragma unroll
    for (int u = 0; u < 8; ++u) dst[e + u * nt] = v[u];
  }
  for (; e < n; e += nt) dst[e] = src[e];
}

// PROG: 0 sv 'n'/'n', 1 sv 'n'/'c', 2 sv 'p', 3 mv 'n'/'n', 4 mv 'n'/'c', 5 mv 'p'
// Shared memory: xs (k x 8), xr (r x 8), tmp (r x 8), red (8 x 256), then
// the stage buffer (bcap elements).  X has nrhs columns, a multiple of 8.
template <typename S, int PROG, bool UP, bool GWS>
__global__ void __launch_bounds__(256) k_sweep(const FacRec* __restrict__ recs,
                                               const S* __restrict__ fval,
                                               const int* __restrict__ fidx, S* X, int64_t ldx,
                                               int nrhs, int kmax, int rmax, int bcap, S* gws,
                                               int64_t gstride) {
  extern __shared__ char smem_raw[];
  // rows of X and the work vectors live in shared memory, or for factors
  // too large for it (GWS) in this workgroup's slice of a global workspace;
  // a compile-time choice, so the shared-memory variant keeps LDS accesses
  S* sm = GWS ? gws + blockIdx.x * gstride : reinterpret_cast<S*>(smem_raw);
  const FacRec f = recs[blockIdx.x];
  const int k = f.k, r = f.r;
  const int* sk = fidx + f.sk;
  const int* rd = fidx + f.rd;
  const int* piv = fidx + f.piv;
  const S* Tg = fval + f.T;
  const S* LUg = fval + f.LU;
  const S* Eg = fval + f.E;
  const S* Gg = fval + f.G;
  const S* Li = fval + f.Li;
  const S* Ui = fval + f.Ui;
  S* xs = sm;
  S* xr = xs + (size_t)kmax * kRhs;
  S* tmp = xr + (size_t)rmax * kRhs;
  S* red = tmp + (size_t)rmax * kRhs;
  S* buf = red + (size_t)kRhs * 256;
  // a factor block is staged in shared memory (one bulk round trip) when it
  // fits in the bcap-element buffer
  auto stage = [&](const S* g, int cnt) -> const S* {
    if (cnt > bcap) return g;
    __syncthreads();
    bulk_copy(buf, g, cnt);
    __syncthreads();
    return buf;
  };
  const int kr = k * r, rr = r * r;
  const S one = 1, mone = -1;
  for (int c0 = 0; c0 < nrhs; c0 += kRhs) {
    // gather X(sk,:), X(rd,:): one index load, then all columns in flight
    for (int a = threadIdx.x; a < k + r; a += blockDim.x) {
      const bool isk = a < k;
      const int64_t row = isk ? sk[a] : rd[a - k];
      const S* src = X + row + c0 * ldx;
      S v[kRhs];
#pragma unroll
      for (int c = 0; c < kRhs; ++c) v[c] = src[c * ldx];
      S* dst = isk ? xs + a : xr + (a - k);
      const int ld = isk ? k : r;
#pragma unroll
      for (int c = 0; c < kRhs; ++c) dst[c * ld] = v[c];
    }
    __syncthreads();
    if (PROG == 0) {            // sv_nn.m
      if (UP) {
        upd(xr, r, xs, k, stage(Tg, kr), k, true, mone, red);
        permrows(xr, tmp, r, piv, true);
        trsolve(xr, r, stage(LUg, rr), Li, true, false);
        upd(xs, k, xr, r, stage(Eg, kr), k, false, mone, red);
      } else {
        upd(xr, r, xs, k, stage(Gg, kr), r, false, mone, red);
        trsolve(xr, r, stage(LUg, rr), Ui, false, false);
        upd(xs, k, xr, r, stage(Tg, kr), k, false, mone, red);
      }
    } else if (PROG == 1) {     // sv_nc.m
      if (UP) {
        upd(xr, r, xs, k, stage(Tg, kr), k, true, mone, red);
        trsolve(xr, r, stage(LUg, rr), Ui, false, true);
        upd(xs, k, xr, r, stage(Gg, kr), r, true, mone, red);
      } else {
        upd(xr, r, xs, k, stage(Eg, kr), k, true, mone, red);
        trsolve(xr, r, stage(LUg, rr), Li, true, true);
        permrows(xr, tmp, r, piv, false);
        upd(xs, k, xr, r, stage(Tg, kr), k, false, mone, red);
      }
    } else if (PROG == 2) {     // sv_p.m
      if (UP) {
        upd(xr, r, xs, k, stage(Tg, kr), k, true, mone, red);
        trsolve(xr, r, stage(LUg, rr), Li, true, false);
        upd(xs, k, xr, r, stage(Eg, kr), k, false, mone, red);
      } else {
        upd(xr, r, xs, k, stage(Eg, kr), k, true, mone, red);
        trsolve(xr, r, stage(LUg, rr), Li, true, true);
        upd(xs, k, xr, r, stage(Tg, kr), k, false, mone, red);
      }
    } else if (PROG == 3) {     // mv_nn.m
      if (UP) {
        upd(xs, k, xr, r, stage(Tg, kr), k, false, one, red);
        trmul(xr, tmp, r, stage(LUg, rr), false, false, false, red);
        upd(xr, r, xs, k, stage(Gg, kr), r, false, one, red);
      } else {
        upd(xs, k, xr, r, stage(Eg, kr), k, false, one, red);
        trmul(xr, tmp, r, stage(LUg, rr), true, true, false, red);
        permrows(xr, tmp, r, piv, false);
        upd(xr, r, xs, k, stage(Tg, kr), k, true, one, red);
      }
    } else if (PROG == 4) {     // mv_nc.m
      if (UP) {
        upd(xs, k, xr, r, stage(Tg, kr), k, false, one, red);
        permrows(xr, tmp, r, piv, true);
        trmul(xr, tmp, r, stage(LUg, rr), true, true, true, red);
        upd(xr, r, xs, k, stage(Eg, kr), k, true, one, red);
      } else {
        upd(xs, k, xr, r, stage(Gg, kr), r, true, one, red);
        trmul(xr, tmp, r, stage(LUg, rr), false, false, true, red);
        upd(xr, r, xs, k, stage(Tg, kr), k, true, one, red);
      }
    } else {                    // mv_p.m; mv_h.m (unit L, then X(rd) = D X(rd))
      const bool unit = f.D >= 0;
      if (UP) {
        upd(xs, k, xr, r, stage(Tg, kr), k, false, one, red);
        trmul(xr, tmp, r, stage(LUg, rr), true, unit, true, red);
        upd(xr, r, xs, k, stage(Eg, kr), k, true, one, red);
        if (unit) {
          __syncthreads();
          const S* D = fval + f.D;
          // X(rd) = D X(rd), D with 1 x 1 and 2 x 2 blocks (subdiagonal at D + r)
          for (int o = threadIdx.x; o < r * kRhs; o += blockDim.x) {
            const int i = o % r, c = o / r;
            const S b = D[r + i];
            if (b != (S)0) {
              const S x0 = xr[i + c * r], x1 = xr[i + 1 + c * r];
              xr[i + c * r] = D[i] * x0 + b * x1;
              xr[i + 1 + c * r] = b * x0 + D[i + 1] * x1;
            } else if (i == 0 || D[r + i - 1] == (S)0) {
              xr[o] *= D[i];
            }
          }
          __syncthreads();
        }
      } else {
        upd(xs, k, xr, r, stage(Eg, kr), k, false, one, red);
        trmul(xr, tmp, r, stage(LUg, rr), true, unit, false, red);
        upd(xr, r, xs, k, stage(Tg, kr), k, true, one, red);
      }
    }
    for (int a = threadIdx.x; a < k + r; a += blockDim.x) {
      const bool isk = a < k;
      const int64_t row = isk ? sk[a] : rd[a - k];
      S* dstg = X + row + c0 * ldx;
      const S* srcl = isk ? xs + a : xr + (a - k);
      const int ld = isk ? k : r;
#pragma unroll
      for (int c = 0; c < kRhs; ++c) dstg[c * ldx] = srcl[c * ld];
    }
    __syncthreads();
  }
}

template <typename S, int PROG>
static void launch_prog(bool up, const FacRec* recs, int nrec, const S* fval, const int* fidx, S* X,
                        int64_t ldx, int nrhs, int kmax, int rmax, int bcap, size_t shmem, S* gws,
                        int64_t gstride, hipStream_t st) {
  auto kern = gws ? (up ? k_sweep<S, PROG, true, true> : k_sweep<S, PROG, false, true>)
                  : (up ? k_sweep<S, PROG, true, false> : k_sweep<S, PROG, false, false>);
  ensure_dyn_lds((const void*)kern, shmem);
  hipLaunchKernelGGL(kern, dim3(nrec), dim3(256), shmem, st, recs, fval, fidx, X, ldx, nrhs, kmax,
                     rmax, bcap, gws, gstride);
}

template <typename S>
void launch_sweep(int prog, bool up, const FacRec* recs, int nrec, const S* fval, const int* fidx,
                  S* X, int64_t ldx, int nrhs, int kmax, int rmax, int bcap, size_t shmem, S* gws,
                  int64_t gstride, hipStream_t st) {
  switch (prog) {
    case 0: launch_prog<S, 0>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
    case 1: launch_prog<S, 1>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
    case 2: launch_prog<S, 2>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
    case 3: launch_prog<S, 3>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
    case 4: launch_prog<S, 4>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
    default: launch_prog<S, 5>(up, recs, nrec, fval, fidx, X, ldx, nrhs, kmax, rmax, bcap, shmem, gws, gstride, st); break;
  }
}

// ------------------------------------------------------- k_pack --------
// Transposed copies of each factor's blocks (FacRec::Tt ...), written once
// after the factorization through 32 x 33 LDS tiles: reads and writes are
// both along columns.
template <typename S>
__device__ void transpose_into(S* __restrict__ dst, const S* __restrict__ src, int R, int C,
                               S (*tile)[33]) {
  // src R x C (leading dimension R) -> dst C x R (leading dimension C); the
  // gridDim.y workgroups of a record share its 32 x 32 tiles
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5, ny = blockDim.x >> 5;
  const int nti = (R + 31) / 32, ntj = (C + 31) / 32;
  for (int64_t t = blockIdx.y; t < (int64_t)nti * ntj; t += gridDim.y) {
    const int i0 = (int)(t % nti) * 32, j0 = (int)(t / nti) * 32;
    for (int q = ty; q < 32; q += ny) {
      const int i = i0 + tx, j = j0 + q;
      tile[q][tx] = (i < R && j < C) ? src[i + (int64_t)j * R] : (S)0;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += ny) {
      const int j = j0 + tx, i = i0 + q;
      if (i < R && j < C) dst[j + (int64_t)i * C] = tile[tx][q];
    }
    __syncthreads();
  }
}

template <typename S, bool LU>
__global__ void __launch_bounds__(256) k_pack(const FacRec* __restrict__ recs, S* fval) {
  __shared__ S tile[32][33];
  const FacRec f = recs[blockIdx.x];
  if (f.packed) return;
  const int k = f.k, r = f.r, nbk = (r + kDiagBlock - 1) / kDiagBlock;
  transpose_into(fval + f.Tt, fval + f.T, k, r, tile);
  transpose_into(fval + f.Et, fval + f.E, k, r, tile);
  if (LU) transpose_into(fval + f.Gt, fval + f.G, r, k, tile);
  transpose_into(fval + f.LUt, fval + f.LU, r, r, tile);
  for (int b = 0; b < nbk; ++b) {
    const int64_t o = (int64_t)b * kDiagBlock * kDiagBlock;
    transpose_into(fval + f.Lti + o, fval + f.Li + o, kDiagBlock, kDiagBlock, tile);
    if (LU) transpose_into(fval + f.Uti + o, fval + f.Ui + o, kDiagBlock, kDiagBlock, tile);
  }
}

template <typename S>
void launch_pack(bool lu, const FacRec* recs, int nrec, int rmax, S* fval, hipStream_t st) {
  if (nrec <= 0) return;
  // workgroups per record: enough for the largest factor's tiles to fill
  // the chip when the level has few records
  const int64_t tiles = (int64_t)((rmax + 31) / 32) * ((rmax + 31) / 32);
  const int ys = (int)std::max<int64_t>(1, std::min<int64_t>(tiles / 16, 2048 / nrec));
  if (lu) hipLaunchKernelGGL((k_pack<S, true>), dim3(nrec, ys), dim3(256), 0, st, recs, fval);
  else hipLaunchKernelGGL((k_pack<S, false>), dim3(nrec, ys), dim3(256), 0, st, recs, fval);
}

// ------------------------------------------------------- k_solve -------
// rskelf_sv sweeps (sv_nn.m, sv_nc.m, sv_p.m).  One workgroup per factor;
// only the factor's rows of X live in LDS, row-major with the kRhs columns
// of a row together (one broadcast read per row), and every matrix is read
// from HBM along its columns: consecutive rows on consecutive lanes.  The
// transposed operands come from k_pack's copies.
constexpr int kSolveThreads = 256;

// Y(i,:) -= sum_l M[i + l ldm] Z(l,:), Y ny x kRhs, Z nz x kRhs (row-major);
// when there are fewer rows than threads each row's sum is split between
// P threads and combined through red (P ny kRhs <= kSolveThreads kRhs).
template <typename S>
__device__ void solve_prod(S* Y, int ny, const S* Z, int nz, const S* __restrict__ M, int ldm,
                           S* red) {
  const int tid = threadIdx.x;
  if (ny == 0 || nz == 0) return;
  int P = kSolveThreads / ny;
  P = P < 1 ? 1 : (P > 16 ? 16 : P);
  if (P > nz) P = nz;
  if (P == 1) {
    for (int i = tid; i < ny; i += kSolveThreads) {
      S acc[kRhs] = {};
#pragma unroll 4
      for (int l = 0; l < nz; ++l) {
        const S a = M[i + (int64_t)l * ldm];
#pragma unroll
        for (int c = 0; c < kRhs; ++c) acc[c] += a * Z[l * kRhs + c];
      }
#pragma unroll
      for (int c = 0; c < kRhs; ++c) Y[i * kRhs + c] -= acc[c];
    }
    __syncthreads();
    return;
  }
  if (tid < ny * P) {
    const int i = tid % ny, part = tid / ny;
    S acc[kRhs] = {};
#pragma unroll 4
    for (int l = part; l < nz; l += P) {
      const S a = M[i + (int64_t)l * ldm];
#pragma unroll
      for (int c = 0; c < kRhs; ++c) acc[c] += a * Z[l * kRhs + c];
    }
#pragma unroll
    for (int c = 0; c < kRhs; ++c) red[(part * ny + i) * kRhs + c] = acc[c];
  }
  __syncthreads();
  for (int o = tid; o < ny * kRhs; o += kSolveThreads) {
    S sum = 0;
    for (int part = 0; part < P; ++part) sum += red[part * ny * kRhs + o];
    Y[o] -= sum;
  }
  __syncthreads();
}

// Y = Op^{-1} Y for a triangular r x r operator Op(i, j) = M[i + j r] whose
// 32 x 32 diagonal blocks have inverses Minv (block b at b * 1024, element
// (i, j) at i + 32 j).  fwd: Op lower (forward substitution by blocks);
// otherwise upper (backward).  Only Op's off-diagonal blocks are read.
template <typename S>
__device__ void solve_tri(S* Y, int r, const S* __restrict__ M, const S* __restrict__ Minv,
                          bool fwd, S* red) {
  const int tid = threadIdx.x;
  const int nbk = (r + kDiagBlock - 1) / kDiagBlock;
  const int bi = tid & (kDiagBlock - 1), bc = tid / kDiagBlock;   // kSolveThreads = 32 kRhs
  for (int bk = 0; bk < nbk; ++bk) {
    const int b = fwd ? bk : nbk - 1 - bk;
    const int j0 = b * kDiagBlock, bs = r - j0 < kDiagBlock ? r - j0 : kDiagBlock;
    const S* Ib = Minv + (int64_t)b * kDiagBlock * kDiagBlock;
    S acc = 0;
    if (bi < bs) {
#pragma unroll 8
      for (int u = 0; u < bs; ++u) acc += Ib[bi + u * kDiagBlock] * Y[(j0 + u) * kRhs + bc];
    }
    __syncthreads();
    if (bi < bs) Y[(j0 + bi) * kRhs + bc] = acc;
    __syncthreads();
    const int lo = fwd ? j0 + bs : 0, hi = fwd ? r : j0, nrow = hi - lo;
    if (nrow <= 0) continue;
    int P = kSolveThreads / nrow;
    P = P < 1 ? 1 : (P > 8 ? 8 : P);
    if (P == 1) {
      for (int i = lo + tid; i < hi; i += kSolveThreads) {
        S a2[kRhs] = {};
#pragma unroll 4
        for (int t = 0; t < bs; ++t) {
          const S a = M[i + (int64_t)(j0 + t) * r];
#pragma unroll
          for (int c = 0; c < kRhs; ++c) a2[c] += a * Y[(j0 + t) * kRhs + c];
        }
#pragma unroll
        for (int c = 0; c < kRhs; ++c) Y[i * kRhs + c] -= a2[c];
      }
      __syncthreads();
      continue;
    }
    if (tid < nrow * P) {
      const int i = lo + tid % nrow, part = tid / nrow;
      S a2[kRhs] = {};
#pragma unroll 4
      for (int t = part; t < bs; t += P) {
        const S a = M[i + (int64_t)(j0 + t) * r];
#pragma unroll
        for (int c = 0; c < kRhs; ++c) a2[c] += a * Y[(j0 + t) * kRhs + c];
      }
#pragma unroll
      for (int c = 0; c < kRhs; ++c) red[(part * nrow + (i - lo)) * kRhs + c] = a2[c];
    }
    __syncthreads();
    for (int o = tid; o < nrow * kRhs; o += kSolveThreads) {
      S sum = 0;
      for (int part = 0; part < P; ++part) sum += red[part * nrow * kRhs + o];
      Y[lo * kRhs + o] -= sum;
    }
    __syncthreads();
  }
}

// rows i of Y (r x kRhs, row-major) permuted: gather Y(i) = Y(p(i)) or
// scatter Y(p(i)) = Y(i), through tmp
template <typename S>
__device__ void solve_perm(S* Y, S* tmp, int r, const int* __restrict__ p, bool gather) {
  for (int o = threadIdx.x; o < r * kRhs; o += kSolveThreads) {
    const int i = o / kRhs, c = o - i * kRhs, pi = p[i];
    if (gather) tmp[o] = Y[pi * kRhs + c];
    else tmp[pi * kRhs + c] = Y[o];
  }
  __syncthreads();
  for (int o = threadIdx.x; o < r * kRhs; o += kSolveThreads) Y[o] = tmp[o];
  __syncthreads();
}

// PROG: 0 sv 'n'/'n', 1 sv 'n'/'c', 2 sv 'p'.  Shared memory: xs (kmax x 8),
// xr (rmax x 8), red (8 x 256), and for PROG < 2 a permutation buffer
// (rmax x 8).  X has nrhs columns, a multiple of 8.
template <typename S, int PROG, bool UP, bool GWS>
__global__ void __launch_bounds__(kSolveThreads) k_solve(const FacRec* __restrict__ recs,
                                                         const S* __restrict__ fval,
                                                         const int* __restrict__ fidx, S* X,
                                                         int64_t ldx, int nrhs, int kmax,
                                                         int rmax, S* gws, int64_t gstride,
                                                         int rlib) {
  extern __shared__ char smem_raw[];
  // shared memory, or (GWS, a compile-time choice that leaves the shared
  // variant its LDS accesses) this workgroup's slice of a global workspace
  // for factors whose rows do not fit
  S* xs = GWS ? gws + blockIdx.x * gstride : reinterpret_cast<S*>(smem_raw);
  S* xr = xs + (size_t)kmax * kRhs;
  S* red = xr + (size_t)rmax * kRhs;
  S* tmp = red + (size_t)kRhs * kSolveThreads;
  const FacRec f = recs[blockIdx.x];
  const int k = f.k, r = f.r;
  if (r >= rlib && f.D < 0) return;      // solved with the whole GPU (library calls)
  const int* sk = fidx + f.sk;
  const int* rd = fidx + f.rd;
  const int* piv = fidx + f.piv;
  for (int c0 = 0; c0 < nrhs; c0 += kRhs) {
    for (int a = threadIdx.x; a < k + r; a += kSolveThreads) {
      const bool isk = a < k;
      const int64_t row = isk ? sk[a] : rd[a - k];
      const S* src = X + row + c0 * ldx;
      S v[kRhs];
#pragma unroll
      for (int c = 0; c < kRhs; ++c) v[c] = src[c * ldx];
      S* dst = isk ? xs + a * kRhs : xr + (a - k) * kRhs;
#pragma unroll
      for (int c = 0; c < kRhs; ++c) dst[c] = v[c];
    }
    __syncthreads();
    if (PROG == 0) {            // sv_nn.m
      if (UP) {
        solve_prod(xr, r, xs, k, fval + f.Tt, r, red);
        solve_perm(xr, tmp, r, piv, true);
        solve_tri(xr, r, fval + f.LU, fval + f.Li, true, red);
        solve_prod(xs, k, xr, r, fval + f.E, k, red);
      } else {
        solve_prod(xr, r, xs, k, fval + f.G, r, red);
        solve_tri(xr, r, fval + f.LU, fval + f.Ui, false, red);
        solve_prod(xs, k, xr, r, fval + f.T, k, red);
      }
    } else if (PROG == 1) {     // sv_nc.m
      if (UP) {
        solve_prod(xr, r, xs, k, fval + f.Tt, r, red);
        solve_tri(xr, r, fval + f.LUt, fval + f.Uti, true, red);
        solve_prod(xs, k, xr, r, fval + f.Gt, k, red);
      } else {
        solve_prod(xr, r, xs, k, fval + f.Et, r, red);
        solve_tri(xr, r, fval + f.LUt, fval + f.Lti, false, red);
        solve_perm(xr, tmp, r, piv, false);
        solve_prod(xs, k, xr, r, fval + f.T, k, red);
      }
    } else {                    // sv_p.m; sv_h.m (unit L, then X(rd) = D \ X(rd))
      if (UP) {
        solve_prod(xr, r, xs, k, fval + f.Tt, r, red);
        solve_tri(xr, r, fval + f.LU, fval + f.Li, true, red);
        solve_prod(xs, k, xr, r, fval + f.E, k, red);
        if (f.D >= 0) {
          __syncthreads();
          const S* D = fval + f.D;
          // X(rd) = D \ X(rd), D with 1 x 1 and 2 x 2 blocks (subdiagonal at D + r)
          for (int o = threadIdx.x; o < r * kRhs; o += kSolveThreads) {
            const int i = o / kRhs, c = o - i * kRhs;
            const S b = D[r + i];
            if (b != (S)0) {
              const S d0 = D[i], d1 = D[i + 1], det = d0 * d1 - b * b;
              const S x0 = xr[i * kRhs + c], x1 = xr[(i + 1) * kRhs + c];
              xr[i * kRhs + c] = (d1 * x0 - b * x1) / det;
              xr[(i + 1) * kRhs + c] = (d0 * x1 - b * x0) / det;
            } else if (i == 0 || D[r + i - 1] == (S)0) {
              xr[o] /= D[i];
            }
          }
          __syncthreads();
        }
      } else {
        solve_prod(xr, r, xs, k, fval + f.Et, r, red);
        solve_tri(xr, r, fval + f.LUt, fval + f.Lti, false, red);
        solve_prod(xs, k, xr, r, fval + f.T, k, red);
      }
    }
    for (int a = threadIdx.x; a < k + r; a += kSolveThreads) {
      const bool isk = a < k;
      const int64_t row = isk ? sk[a] : rd[a - k];
      S* dstg = X + row + c0 * ldx;
      const S* srcl = isk ? xs + a * kRhs : xr + (a - k) * kRhs;
#pragma unroll
      for (int c = 0; c < kRhs; ++c) dstg[c * ldx] = srcl[c];
    }
    __syncthreads();
  }
}

size_t solve_shmem(int prog, int kmax, int rmax, size_t es) {
  return ((size_t)(kmax + rmax) * kRhs + (size_t)kRhs * kSolveThreads +
          (prog < 2 ? (size_t)rmax * kRhs : 0)) * es;
}

template <typename S>
void launch_solve(int prog, bool up, const FacRec* recs, int nrec, const S* fval, const int* fidx,
                  S* X, int64_t ldx, int nrhs, int kmax, int rmax, size_t shmem, S* gws,
                  int64_t gstride, int rlib, hipStream_t st) {
#define RSKX_SOLVE(P)                                                                           \
  {                                                                                             \
    auto kern = gws ? (up ? k_solve<S, P, true, true> : k_solve<S, P, false, true>)            \
                    : (up ? k_solve<S, P, true, false> : k_solve<S, P, false, false>);          \
    ensure_dyn_lds((const void*)kern, shmem);                                                    \
    hipLaunchKernelGGL(kern, dim3(nrec), dim3(kSolveThreads), shmem, st, recs, fval, fidx, X,    \
                       ldx, nrhs, kmax, rmax, gws, gstride, rlib);                              \
  }
  if (prog == 0) RSKX_SOLVE(0) else if (prog == 1) RSKX_SOLVE(1) else RSKX_SOLVE(2)
#undef RSKX_SOLVE
}

// ------------------------------------------------ X to / from tree order --
// Y has ncol >= nrhs columns; the extra ones are zero
template <typename U, typename S>
__global__ void k_to_tree(const U* __restrict__ X, int64_t ldx, S* __restrict__ Y, int64_t N,
                          int nrhs, int ncol, const int* __restrict__ order) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = order[p];
    for (int c = 0; c < ncol; ++c) Y[p + c * N] = c < nrhs ? (S)X[q + c * ldx] : (S)0;
  }
}

template <typename U, typename S>
__global__ void k_from_tree(const S* __restrict__ Y, U* __restrict__ X, int64_t ldx, int64_t N,
                            int nrhs, const int* __restrict__ order) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = order[p];
    for (int c = 0; c < nrhs; ++c) X[q + c * ldx] = (U)Y[p + c * N];
  }
}

template <typename U, typename S>
void launch_perm(bool to, void* X, int64_t ldx, S* Y, int64_t N, int nrhs, int ncol,
                 const int* order, hipStream_t st) {
  int grid = (int)((N + 255) / 256);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  if (to)
    hipLaunchKernelGGL((k_to_tree<U, S>), dim3(grid), dim3(256), 0, st, (const U*)X, ldx, Y, N,
                       nrhs, ncol, order);
  else
    hipLaunchKernelGGL((k_from_tree<U, S>), dim3(grid), dim3(256), 0, st, (const S*)Y, (U*)X, ldx,
                       N, nrhs, order);
}

// ------------------------------------------------------------ log det --
// One workgroup per factor: the sum of log |d| over its r pivots (diag(L)
// for 'p', D for 'h', diag(U) for 'n'), how many are negative, and for 'n'
// the parity of the row permutation, counted by walking its cycles with
// a visited bitmap in LDS (rskelf/rskelf_logdet.m:17-27, core/logdet_ldl.m,
// compat/detperm.m).  out[3 f + {0, 1, 2}].
constexpr int kLogdetMaxR = 65536;

template <typename S>
__global__ void __launch_bounds__(256) k_logdet(const FacRec* __restrict__ recs,
                                                const S* __restrict__ fval,
                                                const int* __restrict__ fidx, int mode,
                                                double* __restrict__ out) {
  __shared__ double rs[4];
  __shared__ int rn[4];
  __shared__ unsigned seen[kLogdetMaxR / 32];
  const FacRec f = recs[blockIdx.x];
  const int r = f.r;
  double s = 0;
  int neg = 0;
  for (int i = threadIdx.x; i < r; i += blockDim.x) {
    double d;
    if (mode == 1) {          // D of L D L': 1 x 1 pivots and 2 x 2 blocks (core/logdet_ldl.m)
      const double b = (double)fval[f.D + r + i];
      if (b != 0.0) d = (double)fval[f.D + i] * (double)fval[f.D + i + 1] - b * b;
      else if (i > 0 && (double)fval[f.D + r + i - 1] != 0.0) continue;   // second row of a block
      else d = (double)fval[f.D + i];
    } else {
      d = (double)fval[f.LU + i + (int64_t)i * r];
    }
    s += log(fabs(d));
    neg += d < 0;
  }
  s = wave_sum(s);
  neg = (int)wave_sum((double)neg);
  if ((threadIdx.x & 63) == 0) { rs[threadIdx.x >> 6] = s; rn[threadIdx.x >> 6] = neg; }
  int parity = 0;
  if (mode == 2) {
    for (int w = threadIdx.x; w < (r + 31) / 32; w += blockDim.x) seen[w] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int* piv = fidx + f.piv;
      int cycles = 0;
      for (int i = 0; i < r; ++i) {
        if (seen[i >> 5] & (1u << (i & 31))) continue;
        ++cycles;
        for (int j = i; !(seen[j >> 5] & (1u << (j & 31))); j = piv[j]) seen[j >> 5] |= 1u << (j & 31);
      }
      parity = (r - cycles) & 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double st = 0;
    int nt = 0;
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) { st += rs[q]; nt += rn[q]; }
    out[3 * (int64_t)blockIdx.x] = st;
    out[3 * (int64_t)blockIdx.x + 1] = nt;
    out[3 * (int64_t)blockIdx.x + 2] = parity;
  }
}

template <typename S>
void launch_logdet(const FacRec* recs, int nrec, const S* fval, const int* fidx, int mode,
                   double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_logdet<S>, dim3(nrec), dim3(256), 0, st, recs, fval, fidx, mode, out);
}

// ------------------------------------------- FP64 on-the-fly A X product --
constexpr double kTwoPiM = 6.283185307179586476925286766559;

// problem of point i: the last j with poff[j] <= i (poff ascending, nprob + 1 entries)
__device__ inline int problem_of(const int64_t* __restrict__ poff, int nprob, int64_t i) {
  int lo = 0, hi = nprob;                     // poff[lo] <= i < poff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (poff[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Y = A X for the block-diagonal matrix of nprob independent problems
// (problem j: points [poff[j], poff[j+1])); poff == nullptr is one problem
__global__ void __launch_bounds__(256) k_matvec(const double* __restrict__ x, int64_t N,
                                                const double* __restrict__ nu,
                                                const double* __restrict__ w,
                                                const double* __restrict__ diag, double scale,
                                                const double* __restrict__ X, int64_t ldx,
                                                double* __restrict__ Y, int64_t ldy, int nrhs,
                                                const int64_t* __restrict__ poff, int nprob) {
  __shared__ double sx[256], sy[256], snx[256], sny[256], sw[256];
  __shared__ double sX[8][256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool act = i < N;
  const double xi = act ? x[2 * i] : 0, yi = act ? x[2 * i + 1] : 0;
  // sources of this thread's problem, and of the block's first and last targets
  int64_t my0 = 0, my1 = N, blk0 = 0, blk1 = N;
  if (poff) {
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x;
    const int64_t i1 = i0 + blockDim.x - 1 < N - 1 ? i0 + blockDim.x - 1 : N - 1;
    const int pa = problem_of(poff, nprob, i0), pb = problem_of(poff, nprob, i1);
    blk0 = poff[pa];
    blk1 = poff[pb + 1];
    if (act) {
      const int pi = problem_of(poff, nprob, i);
      my0 = poff[pi];
      my1 = poff[pi + 1];
    }
  }
  for (int c0 = 0; c0 < nrhs; c0 += 8) {
    const int cc = nrhs - c0 < 8 ? nrhs - c0 : 8;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j0 = blk0; j0 < blk1; j0 += 256) {
      __syncthreads();
      int64_t j = j0 + threadIdx.x;
      if (j < blk1) {
        sx[threadIdx.x] = x[2 * j];
        sy[threadIdx.x] = x[2 * j + 1];
        snx[threadIdx.x] = nu[2 * j];
        sny[threadIdx.x] = nu[2 * j + 1];
        sw[threadIdx.x] = w[j];
        for (int c = 0; c < cc; ++c) sX[c][threadIdx.x] = X[j + (c0 + c) * ldx];
      }
      __syncthreads();
      const int jn = blk1 - j0 < 256 ? (int)(blk1 - j0) : 256;
      for (int t = 0; t < jn; ++t) {
        int64_t jj = j0 + t;
        if (jj < my0 || jj >= my1) continue;   // another problem's source
        double a;
        if (jj == i) a = diag[jj];
        else {
          double dx = xi - sx[t], dy = yi - sy[t];
          a = scale * sw[t] * (dx * snx[t] + dy * sny[t]) / (kTwoPiM * (dx * dx + dy * dy));
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) if (c < cc) acc[c] += a * sX[c][t];
      }
    }
    if (act)
#pragma unroll
      for (int c = 0; c < 8; ++c) if (c < cc) Y[i + (c0 + c) * ldy] = acc[c];
  }
}

// Y = A X for RSKX_KERNEL_LAPLACE3D_SL: A_ij = scale w_j / (4 pi |x_i - x_j|),
// A_ii = diag_i; x column-major 3 x N; blocks of problems as k_matvec
constexpr double kFourPiM = 12.566370614359172953850573533118;

__global__ void __launch_bounds__(256) k_matvec_sl3(const double* __restrict__ x, int64_t N,
                                                    const double* __restrict__ w,
                                                    const double* __restrict__ diag, double scale,
                                                    const double* __restrict__ X, int64_t ldx,
                                                    double* __restrict__ Y, int64_t ldy, int nrhs,
                                                    const int64_t* __restrict__ poff, int nprob) {
  // a staged source: its point and scale w_j / (4 pi), then its 8
  // right-hand-side values together (two and four 16-byte LDS reads)
  __shared__ double4 sp[256];
  __shared__ double4 sX[256][2];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool act = i < N;
  const double xi = act ? x[3 * i] : 0, yi = act ? x[3 * i + 1] : 0, zi = act ? x[3 * i + 2] : 0;
  int64_t my0 = 0, my1 = N, blk0 = 0, blk1 = N;
  if (poff) {
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x;
    const int64_t i1 = i0 + blockDim.x - 1 < N - 1 ? i0 + blockDim.x - 1 : N - 1;
    const int pa = problem_of(poff, nprob, i0), pb = problem_of(poff, nprob, i1);
    blk0 = poff[pa];
    blk1 = poff[pb + 1];
    if (act) {
      const int pi = problem_of(poff, nprob, i);
      my0 = poff[pi];
      my1 = poff[pi + 1];
    }
  }
  const double c0 = scale / kFourPiM;
  for (int c0r = 0; c0r < nrhs; c0r += 8) {
    const int cc = nrhs - c0r < 8 ? nrhs - c0r : 8;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j0 = blk0; j0 < blk1; j0 += 256) {
      __syncthreads();
      int64_t j = j0 + threadIdx.x;
      if (j < blk1) {
        sp[threadIdx.x] = double4{x[3 * j], x[3 * j + 1], x[3 * j + 2], c0 * w[j]};
        double v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = c < cc ? X[j + (c0r + c) * ldx] : 0.0;
        sX[threadIdx.x][0] = double4{v[0], v[1], v[2], v[3]};
        sX[threadIdx.x][1] = double4{v[4], v[5], v[6], v[7]};
      }
      __syncthreads();
      const int jn = blk1 - j0 < 256 ? (int)(blk1 - j0) : 256;
      if (j0 >= my0 && j0 + jn <= my1) {
        // every source of the tile belongs to this row's problem: 1/r from
        // the FP32 reciprocal square root and two Newton-Raphson steps in
        // FP64 (relative error ~1e-16); the diagonal is added after the
        // loop.  Columns past cc hold zeros and are not stored.
#pragma unroll 2
        for (int t = 0; t < jn; ++t) {
          const double4 q = sp[t];
          const double dx = xi - q.x, dy = yi - q.y, dz = zi - q.z;
          const double r2 = dx * dx + dy * dy + dz * dz;
          double y = (double)rsqrtf((float)r2);
          y = y * (1.5 - 0.5 * r2 * y * y);
          y = y * (1.5 - 0.5 * r2 * y * y);
          const double a = r2 > 0 ? q.w * y : 0.0;
          const double4 u = sX[t][0], v = sX[t][1];
          acc[0] += a * u.x; acc[1] += a * u.y; acc[2] += a * u.z; acc[3] += a * u.w;
          acc[4] += a * v.x; acc[5] += a * v.y; acc[6] += a * v.z; acc[7] += a * v.w;
        }
        if (act && i >= j0 && i < j0 + jn) {
          const double d = diag[i];
          const double4 u = sX[i - j0][0], v = sX[i - j0][1];
          acc[0] += d * u.x; acc[1] += d * u.y; acc[2] += d * u.z; acc[3] += d * u.w;
          acc[4] += d * v.x; acc[5] += d * v.y; acc[6] += d * v.z; acc[7] += d * v.w;
        }
        continue;
      }
      for (int t = 0; t < jn; ++t) {
        int64_t jj = j0 + t;
        if (jj < my0 || jj >= my1) continue;   // another problem's source
        const double4 q = sp[t];
        double a;
        if (jj == i) a = diag[jj];
        else {
          const double dx = xi - q.x, dy = yi - q.y, dz = zi - q.z;
          a = q.w / sqrt(dx * dx + dy * dy + dz * dz);
        }
        const double4 u = sX[t][0], v = sX[t][1];
        acc[0] += a * u.x; acc[1] += a * u.y; acc[2] += a * u.z; acc[3] += a * u.w;
        acc[4] += a * v.x; acc[5] += a * v.y; acc[6] += a * v.z; acc[7] += a * v.w;
      }
    }
    if (act)
#pragma unroll
      for (int c = 0; c < 8; ++c) if (c < cc) Y[i + (c0r + c) * ldy] = acc[c];
  }
}

void launch_matvec(int kind, const double* x, int64_t N, const double* nu, const double* w,
                   const double* diag, double scale, const double* X, int64_t ldx, double* Y,
                   int64_t ldy, int nrhs, const int64_t* poff, int nprob, hipStream_t st) {
  int grid = (int)((N + 255) / 256);
  if (kind == RSKX_KERNEL_LAPLACE3D_SL)
    hipLaunchKernelGGL(k_matvec_sl3, dim3(grid), dim3(256), 0, st, x, N, w, diag, scale, X, ldx,
                       Y, ldy, nrhs, poff, nprob);
  else
    hipLaunchKernelGGL(k_matvec, dim3(grid), dim3(256), 0, st, x, N, nu, w, diag, scale, X, ldx,
                       Y, ldy, nrhs, poff, nprob);
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
      pts[(d + i) * N + p] = nu ? nu[d * q + i] : 0.0;   // kernels without normals
    }
    pts[2 * d * N + p] = w ? w[q] : 0.0;             // host-callback matrices carry
    pts[(2 * d + 1) * N + p] = diag ? diag[q] : 0.0; // neither weights nor diagonal
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
                                int64_t, int, int, int, int, size_t, S*, int64_t, hipStream_t); \
  template void launch_perm<double, S>(bool, void*, int64_t, S*, int64_t, int, int,         \
                                       const int*, hipStream_t);                              \
  template void launch_perm<float, S>(bool, void*, int64_t, S*, int64_t, int, int,          \
                                      const int*, hipStream_t);                               \
  template void launch_pack<S>(bool, const FacRec*, int, int, S*, hipStream_t);               \
  template void launch_logdet<S>(const FacRec*, int, const S*, const int*, int, double*,      \
                                 hipStream_t);                                                \
  template void launch_solve<S>(int, bool, const FacRec*, int, const S*, const int*, S*,      \
                                int64_t, int, int, int, size_t, S*, int64_t, int, hipStream_t);
RSKX_SWEEP_INST(double)
RSKX_SWEEP_INST(float)

}  // namespace rskx
