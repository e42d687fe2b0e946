// Per-level factorization kernels (one workgroup per box of the level).
//
//   k_lists     slf = own points or children's skeletons (rskelf.m:179-187),
//               neighbour rows filtered by the proxy ring (ie_circle.m:114)
//   k_assemble  K = [A(nbr,slf); A(slf,nbr)' ('n'); Kpxy] (rskelf.m:201-212)
//   k_id        column ID of K by QRCP truncated at the tolerance rank,
//               T = R11\R12 and Gu-Eisenstat refinement (core/id.m:44-248)
//   k_factor    sparsify, LU / Cholesky of the redundant block, E, G and
//               the skeleton Schur complement (rskelf.m:218-261)
#include "common.hpp"
#include "devutil.hpp"
#include "entries.hpp"
#include "mfma.hpp"
#include "rskx.h"

namespace rskx {

// ------------------------------------------------------- matrix entries --
// A_ij of RSKX_KERNEL_LAPLACE2D_DL, i != j (points in tree order)
__device__ inline double entry_offdiag(const PointsDev& P, int i, int j) {
  double dx = P.x[0][i] - P.x[0][j];
  double dy = P.x[1][i] - P.x[1][j];
  double rdotn = dx * P.nu[0][j] + dy * P.nu[1][j];
  return P.scale * P.w[j] * rdotn / (kTwoPi * (dx * dx + dy * dy));
}

__device__ inline double entry(const PointsDev& P, int i, int j) {
  return i == j ? P.diag[j] : entry_offdiag(P, i, j);
}

__device__ inline bool in_ring(const PointsDev& P, const LevelParams& L, const LvlBox& B, int pt) {
  double s = 0;
  for (int q = 0; q < P.d; ++q) {
    double t = (P.x[q][pt] - B.ctr[q]) / B.l[q];
    s += t * t;
  }
  return s < L.radius * L.radius;
}

// ------------------------------------------------------------- k_lists --
__global__ void __launch_bounds__(256) k_lists(const LvlBox* __restrict__ boxes,
                                               const Seg* __restrict__ nsegs,
                                               const Seg* __restrict__ csegs, int* pool,
                                               int* __restrict__ nbrar, int* __restrict__ mnbr,
                                               PointsDev P, LevelParams L) {
  const LvlBox B = boxes[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(tid / kWave), nw = nt / kWave;
  int* slf = pool + B.slf_off;
  int pos = 0;
  for (int s = 0; s < B.ncs; ++s) {
    Seg g = csegs[B.cs0 + s];
    for (int i = tid; i < g.len; i += nt) slf[pos + i] = pool[g.off + i];
    pos += g.len;
  }
  __shared__ int wtot[16];
  int* nbr = nbrar + B.nbr_off;
  int cnt = 0;
  for (int s = 0; s < B.nseg; ++s) {
    Seg g = nsegs[B.seg0 + s];
    for (int b0 = 0; b0 < g.len; b0 += nt) {
      int i = b0 + tid;
      bool keep = false;
      int pt = 0;
      if (i < g.len) {
        pt = pool[g.off + i];
        keep = !L.filter || in_ring(P, L, B, pt);
      }
      uint64_t mask = __ballot(keep);
      int pre = __popcll(mask & ((1ull << lane) - 1ull));
      if (lane == 0) wtot[wid] = __popcll(mask);
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < nw; ++w) {
        if (w < wid) off += wtot[w];
        tot += wtot[w];
      }
      if (keep) nbr[cnt + off + pre] = pt;
      cnt += tot;
      __syncthreads();
    }
  }
  if (tid == 0) mnbr[blockIdx.x] = cnt;
}

// ----------------------------------------------------------- k_assemble --
// K of the boxes no register ID class takes, in global memory; gridDim.y
// workgroups share a box's columns.  The neighbour (row) fields are staged
// in LDS a chunk of kAsmRows at a time; a column's own fields are the same
// for every thread and are read directly.  No limit on the box or
// neighbourhood size.
constexpr int kAsmRows = 1024;

template <typename S>
__global__ void __launch_bounds__(256) k_assemble(const int* __restrict__ ids,
                                                  const LvlBox* __restrict__ boxes,
                                                  const int* __restrict__ pool,
                                                  const int* __restrict__ nbrar,
                                                  const int* __restrict__ mnbr,
                                                  int* __restrict__ mrows, S* __restrict__ Kar,
                                                  PointsDev P, LevelParams L) {
  __shared__ double rf[5][kAsmRows];
  const int bidx = ids[blockIdx.x];
  const LvlBox B = boxes[bidx];
  const int n = B.n, mn = mnbr[bidx];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  // workgroup (box, y) assembles columns y, y + gridDim.y, ...
  const int j0 = blockIdx.y, js = gridDim.y;
  if (j0 >= n) return;
  if (j0 == 0 && threadIdx.x == 0) mrows[bidx] = m;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  S* K = Kar + B.K_off;
  const int ld = B.mcap;
  const bool sl3 = P.kind == RSKX_KERNEL_LAPLACE3D_SL;
  const double c0 = P.scale / kTwoPi;
  for (int r0 = 0; r0 < mn; r0 += kAsmRows) {
    const int rc = mn - r0 < kAsmRows ? mn - r0 : kAsmRows;
    __syncthreads();
    for (int i = threadIdx.x; i < rc; i += blockDim.x)
      point_fields(P, nbr[r0 + i], rf[0][i], rf[1][i], rf[2][i], rf[3][i], rf[4][i]);
    __syncthreads();
    for (int j = j0; j < n; j += js) {
      double xj, yj, nxj, nyj, wj;
      point_fields(P, slf[j], xj, yj, nxj, nyj, wj);
      S* Kj = K + (int64_t)j * ld + r0;
      for (int i = threadIdx.x; i < rc; i += blockDim.x) {    // A(nbr, slf)
        double v;
        if (sl3) {
          v = sl3_entry(P, rf[0][i], rf[1][i], rf[2][i], xj, yj, nxj, wj);
        } else {
          const double dx = rf[0][i] - xj, dy = rf[1][i] - yj;
          v = c0 * wj * (dx * nxj + dy * nyj) / (dx * dx + dy * dy);
        }
        Kj[i] = (S)v;
      }
      if (L.two_sided)
        for (int i = threadIdx.x; i < rc; i += blockDim.x) {  // A(slf, nbr)'
          double v;
          if (sl3) {
            v = sl3_entry(P, xj, yj, nxj, rf[0][i], rf[1][i], rf[2][i], rf[4][i]);
          } else {
            const double dx = xj - rf[0][i], dy = yj - rf[1][i];
            v = c0 * rf[4][i] * (dx * rf[2][i] + dy * rf[3][i]) / (dx * dx + dy * dy);
          }
          Kj[mn + i] = (S)v;
        }
    }
  }
  for (int j = j0; j < n; j += js) {                         // proxy rows
    double xj, yj, nxj, nyj, wj;
    point_fields(P, slf[j], xj, yj, nxj, nyj, wj);
    S* Kj = K + (int64_t)j * ld;
    for (int i = threadIdx.x; i < L.p; i += blockDim.x) {
      double px, py, pz;
      proxy_point(L, B, i, px, py, pz);
      double v;
      if (L.pkind == RSKX_PROXY_POINTS_SL) {
        v = proxy_sl3(L, B, i, px, py, pz, xj, yj, nxj, wj);
      } else {
        const double dx = px - xj, dy = py - yj;
        const double r2 = dx * dx + dy * dy;
        v = L.pkind == RSKX_PROXY_RING_SL ? -0.5 * log(r2) / kTwoPi * B.pw
                                          : wj * (dx * nxj + dy * nyj) / (kTwoPi * r2);
      }
      Kj[mt + i] = (S)v;
    }
  }
}

// ----------------------------------------------------------------- k_id --
// Householder QR of the k x k matrix R (leading dimension ldr) in place;
// only the triangular factor is kept.  All threads of the block take part.
__device__ void small_qr(double* R, int k, int ldr, double* v, double* red) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int s = 0; s < k - 1; ++s) {
    double part = 0;
    for (int i = s + 1 + tid; i < k; i += nt) part += R[i + s * ldr] * R[i + s * ldr];
    double sigma = block_sum(part, red);
    if (sigma == 0) continue;
    double alpha = R[s + s * ldr];
    double nrm = sqrt(alpha * alpha + sigma);
    double beta = alpha >= 0 ? -nrm : nrm;
    double tau = (beta - alpha) / beta, scal = 1.0 / (alpha - beta);
    for (int i = s + tid; i < k; i += nt) v[i] = i == s ? 1.0 : R[i + s * ldr] * scal;
    __syncthreads();
    for (int j = s + 1 + tid; j < k; j += nt) {
      double dot = 0;
      for (int i = s; i < k; ++i) dot += v[i] * R[i + j * ldr];
      dot *= tau;
      for (int i = s; i < k; ++i) R[i + j * ldr] -= dot * v[i];
    }
    __syncthreads();
    if (tid == 0) R[s + s * ldr] = beta;
    for (int i = s + 1 + tid; i < k; i += nt) R[i + s * ldr] = 0;
    __syncthreads();
  }
}

// squared row norms of inv(R) for upper-triangular k x k R; X is k x k scratch
__device__ void inv_row_norms(const double* R, int k, int ldr, double* X, int ldx, double* r2) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = tid; c < k; c += nt) {
    for (int l = k - 1; l >= 0; --l) {
      if (l > c) { X[l + c * ldx] = 0; continue; }
      double s = (l == c) ? 1.0 : 0.0;
      for (int q = l + 1; q <= c; ++q) s -= R[l + q * ldr] * X[q + c * ldx];
      X[l + c * ldx] = s / R[l + l * ldr];
    }
  }
  __syncthreads();
  for (int l = tid; l < k; l += nt) {
    double s = 0;
    for (int c = l; c < k; ++c) s += X[l + c * ldx] * X[l + c * ldx];
    r2[l] = s;
  }
  __syncthreads();
}

// swap skeleton positions a and b: perm entries, T rows, R columns
__device__ void swap_skel(int* perm, double* T, int ldt, int r, double* R, int ldr, int k, int a,
                          int b) {
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) { int t = perm[a]; perm[a] = perm[b]; perm[b] = t; }
  for (int c = tid; c < r; c += nt) {
    double t = T[a + c * ldt]; T[a + c * ldt] = T[b + c * ldt]; T[b + c * ldt] = t;
  }
  for (int i = tid; i < k; i += nt) {
    double t = R[i + a * ldr]; R[i + a * ldr] = R[i + b * ldr]; R[i + b * ldr] = t;
  }
  __syncthreads();
}

// ---- refinement for large ranks (the tails of the widest boxes) --------
// Exchange columns i and k-1 of the upper-triangular k x k R and make it
// triangular again: what qrupdate returns in id.m:139-141, up to row signs
// (which change neither T nor the inverse row norms).  Rotations on rows
// (rr-1, rr), rr = k-1 .. i+1, clear column i below its diagonal (their
// parameters depend on column i alone, so one thread forms them and every
// other column applies the whole chain on its own); the subdiagonal they
// leave in columns i+1 .. k-2 is cleared top-down, one rotation per step.
__device__ void col_exchange(double* R, int k, int ldr, int i, double* cs, double* sn) {
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double rot[2];
  for (int l = tid; l < k; l += nt) {
    const double a = R[l + (int64_t)i * ldr];
    R[l + (int64_t)i * ldr] = R[l + (int64_t)(k - 1) * ldr];
    R[l + (int64_t)(k - 1) * ldr] = a;
  }
  __syncthreads();
  if (tid == 0) {
    double b = R[(k - 1) + (int64_t)i * ldr];
    for (int rr = k - 1; rr > i; --rr) {
      const double a = R[rr - 1 + (int64_t)i * ldr];
      const double rho = hypot(a, b);
      cs[rr] = rho > 0 ? a / rho : 1.0;
      sn[rr] = rho > 0 ? b / rho : 0.0;
      b = rho;
    }
    R[i + (int64_t)i * ldr] = b;
  }
  __syncthreads();
  for (int rr = i + 1 + tid; rr < k; rr += nt) R[rr + (int64_t)i * ldr] = 0.0;
  for (int col = i + 1 + tid; col < k; col += nt) {
    double* Rc = R + (int64_t)col * ldr;
    const int start = col + 1 < k - 1 ? col + 1 : k - 1;
    double y = Rc[start];
    for (int rr = start; rr > i; --rr) {
      const double x = Rc[rr - 1], c = cs[rr], s = sn[rr];
      Rc[rr] = -s * x + c * y;
      y = c * x + s * y;
    }
    Rc[i] = y;
  }
  __syncthreads();
  // top-down: clear R(c+1, c), c = i+1 .. k-2; the owner of column c+1
  // forms the next rotation once it has applied the current one
  if (tid == 0 && i + 1 < k - 1) {
    const double a = R[(i + 1) + (int64_t)(i + 1) * ldr], b = R[(i + 2) + (int64_t)(i + 1) * ldr];
    const double rho = hypot(a, b);
    rot[0] = rho > 0 ? a / rho : 1.0;
    rot[1] = rho > 0 ? b / rho : 0.0;
  }
  __syncthreads();
  for (int c0 = i + 1; c0 < k - 1; ++c0) {
    const double c = rot[0], s = rot[1];
    __syncthreads();                     // every thread holds this step's rotation
    for (int col = c0 + tid; col < k; col += nt) {
      double* Rc = R + (int64_t)col * ldr;
      const double x = Rc[c0], y = Rc[c0 + 1];
      Rc[c0] = c * x + s * y;
      Rc[c0 + 1] = col == c0 ? 0.0 : -s * x + c * y;
      if (col == c0 + 1 && c0 + 1 < k - 1) {
        const double a = Rc[c0 + 1], b = Rc[c0 + 2];
        const double rho = hypot(a, b);
        rot[0] = rho > 0 ? a / rho : 1.0;
        rot[1] = rho > 0 ? b / rho : 0.0;
      }
    }
    __syncthreads();
  }
}

// x = R(0:kk, 0:kk) \ b (upper triangular, column-oriented: one barrier per
// column); w is work space, b may alias neither
__device__ void trsv_upper(const double* R, int kk, int ldr, const double* b, double* w, double* x) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int l = tid; l < kk; l += nt) w[l] = b[l];
  __syncthreads();
  for (int q = kk - 1; q >= 0; --q) {
    const double xq = w[q] / R[q + (int64_t)q * ldr];
    if (tid == 0) x[q] = xq;
    for (int l = tid; l < q; l += nt) w[l] -= R[l + (int64_t)q * ldr] * xq;
    __syncthreads();
  }
}

// Gu-Eisenstat refinement with R11 in R (upper, leading dimension ldt) and
// T = R11 \ R12 (k x r), in O(k^2 + k r) per exchange like id.m:116-236
// with c2 = 0: the inverse row norms r2 (id.m:124; in vb already when
// have_r2, else formed when the first test fails) are updated by
// triangular solves (id.m:151-152, 191-192, 218); vb holds six n-vectors
__device__ void refine_large(int& k, int n, double* T, double* R, double* X, int ldt, int* perm,
                             const IdParams& prm, double tol_abs, double* vb, double* red,
                             int& iter, bool& conv, bool have_r2) {
  const int tid = threadIdx.x, nt = blockDim.x;
  double* r2 = vb;
  double* u = vb + n;
  double* w = vb + 2 * n;
  double* x = vb + 3 * n;
  double* cs = vb + 4 * n;
  double* sn = vb + 5 * n;
  const double f2 = prm.Tmax * prm.Tmax;
  const double cap = fmin(prm.rrqr_iter, 4.0 * n + 64.0);
  conv = false;
  while ((double)iter < cap) {
    const int r = n - k;
    double bv = -1.0;
    int bi = k * r;
    for (int e = tid; e < k * r; e += nt) {
      const int i = e % k, c = e / k;
      const double a = T[i + (int64_t)c * ldt];
      argmax_merge(bv, bi, a * a, e);
    }
    block_argmax(bv, bi, red);
    if (bv <= f2) { conv = true; break; }
    if (!have_r2) {                       // id.m:124: r2 = sum(inv(R11).^2, 2)
      inv_row_norms(R, k, ldt, X, ldt, r2);
      have_r2 = true;
    }
    ++iter;
    const int i = bi % k, j = bi / k;
    if (i < k - 1) {                      // id.m:136-143
      col_exchange(R, k, ldt, i, cs, sn);
      for (int c = tid; c < r; c += nt) {
        const double a = T[i + (int64_t)c * ldt];
        T[i + (int64_t)c * ldt] = T[k - 1 + (int64_t)c * ldt];
        T[k - 1 + (int64_t)c * ldt] = a;
      }
      if (tid == 0) {
        double a = r2[i]; r2[i] = r2[k - 1]; r2[k - 1] = a;
        int q = perm[i]; perm[i] = perm[k - 1]; perm[k - 1] = q;
      }
      __syncthreads();                    // rows i, k-1 of T before its columns move
    }
    if (j > 0) {                          // id.m:146-150
      for (int l = tid; l < k; l += nt) {
        const double a = T[l]; T[l] = T[l + (int64_t)j * ldt]; T[l + (int64_t)j * ldt] = a;
      }
      if (tid == 0) { int q = perm[k]; perm[k] = perm[k + j]; perm[k + j] = q; }
    }
    __syncthreads();
    // downdate r2 (id.m:152-153)
    const double* Rl = R + (int64_t)(k - 1) * ldt;
    trsv_upper(R, k - 1, ldt, Rl, w, x);
    const double dk = R[k - 1 + (int64_t)(k - 1) * ldt];
    for (int l = tid; l < k - 1; l += nt) r2[l] -= (x[l] / dk) * (x[l] / dk);
    for (int l = tid; l < k; l += nt) u[l] = T[l];
    __syncthreads();
    // the new last skeleton column of R11 is R11 u (id.m:168-170, 207-208)
    for (int l = tid; l < k; l += nt) {
      double acc = 0;
      for (int q = l; q < k; ++q) acc += R[l + (int64_t)q * ldt] * u[q];
      w[l] = acc;
    }
    __syncthreads();
    for (int l = tid; l < k; l += nt) R[l + (int64_t)(k - 1) * ldt] = w[l];
    if (tid == 0) { int q = perm[k - 1]; perm[k - 1] = perm[k]; perm[k] = q; }
    // Sherman-Morrison (id.m:212-213): column 0 of the half-updated T is
    // e_{k-1}; rho_c = T(k-1, c) / u(k-1)
    const double g = 1.0 / u[k - 1];
    for (int c = tid; c < r; c += nt) cs[c] = c == 0 ? g : T[k - 1 + (int64_t)c * ldt] * g;
    __syncthreads();
    for (int e = tid; e < k * r; e += nt) {
      const int l = e % k, c = e / k;
      double* Tc = T + (int64_t)c * ldt;
      if (l == k - 1) Tc[l] = cs[c];
      else Tc[l] = (c == 0 ? 0.0 : Tc[l]) - cs[c] * u[l];
    }
    __syncthreads();
    // update r2 with the new R11 (id.m:208-209)
    trsv_upper(R, k - 1, ldt, Rl, w, x);
    const double dn = R[k - 1 + (int64_t)(k - 1) * ldt];
    for (int l = tid; l < k - 1; l += nt) r2[l] += (x[l] / dn) * (x[l] / dn);
    if (tid == 0) r2[k - 1] = 1.0 / (dn * dn);
    __syncthreads();
    // drop a skeleton within tol of the others' span (id.m:215-231)
    double best = -1.0;
    int ib = k;
    for (int l = tid; l < k; l += nt) argmax_merge(best, ib, r2[l], l);
    block_argmax(best, ib, red);
    if (1.0 / sqrt(best) > tol_abs) continue;
    if (ib < k - 1) {
      col_exchange(R, k, ldt, ib, cs, sn);
      for (int c = tid; c < r; c += nt) {
        const double a = T[ib + (int64_t)c * ldt];
        T[ib + (int64_t)c * ldt] = T[k - 1 + (int64_t)c * ldt];
        T[k - 1 + (int64_t)c * ldt] = a;
      }
      if (tid == 0) {
        double a = r2[ib]; r2[ib] = r2[k - 1]; r2[k - 1] = a;
        int q = perm[ib]; perm[ib] = perm[k - 1]; perm[k - 1] = q;
      }
      __syncthreads();
    }
    // skeleton k-1 becomes the first redundant column (id.m:226-230)
    const int k1 = k - 1;
    trsv_upper(R, k1, ldt, R + (int64_t)k1 * ldt, w, x);
    const double d1 = R[k1 + (int64_t)k1 * ldt];
    for (int l = tid; l < k1; l += nt) r2[l] -= (x[l] / d1) * (x[l] / d1);
    for (int e = tid; e < k1 * r; e += nt) {          // shift columns right by one
      const int l = e % k1, c = e / k1;
      X[l + (int64_t)(c + 1) * ldt] = T[l + (int64_t)c * ldt] + x[l] * T[k1 + (int64_t)c * ldt];
    }
    for (int l = tid; l < k1; l += nt) X[l] = x[l];
    __syncthreads();
    for (int e = tid; e < k1 * (r + 1); e += nt) {
      const int l = e % k1, c = e / k1;
      T[l + (int64_t)c * ldt] = X[l + (int64_t)c * ldt];
    }
    __syncthreads();
    k = k1;
    if (k == 0) break;
  }
}

// Shared memory: vn1[n], vn2[n], v[max(m, n)], red[32]; with vglob, v (the
// QR) or vn1, vn2 and v (the tails) in global memory.
// MODE 0: run the column-pivoted QR on K first.  Otherwise the leading k
// rows of R are already there by pivot position, perm holds the pivot order
// and kout the rank, and only T and the refinement remain; R is in K
// (MODE 1, from k_id_reg) or, in FP64 with leading dimension n, in the
// third n x n slice of the box's T scratch (MODE 2, from k_gram_id; MODE 3
// as MODE 2 with T = R11 \ R12 already in the first slice, formed by a
// library triangular solve for the largest ranks, and the refinement done
// by refine_large in O(k^2 + k r) per exchange, six n-vectors per box in
// vglob).
template <typename S, int MODE>
__global__ void __launch_bounds__(256) k_id(const int* __restrict__ ids,
                                            const LvlBox* __restrict__ boxes,
                                            const int* __restrict__ mrows, S* __restrict__ Kar,
                                            int* __restrict__ permar, double* __restrict__ Tar,
                                            IdParams prm, int* __restrict__ kout,
                                            int* __restrict__ iterout, double* __restrict__ vglob,
                                            int64_t vstride) {
  extern __shared__ double sh[];
  const int b = ids[blockIdx.x];
  const LvlBox B = boxes[b];
  const int n = B.n, m = mrows[b], ld = B.mcap;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(tid / kWave), nw = nt / kWave;
  S* K = Kar + B.K_off;
  int* perm = permar + B.perm_off;
  constexpr bool QR = MODE == 0;
  // shared: norms (2n), then the QR's Householder vector (m, or in this
  // workgroup's slice of vglob when m rows do not fit) or the tails' k-vector
  // (at most n), then the reduction scratch
  // the tails (MODE >= 1) of boxes too wide for shared memory keep their
  // three n-vectors in this workgroup's slice of vglob instead
  const bool gv = !QR && vglob != nullptr;
  double* vb = gv ? vglob + blockIdx.x * vstride : sh;
  double* vn1 = vb;
  double* vn2 = vb + n;
  double* v = QR && vglob ? vglob + blockIdx.x * vstride : vb + 2 * n;
  double* red = gv ? sh : QR && !vglob ? sh + 2 * n + (m > n ? m : n) : sh + 3 * n;
  __shared__ int s_piv;

  if (QR)
    for (int j = tid; j < n; j += nt) perm[j] = j;
  if (m == 0 || n == 0) {           // id.m:64-68: nothing to compress against
    if (tid == 0) { kout[b] = 0; iterout[b] = 0; }
    return;
  }
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  int k = 0;
  double tol_abs = 0;
  if (QR) {

  for (int j = wid; j < n; j += nw) {
    double s = 0;
    for (int i = lane; i < m; i += kWave) { double a = (double)K[i + (int64_t)j * ld]; s += a * a; }
    s = wave_sum(s);
    if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
  }
  __syncthreads();

  const double tol3z = sqrt(Eps<S>::v);
  const int smax = m < n ? m : n;
  for (int s = 0; s < smax; ++s) {
    double bv = -1.0;
    int bi = n;
    for (int j = s + tid; j < n; j += nt) argmax_merge(bv, bi, vn1[j], j);
    block_argmax(bv, bi, red);
    const int piv = bi;
    if (piv != s) {
      for (int i = tid; i < m; i += nt) {
        S t = K[i + (int64_t)s * ld];
        K[i + (int64_t)s * ld] = K[i + (int64_t)piv * ld];
        K[i + (int64_t)piv * ld] = t;
      }
      if (tid == 0) {
        double t = vn1[s]; vn1[s] = vn1[piv]; vn1[piv] = t;
        t = vn2[s]; vn2[s] = vn2[piv]; vn2[piv] = t;
        int q = perm[s]; perm[s] = perm[piv]; perm[piv] = q;
      }
      __syncthreads();
    }
    double part = 0;
    for (int i = s + 1 + tid; i < m; i += nt) { double a = (double)K[i + (int64_t)s * ld]; part += a * a; }
    const double sigma = block_sum(part, red);
    const double alpha = (double)K[s + (int64_t)s * ld];
    const double nrm = sqrt(alpha * alpha + sigma);
    if (s == 0) tol_abs = nrm * prm.tol;      // id.m:106-107: relative to |R11|
    if (!(nrm > tol_abs)) break;               // id.m:108-111
    k = s + 1;
    double beta, tau, scal;
    if (sigma == 0) { beta = alpha; tau = 0; scal = 0; }
    else {
      beta = alpha >= 0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int i = s + tid; i < m; i += nt) v[i] = i == s ? 1.0 : (double)K[i + (int64_t)s * ld] * scal;
    __syncthreads();
    if (tid == 0) K[s + (int64_t)s * ld] = (S)beta;
    const bool last = (k == kmax) || (s + 1 == smax);
    for (int j = s + 1 + wid; j < n; j += nw) {
      S* col = K + (int64_t)j * ld;
      double dot = 0;
      for (int i = s + lane; i < m; i += kWave) dot += v[i] * (double)col[i];
      dot = wave_sum(dot) * tau;
      if (last) {
        if (lane == 0) col[s] = (S)((double)col[s] - dot);
        continue;
      }
      double rsj = 0, tail = 0;
      for (int i = s + lane; i < m; i += kWave) {
        S nvS = (S)((double)col[i] - dot * v[i]);
        col[i] = nvS;
        double nv = (double)nvS;
        if (i == s) rsj = nv; else tail += nv * nv;
      }
      rsj = __shfl(rsj, 0, kWave);
      tail = wave_sum(tail);
      if (lane == 0) {
        double nv1 = vn1[j];
        if (nv1 != 0) {
          double t = fabs(rsj) / nv1;
          t = fmax(0.0, (1.0 + t) * (1.0 - t));
          double q = nv1 / vn2[j];
          if (t * q * q <= tol3z) { vn1[j] = sqrt(tail); vn2[j] = vn1[j]; }
          else vn1[j] = nv1 * sqrt(t);
        }
      }
    }
    __syncthreads();
    if (last) break;
  }
  }

  // T = R11 \ R12 in double (id.m:114); T, R and the work matrix X are
  // slices of ldt x n with leading dimension ldt = B.ldt (n, or the rank
  // cap a rank model sized the box's buffers with)
  const int ldt = B.ldt;
  double* T = Tar + B.T_off;
  double* R = T + (int64_t)ldt * n;
  double* X = R + (int64_t)ldt * n;
  // leading rows of R by position
  auto rv = [&](int i, int j) -> double {
    return MODE >= 2 ? X[i + (int64_t)j * ldt] : (double)K[i + (int64_t)j * ld];
  };
  if (!QR) {
    k = kout[b];
    if (k > ldt) {                  // the ID ran past the buffer's rank cap: redone later
      if (tid == 0) iterout[b] = 0;
      return;
    }
    tol_abs = k > 0 ? fabs(rv(0, 0)) * prm.tol : 0.0;
  }
  const int r0 = n - k;
  for (int c = tid; c < (MODE == 3 ? 0 : r0); c += nt) {   // MODE 3: T is already formed
    for (int l = k - 1; l >= 0; --l) {
      double s = rv(l, k + c);
      for (int q = l + 1; q < k; ++q) s -= rv(l, q) * T[q + (int64_t)c * ldt];
      T[l + (int64_t)c * ldt] = s / rv(l, l);
    }
  }
  int iter = 0;
  bool conv = true;
  if (MODE == 3 && isfinite(prm.Tmax) && prm.rrqr_iter > 0 && k > 0 && k < n) {
    // inv(R11) is in R's slice (a library triangular inverse ran before
    // this kernel): its squared row norms first, then R11 itself
    double* r2 = vglob + blockIdx.x * vstride;
    for (int l = tid; l < k; l += nt) {
      double acc = 0;
      for (int c = l; c < k; ++c) acc += R[l + (int64_t)c * ldt] * R[l + (int64_t)c * ldt];
      r2[l] = acc;
    }
    __syncthreads();
    for (int e = tid; e < k * k; e += nt) {
      int i = e % k, j = e / k;
      R[i + (int64_t)j * ldt] = i <= j ? rv(i, j) : 0.0;
    }
    __syncthreads();
    refine_large(k, n, T, R, X, ldt, perm, prm, tol_abs, vglob + blockIdx.x * vstride, red, iter,
                 conv, true);
  } else if (isfinite(prm.Tmax) && prm.rrqr_iter > 0 && k > 0 && k < n) {
    // Gu-Eisenstat refinement in the fixed-precision form the reference
    // takes when R has no trailing block (id.m:116-236, c2 = 0).
    __syncthreads();
    for (int e = tid; e < k * k; e += nt) {
      int i = e % k, j = e / k;
      R[i + (int64_t)j * ldt] = i <= j ? rv(i, j) : 0.0;
    }
    __syncthreads();
    const double f2 = prm.Tmax * prm.Tmax;
    conv = false;
    double* t = vn1;        // reuse shared scratch: t[k], u[k]
    double* u = vn2;
    // Every exchange multiplies |det R11| by more than Tmax, so the loop
    // ends in exact arithmetic; the device also caps it at 4n + 64 swaps so
    // that rounding can never keep a workgroup spinning.
    const double cap = fmin(prm.rrqr_iter, 4.0 * n + 64.0);
    while ((double)iter < cap) {
      const int r = n - k;
      double bv = -1.0;
      int bi = k * r;
      for (int e = tid; e < k * r; e += nt) {
        int i = e % k, c = e / k;
        double a = T[i + (int64_t)c * ldt];
        argmax_merge(bv, bi, a * a, e);
      }
      block_argmax(bv, bi, red);
      if (bv <= f2) { conv = true; break; }
      ++iter;
      const int i = bi % k, j = bi / k;
      if (i < k - 1) swap_skel(perm, T, ldt, r, R, ldt, k, i, k - 1);
      if (j > 0) {
        if (tid == 0) { int q = perm[k]; perm[k] = perm[k + j]; perm[k + j] = q; }
        for (int l = tid; l < k; l += nt) {
          double a = T[l]; T[l] = T[l + (int64_t)j * ldt]; T[l + (int64_t)j * ldt] = a;
        }
        __syncthreads();
      }
      for (int l = tid; l < k; l += nt) t[l] = T[l];
      __syncthreads();
      // new last skeleton column in R-space: R * t
      for (int l = tid; l < k; l += nt) {
        double s = 0;
        for (int q = 0; q < k; ++q) s += R[l + (int64_t)q * ldt] * t[q];
        u[l] = s;
      }
      __syncthreads();
      // rank-one update of T for the exchange of skeleton k-1 and redundant 0
      const double g = 1.0 / t[k - 1];
      for (int c = tid; c < r; c += nt) {
        double* Tc = T + (int64_t)c * ldt;
        if (c == 0) {
          for (int l = 0; l < k - 1; ++l) Tc[l] = -t[l] * g;
          Tc[k - 1] = g;
        } else {
          double rho = Tc[k - 1] * g;
          for (int l = 0; l < k - 1; ++l) Tc[l] -= rho * t[l];
          Tc[k - 1] = rho;
        }
      }
      for (int l = tid; l < k; l += nt) R[l + (int64_t)(k - 1) * ldt] = u[l];
      if (tid == 0) { int q = perm[k - 1]; perm[k - 1] = perm[k]; perm[k] = q; }
      __syncthreads();
      small_qr(R, k, ldt, v, red);
      // id.m:210-229: drop a skeleton within tol of the others' span
      double* r2 = u;
      inv_row_norms(R, k, ldt, X, ldt, r2);
      double best = -1.0;
      int ib = k;
      for (int l = tid; l < k; l += nt) argmax_merge(best, ib, r2[l], l);
      block_argmax(best, ib, red);
      if (1.0 / sqrt(best) > tol_abs) continue;
      if (ib < k - 1) {
        swap_skel(perm, T, ldt, r, R, ldt, k, ib, k - 1);
        small_qr(R, k, ldt, v, red);
      }
      // skeleton k-1 becomes the first redundant column
      const int k1 = k - 1;
      for (int l = tid; l < k1; l += nt) t[l] = R[l + (int64_t)k1 * ldt];
      __syncthreads();
      if (tid == 0) {
        for (int l = k1 - 1; l >= 0; --l) {
          double s = t[l];
          for (int q = l + 1; q < k1; ++q) s -= R[l + (int64_t)q * ldt] * t[q];
          t[l] = s / R[l + (int64_t)l * ldt];
        }
      }
      __syncthreads();
      for (int e = tid; e < k1 * r; e += nt) {        // shift columns right by one
        int l = e % k1, c = e / k1;
        X[l + (int64_t)(c + 1) * ldt] = T[l + (int64_t)c * ldt] + t[l] * T[k1 + (int64_t)c * ldt];
      }
      for (int l = tid; l < k1; l += nt) X[l] = t[l];
      __syncthreads();
      for (int e = tid; e < k1 * (r + 1); e += nt) {
        int l = e % k1, c = e / k1;
        T[l + (int64_t)c * ldt] = X[l + (int64_t)c * ldt];
      }
      __syncthreads();
      k = k1;
      if (k == 0) break;
    }
  }
  if (tid == 0) { kout[b] = k; iterout[b] = conv ? iter : -1 - iter; }
}

// ------------------------------------------------------------- k_factor --
// Elimination of one box's redundant points (rskelf.m:218-261).
//
// With the skeleton-first order [sk rd] and K = A(slf,slf) + M, the two
// sparsification updates K(rd,:) -= T' K(sk,:) and K(:,rd) -= K(:,sk) T
// leave
//     Ksr' = Ksr - Kss T,   Krs' = Krs - T' Kss,
//     Krr' = Krr - T' Ksr' - Krs T,
// so only these blocks are formed, never the full n x n matrix.  Then
// Krr' = L U (partial pivoting, 'n') or L L' ('p'), E = Ksr'/U or Ksr'/L',
// G = L \ Krs'(p,:), and M(sk,sk) - E G (or - E E') is the new skeleton
// block.  The blocks live in shared memory when SH (one or two boxes per
// CU), else in the box's global scratch.
template <typename S, bool LU, bool SH, bool LDL>
__global__ void __launch_bounds__(1024) k_elim(const int* __restrict__ ids,
                                              const LvlBox* __restrict__ boxes,
                                              const BoxOut* __restrict__ outs,
                                              const Seg* __restrict__ csegs, int* pool,
                                              const int* __restrict__ permar,
                                              const double* __restrict__ Tar,
                                              const S* __restrict__ Mprev, S* __restrict__ Mcur,
                                              S* __restrict__ War, S* __restrict__ fval,
                                              int* __restrict__ fidx, PointsDev P,
                                              int* __restrict__ fail) {
  extern __shared__ double smem_d[];
  const int bidx = ids[blockIdx.x];
  const LvlBox B = boxes[bidx];
  const BoxOut O = outs[bidx];
  const int n = B.n, k = O.k, r = O.r;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ ChildMap cm;
  __shared__ double red[32];
  __shared__ S s_d;
  if (tid == 0) {
    int nc = B.ncs > 8 ? 8 : B.ncs;
    cm.nc = nc;
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      Seg g = csegs[B.cs0 + c];
      cm.start[c] = acc;
      cm.moff[c] = g.moff;
      acc += g.len;
    }
    cm.start[nc] = acc;
  }
  __syncthreads();
  const int* slf = pool + B.slf_off;
  if (r == 0) {                      // rskelf.m:216: no compression, keep M
    S* Mo = Mcur + O.M_off;
    for (int c = 0; c < n; ++c)
      for (int a = tid; a < n; a += nt) Mo[a + c * n] = mval(cm, Mprev, a, c);
    return;
  }
  const int* q = permar + B.perm_off;
  const double* Tg = Tar + B.T_off;  // k x r, leading dimension B.ldt
  // point data of slf in skeleton-first order, cached in shared memory, or
  // for the global-scratch class after the box's 2 n^2 matrix elements
  double* cx = SH ? smem_d : reinterpret_cast<double*>(War + B.W_off + 2 * (int64_t)n * n);
  double* cy = cx + n;
  double* cnx = cy + n;
  double* cny = cnx + n;
  double* cw = cny + n;
  double* cdg = cw + n;
  for (int a = tid; a < n; a += nt) {
    int pt = slf[q[a]];
    point_fields(P, pt, cx[a], cy[a], cnx[a], cny[a], cw[a]);
    cdg[a] = P.diag[pt];
  }
  __syncthreads();
  S* base = SH ? reinterpret_cast<S*>(smem_d + 6 * n) : War + B.W_off;
  S* Krr = base;                     // r x r
  S* Ksr = Krr + r * r;              // k x r  (becomes E)
  S* Krs = Ksr + k * r;              // r x k  (becomes G; LU only)
  S* Kss = Krs + (LU ? r * k : 0);   // k x k
  S* Ts = Kss + k * k;               // k x r
  const int lane = tid & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(tid / kWave), nw = nt / kWave;
  // rows over lanes, columns over waves: no integer division in the loops
#define FOR_COLS_ROWS(i, i0, i1, j, j0, j1) \
  for (int j = (j0) + wid; j < (j1); j += nw) for (int i = (i0) + lane; i < (i1); i += kWave)
  auto kval = [&](int a, int c) -> S {
    double v;
    if (P.kind == RSKX_KERNEL_CALLBACK) v = P.blk[B.A_off + q[a] + (int64_t)q[c] * n];
    else if (a == c) v = cdg[c];
    else if (P.kind == RSKX_KERNEL_LAPLACE3D_SL) v = sl3_entry(P, cx[a], cy[a], cnx[a], cx[c], cy[c], cnx[c], cw[c]);
    else {
      double dx = cx[a] - cx[c], dy = cy[a] - cy[c];
      v = P.scale * cw[c] * (dx * cnx[c] + dy * cny[c]) / (kTwoPi * (dx * dx + dy * dy));
    }
    return (S)v + (cm.moff[0] < 0 && cm.nc == 1 ? (S)0 : mval(cm, Mprev, q[a], q[c]));
  };
  FOR_COLS_ROWS(a, 0, k, c, 0, r) Ts[a + c * k] = (S)Tg[a + (int64_t)c * B.ldt];
  FOR_COLS_ROWS(a, 0, k, c, 0, k) Kss[a + c * k] = kval(a, c);
  // Krs (r x k): own region for LU; for Cholesky it passes through Ksr's
  S* KrsTmp = LU ? Krs : Ksr;
  FOR_COLS_ROWS(i, 0, r, c, 0, k) KrsTmp[i + c * r] = kval(k + i, c);
  __syncthreads();
  // C(i,j) (+)= sum_l A(i,l) B(l,j) over an r x r block, rows i in groups
  // of four per lane (independent accumulators), columns over waves;
  // A(i,l) = A[i*ai + l*al], B(l,j) = B[l*bl + j*bj]
  auto rankk = [&](S* C, const S* A, int ai, int al, const S* Bm, int bl, int bj, S sgn,
                   bool add) {
    for (int j = wid; j < r; j += nw) {
      const S* bcol = Bm + j * bj;
      for (int i0 = lane; i0 < r; i0 += 4 * kWave) {
        S acc[4] = {0, 0, 0, 0};
        const int i1 = i0 + kWave, i2 = i0 + 2 * kWave, i3 = i0 + 3 * kWave;
        const bool v1 = i1 < r, v2 = i2 < r, v3 = i3 < r;
        for (int l = 0; l < k; ++l) {
          const S b = bcol[l * bl];
          const S* a = A + l * al;
          acc[0] += a[i0 * ai] * b;
          if (v1) acc[1] += a[i1 * ai] * b;
          if (v2) acc[2] += a[i2 * ai] * b;
          if (v3) acc[3] += a[i3 * ai] * b;
        }
        S* c = C + j * r;
        if (add) {
          c[i0] += sgn * acc[0];
          if (v1) c[i1] += sgn * acc[1];
          if (v2) c[i2] += sgn * acc[2];
          if (v3) c[i3] += sgn * acc[3];
        } else {
          c[i0] = sgn * acc[0];
          if (v1) c[i1] = sgn * acc[1];
          if (v2) c[i2] = sgn * acc[2];
          if (v3) c[i3] = sgn * acc[3];
        }
      }
    }
  };
  if (!LU) {
    // Krr = -Krs T while Krs sits in Ksr's region
    rankk(Krr, KrsTmp, 1, r, Ts, 1, k, (S)-1, false);
    __syncthreads();
  }
  FOR_COLS_ROWS(a, 0, k, c, 0, r) Ksr[a + c * k] = kval(a, k + c);
  __syncthreads();
  FOR_COLS_ROWS(a, 0, k, c, 0, r) {                 // Ksr' = Ksr - Kss T
    S acc = 0;
    for (int l = 0; l < k; ++l) acc += Kss[a + l * k] * Ts[l + c * k];
    Ksr[a + c * k] -= acc;
  }
  __syncthreads();
  // Krr' = Krr - T' Ksr' - Krs T
  if (LU) {
    FOR_COLS_ROWS(i, 0, r, j, 0, r) Krr[i + j * r] = kval(k + i, k + j);
    __syncthreads();
    rankk(Krr, Krs, 1, r, Ts, 1, k, (S)-1, true);
  } else {
    FOR_COLS_ROWS(i, 0, r, j, 0, r) Krr[i + j * r] += kval(k + i, k + j);
  }
  __syncthreads();
  rankk(Krr, Ts, k, 1, Ksr, 1, k, (S)-1, true);
  if (LU) {                                         // Krs' = Krs - T' Kss
    __syncthreads();
    FOR_COLS_ROWS(i, 0, r, c, 0, k) {
      S acc = 0;
      for (int l = 0; l < k; ++l) acc += Ts[l + i * k] * Kss[l + c * k];
      Krs[i + c * r] -= acc;
    }
  }
  __syncthreads();

  int* piv = fidx + O.piv;
  if (LU) {
    for (int i = tid; i < r; i += nt) piv[i] = i;
    for (int s = 0; s < r; ++s) {    // right-looking LU with partial pivoting
      double bv = -1.0;
      int bi = r;
      for (int i = s + tid; i < r; i += nt) argmax_merge(bv, bi, fabs((double)Krr[i + s * r]), i);
      block_argmax(bv, bi, red);
      if (!(bv > 0)) { if (tid == 0) atomicOr(fail, 2); }
      if (bi != s && bi < r) {      // swap rows s and bi of Krr and Krs
        for (int c = tid; c < r + k; c += nt) {
          S* M = c < r ? Krr + c * r : Krs + (c - r) * r;
          S t = M[s]; M[s] = M[bi]; M[bi] = t;
        }
        if (tid == 0) { int t = piv[s]; piv[s] = piv[bi]; piv[bi] = t; }
        __syncthreads();
      }
      const S dinv = (S)1 / Krr[s + s * r];
      FOR_COLS_ROWS(i, s + 1, r, j, s + 1, r)
        Krr[i + j * r] -= Krr[i + s * r] * dinv * Krr[s + j * r];
      __syncthreads();
      for (int i = s + 1 + tid; i < r; i += nt) Krr[i + s * r] *= dinv;
      __syncthreads();
    }
  } else if (!LDL) {
    // right-looking Cholesky ('p')
    for (int s = 0; s < r; ++s) {
      if (tid == 0) {
        S d = Krr[s + s * r];
        if (!(d > (S)0)) { atomicOr(fail, 1); d = (S)1; }
        d = sqrt(d);
        Krr[s + s * r] = d;
        s_d = (S)1 / d;
      }
      __syncthreads();
      const S d2 = s_d * s_d;
      for (int j = s + 1 + wid; j < r; j += nw) {
        const S ljs = Krr[j + s * r] * d2;
        for (int i = j + lane; i < r; i += kWave) Krr[i + j * r] -= Krr[i + s * r] * ljs;
      }
      __syncthreads();
      for (int i = s + 1 + tid; i < r; i += nt) Krr[i + s * r] *= s_d;
      __syncthreads();
    }
  } else {
    // 'h': [L,U,p] = ldl(K(rd,rd),'vector'), rd = rd(p), T = T(:,p)
    // (rskelf.m:229-231): L D L' with Bunch-Kaufman pivots, 1 x 1 and 2 x 2
    // (the pivot rule of LAPACK's dsytf2 on the lower triangle, alpha =
    // (1 + sqrt(17)) / 8).  A symmetric interchange swaps the computed rows
    // of L, the trailing matrix, T's and Ksr''s columns and the rd order.
    // D goes to fD[0..r) (diagonal) and fD[r..2r) (subdiagonal: D(s+1,s)
    // at the first index s of a 2 x 2 block, 0 elsewhere).
    S* fD = fval + O.D;
    int* frd = fidx + O.rd;
    for (int i = tid; i < r; i += nt) frd[i] = slf[q[k + i]];
    __shared__ S s_e[3];
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    __syncthreads();
    for (int s = 0; s < r;) {
      double lam = -1.0;
      int imax = r;
      for (int i = s + 1 + tid; i < r; i += nt) argmax_merge(lam, imax, fabs((double)Krr[i + s * r]), i);
      block_argmax(lam, imax, red);
      lam = lam > 0 ? lam : 0.0;
      const double absakk = fabs((double)Krr[s + s * r]);
      int kstep = 1, kp = s;
      if (absakk < alpha * lam) {
        // the largest entry of row / column imax off its diagonal
        double sig = -1.0;
        int jm = r;
        for (int j = s + tid; j < r; j += nt)
          if (j != imax)
            argmax_merge(sig, jm, fabs((double)(j < imax ? Krr[imax + j * r] : Krr[j + imax * r])), j);
        block_argmax(sig, jm, red);
        sig = sig > 0 ? sig : 0.0;
        if (absakk * sig >= alpha * lam * lam) kp = s;
        else if (fabs((double)Krr[imax + imax * r]) >= alpha * sig) kp = imax;
        else { kstep = 2; kp = imax; }
      }
      if (s + kstep > r) kstep = 1;             // (cannot happen: imax > s)
      const int p1 = s + kstep - 1, p2 = kp;
      if (p2 != p1) {                           // symmetric interchange p1 <-> p2 (p1 < p2)
        auto swp = [](S& a, S& b) { const S t = a; a = b; b = t; };
        for (int c = tid; c < p1; c += nt) swp(Krr[p1 + c * r], Krr[p2 + c * r]);
        for (int i = p2 + 1 + tid; i < r; i += nt) swp(Krr[i + p1 * r], Krr[i + p2 * r]);
        for (int j = p1 + 1 + tid; j < p2; j += nt) swp(Krr[j + p1 * r], Krr[p2 + j * r]);
        for (int a = tid; a < k; a += nt) {
          swp(Ksr[a + p1 * k], Ksr[a + p2 * k]);
          swp(Ts[a + p1 * k], Ts[a + p2 * k]);
        }
        if (tid == 0) {
          swp(Krr[p1 + p1 * r], Krr[p2 + p2 * r]);
          const int t = frd[p1]; frd[p1] = frd[p2]; frd[p2] = t;
        }
        __syncthreads();
      }
      if (kstep == 1) {
        if (tid == 0) {
          S d = Krr[s + s * r];
          if (!(d != (S)0)) { atomicOr(fail, 4); d = (S)1; }
          fD[s] = d;
          fD[r + s] = (S)0;
          Krr[s + s * r] = (S)1;
          s_d = (S)1 / d;
        }
        __syncthreads();
        for (int j = s + 1 + wid; j < r; j += nw) {
          const S ljs = Krr[j + s * r] * s_d;
          for (int i = j + lane; i < r; i += kWave) Krr[i + j * r] -= Krr[i + s * r] * ljs;
        }
        __syncthreads();
        for (int i = s + 1 + tid; i < r; i += nt) Krr[i + s * r] *= s_d;
        __syncthreads();
      } else {
        if (tid == 0) {
          const S d11 = Krr[s + s * r], d21 = Krr[s + 1 + s * r], d22 = Krr[s + 1 + (s + 1) * r];
          S det = d11 * d22 - d21 * d21;
          if (!(det != (S)0)) { atomicOr(fail, 4); det = (S)1; }
          fD[s] = d11;
          fD[s + 1] = d22;
          fD[r + s] = d21;
          fD[r + s + 1] = (S)0;
          s_e[0] = d22 / det;                   // D^{-1} = [e0 e1; e1 e2]
          s_e[1] = -d21 / det;
          s_e[2] = d11 / det;
          Krr[s + s * r] = (S)1;
          Krr[s + 1 + (s + 1) * r] = (S)1;
          Krr[s + 1 + s * r] = (S)0;
        }
        __syncthreads();
        const S e0 = s_e[0], e1 = s_e[1], e2 = s_e[2];
        // trailing A(i,j) -= W(i,:) L(j,:)' with W = A(:, s:s+1), L = W D^{-1}
        for (int j = s + 2 + wid; j < r; j += nw) {
          const S w0 = Krr[j + s * r], w1 = Krr[j + (s + 1) * r];
          const S l0 = w0 * e0 + w1 * e1, l1 = w0 * e1 + w1 * e2;
          for (int i = j + lane; i < r; i += kWave)
            Krr[i + j * r] -= Krr[i + s * r] * l0 + Krr[i + (s + 1) * r] * l1;
        }
        __syncthreads();
        for (int i = s + 2 + tid; i < r; i += nt) {
          const S w0 = Krr[i + s * r], w1 = Krr[i + (s + 1) * r];
          Krr[i + s * r] = w0 * e0 + w1 * e1;
          Krr[i + (s + 1) * r] = w0 * e1 + w1 * e2;
        }
        __syncthreads();
      }
      s += kstep;
    }
  }
  // inverted 32 x 32 diagonal blocks of L (unit for LU) and U, one wave
  // per block, lane c computing column c of the inverse; the sweeps apply
  // them as small dense products instead of a 32-step substitution
  {
    const int nbk = (r + kDiagBlock - 1) / kDiagBlock;
    for (int bb = wid; bb < 2 * nbk; bb += nw) {
      const bool upper = bb >= nbk;
      if (upper && !LU) continue;
      const int b = upper ? bb - nbk : bb;
      const int j0 = b * kDiagBlock;
      const int bs = r - j0 < kDiagBlock ? r - j0 : kDiagBlock;
      const int c = lane;
      S x[kDiagBlock];
#pragma unroll
      for (int t = 0; t < kDiagBlock; ++t) x[t] = (S)0;
      if (!upper) {
#pragma unroll
        for (int t = 0; t < kDiagBlock; ++t) {
          if (t < bs) {
            S v = (t == c) ? (S)1 : (S)0;
#pragma unroll
            for (int u = 0; u < t; ++u) v -= Krr[(j0 + t) + (j0 + u) * r] * x[u];
            x[t] = (LU || LDL) ? v : v / Krr[(j0 + t) + (j0 + t) * r];
          }
        }
      } else {
#pragma unroll
        for (int tt = kDiagBlock - 1; tt >= 0; --tt) {
          if (tt < bs) {
            S v = (tt == c) ? (S)1 : (S)0;
#pragma unroll
            for (int u = tt + 1; u < kDiagBlock; ++u)
              if (u < bs) v -= Krr[(j0 + tt) + (j0 + u) * r] * x[u];
            x[tt] = v / Krr[(j0 + tt) + (j0 + tt) * r];
          }
        }
      }
      if (c < kDiagBlock) {
        S* dst = fval + (upper ? O.Ui : O.Li) + (int64_t)b * kDiagBlock * kDiagBlock;
#pragma unroll
        for (int t = 0; t < kDiagBlock; ++t) dst[t + c * kDiagBlock] = (t < bs && c < bs) ? x[t] : (S)0;
      }
    }
  }
  // E = Ksr'/U ('n'), Ksr'/L' ('p') or (Ksr'/L')/D ('h'); G = L \ Krs'(p,:)
  // ('n'); column sweeps
  for (int j = 0; j < r; ++j) {
    const S inv = (S)1 / Krr[j + j * r];
    for (int a = tid; a < k; a += nt) Ksr[a + j * k] *= inv;
    __syncthreads();
    FOR_COLS_ROWS(a, 0, k, c, j + 1, r) {
      S u = LU ? Krr[j + c * r] : Krr[c + j * r];
      Ksr[a + c * k] -= Ksr[a + j * k] * u;
    }
    if (LU)
      FOR_COLS_ROWS(i, j + 1, r, c, 0, k) Krs[i + c * r] -= Krr[i + j * r] * Krs[j + c * r];
    __syncthreads();
  }
  if (LDL) {                       // E = (Ksr'/L')/D with D's 1 x 1 and 2 x 2 blocks
    const S* fD = fval + O.D;
    FOR_COLS_ROWS(a, 0, k, c, 0, r) {
      const S b = fD[r + c];
      if (b != (S)0) {                 // block (c, c+1)
        const S d11 = fD[c], d22 = fD[c + 1], det = d11 * d22 - b * b;
        const S x0 = Ksr[a + c * k], x1 = Ksr[a + (c + 1) * k];
        Ksr[a + c * k] = (x0 * d22 - x1 * b) / det;
        Ksr[a + (c + 1) * k] = (x1 * d11 - x0 * b) / det;
      } else if (c == 0 || fD[r + c - 1] == (S)0) {
        Ksr[a + c * k] /= fD[c];
      }
    }
    __syncthreads();
  }
  // M = M(sk,sk) - E G, - E E' or - E D E' (rskelf.m:240-246)
  S* Mo = Mcur + O.M_off;
  FOR_COLS_ROWS(a, 0, k, c, 0, k) {
    S acc = 0;
    if (LU) for (int l = 0; l < r; ++l) acc += Ksr[a + l * k] * Krs[l + c * r];
    else if (LDL) {                   // E D E', D block diagonal
      const S* fD = fval + O.D;
      for (int l = 0; l < r; ++l) {
        acc += Ksr[a + l * k] * fD[l] * Ksr[c + l * k];
        const S b = fD[r + l];
        if (b != (S)0) acc += b * (Ksr[a + l * k] * Ksr[c + (l + 1) * k] + Ksr[a + (l + 1) * k] * Ksr[c + l * k]);
      }
    } else for (int l = 0; l < r; ++l) acc += Ksr[a + l * k] * Ksr[c + l * k];
    Mo[a + c * k] = mval(cm, Mprev, q[a], q[c]) - acc;
  }
#undef FOR_COLS_ROWS
  // store the factor (rskelf.m:249-257) and the new skeleton list
  S* fT = fval + O.T;
  for (int e = tid; e < k * r; e += nt) fT[e] = Ts[e];
  S* fLU = fval + O.LU;
  for (int e = tid; e < r * r; e += nt) fLU[e] = Krr[e];
  S* fE = fval + O.E;
  for (int e = tid; e < k * r; e += nt) fE[e] = Ksr[e];
  if (LU) {
    S* fG = fval + O.G;
    for (int e = tid; e < r * k; e += nt) fG[e] = Krs[e];
  }
  int* fsk = fidx + O.sk;
  int* frd = fidx + O.rd;
  for (int a = tid; a < k; a += nt) {
    int pt = slf[q[a]];
    fsk[a] = pt;
    pool[O.nl_off + a] = pt;
  }
  if (!LDL)                                      // 'h' wrote rd(p) itself
    for (int i = tid; i < r; i += nt) frd[i] = slf[q[k + i]];
}

// ------------------------------------------------------------- k_id_reg --
// Column-pivoted Householder QR of one box's K with K assembled straight
// into registers: thread (wave w, lane l) holds K(l + 64 a, w + 16 c),
// a < RA, c < CB, so m <= 64 RA and n <= 16 CB.  Columns never move: a
// position -> column map plays the role of the column swaps, and the pivot
// at each step is the first position with the largest remaining norm, as
// in k_id / LAPACK.  Arithmetic is FP64 on values stored as S.  On exit K's
// leading k rows hold R by pivot position, perm the pivot order and kout
// the rank; k_id<S, 1> then forms T and refines the ID.
template <typename S, int RA, int CB>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_id_reg(const int* __restrict__ ids, const LvlBox* __restrict__ boxes,
         const int* __restrict__ pool, const int* __restrict__ nbrar,
         const int* __restrict__ mnbr, int* __restrict__ mrows, S* __restrict__ Kar,
         int* __restrict__ permar, PointsDev P, LevelParams L, IdParams prm,
         int* __restrict__ kout) {
  extern __shared__ double sh[];
  const int b = ids[blockIdx.x];
  const LvlBox B = boxes[b];
  const int n = B.n, mn = mnbr[b];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  const int ld = B.mcap;
  const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  S* K = Kar + B.K_off;
  int* perm = permar + B.perm_off;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  S* stg = reinterpret_cast<S*>(sh);   // per-thread assembly staging, RA x 1024
  double* ppx = sh + (RA * 1024 * sizeof(S) + 7) / 8;   // proxy points
  double* ppy = ppx + L.p;
  double* ppz = ppy + L.p;
  double* sx = ppz + L.p;          // column points (fields as point_fields)
  double* sy = sx + n;
  double* snx = sy + n;
  double* sny = snx + n;
  double* sw = sny + n;
  double* rx = sw + n;             // neighbour (row) points
  double* ry = rx + mn;
  double* rnx = ry + mn;
  double* rny = rnx + mn;
  double* rw = rny + mn;
  double* vn1 = rw + mn;
  double* vn2 = vn1 + n;
  double* v = vn2 + n;             // Householder vector, m
  double* red = v + m;             // 32
  int* p2c = reinterpret_cast<int*>(red + 32);   // position -> column
  __shared__ double s_tau, s_tol;
  __shared__ int s_stop;
  for (int j = tid; j < n; j += 1024) {
    point_fields(P, slf[j], sx[j], sy[j], snx[j], sny[j], sw[j]);
    p2c[j] = j;
  }
  for (int i = tid; i < mn; i += 1024)
    point_fields(P, nbr[i], rx[i], ry[i], rnx[i], rny[i], rw[i]);
  for (int q = tid; q < L.p; q += 1024) proxy_point(L, B, q, ppx[q], ppy[q], ppz[q]);
  if (tid == 0) mrows[b] = m;
  __syncthreads();
  if (m == 0 || n == 0) {
    for (int j = tid; j < n; j += 1024) perm[j] = j;
    if (tid == 0) kout[b] = 0;
    return;
  }
  // K = [A(nbr,slf); A(slf,nbr)'; Kpxy] (rskelf.m:201-212), as k_assemble.
  // Each column block is evaluated by a rolled loop into this thread's own
  // LDS slots and then moved to registers, which keeps the entry
  // evaluation from inflating register pressure.
  const double c0 = P.scale / kTwoPi;
  const bool sl3 = P.kind == RSKX_KERNEL_LAPLACE3D_SL;
  S Kr[RA][CB];
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    const int j = wid + 16 * cb;
    const bool cj = j < n;
    const double xj = cj ? sx[j] : 0, yj = cj ? sy[j] : 0, nxj = cj ? snx[j] : 0,
                 nyj = cj ? sny[j] : 0, wj = cj ? sw[j] : 0;
#pragma unroll 1
    for (int ra = 0; ra < RA; ++ra) {
      const int i = lane + 64 * ra;
      double val = 0;
      if (cj && i < m) {
        if (i < mn) {
          if (sl3) {
            val = sl3_entry(P, rx[i], ry[i], rnx[i], xj, yj, nxj, wj);
          } else {
            const double dx = rx[i] - xj, dy = ry[i] - yj;
            val = c0 * wj * (dx * nxj + dy * nyj) / (dx * dx + dy * dy);
          }
        } else if (i < mt) {
          const int q = i - mn;
          if (sl3) {
            val = sl3_entry(P, xj, yj, nxj, rx[q], ry[q], rnx[q], rw[q]);
          } else {
            const double dx = xj - rx[q], dy = yj - ry[q];
            val = c0 * rw[q] * (dx * rnx[q] + dy * rny[q]) / (dx * dx + dy * dy);
          }
        } else {
          const int q = i - mt;
          if (L.pkind == RSKX_PROXY_POINTS_SL) {
            val = proxy_sl3(L, B, q, ppx[q], ppy[q], ppz[q], xj, yj, nxj, wj);
          } else {
            const double dx = ppx[q] - xj, dy = ppy[q] - yj;
            const double r2 = dx * dx + dy * dy;
            val = L.pkind == RSKX_PROXY_RING_SL ? -0.5 * log(r2) / kTwoPi * B.pw
                                                : wj * (dx * nxj + dy * nyj) / (kTwoPi * r2);
          }
        }
      }
      stg[ra * 1024 + tid] = (S)val;
    }
#pragma unroll
    for (int ra = 0; ra < RA; ++ra) Kr[ra][cb] = stg[ra * 1024 + tid];
  }
  // column norms
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    double sq = 0;
#pragma unroll
    for (int ra = 0; ra < RA; ++ra) { const double a = (double)Kr[ra][cb]; sq += a * a; }
    sq = wave_sum(sq);
    const int j = wid + 16 * cb;
    if (lane == 0 && j < n) { vn1[j] = sqrt(sq); vn2[j] = vn1[j]; }
  }
  __syncthreads();
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  const double tol3z = sqrt(Eps<S>::v);
  const int smax = m < n ? m : n;
  unsigned donemask = 0;            // columns of this wave already pivoted (bit cb)
  int k = 0;
  for (int s = 0; s < smax; ++s) {
    // pivot: largest remaining norm over positions s..n-1 (id.m:105)
    double bv = -1.0;
    int bi = n;
    for (int pos = s + tid; pos < n; pos += 1024) argmax_merge(bv, bi, vn1[p2c[pos]], pos);
    block_argmax(bv, bi, red);
    if (tid == 0 && bi != s) { const int t = p2c[s]; p2c[s] = p2c[bi]; p2c[bi] = t; }
    __syncthreads();
    const int pc = p2c[s], pw = pc & 15, pcb = pc >> 4;
    const int sa = s >> 6, sl = s & 63;
    if (wid == pw) {
      double colv[RA];
#pragma unroll
      for (int ra = 0; ra < RA; ++ra) {
        double val = 0;
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) if (cb == pcb) val = (double)Kr[ra][cb];
        colv[ra] = val;
      }
      double part = 0, alpha_l = 0;
#pragma unroll
      for (int ra = 0; ra < RA; ++ra) {
        const int i = lane + 64 * ra;
        if (i > s && i < m) part += colv[ra] * colv[ra];
        if (ra == sa) alpha_l = colv[ra];
      }
      const double sigma = wave_sum(part);
      const double alpha = __shfl(alpha_l, sl, 64);
      const double nrm = sqrt(alpha * alpha + sigma);
      if (s == 0 && lane == 0) s_tol = nrm * prm.tol;   // id.m:106-107
      const double tol_abs = s == 0 ? nrm * prm.tol : s_tol;
      const bool stop = !(nrm > tol_abs);                // id.m:108-111
      double beta, tau, scal;
      if (sigma == 0) { beta = alpha; tau = 0; scal = 0; }
      else { beta = alpha >= 0 ? -nrm : nrm; tau = (beta - alpha) / beta; scal = 1.0 / (alpha - beta); }
      if (!stop) {
#pragma unroll
        for (int ra = 0; ra < RA; ++ra) {
          const int i = lane + 64 * ra;
          if (i >= s && i < m) v[i] = i == s ? 1.0 : colv[ra] * scal;
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          if (cb == pcb)
#pragma unroll
            for (int ra = 0; ra < RA; ++ra) if (ra == sa && lane == sl) Kr[ra][cb] = (S)beta;
      }
      if (lane == 0) { s_tau = tau; s_stop = stop ? 1 : 0; }
    }
    if (wid == pw) donemask |= 1u << pcb;
    __syncthreads();
    if (s_stop) break;
    k = s + 1;
    const bool last = (k == kmax) || (s + 1 == smax);
    const double tau = s_tau;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const int j = wid + 16 * cb;
      if (j >= n || ((donemask >> cb) & 1u)) continue;
      double dot = 0;
#pragma unroll
      for (int ra = 0; ra < RA; ++ra) {
        const int i = lane + 64 * ra;
        if (i >= s && i < m) dot += v[i] * (double)Kr[ra][cb];
      }
      dot = wave_sum(dot) * tau;
      double rsj_l = 0, tail = 0;
#pragma unroll
      for (int ra = 0; ra < RA; ++ra) {
        const int i = lane + 64 * ra;
        if (i >= s && i < m && (!last || i == s)) {
          const S nv = (S)((double)Kr[ra][cb] - dot * v[i]);
          Kr[ra][cb] = nv;
          if (i == s) rsj_l = (double)nv;
          else tail += (double)nv * (double)nv;
        }
      }
      if (last) continue;
      const double rsj = __shfl(rsj_l, sl, 64);
      tail = wave_sum(tail);
      if (lane == 0) {
        const double nv1 = vn1[j];
        if (nv1 != 0) {
          double t = fabs(rsj) / nv1;
          t = fmax(0.0, (1.0 + t) * (1.0 - t));
          const double q = nv1 / vn2[j];
          if (t * q * q <= tol3z) { vn1[j] = sqrt(tail); vn2[j] = vn1[j]; }
          else vn1[j] = nv1 * sqrt(t);
        }
      }
    }
    __syncthreads();
    if (last) break;
  }
  // R's leading k rows by pivot position into K (leading dimension ld)
  int* c2p = p2c + n;                // column -> position
  for (int pos = tid; pos < n; pos += 1024) { c2p[p2c[pos]] = pos; perm[pos] = p2c[pos]; }
  __syncthreads();
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    const int j = wid + 16 * cb;
    if (j >= n) continue;
    const int pos = c2p[j];
#pragma unroll
    for (int ra = 0; ra < RA; ++ra) {
      const int i = lane + 64 * ra;
      if (i < k) K[i + (int64_t)pos * ld] = Kr[ra][cb];
    }
  }
  if (tid == 0) kout[b] = k;
}

// ------------------------------------------------------------ k_gram_id --
// Column ID through the Gram matrix, for FP32 factorizations at tolerances
// of 1e-7 and above.  In exact arithmetic column-pivoted QR of K and the
// diagonally pivoted Cholesky factorization of G = K'K are the same
// process: the column of largest remaining norm is the one of largest
// remaining diagonal of G, and R'R = G(p, p) with R QRCP's triangular factor
// up to row signs (which T = R11 \ R12 does not see).  K's entries are
// evaluated in FP64 and never stored; G is accumulated in FP64, each thread
// holding TPT 4 x 4 tiles of its lower triangle.  The remaining diagonal at
// the stopping threshold, tol^2 G_11 >= 1e-14 G_11, is then resolved to
// better than 1e-2 relative -- closer to the FP64 QRCP than a QRCP on
// FP32-rounded K, whose rounding alone moves |R_jj| by ~1e-7 |R_11|.
// On exit the third n x n slice of the box's T scratch holds R's leading k
// rows by pivot position (leading dimension n), perm the pivot order and
// kout the rank; k_id<S, 2> then forms T and refines the ID.
// rows of K staged in shared memory per accumulation step of k_gram_id
constexpr int kGramRows = 16;

// (classes with two tiles per thread held to 128 VGPRs: four waves per SIMD)
template <int NT, int TPT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TPT <= 2 ? 4 : 1, 8)))
k_gram_id(const int* __restrict__ ids, const LvlBox* __restrict__ boxes,
          const int* __restrict__ pool, const int* __restrict__ nbrar,
          const int* __restrict__ mnbr, int* __restrict__ mrows, double* __restrict__ Tar,
          int* __restrict__ permar, PointsDev P, LevelParams L, IdParams prm,
          int* __restrict__ kout) {
  constexpr int RC = kGramRows;      // rows of K per chunk
  constexpr int NW = NT / 64;
  extern __shared__ double sh[];
  const int b = ids[blockIdx.x];
  const LvlBox B = boxes[b];
  const int n = B.n, mn = mnbr[b];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  const int tid = threadIdx.x;
  const int n4 = (n + 3) >> 2, np = 4 * n4;
  int* perm = permar + B.perm_off;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  double* chunk = sh;                // RC x np, row-major (32-byte rows)
  double* dg = chunk + RC * np;      // remaining diagonal of G (np)
  double* prow = dg + np;            // pivot row of the updated G (np)
  double* red = prow + np;           // 2 NW
  double* ppx = red + 2 * NW;        // proxy points
  double* ppy = ppx + L.p;
  double* ppz = ppy + L.p;
  double* sx = ppz + L.p;            // column points (fields as point_fields)
  double* sy = sx + n;
  double* snx = sy + n;
  double* sny = snx + n;
  double* sw = sny + n;
  double* rx = sw + n;               // neighbour (row) points
  double* ry = rx + mn;
  double* rnx = ry + mn;
  double* rny = rnx + mn;
  double* rw = rny + mn;
  int* p2c = reinterpret_cast<int*>(rw + mn);        // position -> column (n)
  const int ldt = B.ldt;                             // R rows kept (the rank cap)
  double* Rc = Tar + B.T_off + (int64_t)ldt * n;     // R rows by column
  double* Rp = Rc + (int64_t)ldt * n;                // R rows by position
  for (int j = tid; j < n; j += NT) {
    point_fields(P, slf[j], sx[j], sy[j], snx[j], sny[j], sw[j]);
    p2c[j] = j;
  }
  for (int i = tid; i < mn; i += NT)
    point_fields(P, nbr[i], rx[i], ry[i], rnx[i], rny[i], rw[i]);
  for (int q = tid; q < L.p; q += NT) proxy_point(L, B, q, ppx[q], ppy[q], ppz[q]);
  for (int j = tid; j < np; j += NT) prow[j] = 0;
  if (tid == 0) mrows[b] = m;
  __syncthreads();
  if (m == 0 || n == 0) {
    for (int j = tid; j < n; j += NT) perm[j] = j;
    if (tid == 0) kout[b] = 0;
    return;
  }
  // this thread's tiles of the lower triangle, row-major tile order
  const int ntile = n4 * (n4 + 1) / 2;
  int ti[TPT], tj[TPT];
#pragma unroll
  for (int q = 0; q < TPT; ++q) {
    const int t = tid + q * NT;
    ti[q] = -1; tj[q] = 0;
    if (t < ntile) {
      int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
      while (i * (i + 1) / 2 > t) --i;
      while ((i + 1) * (i + 2) / 2 <= t) ++i;
      ti[q] = i; tj[q] = t - i * (i + 1) / 2;
    }
  }
  double acc[TPT][4][4];
#pragma unroll
  for (int q = 0; q < TPT; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[q][v][u] = 0;
  // G = K'K over chunks of RC rows of K = [A(nbr,slf); A(slf,nbr)'; Kpxy]
  // (rskelf.m:201-212), entries as in k_assemble
  const double c0 = P.scale / kTwoPi;
  const bool sl3 = P.kind == RSKX_KERNEL_LAPLACE3D_SL;
  // a thread's entries of a chunk: the first at (tid / np, tid % np), the
  // next ones by steps of NT = sR rows + sJ columns through the row-major
  // chunk (one division per chunk instead of one per entry)
  const int sR = NT / np, sJ = NT - sR * np;
  for (int r0 = 0; r0 < m; r0 += RC) {
    __syncthreads();
    int rr = tid / np, j = tid - rr * np;
    while (rr < RC) {
      const int i = r0 + rr;
      double val = 0;
      if (i < m && j < n) {
        const double xj = sx[j], yj = sy[j];
        if (i < mn) {
          if (sl3) {
            val = sl3_entry(P, rx[i], ry[i], rnx[i], xj, yj, snx[j], sw[j]);
          } else {
            const double dx = rx[i] - xj, dy = ry[i] - yj;
            val = c0 * sw[j] * (dx * snx[j] + dy * sny[j]) / (dx * dx + dy * dy);
          }
        } else if (i < mt) {
          const int q = i - mn;
          if (sl3) {
            val = sl3_entry(P, xj, yj, snx[j], rx[q], ry[q], rnx[q], rw[q]);
          } else {
            const double dx = xj - rx[q], dy = yj - ry[q];
            val = c0 * rw[q] * (dx * rnx[q] + dy * rny[q]) / (dx * dx + dy * dy);
          }
        } else {
          const int q = i - mt;
          if (L.pkind == RSKX_PROXY_POINTS_SL) {
            val = proxy_sl3(L, B, q, ppx[q], ppy[q], ppz[q], xj, yj, snx[j], sw[j]);
          } else {
            const double dx = ppx[q] - xj, dy = ppy[q] - yj;
            const double r2 = dx * dx + dy * dy;
            val = L.pkind == RSKX_PROXY_RING_SL ? -0.5 * log(r2) / kTwoPi * B.pw
                                                : sw[j] * (dx * snx[j] + dy * sny[j]) / (kTwoPi * r2);
          }
        }
      }
      chunk[rr * np + j] = val;
      j += sJ;
      rr += sR;
      if (j >= np) { j -= np; ++rr; }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      if (ti[q] < 0) continue;
      const double* ca = chunk + 4 * ti[q];
      const double* cb = chunk + 4 * tj[q];
#pragma unroll 4
      for (int rr = 0; rr < RC; ++rr) {
        const double4 a = *reinterpret_cast<const double4*>(ca + rr * np);
        const double4 c = *reinterpret_cast<const double4*>(cb + rr * np);
        const double av[4] = {a.x, a.y, a.z, a.w}, cv[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[q][v][u] += av[v] * cv[u];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPT; ++q)
    if (ti[q] >= 0 && ti[q] == tj[q])
#pragma unroll
      for (int v = 0; v < 4; ++v) dg[4 * ti[q] + v] = acc[q][v][v];
  __syncthreads();
  // diagonally pivoted Cholesky, right-looking on the register tiles,
  // stopped as id.m:105-111 stops the QR
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  const int kuser = kmax;
  if (kmax > ldt) kmax = ldt;            // the box's buffers hold ldt rows of R
  const int smax = m < n ? m : n;
  double tol_abs = 0;
  int k = 0;
  for (int s = 0; s < smax; ++s) {
    double bv = -1.0;
    int bi = n;
    for (int pos = s + tid; pos < n; pos += NT) argmax_merge(bv, bi, dg[p2c[pos]], pos);
    block_argmax(bv, bi, red);
    if (bi >= n) break;                // no finite candidate
    const int pc = p2c[bi];
    const double nrm = sqrt(fmax(bv, 0.0));
    if (s == 0) tol_abs = nrm * prm.tol;
    if (!(nrm > tol_abs)) break;
    k = s + 1;
    const int pt = pc >> 2, pl = pc & 3;
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      if (ti[q] < 0) continue;
      if (ti[q] == pt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (v == pl)
#pragma unroll
            for (int u = 0; u < 4; ++u) prow[4 * tj[q] + u] = acc[q][v][u];
      if (tj[q] == pt)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (u == pl)
#pragma unroll
            for (int v = 0; v < 4; ++v) prow[4 * ti[q] + v] = acc[q][v][u];
    }
    __syncthreads();
    if (tid == 0 && bi != s) { p2c[bi] = p2c[s]; p2c[s] = pc; }
    const double rinv = 1.0 / nrm;
    for (int j = tid; j < n; j += NT) {
      const double r = prow[j] * rinv;
      Rc[s + (int64_t)j * ldt] = r;
      dg[j] -= r * r;
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      if (ti[q] < 0) continue;
      const double4 a = *reinterpret_cast<const double4*>(prow + 4 * ti[q]);
      const double4 c = *reinterpret_cast<const double4*>(prow + 4 * tj[q]);
      const double r2 = rinv * rinv;
      const double av[4] = {a.x * r2, a.y * r2, a.z * r2, a.w * r2}, cv[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[q][v][u] -= av[v] * cv[u];
    }
    __syncthreads();
    if (k == kmax) break;
  }
  if (k == ldt && ldt < kuser && k < smax) {
    // stopped by the rank cap: if a remaining column is still above the
    // tolerance the true rank exceeds the cap (an under-predicted rank);
    // kout = ldt + 1 tells the host to redo this box with full buffers
    double bv = -1.0;
    int bi = n;
    for (int pos = k + tid; pos < n; pos += NT) argmax_merge(bv, bi, dg[p2c[pos]], pos);
    block_argmax(bv, bi, red);
    if (bi < n && sqrt(fmax(bv, 0.0)) > tol_abs) {
      if (tid == 0) kout[b] = ldt + 1;
      return;
    }
  }
  // R's leading k rows by pivot position
  for (int e = tid; e < k * n; e += NT) {
    const int i = e % k, pos = e / k;
    Rp[i + (int64_t)pos * ldt] = Rc[i + (int64_t)p2c[pos] * ldt];
  }
  for (int pos = tid; pos < n; pos += NT) perm[pos] = p2c[pos];
  if (tid == 0) kout[b] = k;
}

// k_gram_id on the matrix cores: the same column ID, with G = K'K
// accumulated by v_mfma_f64_16x16x4f64 into 16 x 16 tiles of its lower
// triangle (spread round-robin over the waves, MS per wave) from each
// 16-row chunk of K staged transposed in LDS (row j of the chunk = column j
// of K over the chunk's rows), so a tile's four products read their
// operands as two 32-byte rows per lane.  The pivoted Cholesky then works
// on the accumulator registers: lane l of a tile holds G(16 I + (l >> 4) +
// 4 v, 16 J + (l & 15)), v < 4.
template <int NW, int MS>
__global__ void __launch_bounds__(64 * NW)
k_gram_id_mf(const int* __restrict__ ids, const LvlBox* __restrict__ boxes,
             const int* __restrict__ pool, const int* __restrict__ nbrar,
             const int* __restrict__ mnbr, int* __restrict__ mrows, double* __restrict__ Tar,
             int* __restrict__ permar, PointsDev P, LevelParams L, IdParams prm,
             int* __restrict__ kout) {
  typedef Mf<double>::V V;
  constexpr int NT = 64 * NW;
  constexpr int RC = 16;                          // rows of K per chunk
  constexpr int LD = RC + kMfPad;                 // chunk row length (doubles)
  extern __shared__ double sh[];
  const int b = ids[blockIdx.x];
  const LvlBox B = boxes[b];
  const int n = B.n, mn = mnbr[b];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nt16 = (n + 15) >> 4, np = 16 * nt16;
  int* perm = permar + B.perm_off;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  double* chunk = sh;                // 2 x np x LD: chunk[j LD + rr] = K(r0 + rr, j)
  double* dg = chunk + 2 * np * LD;  // remaining diagonal of G (np)
  double* prow = dg + np;            // pivot row of the updated G (np)
  double* red = prow + np;           // 2 NW values, 2 NW indices
  double* ppx = red + 4 * NW;        // proxy points
  double* ppy = ppx + L.p;
  double* ppz = ppy + L.p;
  double* sx = ppz + L.p;            // column points (fields as point_fields)
  double* sy = sx + n;
  double* snx = sy + n;
  double* sny = snx + n;
  double* sw = sny + n;
  double* rx = sw + n;               // neighbour (row) points
  double* ry = rx + mn;
  double* rnx = ry + mn;
  double* rny = rnx + mn;
  double* rw = rny + mn;
  int* p2c = reinterpret_cast<int*>(rw + mn);        // position -> column (n)
  const int ldt = B.ldt;                             // R rows kept (the rank cap)
  double* Rc = Tar + B.T_off + (int64_t)ldt * n;     // R rows by column
  double* Rp = Rc + (int64_t)ldt * n;                // R rows by position
  for (int j = tid; j < n; j += NT) {
    point_fields(P, slf[j], sx[j], sy[j], snx[j], sny[j], sw[j]);
    p2c[j] = j;
  }
  for (int i = tid; i < mn; i += NT)
    point_fields(P, nbr[i], rx[i], ry[i], rnx[i], rny[i], rw[i]);
  for (int q = tid; q < L.p; q += NT) proxy_point(L, B, q, ppx[q], ppy[q], ppz[q]);
  for (int j = tid; j < np; j += NT) { prow[j] = 0; dg[j] = 0; }
  if (tid == 0) mrows[b] = m;
  __syncthreads();
  if (m == 0 || n == 0) {
    for (int j = tid; j < n; j += NT) perm[j] = j;
    if (tid == 0) kout[b] = 0;
    return;
  }
  // this wave's tiles (column-major order of the lower triangle); lane s of
  // tIv / tJv holds slot s, read back with readlane where used
  int tIv = -1, tJv = -1;
  {
    const int ntile = nt16 * (nt16 + 1) / 2;
    int t = wid + NW * lane;
    if (lane < MS && t < ntile) {
      int J = 0, cnt = nt16;
      while (t >= cnt) { t -= cnt; ++J; --cnt; }
      tIv = J + t;
      tJv = J;
    }
  }
  auto tI = [&](int s) -> int { return __builtin_amdgcn_readlane(tIv, s); };
  auto tJ = [&](int s) -> int { return __builtin_amdgcn_readlane(tJv, s); };
  V acc[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[s][v] = 0;
  // G = K'K over chunks of RC rows of K = [A(nbr,slf); A(slf,nbr)'; Kpxy]
  // (rskelf.m:201-212), entries as in k_assemble
  const double c0 = P.scale / kTwoPi;
  const bool sl3 = P.kind == RSKX_KERNEL_LAPLACE3D_SL;
  // entries of rows r0.. of K into a chunk buffer (transposed)
  auto eval_chunk = [&](int r0, double* ch) {
    for (int e = tid; e < RC * np; e += NT) {
      const int rr = e & (RC - 1), j = e >> 4, i = r0 + rr;
      double val = 0;
      if (i < m && j < n) {
        const double xj = sx[j], yj = sy[j];
        if (i < mn) {
          if (sl3) {
            val = sl3_entry(P, rx[i], ry[i], rnx[i], xj, yj, snx[j], sw[j]);
          } else {
            const double dx = rx[i] - xj, dy = ry[i] - yj;
            val = c0 * sw[j] * (dx * snx[j] + dy * sny[j]) / (dx * dx + dy * dy);
          }
        } else if (i < mt) {
          const int q = i - mn;
          if (sl3) {
            val = sl3_entry(P, xj, yj, snx[j], rx[q], ry[q], rnx[q], rw[q]);
          } else {
            const double dx = xj - rx[q], dy = yj - ry[q];
            val = c0 * rw[q] * (dx * rnx[q] + dy * rny[q]) / (dx * dx + dy * dy);
          }
        } else {
          const int q = i - mt;
          if (L.pkind == RSKX_PROXY_POINTS_SL) {
            val = proxy_sl3(L, B, q, ppx[q], ppy[q], ppz[q], xj, yj, snx[j], sw[j]);
          } else {
            const double dx = ppx[q] - xj, dy = ppy[q] - yj;
            const double r2 = dx * dx + dy * dy;
            val = L.pkind == RSKX_PROXY_RING_SL ? -0.5 * log(r2) / kTwoPi * B.pw
                                                : sw[j] * (dx * snx[j] + dy * sny[j]) / (kTwoPi * r2);
          }
        }
      }
      ch[j * LD + rr] = val;
    }
  };
  // two chunk buffers: the products of one chunk (matrix cores) are issued
  // before the next chunk's entries are evaluated (vector FP64), so the two
  // pipes overlap
  eval_chunk(0, chunk);
  for (int r0 = 0; r0 < m; r0 += RC) {
    asm volatile("" : "+v"(tIv), "+v"(tJv));
    __syncthreads();
    const double* cur = chunk + ((r0 / RC) & 1) * np * LD;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      if (s && !(s & 3)) asm volatile("" ::: "memory");
      const int I = tI(s);
      if (I < 0) continue;
      acc[s] = mac_rows<double, false>(acc[s], cur + 16 * I * LD, LD, cur + 16 * tJ(s) * LD,
                                       LD, RC, lane);
    }
    if (r0 + RC < m) eval_chunk(r0 + RC, chunk + (((r0 / RC) + 1) & 1) * np * LD);
  }
  const int lrow = lane & 15;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int I = tI(s);
    if (I < 0 || I != tJ(s)) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (Mf<double>::row(lane, v) == lrow) dg[16 * I + lrow] = acc[s][v];
  }
  __syncthreads();
  // diagonally pivoted Cholesky, right-looking on the register tiles,
  // stopped as id.m:105-111 stops the QR
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  const int kuser = kmax;
  if (kmax > ldt) kmax = ldt;            // the box's buffers hold ldt rows of R
  const int smax = m < n ? m : n;
  double tol_abs = 0;
  int k = 0;
  for (int s = 0; s < smax; ++s) {
    asm volatile("" : "+v"(tIv), "+v"(tJv));
    double bv = -1.0;
    int bi = n;
    for (int pos = s + tid; pos < n; pos += NT) argmax_merge(bv, bi, dg[p2c[pos]], pos);
    // block argmax through LDS barriers only (R's stores stay in flight);
    // the reduction slots alternate between steps
    wave_argmax(bv, bi);
    double* rv = red + (s & 1) * NW;
    int* ri = reinterpret_cast<int*>(red + 2 * NW) + (s & 1) * NW;
    if (lane == 0) { rv[wid] = bv; ri[wid] = bi; }
    lds_barrier();
    bv = rv[0];
    bi = ri[0];
    for (int w = 1; w < NW; ++w) argmax_merge(bv, bi, rv[w], ri[w]);
    if (bi >= n) break;                // no finite candidate
    const int pc = p2c[bi];
    const double nrm = sqrt(fmax(bv, 0.0));
    if (s == 0) tol_abs = nrm * prm.tol;
    if (!(nrm > tol_abs)) break;
    k = s + 1;
    // row pc of G: from the tiles of block row pc / 16 and block column pc / 16
    const int pt = pc >> 4, pl = pc & 15;
#pragma unroll
    for (int q = 0; q < MS; ++q) {
      const int I = tI(q), J = tJ(q);
      if (I < 0) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rv = Mf<double>::row(lane, v);
        if (I == pt && rv == pl) prow[16 * J + lrow] = acc[q][v];
        if (J == pt && lrow == pl) prow[16 * I + rv] = acc[q][v];
      }
    }
    lds_barrier();
    if (tid == 0 && bi != s) { p2c[bi] = p2c[s]; p2c[s] = pc; }
    const double rinv = 1.0 / nrm;
    for (int j = tid; j < n; j += NT) {
      const double r = prow[j] * rinv;
      Rc[s + (int64_t)j * ldt] = r;
      dg[j] -= r * r;
    }
    const double r2 = rinv * rinv;
#pragma unroll
    for (int q = 0; q < MS; ++q) {
      const int I = tI(q), J = tJ(q);
      if (I < 0) continue;
      const double cj = prow[16 * J + lrow] * r2;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[q][v] -= prow[16 * I + Mf<double>::row(lane, v)] * cj;
    }
    lds_barrier();
    if (k == kmax) break;
  }
  __syncthreads();                       // R's stores, before they are read back
  if (k == ldt && ldt < kuser && k < smax) {
    // stopped by the rank cap: if a remaining column is still above the
    // tolerance the true rank exceeds the cap (an under-predicted rank);
    // kout = ldt + 1 tells the host to redo this box with full buffers
    double bv = -1.0;
    int bi = n;
    for (int pos = k + tid; pos < n; pos += NT) argmax_merge(bv, bi, dg[p2c[pos]], pos);
    block_argmax(bv, bi, red);
    if (bi < n && sqrt(fmax(bv, 0.0)) > tol_abs) {
      if (tid == 0) kout[b] = ldt + 1;
      return;
    }
  }
  // R's leading k rows by pivot position
  for (int e = tid; e < k * n; e += NT) {
    const int i = e % k, pos = e / k;
    Rp[i + (int64_t)pos * ldt] = Rc[i + (int64_t)p2c[pos] * ldt];
  }
  for (int pos = tid; pos < n; pos += NT) perm[pos] = p2c[pos];
  if (tid == 0) kout[b] = k;
}

// the matrix-core Gram ID for every class (set from the host's tuning)
static bool g_gram_mfma = false;
void set_gram_mfma(bool on) { g_gram_mfma = on; }

void launch_gram_id(int cls, const int* ids, const LvlBox* boxes, const int* pool,
                    const int* nbrar, const int* mnbr, int* mrows, double* Tar, int* permar,
                    const PointsDev& P, const LevelParams& L, const IdParams& prm, int* kout,
                    int nb, size_t shmem, hipStream_t st) {
#define RSKX_GID(NTV, TPTV)                                                                      \
  {                                                                                              \
    auto kern = k_gram_id<NTV, TPTV>;                                                            \
    ensure_dyn_lds((const void*)kern, shmem);                                                     \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(NTV), shmem, st, ids, boxes, pool, nbrar, mnbr,      \
                       mrows, Tar, permar, P, L, prm, kout);                                     \
  }
#define RSKX_GIDM(NWV, MSV)                                                                     \
  {                                                                                              \
    auto kern = k_gram_id_mf<NWV, MSV>;                                                          \
    ensure_dyn_lds((const void*)kern, shmem);                                                     \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * NWV), shmem, st, ids, boxes, pool, nbrar, mnbr, \
                       mrows, Tar, permar, P, L, prm, kout);                                     \
  }
  if (g_gram_mfma) {
    if (cls == 0) RSKX_GIDM(4, 9) else if (cls == 1) RSKX_GIDM(8, 9) else RSKX_GIDM(8, 22)
  } else {
    if (cls == 0) RSKX_GID(256, 2) else if (cls == 1) RSKX_GID(512, 2) else RSKX_GID(512, 5)
  }
#undef RSKX_GIDM
#undef RSKX_GID
}

// k_gram_id classes: {threads, 4 x 4 tiles per thread}; the largest n each
// holds (n4 (n4 + 1) / 2 <= threads x tiles, n4 = ceil(n / 4))
int gram_id_nmax(int cls) { return cls == 0 ? 124 : cls == 1 ? 176 : 284; }
int gram_id_class(int n) {
  for (int c = 0; c < 3; ++c)
    if (n <= gram_id_nmax(c)) return c;
  return -1;
}
size_t gram_id_shmem(int cls, int n, int mn, int m, int p) {
  if (g_gram_mfma) {
    // two chunks np16 x (16 + pad), dg, prow, red, proxies, point data, p2c
    const int np = (n + 15) / 16 * 16, nw = cls == 0 ? 4 : 8;
    return (size_t)(2 * np * kMfLdp + 3 * p + 5 * n + 5 * mn + 2 * np + 4 * nw) * sizeof(double) +
           (size_t)n * sizeof(int) + 64;
  }
  const int np = (n + 3) / 4 * 4, nw = (cls == 0 ? 256 : 512) / 64;
  return (size_t)(kGramRows * np + 3 * p + 5 * n + 5 * mn + 2 * np + 2 * nw) * sizeof(double) +
         (size_t)n * sizeof(int) + 64;
}

template <typename S>
void launch_id_reg(int cls, const int* ids, const LvlBox* boxes, const int* pool, const int* nbrar,
                   const int* mnbr, int* mrows, S* Kar, int* permar, const PointsDev& P,
                   const LevelParams& L, const IdParams& prm, int* kout, int nb, size_t shmem,
                   hipStream_t st) {
#define RSKX_IDR(RAV, CBV)                                                                       \
  {                                                                                              \
    auto kern = k_id_reg<S, RAV, CBV>;                                                           \
    ensure_dyn_lds((const void*)kern, shmem);                                                     \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(1024), shmem, st, ids, boxes, pool, nbrar, mnbr,     \
                       mrows, Kar, permar, P, L, prm, kout);                                     \
  }
  if constexpr (sizeof(S) == 4) {
    if (cls == 0) RSKX_IDR(8, 8) else RSKX_IDR(4, 16)
  } else {
    if (cls == 0) RSKX_IDR(8, 4) else RSKX_IDR(4, 8)
  }
#undef RSKX_IDR
}

// register classes of k_id_reg: {rows, cols} limits
void id_reg_limits(size_t es, int cls, int* mlim, int* nlim) {
  if (es == 4) { *mlim = cls == 0 ? 512 : 256; *nlim = cls == 0 ? 128 : 256; }
  else { *mlim = cls == 0 ? 512 : 256; *nlim = cls == 0 ? 64 : 128; }
}

// ------------------------------------------------------------ launchers --
void launch_lists(const LvlBox* boxes, const Seg* nsegs, const Seg* csegs, int* pool, int* nbrar,
                  int* mnbr, const PointsDev& P, const LevelParams& L, int nb, hipStream_t st) {
  hipLaunchKernelGGL(k_lists, dim3(nb), dim3(256), 0, st, boxes, nsegs, csegs, pool, nbrar, mnbr,
                     P, L);
}

template <typename S>
void launch_assemble(const int* ids, const LvlBox* boxes, const int* pool, const int* nbrar,
                     const int* mnbr, int* mrows, S* Kar, const PointsDev& P, const LevelParams& L,
                     int nb, int nmax, hipStream_t st) {
  // enough workgroups per box to fill the chip, at least 8 columns each
  int ys = (nmax + 7) / 8;
  ys = std::max(1, std::min(ys, std::max(1, 2048 / std::max(nb, 1))));
  hipLaunchKernelGGL(k_assemble<S>, dim3(nb, ys), dim3(256), 0, st, ids, boxes, pool, nbrar, mnbr,
                     mrows, Kar, P, L);
}

template <typename S>
void launch_id(int mode, const int* ids, const LvlBox* boxes, const int* mrows, S* Kar,
               int* permar, double* Tar, const IdParams& prm, int* kout, int* iterout, int nb,
               size_t shmem, double* vglob, int64_t vstride, hipStream_t st) {
  auto kern = mode == 0 ? k_id<S, 0> : mode == 1 ? k_id<S, 1> : mode == 2 ? k_id<S, 2> : k_id<S, 3>;
  ensure_dyn_lds((const void*)kern, shmem);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(256), shmem, st, ids, boxes, mrows, Kar, permar, Tar, prm,
                     kout, iterout, vglob, vstride);
}

// elements of the register elimination's column region: two panel buffers
// of nb columns, or the staging of nw * spr columns, whichever is larger
__host__ __device__ inline int elim_reg_colb(int n, int nb, int nw, int spr) {
  const int a = 2 * nb * n, b = nw * spr * n;
  return a > b ? a : b;
}

// value of lane l (uniform l) broadcast to the wave
__device__ inline float lane_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_bcast(double v, int l) { return read_lane(v, l); }

// Elimination with the augmented matrix held in registers.
//
// In the order [rd; sk] the box's sparsified matrix with the skeleton
// block replaced by M(sk,sk) is
//     Z = [ Krr'  Krs' ]        Krr' = Krr - T' Ksr' - Krs T
//         [ Ksr'  Mss  ]        Ksr' = Ksr - Kss T,  Krs' = Krs - T' Kss
// and r steps of right-looking elimination on Z -- pivots chosen only
// among the first r rows for LU, none for Cholesky -- leave L and U
// (rskelf.m:225/235) in the leading block, G = L \ Krs'(p,:) and
// E = Ksr' / U (or Ksr' / L') beside and below it (rskelf.m:226-227,236),
// and Mss - E G (or Mss - E E') in the trailing block (rskelf.m:240-246).
//
// Thread (wave w, lane l) of NW waves holds Z(l + 64 a, w + NW b), a < TA,
// b < TB, so n <= min(64 TA, NW TB).  Two shapes: 16 waves for boxes up to
// the register limit, and 8 waves for smaller boxes, two workgroups per CU
// so that one box's serial steps (a panel's factorization, a pivot search)
// overlap the other's updates.  Pivoting is lazy: rows are never moved;
// each step's pivot row is found among the rows not yet used, ties going to
// the row LAPACK's swapped order would reach first.  (symm 'h', whose
// Bunch-Kaufman interchanges move rows and columns, runs in k_elim.)
template <typename S, bool LU, int TA, int TB, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 4 : 1, 4))) k_elim_reg(const int* __restrict__ ids,
                                                   const LvlBox* __restrict__ boxes,
                                                   const BoxOut* __restrict__ outs,
                                                   const Seg* __restrict__ csegs, int* pool,
                                                   const int* __restrict__ permar,
                                                   const double* __restrict__ Tar,
                                                   const S* __restrict__ Mprev, S* __restrict__ Mcur,
                                                   S* __restrict__ fval, int* __restrict__ fidx,
                                                   PointsDev P, int* __restrict__ fail) {
  extern __shared__ double smem_d[];
  const int bidx = ids[blockIdx.x];
  const LvlBox B = boxes[bidx];
  const BoxOut O = outs[bidx];
  const int n = B.n, k = O.k, r = O.r;
  const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NT = 64 * NW;
  // columns go to waves in panels of NB: slot tb of wave w holds column
  // ((tb / NB) * NW + w) * NB + tb % NB (NB = 1: column w + NW tb)
  constexpr int NB = LU ? 1 : 4;
  static_assert(TB % NB == 0, "TB must be a multiple of the panel width");
  auto colj = [&](int tb) { return ((tb / NB) * NW + wid) * NB + tb % NB; };
  __shared__ ChildMap cm;
  __shared__ int s_piv[2];
  __shared__ S s_pv[2];
  if (tid == 0) {
    int nc = B.ncs > 8 ? 8 : B.ncs;
    cm.nc = nc;
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      Seg g = csegs[B.cs0 + c];
      cm.start[c] = acc;
      cm.moff[c] = g.moff;
      acc += g.len;
    }
    cm.start[nc] = acc;
  }
  const int* slf = pool + B.slf_off;
  const int* q = permar + B.perm_off;           // q[0..k) skeletons, q[k..n) redundant
  const double* Tg = Tar + B.T_off;              // k x r, leading dimension B.ldt
  // shared layout
  double* cx = smem_d;
  double* cy = cx + n;
  double* cnx = cy + n;
  double* cny = cnx + n;
  double* cw = cny + n;
  double* cdg = cw + n;
  int* qp = reinterpret_cast<int*>(cdg + n);     // slf position of Z row/col p
  int* lpos = qp + n;                            // LU: logical position of physical row
  int* prow = lpos + n;                          // LU: physical row at logical position,
                                                 // then 2 x TB panel pivot rows
  // two panel buffers (NB columns each) and the staging of the transposed
  // copies (NW SPR columns) share one region
  constexpr int SPR = TB >= 16 ? 2 : 1;          // slots per staging round
  const int ncolb = elim_reg_colb(n, LU ? 1 : 4, NW, SPR);
  S* colb = reinterpret_cast<S*>(prow + n + 2 * TB + ((n + 2 * TB) & 1));
  S* Ts = colb + ncolb;                          // k x r
  S* Kss = Ts + k * r;                           // k x k
  S* KX = Kss + k * k;                           // k x r, then r x k
  for (int p = tid; p < n; p += NT) {
    const int sp = p < r ? q[k + p] : q[p - r];
    qp[p] = sp;
    const int pt = slf[sp];
    point_fields(P, pt, cx[p], cy[p], cnx[p], cny[p], cw[p]);
    cdg[p] = P.diag[pt];
    if (LU && p < r) { lpos[p] = p; prow[p] = p; }
  }
  __syncthreads();
  const bool leafbox = cm.nc == 1 && cm.moff[0] < 0;
  auto aval = [&](int p1, int p2) -> double {    // A entry between Z positions
    if (P.kind == RSKX_KERNEL_CALLBACK) return P.blk[B.A_off + qp[p1] + (int64_t)qp[p2] * n];
    if (p1 == p2) return cdg[p2];
    if (P.kind == RSKX_KERNEL_LAPLACE3D_SL)
      return sl3_entry(P, cx[p1], cy[p1], cnx[p1], cx[p2], cy[p2], cnx[p2], cw[p2]);
    double dx = cx[p1] - cx[p2], dy = cy[p1] - cy[p2];
    return P.scale * cw[p2] * (dx * cnx[p2] + dy * cny[p2]) / (kTwoPi * (dx * dx + dy * dy));
  };
  auto mv = [&](int p1, int p2) -> S { return leafbox ? (S)0 : mval(cm, Mprev, qp[p1], qp[p2]); };
  auto kv = [&](int p1, int p2) -> S { return (S)aval(p1, p2) + mv(p1, p2); };
  // T, Kss, Ksr (positions: rd j -> j, sk a -> r + a)
  for (int e = tid; e < k * r; e += NT) {
    const int a = e % k, j = e / k;
    Ts[e] = (S)Tg[a + (int64_t)j * B.ldt];
    KX[e] = kv(r + a, j);
  }
  for (int e = tid; e < k * k; e += NT) Kss[e] = kv(r + e % k, r + e / k);
  __syncthreads();
  for (int e = tid; e < k * r; e += NT) {      // Ksr' = Ksr - Kss T
    const int a = e % k, j = e / k;
    S acc = 0;
    for (int l = 0; l < k; ++l) acc += Kss[a + l * k] * Ts[l + j * k];
    KX[e] -= acc;
  }
  __syncthreads();
  S Z[TA][TB];
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int i = lane + 64 * ta, j = colj(tb);
      S z = 0;
      if (i < n && j < n) {
        if (i < r && j < r) z = kv(i, j);                    // Krr
        else if (i >= r && j < r) z = KX[(i - r) + j * k];   // Ksr'
        else if (i >= r && j >= r) z = mv(i, j);             // M(sk,sk)
      }
      Z[ta][tb] = z;
    }
  // Krr -= T' Ksr': per l, four T values per lane and sixteen Ksr' values
  // per wave feed 64 multiply-adds
  for (int l = 0; l < k; ++l) {
    S tv[TA];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) {
      const int i = lane + 64 * ta;
      tv[ta] = i < r ? Ts[l + i * k] : (S)0;
    }
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int j = colj(tb);
      const S kx = j < r ? KX[l + j * k] : (S)0;
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) Z[ta][tb] -= tv[ta] * kx;
    }
  }
  __syncthreads();
  for (int e = tid; e < r * k; e += NT) KX[e] = kv(e % r, r + e / r);   // Krs, r x k
  __syncthreads();
  for (int l = 0; l < k; ++l) {                  // Krr -= Krs T
    S kr[TA];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) {
      const int i = lane + 64 * ta;
      kr[ta] = i < r ? KX[i + l * r] : (S)0;
    }
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int j = colj(tb);
      const S tt = j < r ? Ts[l + j * k] : (S)0;
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) Z[ta][tb] -= kr[ta] * tt;
    }
  }
  __syncthreads();
  for (int e = tid; e < r * k; e += NT) {      // Krs' = Krs - T' Kss
    const int i = e % r, c = e / r;
    S acc = 0;
    for (int l = 0; l < k; ++l) acc += Ts[l + i * k] * Kss[l + c * k];
    KX[e] -= acc;
  }
  __syncthreads();
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int i = lane + 64 * ta, j = colj(tb);
      if (i < r && j >= r && j < n) Z[ta][tb] = KX[i + (j - r) * r];
    }

  // ---- Cholesky: r / NB panels, one barrier each.  The wave owning panel
  // p factors its NB columns in registers (pivots broadcast by readlane),
  // leaves L(:, panel) in LDS, and every wave then applies the rank-NB
  // update to its own trailing columns.
  if constexpr (!LU) {
    for (int p = 0; p * NB < r; ++p) {
      const int s0 = p * NB, s1 = s0 + NB < r ? s0 + NB : r;
      const int wp = p % NW, gp = p / NW;
      S* Lp = colb + (p & 1) * NB * n;           // L(i, s0 + c) at Lp[c * n + i]
      if (wid == wp) {
        S pc[NB][TA];
#pragma unroll
        for (int c = 0; c < NB; ++c)
#pragma unroll
          for (int ta = 0; ta < TA; ++ta) {
            S v = 0;
#pragma unroll
            for (int g = 0; g < TB / NB; ++g) if (g == gp) v = Z[ta][g * NB + c];
            pc[c][ta] = v;
          }
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          const int sc = s0 + c;
          if (sc >= s1) break;
          S dv = 0;
#pragma unroll
          for (int ta = 0; ta < TA; ++ta) if (ta == (sc >> 6)) dv = pc[c][ta];
          dv = lane_bcast(dv, sc & 63);
          // Cholesky pivot sqrt(d)
          if (!(dv > (S)0)) { if (lane == 0) atomicOr(fail, 1); dv = (S)1; }
          const S pv = sqrt(dv), inv = (S)1 / pv;
#pragma unroll
          for (int ta = 0; ta < TA; ++ta) {
            const int i = lane + 64 * ta;
            if (i > sc && i < n) pc[c][ta] *= inv;
            else if (i == sc) pc[c][ta] = pv;
          }
#pragma unroll
          for (int c2 = c + 1; c2 < NB; ++c2) {   // the panel's later columns
            const int j = s0 + c2;
            if (j >= s1) break;
            S u = 0;
#pragma unroll
            for (int ta = 0; ta < TA; ++ta) if (ta == (j >> 6)) u = pc[c][ta];
            u = lane_bcast(u, j & 63);
#pragma unroll
            for (int ta = 0; ta < TA; ++ta) {
              const int i = lane + 64 * ta;
              if (i >= j && i < n) pc[c2][ta] -= pc[c][ta] * u;
            }
          }
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
#pragma unroll
          for (int ta = 0; ta < TA; ++ta) {
            const int i = lane + 64 * ta;
#pragma unroll
            for (int g = 0; g < TB / NB; ++g) if (g == gp) Z[ta][g * NB + c] = pc[c][ta];
            if (i < n) Lp[c * n + i] = (s0 + c < s1 && i > s0 + c) ? pc[c][ta] : (S)0;
          }
      }
      __syncthreads();
      S li[NB][TA];
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
          const int i = lane + 64 * ta;
          li[c][ta] = i < n ? Lp[c * n + i] : (S)0;
        }
#pragma unroll
      for (int tb = 0; tb < TB; ++tb) {
        const int j = colj(tb);
        if (j < s1 || j >= n) continue;
        S u[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) u[c] = Lp[c * n + j];
        // entries above the diagonal of Krr, the Krs' block and padding
        // rows are never read again: update without masks
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
          if (64 * ta >= n) continue;
          S z = Z[ta][tb];
#pragma unroll
          for (int c = 0; c < NB; ++c) z -= li[c][ta] * u[c];
          Z[ta][tb] = z;
        }
      }
    }
  }

  // ---- LU: r elimination steps, one barrier each; column s is held by
  // wave s mod NW, slot s / NW
  bool done[TA];                                 // LU: row already used as a pivot
#pragma unroll
  for (int ta = 0; ta < TA; ++ta) done[ta] = false;
  for (int s = 0; LU && s < r; ++s) {
    S* cb = colb + (s & 1) * n;
    const int ws = s % NW, bs = s / NW;
    if (wid == ws) {
      S colv[TA];
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) {
        S v = 0;
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) if (tb == bs) v = Z[ta][tb];
        colv[ta] = v;
      }
      int pr;
      S pv;
      if (LU) {
        double best = -1.0;
        int bl = 1 << 30, bp = -1;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
          const int i = lane + 64 * ta;
          if (i < r && !done[ta]) {
            const double av = fabs((double)colv[ta]);
            const int lp = lpos[i];
            if (av > best || (av == best && lp < bl)) { best = av; bl = lp; bp = i; }
          }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const double b2 = __shfl_xor(best, o, 64);
          const int l2 = __shfl_xor(bl, o, 64), p2 = __shfl_xor(bp, o, 64);
          if (b2 > best || (b2 == best && l2 < bl)) { best = b2; bl = l2; bp = p2; }
        }
        // no candidate compared (all NaN): take the row at position s
        pr = bp >= 0 ? bp : prow[s];
        S pval = 0;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) if (ta == (pr >> 6)) pval = colv[ta];
        pv = __shfl(pval, pr & 63, 64);
        if (!(best > 0)) { if (lane == 0) atomicOr(fail, 2); pv = (S)1; }
      } else {
        pr = s;
        S dval = 0;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) if (ta == (s >> 6)) dval = colv[ta];
        dval = __shfl(dval, s & 63, 64);
        if (!(dval > (S)0)) { if (lane == 0) atomicOr(fail, 1); dval = (S)1; }
        pv = sqrt(dval);
      }
      const S inv = (S)1 / pv;
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) {
        const int i = lane + 64 * ta;
        bool below;
        if (LU) below = i < n && i != pr && !done[ta];
        else below = i > s && i < n;
        if (below) colv[ta] *= inv;
        else if (i == pr) colv[ta] = pv;
        if (below) cb[i] = colv[ta];
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) if (tb == bs && i < n) Z[ta][tb] = colv[ta];
      }
      if (lane == 0) { s_piv[s & 1] = pr; s_pv[s & 1] = pv; }
      if (LU && lane == 0) {
        // the rows at logical positions s and lpos[pr] trade places
        const int ls = lpos[pr], ps = prow[s];
        prow[ls] = ps; lpos[ps] = ls;
        prow[s] = pr; lpos[pr] = s;
      }
    }
    __syncthreads();
    const int pr = s_piv[s & 1];
    S li[TA];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) {
      const int i = lane + 64 * ta;
      bool live;
      if (LU) {
        live = i < n && i != pr && !done[ta];
        if (i == pr) done[ta] = true;
      } else {
        live = i > s && i < n;
      }
      li[ta] = live ? cb[i] : (S)0;
    }
    const int pa = pr >> 6, pl = pr & 63;
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int j = colj(tb);
      if (j <= s || j >= n) continue;
      S u;
      if (LU) {
        S v = 0;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) if (ta == pa) v = Z[ta][tb];
        u = __shfl(v, pl, 64);
      } else {
        u = cb[j];
      }
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) {
        const int i = lane + 64 * ta;
        if (LU || j <= i || (i >= r && j >= r)) Z[ta][tb] -= li[ta] * u;
      }
    }
  }
  __syncthreads();

  // ---- outputs (rskelf.m:249-257)
  S* fLU = fval + O.LU;
  S* fE = fval + O.E;
  S* fG = fval + O.G;
  S* Mo = Mcur + O.M_off;
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      const int i = lane + 64 * ta, j = colj(tb);
      if (i >= n || j >= n) continue;
      const S z = Z[ta][tb];
      if (i < r) {
        const int li = LU ? lpos[i] : i;
        if (j < r) fLU[li + j * r] = z;
        else if (LU) fG[li + (j - r) * r] = z;
      } else {
        if (j < r) fE[(i - r) + j * k] = z;
        else Mo[(i - r) + (j - r) * k] = z;
      }
    }
  // transposed copies for the solve sweeps (FacRec::LUt, Et, Gt, Tt), a
  // few register slots (NW columns each) at a time through LDS so that the
  // global writes run along rows
  {
    S* fLUt = fval + O.LUt;
    S* fEt = fval + O.Et;
    S* fGt = fval + O.Gt;
    static_assert(TB % SPR == 0, "slot rounds");
    S* sb = colb;                                // NW SPR x n staging
#pragma unroll
    for (int rd = 0; rd < TB / SPR; ++rd) {
      __syncthreads();
#pragma unroll
      for (int u = 0; u < SPR; ++u)
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
          const int i = lane + 64 * ta;
          if (i < n) sb[(SPR * wid + u) * n + i] = Z[ta][SPR * rd + u];
        }
      __syncthreads();
      for (int e = tid; e < NW * SPR * n; e += NT) {
        const int cs = e % (NW * SPR), i = e / (NW * SPR);
        const int w = cs / SPR, tb = SPR * rd + cs % SPR;
        const int j = ((tb / NB) * NW + w) * NB + tb % NB;
        if (j >= n) continue;
        const S v = sb[cs * n + i];
        if (i < r) {
          const int li = LU ? lpos[i] : i;
          if (j < r) fLUt[j + li * r] = v;
          else if (LU) fGt[(j - r) + li * k] = v;
        } else if (j < r) {
          fEt[j + (i - r) * r] = v;
        }
      }
    }
    S* fTt = fval + O.Tt;
    for (int e = tid; e < k * r; e += NT) {
      const int j = e % r, a = e / r;
      fTt[e] = Ts[a + j * k];
    }
  }
  int* piv = fidx + O.piv;
  if (LU)
    for (int p = tid; p < r; p += NT) piv[p] = prow[p];
  S* fT = fval + O.T;
  for (int e = tid; e < k * r; e += NT) fT[e] = Ts[e];
  int* fsk = fidx + O.sk;
  int* frd = fidx + O.rd;
  for (int a = tid; a < k; a += NT) {
    const int pt = slf[q[a]];
    fsk[a] = pt;
    pool[O.nl_off + a] = pt;
  }
  for (int i = tid; i < r; i += NT) frd[i] = slf[q[k + i]];
  __syncthreads();
  // inverted 32 x 32 diagonal blocks of L and U, from copies of the
  // diagonal blocks staged in the (now unused) panel/T/K scratch when they
  // fit, else read back from fLU
  {
    const int nbk = (r + kDiagBlock - 1) / kDiagBlock;
    constexpr int kBB = kDiagBlock * kDiagBlock;
    const bool staged = (size_t)nbk * kBB <= (size_t)ncolb + (size_t)2 * k * r + (size_t)k * k;
    S* stg = colb;
    if (staged) {
      for (int e2 = tid; e2 < nbk * kBB; e2 += NT) {
        const int b = e2 / kBB, t = (e2 % kBB) % kDiagBlock, u = (e2 % kBB) / kDiagBlock;
        const int gi = b * kDiagBlock + t, gj = b * kDiagBlock + u;
        stg[e2] = (gi < r && gj < r) ? fLU[gi + (int64_t)gj * r] : (S)0;
      }
    }
    __syncthreads();
    for (int bb = wid; bb < 2 * nbk; bb += NW) {
      const bool upper = bb >= nbk;
      if (upper && !LU) continue;
      const int b = upper ? bb - nbk : bb;
      const int j0 = b * kDiagBlock;
      const int bs = r - j0 < kDiagBlock ? r - j0 : kDiagBlock;
      const int c = lane;
      S x[kDiagBlock];
#pragma unroll
      for (int t = 0; t < kDiagBlock; ++t) x[t] = (S)0;
      // column c of the block's inverse by substitution; inlined once for
      // the LDS copy and once for the factor in global memory, so that
      // each call site's loads know their address space (a pointer that
      // may be either compiles to flat loads)
      auto invert = [&](const S* Lb, int ldb) __attribute__((always_inline)) {
        if (!upper) {
#pragma unroll
          for (int t = 0; t < kDiagBlock; ++t) {
            if (t < bs) {
              S v = (t == c) ? (S)1 : (S)0;
#pragma unroll
              for (int u = 0; u < t; ++u) v -= Lb[t + u * ldb] * x[u];
              x[t] = LU ? v : v / Lb[t + t * ldb];
            }
          }
        } else {
#pragma unroll
          for (int tt = kDiagBlock - 1; tt >= 0; --tt) {
            if (tt < bs) {
              S v = (tt == c) ? (S)1 : (S)0;
#pragma unroll
              for (int u = tt + 1; u < kDiagBlock; ++u)
                if (u < bs) v -= Lb[tt + u * ldb] * x[u];
              x[tt] = v / Lb[tt + tt * ldb];
            }
          }
        }
      };
      if (staged) invert(stg + (size_t)b * kBB, kDiagBlock);
      else invert(fLU + j0 + (int64_t)j0 * r, r);
      if (c < kDiagBlock) {
        S* dst = fval + (upper ? O.Ui : O.Li) + (int64_t)b * kDiagBlock * kDiagBlock;
        S* dstt = fval + (upper ? O.Uti : O.Lti) + (int64_t)b * kDiagBlock * kDiagBlock;
#pragma unroll
        for (int t = 0; t < kDiagBlock; ++t) {
          const S xv = (t < bs && c < bs) ? x[t] : (S)0;
          dst[t + c * kDiagBlock] = xv;
          dstt[c + t * kDiagBlock] = xv;
        }
      }
    }
  }
}

// register elimination shapes: TA row slots, TB column slots, NW waves
constexpr int kElimTA = 4, kElimTB32 = 16, kElimTB64 = 12;          // 16 waves
constexpr int kElimTA8_32 = 3, kElimTB8_32 = 24;                     // 8 waves, FP32
constexpr int kElimTA8_64 = 2, kElimTB8_64 = 16;                     // 8 waves, FP64

// largest box of each register elimination shape (nw: 8 or 16 waves)
int elim_reg_nmax(size_t es, int nw) {
  if (nw == 8) return es == 4 ? std::min(64 * kElimTA8_32, 8 * kElimTB8_32)
                              : std::min(64 * kElimTA8_64, 8 * kElimTB8_64);
  return es == 4 ? 256 : 192;
}

// shared bytes of k_elim_reg for a box
size_t elim_reg_shmem(int n, int k, int r, size_t es, int nw, bool lu) {
  const int TB = nw == 8 ? (es == 4 ? kElimTB8_32 : kElimTB8_64) : (es == 4 ? kElimTB32 : kElimTB64);
  const int spr = TB >= 16 ? 2 : 1;
  size_t b = (size_t)6 * n * sizeof(double) + (size_t)(3 * n + 2 * TB + 2) * sizeof(int);
  b += ((size_t)elim_reg_colb(n, lu ? 1 : 4, nw, spr) + (size_t)k * r + (size_t)k * k +
        (size_t)k * r) * es + 16;
  return b;
}

template <typename S>
void launch_elim_reg(bool lu, int nw, size_t shmem, const int* ids, int nb, const LvlBox* boxes,
                     const BoxOut* outs, const Seg* csegs, int* pool, const int* permar,
                     const double* Tar, const S* Mprev, S* Mcur, S* fval, int* fidx,
                     const PointsDev& P, int* fail, hipStream_t st) {
  constexpr bool F32 = sizeof(S) == 4;
  constexpr int TA16 = kElimTA, TB16 = F32 ? kElimTB32 : kElimTB64;
  constexpr int TA8 = F32 ? kElimTA8_32 : kElimTA8_64, TB8 = F32 ? kElimTB8_32 : kElimTB8_64;
#define RSKX_ELIMR(LUV, TAV, TBV, NWV)                                                          \
  {                                                                                             \
    auto kern = k_elim_reg<S, LUV, TAV, TBV, NWV>;                                              \
    ensure_dyn_lds((const void*)kern, shmem);                                                    \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * NWV), shmem, st, ids, boxes, outs, csegs, pool, \
                       permar, Tar, Mprev, Mcur, fval, fidx, P, fail);                          \
  }
  if (nw == 8) {
    if (lu) RSKX_ELIMR(true, TA8, TB8, 8) else RSKX_ELIMR(false, TA8, TB8, 8)
  } else {
    if (lu) RSKX_ELIMR(true, TA16, TB16, 16) else RSKX_ELIMR(false, TA16, TB16, 16)
  }
#undef RSKX_ELIMR
}

// shared-memory elements of one box's elimination blocks
size_t elim_elems(bool lu, int k, int r) {
  return (size_t)r * r + (size_t)k * r * (lu ? 3 : 2) + (size_t)k * k;
}

template <typename S>
void launch_elim(bool lu, bool ldl, int mode, size_t shmem, const int* ids, int nb, const LvlBox* boxes,
                 const BoxOut* outs, const Seg* csegs, int* pool, const int* permar,
                 const double* Tar, const S* Mprev, S* Mcur, S* War, S* fval, int* fidx,
                 const PointsDev& P, int* fail, hipStream_t st) {
  // mode 1: blocks in shared memory, mode 0: blocks in global scratch
#define RSKX_ELIM(LUV, SHV)                                                                   \
  {                                                                                           \
    auto kern = ldl ? k_elim<S, false, SHV, true> : k_elim<S, LUV, SHV, false>;               \
    ensure_dyn_lds((const void*)kern, shmem);                                                  \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(1024), shmem, st, ids, boxes, outs, csegs, pool,  \
                       permar, Tar, Mprev, Mcur, War, fval, fidx, P, fail);                   \
  }
  if (lu) { if (mode) RSKX_ELIM(true, true) else RSKX_ELIM(true, false) }
  else    { if (mode) RSKX_ELIM(false, true) else RSKX_ELIM(false, false) }
#undef RSKX_ELIM
}

#define RSKX_INST(S)                                                                          \
  template void launch_assemble<S>(const int*, const LvlBox*, const int*, const int*,          \
                                   const int*, int*, S*, const PointsDev&, const LevelParams&,  \
                                   int, int, hipStream_t);                                      \
  template void launch_id<S>(int, const int*, const LvlBox*, const int*, S*, int*, double*,    \
                             const IdParams&, int*, int*, int, size_t, double*, int64_t,       \
                             hipStream_t);                                                     \
  template void launch_id_reg<S>(int, const int*, const LvlBox*, const int*, const int*,        \
                                 const int*, int*, S*, int*, const PointsDev&, const LevelParams&, \
                                 const IdParams&, int*, int, size_t, hipStream_t);             \
  template void launch_elim_reg<S>(bool, int, size_t, const int*, int, const LvlBox*, const BoxOut*, \
                                   const Seg*, int*, const int*, const double*, const S*, S*, S*, \
                                   int*, const PointsDev&, int*, hipStream_t);                   \
  template void launch_elim<S>(bool, bool, int, size_t, const int*, int, const LvlBox*,       \
                               const BoxOut*, const Seg*, int*, const int*, const double*,     \
                               const S*, S*, S*, S*, int*, const PointsDev&, int*, hipStream_t);
RSKX_INST(double)
RSKX_INST(float)

}  // namespace rskx
