// Complex factorization kernels, for RSKX_KERNEL_HELMHOLTZ2D_CFIE: one
// workgroup per box of the level, the complex counterparts of k_assemble,
// k_id and k_elim (factor_kernels.hip).  Storage S is zf (complex FP32 path)
// or zd; every arithmetic step runs in zd, and the ID's pivots, norms and
// rank decisions are FP64 on K's entries as stored.
//
//   k_zassemble  K = [A(nbr,slf); A(slf,nbr)' ('n'); Kpxy] (rskelf.m:201-212)
//                with the combined-field kernel and ring proxies
//   k_zid        column ID of K by complex QRCP truncated at the tolerance
//                rank, T = R11 \ R12 and Gu-Eisenstat refinement (core/id.m)
//   k_zelim      sparsification with T' ('n') or T.' ('s', rskelf.m:220-221),
//                LU with partial pivoting, E, G and the Schur update
//                (rskelf.m:218-246)
#include "common.hpp"
#include "devutil.hpp"
#include "rskx.h"
#include "zcomplex.hpp"

namespace rskx {

namespace {
constexpr double kTwoPiZ = 6.283185307179586476925286766559;

// w (dG/dn_y(t, s) - i eta G(t, s)) for target t and source s with unit
// normal n and weight w, G = (i/4) H0^(1)(k r):
//   G = (-Y0(kr) + i J0(kr)) / 4,
//   dG/dn_y = (i k / 4) H1^(1)(kr) ((t - s) . n) / r
//           = (k c / 4) (-Y1(kr) + i J1(kr)),  c = ((t - s) . n) / r,
// so the entry is w ((eta J0 - k c Y1) + i (eta Y0 + k c J1)) / 4 (the
// Bessel functions by bessel01).
__device__ inline zd cfie_entry(double k, double eta, double tx, double ty, double sx, double sy,
                                double nx, double ny, double w) {
  const double dx = tx - sx, dy = ty - sy;
  const double r = sqrt(dx * dx + dy * dy);
  const double z = k * r, c = (dx * nx + dy * ny) / r;
  double J0, Y0, J1, Y1;
  bessel01(z, J0, Y0, J1, Y1);
  const double q = 0.25 * w;
  return zd{q * (eta * J0 - k * c * Y1), q * (eta * Y0 + k * c * J1)};
}

// proxy point q of box B on the ring of L.np points at L.radius box sizes
// (ie_circle.m:35,107) and the ring's outward unit normal there
__device__ inline void ring_point(const LevelParams& L, const LvlBox& B, int q, double& px,
                                  double& py, double& nx, double& ny) {
  const double th = (double)(q + 1) * kTwoPiZ / (double)L.np;
  const double c = cos(th), s = sin(th);
  px = L.radius * c * B.l[0] + B.ctr[0];
  py = L.radius * s * B.l[1] + B.ctr[1];
  const double ax = c / B.l[0], ay = s / B.l[1], a = sqrt(ax * ax + ay * ay);
  nx = ax / a;
  ny = ay / a;
}

__device__ inline zd wave_zsum(zd v) { return zd{wave_sum(v.re), wave_sum(v.im)}; }
}  // namespace

// ---------------------------------------------------------- k_zassemble --
constexpr int kZAsmRows = 512;

template <typename S>
__global__ void __launch_bounds__(256) k_zassemble(const int* __restrict__ ids,
                                                   const LvlBox* __restrict__ boxes,
                                                   const int* __restrict__ pool,
                                                   const int* __restrict__ nbrar,
                                                   const int* __restrict__ mnbr,
                                                   int* __restrict__ mrows, S* __restrict__ Kar,
                                                   PointsDev P, LevelParams L) {
  __shared__ double rf[5][kZAsmRows];
  const int bidx = ids[blockIdx.x];
  const LvlBox B = boxes[bidx];
  const int n = B.n, mn = mnbr[bidx];
  const int mt = L.two_sided ? 2 * mn : mn;
  const int m = mt + L.p;
  if (threadIdx.x == 0) mrows[bidx] = m;
  const int* slf = pool + B.slf_off;
  const int* nbr = nbrar + B.nbr_off;
  S* K = Kar + B.K_off;
  const int64_t ldk = B.mcap;
  const double k = P.k, eta = P.eta;
  for (int r0 = 0; r0 < mn; r0 += kZAsmRows) {
    const int rc = mn - r0 < kZAsmRows ? mn - r0 : kZAsmRows;
    __syncthreads();
    for (int i = threadIdx.x; i < rc; i += blockDim.x) {
      const int pt = nbr[r0 + i];
      rf[0][i] = P.x[0][pt];
      rf[1][i] = P.x[1][pt];
      rf[2][i] = P.nu[0][pt];
      rf[3][i] = P.nu[1][pt];
      rf[4][i] = P.w[pt];
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
      const int pj = slf[j];
      const double xj = P.x[0][pj], yj = P.x[1][pj], nxj = P.nu[0][pj], nyj = P.nu[1][pj], wj = P.w[pj];
      S* Kj = K + j * ldk + r0;
      for (int i = threadIdx.x; i < rc; i += blockDim.x)      // A(nbr, slf)
        Kj[i] = st<S>(cfie_entry(k, eta, rf[0][i], rf[1][i], xj, yj, nxj, nyj, wj));
      if (L.two_sided)
        for (int i = threadIdx.x; i < rc; i += blockDim.x)    // A(slf, nbr)'
          Kj[mn + i] = st<S>(conj(cfie_entry(k, eta, xj, yj, rf[0][i], rf[1][i], rf[2][i],
                                             rf[3][i], rf[4][i])));
    }
  }
  for (int j = 0; j < n; ++j) {                               // proxy rows
    const int pj = slf[j];
    const double xj = P.x[0][pj], yj = P.x[1][pj], nxj = P.nu[0][pj], nyj = P.nu[1][pj], wj = P.w[pj];
    S* Kj = K + j * ldk + mt;
    for (int i = threadIdx.x; i < L.p; i += blockDim.x) {
      double px, py, nx, ny;
      ring_point(L, B, i % L.np, px, py, nx, ny);
      const zd v = i < L.np ? cfie_entry(k, eta, px, py, xj, yj, nxj, nyj, wj)
                            : conj(cfie_entry(k, eta, xj, yj, px, py, nx, ny, B.pw));
      Kj[i] = st<S>(v);
    }
  }
}

// --------------------------------------------------------------- k_zid --
namespace {
// Householder reflector of x = (alpha, tail) with |tail|^2 = sigma:
// H = I - tau v v', v = (1, tail / (alpha - beta)), H' x = beta e_1 with
// beta = -sign(alpha) ||x|| (sign = alpha / |alpha|, 1 for alpha = 0) and
// real tau = 1 + |alpha| / ||x||.
struct Refl {
  zd beta, scal;
  double tau;
};
__device__ inline Refl make_refl(zd alpha, double sigma) {
  Refl h;
  const double aa = zabs(alpha), nrm = sqrt(aa * aa + sigma);
  if (sigma == 0) {
    h.beta = alpha;
    h.tau = 0;
    h.scal = zd{0, 0};
    return h;
  }
  const zd ph = aa > 0 ? zd{alpha.re / aa, alpha.im / aa} : zd{1, 0};
  h.beta = -nrm * ph;
  h.tau = 1.0 + aa / nrm;
  h.scal = zrecip(alpha - h.beta);
  return h;
}

// Householder QR of the k x k complex R (leading dimension ldr) in place;
// only the triangular factor is kept.
__device__ void zsmall_qr(zd* R, int k, int ldr, zd* v, double* red) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int s = 0; s < k - 1; ++s) {
    double part = 0;
    for (int i = s + 1 + tid; i < k; i += nt) part += abs2(R[i + s * ldr]);
    const double sigma = block_sum(part, red);
    if (sigma == 0) continue;
    const Refl h = make_refl(R[s + s * ldr], sigma);
    for (int i = s + tid; i < k; i += nt) v[i] = i == s ? zd{1, 0} : R[i + s * ldr] * h.scal;
    __syncthreads();
    for (int j = s + 1 + tid; j < k; j += nt) {
      zd dot{0, 0};
      for (int i = s; i < k; ++i) dot += cmul(v[i], R[i + j * ldr]);
      dot = h.tau * dot;
      for (int i = s; i < k; ++i) R[i + j * ldr] -= dot * v[i];
    }
    __syncthreads();
    if (tid == 0) R[s + s * ldr] = h.beta;
    for (int i = s + 1 + tid; i < k; i += nt) R[i + s * ldr] = zd{0, 0};
    __syncthreads();
  }
}

// squared row norms of inv(R), R upper triangular k x k; X is k x k scratch
__device__ void zinv_row_norms(const zd* R, int k, int ldr, zd* X, int ldx, double* r2) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = tid; c < k; c += nt) {
    for (int l = k - 1; l >= 0; --l) {
      if (l > c) { X[l + c * ldx] = zd{0, 0}; continue; }
      zd s = (l == c) ? zd{1, 0} : zd{0, 0};
      for (int q = l + 1; q <= c; ++q) s -= R[l + q * ldr] * X[q + c * ldx];
      X[l + c * ldx] = zdiv(s, R[l + l * ldr]);
    }
  }
  __syncthreads();
  for (int l = tid; l < k; l += nt) {
    double s = 0;
    for (int c = l; c < k; ++c) s += abs2(X[l + c * ldx]);
    r2[l] = s;
  }
  __syncthreads();
}

// swap skeleton positions a and b: perm entries, T rows, R columns
__device__ void zswap_skel(int* perm, zd* T, int ldt, int r, zd* R, int ldr, int k, int a, int b) {
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) { int t = perm[a]; perm[a] = perm[b]; perm[b] = t; }
  for (int c = tid; c < r; c += nt) {
    zd t = T[a + c * ldt]; T[a + c * ldt] = T[b + c * ldt]; T[b + c * ldt] = t;
  }
  for (int i = tid; i < k; i += nt) {
    zd t = R[i + a * ldr]; R[i + a * ldr] = R[i + b * ldr]; R[i + b * ldr] = t;
  }
  __syncthreads();
}
}  // namespace

// Complex column-pivoted QR of K (m x n in global memory, storage S),
// stopped at the tolerance rank (core/id.m:105-111: |R_jj| > tol |R_11|),
// then T = R11 \ R12 (id.m:114) and the Gu-Eisenstat refinement in its
// tol-mode form (id.m:116-236, max |T_ij| <= Tmax).  The pivot is the first
// column of largest remaining norm, as LAPACK's zgeqp3 takes it.  The box's
// T scratch holds, as zd with leading dimension n, T, R and a work matrix
// (3 n^2 zd).  Shared memory: vn1[n], vn2[n] (double), the Householder
// vector v[max(m, n)] (zd; in this workgroup's slice of vglob when m rows do
// not fit; at least 3 n entries), red[64].
template <typename S>
__global__ void __launch_bounds__(256) k_zid(const int* __restrict__ ids,
                                             const LvlBox* __restrict__ boxes,
                                             const int* __restrict__ mrows, S* __restrict__ Kar,
                                             int* __restrict__ permar, double* __restrict__ Tar,
                                             IdParams prm, int* __restrict__ kout,
                                             int* __restrict__ iterout, zd* __restrict__ vglob,
                                             int64_t vstride) {
  extern __shared__ double sh[];
  const int b = ids[blockIdx.x];
  const LvlBox B = boxes[b];
  const int n = B.n, m = mrows[b];
  const int64_t ldk = B.mcap;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(tid / kWave), nw = nt / kWave;
  S* K = Kar + B.K_off;
  int* perm = permar + B.perm_off;
  double* vn1 = sh;
  double* vn2 = sh + n;
  const int mv = m > 3 * n ? m : 3 * n;     // the refinement's k-vectors use 3 n
  zd* v = vglob ? vglob + blockIdx.x * vstride : reinterpret_cast<zd*>(sh + 2 * n);
  double* red = vglob ? sh + 2 * n : sh + 2 * n + 2 * mv;

  for (int j = tid; j < n; j += nt) perm[j] = j;
  if (m == 0 || n == 0) {           // id.m:64-68: nothing to compress against
    if (tid == 0) { kout[b] = 0; iterout[b] = 0; }
    return;
  }
  int kmax = prm.kmax;
  if (kmax <= 0 || kmax > n) kmax = n;
  int k = 0;
  double tol_abs = 0;
  for (int j = wid; j < n; j += nw) {
    double s = 0;
    for (int i = lane; i < m; i += kWave) s += abs2(ld(K[i + j * ldk]));
    s = wave_sum(s);
    if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
  }
  __syncthreads();
  const double tol3z = sqrt(Eps<typename cplx_base<S>::type>::v);
  const int smax = m < n ? m : n;
  for (int s = 0; s < smax; ++s) {
    double bv = -1.0;
    int bi = n;
    for (int j = s + tid; j < n; j += nt) argmax_merge(bv, bi, vn1[j], j);
    block_argmax(bv, bi, red);
    const int piv = bi;
    if (piv != s) {
      for (int i = tid; i < m; i += nt) {
        const S t = K[i + s * ldk];
        K[i + s * ldk] = K[i + piv * ldk];
        K[i + piv * ldk] = t;
      }
      if (tid == 0) {
        double t = vn1[s]; vn1[s] = vn1[piv]; vn1[piv] = t;
        t = vn2[s]; vn2[s] = vn2[piv]; vn2[piv] = t;
        const int q = perm[s]; perm[s] = perm[piv]; perm[piv] = q;
      }
      __syncthreads();
    }
    double part = 0;
    for (int i = s + 1 + tid; i < m; i += nt) part += abs2(ld(K[i + s * ldk]));
    const double sigma = block_sum(part, red);
    const zd alpha = ld(K[s + s * ldk]);
    const double nrm = sqrt(abs2(alpha) + sigma);
    if (s == 0) tol_abs = nrm * prm.tol;      // id.m:106-107: relative to |R11|
    if (!(nrm > tol_abs)) break;               // id.m:108-111
    k = s + 1;
    const Refl h = make_refl(alpha, sigma);
    for (int i = s + tid; i < m; i += nt) v[i] = i == s ? zd{1, 0} : ld(K[i + s * ldk]) * h.scal;
    __syncthreads();
    if (tid == 0) K[s + s * ldk] = st<S>(h.beta);
    const bool last = (k == kmax) || (s + 1 == smax);
    for (int j = s + 1 + wid; j < n; j += nw) {
      S* col = K + j * ldk;
      zd dot{0, 0};
      for (int i = s + lane; i < m; i += kWave) dot += cmul(v[i], ld(col[i]));
      dot = h.tau * wave_zsum(dot);
      if (last) {
        if (lane == 0) col[s] = st<S>(ld(col[s]) - dot);
        continue;
      }
      zd rsj{0, 0};
      double tail = 0;
      for (int i = s + lane; i < m; i += kWave) {
        const S nvS = st<S>(ld(col[i]) - dot * v[i]);
        col[i] = nvS;
        const zd nv = ld(nvS);
        if (i == s) rsj = nv; else tail += abs2(nv);
      }
      rsj = zd{__shfl(rsj.re, 0, kWave), __shfl(rsj.im, 0, kWave)};
      tail = wave_sum(tail);
      if (lane == 0) {
        const double nv1 = vn1[j];
        if (nv1 != 0) {                     // LAPACK's norm downdate (zlaqp2)
          double t = zabs(rsj) / nv1;
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

  // T = R11 \ R12 (id.m:114), leading dimension n
  zd* T = reinterpret_cast<zd*>(Tar + B.T_off);
  zd* R = T + (int64_t)n * n;
  zd* X = R + (int64_t)n * n;
  const int r0 = n - k;
  for (int c = tid; c < r0; c += nt) {
    for (int l = k - 1; l >= 0; --l) {
      zd s = ld(K[l + (k + c) * ldk]);
      for (int q = l + 1; q < k; ++q) s -= ld(K[l + q * ldk]) * T[q + (int64_t)c * n];
      T[l + (int64_t)c * n] = zdiv(s, ld(K[l + l * ldk]));
    }
  }
  int iter = 0;
  bool conv = true;
  if (isfinite(prm.Tmax) && prm.rrqr_iter > 0 && k > 0 && k < n) {
    __syncthreads();
    for (int e = tid; e < k * k; e += nt) {
      const int i = e % k, j = e / k;
      R[i + (int64_t)j * n] = i <= j ? ld(K[i + j * ldk]) : zd{0, 0};
    }
    __syncthreads();
    const double f2 = prm.Tmax * prm.Tmax;
    conv = false;
    zd* t = v;              // k-vectors in the Householder vector's storage
    zd* u = v + n;
    double* r2 = vn1;
    const double cap = fmin(prm.rrqr_iter, 4.0 * n + 64.0);
    while ((double)iter < cap) {
      const int r = n - k;
      double bv = -1.0;
      int bi = k * r;
      for (int e = tid; e < k * r; e += nt) {
        const int i = e % k, c = e / k;
        argmax_merge(bv, bi, abs2(T[i + (int64_t)c * n]), e);
      }
      block_argmax(bv, bi, red);
      if (bv <= f2) { conv = true; break; }
      ++iter;
      const int i = bi % k, j = bi / k;
      if (i < k - 1) zswap_skel(perm, T, n, r, R, n, k, i, k - 1);
      if (j > 0) {
        if (tid == 0) { const int q = perm[k]; perm[k] = perm[k + j]; perm[k + j] = q; }
        for (int l = tid; l < k; l += nt) {
          const zd a = T[l]; T[l] = T[l + (int64_t)j * n]; T[l + (int64_t)j * n] = a;
        }
        __syncthreads();
      }
      for (int l = tid; l < k; l += nt) t[l] = T[l];
      __syncthreads();
      for (int l = tid; l < k; l += nt) {     // new last skeleton column in R-space: R t
        zd s{0, 0};
        for (int q = 0; q < k; ++q) s += R[l + (int64_t)q * n] * t[q];
        u[l] = s;
      }
      __syncthreads();
      // rank-one update of T for the exchange of skeleton k-1 and redundant 0
      const zd g = zrecip(t[k - 1]);
      for (int c = tid; c < r; c += nt) {
        zd* Tc = T + (int64_t)c * n;
        if (c == 0) {
          for (int l = 0; l < k - 1; ++l) Tc[l] = -(t[l] * g);
          Tc[k - 1] = g;
        } else {
          const zd rho = Tc[k - 1] * g;
          for (int l = 0; l < k - 1; ++l) Tc[l] -= rho * t[l];
          Tc[k - 1] = rho;
        }
      }
      for (int l = tid; l < k; l += nt) R[l + (int64_t)(k - 1) * n] = u[l];
      if (tid == 0) { const int q = perm[k - 1]; perm[k - 1] = perm[k]; perm[k] = q; }
      __syncthreads();
      zsmall_qr(R, k, n, v + 2 * n, red);
      // id.m:210-229: drop a skeleton within tol of the others' span
      zinv_row_norms(R, k, n, X, n, r2);
      double best = -1.0;
      int ib = k;
      for (int l = tid; l < k; l += nt) argmax_merge(best, ib, r2[l], l);
      block_argmax(best, ib, red);
      if (1.0 / sqrt(best) > tol_abs) continue;
      if (ib < k - 1) {
        zswap_skel(perm, T, n, r, R, n, k, ib, k - 1);
        zsmall_qr(R, k, n, v + 2 * n, red);
      }
      const int k1 = k - 1;                   // skeleton k-1 becomes the first redundant
      for (int l = tid; l < k1; l += nt) t[l] = R[l + (int64_t)k1 * n];
      __syncthreads();
      if (tid == 0) {
        for (int l = k1 - 1; l >= 0; --l) {
          zd s = t[l];
          for (int q = l + 1; q < k1; ++q) s -= R[l + (int64_t)q * n] * t[q];
          t[l] = zdiv(s, R[l + (int64_t)l * n]);
        }
      }
      __syncthreads();
      for (int e = tid; e < k1 * r; e += nt) {
        const int l = e % k1, c = e / k1;
        X[l + (int64_t)(c + 1) * n] = T[l + (int64_t)c * n] + t[l] * T[k1 + (int64_t)c * n];
      }
      for (int l = tid; l < k1; l += nt) X[l] = t[l];
      __syncthreads();
      for (int e = tid; e < k1 * (r + 1); e += nt) {
        const int l = e % k1, c = e / k1;
        T[l + (int64_t)c * n] = X[l + (int64_t)c * n];
      }
      __syncthreads();
      k = k1;
      if (k == 0) break;
    }
  }
  if (tid == 0) { kout[b] = k; iterout[b] = conv ? iter : -1 - iter; }
}

// --------------------------------------------------------------- k_zelim --
namespace {
struct ZChildMap {
  int nc;
  int start[9];
  int64_t moff[8];
};

template <typename S>
__device__ inline zd zmval(const ZChildMap& cm, const S* Mprev, int a, int b) {
  for (int c = 0; c < cm.nc; ++c) {
    const int s0 = cm.start[c], s1 = cm.start[c + 1];
    if (a >= s0 && a < s1) {
      if (b < s0 || b >= s1 || cm.moff[c] < 0) return zd{0, 0};
      return ld(Mprev[cm.moff[c] + (a - s0) + (int64_t)(b - s0) * (s1 - s0)]);
    }
  }
  return zd{0, 0};
}
}  // namespace

// Elimination of one box's redundant points (rskelf.m:218-261), complex:
// with the skeleton-first order [sk rd], K = A(slf,slf) + M and T* = T'
// ('n') or T.' ('s'), the updates K(rd,:) -= T* K(sk,:), K(:,rd) -= K(:,sk) T
// leave Ksr' = Ksr - Kss T, Krs' = Krs - T* Kss, Krr' = Krr - T* Ksr' - Krs T;
// then Krr' = L U with partial pivoting by |re| + |im| (zgetrf's measure),
// E = Ksr' / U, G = L \ Krs'(p,:) and M(sk,sk) - E G.  Blocks in shared
// memory (SH) or in the box's global scratch; the point data of slf (x, y,
// normal, weight, complex diagonal) cached beside them.
template <typename S, bool SYM, bool SH>
__global__ void __launch_bounds__(1024) k_zelim(const int* __restrict__ ids,
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
  const int lane = tid & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(tid / kWave), nw = nt / kWave;
  __shared__ ZChildMap cm;
  __shared__ double red[32];
  if (tid == 0) {
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
  const int* slf = pool + B.slf_off;
  if (r == 0) {                      // rskelf.m:216: no compression, keep M
    S* Mo = Mcur + O.M_off;
    for (int c = 0; c < n; ++c)
      for (int a = tid; a < n; a += nt) Mo[a + c * n] = st<S>(zmval(cm, Mprev, a, c));
    return;
  }
  const int* q = permar + B.perm_off;
  const zd* Tg = reinterpret_cast<const zd*>(Tar + B.T_off);   // k x r, leading dimension n
  double* cx = SH ? smem_d : reinterpret_cast<double*>(War + B.W_off + 2 * (int64_t)n * n);
  double* cy = cx + n;
  double* cnx = cy + n;
  double* cny = cnx + n;
  double* cw = cny + n;
  double* cdr = cw + n;
  double* cdi = cdr + n;
  for (int a = tid; a < n; a += nt) {
    const int pt = slf[q[a]];
    cx[a] = P.x[0][pt];
    cy[a] = P.x[1][pt];
    cnx[a] = P.nu[0][pt];
    cny[a] = P.nu[1][pt];
    cw[a] = P.w[pt];
    cdr[a] = P.diag[pt];
    cdi[a] = P.diag_im[pt];
  }
  __syncthreads();
  S* base = SH ? reinterpret_cast<S*>(smem_d + 8 * n) : War + B.W_off;
  S* Krr = base;                     // r x r
  S* Ksr = Krr + r * r;              // k x r  (becomes E)
  S* Krs = Ksr + k * r;              // r x k  (becomes G)
  S* Kss = Krs + r * k;              // k x k
  S* Ts = Kss + k * k;               // k x r
  const double kk = P.k, eta = P.eta;
  auto kval = [&](int a, int c) -> zd {
    const zd v = a == c ? zd{cdr[c], cdi[c]}
                        : cfie_entry(kk, eta, cx[a], cy[a], cx[c], cy[c], cnx[c], cny[c], cw[c]);
    return v + zmval(cm, Mprev, q[a], q[c]);
  };
  // T* (i, l) = T(l, i)' ('n') or T(l, i) ('s')
  auto tstar = [&](int l, int i) -> zd {
    const zd t = ld(Ts[l + i * k]);
    return SYM ? t : conj(t);
  };
#define FOR_COLS_ROWS(i, i0, i1, j, j0, j1) \
  for (int j = (j0) + wid; j < (j1); j += nw) for (int i = (i0) + lane; i < (i1); i += kWave)
  FOR_COLS_ROWS(a, 0, k, c, 0, r) Ts[a + c * k] = st<S>(Tg[a + (int64_t)c * n]);
  FOR_COLS_ROWS(a, 0, k, c, 0, k) Kss[a + c * k] = st<S>(kval(a, c));
  FOR_COLS_ROWS(i, 0, r, c, 0, k) Krs[i + c * r] = st<S>(kval(k + i, c));
  FOR_COLS_ROWS(a, 0, k, c, 0, r) Ksr[a + c * k] = st<S>(kval(a, k + c));
  FOR_COLS_ROWS(i, 0, r, j, 0, r) Krr[i + j * r] = st<S>(kval(k + i, k + j));
  __syncthreads();
  FOR_COLS_ROWS(a, 0, k, c, 0, r) {                 // Ksr' = Ksr - Kss T
    zd acc{0, 0};
    for (int l = 0; l < k; ++l) acc += ld(Kss[a + l * k]) * ld(Ts[l + c * k]);
    Ksr[a + c * k] = st<S>(ld(Ksr[a + c * k]) - acc);
  }
  __syncthreads();
  FOR_COLS_ROWS(i, 0, r, j, 0, r) {                 // Krr' = Krr - T* Ksr' - Krs T
    zd acc{0, 0};
    for (int l = 0; l < k; ++l)
      acc += tstar(l, i) * ld(Ksr[l + j * k]) + ld(Krs[i + l * r]) * ld(Ts[l + j * k]);
    Krr[i + j * r] = st<S>(ld(Krr[i + j * r]) - acc);
  }
  __syncthreads();
  FOR_COLS_ROWS(i, 0, r, c, 0, k) {                 // Krs' = Krs - T* Kss
    zd acc{0, 0};
    for (int l = 0; l < k; ++l) acc += tstar(l, i) * ld(Kss[l + c * k]);
    Krs[i + c * r] = st<S>(ld(Krs[i + c * r]) - acc);
  }
  __syncthreads();

  int* piv = fidx + O.piv;
  for (int i = tid; i < r; i += nt) piv[i] = i;
  __shared__ zd s_dinv;
  for (int s = 0; s < r; ++s) {      // right-looking LU with partial pivoting
    double bv = -1.0;
    int bi = r;
    for (int i = s + tid; i < r; i += nt) argmax_merge(bv, bi, cabs1(ld(Krr[i + s * r])), i);
    block_argmax(bv, bi, red);
    if (!(bv > 0)) { if (tid == 0) atomicOr(fail, 2); }
    if (bi != s && bi < r) {         // swap rows s and bi of Krr and Krs
      for (int c = tid; c < r + k; c += nt) {
        S* M = c < r ? Krr + c * r : Krs + (c - r) * r;
        const S t = M[s]; M[s] = M[bi]; M[bi] = t;
      }
      if (tid == 0) { const int t = piv[s]; piv[s] = piv[bi]; piv[bi] = t; }
      __syncthreads();
    }
    if (tid == 0) {
      const zd d = ld(Krr[s + s * r]);
      s_dinv = bv > 0 ? zrecip(d) : zd{1, 0};
    }
    __syncthreads();
    const zd dinv = s_dinv;
    FOR_COLS_ROWS(i, s + 1, r, j, s + 1, r)
      Krr[i + j * r] = st<S>(ld(Krr[i + j * r]) - ld(Krr[i + s * r]) * dinv * ld(Krr[s + j * r]));
    __syncthreads();
    for (int i = s + 1 + tid; i < r; i += nt) Krr[i + s * r] = st<S>(ld(Krr[i + s * r]) * dinv);
    __syncthreads();
  }
  // E = Ksr' / U, G = L \ Krs'(p,:): column sweeps
  for (int j = 0; j < r; ++j) {
    const zd inv = zrecip(ld(Krr[j + j * r]));
    for (int a = tid; a < k; a += nt) Ksr[a + j * k] = st<S>(ld(Ksr[a + j * k]) * inv);
    __syncthreads();
    FOR_COLS_ROWS(a, 0, k, c, j + 1, r)
      Ksr[a + c * k] = st<S>(ld(Ksr[a + c * k]) - ld(Ksr[a + j * k]) * ld(Krr[j + c * r]));
    FOR_COLS_ROWS(i, j + 1, r, c, 0, k)
      Krs[i + c * r] = st<S>(ld(Krs[i + c * r]) - ld(Krr[i + j * r]) * ld(Krs[j + c * r]));
    __syncthreads();
  }
  // M = M(sk,sk) - E G (rskelf.m:240-246)
  S* Mo = Mcur + O.M_off;
  FOR_COLS_ROWS(a, 0, k, c, 0, k) {
    zd acc{0, 0};
    for (int l = 0; l < r; ++l) acc += ld(Ksr[a + l * k]) * ld(Krs[l + c * r]);
    Mo[a + c * k] = st<S>(zmval(cm, Mprev, q[a], q[c]) - acc);
  }
#undef FOR_COLS_ROWS
  // store the factor (rskelf.m:249-257) and the new skeleton list
  S* fT = fval + O.T;
  for (int e = tid; e < k * r; e += nt) fT[e] = Ts[e];
  S* fLU = fval + O.LU;
  for (int e = tid; e < r * r; e += nt) fLU[e] = Krr[e];
  S* fE = fval + O.E;
  for (int e = tid; e < k * r; e += nt) fE[e] = Ksr[e];
  S* fG = fval + O.G;
  for (int e = tid; e < r * k; e += nt) fG[e] = Krs[e];
  int* fsk = fidx + O.sk;
  int* frd = fidx + O.rd;
  for (int a = tid; a < k; a += nt) {
    const int pt = slf[q[a]];
    fsk[a] = pt;
    pool[O.nl_off + a] = pt;
  }
  for (int i = tid; i < r; i += nt) frd[i] = slf[q[k + i]];
}

// ---------------------------------------------------------- gather diag --
// the complex diagonal (interleaved re, im, caller's order) into tree order
__global__ void k_zgather_diag(const double* __restrict__ diag, const int* __restrict__ order,
                               int64_t N, double* __restrict__ dre, double* __restrict__ dim) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = order[p];
    dre[p] = diag[2 * q];
    dim[p] = diag[2 * q + 1];
  }
}

// ------------------------------------------------------------ launchers --
size_t zelim_elems(int k, int r) {
  return (size_t)r * r + 3 * (size_t)k * r + (size_t)k * k;
}

template <typename S>
void launch_zassemble(const int* ids, const LvlBox* boxes, const int* pool, const int* nbr,
                      const int* mnbr, int* mrows, S* K, const PointsDev& P, const LevelParams& L,
                      int nb, hipStream_t st) {
  hipLaunchKernelGGL(k_zassemble<S>, dim3(nb), dim3(256), 0, st, ids, boxes, pool, nbr, mnbr,
                     mrows, K, P, L);
}

template <typename S>
void launch_zid(const int* ids, const LvlBox* boxes, const int* mrows, S* K, int* perm, double* T,
                const IdParams& prm, int* kout, int* iter, int nb, size_t shmem, zd* vglob,
                int64_t vstride, hipStream_t st) {
  ensure_dyn_lds((const void*)k_zid<S>, shmem);
  hipLaunchKernelGGL(k_zid<S>, dim3(nb), dim3(256), shmem, st, ids, boxes, mrows, K, perm, T, prm,
                     kout, iter, vglob, vstride);
}

template <typename S>
void launch_zelim(bool sym, bool sh, size_t shmem, const int* ids, int nb, const LvlBox* boxes,
                  const BoxOut* outs, const Seg* csegs, int* pool, const int* perm, const double* T,
                  const S* Mprev, S* Mcur, S* W, S* fval, int* fidx, const PointsDev& P, int* fail,
                  hipStream_t st) {
  auto kern = sym ? (sh ? k_zelim<S, true, true> : k_zelim<S, true, false>)
                  : (sh ? k_zelim<S, false, true> : k_zelim<S, false, false>);
  ensure_dyn_lds((const void*)kern, shmem);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(1024), shmem, st, ids, boxes, outs, csegs, pool, perm, T,
                     Mprev, Mcur, W, fval, fidx, P, fail);
}

void launch_zgather_diag(const double* diag, const int* order, int64_t N, double* dre, double* dim,
                         hipStream_t st) {
  int grid = (int)std::min<int64_t>((N + 255) / 256, 16384);
  hipLaunchKernelGGL(k_zgather_diag, dim3(grid < 1 ? 1 : grid), dim3(256), 0, st, diag, order, N,
                     dre, dim);
}

#define RSKX_ZFACTOR_INST(S)                                                                    \
  template void launch_zassemble<S>(const int*, const LvlBox*, const int*, const int*, const int*, \
                                    int*, S*, const PointsDev&, const LevelParams&, int,          \
                                    hipStream_t);                                               \
  template void launch_zid<S>(const int*, const LvlBox*, const int*, S*, int*, double*,          \
                              const IdParams&, int*, int*, int, size_t, zd*, int64_t,          \
                              hipStream_t);                                                     \
  template void launch_zelim<S>(bool, bool, size_t, const int*, int, const LvlBox*,              \
                                const BoxOut*, const Seg*, int*, const int*, const double*,    \
                                const S*, S*, S*, S*, int*, const PointsDev&, int*, hipStream_t);
RSKX_ZFACTOR_INST(zd)
RSKX_ZFACTOR_INST(zf)

}  // namespace rskx
