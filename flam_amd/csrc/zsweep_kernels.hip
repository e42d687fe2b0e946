// Complex sweeps and products for the Helmholtz path.
//
//   k_zsweep    one workgroup per factor of a level: the upward or downward
//               sweep of rskelf_sv / rskelf_mv for symm 'n' and 's', trans
//               'n' and 'c' (rskelf/sv/rskelf_sv_nn.m, _nc.m, _sn.m, _sc.m;
//               rskelf/mv/rskelf_mv_nn.m, _nc.m, _sn.m, _sc.m) on up to 8
//               right-hand-side columns held in shared memory
//   k_zto_tree / k_zfrom_tree   X (interleaved complex, caller's order) to
//               and from the tree-order work copy
//   k_zmatvec   FP64 complex product with the combined-field matrix,
//               evaluated on the fly (refinement residuals)
#include "common.hpp"
#include "devutil.hpp"
#include "rskx.h"
#include "zcomplex.hpp"

namespace rskx {

namespace {
constexpr int kZSweepThreads = 256;

// the storage-to-arithmetic load (a function name the kernels' leading
// dimensions do not shadow)
template <typename S> __device__ inline zd ld_(S a) { return ld(a); }

// dst(i, c) += sgn sum_l op(M)(i, l) src(l, c) for i < nd, c < nc, l < ns;
// rows of dst / src are nc consecutive zd (row-major in shared memory).
// op 0: M(i + l ld); 1: M(l + i ld) (transpose); 2: conj(M(l + i ld))
// (conjugate transpose); 3: conj(M(i + l ld)).
template <typename S>
__device__ void zgemv(zd* dst, int nd, const zd* src, int ns, const S* __restrict__ M, int ld,
                      int op, int nc, double sgn) {
  for (int e = threadIdx.x; e < nd * nc; e += blockDim.x) {
    const int i = e / nc, c = e % nc;
    zd acc{0, 0};
    if (op == 0) for (int l = 0; l < ns; ++l) acc += ld_(M[i + l * ld]) * src[l * nc + c];
    else if (op == 1) for (int l = 0; l < ns; ++l) acc += ld_(M[l + i * ld]) * src[l * nc + c];
    else if (op == 2) for (int l = 0; l < ns; ++l) acc += cmul(ld_(M[l + i * ld]), src[l * nc + c]);
    else for (int l = 0; l < ns; ++l) acc += cmul(ld_(M[i + l * ld]), src[l * nc + c]);
    dst[e] += sgn * acc;
  }
  __syncthreads();
}

// x = op(T) \ x for the triangles of the packed r x r LU (leading
// dimension r): which 0 = unit lower L, 1 = upper U, 2 = U' (lower,
// conjugated), 3 = L' (unit upper, conjugated); column-oriented, one
// barrier pair per step.
template <typename S>
__device__ void ztrsv(zd* x, int r, const S* __restrict__ LU, int which, int nc) {
  const bool fwd = which == 0 || which == 2;
  for (int q = 0; q < r; ++q) {
    const int j = fwd ? q : r - 1 - q;
    if (which == 1 || which == 2) {
      const zd d = which == 1 ? ld_(LU[j + j * r]) : conj(ld_(LU[j + j * r]));
      const zd inv = zrecip(d);
      for (int c = threadIdx.x; c < nc; c += blockDim.x) x[j * nc + c] = x[j * nc + c] * inv;
      __syncthreads();
    }
    const int i0 = fwd ? j + 1 : 0, i1 = fwd ? r : j;
    for (int e = threadIdx.x; e < (i1 - i0) * nc; e += blockDim.x) {
      const int i = i0 + e / nc, c = e % nc;
      zd a;
      if (which == 0) a = ld_(LU[i + j * r]);              // L(i, j), i > j
      else if (which == 1) a = ld_(LU[i + j * r]);         // U(i, j), i < j
      else if (which == 2) a = conj(ld_(LU[j + i * r]));   // U'(i, j) = conj U(j, i), i > j
      else a = conj(ld_(LU[j + i * r]));                   // L'(i, j) = conj L(j, i), i < j
      x[i * nc + c] -= a * x[j * nc + c];
    }
    __syncthreads();
  }
}

// y = op(T) x with the same triangles (y and x distinct)
template <typename S>
__device__ void ztrmv(zd* y, const zd* x, int r, const S* __restrict__ LU, int which, int nc) {
  for (int e = threadIdx.x; e < r * nc; e += blockDim.x) {
    const int i = e / nc, c = e % nc;
    zd acc{0, 0};
    if (which == 0) {                                      // unit lower L
      acc = x[i * nc + c];
      for (int j = 0; j < i; ++j) acc += ld_(LU[i + j * r]) * x[j * nc + c];
    } else if (which == 1) {                               // upper U
      for (int j = i; j < r; ++j) acc += ld_(LU[i + j * r]) * x[j * nc + c];
    } else if (which == 2) {                               // U'
      for (int j = 0; j <= i; ++j) acc += cmul(ld_(LU[j + i * r]), x[j * nc + c]);
    } else {                                               // L' (unit)
      acc = x[i * nc + c];
      for (int j = i + 1; j < r; ++j) acc += cmul(ld_(LU[j + i * r]), x[j * nc + c]);
    }
    y[e] = acc;
  }
  __syncthreads();
}
}  // namespace

// PROG 0: solve 'n', 1: solve 'c', 2: apply 'n', 3: apply 'c'; SYM: symm 's'
// (T.' for T', conj(T) for T where the 's' files differ).  One workgroup per
// factor; Y (tree order, leading dimension ldy) columns c0 .. c0 + nc - 1.
// Shared memory: xs (k x nc), xr (r x nc), tmp (r x nc), all zd.
template <typename S, int PROG, bool SYM>
__global__ void __launch_bounds__(kZSweepThreads) k_zsweep(const FacRec* __restrict__ recs, bool up,
                                                           const S* __restrict__ fval,
                                                           const int* __restrict__ fidx, S* Y,
                                                           int64_t ldy, int c0, int nc) {
  extern __shared__ double zsh[];
  const FacRec f = recs[blockIdx.x];
  const int k = f.k, r = f.r;
  zd* xs = reinterpret_cast<zd*>(zsh);
  zd* xr = xs + (size_t)k * nc;
  zd* tmp = xr + (size_t)r * nc;
  const int* sk = fidx + f.sk;
  const int* rd = fidx + f.rd;
  const int* p = fidx + f.piv;
  const S* T = fval + f.T;
  const S* LU = fval + f.LU;
  const S* E = fval + f.E;
  const S* G = fval + f.G;
  for (int e = threadIdx.x; e < k * nc; e += blockDim.x)
    xs[e] = ld(Y[sk[e / nc] + (c0 + e % nc) * ldy]);
  for (int e = threadIdx.x; e < r * nc; e += blockDim.x)
    xr[e] = ld(Y[rd[e / nc] + (c0 + e % nc) * ldy]);
  __syncthreads();
  const int tstar = SYM ? 1 : 2;                 // T.' or T' as an op on T
  if (PROG == 0) {
    if (up) {                                    // sv_nn.m:9-15, sv_sn.m
      zgemv(xr, r, xs, k, T, k, tstar, nc, -1.0);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) tmp[e] = xr[p[e / nc] * nc + e % nc];
      __syncthreads();
      ztrsv(tmp, r, LU, 0, nc);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) xr[e] = tmp[e];
      __syncthreads();
      zgemv(xs, k, xr, r, E, k, 0, nc, -1.0);
    } else {                                     // sv_nn.m:19-27
      zgemv(xr, r, xs, k, G, r, 0, nc, -1.0);
      ztrsv(xr, r, LU, 1, nc);
      zgemv(xs, k, xr, r, T, k, 0, nc, -1.0);
    }
  } else if (PROG == 1) {
    if (up) {                                    // sv_nc.m:9-15, sv_sc.m
      zgemv(xr, r, xs, k, T, k, 2, nc, -1.0);
      ztrsv(xr, r, LU, 2, nc);
      zgemv(xs, k, xr, r, G, r, 2, nc, -1.0);
    } else {                                     // sv_nc.m:19-27, sv_sc.m (conj(T))
      zgemv(xr, r, xs, k, E, k, 2, nc, -1.0);
      ztrsv(xr, r, LU, 3, nc);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) tmp[p[e / nc] * nc + e % nc] = xr[e];
      __syncthreads();
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) xr[e] = tmp[e];
      __syncthreads();
      zgemv(xs, k, xr, r, T, k, SYM ? 3 : 0, nc, -1.0);
    }
  } else if (PROG == 2) {
    if (up) {                                    // mv_nn.m:9-15, mv_sn.m
      zgemv(xs, k, xr, r, T, k, 0, nc, 1.0);
      ztrmv(tmp, xr, r, LU, 1, nc);
      zgemv(tmp, r, xs, k, G, r, 0, nc, 1.0);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) xr[e] = tmp[e];
      __syncthreads();
    } else {                                     // mv_nn.m:19-27, mv_sn.m (T.')
      zgemv(xs, k, xr, r, E, k, 0, nc, 1.0);
      ztrmv(tmp, xr, r, LU, 0, nc);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) xr[p[e / nc] * nc + e % nc] = tmp[e];
      __syncthreads();
      zgemv(xr, r, xs, k, T, k, tstar, nc, 1.0);
    }
  } else {
    if (up) {                                    // mv_nc.m:9-15, mv_sc.m (conj(T))
      zgemv(xs, k, xr, r, T, k, SYM ? 3 : 0, nc, 1.0);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) tmp[e] = xr[p[e / nc] * nc + e % nc];
      __syncthreads();
      ztrmv(xr, tmp, r, LU, 3, nc);
      zgemv(xr, r, xs, k, E, k, 2, nc, 1.0);
    } else {                                     // mv_nc.m:19-27
      zgemv(xs, k, xr, r, G, r, 2, nc, 1.0);
      ztrmv(tmp, xr, r, LU, 2, nc);
      zgemv(tmp, r, xs, k, T, k, 2, nc, 1.0);
      for (int e = threadIdx.x; e < r * nc; e += blockDim.x) xr[e] = tmp[e];
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < k * nc; e += blockDim.x)
    Y[sk[e / nc] + (c0 + e % nc) * ldy] = st<S>(xs[e]);
  for (int e = threadIdx.x; e < r * nc; e += blockDim.x)
    Y[rd[e / nc] + (c0 + e % nc) * ldy] = st<S>(xr[e]);
}

// ------------------------------------------------ X to / from tree order --
// U: the caller's element (zd or zf); Y has ncol >= nrhs columns, the extra
// ones zero
// cj: conjugate on the way (trans 't' runs 'c' on conj(X) and conjugates
// the result, rskelf_sv.m:22 / rskelf_mv.m:22)
template <typename U, typename S>
__global__ void k_zto_tree(const U* __restrict__ X, int64_t ldx, S* __restrict__ Y, int64_t N,
                           int nrhs, int ncol, const int* __restrict__ order, bool cj) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = order[p];
    for (int c = 0; c < ncol; ++c) {
      zd v = c < nrhs ? ld(X[q + c * ldx]) : zd{0, 0};
      Y[p + c * N] = st<S>(cj ? conj(v) : v);
    }
  }
}

template <typename U, typename S>
__global__ void k_zfrom_tree(const S* __restrict__ Y, U* __restrict__ X, int64_t ldx, int64_t N,
                             int nrhs, const int* __restrict__ order, bool cj) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = order[p];
    for (int c = 0; c < nrhs; ++c) {
      const zd v = ld(Y[p + c * N]);
      X[q + c * ldx] = st<U>(cj ? conj(v) : v);
    }
  }
}

// ----------------------------------------------------------- k_zmatvec --
// Y = A X with the combined-field matrix of the points (x, nu, w, complex
// diagonal), FP64, evaluated on the fly: a workgroup of 256 targets loops
// over all sources, 256 at a time staged in shared memory with their
// right-hand-side values; each thread accumulates NC columns.  Two passes:
// the far pass (NEAR false) takes the pairs with (k r)^2 >= kHankelZ^2, on
// the Hankel asymptotic expansion alone (1 / r and 1 / sqrt(r) by
// reciprocal square roots, sin and cos by sincos_cw), and writes Y; the
// near pass adds the other pairs (HIP's Bessel functions, whose registers
// would otherwise hold the far pass to one wave per SIMD) and the diagonal.
// Both passes classify a pair by the same arithmetic, so each pair is
// counted once.
constexpr int kZMvTile = 256;
constexpr int kZMvCols = 8;

namespace {
// near pairs: ik = 1 / k, one division per pair (1 / r) serves c and 1 / z
__device__ inline zd cfie_entry_near(double k, double ik, double eta, double dx, double dy,
                                     double r2, double nx, double ny, double w) {
  const double r = sqrt(r2), ir = 1.0 / r;
  const double z = k * r, c = (dx * nx + dy * ny) * ir;
  double J0, Y0, J1, Y1;
  bessel01(z, ir * ik, J0, Y0, J1, Y1);
  const double q = 0.25 * w;
  return zd{q * (eta * J0 - k * c * Y1), q * (eta * Y0 + k * c * J1)};
}
// far pairs: 1 / r and sqrt(2 / (pi z)) = sqrt(2 / (pi k)) / sqrt(r) by
// reciprocal square roots (sk = sqrt(2 / (pi k)))
__device__ inline zd cfie_entry_far(double k, double ik, double sk, double eta, double dx,
                                    double dy, double r2, double nx, double ny, double w) {
  const double ir = rsqrt_nr(r2), r = r2 * ir;
  const double z = k * r, c = (dx * nx + dy * ny) * ir;
  double J0, Y0, J1, Y1;
  bessel01_far_sq(z, ir * ik, sk * rsqrt_nr(r), J0, Y0, J1, Y1);
  const double q = 0.25 * w;
  return zd{q * (eta * J0 - k * c * Y1), q * (eta * Y0 + k * c * J1)};
}
}  // namespace

template <bool NEAR>
__device__ __forceinline__ void zmatvec_body(const double* __restrict__ x, int64_t N,
                                             const double* __restrict__ nu,
                                             const double* __restrict__ w,
                                             const double* __restrict__ diag, double k, double eta,
                                             const zd* __restrict__ X, int64_t ldx,
                                             zd* __restrict__ Y, int64_t ldy, int c0, int nc) {
  __shared__ double4 sp[kZMvTile];          // x, y, nx, ny
  __shared__ double sw[kZMvTile];
  __shared__ zd sX[kZMvCols][kZMvTile];
  const int64_t i = blockIdx.x * (int64_t)kZMvTile + threadIdx.x;
  const bool act = i < N;
  const double tx = act ? x[2 * i] : 0.0, ty = act ? x[2 * i + 1] : 0.0;
  const double ik = 1.0 / k, sk = sqrt(0.63661977236758134308 * ik);   // sqrt(2 / (pi k))
  const double z2near = kHankelZ * kHankelZ * ik * ik;                  // k r < kHankelZ
  zd acc[kZMvCols];
#pragma unroll
  for (int c = 0; c < kZMvCols; ++c) acc[c] = zd{0, 0};
  for (int64_t j0s = 0; j0s < N; j0s += kZMvTile) {
    const int64_t j = j0s + threadIdx.x;
    __syncthreads();
    if (j < N) {
      sp[threadIdx.x] = double4{x[2 * j], x[2 * j + 1], nu[2 * j], nu[2 * j + 1]};
      sw[threadIdx.x] = w[j];
#pragma unroll
      for (int c = 0; c < kZMvCols; ++c)
        sX[c][threadIdx.x] = c < nc ? X[j + (c0 + c) * ldx] : zd{0, 0};
    }
    __syncthreads();
    const int cnt = N - j0s < kZMvTile ? (int)(N - j0s) : kZMvTile;
    if (!act) continue;
    for (int t = 0; t < cnt; ++t) {
      const int64_t jj = j0s + t;
      const double4 q = sp[t];
      const double dx = tx - q.x, dy = ty - q.y;
      const double r2 = dx * dx + dy * dy;
      const bool near = r2 < z2near;             // the same test in both passes
      zd a;
      if (jj == i) {
        if (!NEAR) continue;
        a = zd{diag[2 * i], diag[2 * i + 1]};
      } else if (near != NEAR) {
        continue;
      } else {
        a = NEAR ? cfie_entry_near(k, ik, eta, dx, dy, r2, q.z, q.w, sw[t])
                 : cfie_entry_far(k, ik, sk, eta, dx, dy, r2, q.z, q.w, sw[t]);
      }
#pragma unroll
      for (int c = 0; c < kZMvCols; ++c) acc[c] += a * sX[c][t];
    }
  }
  if (act)
    for (int c = 0; c < nc; ++c) {
      zd* y = Y + i + (c0 + c) * ldy;
      *y = NEAR ? *y + acc[c] : acc[c];
    }
}

// the far pass at four waves per SIMD: the series and Taylor constants are
// rematerialized instead of held in registers across the source loop
__global__ void __launch_bounds__(kZMvTile) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_zmatvec_far(const double* __restrict__ x, int64_t N, const double* __restrict__ nu,
              const double* __restrict__ w, const double* __restrict__ diag, double k, double eta,
              const zd* __restrict__ X, int64_t ldx, zd* __restrict__ Y, int64_t ldy, int c0,
              int nc) {
  zmatvec_body<false>(x, N, nu, w, diag, k, eta, X, ldx, Y, ldy, c0, nc);
}

__global__ void __launch_bounds__(kZMvTile)
k_zmatvec_near(const double* __restrict__ x, int64_t N, const double* __restrict__ nu,
               const double* __restrict__ w, const double* __restrict__ diag, double k, double eta,
               const zd* __restrict__ X, int64_t ldx, zd* __restrict__ Y, int64_t ldy, int c0,
               int nc) {
  zmatvec_body<true>(x, N, nu, w, diag, k, eta, X, ldx, Y, ldy, c0, nc);
}

// ------------------------------------------------------------ launchers --
size_t zsweep_shmem(int kmax, int rmax, int nc) {
  return (size_t)(kmax + 2 * rmax) * nc * sizeof(zd);
}

template <typename S>
void launch_zsweep(int prog, bool sym, bool up, const FacRec* recs, int nrec, const S* fval,
                   const int* fidx, S* Y, int64_t ldy, int c0, int nc, size_t shmem, hipStream_t st) {
#define RSKX_ZS(P, SY)                                                                          \
  {                                                                                             \
    auto kern = k_zsweep<S, P, SY>;                                                             \
    ensure_dyn_lds((const void*)kern, shmem);                                                   \
    hipLaunchKernelGGL(kern, dim3(nrec), dim3(kZSweepThreads), shmem, st, recs, up, fval, fidx, \
                       Y, ldy, c0, nc);                                                         \
  }
  if (sym) {
    if (prog == 0) RSKX_ZS(0, true) else if (prog == 1) RSKX_ZS(1, true)
    else if (prog == 2) RSKX_ZS(2, true) else RSKX_ZS(3, true)
  } else {
    if (prog == 0) RSKX_ZS(0, false) else if (prog == 1) RSKX_ZS(1, false)
    else if (prog == 2) RSKX_ZS(2, false) else RSKX_ZS(3, false)
  }
#undef RSKX_ZS
}

template <typename U, typename S>
void launch_zperm(bool to, void* X, int64_t ldx, S* Y, int64_t N, int nrhs, int ncol,
                  const int* order, bool cj, hipStream_t st) {
  int grid = (int)((N + 255) / 256);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  if (to)
    hipLaunchKernelGGL((k_zto_tree<U, S>), dim3(grid), dim3(256), 0, st, (const U*)X, ldx, Y, N,
                       nrhs, ncol, order, cj);
  else
    hipLaunchKernelGGL((k_zfrom_tree<U, S>), dim3(grid), dim3(256), 0, st, (const S*)Y, (U*)X,
                       ldx, N, nrhs, order, cj);
}

void launch_zmatvec(const double* x, int64_t N, const double* nu, const double* w,
                    const double* diag, double k, double eta, const double* X, int64_t ldx,
                    double* Y, int64_t ldy, int nrhs, hipStream_t st) {
  const int grid = (int)((N + kZMvTile - 1) / kZMvTile);
  for (int c0 = 0; c0 < nrhs; c0 += kZMvCols) {
    const int nc = nrhs - c0 < kZMvCols ? nrhs - c0 : kZMvCols;
    hipLaunchKernelGGL(k_zmatvec_far, dim3(grid), dim3(kZMvTile), 0, st, x, N, nu, w, diag, k,
                       eta, reinterpret_cast<const zd*>(X), ldx, reinterpret_cast<zd*>(Y), ldy, c0,
                       nc);
    hipLaunchKernelGGL(k_zmatvec_near, dim3(grid), dim3(kZMvTile), 0, st, x, N, nu, w, diag, k,
                       eta, reinterpret_cast<const zd*>(X), ldx, reinterpret_cast<zd*>(Y), ldy, c0,
                       nc);
  }
}

#define RSKX_ZSWEEP_INST(S)                                                                     \
  template void launch_zsweep<S>(int, bool, bool, const FacRec*, int, const S*, const int*, S*,  \
                                 int64_t, int, int, size_t, hipStream_t);                      \
  template void launch_zperm<zd, S>(bool, void*, int64_t, S*, int64_t, int, int, const int*,    \
                                    bool, hipStream_t);                                         \
  template void launch_zperm<zf, S>(bool, void*, int64_t, S*, int64_t, int, int, const int*,    \
                                    bool, hipStream_t);
RSKX_ZSWEEP_INST(zd)
RSKX_ZSWEEP_INST(zf)

}  // namespace rskx
