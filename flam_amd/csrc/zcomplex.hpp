// Complex scalars of the complex (Helmholtz) path: storage types zf / zd
// (interleaved re, im, the layout of MATLAB's complex arrays and of C99
// complex), arithmetic always in zd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rskx {

template <typename T>
struct cplx {
  T re, im;
};
using zf = cplx<float>;
using zd = cplx<double>;

template <typename S> struct is_cplx { static constexpr bool v = false; };
template <typename T> struct is_cplx<cplx<T>> { static constexpr bool v = true; };
template <typename S> struct cplx_base { using type = S; };
template <typename T> struct cplx_base<cplx<T>> { using type = T; };

__host__ __device__ inline zd zmk(double r, double i) { return zd{r, i}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return zd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zd operator-(zd a) { return zd{-a.re, -a.im}; }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zd operator*(double s, zd a) { return zd{s * a.re, s * a.im}; }
__host__ __device__ inline zd operator*(zd a, double s) { return zd{s * a.re, s * a.im}; }
__host__ __device__ inline zd& operator+=(zd& a, zd b) { a.re += b.re; a.im += b.im; return a; }
__host__ __device__ inline zd& operator-=(zd& a, zd b) { a.re -= b.re; a.im -= b.im; return a; }
__host__ __device__ inline zd conj(zd a) { return zd{a.re, -a.im}; }
__host__ __device__ inline double abs2(zd a) { return a.re * a.re + a.im * a.im; }
// |re| + |im|: LAPACK's cabs1, the pivot measure of zgetrf / izamax
__host__ __device__ inline double cabs1(zd a) { return fabs(a.re) + fabs(a.im); }
__device__ inline double zabs(zd a) { return hypot(a.re, a.im); }
// a / b by Smith's scaling (no overflow for large |b|)
__device__ inline zd zdiv(zd a, zd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, d = b.re + b.im * t;
    return zd{(a.re + a.im * t) / d, (a.im - a.re * t) / d};
  }
  const double t = b.re / b.im, d = b.re * t + b.im;
  return zd{(a.re * t + a.im) / d, (a.im * t - a.re) / d};
}
__device__ inline zd zrecip(zd b) { return zdiv(zd{1.0, 0.0}, b); }
// conj(a) * b
__host__ __device__ inline zd cmul(zd a, zd b) {
  return zd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

// storage <-> arithmetic
__host__ __device__ inline zd ld(zd a) { return a; }
__host__ __device__ inline zd ld(zf a) { return zd{(double)a.re, (double)a.im}; }
template <typename S> __host__ __device__ inline S st(zd a);
template <> __host__ __device__ inline zd st<zd>(zd a) { return a; }
template <> __host__ __device__ inline zf st<zf>(zd a) { return zf{(float)a.re, (float)a.im}; }

// ------------------------------------------- Bessel functions J0 Y0 J1 Y1 --
// For z >= kHankelZ, the Hankel asymptotic expansion (DLMF 10.17.3-10.17.4,
// Abramowitz & Stegun 9.2.5-9.2.10): with w = z - (2 nu + 1) pi / 4,
//   J_nu(z) = sqrt(2 / (pi z)) (P cos w - Q sin w),
//   Y_nu(z) = sqrt(2 / (pi z)) (P sin w + Q cos w),
//   P ~ sum_k (-1)^k a_2k(nu) / z^2k,  Q ~ sum_k (-1)^k a_2k+1(nu) / z^2k+1,
//   a_k(nu) = (4 nu^2 - 1^2)(4 nu^2 - 3^2) ... (4 nu^2 - (2k-1)^2) / (k! 8^k).
// For real z > 0 and nu = 0, 1 the error of each truncated series is below
// its first omitted term (DLMF 10.17(iii)); at z >= 30, kHankelTerms = 16
// terms leave it below 3e-16 relative.  One sincos serves both orders:
// cos w0 = (cos z + sin z) / sqrt 2, sin w0 = (sin z - cos z) / sqrt 2,
// cos w1 = (sin z - cos z) / sqrt 2, sin w1 = -(sin z + cos z) / sqrt 2.
// Below kHankelZ, HIP's device j0, y0, j1, y1.
constexpr double kHankelZ = 30.0;
constexpr int kHankelTerms = 16;

struct HankelCoef {
  double a0[kHankelTerms], a1[kHankelTerms];
};
constexpr HankelCoef hankel_coef() {
  HankelCoef h{};
  double c0 = 1.0, c1 = 1.0;
  for (int k = 0; k < kHankelTerms; ++k) {
    if (k > 0) {
      const double t = (double)(2 * k - 1) * (double)(2 * k - 1);
      c0 *= (0.0 - t) / (8.0 * k);
      c1 *= (4.0 - t) / (8.0 * k);
    }
    h.a0[k] = c0;
    h.a1[k] = c1;
  }
  return h;
}

// iz = 1 / z (callers that have it spare a division)
__device__ inline void bessel01(double z, double iz, double& J0, double& Y0, double& J1,
                                double& Y1) {
  if (z < kHankelZ) {
    J0 = j0(z); Y0 = y0(z); J1 = j1(z); Y1 = y1(z);
    return;
  }
  constexpr HankelCoef h = hankel_coef();
  const double u = iz * iz;
  // P = sum_k (-1)^k a_2k u^k, Q = iz sum_k (-1)^k a_2k+1 u^k (Horner, highest first)
  double p0 = 0, q0 = 0, p1 = 0, q1 = 0;
#pragma unroll
  for (int k = kHankelTerms / 2 - 1; k >= 0; --k) {
    const double sg = (k & 1) ? -1.0 : 1.0;
    p0 = p0 * u + sg * h.a0[2 * k];
    q0 = q0 * u + sg * h.a0[2 * k + 1];
    p1 = p1 * u + sg * h.a1[2 * k];
    q1 = q1 * u + sg * h.a1[2 * k + 1];
  }
  q0 *= iz;
  q1 *= iz;
  double sz, cz;
  sincos(z, &sz, &cz);
  const double rs2 = 0.70710678118654752440;     // 1 / sqrt 2
  const double cw0 = (cz + sz) * rs2, sw0 = (sz - cz) * rs2;
  const double cw1 = (sz - cz) * rs2, sw1 = -(sz + cz) * rs2;
  const double s = sqrt(0.63661977236758134308 * iz);   // sqrt(2 / (pi z))
  J0 = s * (p0 * cw0 - q0 * sw0);
  Y0 = s * (p0 * sw0 + q0 * cw0);
  J1 = s * (p1 * cw1 - q1 * sw1);
  Y1 = s * (p1 * sw1 + q1 * cw1);
}

__device__ inline void bessel01(double z, double& J0, double& Y0, double& J1, double& Y1) {
  bessel01(z, 1.0 / z, J0, Y0, J1, Y1);
}

}  // namespace rskx
