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

}  // namespace rskx
