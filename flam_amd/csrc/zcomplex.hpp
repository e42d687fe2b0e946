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

// sin z and cos z for 0 <= z < 2^20 (Cody-Waite reduction by pi/2 split in
// three parts, the first two with 32 significant bits so q times them is
// exact, then the Taylor series of sin and cos on |r| <= pi/4 + 1e-6 up to
// r^17 and r^16, whose first omitted terms are below 2e-17): max error
// 1.1e-16 on [0, 1e4] against the libm values
__device__ inline void sincos_cw(double z, double& sz, double& cz) {
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double kP1 = 1.5707963267341256, kP2 = 6.077100506303966e-11,
                   kP3 = 2.0222662487959506e-21;
  const double q = rint(z * kTwoOverPi);
  const double r = ((z - q * kP1) - q * kP2) - q * kP3;
  const double r2 = r * r;
  // Horner in r^2 with (-1)^k / (2k+1)! and (-1)^k / (2k)!
  double sp = 1.0 / 355687428096000.0, cp = 1.0 / 20922789888000.0;   // 1/17!, 1/16!
  sp = sp * r2 - 1.0 / 1307674368000.0;                               // 1/15!
  cp = cp * r2 - 1.0 / 87178291200.0;                                 // 1/14!
  sp = sp * r2 + 1.0 / 6227020800.0;                                  // 1/13!
  cp = cp * r2 + 1.0 / 479001600.0;                                   // 1/12!
  sp = sp * r2 - 1.0 / 39916800.0;                                    // 1/11!
  cp = cp * r2 - 1.0 / 3628800.0;                                     // 1/10!
  sp = sp * r2 + 1.0 / 362880.0;                                      // 1/9!
  cp = cp * r2 + 1.0 / 40320.0;                                       // 1/8!
  sp = sp * r2 - 1.0 / 5040.0;                                        // 1/7!
  cp = cp * r2 - 1.0 / 720.0;                                         // 1/6!
  sp = sp * r2 + 1.0 / 120.0;                                         // 1/5!
  cp = cp * r2 + 1.0 / 24.0;                                          // 1/4!
  sp = sp * r2 - 1.0 / 6.0;                                           // 1/3!
  cp = cp * r2 - 0.5;                                                 // 1/2!
  const double s = r + r * r2 * sp, c = 1.0 + r2 * cp;
  const int qi = (int)q & 3;
  sz = qi == 0 ? s : qi == 1 ? c : qi == 2 ? -s : -c;
  cz = qi == 0 ? c : qi == 1 ? -s : qi == 2 ? -c : s;
}

// 1 / sqrt(a) for a > 0: the FP32 reciprocal square root refined by two
// Newton-Raphson steps in FP64 (relative error ~3e-16)
__device__ inline double rsqrt_nr(double a) {
  double y = (double)rsqrtf((float)a);
  y = y * (1.5 - 0.5 * a * y * y);
  return y * (1.5 - 0.5 * a * y * y);
}

// the asymptotic expansion alone (z >= kHankelZ) with iz = 1 / z and
// sq = sqrt(2 / (pi z)) given, sin and cos from sincos_cw
__device__ inline void bessel01_far_sq(double z, double iz, double sq, double& J0, double& Y0,
                                       double& J1, double& Y1) {
  constexpr HankelCoef h = hankel_coef();
  const double u = iz * iz;
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
  sincos_cw(z, sz, cz);
  const double rs2 = 0.70710678118654752440;     // 1 / sqrt 2
  const double cw0 = (cz + sz) * rs2, sw0 = (sz - cz) * rs2;
  const double cw1 = (sz - cz) * rs2, sw1 = -(sz + cz) * rs2;
  J0 = sq * (p0 * cw0 - q0 * sw0);
  Y0 = sq * (p0 * sw0 + q0 * cw0);
  J1 = sq * (p1 * cw1 - q1 * sw1);
  Y1 = sq * (p1 * sw1 + q1 * cw1);
}

// the asymptotic expansion alone (z >= kHankelZ); iz = 1 / z
__device__ inline void bessel01_far(double z, double iz, double& J0, double& Y0, double& J1,
                                    double& Y1) {
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

// iz = 1 / z (callers that have it spare a division)
__device__ inline void bessel01(double z, double iz, double& J0, double& Y0, double& J1,
                                double& Y1) {
  if (z < kHankelZ) {
    J0 = j0(z); Y0 = y0(z); J1 = j1(z); Y1 = y1(z);
    return;
  }
  bessel01_far(z, iz, J0, Y0, J1, Y1);
}

__device__ inline void bessel01(double z, double& J0, double& Y0, double& J1, double& Y1) {
  bessel01(z, 1.0 / z, J0, Y0, J1, Y1);
}

}  // namespace rskx
