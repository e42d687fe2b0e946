// Device helpers shared by the factorization kernels: point fields, the
// 3D single-layer entry, and the lookup of the children's M blocks.
#pragma once
#include "common.hpp"
#include "rskx.h"

namespace rskx {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kFourPi = 12.566370614359172953850573533118;

// ------------------------------------------------- point fields, kernels --
// Every kernel stages five fields per point: x, y, f2, f3, w, where
// (f2, f3) = (nu_x, nu_y) for the 2D double layer and (z, 0) for the 3D
// single layer, which needs no normals.
__device__ inline void point_fields(const PointsDev& P, int pt, double& f0, double& f1, double& f2,
                                    double& f3, double& fw) {
  f0 = P.x[0][pt];
  f1 = P.x[1][pt];
  if (P.kind == RSKX_KERNEL_LAPLACE3D_SL) {
    f2 = P.x[2][pt];
    f3 = 0;
  } else {
    f2 = P.nu[0][pt];
    f3 = P.nu[1][pt];
  }
  fw = P.w[pt];
}

// A_ij = scale w_j / (4 pi |x_i - x_j|) (RSKX_KERNEL_LAPLACE3D_SL, i != j)
// from target (t0, t1, t2) and source (s0, s1, s2) coordinates
__device__ inline double sl3_entry(const PointsDev& P, double t0, double t1, double t2, double s0,
                                   double s1, double s2, double sw) {
  const double dx = t0 - s0, dy = t1 - s1, dz = t2 - s2;
  return P.scale * sw / (kFourPi * sqrt(dx * dx + dy * dy + dz * dz));
}

// proxy point q of box B: the ring of ie_circle.m:35,107 / ie_ellipse.m:127,
// or the caller's points scaled by the box size (ie_sphere.m:175)
__device__ inline void proxy_point(const LevelParams& L, const LvlBox& B, int q, double& px,
                                   double& py, double& pz) {
  if (L.pkind == RSKX_PROXY_POINTS_SL) {
    const int u = q % L.np;
    px = L.ppts[3 * u] * B.l[0] + B.ctr[0];
    py = L.ppts[3 * u + 1] * B.l[1] + B.ctr[1];
    pz = L.ppts[3 * u + 2] * B.l[2] + B.ctr[2];
  } else {
    const double th = (double)(q + 1) * kTwoPi / (double)L.p;
    px = L.radius * cos(th) * B.l[0] + B.ctr[0];
    py = L.radius * sin(th) * B.l[1] + B.ctr[1];
    pz = 0;
  }
}

// RSKX_PROXY_POINTS_SL row q against source point (s0, s1, s2) of weight sw:
// rows q < np are the outgoing interactions w_j G(pxy, x_j), rows q >= np
// (symm 'n') the incoming ones, weighted by the problem's proxy weight
// (ie_sphere.m:178-183)
__device__ inline double proxy_sl3(const LevelParams& L, const LvlBox& B, int q, double px,
                                   double py, double pz, double s0, double s1, double s2,
                                   double sw) {
  const double dx = px - s0, dy = py - s1, dz = pz - s2;
  return (q < L.np ? sw : B.pw) / (kFourPi * sqrt(dx * dx + dy * dy + dz * dz));
}

// child block lookup for the modified diagonal M (rskelf.m:190-199)
struct ChildMap {
  int nc;
  int start[9];
  int64_t moff[8];
};

template <typename S>
__device__ inline S mval(const ChildMap& cm, const S* Mprev, int a, int b) {
  for (int c = 0; c < cm.nc; ++c) {
    int s0 = cm.start[c], s1 = cm.start[c + 1];
    if (a >= s0 && a < s1) {
      if (b < s0 || b >= s1 || cm.moff[c] < 0) return (S)0;
      int len = s1 - s0;
      return Mprev[cm.moff[c] + (a - s0) + (int64_t)(b - s0) * len];
    }
  }
  return (S)0;
}

}  // namespace rskx
