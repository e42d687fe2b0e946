// Blocked Cholesky elimination of one box on the matrix cores (symm 'p',
// rskelf.m:218-246 with L = chol(K(rd,rd),'lower'), E = K(sk,rd)/L',
// M(sk,sk) - E E').
//
// The box's matrix in the order [rd; sk] is cut into 16 x 16 tiles after
// padding each part to a multiple of 16: rd rows r..R16 are identity rows
// (pivot 1, no coupling, so they factor trivially and leave every other
// entry unchanged), sk rows k..K16 are zero.  The lower triangle of tiles
// is spread round-robin over the workgroup's waves and held in MFMA
// accumulator registers (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64);
// every update is a product of two row-major 16-row blocks staged in LDS:
//
//   Ksr' = Ksr - Kss T                     (sk x rd tiles)
//   Krr' = Krr - T' Ksr' - Krs T           (rd x rd tiles)
//   per 16-column panel p of rd: the panel's column of tiles goes to LDS,
//   one wave factors the diagonal tile (16 steps in registers), inverts
//   it, and forms the tiles below as Z Linv' on the matrix cores; after a
//   barrier every wave subtracts L(I,p) L(J,p)' from its tiles right of
//   the panel.  The sk rows of the panels are E; the sk x sk tiles end as
//   M(sk,sk) - E E'.
//
// Outputs are the register elimination's: L (lower triangle of LU and its
// transposed copy LUt), E and E', T and T', M(sk,sk), the skeleton and
// redundant point lists and the inverted 32 x 32 diagonal blocks of L (and
// of L'), each assembled from two inverted 16 x 16 panel blocks A, B and the
// tile C between them as [inv(A) 0; -inv(B) C inv(A) inv(B)].
#include <algorithm>

#include "common.hpp"
#include "devutil.hpp"
#include "entries.hpp"
#include "mfma.hpp"
#include "rskx.h"

namespace rskx {

// 1 / sqrt(d) for a positive pivot: the hardware reciprocal square root in
// FP32 (about 1 ulp, below the factor's own rounding), the correctly rounded
// quotient in FP64
__device__ inline float mf_rsqrt(float d) { return __builtin_amdgcn_rsqf(d); }
__device__ inline double mf_rsqrt(double d) { return 1.0 / sqrt(d); }

__device__ inline float mf_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double mf_bcast(double v, int l) { return read_lane(v, l); }

// layout of the shared memory (shared with the host's size computation)
struct MfLayout {
  int TR, TK, TT, R16, K16, DP, LDK;
  size_t pts_bytes;                             // point fields, positions, child maps
  size_t reg;                                   // elements of the phase region
  __host__ __device__ MfLayout(int n, int k, int r) {
    TR = (r + 15) >> 4;
    TK = (k + 15) >> 4;
    TT = TR + TK;
    R16 = 16 * TR;
    K16 = 16 * TK;
    DP = R16 + K16;
    LDK = K16 + kMfPad;
    pts_bytes = ((size_t)6 * n * sizeof(double) + (size_t)3 * n * sizeof(int) + 31) & ~(size_t)31;
    // entry staging: two columns of tiles (2 DP x LDP)
    const size_t st = (size_t)2 * DP * kMfLdp;
    // sparsification: T' (R16 rows), Ksr'' then Krs (R16 rows), Kss (K16 rows)
    const size_t r1 = (size_t)(2 * R16 + K16) * LDK;
    // elimination: two panel buffers, two inverted diagonal tiles, the tile
    // below a 32-block's first diagonal tile, one product
    const size_t r2 = (size_t)2 * DP * kMfLdp + (size_t)5 * 16 * kMfLdp + 16;
    reg = st > r1 ? st : r1;
    reg = reg > r2 ? reg : r2;
  }
  __host__ __device__ size_t bytes(size_t es) const { return pts_bytes + reg * es; }
};

template <typename S, int NW, int MS>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(sizeof(S) == 4 ? 4 : 1, 8))) k_elim_mfma(const int* __restrict__ ids,
                                                       const LvlBox* __restrict__ boxes,
                                                       const BoxOut* __restrict__ outs,
                                                       const Seg* __restrict__ csegs, int* pool,
                                                       const int* __restrict__ permar,
                                                       const double* __restrict__ Tar,
                                                       const S* __restrict__ Mprev,
                                                       S* __restrict__ Mcur, S* __restrict__ fval,
                                                       int* __restrict__ fidx, PointsDev P,
                                                       int* __restrict__ fail) {
  typedef typename Mf<S>::V V;
  extern __shared__ double smem_d[];
  const int bidx = ids[blockIdx.x];
  const LvlBox B = boxes[bidx];
  const BoxOut O = outs[bidx];
  const int n = B.n, k = O.k, r = O.r;
  const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lgrp = lane >> 4;
  constexpr int NT = 64 * NW;
  const MfLayout Ly(n, k, r);
  const int TR = Ly.TR, TT = Ly.TT, R16 = Ly.R16, K16 = Ly.K16, DP = Ly.DP, LDK = Ly.LDK;
  constexpr int LDP = kMfLdp;
  __shared__ ChildMap cm;
  __shared__ int clen[8];
  if (tid == 0) {
    const int nc = B.ncs > 8 ? 8 : B.ncs;
    cm.nc = nc;
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      const Seg g = csegs[B.cs0 + c];
      cm.start[c] = acc;
      cm.moff[c] = g.moff;
      clen[c] = g.len;
      acc += g.len;
    }
    cm.start[nc] = acc;
  }
  const int* slf = pool + B.slf_off;
  const int* q = permar + B.perm_off;           // q[0..k) skeletons, q[k..n) redundant
  const double* Tg = Tar + B.T_off;              // k x r, leading dimension B.ldt
  double* cx = smem_d;
  double* cy = cx + n;
  double* cnx = cy + n;
  double* cny = cnx + n;
  double* cw = cny + n;
  double* cdg = cw + n;
  int* qp = reinterpret_cast<int*>(cdg + n);     // slf position of box position p
  int* cch = qp + n;                             // child whose M block holds position p (-1: none)
  int* cps = cch + n;                            // position within that child's block
  S* R1 = reinterpret_cast<S*>(reinterpret_cast<char*>(smem_d) + Ly.pts_bytes);
  S* ST = R1;                                    // entry staging (DP x LDP)
  S* Tt = R1;                                    // T' (R16 x K16, row j = column j of T)
  S* KX = Tt + (size_t)R16 * LDK;                // Ksr'' (row j = column j of Ksr'), then Krs
  S* Kss = KX + (size_t)R16 * LDK;               // K16 x K16
  S* PB = R1;                                    // panel buffers (DP x LDP each)
  S* LI = PB + (size_t)2 * DP * LDP;             // inverted diagonal tiles (three, row-major)
  S* CS = LI + 3 * 16 * LDP;                     // L tile below a 32-block's first diagonal tile
  S* US = CS + 16 * LDP;                         // inv(B) C of that block
  S* CB = US + 16 * LDP;                         // the diagonal wave's column broadcast

  // tiles of this wave: t = wid + NW s in the column-major order of the
  // lower triangle of TT x TT tiles; lane s of tIv / tJv holds slot s
  int tIv = -1, tJv = -1;
  {
    const int ntile = TT * (TT + 1) / 2;
    int t = wid + NW * lane;
    if (lane < MS && t < ntile) {
      int J = 0, cnt = TT;
      while (t >= cnt) { t -= cnt; ++J; --cnt; }
      tIv = J + t;
      tJv = J;
    }
  }
  auto tI = [&](int s) -> int { return __builtin_amdgcn_readlane(tIv, s); };
  auto tJ = [&](int s) -> int { return __builtin_amdgcn_readlane(tJv, s); };
  // the tables are re-read at every loop iteration instead of being kept in
  // scalar registers per slot (which would spill)
  auto opaque_tiles = [&]() { asm volatile("" : "+v"(tIv), "+v"(tJv)); };
  // every fourth slot no LDS access moves across: the unrolled slot loops
  // would otherwise issue all their operand loads at once, and the registers
  // they hold decide how many boxes share a CU
  auto slot_fence = [&](int s) { if (s && !(s & 3)) asm volatile("" ::: "memory"); };

  // ---- phase A: point data and child maps, the tiles' entries, T, Kss
  __syncthreads();                               // cm, clen
  const bool leafbox = cm.nc == 1 && cm.moff[0] < 0;
  for (int p = tid; p < n; p += NT) {
    const int sp = p < r ? q[k + p] : q[p - r];
    qp[p] = sp;
    const int pt = slf[sp];
    point_fields(P, pt, cx[p], cy[p], cnx[p], cny[p], cw[p]);
    cdg[p] = P.diag[pt];
    int c = -1, o = 0;
    if (!leafbox)
      for (int u = 0; u < cm.nc; ++u)
        if (sp >= cm.start[u] && sp < cm.start[u + 1]) {
          if (cm.moff[u] >= 0) { c = u; o = sp - cm.start[u]; }
          break;
        }
    cch[p] = c;
    cps[p] = o;
  }
  lds_barrier();
  auto aval = [&](int p1, int p2) -> double {    // A entry between box positions
    if (P.kind == RSKX_KERNEL_CALLBACK) return P.blk[B.A_off + qp[p1] + (int64_t)qp[p2] * n];
    if (p1 == p2) return cdg[p2];
    if (P.kind == RSKX_KERNEL_LAPLACE3D_SL)
      return sl3_entry(P, cx[p1], cy[p1], cnx[p1], cx[p2], cy[p2], cnx[p2], cw[p2]);
    const double dx = cx[p1] - cx[p2], dy = cy[p1] - cy[p2];
    return P.scale * cw[p2] * (dx * cnx[p2] + dy * cny[p2]) / (kTwoPi * (dx * dx + dy * dy));
  };
  auto mv = [&](int p1, int p2) -> S {           // M(slf,slf) entry (rskelf.m:190-199)
    const int c = cch[p1];
    if (c < 0 || c != cch[p2]) return (S)0;
    return Mprev[cm.moff[c] + cps[p1] + (int64_t)cps[p2] * clen[c]];
  };
  auto kv = [&](int p1, int p2) -> S { return (S)aval(p1, p2) + mv(p1, p2); };
  // padded row -> box position (rd j -> j, sk a -> r + a), -1 for padding
  auto zpos = [&](int i) -> int {
    if (i < R16) return i < r ? i : -1;
    const int a = i - R16;
    return a < k ? r + a : -1;
  };
  V acc[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[s][v] = (S)0;
  // the entries of two columns of tiles at a time, staged in LDS
  for (int J0 = 0; J0 < TT; J0 += 2) {
    opaque_tiles();
    int lane_o = lane;
    asm volatile("" : "+v"(lane_o));
    const int lane = lane_o, lrow = lane & 15;
    const int rows = DP - 16 * J0;               // column J0 + 1 stages its rows from 16 J0 too
    const int lcol = J0 + 1 < TT ? 5 : 4;        // log2 of the staged columns
    for (int e = tid; e < (rows << lcol); e += NT) {
      const int x = e >> lcol, c = e & ((1 << lcol) - 1);
      const int i = 16 * J0 + x, j = 16 * J0 + c;
      const int pi = zpos(i), pj = zpos(j);
      S z;
      if (pi < 0 || pj < 0) z = (i == j && i < R16) ? (S)1 : (S)0;
      else if (i >= R16 && j >= R16) z = mv(pi, pj);          // M(sk,sk)
      else z = kv(pi, pj);                                      // Krr, Ksr (and a few
                                                                // upper entries, unread)
      ST[(c >> 4) * DP * LDP + x * LDP + (c & 15)] = z;
    }
    lds_barrier();
#pragma unroll
    for (int s = 0; s < MS; ++s) {
    slot_fence(s);
      const int J = tJ(s);
      if (J != J0 && J != J0 + 1) continue;
      const int I = tI(s);
      const S* st = ST + (J - J0) * DP * LDP;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[s][v] = st[(16 * (I - J0) + Mf<S>::row(lane, v)) * LDP + lrow];
    }
    lds_barrier();
  }
  S* fT = fval + O.T;
  for (int e = tid; e < R16 * K16; e += NT) {
    const int a = e % K16, j = e / K16;
    S v = 0;
    if (j < r && a < k) {
      v = (S)Tg[a + (int64_t)j * B.ldt];
      fT[a + j * k] = v;
    }
    Tt[j * LDK + a] = v;
  }
  for (int e = tid; e < K16 * K16; e += NT) {
    const int b = e % K16, a = e / K16;
    Kss[a * LDK + b] = (a < k && b < k) ? kv(r + a, r + b) : (S)0;
  }
  lds_barrier();

  // ---- phase B: Ksr' = Ksr - Kss T on the sk x rd tiles; Ksr'' to LDS
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    slot_fence(s);
    const int I = tI(s), J = tJ(s);
    if (I < TR || J >= TR) continue;             // also skips empty slots (I = -1)
    acc[s] = mac_rows<S, true>(acc[s], Kss + 16 * (I - TR) * LDK, LDK, Tt + 16 * J * LDK, LDK, K16, lane);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = 16 * (I - TR) + Mf<S>::row(lane, v), j = 16 * J + lrow;
      KX[j * LDK + a] = acc[s][v];
    }
  }
  {
    S* fTt = fval + O.Tt;                         // T' (r x k) for the sweeps
    for (int e = tid; e < r * k; e += NT) {
      const int j = e % r, a = e / r;
      fTt[e] = Tt[j * LDK + a];
    }
  }
  lds_barrier();

  // ---- phase C: Krr' = Krr - T' Ksr' - Krs T on the rd x rd tiles
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    slot_fence(s);
    const int I = tI(s), J = tJ(s);
    if (I < 0 || I >= TR) continue;
    acc[s] = mac_rows<S, true>(acc[s], Tt + 16 * I * LDK, LDK, KX + 16 * J * LDK, LDK, K16, lane);
  }
  lds_barrier();
  for (int e = tid; e < R16 * K16; e += NT) {    // Krs over Ksr''
    const int a = e % K16, i = e / K16;
    KX[i * LDK + a] = (i < r && a < k) ? kv(i, r + a) : (S)0;
  }
  lds_barrier();
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    slot_fence(s);
    const int I = tI(s), J = tJ(s);
    if (I < 0 || I >= TR) continue;
    acc[s] = mac_rows<S, true>(acc[s], KX + 16 * I * LDK, LDK, Tt + 16 * J * LDK, LDK, K16, lane);
  }
  lds_barrier();

  // ---- phase D: blocked Cholesky over the TR panels of rd.  Iteration p:
  //   1. the column-p tiles take the update of panel p-1 and go to PB[p&1]
  //   2. the diagonal wave factors tile (p,p) and inverts it into LI[p&1];
  //      the other waves meanwhile apply panel p-1 to the rest of their tiles
  //   3. the owners of the tiles below form L(I,p) = Z(I,p) Linv' into PB
  S* fLU = fval + O.LU;
  S* fLUt = fval + O.LUt;
  S* fE = fval + O.E;
  S* fEt = fval + O.Et;
  // inverted 32 x 32 diagonal block q / 2 of L (and of L') once its panels
  // are factored (q odd: its second panel; q = TR - 1 even: a half block):
  // [inv(A) 0; -inv(B) C inv(A)  inv(B)], by one wave
  auto block32 = [&](int q, int lane) {
    const int lrow = lane & 15, lgrp = lane >> 4;
    const bool full = q & 1;
    const int b = q >> 1, bs = r - 32 * b < 32 ? r - 32 * b : 32;
    S* dst = fval + O.Li + b * 1024;
    S* dstt = fval + O.Lti + b * 1024;
    const S* LIa = LI + ((full ? q - 1 : q) % 3) * 16 * LDP;
    const S* LIb = LI + (q % 3) * 16 * LDP;
    if (full) {
      // U = inv(B) C (C's columns read down its rows), then X = -U inv(A)
      V u = {0, 0, 0, 0};
      const V ib = *reinterpret_cast<const V*>(LIb + lrow * LDP + 4 * lgrp);
#pragma unroll
      for (int c = 0; c < 4; ++c) u = Mf<S>::mac(ib[c], CS[(4 * lgrp + c) * LDP + lrow], u);
#pragma unroll
      for (int v = 0; v < 4; ++v) US[Mf<S>::row(lane, v) * LDP + lrow] = u[v];
      wave_lds_sync();
      const V ur = *reinterpret_cast<const V*>(US + lrow * LDP + 4 * lgrp);
      V xq = {0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < 4; ++c) xq = Mf<S>::mac(-ur[c], LIa[(4 * lgrp + c) * LDP + lrow], xq);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = 16 + Mf<S>::row(lane, v), c = lrow;
        const S val = (t < bs && c < bs) ? xq[v] : (S)0;
        dst[t + 32 * c] = val;
        dstt[c + 32 * t] = val;
      }
    }
    for (int e = lane; e < 1024; e += 64) {
      const int t = e & 31, c = e >> 5;
      if (full && t >= 16 && c < 16) continue;
      S val = 0;
      if (t < 16 && c < 16) val = LIa[t * LDP + c];
      else if (t >= 16 && c >= 16 && full) val = LIb[(t - 16) * LDP + (c - 16)];
      if (t >= bs || c >= bs) val = 0;
      dst[t + 32 * c] = val;
      dstt[c + 32 * t] = val;
    }
  };
  for (int p = 0; p <= TR; ++p) {
    opaque_tiles();
    // lane-derived addresses are recomputed per panel rather than hoisted
    // out of the loop, where they would stay live for all of it
    int lane_o = lane;
    asm volatile("" : "+v"(lane_o));
    const int lane = lane_o, lrow = lane & 15, lgrp = lane >> 4;
    S* PBp = PB + (size_t)(p & 1) * DP * LDP;    // row x = padded row 16 p + x
    S* PBq = PB + (size_t)((p + 1) & 1) * DP * LDP;   // panel p-1
    // 1. column p: the previous panel's update, then to PB
    if (p > 0) {
#pragma unroll
      for (int s = 0; s < MS; ++s) {
    slot_fence(s);
        const int I = tI(s), J = tJ(s);
        if (p < TR ? J != p : J < TR) continue;   // after the last panel: every sk tile
        acc[s] = mac_rows<S, true>(acc[s], PBq + 16 * (I - p + 1) * LDP, LDP,
                                   PBq + 16 * (J - p + 1) * LDP, LDP, 16, lane);
      }
    }
    if (p == TR) break;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
    slot_fence(s);
      if (tJ(s) != p) continue;
      const int I = tI(s);
#pragma unroll
      for (int v = 0; v < 4; ++v)
        PBp[(16 * (I - p) + Mf<S>::row(lane, v)) * LDP + lrow] = acc[s][v];
    }
    lds_barrier();
    // the 32-block completed at panel p - 1, by a wave other than the
    // diagonal one (inverses triple-buffered: the diagonal wave writes the
    // third)
    if (p >= 2 && !(p & 1) && wid == (p + NW / 2) % NW) block32(p - 1, lane);
    // 2. diagonal tile (wave p mod NW): lane l holds row l & 15 of the tile
    //    and column l & 15 of its inverse, formed alongside the factorization
    //    from the same broadcasts: x(c) -= L(c,j) x(j) after step j
    if (wid == p % NW) {
      // Lane broadcasts of column j (readlane) feed the next updates; the
      // updates are materialized at the end of every step, so the broadcast
      // values of one step only are live.  Entries right of a lane's
      // diagonal are updated too, without masks: no later step reads them.
      // The pivot's own lane scales d by 1/sqrt(d), its square root.
      int lr = lrow;
      asm volatile("" : "+v"(lr));
      S rw[16], xs[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        rw[c] = PBp[lr * LDP + c];
        xs[c] = (S)(c == lr);
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        S d = mf_bcast(rw[j], j);
        if (!(d > (S)0)) {
          if (lane == 0) atomicOr(fail, 1);
          d = (S)1;
        }
        const S inv = mf_rsqrt(d);
        rw[j] *= inv;
        const S xj = xs[j] * inv;
        xs[j] = xj;
#pragma unroll
        for (int c = j + 1; c < 16; ++c) {
          const S lc = mf_bcast(rw[j], c);
          rw[c] -= rw[j] * lc;
          xs[c] -= lc * xj;
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) asm volatile("" : "+v"(rw[c]), "+v"(xs[c]));
      }
      S* LIp = LI + (p % 3) * 16 * LDP;
      if (lane < 16) {
#pragma unroll
        for (int t = 0; t < 16; ++t) LIp[t * LDP + lrow] = xs[t];
      }
      // L(p,p) to the factor: lane group g its columns 4g..4g+3
      const int i = 16 * p + lrow;
      if (i < r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int cc = 4 * lgrp + c, j = 16 * p + cc;
          if (cc <= lrow) {
            const S v = c == 0 ? rw[4 * lgrp] : c == 1 ? rw[4 * lgrp + 1] : c == 2 ? rw[4 * lgrp + 2] : rw[4 * lgrp + 3];
            fLU[i + j * r] = v;
            fLUt[j + i * r] = v;
          }
        }
      }
    }
    if (p > 0) {
#pragma unroll
      for (int s = 0; s < MS; ++s) {
    slot_fence(s);
        const int I = tI(s), J = tJ(s);
        if (J <= p) continue;
        acc[s] = mac_rows<S, true>(acc[s], PBq + 16 * (I - p + 1) * LDP, LDP,
                                   PBq + 16 * (J - p + 1) * LDP, LDP, 16, lane);
      }
    }
    lds_barrier();
    // 3. tiles below the diagonal: L(I,p) = Z(I,p) Linv'
    {
      const S* LIp = LI + (p % 3) * 16 * LDP;
      const V bv = *reinterpret_cast<const V*>(LIp + lrow * LDP + 4 * lgrp);
#pragma unroll
      for (int s = 0; s < MS; ++s) {
    slot_fence(s);
        const int I = tI(s);
        if (tJ(s) != p || I == p) continue;
        const V z = *reinterpret_cast<const V*>(PBp + (16 * (I - p) + lrow) * LDP + 4 * lgrp);
        V c4 = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 4; ++c) c4 = Mf<S>::mac(z[c], bv[c], c4);
        const int j = 16 * p + lrow;
        const bool rdt = I < TR;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int x = 16 * (I - p) + Mf<S>::row(lane, v), i = 16 * I + Mf<S>::row(lane, v);
          PBp[x * LDP + lrow] = c4[v];
          if (I == p + 1 && !(p & 1)) CS[Mf<S>::row(lane, v) * LDP + lrow] = c4[v];
          if (j < r) {
            if (rdt) {
              if (i < r) {
                fLU[i + j * r] = c4[v];
                fLUt[j + i * r] = c4[v];
              }
            } else if (i - R16 < k) {
              fE[(i - R16) + j * k] = c4[v];
              fEt[j + (i - R16) * r] = c4[v];
            }
          }
        }
      }
    }
    lds_barrier();
  }

  if (wid == (NW / 2) % NW) block32(TR - 1, lane);

  // ---- outputs: M(sk,sk) - E E' (full k x k), point lists
  S* Mo = Mcur + O.M_off;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    slot_fence(s);
    const int I = tI(s), J = tJ(s);
    if (J < TR) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = 16 * (I - TR) + Mf<S>::row(lane, v), b = 16 * (J - TR) + lrow;
      if (a < k && b < k) {
        Mo[a + b * k] = acc[s][v];
        if (I != J) Mo[b + a * k] = acc[s][v];
      }
    }
  }
  int* fsk = fidx + O.sk;
  int* frd = fidx + O.rd;
  for (int a = tid; a < k; a += NT) {
    const int pt = slf[q[a]];
    fsk[a] = pt;
    pool[O.nl_off + a] = pt;
  }
  for (int i = tid; i < r; i += NT) frd[i] = slf[q[k + i]];
}

// shapes: {waves, tile slots per wave}; FP32 boxes take shapes 0-2 (four
// waves per SIMD), FP64 boxes shapes 0 and 3 (their accumulators are twice
// the registers)
constexpr int kMfShapes = 4;
constexpr int kMfNW[kMfShapes] = {8, 16, 16, 8};
constexpr int kMfMS[kMfShapes] = {8, 8, 12, 16};

// shape of the MFMA elimination for a box (-1: none), and its shared bytes
int elim_mfma_shape(int n, int k, int r, size_t es, size_t* shmem) {
  const MfLayout Ly(n, k, r);
  const size_t by = Ly.bytes(es);
  if (by > 160 * 1024) return -1;
  const int ntile = Ly.TT * (Ly.TT + 1) / 2;
  static const int order32[] = {0, 1, 2}, order64[] = {0, 3};
  const int* ord = es == 4 ? order32 : order64;
  const int nord = es == 4 ? 3 : 2;
  for (int u = 0; u < nord; ++u) {
    const int c = ord[u];
    if (ntile <= kMfNW[c] * kMfMS[c]) {
      *shmem = by;
      return c;
    }
  }
  return -1;
}

template <typename S>
void launch_elim_mfma(int shape, size_t shmem, const int* ids, int nb, const LvlBox* boxes,
                      const BoxOut* outs, const Seg* csegs, int* pool, const int* permar,
                      const double* Tar, const S* Mprev, S* Mcur, S* fval, int* fidx,
                      const PointsDev& P, int* fail, hipStream_t st) {
#define RSKX_ELIMM(C)                                                                          \
  {                                                                                            \
    auto kern = k_elim_mfma<S, kMfNW[C], kMfMS[C]>;                                            \
    ensure_dyn_lds((const void*)kern, shmem);                                                   \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * kMfNW[C]), shmem, st, ids, boxes, outs, csegs, \
                       pool, permar, Tar, Mprev, Mcur, fval, fidx, P, fail);                   \
  }
  if constexpr (sizeof(S) == 4) {
    if (shape == 0) RSKX_ELIMM(0) else if (shape == 1) RSKX_ELIMM(1) else RSKX_ELIMM(2)
  } else {
    if (shape == 0) RSKX_ELIMM(0) else RSKX_ELIMM(3)
  }
#undef RSKX_ELIMM
}

template void launch_elim_mfma<float>(int, size_t, const int*, int, const LvlBox*, const BoxOut*,
                                      const Seg*, int*, const int*, const double*, const float*,
                                      float*, float*, int*, const PointsDev&, int*, hipStream_t);
template void launch_elim_mfma<double>(int, size_t, const int*, int, const LvlBox*, const BoxOut*,
                                       const Seg*, int*, const int*, const double*, const double*,
                                       double*, double*, int*, const PointsDev&, int*, hipStream_t);

}  // namespace rskx
