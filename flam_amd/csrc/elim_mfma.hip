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
#include "rskx.h"

namespace rskx {

typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef double mf_d4 __attribute__((ext_vector_type(4)));

// 16 x 16 x 4 MFMA of each precision.  A (16 x 4): lane l holds A(l & 15,
// l >> 4); B (4 x 16): lane l holds B(l >> 4, l & 15); D (16 x 16): lane l
// holds column l & 15, its element v in row row(l, v) (the f32 and f64
// instructions order the rows differently).
template <typename S> struct Mf;
template <> struct Mf<float> {
  typedef mf_f4 V;
  __device__ static V mac(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mf<double> {
  typedef mf_d4 V;
  __device__ static V mac(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

constexpr int kMfPad = 4;                     // row padding of the LDS blocks (elements)
constexpr int kMfLdp = 16 + kMfPad;           // leading dimension of a panel row

// acc +/-= A B' over `inner` columns (a multiple of 16) of two 16-row
// row-major blocks (leading dimensions multiples of 4, rows 16-byte
// aligned for f32 and 32-byte aligned for f64).  Per 16 columns lane l
// reads the four consecutive entries 4 (l >> 4) + [0, 4) of row l & 15 of
// each block, and MFMA c multiplies their c-th entries: A and B pair the
// same inner index, and the four products cover all 16.
template <typename S, bool NEG>
__device__ inline typename Mf<S>::V mac_rows(typename Mf<S>::V acc, const S* A, int lda,
                                             const S* B, int ldb, int inner) {
  typedef typename Mf<S>::V V;
  const int lane = threadIdx.x & 63;
  const S* a = A + (lane & 15) * lda + 4 * (lane >> 4);
  const S* b = B + (lane & 15) * ldb + 4 * (lane >> 4);
  for (int c0 = 0; c0 < inner; c0 += 16) {
    const V av = *reinterpret_cast<const V*>(a + c0);
    const V bv = *reinterpret_cast<const V*>(b + c0);
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = Mf<S>::mac(NEG ? -av[c] : av[c], bv[c], acc);
  }
  return acc;
}

__device__ inline float mf_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double mf_bcast(double v, int l) { return read_lane(v, l); }

// LDS written by some lanes of a wave and read by others of the same wave:
// a wave's LDS operations complete in order, so only the compiler must be
// kept from moving accesses across this point
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// layout of the shared memory (shared with the host's size computation)
struct MfLayout {
  int TR, TK, TT, R16, K16, DP, LDK;
  size_t pts_bytes;                             // point fields and positions
  size_t r1, r2;                                // element counts of the two phases' regions
  __host__ __device__ MfLayout(int n, int k, int r) {
    TR = (r + 15) >> 4;
    TK = (k + 15) >> 4;
    TT = TR + TK;
    R16 = 16 * TR;
    K16 = 16 * TK;
    DP = R16 + K16;
    LDK = K16 + kMfPad;
    pts_bytes = ((size_t)6 * n * sizeof(double) + (size_t)n * sizeof(int) + 31) & ~(size_t)31;
    // phases A-C: T' (R16 rows), Ksr'' (R16 rows), Kss then Krs
    r1 = (size_t)(2 * R16 + (R16 > K16 ? R16 : K16)) * LDK;
    // phase D: two panel buffers, two Linv and two Linv' blocks, one W' block
    r2 = (size_t)2 * DP * kMfLdp + (size_t)5 * 16 * kMfLdp;
  }
  __host__ __device__ size_t bytes(size_t es) const { return pts_bytes + (r1 > r2 ? r1 : r2) * es; }
};

template <typename S, int NW, int MS>
__global__ void __launch_bounds__(64 * NW) k_elim_mfma(const int* __restrict__ ids,
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
  constexpr int NT = 64 * NW;
  const MfLayout Ly(n, k, r);
  const int TR = Ly.TR, TT = Ly.TT, R16 = Ly.R16, K16 = Ly.K16, DP = Ly.DP, LDK = Ly.LDK;
  constexpr int LDP = kMfLdp;
  __shared__ ChildMap cm;
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
  S* R1 = reinterpret_cast<S*>(reinterpret_cast<char*>(smem_d) + Ly.pts_bytes);
  S* Tt = R1;                                    // T' (R16 x K16, row j = column j of T)
  S* KsT = Tt + (size_t)R16 * LDK;               // Ksr'' (row j = column j of Ksr')
  S* X = KsT + (size_t)R16 * LDK;                // Kss (K16 rows), then Krs (R16 rows)
  S* PB = R1;                                    // phase D: panel buffers (DP x LDP each)
  S* LI = PB + (size_t)2 * DP * LDP;             // inverted diagonal tiles, row-major
  S* LIT = LI + 2 * 16 * LDP;                    // and their transposes
  S* WT = LIT + 2 * 16 * LDP;

  // tiles of this wave: t = wid + NW s in the column-major order of the
  // lower triangle of TT x TT tiles
  int tI[MS], tJ[MS];
  {
    const int ntile = TT * (TT + 1) / 2;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      int t = wid + NW * s, I = -1, J = -1;
      if (t < ntile) {
        J = 0;
        int cnt = TT;
        while (t >= cnt) { t -= cnt; ++J; --cnt; }
        I = J + t;
      }
      tI[s] = I;
      tJ[s] = J;
    }
  }

  // ---- phase A: point data, T, Kss, the tiles' entries
  for (int p = tid; p < n; p += NT) {
    const int sp = p < r ? q[k + p] : q[p - r];
    qp[p] = sp;
    const int pt = slf[sp];
    point_fields(P, pt, cx[p], cy[p], cnx[p], cny[p], cw[p]);
    cdg[p] = P.diag[pt];
  }
  __syncthreads();
  const bool leafbox = cm.nc == 1 && cm.moff[0] < 0;
  auto aval = [&](int p1, int p2) -> double {    // A entry between box positions
    if (P.kind == RSKX_KERNEL_CALLBACK) return P.blk[B.A_off + qp[p1] + (int64_t)qp[p2] * n];
    if (p1 == p2) return cdg[p2];
    if (P.kind == RSKX_KERNEL_LAPLACE3D_SL)
      return sl3_entry(P, cx[p1], cy[p1], cnx[p1], cx[p2], cy[p2], cnx[p2], cw[p2]);
    const double dx = cx[p1] - cx[p2], dy = cy[p1] - cy[p2];
    return P.scale * cw[p2] * (dx * cnx[p2] + dy * cny[p2]) / (kTwoPi * (dx * dx + dy * dy));
  };
  auto mv = [&](int p1, int p2) -> S { return leafbox ? (S)0 : mval(cm, Mprev, qp[p1], qp[p2]); };
  auto kv = [&](int p1, int p2) -> S { return (S)aval(p1, p2) + mv(p1, p2); };
  // padded row -> box position (rd j -> j, sk a -> r + a), -1 for padding
  auto zpos = [&](int i) -> int {
    if (i < R16) return i < r ? i : -1;
    const int a = i - R16;
    return a < k ? r + a : -1;
  };
  S* fT = fval + O.T;
  for (int e = tid; e < R16 * K16; e += NT) {
    const int a = e % K16, j = e / K16;
    S v = 0;
    if (j < r && a < k) {
      v = (S)Tg[a + (int64_t)j * B.ldt];
      fT[a + (int64_t)j * k] = v;
    }
    Tt[j * LDK + a] = v;
  }
  for (int e = tid; e < K16 * K16; e += NT) {
    const int b = e % K16, a = e / K16;
    X[a * LDK + b] = (a < k && b < k) ? kv(r + a, r + b) : (S)0;
  }
  V acc[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) {
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[s][v] = (S)0;
    if (tI[s] < 0) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * tI[s] + Mf<S>::row(lane, v), j = 16 * tJ[s] + (lane & 15);
      const int pi = zpos(i), pj = zpos(j);
      S z;
      if (pi < 0 || pj < 0) z = (i == j && i < R16) ? (S)1 : (S)0;
      else if (i >= R16 && j >= R16) z = mv(pi, pj);          // M(sk,sk)
      else z = kv(pi, pj);                                      // Krr, Ksr
      acc[s][v] = z;
    }
  }
  __syncthreads();

  // ---- phase B: Ksr' = Ksr - Kss T on the sk x rd tiles; Ksr'' to LDS
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (tI[s] < TR || tJ[s] >= TR) continue;     // also skips empty slots (tI = -1)
    acc[s] = mac_rows<S, true>(acc[s], X + 16 * (tI[s] - TR) * LDK, LDK, Tt + 16 * tJ[s] * LDK,
                               LDK, K16);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = 16 * (tI[s] - TR) + Mf<S>::row(lane, v), j = 16 * tJ[s] + (lane & 15);
      KsT[j * LDK + a] = acc[s][v];
    }
  }
  {
    S* fTt = fval + O.Tt;                         // T' (r x k) for the sweeps
    for (int e = tid; e < r * k; e += NT) {
      const int j = e % r, a = e / r;
      fTt[e] = Tt[j * LDK + a];
    }
  }
  __syncthreads();

  // ---- phase C: Krr' = Krr - T' Ksr' - Krs T on the rd x rd tiles
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (tI[s] < 0 || tI[s] >= TR) continue;
    acc[s] = mac_rows<S, true>(acc[s], Tt + 16 * tI[s] * LDK, LDK, KsT + 16 * tJ[s] * LDK, LDK,
                               K16);
  }
  for (int e = tid; e < R16 * K16; e += NT) {    // Krs over the dead Kss
    const int a = e % K16, i = e / K16;
    X[i * LDK + a] = (i < r && a < k) ? kv(i, r + a) : (S)0;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (tI[s] < 0 || tI[s] >= TR) continue;
    acc[s] = mac_rows<S, true>(acc[s], X + 16 * tI[s] * LDK, LDK, Tt + 16 * tJ[s] * LDK, LDK,
                               K16);
  }
  __syncthreads();

  // ---- phase D: blocked Cholesky over the TR panels of rd
  S* fLU = fval + O.LU;
  S* fLUt = fval + O.LUt;
  S* fE = fval + O.E;
  S* fEt = fval + O.Et;
  for (int p = 0; p < TR; ++p) {
    S* PBp = PB + (size_t)(p & 1) * DP * LDP;    // row x = padded row 16 p + x
    S* LIp = LI + (p & 1) * 16 * LDP;
    S* LITp = LIT + (p & 1) * 16 * LDP;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      if (tJ[s] != p) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        PBp[(16 * (tI[s] - p) + Mf<S>::row(lane, v)) * LDP + (lane & 15)] = acc[s][v];
    }
    __syncthreads();
    if (wid == p % NW) {
      // diagonal tile: lane i < 16 holds its row i
      S rw[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) rw[c] = lane < 16 ? PBp[lane * LDP + c] : (S)0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        S d = mf_bcast(rw[j], j);
        if (!(d > (S)0)) {
          if (lane == 0) atomicOr(fail, 1);
          d = (S)1;
        }
        const S pv = sqrt(d), inv = (S)1 / pv;
        if (lane > j) rw[j] *= inv;
        else if (lane == j) rw[j] = pv;
#pragma unroll
        for (int c = j + 1; c < 16; ++c) {
          const S lc = mf_bcast(rw[j], c);
          if (lane >= c) rw[c] -= rw[j] * lc;
        }
      }
      if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) PBp[lane * LDP + c] = c <= lane ? rw[c] : (S)0;
      }
      wave_lds_sync();
      // inverse of the diagonal tile: lane c < 16 forms its column c
      if (lane < 16) {
        S x[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          S v = t == lane ? (S)1 : (S)0;
#pragma unroll
          for (int u = 0; u < t; ++u) v -= PBp[t * LDP + u] * x[u];
          x[t] = v / PBp[t * LDP + t];
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          LIp[t * LDP + lane] = x[t];
          LITp[lane * LDP + t] = x[t];
        }
      }
      wave_lds_sync();
      // tiles below the diagonal: L(I,p) = Z(I,p) Linv'
      for (int xb = 1; xb < TT - p; ++xb) {
        V c = {0, 0, 0, 0};
        c = mac_rows<S, false>(c, PBp + 16 * xb * LDP, LDP, LIp, LDP, 16);
#pragma unroll
        for (int v = 0; v < 4; ++v) PBp[(16 * xb + Mf<S>::row(lane, v)) * LDP + (lane & 15)] = c[v];
      }
      // inverted 32 x 32 diagonal block b of L (and of L'), complete after
      // its second panel or at the last panel
      if ((p & 1) || p == TR - 1) {
        const int b = p >> 1, bs = r - 32 * b < 32 ? r - 32 * b : 32;
        S* dst = fval + O.Li + (int64_t)b * 1024;
        S* dstt = fval + O.Lti + (int64_t)b * 1024;
        const S* LIa = (p & 1) ? LI + ((p - 1) & 1) * 16 * LDP : LIp;
        if (p & 1) {
          // -inv(B) C inv(A): C = L(16 p.., 16 (p-1)..) in the previous panel
          const S* PBq = PB + (size_t)((p - 1) & 1) * DP * LDP;
          const S* LITa = LIT + ((p - 1) & 1) * 16 * LDP;
          V w = {0, 0, 0, 0};
          w = mac_rows<S, false>(w, PBq + 16 * LDP, LDP, LITa, LDP, 16);    // C inv(A)
#pragma unroll
          for (int v = 0; v < 4; ++v) WT[(lane & 15) * LDP + Mf<S>::row(lane, v)] = w[v];
          wave_lds_sync();
          V xq = {0, 0, 0, 0};
          xq = mac_rows<S, true>(xq, LIp, LDP, WT, LDP, 16);              // -inv(B) (C inv(A))
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int t = 16 + Mf<S>::row(lane, v), c = lane & 15;
            const S val = (t < bs && c < bs) ? xq[v] : (S)0;
            dst[t + 32 * c] = val;
            dstt[c + 32 * t] = val;
          }
        }
        for (int e = lane; e < 1024; e += 64) {
          const int t = e & 31, c = e >> 5;
          if ((p & 1) && t >= 16 && c < 16) continue;
          S val = 0;
          if (t < 16 && c < 16) val = LIa[t * LDP + c];
          else if (t >= 16 && c >= 16 && (p & 1)) val = LIp[(t - 16) * LDP + (c - 16)];
          if (t >= bs || c >= bs) val = 0;
          dst[t + 32 * c] = val;
          dstt[c + 32 * t] = val;
        }
      }
    }
    __syncthreads();
    // the panel's L and E rows to the factor, column-major copies with the
    // row index fastest, the transposed copies with the column index fastest
    const int prow = DP - 16 * p;
    for (int e = tid; e < prow * 16; e += NT) {
      const int xr = e % prow, c = e / prow;
      const int i = 16 * p + xr, j = 16 * p + c;
      if (j >= r) continue;
      const S v = PBp[xr * LDP + c];
      if (i < r) {
        if (i >= j) fLU[i + (int64_t)j * r] = v;
      } else if (i >= R16 && i - R16 < k) {
        fE[(i - R16) + (int64_t)j * k] = v;
      }
    }
    for (int e = tid; e < prow * 16; e += NT) {
      const int c = e & 15, xr = e >> 4;
      const int i = 16 * p + xr, j = 16 * p + c;
      if (j >= r) continue;
      const S v = PBp[xr * LDP + c];
      if (i < r) {
        if (i >= j) fLUt[j + (int64_t)i * r] = v;
      } else if (i >= R16 && i - R16 < k) {
        fEt[j + (int64_t)(i - R16) * r] = v;
      }
    }
    // trailing update of the tiles right of the panel
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      if (tJ[s] <= p) continue;
      acc[s] = mac_rows<S, true>(acc[s], PBp + 16 * (tI[s] - p) * LDP, LDP,
                                 PBp + 16 * (tJ[s] - p) * LDP, LDP, 16);
    }
  }

  // ---- outputs: M(sk,sk) - E E' (full k x k), point lists
  S* Mo = Mcur + O.M_off;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (tJ[s] < TR) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = 16 * (tI[s] - TR) + Mf<S>::row(lane, v), b = 16 * (tJ[s] - TR) + (lane & 15);
      if (a < k && b < k) {
        Mo[a + (int64_t)b * k] = acc[s][v];
        if (tI[s] != tJ[s]) Mo[b + (int64_t)a * k] = acc[s][v];
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

// shapes: {waves, tile slots per wave}
constexpr int kMfShapes = 3;
constexpr int kMfNW[kMfShapes] = {8, 8, 16};
constexpr int kMfMS[kMfShapes] = {8, 16, 12};

// shape of the MFMA elimination for a box (-1: none), and its shared bytes
int elim_mfma_shape(int n, int k, int r, size_t es, size_t* shmem) {
  const MfLayout Ly(n, k, r);
  const size_t by = Ly.bytes(es);
  if (by > 160 * 1024) return -1;
  const int ntile = Ly.TT * (Ly.TT + 1) / 2;
  for (int c = 0; c < kMfShapes; ++c)
    // (the 16-wave shape's 12 FP64 tiles per wave exceed the registers of
    // four waves per SIMD)
    if (ntile <= kMfNW[c] * kMfMS[c] && !(es == 8 && kMfNW[c] == 16)) {
      *shmem = by;
      return c;
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
  if (shape == 0) RSKX_ELIMM(0) else if (shape == 1) RSKX_ELIMM(1) else RSKX_ELIMM(2)
#undef RSKX_ELIMM
}

template void launch_elim_mfma<float>(int, size_t, const int*, int, const LvlBox*, const BoxOut*,
                                      const Seg*, int*, const int*, const double*, const float*,
                                      float*, float*, int*, const PointsDev&, int*, hipStream_t);
template void launch_elim_mfma<double>(int, size_t, const int*, int, const LvlBox*, const BoxOut*,
                                       const Seg*, int*, const int*, const double*, const double*,
                                       double*, double*, int*, const PointsDev&, int*, hipStream_t);

}  // namespace rskx
