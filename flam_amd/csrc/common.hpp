// Shared host/device definitions of the rskx library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>

namespace rskx {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
#define RSKX_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::rskx::set_error(std::string("FLAM:rskx:hip: ") + #call + ": " +        \
                        hipGetErrorString(e_));                                \
      return RSKX_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

// ------------------------------------------------- point data (tree order) --
// Structure-of-arrays copy of the points and the matrix description,
// permuted into tree order so a leaf's points are contiguous.
struct PointsDev {
  const double* x[3];     // coordinates
  const double* nu[3];    // unit normals
  const double* w;        // quadrature weights
  const double* diag;     // diagonal entries (real parts for a complex kernel)
  const double* diag_im;  // imaginary parts of the diagonal (complex kernels)
  int d;
  int kind;               // RSKX_KERNEL_*
  double scale;           // off-diagonal multiplier
  double k, eta;          // wavenumber and coupling of RSKX_KERNEL_HELMHOLTZ2D_CFIE
  const double* blk;      // RSKX_KERNEL_CALLBACK: every box's A(slf,slf) (n x n,
                          // slf order) at LvlBox::A_off, evaluated by the host
};

// proxy / filter parameters of one level (box sizes and proxy weights are
// per box: the boxes of one level may belong to different trees)
struct LevelParams {
  double radius;          // proxy ring radius = neighbour filter radius (box units)
  const double* ppts;     // RSKX_PROXY_POINTS_SL: 3 x np points in box units (device)
  int p;                  // proxy rows of K (0 at levels <= 2)
  int np;                 // distinct proxy points (p, or p / 2 with incoming rows)
  int pkind;              // RSKX_PROXY_*
  int filter;             // apply the neighbour filter (levels > 2)
  int two_sided;          // symm 'n': stack A(slf,nbr)' under A(nbr,slf)
};

// one box of the level being processed
struct LvlBox {
  int n;                  // |slf|
  int mc;                 // neighbour candidates
  int mcap;               // K row capacity (leading dimension of K)
  int nseg, seg0;         // neighbour candidate segments
  int ncs, cs0;           // child segments forming slf (own points for a leaf)
  int ldt;                // rows (leading dimension) of the box's T / R scratch:
                          // n, or a rank cap when a rank model sized it
  int64_t slf_off;        // pool offset of slf
  int64_t nbr_off;        // offset in the neighbour list arena
  int64_t K_off;          // offset in the K arena
  int64_t perm_off;       // offset in the ID permutation arena (n)
  int64_t T_off;          // offset in the double T/R scratch (3 ldt n)
  int64_t W_off;          // offset in the factor scratch (n^2)
  int64_t A_off;          // RSKX_KERNEL_CALLBACK: offset of A(slf,slf) in PointsDev::blk
  double ctr[3];
  double l[3];            // box size (its tree's size at this level)
  double pw;              // single-layer proxy multiplier of its problem
};

// list segment: `len` point indices at pool[off]; for child segments moff
// is the child's M block offset in the previous level's M arena (-1: none)
struct Seg {
  int64_t off;
  int64_t moff;
  int len;
  int pad;
};

// outputs of one box's elimination, filled by the host after the ID
struct BoxOut {
  int k, r;
  int fac;                // factor index (global), -1 if no compression
  int pad;
  int64_t M_off;          // M block output (k x k, or n x n if r == 0)
  int64_t nl_off;         // pool offset of the new current list (k)
  int64_t T, LU, E, G;    // offsets into the factor value arena
  int64_t Li, Ui;         // inverted 32 x 32 diagonal blocks of L and U
  int64_t sk, rd, piv;    // offsets into the factor index arena
  int64_t Tt, Et, Gt, LUt, Lti, Uti;   // transposed copies (FacRec)
  int64_t D;              // symm 'h': the r pivots of L D L' (-1 otherwise)
};

// per-factor record used by the sweeps (values in the factor arena)
struct FacRec {
  int k, r;
  int packed, pad_;       // transposed copies already written by the elimination
  int64_t T, LU, E, G;    // T k x r, LU r x r packed, E k x r, G r x k
  int64_t Li, Ui;         // ceil(r/32) inverted diagonal blocks of L (and U), 32 x 32 each
  int64_t sk, rd, piv;    // index arena offsets (tree order point indices)
  // transposed copies written by k_pack so that every product of the
  // solve sweeps reads its matrix with consecutive rows on consecutive
  // lanes: T' (r x k), E' (r x k), G' (k x r, 'n'), LU' (r x r), and the
  // diagonal-block inverses transposed block by block
  int64_t Tt, Et, Gt, LUt, Lti, Uti;
  int64_t D;              // symm 'h': D of K(rd,rd) = L D L' (rskelf.m:229-233), -1 otherwise
};

// diagonal block size of the blocked triangular solves in the sweeps
constexpr int kDiagBlock = 32;

struct IdParams {
  double tol;             // relative tolerance
  int kmax;               // rank cap (0: none)
  double Tmax;            // interpolation bound (inf: no refinement)
  double rrqr_iter;       // max refinement iterations
};

// one chunk of a splitting box's point range (device tree build)
struct TreeChunk {
  int64_t start;          // first point of the chunk (current order)
  int len;                // points in the chunk
  int box;                // splitting box index within this level's list
  int divmask;            // dimensions its tree halves at this level
  double ctr[3];          // that box's centre
  int64_t off[8];         // scatter: first destination of each child code
};

// Raise a kernel's dynamic shared-memory limit to `bytes` (above 64 KB),
// calling into HIP only when the kernel has not yet been given that much:
// the per-level launches of a factorization and the sweeps would otherwise
// repeat the same attribute call every launch.
inline void ensure_dyn_lds(const void* kern, size_t bytes) {
  if (bytes <= 65536) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> set[64];   // per device
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = set[dev & 63][kern];
  if (cur >= bytes) return;
  if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess)
    cur = bytes;
  else
    (void)hipGetLastError();
}

}  // namespace rskx
