/*
 * rskx.h -- C ABI of the MI355X recursive skeletonization factorization.
 *
 * Drop-in boundary for FLAM's rskelf path.  The reference is MATLAB; the
 * entry points below are what a MEX gateway (or Python ctypes) binds in
 * place of the MATLAB functions cited next to each one.  Conventions:
 *
 *   - Every array argument is a DEVICE pointer (hipMalloc'd or a torch
 *     tensor's data_ptr) unless stated otherwise; the caller keeps
 *     ownership.  Matrices are column-major, like MATLAB arrays.
 *   - Indices handed back to the host (rskx_export_factor) are 0-based;
 *     the gateway adds 1 for MATLAB.
 *   - Work runs on the library's stream (rskx_set_stream, default: the
 *     null stream).  The factorizations and every call that hands data
 *     back to the host return when their work is done; the sweeps
 *     (rskx_rskelf_sv, _mv, _cholmv, _cholsv), the on-the-fly products
 *     (rskx_kernel_matvec, _batch) and rskx_cgs2 only queue theirs: the
 *     results are ready for later work on that stream (synchronize it
 *     before reading them on the host).  rskx_set_stream first drains the
 *     old stream when the stream changes.
 *   - Every call returns RSKX_OK (0) or a negative code; rskx_last_error()
 *     then gives a message that starts with the reference's error
 *     identifier ("FLAM:chksymm:invalidSymm: ...").
 *   - A factor is device-resident behind an opaque handle.  The library's
 *     scratch (workspaces, staging buffers, the stream) is shared by all
 *     factors, so calls must come from one thread at a time, like the
 *     single-threaded MATLAB functions they replace.
 */
#ifndef RSKX_H
#define RSKX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSKX_OK 0
#define RSKX_ERR_ARG (-1)     /* invalid argument (FLAM:* identifier)       */
#define RSKX_ERR_HIP (-2)     /* HIP runtime failure                        */
#define RSKX_ERR_NUM (-3)     /* numerical failure (e.g. chol not SPD)      */
#define RSKX_ERR_UNSUP (-4)   /* option outside what this library builds   */
#define RSKX_ERR_MEM (-5)     /* allocation failure                         */

/* precision of the factorization (rskx_opts.precision) */
#define RSKX_PREC_FP64 0      /* everything in FP64                         */
#define RSKX_PREC_FP32 1      /* FP32 storage/assembly/LU/Schur/sweeps, FP64 ID
                                 norms and pivots, FP64 root factorization    */

/* element type of sweep right-hand sides */
#define RSKX_F64 0
#define RSKX_F32 1
#define RSKX_Z64 2            /* complex double, interleaved (re, im)        */
#define RSKX_Z32 3            /* complex float, interleaved                  */

/* built-in matrices (the reference passes a MATLAB function handle A(I,J);
 * a handle cannot run on the GPU, so the kernel is described by data) */
#define RSKX_KERNEL_LAPLACE2D_DL 1
/*   A_ij = scale * w_j * ((x_i - x_j) . nu_j) / (2 pi |x_i - x_j|^2), i != j
 *   A_ii = diag_i
 * rskelf/test/ie_circle.m:85-103 (scale 1, diag -1/2(1+1/N)),
 * rskelf/test/ie_ellipse.m:103-123 (scale 1, diag -1/2(1 + h kappa/2pi)),
 * and the benchmark circle (scale -1, diag 1/2 - w/4pi).  d = 2. */
#define RSKX_KERNEL_LAPLACE3D_SL 2
/*   A_ij = scale * w_j / (4 pi |x_i - x_j|), i != j;  A_ii = diag_i
 * the single-layer kernel of rskelf/test/ie_sphere.m:149-160 with
 * one-point quadrature (w = triangle areas) and the self-panel integral on
 * the diagonal.  d = 3; nu is not read (may be NULL). */
#define RSKX_KERNEL_HELMHOLTZ2D_CFIE 3
/*   exterior Dirichlet Helmholtz combined field, complex:
 *   A_ij = w_j (dG/dn_j(x_i, x_j) - i eta G(x_i, x_j)), i != j,
 *   G(x, y) = (i/4) H0^(1)(k |x - y|) (the kernel of
 *   rskelf/test/ie_square3.m:44,107-110), dG/dn_y = (i k / 4) H1^(1)(k r)
 *   ((x - y) . n_y) / r;  A_ii = diag_i, complex: diag holds 2 N doubles
 *   (re, im interleaved).  d = 2; scale is not read.  The factor is complex
 *   (FP32 precision: complex float storage), symm 'n' or 's'. */
#define RSKX_KERNEL_CALLBACK 4
/*   any real matrix, its entries evaluated on the host by `entries` (the
 *   reference's A(I,J) handle, called where rskelf.m:210-219 evaluates
 *   full(A(nbr,slf)), full(A(slf,nbr)) and full(A(slf,slf))).  The tree is
 *   built on the device from x as for the built-in kernels; every box's K
 *   and A(slf,slf) are evaluated on the host, uploaded, and compressed and
 *   eliminated by the same device kernels (the QRCP route).  nu, w, diag
 *   are not read; the proxy must be RSKX_PROXY_CALLBACK; d = 2 or 3; real
 *   data, one problem per factorization.  A slow path for
 *   matrices no built-in kernel describes: the callbacks run one box at a
 *   time on the calling thread. */
/* Fill out (nI x nJ, column-major, leading dimension nI) with A(I,J) at
 * 0-based caller point indices.  Return 0, or nonzero to abort the
 * factorization (it then fails with FLAM:rskx:callback). */
typedef int (*rskx_entries_fn)(void* ctx, const int64_t* I, int64_t nI, const int64_t* J,
                               int64_t nJ, double* out);
typedef struct {
  int kind;             /* RSKX_KERNEL_*                                    */
  const double* nu;     /* d x N unit normals (device)                      */
  const double* w;      /* N quadrature weights (device)                    */
  const double* diag;   /* N diagonal entries (device); 2 N for complex    */
  double scale;         /* off-diagonal multiplier                          */
  double wavenumber;    /* HELMHOLTZ2D_CFIE: k > 0                          */
  double eta;           /* HELMHOLTZ2D_CFIE: coupling of the single layer   */
  rskx_entries_fn entries;  /* CALLBACK: A(I,J) on the host               */
  void* ctx;            /* CALLBACK: passed to entries unchanged            */
} rskx_kernel;

/* proxy surfaces (the reference passes pxyfun(x,slf,nbr,l,ctr)) */
#define RSKX_PROXY_NONE 0     /* naive global compression (rskelf.m:204)     */
#define RSKX_PROXY_RING_SL 1  /* ring, single layer -w/2pi log r (ie_circle.m:106-115) */
#define RSKX_PROXY_RING_DL 2  /* ring, double layer of slf (ie_ellipse.m:126-132) */
#define RSKX_PROXY_POINTS_SL 3 /* given points, single layer (ie_sphere.m:174-187):
                                 pxy = pts * l + ctr; rows w_j G(pxy, x_j) and,
                                 for symm 'n', weight * G(x_j, pxy) (incoming);
                                 d = 3 */
#define RSKX_PROXY_RING_CFIE 4 /* ring for HELMHOLTZ2D_CFIE (the pxyfun pattern
                                 of ie_square3.m:132-143): p_l = max(p,
                                 8 ceil((1.5 k l + 48) / 8)) points at radius
                                 box sizes (l: the box's longest side); rows
                                 A(pxy, x_j) and, for symm 'n', the incoming
                                 conj(weight (dG/dn_q - i eta G)(x_j, pxy)) with
                                 the ring's normal n_q; d = 2 */
#define RSKX_PROXY_CALLBACK 5 /* [Kpxy,nbr] = pxyfun(x,slf,nbr,l,ctr) on the host
                                 (rskelf.m:205, levels > 2), for
                                 RSKX_KERNEL_CALLBACK; p = the first size of
                                 the Kpxy buffer in rows (grown on request) */
/* The proxy callback: given a box's points slf (nslf 0-based caller
 * indices), its neighbour candidates nbr (*nnbr entries: the current
 * skeletons of the neighbouring boxes, rskelf.m:186), the box size l and
 * centre ctr (d entries each), write the proxy rows Kpxy (*rows x nslf,
 * column-major, leading dimension maxrows) and compact nbr in place to the
 * neighbours kept (their count in *nnbr, in their given order).  A callback
 * with more than maxrows rows sets *rows to its count and returns 0
 * without writing Kpxy: it is called once more, for the same box, with
 * maxrows = *rows and the original nbr.  Return 0, or nonzero to abort. */
typedef int (*rskx_proxy_fn)(void* ctx, const int64_t* slf, int64_t nslf, int64_t* nbr,
                             int64_t* nnbr, const double* l, const double* ctr, double* Kpxy,
                             int64_t maxrows, int64_t* rows);
typedef struct {
  int kind;             /* RSKX_PROXY_*                                     */
  int p;                /* number of proxy points (64 in the demos)         */
  double radius;        /* ring radius and neighbour filter radius, in units
                           of the box size (1.5 in ie_circle.m:35,114);
                           POINTS_SL: the filter radius only (1.5 in
                           ie_sphere.m:186)                                  */
  double weight;        /* RING_SL multiplier (2 pi / N in ie_circle.m:111);
                           POINTS_SL: incoming-row multiplier (4 pi / N,
                           the mean triangle area, ie_sphere.m:179)          */
  const double* pts;    /* POINTS_SL: 3 x p points (device, column-major) in
                           box-size units about the box centre (the unit
                           sphere of ie_sphere.m:39-40 and :175); else NULL  */
  rskx_proxy_fn fn;     /* CALLBACK: pxyfun on the host                     */
  void* ctx;            /* CALLBACK: passed to fn unchanged                 */
} rskx_proxy;

/* A learned rank model (rskx_rank_model_*): a small multilayer perceptron
 * that predicts a box's skeleton count from its features. */
typedef struct rskx_rank_model rskx_rank_model;

/* options: rskelf.m:112-128 */
typedef struct {
  int lvlmax;           /* maximum tree depth, <= 0 means Inf               */
  const double* ext;    /* HOST d x 2 root extent, or NULL (from the data)  */
  double Tmax;          /* ID interpolation bound, default 2                */
  double rrqr_iter;     /* max RRQR iterations, default Inf                 */
  char symm;            /* 'n' (LU), 's' (LU, one-sided compression: a    */
                        /* symmetric matrix), 'p' (Cholesky), 'h' (LDL') */
  int verb;             /* 1: print the per-level table to stdout           */
  int precision;        /* RSKX_PREC_*                                      */
  int profile;          /* 1: time each kernel class with HIP events;       */
                        /* 2: time only the elimination class ("factor")  */
  int stop;             /* factor at most this many levels, leaves first
                           (rskelf.m:128,172: numeric opts.stop); < 0: all   */
  int id_gram;          /* ID route at tol >= 1e-7: 0 the Gram-matrix ID for
                           FP32 factors and column-pivoted QR for FP64 (so
                           FP64 rank tables follow the reference's QRCP
                           step by step) except boxes beyond the register
                           Gram classes (n > 284), which take the
                           whole-GPU Gram ID; 1 the Gram ID for both; -1
                           QRCP */
  const rskx_rank_model* rank_model;
                        /* non-NULL: predict every box's rank before the
                           level loop and size the Gram ID's R / T work
                           buffers to a cap of rank_scale (1.25 k^ + 4)
                           rows instead of n; a box whose rank exceeds its
                           cap is redone with full buffers (the exact
                           pivoted ID is the judge)                        */
  double rank_scale;    /* multiplier of the predicted caps (default 1)     */
  int record_boxes;     /* 1: keep every box's features and exact rank for
                           rskx_box_features (training data)               */
} rskx_opts;

typedef struct rskx_factor rskx_factor;

typedef struct {
  int64_t N;            /* matrix order                                     */
  int d;                /* point dimension                                  */
  int nlvl;             /* tree levels                                      */
  int nlvl_done;        /* levels factored                                  */
  int64_t nfactor;      /* number of per-box factors (F.lvp(end))           */
  int64_t bytes;        /* device bytes held by the factor                  */
  int precision;        /* RSKX_PREC_*                                      */
  char symm;
  int64_t id_scratch;   /* largest per-level work buffer of the ID (bytes)  */
  int64_t rank_fallbacks; /* boxes redone with full buffers (rank model)    */
} rskx_info;

/* host copy of one factor, F.factors(i) of rskelf.m:249-257.  Call with
 * NULL arrays to get the sizes k = |sk|, r = |rd|, then again with
 * buffers of k, r, k*r, r*r, r*r, r, k*r, r*k elements.  Matrices are
 * column-major doubles; U and F are written for symm 'n' and 's', p (0-based
 * pivot vector, K(rd(p),rd) = L U) for 'n' and 's'; for 'h', U is D of
 * K(rd,rd) = L D L' (1 x 1 and 2 x 2 Bunch-Kaufman blocks, rd already in
 * pivot order as rskelf.m:230 leaves it). */
typedef struct {
  int64_t k, r;
  int64_t* sk; int64_t* rd;
  double* T; double* L; double* U; int64_t* p; double* E; double* F;
} rskx_factor_view;

/* defaults of rskelf.m:112-120 */
int rskx_opts_default(rskx_opts* opts);

/* F = rskelf(A,x,occ,rank_or_tol,pxyfun,opts)   -- rskelf/rskelf.m:109
 * x: device d x N coordinates.  d must be 2 for RSKX_KERNEL_LAPLACE2D_DL
 * with a RING proxy, 3 for RSKX_KERNEL_LAPLACE3D_SL with POINTS_SL. */
int rskx_rskelf(const double* x, int d, int64_t N, int occ, double rank_or_tol,
                const rskx_kernel* A, const rskx_proxy* pxy, const rskx_opts* opts,
                rskx_factor** out);

/* Many independent problems factored together.  No single reference entry
 * corresponds: it is what calling rskelf (rskelf/rskelf.m:109) once per
 * problem computes, for the block-diagonal matrix diag(A_1, ..., A_P).
 * Problem j owns the points [offsets[j], offsets[j+1]) of x (device,
 * d x N with N = offsets[nprob]) and of the kernel's nu / w / diag arrays;
 * offsets is a HOST array of nprob + 1 ascending entries, offsets[0] = 0,
 * every problem non-empty.  Each problem keeps its own tree (its own root
 * box, core/hypoct.m:65); level l of every tree is factored by the same
 * kernel launches.  pxy is a HOST array of nprob proxy rings that share
 * kind, p and radius (the weights may differ); opts.ext must be NULL.
 * The handle is one factor of order N: rskx_rskelf_sv / _mv apply it to
 * the stacked right-hand sides of all problems, and each problem's block
 * equals the factor rskx_rskelf computes for that problem alone. */
int rskx_rskelf_batch(int nprob, const int64_t* offsets, const double* x, int d, int occ,
                      double rank_or_tol, const rskx_kernel* A, const rskx_proxy* pxy,
                      const rskx_opts* opts, rskx_factor** out);

/* Y = rskelf_sv(F,X,trans)                      -- rskelf/rskelf_sv.m:13
 * In place on device X (N x nrhs, leading dimension ldx, type dtype).
 * trans 'n', 't' or 'c' (core/chktrans.m:8). */
int rskx_rskelf_sv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                   int dtype);

/* Y = rskelf_mv(F,X,trans)                      -- rskelf/rskelf_mv.m:13 */
int rskx_rskelf_mv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                   int dtype);

/* rskelf_sv captured in a HIP graph: the solve of F on the fixed device
 * block X (N x nrhs, leading dimension ldx) is recorded once
 * (rskx_sv_graph, which runs it once on its own rows first and leaves X
 * untouched) and replayed in place on X by rskx_graph_launch, on the
 * library stream, as often as needed (an FGMRES preconditioner: same
 * launches every step).  The graph owns its work rows; F must outlive it.
 * No reference counterpart (MATLAB has no graphs).  Real factors. */
typedef struct rskx_graph rskx_graph;
int rskx_sv_graph(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  rskx_graph** out);
int rskx_graph_launch(rskx_graph* G);
void rskx_graph_free(rskx_graph* G);

/* Launch-shape parameters chosen by measurement (flam_amd/tune.py): the
 * largest box of the 8-wave register elimination ("elim8_nmax") and its
 * shared-memory budget ("elim8_lds"), the largest box of Gram-ID classes 0
 * and 1 ("gram_n0", "gram_n1"), concurrent size classes ("concurrent"),
 * the sweep's row / stage buffer budget ("sweep_stage"), the shared
 * bytes above which the QR ID keeps Householder vectors in global memory
 * ("id_lds"), the register QRCP classes on or off ("id_reg"), the
 * matrix-core elimination of symm 'p' boxes ("elim_mfma", on), the
 * matrix-core accumulation of G in the Gram ID ("gram_mfma", off), the box
 * size above which a box is eliminated by the batched whole-GPU library
 * path ("big_n", 384) and the redundant count from which a factor's solve
 * sweep uses library calls ("lib_solve_r", 4096; 0 turns it off).  Values
 * apply to later calls; results do not depend on them beyond rounding (only
 * which kernel shape runs). */
int rskx_set_tuning(const char* name, double value);
int rskx_get_tuning(const char* name, double* value);

/* rskelf_sv without allocating: no counterpart in the reference (MATLAB
 * allocates freely).  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same result as rskx_rskelf_sv, with the work copy of X
 * and the global row slices of large factors in a static 64 MiB device
 * array of the library instead of its workspaces; no hipMalloc, hipFree or
 * event is created.  The sweep needs N * ceil(nrhs / 8) * 8 elements of
 * the factor's precision (plus slices for factors beyond shared memory);
 * a larger sweep returns RSKX_ERR_MEM (FLAM:rskx:scratch) without touching
 * X.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (another thread, or work still queued on another stream) race.
 * Asynchronous on the library stream, like rskx_rskelf_sv. */
int rskx_rskelf_sv_noalloc(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                           int dtype);

/* LD = rskelf_logdet(F)                         -- rskelf/rskelf_logdet.m:11
 * log det of the factored matrix (of its factored part, for a partial
 * factorization) as *re + i *im with 0 <= *im < 2 pi: 2 sum log diag(L)
 * for 'p', sum log D for 'h', sum log diag(U) + log detperm(p) for 'n';
 * each negative pivot or odd row permutation adds i pi.  Host outputs. */
int rskx_rskelf_logdet(const rskx_factor* F, double* re, double* im);

/* Y = rskelf_cholmv(F,X,trans), Y = rskelf_cholsv(F,X,trans)
 *                                 -- rskelf/rskelf_cholmv.m:15, rskelf_cholsv.m:15
 * The generalized Cholesky factor C of F = C C' (symm 'p' only): C X, C' X
 * ('t' = 'c' for real data), C \ X, C' \ X, in place on device X as
 * rskx_rskelf_sv.  Error FLAM:rskelf_cholmv:invalidSymm (cholsv) for
 * other symmetries. */
int rskx_rskelf_cholmv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                       int dtype);
int rskx_rskelf_cholsv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                       int dtype);

/* The upward (mode 1) or downward (mode 2) sweep alone, or both (mode 3),
 * of rskelf_sv / rskelf_mv: the sweeps' mode argument
 * (rskelf/sv/rskelf_sv_nn.m:3, rskelf/mv/rskelf_mv_nn.m:3), as
 * rskelf_partial_sv.m:35-61 and rskelf_partial_mv.m:35-61 call them around
 * the remaining skeleton system of a partial factorization. */
int rskx_sweep_sv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  int mode);
int rskx_sweep_mv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  int mode);

/* [sk, S] = rskelf_partial_info(F)              -- rskelf/rskelf_partial_info.m:14
 * The points left by a partial factorization (opts.stop) and the sparse
 * modification S of their system (rskelf.m:279-308): the remaining matrix
 * is A(sk, sk) + S.  Call with NULL arrays to get nsk and nnz, then with
 * host arrays of nsk (sk, ascending original indices, 0-based) and nnz
 * (I, J: 0-based positions in sk; V) elements. */
int rskx_partial_info(const rskx_factor* F, int64_t* nsk, int64_t* nnz, int64_t* sk, int64_t* I,
                      int64_t* J, double* V);

void rskx_free(rskx_factor* F);

/* Return the library's idle device memory to the device: the workspaces
 * reused across calls and the cache of device blocks released by freed
 * factors (up to 48 GiB is kept for reuse, so that factorizations do not
 * call hipMalloc / hipFree, which synchronizes the device).  Live factors
 * are untouched.  The library also empties the cache by itself when an
 * allocation fails.  No counterpart in the reference: MATLAB frees F's
 * memory when the variable is cleared. */
int rskx_release_cache(void);
const char* rskx_last_error(void);
int rskx_set_stream(void* hip_stream);

int rskx_factor_info(const rskx_factor* F, rskx_info* info);

/* The verbose table of rskelf.m:264-272: for each factored level, rows of
 * (lvl, nblk, start npts, end npts) written to host rows[4*i..4*i+3].
 * Returns the number of rows. */
int rskx_level_table(const rskx_factor* F, int64_t* rows, int maxrows);

/* F.lvp (rskelf.m:263): the cumulative factor counts per factored level,
 * leaves first, nlvl_done + 1 host entries.  Returns the count. */
int rskx_factor_lvp(const rskx_factor* F, int64_t* lvp, int maxn);

/* Host copy of F.factors(i) (0-based i), indices in the caller's order. */
int rskx_export_factor(const rskx_factor* F, int64_t i, rskx_factor_view* view);

/* F from the reference's layout: the inverse of rskx_export_factor, for a
 * factor computed elsewhere (a MATLAB F through the gateway, which
 * subtracts 1 from its indices; the oracle's factors in the tests).
 * factors[f] holds F.factors(f+1) as rskx_export_factor writes it (host
 * arrays; indices 0-based in the caller's point order; T, L, U, E, F
 * column-major doubles; symm 'h': U = D, diagonal only); lvp (host, nlvl+1
 * entries) the cumulative factor counts of F.lvp, leaves first.  The
 * factor is copied to the device in the given precision; the sweeps and
 * rskx_export_factor work on it as on a factor of rskx_rskelf (the level
 * table lists each level's factor count and the points it eliminates).
 * Real data only. */
int rskx_import(int64_t N, char symm, int precision, int nlvl, const int64_t* lvp,
                const rskx_factor_view* factors, rskx_factor** out);

/* Accumulated HIP-event time (ms) and launch count per kernel class of the
 * last rskx_rskelf / sweep calls when opts.profile was set.  names[i] is a
 * static string.  Returns the number of classes. */
int rskx_kernel_times(const rskx_factor* F, const char** names, double* ms, int64_t* launches,
                      int64_t* bytes, double* flops, int maxn);

/* FP64 product Y = A X with the built-in matrix, evaluated on the fly
 * (refinement residuals and tests; the reference's exact matvecs,
 * e.g. rskelf/test/ie_ellipse.m:56).  All pointers device. */
int rskx_kernel_matvec(const double* x, int d, int64_t N, const rskx_kernel* A,
                       const double* X, int64_t ldx, double* Y, int64_t ldy, int nrhs);

/* The same product for the block-diagonal matrix of rskx_rskelf_batch's
 * problems (offsets: HOST, nprob + 1 entries as there): each problem's
 * rows see only its own points. */
int rskx_kernel_matvec_batch(int nprob, const int64_t* offsets, const double* x, int d,
                             const rskx_kernel* A, const double* X, int64_t ldx, double* Y,
                             int64_t ldy, int nrhs);

/* The learned rank model.  No reference counterpart: FLAM sizes every
 * box's ID by its column count.  A network of RSKX_RANK_FEATURES inputs,
 * two tanh layers of 32 and a linear output, trained by its own forward /
 * backward kernels with Adam (FP64) on (features, log2(1 + rank) / 10)
 * pairs harvested from factorizations (rskx_box_features).  Features of a
 * box, from the tree and its points: reference level / 16, log2(points in
 * its subtree) / 20, log2(longest side) / 8, shortest / longest side, RMS
 * distance of its points from their mean / longest side, distance of that
 * mean from the box centre / longest side, -log10(tol) / 16 and
 * log2(1 + k * longest side) / 10 (k the wavenumber, 0 for Laplace).
 * X: device n x RSKX_RANK_FEATURES row-major; t, y, g: device. */
#define RSKX_RANK_FEATURES 8
int rskx_rank_model_create(uint64_t seed, rskx_rank_model** out);
void rskx_rank_model_free(rskx_rank_model* M);
int rskx_rank_model_nparams(void);
/* host copies of the parameter vector: get into out (if not NULL), then set
 * from in (if not NULL); nparams doubles each */
int rskx_rank_model_params(rskx_rank_model* M, double* out, const double* in);
/* loss = mean (model(X_i) - t_i)^2 and its gradient g (device, nparams) */
int rskx_rank_model_grad(rskx_rank_model* M, const double* X, const double* t, int64_t n,
                         double* g, double* loss);
/* Adam over minibatches of `batch` consecutive samples, `epochs` passes;
 * *loss: the mean squared error over all samples afterwards (host) */
int rskx_rank_model_train(rskx_rank_model* M, const double* X, const double* t, int64_t n,
                          int epochs, int64_t batch, double lr, double* loss);
int rskx_rank_model_predict(const rskx_rank_model* M, const double* X, int64_t n, double* y);
/* features and exact ranks of every box a factorization compressed
 * (opts.record_boxes = 1): call with NULL arrays for *nbox, then with host
 * arrays of nbox * RSKX_RANK_FEATURES and nbox entries */
int rskx_box_features(const rskx_factor* F, int64_t* nbox, double* X, int64_t* rank);

/* Orthogonalization step of the flexible GMRES refinement that uses a
 * factor as preconditioner (the reference preconditions MATLAB's gmres with
 * rskelf_sv, rskelf/test/ie_square3.m:97): classical Gram-Schmidt applied
 * twice to w against nv basis vectors, for nrhs right-hand sides at once.
 * Right-hand side r of the basis starts at V + r * ldr, its vector i at
 * + i * ldv; w's at w + r * ldw; all vectors hold N doubles (device).  On
 * return w is orthogonalized in place and hout (device, nrhs x (nv + 1),
 * row-major) holds each right-hand side's projections (both passes
 * summed) followed by the norm of the new w.  1 <= nv <= 32.  Asynchronous
 * on the library stream. */
int rskx_cgs2(const double* V, int64_t ldv, int64_t ldr, int nv, double* w, int64_t ldw,
              int64_t N, int nrhs, double* hout);

#ifdef __cplusplus
}
#endif
#endif /* RSKX_H */
