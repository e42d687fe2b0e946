// Host side of the rskx library: the C ABI of include/rskx.h, the level
// loop of the factorization (rskelf/rskelf.m:109-309) and the sweep drivers
// (rskelf/rskelf_sv.m:13-36, rskelf/rskelf_mv.m:13-36).
//
// The tree is built on the host; every per-box numerical step runs on the
// device.  The host keeps only index bookkeeping: per box, where its current
// point list (xi) lives in the device pool and how long it is.  After each
// level's ID it reads back the ranks (one int per box) to lay out the
// factor storage of that level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "rskx.h"
#include "tree.hpp"
#include "zcomplex.hpp"

namespace rskx {

// ------------------------------------------------------------ launchers --
void launch_lists(const LvlBox*, const Seg*, const Seg*, int*, int*, int*, const PointsDev&,
                  const LevelParams&, int, hipStream_t);
template <typename S>
void launch_assemble(const int*, const LvlBox*, const int*, const int*, const int*, int*, S*,
                     const PointsDev&, const LevelParams&, int, int, hipStream_t);
template <typename S>
void launch_id(int, const int*, const LvlBox*, const int*, S*, int*, double*, const IdParams&, int*,
               int*, int, size_t, double*, int64_t, hipStream_t);
void launch_gram_id(int, const int*, const LvlBox*, const int*, const int*, const int*, int*,
                    double*, int*, const PointsDev&, const LevelParams&, const IdParams&, int*, int,
                    size_t, hipStream_t);
int gram_id_nmax(int cls);
int gram_id_class(int n);
size_t gram_id_shmem(int cls, int n, int mn, int m, int p);
void set_gram_mfma(bool on);
constexpr int kGramClasses = 3;
template <typename S>
void launch_id_reg(int, const int*, const LvlBox*, const int*, const int*, const int*, int*, S*,
                   int*, const PointsDev&, const LevelParams&, const IdParams&, int*, int, size_t,
                   hipStream_t);
void id_reg_limits(size_t es, int cls, int* mlim, int* nlim);
size_t elim_elems(bool lu, int k, int r);
size_t elim_reg_shmem(int n, int k, int r, size_t es, int nw, bool lu);
int elim_reg_nmax(size_t es, int nw);
template <typename S>
void launch_elim_reg(bool, int, size_t, const int*, int, const LvlBox*, const BoxOut*, const Seg*, int*,
                     const int*, const double*, const S*, S*, S*, int*, const PointsDev&, int*,
                     hipStream_t);
template <typename S>
void launch_elim(bool, bool, int, size_t, const int*, int, const LvlBox*, const BoxOut*, const Seg*,
                 int*, const int*, const double*, const S*, S*, S*, S*, int*, const PointsDev&,
                 int*, hipStream_t);
int elim_mfma_shape(int n, int k, int r, size_t es, size_t* shmem);
int64_t gram_wide_stride(int nmax);
template <typename S>
void gram_wide_gram(const int* ids, int nw, int nmax, const LvlBox* boxes, const int* mrows,
                    const S* Kar, double* Tar, hipStream_t st);
int gram_wide_chol(const int* ids, int nw, int nmax, const LvlBox* boxes, const int* mrows,
                   double* Tar, int* permar, int* kout, const IdParams& prm, double* wsp,
                   hipStream_t st);
int gram_wide_tsolve(const LvlBox* boxes, int b, const LvlBox& hbox, int k, double* Tar,
                     hipStream_t st);
size_t big_elim_scratch(int nbx, int rM, int kM, size_t es);
template <typename S>
int big_elim(bool lu, const int* ids, int nbx, int rM, int kM, const LvlBox* boxes,
             const BoxOut* outs, const Seg* csegs, int* pool, const int* permar, const double* Tar,
             const S* Mprev, S* Mcur, S* fval, int* fidx, const PointsDev& P, int* fail,
             void* scratch, hipStream_t st);
constexpr int kMfmaShapes = 4;
template <typename S>
void launch_elim_mfma(int, size_t, const int*, int, const LvlBox*, const BoxOut*, const Seg*, int*,
                      const int*, const double*, const S*, S*, S*, int*, const PointsDev&, int*,
                      hipStream_t);
template <typename S>
void launch_sweep(int, bool, const FacRec*, int, const S*, const int*, S*, int64_t, int, int, int,
                  int, size_t, S*, int64_t, hipStream_t);
template <typename U, typename S>
void launch_perm(bool, void*, int64_t, S*, int64_t, int, int, const int*, hipStream_t);
template <typename S>
void launch_pack(bool, const FacRec*, int, int, S*, hipStream_t);
template <typename S>
void launch_solve(int, bool, const FacRec*, int, const S*, const int*, S*, int64_t, int, int, int,
                  size_t, S*, int64_t, int, hipStream_t);
size_t big_solve_scratch(int k, int r, int ncol);
template <typename S>
int big_solve(int prog, bool up, const FacRec& f, const S* fval, const int* fidx, S* X,
              int64_t ldx, int ncol, S* scratch, hipStream_t st);
size_t solve_shmem(int prog, int kmax, int rmax, size_t es);
template <typename S>
void launch_logdet(const FacRec*, int, const S*, const int*, int, double*, hipStream_t);
void launch_matvec(int, const double*, int64_t, const double*, const double*, const double*, double,
                   const double*, int64_t, double*, int64_t, int, const int64_t*, int, hipStream_t);
void launch_gather_points(const double*, const double*, const double*, const double*, const int*,
                          int64_t, int, double*, hipStream_t);
void launch_tree_init(const double*, int, int64_t, const int64_t*, int, double*, int*, double*, int,
                      hipStream_t);
void launch_iota(int*, int64_t, hipStream_t);
int cgs_max_vectors();
int cgs_blocks(int64_t N);
void launch_cgs(const double*, int64_t, int64_t, int, double*, int64_t, int64_t, int, double*,
                double*, double*, double*, hipStream_t);
void launch_tree_split(bool, const TreeChunk*, int, int, int64_t, double*, int*, double*, int*,
                       int*, hipStream_t);

// complex (Helmholtz) path: zfactor_kernels.hip, zsweep_kernels.hip
template <typename S>
void launch_zassemble(const int*, const LvlBox*, const int*, const int*, const int*, int*, S*,
                      const PointsDev&, const LevelParams&, int, hipStream_t);
template <typename S>
void launch_zid(const int*, const LvlBox*, const int*, S*, int*, double*, const IdParams&, int*,
                int*, int, size_t, zd*, int64_t, hipStream_t);
template <typename S>
void launch_zelim(bool, bool, size_t, const int*, int, const LvlBox*, const BoxOut*, const Seg*,
                  int*, const int*, const double*, const S*, S*, S*, S*, int*, const PointsDev&,
                  int*, hipStream_t);
size_t zelim_elems(int k, int r);
void launch_zgather_diag(const double*, const int*, int64_t, double*, double*, hipStream_t);
template <typename S>
void launch_zsweep(int, bool, bool, const FacRec*, int, const S*, const int*, S*, int64_t, int, int,
                   size_t, hipStream_t);
template <typename U, typename S>
void launch_zperm(bool, void*, int64_t, S*, int64_t, int, int, const int*, bool, hipStream_t);
size_t zsweep_shmem(int kmax, int rmax, int nc);
void launch_zmatvec(const double*, int64_t, const double*, const double*, const double*, double,
                    double, const double*, int64_t, double*, int64_t, int, hipStream_t);

// rank model (mlp_kernels.hip)
int mlp_nparams();
int mlp_grad_blocks(int64_t n);
void launch_mlp_forward(const double*, const double*, int64_t, double*, hipStream_t);
void launch_mlp_grad(const double*, const double*, const double*, int64_t, double*, double*,
                     double*, double*, hipStream_t);
void launch_mlp_adam(double*, const double*, double*, double*, double, double, double, double, int,
                     hipStream_t);
void launch_box_moments(const int64_t*, const int64_t*, int, const double*, const double*,
                        const double*, double*, hipStream_t);

// proxies on the ring of a combined-field box of longest side l: the ring
// holds the box's outgoing field, whose angular bandwidth is about
// k l / sqrt 2 plus a margin for the tolerance, so the count grows with k l
inline int cfie_proxy_count(int pmin, double k, double l) {
  const int p = 8 * (int)std::ceil((1.5 * k * l + 48.0) / 8.0);
  return std::max(pmin, p);
}

// ------------------------------------------------------------- tuning --
// Launch-shape choices the autotuner (flam_amd/tune.py) sets by measuring
// (rskx_set_tuning): which register shape and Gram-ID class a box of a
// given size goes to, the shared-memory budgets that route boxes between
// kernels, and whether a phase's kernel classes run concurrently.  The
// defaults are the values measured best on the benchmark's workload.
struct Tuning {
  int elim8_nmax = 1 << 30;        // largest box of the 8-wave register elimination (capped
                                   // at what its tile holds)
  int elim8_lds = 80 * 1024;       // its shared-memory budget (two workgroups per CU)
  int gram_n0 = 1 << 30;           // largest box of Gram-ID class 0 / 1 (capped at the
  int gram_n1 = 1 << 30;           // classes' register tiles)
  int concurrent = 1;              // size classes of one phase on side streams
  int sweep_stage = 76 * 1024;     // shared bytes of a sweep workgroup's row/stage buffers
  int id_lds = 160 * 1024;         // shared bytes above which the QR ID keeps its Householder
                                   // vectors in global memory
  int id_reg = 1;                  // register QRCP classes for the boxes that fit them
  int elim_mfma = 1;               // matrix-core elimination for the Cholesky ('p') boxes
  int lib_solve_r = 4096;          // solve sweeps: factors with this many redundant points or
                                   // more take library triangular solves and products
  int big_n = 384;                 // boxes with more points take the whole-GPU path (library
                                   // GEMM / LU / Cholesky, grid-wide pivoted Cholesky of G)
  int gram_mfma = 0;               // matrix-core accumulation of G in the Gram ID (measured
                                   // slower than the vector FP64 one: 5.3 vs 4.9 ms)
};
static Tuning g_tune;

// the Gram-ID class of a box of n columns under the tuned class limits
static int gram_class_tuned(int n) {
  const int c = gram_id_class(n);
  if (c == 0 && n > g_tune.gram_n0) return gram_id_nmax(1) >= n ? 1 : 2;
  if (c == 1 && n > g_tune.gram_n1) return 2;
  return c;
}

// --------------------------------------------------------------- errors --
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
static int fail(int code, const std::string& msg) { set_error(msg); return code; }

static hipStream_t g_stream = nullptr;

// ------------------------------------------------------- device buffers --
// Device blocks released by factors and workspaces are kept for reuse
// (hipMalloc of large blocks and hipFree, which synchronizes the device,
// would otherwise sit inside every factorization).  A request takes the
// smallest cached block of at least its size and at most 4/3 of it; the
// cache is trimmed, largest blocks first, above kCacheBytes.  When
// hipMalloc fails, every idle cached block is freed and the allocation
// tried once more, so the cache never makes a factorization fail that
// would fit without it; rskx_release_cache returns the cache and the
// workspaces to the device on request.
struct BlockCache {
  std::multimap<size_t, void*> free;
  size_t bytes = 0;
  std::mutex mu;
};
static BlockCache g_blocks;
constexpr size_t kCacheBytes = size_t(48) << 30;

// hipFree every idle cached block, largest first
static void block_trim_all() {
  std::lock_guard<std::mutex> lk(g_blocks.mu);
  while (!g_blocks.free.empty()) {
    auto it = std::prev(g_blocks.free.end());
    (void)hipFree(it->second);
    g_blocks.bytes -= it->first;
    g_blocks.free.erase(it);
  }
}

static void* block_get(size_t want, size_t* got) {
  {
    std::lock_guard<std::mutex> lk(g_blocks.mu);
    auto it = g_blocks.free.lower_bound(want);
    if (it != g_blocks.free.end() && it->first <= want + want / 3) {
      void* p = it->second;
      *got = it->first;
      g_blocks.bytes -= it->first;
      g_blocks.free.erase(it);
      return p;
    }
  }
  void* q = nullptr;
  if (hipMalloc(&q, want) != hipSuccess) {
    (void)hipGetLastError();
    // cached blocks may hold the memory this request needs: release them
    // (no kernel uses an idle block; hipFree waits for the device) and retry
    if (g_blocks.bytes == 0) return nullptr;
    (void)hipStreamSynchronize(g_stream);
    block_trim_all();
    if (hipMalloc(&q, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  *got = want;
  return q;
}

static void block_put(void* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_blocks.mu);
  g_blocks.free.emplace(cap, p);
  g_blocks.bytes += cap;
  while (g_blocks.bytes > kCacheBytes && !g_blocks.free.empty()) {
    auto it = std::prev(g_blocks.free.end());
    (void)hipFree(it->second);
    g_blocks.bytes -= it->first;
    g_blocks.free.erase(it);
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes, bool keep = false) {
    if (bytes <= cap) return RSKX_OK;
    size_t nc = std::max(bytes, cap + cap / 2), got = 0;
    void* q = block_get(nc, &got);
    if (!q)
      return fail(RSKX_ERR_MEM, "FLAM:rskx:oom: device allocation of " + std::to_string(nc) + " bytes failed");
    if (keep && p && cap) {
      if (hipMemcpyAsync(q, p, cap, hipMemcpyDeviceToDevice, g_stream) != hipSuccess)
        return fail(RSKX_ERR_HIP, "FLAM:rskx:hip: arena copy failed");
      hipStreamSynchronize(g_stream);
    }
    if (p) {
      hipStreamSynchronize(g_stream);    // no kernel may still use the old block
      block_put(p, cap);
    }
    p = q;
    cap = got;
    return RSKX_OK;
  }
  void release() {
    if (p) {
      hipStreamSynchronize(g_stream);
      block_put(p, cap);
    }
    p = nullptr;
    cap = 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// scratch reused across factorizations of this process
struct Workspace {
  DevBuf boxes, nsegs, csegs, outs, ids, pool, nbr, mnbr, mrows, kout, iter, K, perm, T, W, M[2], flag;
  DevBuf V;                     // Householder vectors of the global QR when rows exceed shared memory
  DevBuf lrecs;                 // factor records of the level being packed
  DevBuf sweep;                 // sweep work rows of factors too large for shared memory
  DevBuf cgs;                   // Gram-Schmidt partial sums
  DevBuf pts, xw, txs, tord, tpart, tchunks, tcounts, xs, poff;
  DevBuf Ablk;                  // host-callback matrices: A(slf,slf) of the level's boxes
  DevBuf big;                   // scratch of the wide-box ID and of the whole-GPU elimination
  DevBuf wids;                  // the wide-box ID's tail lists
  DevBuf wv;                    // the wide-box tails' vectors when shared memory is too small
  DevBuf bsv;                   // the library solve's gathered rows
  int* hflag = nullptr;
};
static Workspace g_ws;

// the wide-box Gram ID takes boxes up to kWideNmax points (G and the three
// T slices are n^2 doubles each); their K is assembled in groups of at most
// kWideKBytes
constexpr int kWideNmax = 32768;
constexpr int64_t kWideKBytes = int64_t(32) << 30;

#define RSKX_TRY(expr) do { int rc_ = (expr); if (rc_ != RSKX_OK) return rc_; } while (0)

// Host-to-device uploads go through a pinned staging ring: a copy from
// pageable memory returns only once the stream has reached it (it waits for
// every kernel queued before it), while a copy from pinned memory is queued
// and returns at once, so the host lays out the next level while the device
// still works.  A staged region is reused only after a stream
// synchronization (sync_stream), which every queued copy precedes.
struct PinnedRing {
  char* p = nullptr;
  size_t cap = 0, top = 0;
};
static PinnedRing g_pin;

// Device-to-host readbacks land in a pinned bounce area too: a copy into
// pageable memory waits for the stream on its own, so three of them cost
// three round trips, while copies into pinned memory are queued and the
// next sync_stream moves every one to its host destination after a single
// wait.  A destination may be read only after that synchronization.
struct PinnedDown {
  char* p = nullptr;
  size_t cap = 0, top = 0;
  struct Pending {
    void* dst;
    size_t off, bytes;
  };
  std::vector<Pending> pending;
};
static PinnedDown g_down;

static int sync_stream() {
  RSKX_HIP(hipStreamSynchronize(g_stream));
  g_pin.top = 0;
  for (const auto& q : g_down.pending) std::memcpy(q.dst, g_down.p + q.off, q.bytes);
  g_down.pending.clear();
  g_down.top = 0;
  return RSKX_OK;
}

static int download_queue(void* dst, const void* src, size_t bytes) {
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (g_down.top + need > g_down.cap) {
    RSKX_TRY(sync_stream());                 // earlier readbacks reach their destinations
    if (need > g_down.cap) {
      const size_t cap = std::max(need, std::max(g_down.cap * 2, (size_t)1 << 20));
      if (g_down.p) RSKX_HIP(hipHostFree(g_down.p));
      g_down.p = nullptr;
      g_down.cap = 0;
      RSKX_HIP(hipHostMalloc((void**)&g_down.p, cap, hipHostMallocDefault));
      g_down.cap = cap;
    }
  }
  RSKX_HIP(hipMemcpyAsync(g_down.p + g_down.top, src, bytes, hipMemcpyDeviceToHost, g_stream));
  g_down.pending.push_back({dst, g_down.top, bytes});
  g_down.top += need;
  return RSKX_OK;
}

static int download(void* dst, const void* src, size_t bytes) {
  if (!bytes) return RSKX_OK;
  const int rc = download_queue(dst, src, bytes);
  if (rc != RSKX_OK) {                       // the caller returns: forget its destinations
    g_down.pending.clear();
    g_down.top = 0;
  }
  return rc;
}

// Side streams for the size classes of one phase.  A level's boxes go to
// a few kernels by size, independent of each other; on one stream each
// class's launch waits for the last workgroups of the previous one, which
// leaves CUs idle.  The classes of the ID and of the elimination run
// concurrently on two extra non-blocking streams (so the legacy null
// stream's implicit ordering does not serialize them; default priority),
// forked from and joined back into g_stream with events.
struct SideStreams {
  int dev = -1;
  hipStream_t s[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
};
static SideStreams g_side;

static int side_init() {
  int dev = 0;
  RSKX_HIP(hipGetDevice(&dev));
  if (g_side.dev == dev) return RSKX_OK;
  if (g_side.dev >= 0) {                        // another device: new streams there
    for (int i = 0; i < 2; ++i) {
      hipStreamDestroy(g_side.s[i]);
      hipEventDestroy(g_side.join[i]);
    }
    hipEventDestroy(g_side.fork);
    g_side.dev = -1;
  }
  for (int i = 0; i < 2; ++i) {
    RSKX_HIP(hipStreamCreateWithFlags(&g_side.s[i], hipStreamNonBlocking));
    RSKX_HIP(hipEventCreateWithFlags(&g_side.join[i], hipEventDisableTiming));
  }
  RSKX_HIP(hipEventCreateWithFlags(&g_side.fork, hipEventDisableTiming));
  g_side.dev = dev;
  return RSKX_OK;
}

// the side streams wait for everything queued on g_stream so far
static int side_fork() {
  RSKX_HIP(hipEventRecord(g_side.fork, g_stream));
  for (int i = 0; i < 2; ++i) RSKX_HIP(hipStreamWaitEvent(g_side.s[i], g_side.fork, 0));
  return RSKX_OK;
}

// g_stream waits for everything queued on the side streams so far
static int side_join() {
  for (int i = 0; i < 2; ++i) {
    RSKX_HIP(hipEventRecord(g_side.join[i], g_side.s[i]));
    RSKX_HIP(hipStreamWaitEvent(g_stream, g_side.join[i], 0));
  }
  return RSKX_OK;
}

template <typename T>
static int upload(DevBuf& b, const std::vector<T>& v) {
  const size_t bytes = v.size() * sizeof(T);
  RSKX_TRY(b.ensure(std::max<size_t>(bytes, 16)));
  if (!bytes) return RSKX_OK;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (g_pin.top + need > g_pin.cap) {
    RSKX_TRY(sync_stream());                 // every staged copy has run
    if (need > g_pin.cap) {
      if (g_pin.p) RSKX_HIP(hipHostFree(g_pin.p));
      g_pin.p = nullptr;
      g_pin.cap = std::max(need, std::max(g_pin.cap * 2, (size_t)16 << 20));
      RSKX_HIP(hipHostMalloc((void**)&g_pin.p, g_pin.cap, hipHostMallocDefault));
    }
  }
  char* src = g_pin.p + g_pin.top;
  std::memcpy(src, v.data(), bytes);
  g_pin.top += need;
  RSKX_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, g_stream));
  return RSKX_OK;
}

// -------------------------------------------------------------- profile --
struct KStat { double ms = 0; int64_t launches = 0; int64_t bytes = 0; double flops = 0; };
// every kernel class (profile 1) or only the elimination class "factor"
// (profile 2, so the other launches run without event markers between them)
struct Prof {
  bool on = false;
  bool elim_only = false;
  std::map<std::string, KStat> st;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> spare;           // events reused after a flush
  Prof() = default;
  Prof(const Prof&) = delete;
  Prof& operator=(const Prof&) = delete;
  ~Prof() {
    for (auto& p : pending) { hipEventDestroy(p.second.first); hipEventDestroy(p.second.second); }
    for (hipEvent_t e : spare) hipEventDestroy(e);
  }
  hipEvent_t event() {
    hipEvent_t e = nullptr;
    if (!spare.empty()) { e = spare.back(); spare.pop_back(); }
    else hipEventCreate(&e);
    return e;
  }
  bool wants(const char* name) const { return on && (!elim_only || !strcmp(name, "factor")); }
  void begin(const char* name, hipEvent_t& a) {
    if (!wants(name)) return;
    a = event();
    hipEventRecord(a, g_stream);
  }
  void end(const char* name, hipEvent_t a, int64_t bytes = 0, double flops = 0) {
    if (!wants(name)) return;
    hipEvent_t b = event();
    hipEventRecord(b, g_stream);
    pending.push_back({name, {a, b}});
    auto& s = st[name];
    s.launches++;
    s.bytes += bytes;
    s.flops += flops;
  }
  void flush() {
    if (!on) return;
    for (auto& p : pending) {
      hipEventSynchronize(p.second.second);
      float ms = 0;
      hipEventElapsedTime(&ms, p.second.first, p.second.second);
      st[p.first].ms += ms;
      spare.push_back(p.second.first);
      spare.push_back(p.second.second);
    }
    pending.clear();
  }
  void host(const char* name, double ms) {
    if (!wants(name)) return;
    auto& s = st[name];
    s.ms += ms;
    s.launches++;
  }
};

}  // namespace rskx

using namespace rskx;

// -------------------------------------------------------- rank model --
struct rskx_rank_model {
  DevBuf params, m, v, g, part, lpart, loss;
  int step = 0;
  ~rskx_rank_model() {
    for (DevBuf* b : {&params, &m, &v, &g, &part, &lpart, &loss}) b->release();
  }
};

// ---------------------------------------------------------- the factor --
struct rskx_factor {
  int64_t N = 0;
  int d = 0;
  int nprob = 1;                          // independent problems (rskx_rskelf_batch)
  char symm = 'n';
  int prec = RSKX_PREC_FP64;
  bool cplx = false;                      // complex factor (RSKX_KERNEL_HELMHOLTZ2D_CFIE)
  int nlvl = 0, nlvl_done = 0;
  std::vector<int64_t> lvp;               // cumulative factor counts per processed level
  std::vector<int> lev_kmax, lev_rmax;
  std::vector<FacRec> recs;               // host copy
  DevBuf d_recs, order;                   // device data
  // factor values and indices, one block per processed level (FacRec
  // offsets are relative to their level's blocks)
  std::vector<DevBuf> lval, lidx;
  mutable std::vector<int64_t> h_order;   // tree order -> original index (filled on export)
  std::vector<std::array<int64_t, 4>> table;
  Prof prof;
  int64_t fval_used = 0, fidx_used = 0;   // totals over the levels
  // partial factorization (rskelf.m:279-308): remaining points (ascending
  // original indices) and the sparse modification of their system
  std::vector<int64_t> Si, SI, SJ;
  std::vector<double> SV;
  // rank model: features of every box (8 each) and the exact ranks of the
  // boxes the factorization compressed (-1: none), when recorded
  std::vector<double> boxfeat;
  std::vector<int64_t> boxrank;
  int64_t id_scratch = 0, rank_fallbacks = 0;
  int level_of(int64_t f) const {         // processed level holding factor f
    return (int)(std::upper_bound(lvp.begin(), lvp.end(), f) - lvp.begin()) - 1;
  }
  ~rskx_factor() {
    d_recs.release();
    order.release();
    for (auto& b : lval) b.release();
    for (auto& b : lidx) b.release();
  }
};

namespace {

// Forest on the device: one tree per problem (problem j owns the points
// [poff[j], poff[j+1]) of x), each with the boxes, centres and point order
// that core/hypoct.m:51-171 builds for that problem alone; the point
// partition of every splitting box of a level, whatever its tree, is one
// k_tree_split pass.  The tree order (tree position -> original index) is
// left in `order` on the device.  A single problem is nprob = 1.
int build_tree_dev(const double* x, int d, int64_t N, int nprob, const int64_t* poff, int occ,
                   int lvlmax, const double* ext, HostTree& t, DevBuf& order) {
  Workspace& ws = g_ws;
  constexpr int kChunk = 2048;           // points per split workgroup
  constexpr int kMaxLevels = 64;
  RSKX_TRY(ws.xs.ensure((size_t)d * N * sizeof(double)));
  RSKX_TRY(ws.txs.ensure((size_t)d * N * sizeof(double)));
  RSKX_TRY(order.ensure((size_t)N * sizeof(int)));
  RSKX_TRY(ws.tord.ensure((size_t)N * sizeof(int)));
  int64_t nmax = 0;
  for (int j = 0; j < nprob; ++j) nmax = std::max<int64_t>(nmax, poff[j + 1] - poff[j]);
  int nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (nmax + 255) / 256));
  nblk = std::max(1, std::min(nblk, 65536 / nprob));
  RSKX_TRY(ws.tpart.ensure((size_t)nprob * nblk * 6 * sizeof(double)));
  const std::vector<int64_t> hoff(poff, poff + nprob + 1);   // alive until the sync below
  RSKX_TRY(upload(ws.poff, hoff));
  double* xs = ws.xs.as<double>();
  int* ord = order.as<int>();
  launch_tree_init(x, d, N, ws.poff.as<int64_t>(), nprob, xs, ord, ws.tpart.as<double>(), nblk,
                   g_stream);
  std::vector<double> part((size_t)nprob * nblk * 6);
  RSKX_TRY(download(part.data(), ws.tpart.p, part.size() * 8));
  RSKX_TRY(sync_stream());
  t = HostTree();
  t.d = d;
  t.nprob = nprob;
  std::vector<double> l((size_t)nprob * d);     // each tree's current box size
  t.cfirst.assign(nprob, -1);
  t.ccnt.assign(nprob, 0);
  for (int j = 0; j < nprob; ++j) {
    double lo[3], hi[3];
    for (int i = 0; i < d; ++i) {
      if (ext) { lo[i] = ext[i]; hi[i] = ext[d + i]; continue; }
      lo[i] = INFINITY; hi[i] = -INFINITY;
      for (int b = 0; b < nblk; ++b) {
        const double* q = &part[((size_t)j * nblk + b) * 6];
        lo[i] = std::min(lo[i], q[i]);
        hi[i] = std::max(hi[i], q[3 + i]);
      }
    }
    for (int i = 0; i < d; ++i) {
      l[(size_t)j * d + i] = hi[i] - lo[i];
      t.ctr.push_back(0.5 * (lo[i] + hi[i]));
      t.bsize.push_back(hi[i] - lo[i]);
    }
    t.beg.push_back(poff[j]);
    t.cnt.push_back(poff[j + 1] - poff[j]);
    t.prnt.push_back(-1);
    t.prob.push_back(j);
  }
  t.lvp = {0, nprob};
  int nlvl = 1;
  const int ncode = 1 << d;
  std::vector<TreeChunk> chunks;
  std::vector<int> counts, divmask(nprob);
  while (lvlmax <= 0 || nlvl < lvlmax) {
    if (nlvl >= kMaxLevels)
      return fail(RSKX_ERR_UNSUP, "FLAM:rskx:tree: more than 64 levels (occ too small for the data)");
    for (int j = 0; j < nprob; ++j) {       // hypoct.m:81-82, per tree
      double* lj = &l[(size_t)j * d];
      const double lmax = *std::max_element(lj, lj + d);
      divmask[j] = 0;
      for (int i = 0; i < d; ++i)
        if (lj[i] >= lmax / std::sqrt(2.0)) { divmask[j] |= 1 << i; lj[i] *= 0.5; }
    }
    chunks.clear();
    std::vector<int> split;   // boxes of this level being split
    for (int64_t p = t.lvp[nlvl - 1]; p < t.lvp[nlvl]; ++p) {
      if (t.cnt[p] <= occ) continue;
      const int sb = (int)split.size();
      split.push_back((int)p);
      for (int64_t s0 = 0; s0 < t.cnt[p]; s0 += kChunk) {
        TreeChunk c{};
        c.start = t.beg[p] + s0;
        c.len = (int)std::min<int64_t>(kChunk, t.cnt[p] - s0);
        c.box = sb;
        c.divmask = divmask[t.prob[p]];
        for (int i = 0; i < d; ++i) c.ctr[i] = t.ctr[(size_t)d * p + i];
        chunks.push_back(c);
      }
    }
    if (split.empty()) break;
    const int nch = (int)chunks.size();
    RSKX_TRY(upload(ws.tchunks, chunks));
    RSKX_TRY(ws.tcounts.ensure((size_t)nch * 8 * sizeof(int)));
    launch_tree_split(false, ws.tchunks.as<TreeChunk>(), nch, d, N, xs, ord,
                      ws.txs.as<double>(), ws.tord.as<int>(), ws.tcounts.as<int>(), g_stream);
    counts.resize((size_t)nch * 8);
    RSKX_TRY(download(counts.data(), ws.tcounts.p, counts.size() * sizeof(int)));
    RSKX_TRY(sync_stream());
    // child boxes and each chunk's destinations (stable within a child)
    int c0 = 0;
    for (int sb = 0; sb < (int)split.size(); ++sb) {
      const int p = split[sb];
      const int j = t.prob[p];
      const int dm = divmask[j];
      const double* lj = &l[(size_t)j * d];
      int c1 = c0;
      while (c1 < nch && chunks[c1].box == sb) ++c1;
      int64_t tot[8] = {0};
      for (int c = c0; c < c1; ++c)
        for (int u = 0; u < ncode; ++u) tot[u] += counts[(size_t)c * 8 + u];
      int64_t start[8];
      int64_t acc = t.beg[p];
      for (int u = 0; u < ncode; ++u) { start[u] = acc; acc += tot[u]; }
      for (int c = c0; c < c1; ++c)
        for (int u = 0; u < ncode; ++u) {
          chunks[c].off[u] = start[u];
          start[u] += counts[(size_t)c * 8 + u];
        }
      const double pc[3] = {t.ctr[(size_t)d * p], d > 1 ? t.ctr[(size_t)d * p + 1] : 0,
                            d > 2 ? t.ctr[(size_t)d * p + 2] : 0};
      int64_t cb = t.beg[p];
      for (int u = 0; u < ncode; ++u) {
        if (tot[u] == 0) continue;
        for (int i = 0; i < d; ++i) {
          double bit = (double)((u >> i) & 1);
          t.ctr.push_back(pc[i] + (((dm >> i) & 1) ? lj[i] : 0.0) * (bit - 0.5));
          t.bsize.push_back(lj[i]);
        }
        t.beg.push_back(cb);
        t.cnt.push_back(tot[u]);
        cb += tot[u];
        t.prnt.push_back(p);
        t.prob.push_back(j);
        t.cfirst.push_back(-1);
        t.ccnt.push_back(0);
        if (t.ccnt[p]++ == 0) t.cfirst[p] = t.nbox() - 1;
      }
      t.cnt[p] = 0;
      c0 = c1;
    }
    RSKX_TRY(upload(ws.tchunks, chunks));
    launch_tree_split(true, ws.tchunks.as<TreeChunk>(), nch, d, N, xs, ord,
                      ws.txs.as<double>(), ws.tord.as<int>(), ws.tcounts.as<int>(), g_stream);
    // the parent level's split status is now final: its neighbour lists
    // are built while the device scatters this level's points
    append_level_neighbours(t, nlvl - 1);
    ++nlvl;
    t.lvp.push_back(t.nbox());
  }
  RSKX_HIP(hipGetLastError());
  t.nlvl = nlvl;
  finish_tree(t);
  return RSKX_OK;
}

// per-box host state during the factorization
struct BoxState {
  int64_t cur_off = 0;   // pool offset of the current point list
  int cur_len = 0;
  int64_t moff = -1;     // M block offset in the arena of its level (-1: zero)
};

constexpr int kBlock = 256;

// Remaining points and skeleton system of a partial factorization
// (rskelf.m:279-308): the points still listed by the boxes of the last
// factored level and by the leaves above it, and the M blocks of the last
// level's boxes placed at those points' positions.  Host side, once.
template <typename S>
int partial_assembly(const HostTree& t, const std::vector<BoxState>& bs, int lv_last,
                     const DevBuf& Marena, rskx_factor* F) {
  Workspace& ws = g_ws;
  const int64_t N = F->N;
  std::vector<int> ord(N), pool_h;
  RSKX_HIP(hipMemcpy(ord.data(), F->order.p, N * sizeof(int), hipMemcpyDeviceToHost));
  int64_t pool_n = 0;
  std::vector<int> boxes;                  // boxes holding remaining points
  for (int b = 0; b < t.nbox(); ++b) {
    const int lvb = t.level_of(b);
    const bool frontier = lvb == lv_last || (lvb < lv_last && t.leaf(b));
    if (frontier && bs[b].cur_len > 0) {
      boxes.push_back(b);
      pool_n = std::max<int64_t>(pool_n, bs[b].cur_off + bs[b].cur_len);
    }
  }
  pool_h.resize(pool_n);
  if (pool_n)
    RSKX_HIP(hipMemcpy(pool_h.data(), ws.pool.p, pool_n * sizeof(int), hipMemcpyDeviceToHost));
  F->Si.clear();
  for (int b : boxes)
    for (int q = 0; q < bs[b].cur_len; ++q) F->Si.push_back(ord[pool_h[bs[b].cur_off + q]]);
  std::sort(F->Si.begin(), F->Si.end());
  std::vector<int64_t> pos(N, -1);
  for (size_t q = 0; q < F->Si.size(); ++q) pos[F->Si[q]] = (int64_t)q;
  // M blocks of the last factored level
  int64_t mmax = 0;
  for (int b : boxes)
    if (t.level_of(b) == lv_last && bs[b].moff >= 0)
      mmax = std::max<int64_t>(mmax, bs[b].moff + (int64_t)bs[b].cur_len * bs[b].cur_len);
  std::vector<S> Mh(mmax);
  if (mmax) RSKX_HIP(hipMemcpy(Mh.data(), Marena.p, mmax * sizeof(S), hipMemcpyDeviceToHost));
  F->SI.clear(); F->SJ.clear(); F->SV.clear();
  for (int b : boxes) {
    if (t.level_of(b) != lv_last || bs[b].moff < 0) continue;
    const int nb = bs[b].cur_len;
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        const double v = (double)Mh[bs[b].moff + i + (int64_t)j * nb];
        if (v == 0) continue;                // sparse() keeps no zeros
        F->SI.push_back(pos[ord[pool_h[bs[b].cur_off + i]]]);
        F->SJ.push_back(pos[ord[pool_h[bs[b].cur_off + j]]]);
        F->SV.push_back(v);
      }
  }
  return RSKX_OK;
}

// Features of every box for the rank model (8 per box, see rskx.h): the
// point moments of the leaves on the device, summed up the tree on the host.
int box_features(const HostTree& t, const PointsDev& P, int d, double tol, double kwave,
                 std::vector<double>& feat) {
  const int nbox = t.nbox();
  std::vector<int64_t> lb, lc;
  std::vector<int> leaves;
  for (int b = 0; b < nbox; ++b)
    if (t.leaf(b) && t.cnt[b] > 0) { leaves.push_back(b); lb.push_back(t.beg[b]); lc.push_back(t.cnt[b]); }
  static DevBuf dbeg, dcnt, dmom;
  RSKX_TRY(upload(dbeg, lb));
  RSKX_TRY(upload(dcnt, lc));
  RSKX_TRY(dmom.ensure(std::max<size_t>(leaves.size(), 1) * 5 * sizeof(double)));
  launch_box_moments(dbeg.as<int64_t>(), dcnt.as<int64_t>(), (int)leaves.size(), P.x[0], P.x[1],
                     d > 2 ? P.x[2] : nullptr, dmom.as<double>(), g_stream);
  std::vector<double> lm(leaves.size() * 5), mom((size_t)nbox * 5, 0.0);
  RSKX_TRY(download(lm.data(), dmom.p, lm.size() * sizeof(double)));
  RSKX_TRY(sync_stream());
  for (size_t i = 0; i < leaves.size(); ++i)
    for (int j = 0; j < 5; ++j) mom[5 * (size_t)leaves[i] + j] = lm[5 * i + j];
  for (int b = nbox - 1; b > 0; --b)          // children follow their parents
    if (t.prnt[b] >= 0)
      for (int j = 0; j < 5; ++j) mom[5 * (size_t)t.prnt[b] + j] += mom[5 * (size_t)b + j];
  feat.assign((size_t)nbox * 8, 0.0);
  for (int b = 0; b < nbox; ++b) {
    const double* q = &mom[5 * (size_t)b];
    double lmax = 0, lmin = INFINITY, c2 = 0;
    for (int i = 0; i < d; ++i) {
      const double li = t.bsize[(size_t)d * b + i];
      lmax = std::max(lmax, li);
      lmin = std::min(lmin, li);
    }
    const double cntb = std::max(q[0], 1.0);
    double mean[3] = {q[1] / cntb, q[2] / cntb, q[3] / cntb}, m2 = 0;
    for (int i = 0; i < 3; ++i) m2 += mean[i] * mean[i];
    for (int i = 0; i < d; ++i) {
      const double dc = mean[i] - t.ctr[(size_t)d * b + i];
      c2 += dc * dc;
    }
    const double var = std::max(q[4] / cntb - m2, 0.0), L = lmax > 0 ? lmax : 1.0;
    double* f = &feat[8 * (size_t)b];
    f[0] = (t.level_of(b) + 1) / 16.0;
    f[1] = std::log2(cntb) / 20.0;
    f[2] = std::log2(L) / 8.0;
    f[3] = lmax > 0 ? lmin / lmax : 1.0;
    f[4] = std::sqrt(var) / L;
    f[5] = std::sqrt(c2) / L;
    f[6] = tol > 0 ? -std::log10(tol) / 16.0 : 1.0;
    f[7] = std::log2(1.0 + kwave * L) / 10.0;
  }
  return RSKX_OK;
}

// One level of a host-callback factorization (RSKX_KERNEL_CALLBACK): read
// back each box's slf list and its neighbour candidates (k_lists ran without
// a filter), call pxyfun for the filtered neighbours and proxy rows
// (rskelf.m:205, levels > 2), evaluate K = [A(nbr,slf); A(slf,nbr)' (symm
// 'n'); Kpxy] (rskelf.m:210-212) and A(slf,slf) (rskelf.m:219) through the
// entries callback, and upload them: K into the K arena for the QRCP route
// (its rows in mrows, leading dimension LvlBox::mcap), A(slf,slf) into
// Workspace::Ablk for the elimination.  hmn receives the kept neighbours.
template <typename S>
static int callback_level(const rskx_kernel& A, const rskx_proxy& PX, int d, const HostTree& t,
                          int64_t b0, int64_t nb, bool use_proxy, bool two, int64_t pool_top,
                          int64_t nslf_all, int64_t nnbr_all, const std::vector<int>& ord,
                          std::vector<LvlBox>& hb, std::vector<int>& hmn) {
  Workspace& ws = g_ws;
  std::vector<int> slf_all(std::max<int64_t>(nslf_all, 1)), nbr_all(std::max<int64_t>(nnbr_all, 1));
  std::vector<int> mnb(nb);
  if (nslf_all)
    RSKX_TRY(download(slf_all.data(), ws.pool.as<int>() + pool_top, nslf_all * sizeof(int)));
  if (nnbr_all) RSKX_TRY(download(nbr_all.data(), ws.nbr.p, nnbr_all * sizeof(int)));
  RSKX_TRY(download(mnb.data(), ws.mnbr.p, nb * sizeof(int)));
  RSKX_TRY(sync_stream());
  std::vector<int64_t> I, J;
  std::vector<double> blk, Kh, Kp, ent;
  std::vector<S> Kall;
  std::vector<double> Aall;
  std::vector<int> mrow(nb, 0);
  int64_t maxp = std::max(PX.p, 0);      // proxy-row buffer, grown on request
  int64_t offK = 0, offA = 0;
  auto bad = [](const char* what, int rc) {
    return fail(RSKX_ERR_ARG, std::string("FLAM:rskx:callback: ") + what + " returned " + std::to_string(rc));
  };
  for (int64_t ib = 0; ib < nb; ++ib) {
    LvlBox& B = hb[ib];
    const int n = B.n;
    const int b = (int)(b0 + ib);
    B.A_off = offA;
    if (n == 0) {
      B.mcap = 1;
      B.K_off = offK;
      offK += 0;
      hmn[ib] = 0;
      continue;
    }
    I.resize(n);
    for (int i = 0; i < n; ++i) I[i] = ord[slf_all[B.slf_off - pool_top + i]];
    int64_t nn = mnb[ib];
    J.resize(std::max<int64_t>(nn, 1));
    for (int64_t j = 0; j < nn; ++j) J[j] = ord[nbr_all[B.nbr_off + j]];
    int64_t rows = 0;
    if (use_proxy) {
      // a callback that needs more rows than the buffer holds says so in
      // *rows and is called once more with a buffer of that many rows
      const int64_t nn0 = nn;
      for (int attempt = 0; attempt < 2; ++attempt) {
        Kp.assign((size_t)std::max<int64_t>(maxp, 1) * n, 0.0);
        nn = nn0;
        for (int64_t j = 0; j < nn0; ++j) J[j] = ord[nbr_all[B.nbr_off + j]];
        rows = 0;
        const int rc = PX.fn(PX.ctx, I.data(), n, J.data(), &nn, &t.bsize[(size_t)d * b],
                             &t.ctr[(size_t)d * b], Kp.data(), maxp, &rows);
        if (rc) return bad("pxyfun", rc);
        if (rows <= maxp || attempt == 1) break;
        maxp = rows;
      }
      if (nn < 0 || nn > nn0 || rows < 0 || rows > maxp)
        return fail(RSKX_ERR_ARG, "FLAM:rskx:callback: pxyfun returned " + std::to_string(rows) +
                                      " proxy rows (buffer " + std::to_string(maxp) + ") and " +
                                      std::to_string(nn) + " of " + std::to_string(nn0) + " neighbours");
    }
    const int64_t m = nn * (two ? 2 : 1) + rows;
    hmn[ib] = (int)nn;
    mrow[ib] = (int)m;
    B.mcap = (int)std::max<int64_t>(m, 1);
    B.K_off = offK;
    Kall.resize(offK + (int64_t)B.mcap * n);
    S* K = Kall.data() + offK;
    if (nn) {
      ent.resize((size_t)nn * n);
      int rc = A.entries(A.ctx, J.data(), nn, I.data(), n, ent.data());      // A(nbr,slf)
      if (rc) return bad("A(I,J)", rc);
      for (int c = 0; c < n; ++c)
        for (int64_t i = 0; i < nn; ++i) K[i + (int64_t)c * m] = (S)ent[i + (size_t)c * nn];
      if (two) {
        rc = A.entries(A.ctx, I.data(), n, J.data(), nn, ent.data());         // A(slf,nbr)'
        if (rc) return bad("A(I,J)", rc);
        for (int c = 0; c < n; ++c)
          for (int64_t i = 0; i < nn; ++i) K[nn + i + (int64_t)c * m] = (S)ent[c + (size_t)i * n];
      }
    }
    for (int c = 0; c < n; ++c)
      for (int64_t i = 0; i < rows; ++i)
        K[m - rows + i + (int64_t)c * m] = (S)Kp[i + (size_t)c * maxp];
    offK += (int64_t)B.mcap * n;
    Aall.resize(offA + (int64_t)n * n);
    int rc = A.entries(A.ctx, I.data(), n, I.data(), n, Aall.data() + offA);   // A(slf,slf)
    if (rc) return bad("A(I,J)", rc);
    offA += (int64_t)n * n;
  }
  if (Kall.empty()) Kall.assign(1, (S)0);
  if (Aall.empty()) Aall.assign(1, 0.0);
  RSKX_TRY(upload(ws.K, Kall));
  RSKX_TRY(upload(ws.Ablk, Aall));
  RSKX_TRY(upload(ws.mrows, mrow));
  RSKX_TRY(upload(ws.boxes, hb));
  return RSKX_OK;
}

template <typename S>
int factor_impl(const double* x, int d, int64_t N, int nprob, const int64_t* poff, int occ,
                double rank_or_tol, const rskx_kernel& A, const rskx_proxy& PX,
                const double* pweight, const rskx_opts& o, rskx_factor* F) {
  using clk = std::chrono::steady_clock;
  Workspace& ws = g_ws;
  Prof& prof = F->prof;
  prof.on = o.profile != 0;
  prof.elim_only = o.profile == 2;
  auto t0 = clk::now();

  // ---- tree (device partition, host box bookkeeping)
  HostTree t;
  RSKX_TRY(build_tree_dev(x, d, N, nprob, poff, occ, o.lvlmax, o.ext, t, F->order));
  F->nlvl = t.nlvl;
  prof.host("tree", std::chrono::duration<double, std::milli>(clk::now() - t0).count());
  if (o.verb) {
    printf("%s\n", std::string(69, '-').c_str());
    printf("%3s | %6s | %19s | %19s | %10s\n", "lvl", "nblk", "start/end npts", "start/end npts/blk", "time (s)");
    printf("%s\n", std::string(69, '-').c_str());
    printf("%3s | %63.2e\n", "t", std::chrono::duration<double>(clk::now() - t0).count());
  }

  // ---- points in tree order (SoA)
  constexpr bool CPLX = is_cplx<S>::v;
  RSKX_TRY(ws.pts.ensure((size_t)N * (2 * d + (CPLX ? 3 : 2)) * sizeof(double)));
  double* pts = ws.pts.as<double>();
  // gather points, normals, weights and diagonal into tree order on the device
  launch_gather_points(x, A.nu, A.w, A.diag, F->order.as<int>(), N, d, pts, g_stream);
  PointsDev P{};
  for (int i = 0; i < 3; ++i) {
    P.x[i] = i < d ? pts + (size_t)i * N : nullptr;
    P.nu[i] = i < d ? pts + (size_t)(d + i) * N : nullptr;
  }
  P.w = pts + (size_t)2 * d * N;
  P.diag = pts + (size_t)(2 * d + 1) * N;
  P.d = d;
  P.kind = A.kind;
  P.scale = A.scale;
  if (CPLX) {            // the complex diagonal (interleaved) in two real rows
    P.diag_im = pts + (size_t)(2 * d + 2) * N;
    launch_zgather_diag(A.diag, F->order.as<int>(), N, pts + (size_t)(2 * d + 1) * N,
                        pts + (size_t)(2 * d + 2) * N, g_stream);
    P.k = A.wavenumber;
    P.eta = A.eta;
  }
  RSKX_HIP(hipGetLastError());
  // host-callback matrix: caller indices of the tree order, for A(I,J) and pxyfun
  const bool cb = A.kind == RSKX_KERNEL_CALLBACK;
  std::vector<int> cb_ord;
  if (cb) {
    cb_ord.resize(N);
    RSKX_TRY(download(cb_ord.data(), F->order.p, N * sizeof(int)));
    RSKX_TRY(sync_stream());
  }

  // ---- rank model: features of every box, predicted caps of the ranks
  const double tol_id = std::fmod(rank_or_tol, 1.0);
  std::vector<int> kcap;                 // per box (0: no cap)
  if (o.rank_model || o.record_boxes) {
    RSKX_TRY(box_features(t, P, d, tol_id, A.kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE ? A.wavenumber : 0.0,
                          F->boxfeat));
    F->boxrank.assign(t.nbox(), -1);
  }
  if (o.rank_model) {
    static DevBuf dX, dy;
    const int nb_all = t.nbox();
    RSKX_TRY(upload(dX, F->boxfeat));
    RSKX_TRY(dy.ensure(std::max(nb_all, 1) * sizeof(double)));
    launch_mlp_forward(o.rank_model->params.as<double>(), dX.as<double>(), nb_all, dy.as<double>(),
                       g_stream);
    std::vector<double> yh(nb_all);
    RSKX_TRY(download(yh.data(), dy.p, nb_all * sizeof(double)));
    RSKX_TRY(sync_stream());
    const double sc = o.rank_scale > 0 ? o.rank_scale : 1.0;
    kcap.assign(nb_all, 0);
    for (int b = 0; b < nb_all; ++b) {
      const double kh = std::exp2(10.0 * std::min(yh[b], 3.0)) - 1.0;   // predicted rank
      kcap[b] = std::max(1, (int)std::ceil(sc * (1.25 * std::max(kh, 0.0) + 4.0)));
    }
  }

  // ---- pool of point lists: leaves own their tree-order ranges
  int nbox = t.nbox();
  std::vector<BoxState> bs(nbox);
  for (int b = 0; b < nbox; ++b)
    if (t.cnt[b] > 0) { bs[b].cur_off = t.beg[b]; bs[b].cur_len = (int)t.cnt[b]; }
  int64_t pool_cap = N * (int64_t)(2 * t.nlvl + 2) + 64;
  RSKX_TRY(ws.pool.ensure(pool_cap * sizeof(int)));
  launch_iota(ws.pool.as<int>(), N, g_stream);
  int64_t pool_top = N;
  if (!ws.hflag) RSKX_HIP(hipHostMalloc((void**)&ws.hflag, 64 * sizeof(int), hipHostMallocDefault));
  ws.hflag[0] = 0;
  // elimination failures flagged on the device (read after a synchronization)
  auto elim_status = [&]() -> int {
    if (ws.hflag[0] & 1)
      return fail(RSKX_ERR_NUM, "FLAM:rskx:chol: matrix is not positive definite (symm 'p')");
    if (ws.hflag[0] & 4)
      return fail(RSKX_ERR_NUM, "FLAM:rskx:ldl: singular redundant block (a zero 1 x 1 or 2 x 2 pivot "
                                "of L D L', symm 'h')");
    if (ws.hflag[0] & 2)
      return fail(RSKX_ERR_NUM, "FLAM:rskx:lu: exactly singular redundant block");
    return RSKX_OK;
  };
  RSKX_TRY(ws.flag.ensure(64));
  RSKX_HIP(hipMemsetAsync(ws.flag.p, 0, 64, g_stream));
  RSKX_TRY(side_init());

  // 's' (rskelf.m:210,220-227): LU like 'n' but one-sided compression;
  // with real data T.' = T' and its sweeps are the 'n' sweeps
  const bool lu = (o.symm == 'n' || o.symm == 's');
  const bool two = (o.symm == 'n');
  const bool ldl = (o.symm == 'h');
  IdParams prm;
  prm.tol = std::fmod(rank_or_tol, 1.0);
  prm.kmax = (int)std::floor(rank_or_tol);
  prm.Tmax = o.Tmax;
  prm.rrqr_iter = o.rrqr_iter;

  int64_t nrem = N;
  int64_t nfac = 0;
  int mpar = 0;                 // M arena of the level below: ws.M[mpar]
  std::vector<int64_t> pblk(t.nlvl + 1, 0);
  std::vector<FacRec> lrec;
  for (int lv = 0; lv < t.nlvl; ++lv) {
    int64_t c = 0;
    for (int64_t b = t.lvp[lv]; b < t.lvp[lv + 1]; ++b) c += t.cnt[b] > 0;
    pblk[lv + 1] = pblk[lv] + c;
  }
  F->lvp.assign(1, 0);

  std::vector<LvlBox> hb;
  std::vector<Seg> hn, hc;
  std::vector<BoxOut> ho;
  std::vector<int> hk, hit;
  int lv_last = t.nlvl;           // last level factored
  for (int lv = t.nlvl - 1; lv >= 0; --lv) {
    // rskelf.m:172: opts.stop(t.nlvl - lvl, l), numeric stop = levels done
    if (o.stop >= 0 && F->nlvl_done >= o.stop) break;
    lv_last = lv;
    auto tl = clk::now();
    // host time per phase of the level ("h:" classes, full profiling only)
    auto th = tl;
    auto hlap = [&](const char* name) {
      const auto now = clk::now();
      prof.host(name, std::chrono::duration<double, std::milli>(now - th).count());
      th = now;
    };
    const int lvl_ref = lv + 1;   // the reference's level number
    const int64_t b0 = t.lvp[lv], nb = t.lvp[lv + 1] - b0;
    const int64_t nrem1 = nrem;
    LevelParams L{};
    L.radius = PX.radius;
    L.pkind = PX.kind;
    L.ppts = PX.pts;
    const bool use_proxy = lvl_ref > 2 && PX.kind != RSKX_PROXY_NONE;
    // point proxies of an unsymmetric matrix add the incoming rows
    L.np = use_proxy ? PX.p : 0;
    if (use_proxy && PX.kind == RSKX_PROXY_RING_CFIE) {
      double lmax = 0;
      for (int64_t b = t.lvp[lv]; b < t.lvp[lv + 1]; ++b)
        for (int i = 0; i < d; ++i) lmax = std::max(lmax, t.bsize[(size_t)d * b + i]);
      L.np = cfie_proxy_count(PX.p, A.wavenumber, lmax);
    }
    // point and combined-field proxies of an unsymmetric matrix add the
    // incoming rows
    L.p = L.np * ((PX.kind == RSKX_PROXY_POINTS_SL || PX.kind == RSKX_PROXY_RING_CFIE) && two ? 2 : 1);
    L.filter = use_proxy ? 1 : 0;
    L.two_sided = two ? 1 : 0;
    if (cb) {              // pxyfun filters and forms the proxy rows on the host
      L.np = L.p = 0;
      L.filter = 0;
    }

    hb.assign(nb, LvlBox{});
    hn.clear();
    hc.clear();
    int64_t off_slf = 0, off_nbr = 0, off_K = 0, off_perm = 0, off_T = 0;
    int nmax = 0, mcapmax = 0, mcmax = 0;
    for (int64_t ib = 0; ib < nb; ++ib) {
      const int b = (int)(b0 + ib);
      LvlBox& B = hb[ib];
      B.cs0 = (int)hc.size();
      if (t.leaf(b)) {
        hc.push_back(Seg{bs[b].cur_off, -1, bs[b].cur_len, 0});
      } else {
        for (int q = t.chld_off[b]; q < t.chld_off[b + 1]; ++q) {
          int c = t.chld[q];
          hc.push_back(Seg{bs[c].cur_off, bs[c].moff, bs[c].cur_len, 0});
        }
      }
      B.ncs = (int)hc.size() - B.cs0;
      int n = 0;
      for (int q = B.cs0; q < (int)hc.size(); ++q) n += hc[q].len;
      B.n = n;
      B.seg0 = (int)hn.size();
      for (int q = t.nbor_off[b]; q < t.nbor_off[b + 1]; ++q) {
        int j = t.nbor[q];
        if (t.level_of(j) == lv && !t.leaf(j)) {
          for (int q2 = t.chld_off[j]; q2 < t.chld_off[j + 1]; ++q2) {
            int c = t.chld[q2];
            if (bs[c].cur_len) hn.push_back(Seg{bs[c].cur_off, -1, bs[c].cur_len, 0});
          }
        } else if (bs[j].cur_len) {
          hn.push_back(Seg{bs[j].cur_off, -1, bs[j].cur_len, 0});
        }
      }
      B.nseg = (int)hn.size() - B.seg0;
      int mc = 0;
      for (int q = B.seg0; q < (int)hn.size(); ++q) mc += hn[q].len;
      B.mc = mc;
      B.mcap = std::max(1, mc * (two ? 2 : 1) + L.p);
      B.slf_off = pool_top + off_slf;  off_slf += n;
      B.nbr_off = off_nbr;             off_nbr += mc;
      B.K_off = off_K;                 off_K += (int64_t)B.mcap * n;
      B.perm_off = off_perm;           off_perm += n;
      B.ldt = n;
      B.T_off = off_T;                 off_T += (CPLX ? 6 : 3) * (int64_t)n * n;
      B.W_off = 0;                     // laid out by the elimination routing
      for (int i = 0; i < d; ++i) {
        B.ctr[i] = t.ctr[(size_t)d * b + i];
        B.l[i] = t.bsize[(size_t)d * b + i];
      }
      B.pw = pweight[t.prob[b]];
      nmax = std::max(nmax, n);
      mcapmax = std::max(mcapmax, B.mcap);
      mcmax = std::max(mcmax, mc);
    }
    if (pool_top + off_slf > pool_cap) return fail(RSKX_ERR_MEM, "FLAM:rskx:pool: point-list pool exhausted");
    RSKX_TRY(upload(ws.boxes, hb));
    RSKX_TRY(upload(ws.nsegs, hn));
    RSKX_TRY(upload(ws.csegs, hc));
    RSKX_TRY(ws.nbr.ensure(std::max<int64_t>(off_nbr, 1) * sizeof(int)));
    RSKX_TRY(ws.mnbr.ensure(nb * sizeof(int)));
    RSKX_TRY(ws.mrows.ensure(nb * sizeof(int)));
    RSKX_TRY(ws.kout.ensure(nb * sizeof(int)));
    RSKX_TRY(ws.iter.ensure(nb * sizeof(int)));
    // real factorizations without callbacks lay K out after the ID routing,
    // for the boxes that use it
    if (CPLX || cb) RSKX_TRY(ws.K.ensure(std::max<int64_t>(off_K, 1) * sizeof(S)));
    RSKX_TRY(ws.perm.ensure(std::max<int64_t>(off_perm, 1) * sizeof(int)));
    // with a rank model the real ID's buffers are laid out after routing
    const bool sized_by_model = !kcap.empty() && !CPLX;
    if (!sized_by_model) RSKX_TRY(ws.T.ensure(std::max<int64_t>(off_T, 1) * sizeof(double)));
    hlap("h:segments");

    hipEvent_t ev;
    prof.begin("lists", ev);
    launch_lists(ws.boxes.as<LvlBox>(), ws.nsegs.as<Seg>(), ws.csegs.as<Seg>(), ws.pool.as<int>(),
                 ws.nbr.as<int>(), ws.mnbr.as<int>(), P, L, (int)nb, g_stream);
    prof.end("lists", ev);
    std::vector<int> hmn(nb);                  // neighbour rows the routing sized K for
    std::vector<char> wide(nb, 0);             // boxes of the wide-box Gram ID
    const bool gram = !cb && !CPLX && prm.tol >= 1e-7 &&
                      (o.id_gram > 0 || (o.id_gram == 0 && sizeof(S) == 4));
    // boxes beyond the register classes take the wide-box Gram ID on the
    // automatic route in FP64 too (a QRCP over their K would run on one
    // workgroup each; rocSOLVER has no column-pivoted QR)
    const bool wide_ok = gram || (!cb && !CPLX && prm.tol >= 1e-7 && o.id_gram == 0);
    if constexpr (!CPLX) {
      if (cb) {
        RSKX_TRY(callback_level<S>(A, PX, d, t, b0, nb, use_proxy, two, pool_top, off_slf, off_nbr,
                                   cb_ord, hb, hmn));
        P.blk = ws.Ablk.as<double>();
      }
    }
    if constexpr (CPLX) {
      // complex: every box assembles K in global memory and runs the
      // complex QRCP ID (k_zid)
      std::vector<int> ida;
      int nmx = 0;
      int64_t mvmax = 0, Kbytes = 0;
      for (int64_t ib = 0; ib < nb; ++ib) {
        ida.push_back((int)ib);
        nmx = std::max(nmx, hb[ib].n);
        mvmax = std::max<int64_t>(mvmax, std::max(hb[ib].mcap, 3 * hb[ib].n));
        Kbytes += (int64_t)hb[ib].mcap * hb[ib].n * (int64_t)sizeof(S);
      }
      RSKX_TRY(upload(ws.ids, ida));
      prof.begin("assemble", ev);
      launch_zassemble<S>(ws.ids.as<int>(), ws.boxes.as<LvlBox>(), ws.pool.as<int>(),
                          ws.nbr.as<int>(), ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.K.as<S>(), P,
                          L, (int)nb, g_stream);
      prof.end("assemble", ev, Kbytes);
      size_t shz = (size_t)(2 * nmx + 2 * mvmax + 64) * sizeof(double);
      zd* vglob = nullptr;
      if (shz > 160 * 1024) {               // Householder vectors in global memory
        shz = (size_t)(2 * nmx + 64) * sizeof(double);
        RSKX_TRY(ws.V.ensure((size_t)nb * mvmax * sizeof(zd)));
        vglob = ws.V.as<zd>();
      }
      if (shz > 160 * 1024)
        return fail(RSKX_ERR_UNSUP, "FLAM:rskx:size: box too large for the complex ID kernel (n=" +
                                         std::to_string(nmx) + ")");
      prof.begin("id", ev);
      launch_zid<S>(ws.ids.as<int>(), ws.boxes.as<LvlBox>(), ws.mrows.as<int>(), ws.K.as<S>(),
                    ws.perm.as<int>(), ws.T.as<double>(), prm, ws.kout.as<int>(), ws.iter.as<int>(),
                    (int)nb, shz, vglob, mvmax, g_stream);
      prof.end("id", ev);
    } else {
    // ID: boxes whose K fits a register tile assemble and factor it in
    // registers (k_id_reg); the rest assemble K in global memory first
    // FP32 factorizations at tol >= 1e-7 take the Gram route (k_gram_id),
    // whose shared memory is sized by the unfiltered neighbour count, so
    // the host need not wait for k_lists; the QR routes read back the
    // filtered counts to choose register classes
    bool exact = !gram && !cb;
    for (int64_t ib = 0; ib < nb && gram && !exact; ++ib) {
      const int n = hb[ib].n, mc = hb[ib].mc;
      if (n > 0 && gram_id_class(n) >= 0 &&
          gram_id_shmem(gram_class_tuned(n), n, mc, mc * (two ? 2 : 1) + L.p, L.p) > 160 * 1024)
        exact = true;                        // the bound is too loose: use the filtered counts
    }
    if (cb) {
      // hmn holds the neighbours pxyfun kept
    } else if (!exact) {
      for (int64_t ib = 0; ib < nb; ++ib) hmn[ib] = hb[ib].mc;
    } else {
      RSKX_TRY(download(hmn.data(), ws.mnbr.p, nb * sizeof(int)));
      RSKX_TRY(sync_stream());
    }
    std::vector<int> idg, idr[2], idq[kGramClasses], idall, idw;
    size_t shm_r[2] = {0, 0}, shm_q[kGramClasses] = {0, 0, 0};
    int nmax_g = 0, mcapmax_g = 0, nmax_w = 0;
    // a box whose point data would not fit a class's shared memory goes to
    // the next route: Gram -> register QRCP -> QRCP on K in global memory;
    // boxes beyond the register-resident Gram classes (every such box on
    // the Gram route, and on FP64's automatic route too) take the batched
    // wide-box Gram ID (K in global memory, G by MFMA tiles, blocked
    // pivoted Cholesky) up to kWideNmax points
    for (int64_t ib = 0; ib < nb; ++ib) {
      const int n = hb[ib].n, mb = hmn[ib] * (two ? 2 : 1) + L.p;
      if (gram && n > 0 && mb > 0 && gram_id_class(n) >= 0) {
        const int c = gram_class_tuned(n);
        const size_t by = gram_id_shmem(c, n, hmn[ib], mb, L.p);
        if (by <= 160 * 1024) {
          idq[c].push_back((int)ib);
          shm_q[c] = std::max(shm_q[c], by);
          continue;
        }
      }
      if (wide_ok && n > 0 && mb > 0 && n <= kWideNmax && (gram || gram_id_class(n) < 0)) {
        idw.push_back((int)ib);
        wide[ib] = 1;
        nmax_w = std::max(nmax_w, n);
        continue;
      }
      int cls = -1;
      for (int c = 1; c >= 0 && cls < 0; --c) {
        int ml, nl;
        id_reg_limits(sizeof(S), c, &ml, &nl);
        if (mb <= ml && n <= nl) cls = c;
      }
      // staging (8 x 1024 S), proxy points, point data, norms, v, red, maps
      const size_t by = (size_t)8 * 1024 * sizeof(S) +
                        (size_t)(3 * L.p + 7 * n + 5 * hmn[ib] + mb + 32) * sizeof(double) +
                        (size_t)2 * n * sizeof(int) + 64;
      if (!cb && cls >= 0 && g_tune.id_reg && n > 0 && mb > 0 && by <= 160 * 1024) {
        idr[cls].push_back((int)ib);
        shm_r[cls] = std::max(shm_r[cls], by);
      } else {
        idg.push_back((int)ib);
        if (mb > 0 || cb) {              // without rows (the root) k_id returns rank 0 at once
          nmax_g = std::max(nmax_g, n);
          mcapmax_g = std::max(mcapmax_g, hb[ib].mcap);
        }
      }
    }
    // K for the boxes that use it (QR, register QR, wide Gram); the wide
    // boxes' K in groups of at most kWideKBytes, each assembled once the
    // previous group's G is formed
    std::vector<int> wgrp{0};
    if (!cb) {
      int64_t oK = 0;
      auto place = [&](int ib) { hb[ib].K_off = oK; oK += (int64_t)hb[ib].mcap * hb[ib].n; };
      for (int ib : idg) place(ib);
      for (int c = 0; c < 2; ++c) for (int ib : idr[c]) place(ib);
      // wide boxes: K and its leading dimension aligned to 4 elements (the
      // Gram kernel reads 4-element vectors)
      const int64_t kbase = (oK + 3) & ~int64_t(3), cap = kWideKBytes / (int64_t)sizeof(S);
      int64_t acc = 0, gmax = 0;
      for (size_t i = 0; i < idw.size(); ++i) {
        const int ib = idw[i];
        hb[ib].mcap = (hb[ib].mcap + 3) & ~3;
        const int64_t sz = (int64_t)hb[ib].mcap * hb[ib].n;
        if (acc > 0 && acc + sz > cap) { wgrp.push_back((int)i); acc = 0; }
        hb[ib].K_off = kbase + acc;
        acc += sz;
        gmax = std::max(gmax, acc);
      }
      RSKX_TRY(ws.K.ensure(std::max<int64_t>(kbase + gmax, 1) * sizeof(S)));
      RSKX_TRY(upload(ws.boxes, hb));
    }
    wgrp.push_back((int)idw.size());
    if (sized_by_model) {
      // the Gram ID keeps only its rank cap of R rows: size every Gram
      // box's R / T buffers (3 ldt n doubles) by the predicted cap
      std::vector<char> isg(nb, 0);
      for (int c = 0; c < kGramClasses; ++c)
        for (int ib : idq[c]) isg[ib] = 1;
      off_T = 0;
      for (int64_t ib = 0; ib < nb; ++ib) {
        LvlBox& B = hb[ib];
        B.ldt = isg[ib] ? std::min(B.n, kcap[b0 + ib]) : B.n;
        B.T_off = off_T;
        off_T += 3 * (int64_t)B.ldt * B.n;
      }
      RSKX_TRY(ws.T.ensure(std::max<int64_t>(off_T, 1) * sizeof(double)));
      RSKX_TRY(upload(ws.boxes, hb));
    }
    F->id_scratch = std::max<int64_t>(F->id_scratch, off_T * (int64_t)sizeof(double));
    // QR boxes and wide Gram boxes first (both assemble K in global memory)
    idall = idg;
    idall.insert(idall.end(), idw.begin(), idw.end());
    idall.insert(idall.end(), idr[0].begin(), idr[0].end());
    idall.insert(idall.end(), idr[1].begin(), idr[1].end());
    for (int c = kGramClasses - 1; c >= 0; --c) idall.insert(idall.end(), idq[c].begin(), idq[c].end());
    RSKX_TRY(upload(ws.ids, idall));
    const int ng = (int)idg.size(), nwd = (int)idw.size(), nr0 = (int)idr[0].size(),
              nr1 = (int)idr[1].size();
    const int* dids = ws.ids.as<int>();
    const int* wids = dids + ng;
    const int* rids = wids + nwd;          // register QR boxes
    int nq = 0;
    for (int c = 0; c < kGramClasses; ++c) nq += (int)idq[c].size();
    int64_t Kbytes = 0;
    int nmax_a = 0;
    for (int ib : idg) {
      Kbytes += (int64_t)hb[ib].mcap * hb[ib].n;
      nmax_a = std::max(nmax_a, hb[ib].n);
    }
    if (ng && !cb) {                       // callbacks uploaded K already
      prof.begin("assemble", ev);
      launch_assemble<S>(dids, ws.boxes.as<LvlBox>(), ws.pool.as<int>(), ws.nbr.as<int>(),
                         ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.K.as<S>(), P, L, ng, nmax_a,
                         g_stream);
      prof.end("assemble", ev, Kbytes * (int64_t)sizeof(S));
    }
    // wide boxes: K by groups, each group's G = K'K before the next group's
    // K overwrites it
    for (size_t g = 0; g + 1 < wgrp.size(); ++g) {
      const int a = wgrp[g], e = wgrp[g + 1];
      if (e == a) continue;
      int nm = 0;
      int64_t kb = 0;
      for (int i = a; i < e; ++i) {
        nm = std::max(nm, hb[idw[i]].n);
        kb += (int64_t)hb[idw[i]].mcap * hb[idw[i]].n;
      }
      prof.begin("assemble", ev);
      launch_assemble<S>(wids + a, ws.boxes.as<LvlBox>(), ws.pool.as<int>(), ws.nbr.as<int>(),
                         ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.K.as<S>(), P, L, e - a, nm,
                         g_stream);
      prof.end("assemble", ev, kb * (int64_t)sizeof(S));
      prof.begin("gram", ev);             // flops added once the row counts are back
      gram_wide_gram<S>(wids + a, e - a, nm, ws.boxes.as<LvlBox>(), ws.mrows.as<int>(),
                        ws.K.as<S>(), ws.T.as<double>(), g_stream);
      prof.end("gram", ev);
    }
    prof.begin("id", ev);
    if (ng) {
      size_t shg = (size_t)(2 * nmax_g + std::max(nmax_g, mcapmax_g) + 32) * sizeof(double);
      double* vglob = nullptr;             // Householder vectors too long for shared memory
      const int64_t vstride = std::max(nmax_g, mcapmax_g);
      if (shg > (size_t)g_tune.id_lds) {
        shg = (size_t)(3 * nmax_g + 32) * sizeof(double);
        RSKX_TRY(ws.V.ensure((size_t)ng * vstride * sizeof(double)));
        vglob = ws.V.as<double>();
      }
      if (shg > 160 * 1024)
        return fail(RSKX_ERR_UNSUP, "FLAM:rskx:size: box too large for the ID kernel (n=" +
                                         std::to_string(nmax_g) + ")");
      launch_id<S>(0, dids, ws.boxes.as<LvlBox>(), ws.mrows.as<int>(), ws.K.as<S>(),
                   ws.perm.as<int>(), ws.T.as<double>(), prm, ws.kout.as<int>(), ws.iter.as<int>(),
                   ng, shg, vglob, vstride, g_stream);
    }
    for (int c = 0; c < 2; ++c) {
      const int cnt = c == 0 ? nr0 : nr1;
      if (!cnt) continue;
      launch_id_reg<S>(c, rids + (c == 0 ? 0 : nr0), ws.boxes.as<LvlBox>(), ws.pool.as<int>(),
                       ws.nbr.as<int>(), ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.K.as<S>(),
                       ws.perm.as<int>(), P, L, prm, ws.kout.as<int>(), cnt, shm_r[c], g_stream);
    }
    if (nr0 + nr1) {
      int nmx = 0;
      for (int c = 0; c < 2; ++c) for (int ib : idr[c]) nmx = std::max(nmx, hb[ib].n);
      launch_id<S>(1, rids, ws.boxes.as<LvlBox>(), ws.mrows.as<int>(), ws.K.as<S>(),
                   ws.perm.as<int>(), ws.T.as<double>(), prm, ws.kout.as<int>(), ws.iter.as<int>(),
                   nr0 + nr1, (size_t)(3 * nmx + 32) * sizeof(double), nullptr, 0, g_stream);
    }
    const int* qids = rids + nr0 + nr1;
    {
      // the Gram classes run concurrently (largest boxes on a side stream,
      // the smallest on the other, the middle class on g_stream); the
      // tails below read every class's output, so they follow the join
      int used = 0;
      for (int c = 0; c < kGramClasses; ++c) used += idq[c].empty() ? 0 : 1;
      const bool conc = used > 1 && g_tune.concurrent;
      if (conc) RSKX_TRY(side_fork());
      int off = 0;
      for (int c = kGramClasses - 1; c >= 0; --c) {     // largest boxes first
        const int cnt = (int)idq[c].size();
        if (!cnt) continue;
        hipStream_t st = !conc ? g_stream : c == 2 ? g_side.s[0] : c == 0 ? g_side.s[1] : g_stream;
        launch_gram_id(c, qids + off, ws.boxes.as<LvlBox>(), ws.pool.as<int>(), ws.nbr.as<int>(),
                       ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.T.as<double>(), ws.perm.as<int>(),
                       P, L, prm, ws.kout.as<int>(), cnt, shm_q[c], st);
        off += cnt;
      }
      if (conc) RSKX_TRY(side_join());
    }
    // wide Gram boxes: the blocked pivoted Cholesky of G, R by position
    if (nwd) {
      RSKX_TRY(ws.big.ensure((size_t)nwd * gram_wide_stride(nmax_w) * sizeof(double)));
      RSKX_TRY(gram_wide_chol(wids, nwd, nmax_w, ws.boxes.as<LvlBox>(), ws.mrows.as<int>(),
                              ws.T.as<double>(), ws.perm.as<int>(), ws.kout.as<int>(), prm,
                              ws.big.as<double>(), g_stream));
    }
    // T = R11 \ R12: the Gram tail for the register-class boxes and the
    // wide boxes of modest rank; rocBLAS dtrsm, then the tail's refinement
    // (MODE 3), for wide boxes of rank >= kGwTrsm (the host reads their
    // ranks first)
    constexpr int kGwTrsm = 256;
    std::vector<int> wt2, wt3;
    if (nwd) {
      bool any = false;
      for (int ib : idw) any = any || hb[ib].n > kGwTrsm;
      if (any) {
        std::vector<int> kw(nb);
        RSKX_TRY(download(kw.data(), ws.kout.p, nb * sizeof(int)));
        RSKX_TRY(sync_stream());
        for (int ib : idw) (kw[ib] >= kGwTrsm && kw[ib] < hb[ib].n ? wt3 : wt2).push_back(ib);
        for (int ib : wt3)
          RSKX_TRY(gram_wide_tsolve(ws.boxes.as<LvlBox>(), ib, hb[ib], kw[ib], ws.T.as<double>(),
                                    g_stream));
      } else {
        wt2 = idw;
      }
    }
    if (nq) {
      int nmx = 0;
      for (int c = 0; c < kGramClasses; ++c) for (int ib : idq[c]) nmx = std::max(nmx, hb[ib].n);
      launch_id<S>(2, qids, ws.boxes.as<LvlBox>(), ws.mrows.as<int>(), ws.K.as<S>(),
                   ws.perm.as<int>(), ws.T.as<double>(), prm, ws.kout.as<int>(), ws.iter.as<int>(),
                   nq, (size_t)(3 * nmx + 32) * sizeof(double), nullptr, 0, g_stream);
    }
    if (!wt2.empty() || !wt3.empty()) {
      std::vector<int> wl(wt2);
      wl.insert(wl.end(), wt3.begin(), wt3.end());
      RSKX_TRY(upload(ws.wids, wl));
      // the tails' three n-vectors in shared memory, or, for boxes too wide
      // for it, in global memory (32 doubles of shared reduction scratch)
      for (int mode = 2; mode <= 3; ++mode) {
        const std::vector<int>& wt = mode == 2 ? wt2 : wt3;
        if (wt.empty()) continue;
        int nm = 0;
        for (int ib : wt) nm = std::max(nm, hb[ib].n);
        // MODE 3's refinement keeps six n-vectors in global memory always
        const int nv = mode == 3 ? 6 : 3;
        size_t shm = (size_t)(3 * nm + 32) * sizeof(double);
        double* vg = nullptr;
        if (mode == 3 || shm > 160 * 1024) {
          RSKX_TRY(ws.wv.ensure((size_t)wt.size() * nv * nm * sizeof(double)));
          vg = ws.wv.as<double>();
          shm = 32 * sizeof(double);
        }
        launch_id<S>(mode, ws.wids.as<int>() + (mode == 2 ? 0 : wt2.size()), ws.boxes.as<LvlBox>(),
                     ws.mrows.as<int>(), ws.K.as<S>(), ws.perm.as<int>(), ws.T.as<double>(), prm,
                     ws.kout.as<int>(), ws.iter.as<int>(), (int)wt.size(), shm, vg,
                     (int64_t)nv * nm, g_stream);
      }
    }
    prof.end("id", ev);
    }
    RSKX_HIP(hipGetLastError());
    hlap("h:routing");
    hk.resize(nb);
    hit.resize(nb);
    std::vector<int> hm(nb);
    RSKX_TRY(download(hk.data(), ws.kout.p, nb * sizeof(int)));
    RSKX_TRY(download(hit.data(), ws.iter.p, nb * sizeof(int)));
    RSKX_TRY(download(hm.data(), ws.mrows.p, nb * sizeof(int)));
    RSKX_TRY(sync_stream());
    hlap("h:wait");
    RSKX_TRY(elim_status());                  // the previous level's elimination
    prof.flush();
    if constexpr (!CPLX) {
      // boxes whose rank exceeded their predicted cap (kout = cap + 1): the
      // plain Gram ID again with full buffers
      std::vector<int> redo;
      for (int64_t ib = 0; ib < nb; ++ib)
        if (hb[ib].ldt < hb[ib].n && hk[ib] > hb[ib].ldt) redo.push_back((int)ib);
      if (!redo.empty()) {
        F->rank_fallbacks += (int64_t)redo.size();
        for (int ib : redo) {
          hb[ib].ldt = hb[ib].n;
          hb[ib].T_off = off_T;
          off_T += 3 * (int64_t)hb[ib].n * hb[ib].n;
        }
        RSKX_TRY(ws.T.ensure(std::max<int64_t>(off_T, 1) * sizeof(double), true));
        RSKX_TRY(upload(ws.boxes, hb));
        std::vector<int> rid[kGramClasses];
        size_t rsh[kGramClasses] = {0, 0, 0};
        int nmx = 0;
        for (int ib : redo) {
          const int n = hb[ib].n, c = gram_class_tuned(n);
          const int mb = hmn[ib] * (two ? 2 : 1) + L.p;
          rid[c].push_back(ib);
          rsh[c] = std::max(rsh[c], gram_id_shmem(c, n, hmn[ib], mb, L.p));
          nmx = std::max(nmx, n);
        }
        std::vector<int> rall;
        for (int c = 0; c < kGramClasses; ++c) rall.insert(rall.end(), rid[c].begin(), rid[c].end());
        RSKX_TRY(upload(ws.ids, rall));
        int off = 0;
        for (int c = 0; c < kGramClasses; ++c) {
          if (rid[c].empty()) continue;
          launch_gram_id(c, ws.ids.as<int>() + off, ws.boxes.as<LvlBox>(), ws.pool.as<int>(),
                         ws.nbr.as<int>(), ws.mnbr.as<int>(), ws.mrows.as<int>(), ws.T.as<double>(),
                         ws.perm.as<int>(), P, L, prm, ws.kout.as<int>(), (int)rid[c].size(), rsh[c],
                         g_stream);
          off += (int)rid[c].size();
        }
        launch_id<S>(2, ws.ids.as<int>(), ws.boxes.as<LvlBox>(), ws.mrows.as<int>(), ws.K.as<S>(),
                     ws.perm.as<int>(), ws.T.as<double>(), prm, ws.kout.as<int>(), ws.iter.as<int>(),
                     (int)rall.size(), (size_t)(3 * nmx + 32) * sizeof(double), nullptr, 0, g_stream);
        RSKX_HIP(hipGetLastError());
        RSKX_TRY(download(hk.data(), ws.kout.p, nb * sizeof(int)));
        RSKX_TRY(download(hit.data(), ws.iter.p, nb * sizeof(int)));
        RSKX_TRY(sync_stream());
        F->id_scratch = std::max<int64_t>(F->id_scratch, off_T * (int64_t)sizeof(double));
      }
    }
    if (!F->boxrank.empty())
      for (int64_t ib = 0; ib < nb; ++ib)
        if (hb[ib].n > 0) F->boxrank[b0 + ib] = hk[ib];
    for (int64_t ib = 0; ib < nb; ++ib)
      if (hit[ib] < 0) {
        fprintf(stderr, "Warning: FLAM:id:maxIterCount: Maximum number of RRQR iterations reached.\n");
        break;
      }

    // ---- lay out this level's factors
    ho.assign(nb, BoxOut{});
    int64_t offM = 0, off_nl = 0, need_v = 0, need_i = 0;
    int kmx = 0, rmx = 0;
    double id_flops = 0, el_flops = 0, gram_flops = 0;
    for (int64_t ib = 0; ib < nb; ++ib) {
      const LvlBox& B = hb[ib];
      BoxOut& O = ho[ib];
      O.k = hk[ib];
      O.r = B.n - O.k;
      const int64_t k = O.k, r = O.r, m = hm[ib], n = B.n;
      if (wide[ib] && n > 0 && m > 0) {   // G = K'K in its own class, then the rest
        gram_flops += (double)m * n * (n + 1);
        id_flops += 2.0 * (double)n * n * k + (double)k * k * r;
      } else if (gram && n > 0 && m > 0)   // G = K'K, pivoted Cholesky, T
        id_flops += (double)m * n * (n + 1) + 2.0 * (double)n * n * k + (double)k * k * r;
      else                                                  // truncated QRCP, T (x4 complex)
        id_flops += (CPLX ? 4.0 : 1.0) * (4.0 * m * n * k - 2.0 * k * k * (m + n) +
                                         4.0 / 3.0 * k * k * k + (double)k * k * r);
      if (r > 0 && CPLX) {                // complex: T, LU, E, G and the pivots only
        O.fac = (int)(nfac++);
        O.T = need_v; need_v += k * r;
        O.LU = need_v; need_v += r * r;
        O.E = need_v; need_v += k * r;
        O.G = need_v; need_v += r * k;
        O.Li = O.Ui = O.Tt = O.Et = O.Gt = O.LUt = O.Lti = O.Uti = -1;
        O.sk = need_i; need_i += k;
        O.rd = need_i; need_i += r;
        O.piv = need_i; need_i += r;
        O.D = -1;
        O.M_off = offM; offM += k * k;
        O.nl_off = pool_top + off_slf + off_nl; off_nl += k;
        kmx = std::max<int>(kmx, (int)k);
        rmx = std::max<int>(rmx, (int)r);
        el_flops += 4.0 * (4.0 * n * k * r + 2.0 / 3.0 * r * r * r + 2.0 * k * r * r + 2.0 * k * k * r);
      } else if (r > 0) {
        O.fac = (int)(nfac++);
        O.T = need_v; need_v += k * r;
        O.LU = need_v; need_v += r * r;
        O.E = need_v; need_v += k * r;
        O.G = need_v; if (lu) need_v += r * k;
        const int64_t nbk = (r + kDiagBlock - 1) / kDiagBlock;
        O.Li = need_v; need_v += nbk * kDiagBlock * kDiagBlock;
        O.Ui = need_v; if (lu) need_v += nbk * kDiagBlock * kDiagBlock;
        O.Tt = need_v; need_v += k * r;
        O.Et = need_v; need_v += k * r;
        O.Gt = need_v; if (lu) need_v += r * k;
        O.LUt = need_v; need_v += r * r;
        O.Lti = need_v; need_v += nbk * kDiagBlock * kDiagBlock;
        O.Uti = need_v; if (lu) need_v += nbk * kDiagBlock * kDiagBlock;
        O.sk = need_i; need_i += k;
        O.rd = need_i; need_i += r;
        O.piv = need_i; if (lu) need_i += r;
        O.D = -1;
        if (ldl) { O.D = need_v; need_v += 2 * r; }   // D's diagonal, then its subdiagonal
        O.M_off = offM; offM += k * k;
        O.nl_off = pool_top + off_slf + off_nl; off_nl += k;
        kmx = std::max<int>(kmx, (int)k);
        rmx = std::max<int>(rmx, (int)r);
        el_flops += 4.0 * n * k * r + (lu ? 2.0 / 3.0 : 1.0 / 3.0) * r * r * r +
                    (lu ? 2.0 : 1.0) * k * r * r + 2.0 * k * k * r;
      } else {
        O.fac = -1;
        O.D = -1;
        O.M_off = offM; offM += n * n;
        O.nl_off = -1;
      }
    }
    if (prof.wants("id")) prof.st["id"].flops += id_flops;
    if (prof.wants("gram") && gram_flops > 0) prof.st["gram"].flops += gram_flops;
    F->lval.emplace_back();
    F->lidx.emplace_back();
    RSKX_TRY(F->lval.back().ensure((need_v + 1) * sizeof(S)));
    RSKX_TRY(F->lidx.back().ensure((need_i + 1) * sizeof(int)));
    S* fval = F->lval.back().as<S>();
    int* fidx = F->lidx.back().as<int>();
    const int mcur = 1 - mpar;
    RSKX_TRY(ws.M[mcur].ensure(std::max<int64_t>(offM, 1) * sizeof(S)));
    RSKX_TRY(ws.M[mpar].ensure(sizeof(S)));
    RSKX_TRY(upload(ws.outs, ho));
    // k_elim_reg writes the transposed copies itself; the complex sweeps
    // read the factor as stored (nothing to pack)
    std::vector<char> regbox(nb, CPLX ? 1 : 0);
    if constexpr (CPLX) {
      // complex: boxes whose blocks and point cache fit shared memory, and
      // the rest (global scratch; also the no-compression copies)
      std::vector<int> zc[2];
      size_t shz = 0;
      for (int64_t ib = 0; ib < nb; ++ib) {
        const BoxOut& O = ho[ib];
        const size_t by = O.r > 0 ? zelim_elems(O.k, O.r) * sizeof(S) + 8 * sizeof(double) * (size_t)hb[ib].n : 0;
        if (O.r > 0 && by <= 158 * 1024) { zc[0].push_back((int)ib); shz = std::max(shz, by); }
        else zc[1].push_back((int)ib);
      }
      if (!zc[1].empty()) {                // global scratch for these boxes only
        int64_t oW = 0;
        for (int ib : zc[1]) {
          hb[ib].W_off = oW;
          oW += 2 * (int64_t)hb[ib].n * hb[ib].n + 12 * (int64_t)hb[ib].n;
        }
        RSKX_TRY(ws.W.ensure(oW * sizeof(S)));
        RSKX_TRY(upload(ws.boxes, hb));
      }
      std::vector<int> zall(zc[0]);
      zall.insert(zall.end(), zc[1].begin(), zc[1].end());
      RSKX_TRY(upload(ws.ids, zall));
      prof.begin("factor", ev);
      const bool sym = o.symm == 's';
      if (!zc[0].empty())
        launch_zelim<S>(sym, true, shz, ws.ids.as<int>(), (int)zc[0].size(), ws.boxes.as<LvlBox>(),
                        ws.outs.as<BoxOut>(), ws.csegs.as<Seg>(), ws.pool.as<int>(), ws.perm.as<int>(),
                        ws.T.as<double>(), ws.M[mpar].as<S>(), ws.M[mcur].as<S>(), ws.W.as<S>(), fval,
                        fidx, P, ws.flag.as<int>(), g_stream);
      if (!zc[1].empty())
        launch_zelim<S>(sym, false, 0, ws.ids.as<int>() + zc[0].size(), (int)zc[1].size(),
                        ws.boxes.as<LvlBox>(), ws.outs.as<BoxOut>(), ws.csegs.as<Seg>(),
                        ws.pool.as<int>(), ws.perm.as<int>(), ws.T.as<double>(), ws.M[mpar].as<S>(),
                        ws.M[mcur].as<S>(), ws.W.as<S>(), fval, fidx, P, ws.flag.as<int>(), g_stream);
      prof.end("factor", ev, 0, el_flops);
    } else {
    // elimination: boxes whose blocks fit two per CU in shared memory, one
    // per CU, or neither (global scratch; also the no-compression copies)
    const size_t kHalf = 78 * 1024, kFull = 158 * 1024;   // 1024-thread workgroups
    // Boxes whose augmented matrix fits registers go to k_elim_reg: 8-wave
    // workgroups (two per CU, n <= 192 in FP32, 128 in FP64) when the box
    // and its shared memory allow, else 16-wave workgroups
    std::vector<int> cls[3], creg[2], cmf[kMfmaShapes], cbig;
    size_t shm_cls[3] = {0, 0, 0}, shm_reg[2] = {0, 0}, shm_mf[kMfmaShapes] = {0, 0, 0, 0};
    const int nreg8 = std::min(elim_reg_nmax(sizeof(S), 8), g_tune.elim8_nmax),
              nreg16 = elim_reg_nmax(sizeof(S), 16);
    for (int64_t ib = 0; ib < nb; ++ib) {
      const BoxOut& O = ho[ib];
      const int n = hb[ib].n;
      // the largest boxes: dense LU / Cholesky with the whole GPU
      if (O.r > 0 && !ldl && n > g_tune.big_n) {
        cbig.push_back((int)ib);
        continue;
      }
      // Cholesky boxes go to the matrix-core elimination when its tiles and
      // shared memory hold them
      if (O.r > 0 && !lu && !ldl && g_tune.elim_mfma) {
        size_t by = 0;
        const int c = elim_mfma_shape(n, O.k, O.r, sizeof(S), &by);
        if (c >= 0) {
          cmf[c].push_back((int)ib);
          regbox[ib] = 1;
          shm_mf[c] = std::max(shm_mf[c], by);
          continue;
        }
      }
      // 'h' (Bunch-Kaufman interchanges and 2 x 2 pivots) always takes k_elim
      if (O.r > 0 && n <= nreg8 && !ldl) {
        const size_t by = elim_reg_shmem(n, O.k, O.r, sizeof(S), 8, lu);
        if (by <= (size_t)g_tune.elim8_lds) {
          creg[0].push_back((int)ib);
          regbox[ib] = 1;
          shm_reg[0] = std::max(shm_reg[0], by);
          continue;
        }
      }
      if (O.r > 0 && n <= nreg16 && !ldl) {
        const size_t by = elim_reg_shmem(n, O.k, O.r, sizeof(S), 16, lu);
        if (by <= 160 * 1024) {
          creg[1].push_back((int)ib);
          regbox[ib] = 1;
          shm_reg[1] = std::max(shm_reg[1], by);
          continue;
        }
      }
      const size_t cache = 6 * sizeof(double) * (size_t)hb[ib].n;
      size_t by = O.r > 0 ? elim_elems(lu, O.k, O.r) * sizeof(S) + cache : 0;
      int c = (O.r > 0 && by <= kHalf) ? 0 : (O.r > 0 && by <= kFull) ? 1 : 2;
      cls[c].push_back((int)ib);
      if (c < 2) shm_cls[c] = std::max(shm_cls[c], by);   // class 2: all in global scratch
    }
    // global-scratch class: 2 n^2 elements of W, then a 6 n double point
    // cache (12 n elements keep the next box 8-byte aligned)
    if (!cls[2].empty()) {
      int64_t oW = 0;
      for (int ib : cls[2]) {
        hb[ib].W_off = oW;
        oW += 2 * (int64_t)hb[ib].n * hb[ib].n + 12 * (int64_t)hb[ib].n;
      }
      RSKX_TRY(ws.W.ensure(oW * sizeof(S)));
      RSKX_TRY(upload(ws.boxes, hb));
    }
    std::vector<int> allids;
    for (int c = kMfmaShapes - 1; c >= 0; --c) allids.insert(allids.end(), cmf[c].begin(), cmf[c].end());
    allids.insert(allids.end(), creg[1].begin(), creg[1].end());
    allids.insert(allids.end(), creg[0].begin(), creg[0].end());
    for (int c = 2; c >= 0; --c) allids.insert(allids.end(), cls[c].begin(), cls[c].end());
    // the largest boxes in batches of similar sizes (n within 25% of the
    // batch's largest, scratch within 8 GB), padded to the batch's largest
    // r and k
    struct BigBatch { int off, cnt, rM, kM; };
    std::vector<BigBatch> bb;
    size_t big_need = 0;
    if (!cbig.empty()) {
      std::sort(cbig.begin(), cbig.end(), [&](int a, int b) { return hb[a].n > hb[b].n; });
      for (size_t i = 0; i < cbig.size();) {
        BigBatch g{(int)(allids.size() + i), 0, 0, 0};
        const int n0 = hb[cbig[i]].n;
        while (i < cbig.size()) {
          const int ib = cbig[i];
          const int rM = std::max(g.rM, ho[ib].r), kM = std::max(g.kM, ho[ib].k);
          if (g.cnt > 0 && (4 * hb[ib].n < 3 * n0 ||
                            big_elim_scratch(g.cnt + 1, rM, kM, sizeof(S)) > (size_t(8) << 30)))
            break;
          g.rM = rM;
          g.kM = kM;
          ++g.cnt;
          ++i;
        }
        big_need = std::max(big_need, big_elim_scratch(g.cnt, g.rM, g.kM, sizeof(S)));
        bb.push_back(g);
      }
      allids.insert(allids.end(), cbig.begin(), cbig.end());
    }
    RSKX_TRY(upload(ws.ids, allids));
    if (!cbig.empty()) RSKX_TRY(ws.big.ensure(big_need));
    prof.begin("factor", ev);
    {
      // the classes run concurrently: the largest matrix-core shape and the
      // 16-wave boxes on one side stream, the other matrix-core shapes and
      // the 8-wave boxes on g_stream, the shared-memory / global classes of
      // k_elim on the other side stream; joined before anything reads them
      int nmf = 0;
      for (int c = 0; c < kMfmaShapes; ++c) nmf += cmf[c].empty() ? 0 : 1;
      const int nclass = nmf + (creg[1].empty() ? 0 : 1) + (creg[0].empty() ? 0 : 1) +
                         ((cls[0].empty() && cls[1].empty() && cls[2].empty()) ? 0 : 1);
      const bool conc = nclass > 1 && g_tune.concurrent;
      if (conc) RSKX_TRY(side_fork());
      int off = 0;
      for (int c = kMfmaShapes - 1; c >= 0; --c) {
        if (cmf[c].empty()) continue;
        launch_elim_mfma<S>(c, shm_mf[c], ws.ids.as<int>() + off, (int)cmf[c].size(),
                            ws.boxes.as<LvlBox>(), ws.outs.as<BoxOut>(), ws.csegs.as<Seg>(),
                            ws.pool.as<int>(), ws.perm.as<int>(), ws.T.as<double>(),
                            ws.M[mpar].as<S>(), ws.M[mcur].as<S>(), fval, fidx, P,
                            ws.flag.as<int>(),
                            conc && c == kMfmaShapes - 1 ? g_side.s[0] : g_stream);
        off += (int)cmf[c].size();
      }
      for (int c = 1; c >= 0; --c) {           // 16-wave boxes first, then 8-wave
        if (creg[c].empty()) continue;
        launch_elim_reg<S>(lu, c == 0 ? 8 : 16, shm_reg[c], ws.ids.as<int>() + off,
                           (int)creg[c].size(), ws.boxes.as<LvlBox>(), ws.outs.as<BoxOut>(),
                           ws.csegs.as<Seg>(), ws.pool.as<int>(), ws.perm.as<int>(),
                           ws.T.as<double>(), ws.M[mpar].as<S>(), ws.M[mcur].as<S>(), fval, fidx,
                           P, ws.flag.as<int>(), conc && c == 1 ? g_side.s[0] : g_stream);
        off += (int)creg[c].size();
      }
      for (int c = 2; c >= 0; --c) {
        if (cls[c].empty()) continue;
        launch_elim<S>(lu, ldl, c < 2 ? 1 : 0, shm_cls[c], ws.ids.as<int>() + off,
                       (int)cls[c].size(), ws.boxes.as<LvlBox>(), ws.outs.as<BoxOut>(),
                       ws.csegs.as<Seg>(), ws.pool.as<int>(), ws.perm.as<int>(), ws.T.as<double>(),
                       ws.M[mpar].as<S>(), ws.M[mcur].as<S>(), ws.W.as<S>(), fval,
                       fidx, P, ws.flag.as<int>(), conc ? g_side.s[1] : g_stream);
        off += (int)cls[c].size();
      }
      if (conc) RSKX_TRY(side_join());
      // the largest boxes batch after batch, each batch with the whole GPU
      for (const BigBatch& g : bb)
        RSKX_TRY(big_elim<S>(lu, ws.ids.as<int>() + g.off, g.cnt, g.rM, g.kM, ws.boxes.as<LvlBox>(),
                             ws.outs.as<BoxOut>(), ws.csegs.as<Seg>(), ws.pool.as<int>(),
                             ws.perm.as<int>(), ws.T.as<double>(), ws.M[mpar].as<S>(),
                             ws.M[mcur].as<S>(), fval, fidx, P, ws.flag.as<int>(), ws.big.p,
                             g_stream));
    }
    prof.end("factor", ev, 0, el_flops);
    }
    RSKX_HIP(hipGetLastError());
    hlap("h:layout");
    // the failure flag comes back with the next synchronization; the host
    // meanwhile lays out the next level
    RSKX_HIP(hipMemcpyAsync(ws.hflag, ws.flag.p, sizeof(int), hipMemcpyDeviceToHost, g_stream));

    // ---- commit the level
    int64_t elim = 0;
    for (int64_t ib = 0; ib < nb; ++ib) {
      const int b = (int)(b0 + ib);
      const LvlBox& B = hb[ib];
      const BoxOut& O = ho[ib];
      if (O.r > 0) {
        bs[b].cur_off = O.nl_off;
        bs[b].cur_len = O.k;
        FacRec fr;
        fr.k = O.k; fr.r = O.r;
        fr.T = O.T; fr.LU = O.LU; fr.E = O.E; fr.G = O.G; fr.Li = O.Li; fr.Ui = O.Ui;
        fr.sk = O.sk; fr.rd = O.rd; fr.piv = O.piv;
        fr.Tt = O.Tt; fr.Et = O.Et; fr.Gt = O.Gt; fr.LUt = O.LUt; fr.Lti = O.Lti; fr.Uti = O.Uti;
        fr.packed = regbox[ib];
        fr.D = O.D;
        fr.pad_ = 0;
        F->recs.push_back(fr);
        elim += O.r;
      } else {
        bs[b].cur_off = B.slf_off;
        bs[b].cur_len = B.n;
      }
      bs[b].moff = O.M_off;
    }
    // transposed copies for the solve sweeps; runs while the host lays
    // out the next level
    if constexpr (!CPLX) {
      const int64_t f0 = F->lvp.back(), nl = (int64_t)F->recs.size() - f0;
      bool any = false;
      for (int64_t f = f0; f < f0 + nl; ++f) any = any || !F->recs[f].packed;
      if (nl > 0 && any) {
        lrec.assign(F->recs.begin() + f0, F->recs.end());   // kept until the next sync
        RSKX_TRY(upload(ws.lrecs, lrec));
        int64_t bytes = 0;
        for (const FacRec& fr : lrec) {
          if (fr.packed) continue;
          const int64_t kr = (int64_t)fr.k * fr.r, nbk = (fr.r + kDiagBlock - 1) / kDiagBlock;
          bytes += 2 * (int64_t)sizeof(S) * ((lu ? 3 : 2) * kr + (int64_t)fr.r * fr.r +
                                             (lu ? 2 : 1) * nbk * kDiagBlock * kDiagBlock);
        }
        prof.begin("pack", ev);
        launch_pack<S>(lu, ws.lrecs.as<FacRec>(), (int)nl, rmx, fval, g_stream);
        prof.end("pack", ev, bytes);
      }
    }
    F->fval_used += need_v;
    F->fidx_used += need_i;
    pool_top += off_slf + off_nl;
    mpar = mcur;
    nrem -= elim;
    F->lvp.push_back((int64_t)F->recs.size());
    F->lev_kmax.push_back(kmx);
    F->lev_rmax.push_back(rmx);
    const int64_t nblk = pblk[lv] + nb;
    F->table.push_back({lvl_ref, nblk, nrem1, nrem});
    F->nlvl_done++;
    hlap("h:commit");
    if (o.verb)
      printf("%3d | %6lld | %8lld | %8lld | %8.2f | %8.2f | %10.2e\n", lvl_ref, (long long)nblk,
             (long long)nrem1, (long long)nrem, (double)nrem1 / nblk, (double)nrem / nblk,
             std::chrono::duration<double>(clk::now() - tl).count());
  }
  if (o.verb) printf("%s\n", std::string(69, '-').c_str());
  if constexpr (!CPLX) {
    if (nrem > 0) RSKX_TRY(partial_assembly<S>(t, bs, lv_last, ws.M[mpar], F));
  }
  RSKX_TRY(upload(F->d_recs, F->recs));
  RSKX_HIP(hipGetLastError());
  RSKX_TRY(sync_stream());
  RSKX_TRY(elim_status());
  prof.flush();
  return RSKX_OK;
}

// Device bytes of a sweep's workspaces: the work copy of X in tree order
// (N x ncol, padded to 256 bytes) and the largest per-level global slice
// set of the levels whose rows do not fit shared memory (k_solve /
// k_sweep's gws, same sizes as sweep_impl takes them).
template <typename S>
static void sweep_ws_bytes(const rskx_factor* F, int ncol, bool solve, int mode, size_t* ybytes,
                           size_t* gbytes) {
  constexpr int kRhs = 8;
  *ybytes = ((size_t)F->N * ncol * sizeof(S) + 255) & ~(size_t)255;
  *gbytes = 0;
  const int prog = (F->symm == 'p' || F->symm == 'h') ? (solve ? 2 : 5) : (solve ? 0 : 3);
  for (int lv = 0; lv + 1 < (int)F->lvp.size(); ++lv) {
    const int64_t nf = F->lvp[lv + 1] - F->lvp[lv];
    if (!nf || !(mode & 3)) continue;
    const int kmx = F->lev_kmax[lv], rmx = F->lev_rmax[lv];
    size_t g = 0;
    if (solve) {
      const size_t shs = solve_shmem(prog, kmx, rmx, sizeof(S));
      if (shs > 160 * 1024) g = (shs / sizeof(S) + 63) / 64 * 64 * nf * sizeof(S);
    } else {
      const size_t xb = ((size_t)(kmx + 2 * rmx) * kRhs + (size_t)kRhs * 256) * sizeof(S);
      const size_t want = (size_t)std::max(rmx * rmx, kmx * rmx) * sizeof(S);
      const size_t stage = (size_t)g_tune.sweep_stage;
      const size_t room = xb < stage ? stage - xb : 0;
      if (xb + std::min(want, room) > 160 * 1024) g = (xb / sizeof(S) + 63) / 64 * 64 * nf * sizeof(S);
    }
    *gbytes = std::max(*gbytes, g);
  }
}

// profile state of calls that must not create events (no profiling)
static Prof g_noprof;

// ext: when not NULL, the caller's workspace of ext_cap bytes (at least
// what sweep_ws_bytes reports) holds the work copy and the global slices,
// and nothing is allocated and no event is recorded
template <typename S, typename U>
int sweep_impl(const rskx_factor* Fc, void* X, int64_t ldx, int nrhs, char trans, bool solve,
               int mode, void* ext = nullptr, size_t ext_cap = 0) {
  rskx_factor* F = const_cast<rskx_factor*>(Fc);
  const int64_t N = F->N;
  if (nrhs <= 0 || N == 0) return RSKX_OK;
  // work copy in tree order, padded with zero columns to a multiple of 8
  constexpr int kRhs = 8;
  const int ncol = (nrhs + kRhs - 1) / kRhs * kRhs;
  S* Y = nullptr;
  size_t ybytes = 0, gbytes = 0;
  if (ext) {
    sweep_ws_bytes<S>(F, ncol, solve, mode, &ybytes, &gbytes);
    if (ybytes + gbytes > ext_cap)
      return fail(RSKX_ERR_MEM, "FLAM:rskx:scratch: this sweep needs " + std::to_string(ybytes + gbytes) +
                                    " bytes of work space; the allocation-free entry point holds " +
                                    std::to_string(ext_cap));
    Y = reinterpret_cast<S*>(ext);
  } else {
    RSKX_TRY(g_ws.xw.ensure((size_t)N * ncol * sizeof(S)));
    Y = g_ws.xw.as<S>();
  }
  // a global slice set for the rows of one level's factors
  auto slices = [&](size_t bytes, S** out) -> int {
    if (ext) {
      *out = reinterpret_cast<S*>(reinterpret_cast<char*>(ext) + ybytes);
      return RSKX_OK;
    }
    RSKX_TRY(g_ws.sweep.ensure(bytes));
    *out = g_ws.sweep.as<S>();
    return RSKX_OK;
  };
  hipEvent_t ev;
  Prof& prof = ext ? g_noprof : F->prof;
  // factors with r >= rlib are solved by library calls with the whole GPU
  // (not in the allocation-free entry point, nor under graph capture)
  int rlib = INT_MAX;
  if (!ext && solve) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(g_stream, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone)
      rlib = g_tune.lib_solve_r > 0 ? g_tune.lib_solve_r : INT_MAX;
  }
  prof.begin("permute", ev);
  launch_perm<U, S>(true, X, ldx, Y, N, nrhs, ncol, F->order.as<int>(), g_stream);
  prof.end("permute", ev, (int64_t)N * nrhs * sizeof(U) + N * ncol * sizeof(S) + N * 4);
  int prog;
  // 'h' runs the 'p' programs with a unit L; the kernels apply D (FacRec::D)
  if (F->symm == 'p' || F->symm == 'h') prog = solve ? 2 : 5;
  else prog = solve ? (trans == 'n' ? 0 : 1) : (trans == 'n' ? 3 : 4);
  const int nl = (int)F->lvp.size() - 1;
  // rskelf_sv.m:25-35 / sv/*.m: mode 3 = upward then downward sweep
  for (int pass = 0; pass < 2; ++pass) {
    const bool up = pass == 0;
    if (!(mode & (up ? 1 : 2))) continue;     // mode 1: upward only, 2: downward only
    for (int q = 0; q < nl; ++q) {
      const int lv = up ? q : nl - 1 - q;
      const int64_t f0 = F->lvp[lv], f1 = F->lvp[lv + 1];
      if (f1 == f0) continue;
      const int kmx = F->lev_kmax[lv], rmx = F->lev_rmax[lv];
      // xs, xr, tmp, red, then a stage buffer for one factor block: the
      // level's largest block if the total stays under 76 KB (two
      // workgroups per CU), else what is left; larger blocks are read from
      // global memory directly
      if (solve) {
        size_t shs = solve_shmem(prog, kmx, rmx, sizeof(S));
        // rows too many for shared memory: each workgroup works in its own
        // slice of a global workspace instead
        S* gws = nullptr;
        int64_t gstride = 0;
        if (shs > 160 * 1024) {
          gstride = (int64_t)((shs / sizeof(S) + 63) / 64 * 64);
          RSKX_TRY(slices((size_t)gstride * (f1 - f0) * sizeof(S), &gws));
          shs = 0;
        }
        int64_t bytes = 0;
        for (int64_t f = f0; f < f1; ++f) {
          const FacRec& fr = F->recs[f];
          int64_t kr = (int64_t)fr.k * fr.r, rr = (int64_t)fr.r * fr.r;
          bytes += (int64_t)sizeof(S) * (2 * kr + (rr + fr.r) / 2 + 2 * (int64_t)(fr.k + fr.r) * nrhs) +
                   4 * (fr.k + fr.r);
        }
        prof.begin("sweep", ev);
        launch_solve<S>(prog, up, F->d_recs.as<FacRec>() + f0, (int)(f1 - f0), F->lval[lv].as<S>(),
                        F->lidx[lv].as<int>(), Y, N, ncol, kmx, rmx, shs, gws, gstride, rlib,
                        g_stream);
        if (rmx >= rlib)
          for (int64_t f = f0; f < f1; ++f) {
            const FacRec& fr = F->recs[f];
            if (fr.r < rlib || fr.D >= 0) continue;
            RSKX_TRY(g_ws.bsv.ensure(big_solve_scratch(fr.k, fr.r, ncol) * sizeof(S)));
            RSKX_TRY(big_solve<S>(prog, up, fr, F->lval[lv].as<S>(), F->lidx[lv].as<int>(), Y, N,
                                  ncol, g_ws.bsv.as<S>(), g_stream));
          }
        prof.end("sweep", ev, bytes);
        continue;
      }
      const size_t xb = ((size_t)(kmx + 2 * rmx) * kRhs + (size_t)kRhs * 256) * sizeof(S);
      const size_t want = (size_t)std::max(rmx * rmx, kmx * rmx) * sizeof(S);
      const size_t stage = (size_t)g_tune.sweep_stage;
      const size_t room = xb < stage ? stage - xb : 0;
      int bcap = (int)(std::min(want, room) / sizeof(S));
      size_t shm = xb + (size_t)bcap * sizeof(S);
      S* gws = nullptr;
      int64_t gstride = 0;
      if (shm > 160 * 1024) {                 // as the solve: a global workspace slice
        gstride = (int64_t)((xb / sizeof(S) + 63) / 64 * 64);
        RSKX_TRY(slices((size_t)gstride * (f1 - f0) * sizeof(S), &gws));
        bcap = 0;
        shm = 0;
      }
      int64_t bytes = 0;
      for (int64_t f = f0; f < f1; ++f) {
        const FacRec& fr = F->recs[f];
        int64_t kr = (int64_t)fr.k * fr.r, rr = (int64_t)fr.r * fr.r;
        bytes += (int64_t)sizeof(S) * (2 * kr + (rr + fr.r) / 2 + 2 * (int64_t)(fr.k + fr.r) * nrhs) +
                 4 * (fr.k + fr.r);
      }
      prof.begin("sweep", ev);
      launch_sweep<S>(prog, up, F->d_recs.as<FacRec>() + f0, (int)(f1 - f0), F->lval[lv].as<S>(),
                      F->lidx[lv].as<int>(), Y, N, ncol, kmx, rmx, bcap, shm, gws, gstride, g_stream);
      prof.end("sweep", ev, bytes);
    }
  }
  prof.begin("permute", ev);
  launch_perm<U, S>(false, X, ldx, Y, N, nrhs, ncol, F->order.as<int>(), g_stream);
  prof.end("permute", ev, (int64_t)N * nrhs * (sizeof(S) + sizeof(U)) + N * 4);
  RSKX_HIP(hipGetLastError());
  prof.flush();                              // waits only when profiling
  return RSKX_OK;
}

// Complex sweeps (rskelf_sv.m / rskelf_mv.m with sv/*_nn, _nc, _sn, _sc and
// the mv counterparts): trans 't' is 'c' applied to conj(X), conjugated
// back (rskelf_sv.m:22, rskelf_mv.m:22).  Columns go through the level
// kernels in groups whose rows fit shared memory (at most 8).
template <typename S, typename U>
int zsweep_impl(const rskx_factor* Fc, void* X, int64_t ldx, int nrhs, char trans, bool solve,
                int mode) {
  rskx_factor* F = const_cast<rskx_factor*>(Fc);
  const int64_t N = F->N;
  if (nrhs <= 0 || N == 0) return RSKX_OK;
  const bool cj = trans == 't';
  const char tr = cj ? 'c' : trans;
  RSKX_TRY(g_ws.xw.ensure((size_t)N * nrhs * sizeof(S)));
  S* Y = g_ws.xw.as<S>();
  hipEvent_t ev;
  Prof& prof = F->prof;
  prof.begin("permute", ev);
  launch_zperm<U, S>(true, X, ldx, Y, N, nrhs, nrhs, F->order.as<int>(), cj, g_stream);
  prof.end("permute", ev, (int64_t)N * nrhs * (sizeof(S) + sizeof(U)) + N * 4);
  const int prog = solve ? (tr == 'n' ? 0 : 1) : (tr == 'n' ? 2 : 3);
  const bool sym = F->symm == 's';
  const int nl = (int)F->lvp.size() - 1;
  for (int pass = 0; pass < 2; ++pass) {
    const bool up = pass == 0;
    if (!(mode & (up ? 1 : 2))) continue;
    for (int q = 0; q < nl; ++q) {
      const int lv = up ? q : nl - 1 - q;
      const int64_t f0 = F->lvp[lv], f1 = F->lvp[lv + 1];
      if (f1 == f0) continue;
      const int kmx = F->lev_kmax[lv], rmx = F->lev_rmax[lv];
      int nc = std::min(nrhs, 8);
      while (nc > 1 && zsweep_shmem(kmx, rmx, nc) > 160 * 1024) nc = (nc + 1) / 2;
      const size_t shm = zsweep_shmem(kmx, rmx, nc);
      if (shm > 160 * 1024)
        return fail(RSKX_ERR_UNSUP, "FLAM:rskx:size: a complex factor too large for the sweep's shared memory (k=" +
                                        std::to_string(kmx) + ", r=" + std::to_string(rmx) + ")");
      int64_t bytes = 0;
      for (int64_t f = f0; f < f1; ++f) {
        const FacRec& fr = F->recs[f];
        const int64_t kr = (int64_t)fr.k * fr.r, rr = (int64_t)fr.r * fr.r;
        bytes += (int64_t)sizeof(S) * (2 * kr + (rr + fr.r) / 2 + 2 * (int64_t)(fr.k + fr.r) * nrhs) +
                 4 * (fr.k + fr.r);
      }
      prof.begin("sweep", ev);
      for (int c0 = 0; c0 < nrhs; c0 += nc)
        launch_zsweep<S>(prog, sym, up, F->d_recs.as<FacRec>() + f0, (int)(f1 - f0),
                         F->lval[lv].as<S>(), F->lidx[lv].as<int>(), Y, N, c0,
                         std::min(nc, nrhs - c0), shm, g_stream);
      prof.end("sweep", ev, bytes);
    }
  }
  prof.begin("permute", ev);
  launch_zperm<U, S>(false, X, ldx, Y, N, nrhs, nrhs, F->order.as<int>(), cj, g_stream);
  prof.end("permute", ev, (int64_t)N * nrhs * (sizeof(S) + sizeof(U)) + N * 4);
  RSKX_HIP(hipGetLastError());
  prof.flush();
  return RSKX_OK;
}

}  // namespace

template <typename S>
static int logdet_impl(const rskx_factor* F, double* re, double* im) {
  const int mode = F->symm == 'p' ? 0 : F->symm == 'h' ? 1 : 2;
  const int64_t nf = (int64_t)F->recs.size();
  int rmax = 0;
  for (const FacRec& f : F->recs) rmax = std::max(rmax, f.r);
  if (mode == 2 && rmax > 65536)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:size: a factor beyond 65536 redundant points for the permutation sign");
  DevBuf out;
  RSKX_TRY(out.ensure(std::max<int64_t>(nf, 1) * 3 * sizeof(double)));
  for (size_t lv = 0; lv + 1 < F->lvp.size(); ++lv) {
    const int64_t f0 = F->lvp[lv], f1 = F->lvp[lv + 1];
    if (f1 > f0)
      launch_logdet<S>(F->d_recs.as<FacRec>() + f0, (int)(f1 - f0), F->lval[lv].as<S>(),
                       F->lidx[lv].as<int>(), mode, out.as<double>() + 3 * f0, g_stream);
  }
  RSKX_HIP(hipGetLastError());
  std::vector<double> h((size_t)nf * 3);
  if (nf) RSKX_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  RSKX_TRY(sync_stream());
  out.release();
  double lr = 0;
  int64_t nneg = 0;
  for (int64_t f = 0; f < nf; ++f) {
    lr += h[3 * f];
    nneg += (int64_t)h[3 * f + 1] + (int64_t)h[3 * f + 2];
  }
  // 'p': 2 sum log diag(L); each negative pivot or odd permutation adds i pi
  *re = mode == 0 ? 2 * lr : lr;
  *im = (nneg & 1) ? 3.141592653589793238462643383279502884 : 0.0;
  return RSKX_OK;
}

// ----------------------------------------------- allocation-free solve --
// The work rows of rskx_rskelf_sv_noalloc: a file-scope device array that
// every GPU the library runs on holds until the process ends.  Its name is
// a host-side stand-in in host code; the device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_address(void** out) {
  int dev = 0;
  RSKX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:scratch: device index beyond " + std::to_string(kMaxDevices));
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    RSKX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RSKX_OK;
}

// ------------------------------------------------------- captured solve --
// rskelf_sv captured once into a HIP graph for a fixed X and replayed: the
// solve of an FGMRES step is the same launch sequence every time, so the
// host issues one graph launch instead of 2 + 2 nlvl kernel launches.  The
// graph owns its work rows (no allocation, no event, no synchronization
// inside the captured region).
struct rskx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  DevBuf ws;
  ~rskx_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    ws.release();
  }
};

// ================================================================ C ABI ==
extern "C" {

int rskx_opts_default(rskx_opts* o) {
  if (!o) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: opts is NULL");
  o->lvlmax = 0;
  o->ext = nullptr;
  o->Tmax = 2.0;
  o->rrqr_iter = INFINITY;
  o->symm = 'n';
  o->verb = 0;
  o->precision = RSKX_PREC_FP64;
  o->profile = 0;
  o->stop = -1;
  o->id_gram = 0;
  o->rank_model = nullptr;
  o->rank_scale = 1.0;
  o->record_boxes = 0;
  return RSKX_OK;
}

const char* rskx_last_error(void) { return g_err.c_str(); }

int rskx_set_stream(void* s) {
  // the asynchronous calls leave work queued on the old stream that still
  // uses the library's workspaces: drain it before switching
  if ((hipStream_t)s != g_stream) RSKX_TRY(sync_stream());
  g_stream = (hipStream_t)s;
  return RSKX_OK;
}

static int* tuning_slot(const char* name) {
  if (!name) return nullptr;
  const std::string s(name);
  if (s == "elim8_nmax") return &g_tune.elim8_nmax;
  if (s == "elim8_lds") return &g_tune.elim8_lds;
  if (s == "gram_n0") return &g_tune.gram_n0;
  if (s == "gram_n1") return &g_tune.gram_n1;
  if (s == "concurrent") return &g_tune.concurrent;
  if (s == "sweep_stage") return &g_tune.sweep_stage;
  if (s == "id_lds") return &g_tune.id_lds;
  if (s == "id_reg") return &g_tune.id_reg;
  if (s == "elim_mfma") return &g_tune.elim_mfma;
  if (s == "gram_mfma") return &g_tune.gram_mfma;
  if (s == "big_n") return &g_tune.big_n;
  if (s == "lib_solve_r") return &g_tune.lib_solve_r;
  return nullptr;
}

int rskx_set_tuning(const char* name, double value) {
  int* slot = tuning_slot(name);
  if (!slot) return fail(RSKX_ERR_ARG, std::string("FLAM:rskx:tuning: unknown parameter ") + (name ? name : "NULL"));
  if (!(value >= 0) || value > 2147483647.0) return fail(RSKX_ERR_ARG, "FLAM:rskx:tuning: value out of range");
  const std::string s(name);
  if ((s == "elim8_lds" || s == "sweep_stage" || s == "id_lds") && (value < 4096 || value > 160 * 1024))
    return fail(RSKX_ERR_ARG, "FLAM:rskx:tuning: shared-memory budgets lie in [4096, 163840] bytes");
  RSKX_TRY(sync_stream());                 // queued work keeps the values it was launched with
  *slot = (int)value;
  set_gram_mfma(g_tune.gram_mfma != 0);
  return RSKX_OK;
}

int rskx_get_tuning(const char* name, double* value) {
  int* slot = tuning_slot(name);
  if (!slot || !value) return fail(RSKX_ERR_ARG, "FLAM:rskx:tuning: unknown parameter or NULL value");
  *value = *slot;
  return RSKX_OK;
}

int rskx_release_cache(void) {
  RSKX_TRY(sync_stream());
  Workspace& ws = g_ws;
  for (DevBuf* b : {&ws.boxes, &ws.nsegs, &ws.csegs, &ws.outs, &ws.ids, &ws.pool, &ws.nbr,
                    &ws.mnbr, &ws.mrows, &ws.kout, &ws.iter, &ws.K, &ws.perm, &ws.T, &ws.W,
                    &ws.M[0], &ws.M[1], &ws.flag, &ws.V, &ws.lrecs, &ws.sweep, &ws.cgs, &ws.pts,
                    &ws.xw, &ws.txs, &ws.tord, &ws.tpart, &ws.tchunks, &ws.tcounts, &ws.xs,
                    &ws.poff, &ws.Ablk, &ws.big, &ws.wids, &ws.wv, &ws.bsv})
    b->release();
  block_trim_all();
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

// the built-in matrices and their dimensions
static int check_kernel(const rskx_kernel* A, int d) {
  if (A && A->kind == RSKX_KERNEL_LAPLACE2D_DL && d == 2) {
    if (!A->nu) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: the double layer needs normals");
    return RSKX_OK;
  }
  if (A && A->kind == RSKX_KERNEL_LAPLACE3D_SL && d == 3) return RSKX_OK;
  if (A && A->kind == RSKX_KERNEL_CALLBACK && (d == 2 || d == 3)) {
    if (!A->entries) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: the callback kernel needs its entries function");
    return RSKX_OK;
  }
  if (A && A->kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE && d == 2) {
    if (!A->nu) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: the combined field needs normals");
    if (!(A->wavenumber > 0) || !std::isfinite(A->eta))
      return fail(RSKX_ERR_ARG, "FLAM:rskx:kernel: the combined field needs a wavenumber > 0 and a finite eta");
    return RSKX_OK;
  }
  return fail(RSKX_ERR_UNSUP, "FLAM:rskx:kernel: built kernels are the 2D Laplace double layer (d = 2), "
                              "the 3D Laplace single layer (d = 3), the 2D Helmholtz combined field (d = 2) "
                              "and host callbacks (d = 2, 3)");
}

// argument checks shared by rskx_rskelf and rskx_rskelf_batch
// (rskelf.m:112-128, core/chksymm.m, core/id.m:52-57, core/hypoct.m:58-61)
static int check_factor_args(const double* x, int d, int64_t N, int occ, double rank_or_tol,
                             const rskx_kernel* A, const rskx_opts* opts, rskx_opts& o) {
  rskx_opts_default(&o);
  if (opts) o = *opts;
  char s = (char)tolower(o.symm);
  if (s != 'n' && s != 's' && s != 'h' && s != 'p')
    return fail(RSKX_ERR_ARG, "FLAM:chksymm:invalidSymm: Symmetry parameter must be one of 'N', 'S', 'H', or 'P'.");
  o.symm = s;
  if (occ < 0) return fail(RSKX_ERR_ARG, "FLAM:hypoct:invalidOcc: Leaf occupancy must be nonnegative.");
  if (o.lvlmax < 0) return fail(RSKX_ERR_ARG, "FLAM:hypoct:invalidLvlmax: Maximum tree depth must be at least 1.");
  if (!(rank_or_tol >= 0)) return fail(RSKX_ERR_ARG, "FLAM:id:invalidRankOrTol: Rank or tolerance must be nonnegative.");
  if (!(o.Tmax >= 1)) return fail(RSKX_ERR_ARG, "FLAM:id:invalidTmax: Interpolation matrix entry bound must be >= 1.");
  if (!(o.rrqr_iter >= 0)) return fail(RSKX_ERR_ARG, "FLAM:id:invalidRRQRIter: Maximum number of RRQR iterations must be nonnegative.");
  if (rank_or_tol >= 1 && std::isfinite(o.Tmax) && o.rrqr_iter > 0)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:rank: fixed-rank ID with RRQR refinement is not built (pass a tolerance < 1, or Tmax = Inf)");
  RSKX_TRY(check_kernel(A, d));
  if (A->kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE) {
    if (o.symm != 'n' && o.symm != 's')
      return fail(RSKX_ERR_UNSUP, "FLAM:rskx:symm: complex factors are built for symm 'n' and 's' "
                                  "('h' and 'p' need a Hermitian matrix)");
    if (o.stop >= 0)
      return fail(RSKX_ERR_UNSUP, "FLAM:rskx:stop: partial factorization is built for real data");
  }
  if (!x || (A->kind != RSKX_KERNEL_CALLBACK && (!A->w || !A->diag)))
    return fail(RSKX_ERR_ARG, "FLAM:rskx:null: point or kernel array is NULL");
  if (N < 1) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: N must be positive");
  if (N > (int64_t)1 << 30) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:size: N above 2^30");
  if (o.precision != RSKX_PREC_FP64 && o.precision != RSKX_PREC_FP32)
    return fail(RSKX_ERR_ARG, "FLAM:rskx:precision: unknown precision");
  return RSKX_OK;
}

static int check_proxy(const rskx_proxy* pxy, int d) {
  if (pxy && pxy->kind == RSKX_PROXY_CALLBACK) {
    if (!pxy->fn) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: the callback proxy needs its function");
    if (pxy->p < 0) return fail(RSKX_ERR_ARG, "FLAM:rskx:proxy: p (the most proxy rows) must be >= 0");
    return RSKX_OK;
  }
  if (!pxy || (pxy->kind != RSKX_PROXY_RING_SL && pxy->kind != RSKX_PROXY_RING_DL &&
               pxy->kind != RSKX_PROXY_POINTS_SL && pxy->kind != RSKX_PROXY_RING_CFIE))
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:proxy: naive global compression (no pxyfun) is not built; pass a proxy surface");
  if (pxy->p < 1 || !(pxy->radius > 0)) return fail(RSKX_ERR_ARG, "FLAM:rskx:proxy: invalid proxy surface");
  if (pxy->kind == RSKX_PROXY_POINTS_SL) {
    if (d != 3) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:proxy: point proxies are built for d = 3");
    if (!pxy->pts) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: proxy points are NULL");
  } else if (d != 2) {
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:proxy: proxy rings are built for d = 2");
  }
  return RSKX_OK;
}

static int run_factor(const double* x, int d, int64_t N, int nprob, const int64_t* poff, int occ,
                      double rank_or_tol, const rskx_kernel* A, const rskx_proxy& px,
                      const double* pw, const rskx_opts& o, rskx_factor** out) {
  rskx_factor* F = new rskx_factor();
  F->N = N;
  F->d = d;
  F->nprob = nprob;
  F->symm = o.symm;
  F->prec = o.precision;
  const bool cfie = A->kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE;
  if (cfie != (px.kind == RSKX_PROXY_RING_CFIE)) {
    delete F;
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:proxy: the combined field takes RSKX_PROXY_RING_CFIE, "
                                "and that ring only the combined field");
  }
  const bool cb = A->kind == RSKX_KERNEL_CALLBACK;
  if (cb != (px.kind == RSKX_PROXY_CALLBACK)) {
    delete F;
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:proxy: a callback matrix takes RSKX_PROXY_CALLBACK, "
                                "and that proxy only a callback matrix");
  }
  if (cb && nprob > 1) {
    delete F;
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:batch: callback matrices are factored one problem at a time");
  }
  F->cplx = cfie;
  int rc;
  if (cfie)
    rc = o.precision == RSKX_PREC_FP64
             ? factor_impl<zd>(x, d, N, nprob, poff, occ, rank_or_tol, *A, px, pw, o, F)
             : factor_impl<zf>(x, d, N, nprob, poff, occ, rank_or_tol, *A, px, pw, o, F);
  else
    rc = o.precision == RSKX_PREC_FP64
             ? factor_impl<double>(x, d, N, nprob, poff, occ, rank_or_tol, *A, px, pw, o, F)
             : factor_impl<float>(x, d, N, nprob, poff, occ, rank_or_tol, *A, px, pw, o, F);
  if (rc != RSKX_OK) { delete F; return rc; }
  *out = F;
  return RSKX_OK;
}

int rskx_rskelf(const double* x, int d, int64_t N, int occ, double rank_or_tol,
                const rskx_kernel* A, const rskx_proxy* pxy, const rskx_opts* opts,
                rskx_factor** out) {
  if (!out) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: output handle is NULL");
  *out = nullptr;
  rskx_opts o;
  RSKX_TRY(check_factor_args(x, d, N, occ, rank_or_tol, A, opts, o));
  RSKX_TRY(check_proxy(pxy, d));
  const int64_t poff[2] = {0, N};
  return run_factor(x, d, N, 1, poff, occ, rank_or_tol, A, *pxy, &pxy->weight, o, out);
}

int rskx_rskelf_batch(int nprob, const int64_t* offsets, const double* x, int d, int occ,
                      double rank_or_tol, const rskx_kernel* A, const rskx_proxy* pxy,
                      const rskx_opts* opts, rskx_factor** out) {
  if (!out) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: output handle is NULL");
  *out = nullptr;
  if (nprob < 1 || !offsets || !pxy) return fail(RSKX_ERR_ARG, "FLAM:rskx:batch: need nprob >= 1, offsets and proxies");
  if (nprob > 65536) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:batch: more than 65536 problems");
  if (offsets[0] != 0) return fail(RSKX_ERR_ARG, "FLAM:rskx:batch: offsets[0] must be 0");
  for (int j = 0; j < nprob; ++j)
    if (offsets[j + 1] <= offsets[j]) return fail(RSKX_ERR_ARG, "FLAM:rskx:batch: every problem needs at least one point");
  const int64_t N = offsets[nprob];
  rskx_opts o;
  RSKX_TRY(check_factor_args(x, d, N, occ, rank_or_tol, A, opts, o));
  if (o.ext) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:batch: opts.ext (one root extent) is not taken by a batch");
  std::vector<double> pw(nprob);
  for (int j = 0; j < nprob; ++j) {
    RSKX_TRY(check_proxy(&pxy[j], d));
    if (pxy[j].kind != pxy[0].kind || pxy[j].p != pxy[0].p || pxy[j].radius != pxy[0].radius ||
        pxy[j].pts != pxy[0].pts)
      return fail(RSKX_ERR_UNSUP, "FLAM:rskx:batch: the problems' proxy surfaces must share kind, p, radius and points");
    pw[j] = pxy[j].weight;
  }
  return run_factor(x, d, N, nprob, offsets, occ, rank_or_tol, A, pxy[0], pw.data(), o, out);
}

static int sweep_entry(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                       bool solve, int mode = 3) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  if (mode < 1 || mode > 3) return fail(RSKX_ERR_ARG, "FLAM:rskx:mode: sweep mode must be 1, 2 or 3");
  char t = (char)tolower(trans);
  if (t != 'n' && t != 't' && t != 'c')
    return fail(RSKX_ERR_ARG, "FLAM:chktrans:invalidTrans: Transpose parameter must be one of 'N', 'T', or 'C'.");
  if (nrhs < 0 || ldx < F->N) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad nrhs or leading dimension");
  if (nrhs > 0 && !X) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: X is NULL");
  if (F->cplx) {
    if (dtype != RSKX_Z64 && dtype != RSKX_Z32)
      return fail(RSKX_ERR_ARG, "FLAM:rskx:dtype: a complex factor takes complex right-hand sides (RSKX_Z64, RSKX_Z32)");
    if (F->prec == RSKX_PREC_FP64)
      return dtype == RSKX_Z64 ? zsweep_impl<zd, zd>(F, X, ldx, nrhs, t, solve, mode)
                               : zsweep_impl<zd, zf>(F, X, ldx, nrhs, t, solve, mode);
    return dtype == RSKX_Z64 ? zsweep_impl<zf, zd>(F, X, ldx, nrhs, t, solve, mode)
                             : zsweep_impl<zf, zf>(F, X, ldx, nrhs, t, solve, mode);
  }
  if (t == 't') t = 'c';   // real data: transpose = conjugate transpose (rskelf_sv.m:22)
  if (dtype != RSKX_F64 && dtype != RSKX_F32) return fail(RSKX_ERR_ARG, "FLAM:rskx:dtype: unknown element type");
  if (F->prec == RSKX_PREC_FP64)
    return dtype == RSKX_F64 ? sweep_impl<double, double>(F, X, ldx, nrhs, t, solve, mode)
                             : sweep_impl<double, float>(F, X, ldx, nrhs, t, solve, mode);
  return dtype == RSKX_F64 ? sweep_impl<float, double>(F, X, ldx, nrhs, t, solve, mode)
                           : sweep_impl<float, float>(F, X, ldx, nrhs, t, solve, mode);
}

int rskx_sweep_sv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  int mode) {
  return sweep_entry(F, X, ldx, nrhs, trans, dtype, true, mode);
}

int rskx_sweep_mv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  int mode) {
  return sweep_entry(F, X, ldx, nrhs, trans, dtype, false, mode);
}

int rskx_partial_info(const rskx_factor* F, int64_t* nsk, int64_t* nnz, int64_t* sk, int64_t* I,
                      int64_t* J, double* V) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  if (nsk) *nsk = (int64_t)F->Si.size();
  if (nnz) *nnz = (int64_t)F->SV.size();
  if (sk) std::copy(F->Si.begin(), F->Si.end(), sk);
  if (I) std::copy(F->SI.begin(), F->SI.end(), I);
  if (J) std::copy(F->SJ.begin(), F->SJ.end(), J);
  if (V) std::copy(F->SV.begin(), F->SV.end(), V);
  return RSKX_OK;
}

int rskx_rskelf_logdet(const rskx_factor* F, double* re, double* im) {
  if (!F || !re || !im) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  if (F->cplx) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:logdet: built for real factors");
  return F->prec == RSKX_PREC_FP64 ? logdet_impl<double>(F, re, im) : logdet_impl<float>(F, re, im);
}

// the generalized Cholesky factor C of F = C C' is half a sweep
// (rskelf_cholmv.m, rskelf_cholsv.m): C X = downward apply, C' X = upward
// apply, C \ X = upward solve, C' \ X = downward solve
static int chol_entry(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                      bool solve) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  if (F->symm != 'p')
    return fail(RSKX_ERR_ARG, std::string("FLAM:rskelf_chol") + (solve ? "sv" : "mv") +
                                  ":invalidSymm: Symmetry parameter must be 'P'.");
  const char t = (char)tolower(trans);
  if (t != 'n' && t != 't' && t != 'c')
    return fail(RSKX_ERR_ARG, "FLAM:chktrans:invalidTrans: Transpose parameter must be one of 'N', 'T', or 'C'.");
  const bool fwd = t == 'n';
  return sweep_entry(F, X, ldx, nrhs, 'n', dtype, solve, solve == fwd ? 1 : 2);
}

int rskx_rskelf_cholmv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype) {
  return chol_entry(F, X, ldx, nrhs, trans, dtype, false);
}

int rskx_rskelf_cholsv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype) {
  return chol_entry(F, X, ldx, nrhs, trans, dtype, true);
}

int rskx_rskelf_sv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype) {
  return sweep_entry(F, X, ldx, nrhs, trans, dtype, true);
}

int rskx_rskelf_mv(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype) {
  return sweep_entry(F, X, ldx, nrhs, trans, dtype, false);
}

int rskx_sv_graph(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans, int dtype,
                  rskx_graph** out) {
  if (!F || !out) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  *out = nullptr;
  char t = (char)tolower(trans);
  if (t != 'n' && t != 't' && t != 'c')
    return fail(RSKX_ERR_ARG, "FLAM:chktrans:invalidTrans: Transpose parameter must be one of 'N', 'T', or 'C'.");
  if (t == 't') t = 'c';
  if (nrhs < 1 || ldx < F->N || !X) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad X, nrhs or leading dimension");
  if (F->cplx) return fail(RSKX_ERR_UNSUP, "FLAM:rskx:graph: built for real factors");
  if (dtype != RSKX_F64 && dtype != RSKX_F32) return fail(RSKX_ERR_ARG, "FLAM:rskx:dtype: unknown element type");
  const int ncol = (nrhs + 7) / 8 * 8;
  size_t yb = 0, gb = 0;
  if (F->prec == RSKX_PREC_FP64) sweep_ws_bytes<double>(F, ncol, true, 3, &yb, &gb);
  else sweep_ws_bytes<float>(F, ncol, true, 3, &yb, &gb);
  rskx_graph* G = new rskx_graph();
  auto run = [&](void* Xp) -> int {
    if (F->prec == RSKX_PREC_FP64)
      return dtype == RSKX_F64 ? sweep_impl<double, double>(F, Xp, ldx, nrhs, t, true, 3, G->ws.p, yb + gb)
                               : sweep_impl<double, float>(F, Xp, ldx, nrhs, t, true, 3, G->ws.p, yb + gb);
    return dtype == RSKX_F64 ? sweep_impl<float, double>(F, Xp, ldx, nrhs, t, true, 3, G->ws.p, yb + gb)
                             : sweep_impl<float, float>(F, Xp, ldx, nrhs, t, true, 3, G->ws.p, yb + gb);
  };
  int rc = G->ws.ensure(std::max<size_t>(yb + gb, 256));
  // one uncaptured run on a zeroed copy of X's shape (X untouched) sets the
  // kernels' shared-memory attributes before capture
  if (rc == RSKX_OK) {
    DevBuf tmp;
    const size_t xb = (size_t)ldx * nrhs * (dtype == RSKX_F64 ? 8 : 4);
    rc = tmp.ensure(xb);
    if (rc == RSKX_OK && hipMemsetAsync(tmp.p, 0, xb, g_stream) != hipSuccess)
      rc = fail(RSKX_ERR_HIP, "FLAM:rskx:hip: hipMemsetAsync failed");
    if (rc == RSKX_OK) rc = run(tmp.p);
    if (rc == RSKX_OK) rc = sync_stream();
    tmp.release();
  }
  if (rc == RSKX_OK && hipStreamCreateWithFlags(&G->cap, hipStreamNonBlocking) != hipSuccess)
    rc = fail(RSKX_ERR_HIP, "FLAM:rskx:hip: capture stream");
  if (rc == RSKX_OK) {
    const hipStream_t keep = g_stream;
    g_stream = G->cap;
    if (hipStreamBeginCapture(G->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = fail(RSKX_ERR_HIP, "FLAM:rskx:hip: hipStreamBeginCapture failed");
    } else {
      const int rr = run(X);
      const hipError_t e = hipStreamEndCapture(G->cap, &G->graph);
      if (rr != RSKX_OK) rc = rr;
      else if (e != hipSuccess) rc = fail(RSKX_ERR_HIP, "FLAM:rskx:hip: hipStreamEndCapture failed");
    }
    g_stream = keep;
  }
  if (rc == RSKX_OK && hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(RSKX_ERR_HIP, "FLAM:rskx:hip: hipGraphInstantiate failed");
  if (rc != RSKX_OK) {
    (void)hipGetLastError();
    delete G;
    return rc;
  }
  *out = G;
  return RSKX_OK;
}

int rskx_graph_launch(rskx_graph* G) {
  if (!G) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: graph is NULL");
  RSKX_HIP(hipGraphLaunch(G->exec, g_stream));
  return RSKX_OK;
}

void rskx_graph_free(rskx_graph* G) {
  if (G) (void)hipStreamSynchronize(g_stream);
  delete G;
}

int rskx_rskelf_sv_noalloc(const rskx_factor* F, void* X, int64_t ldx, int nrhs, char trans,
                           int dtype) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  char t = (char)tolower(trans);
  if (t != 'n' && t != 't' && t != 'c')
    return fail(RSKX_ERR_ARG, "FLAM:chktrans:invalidTrans: Transpose parameter must be one of 'N', 'T', or 'C'.");
  if (t == 't') t = 'c';
  if (nrhs < 0 || ldx < F->N) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad nrhs or leading dimension");
  if (nrhs > 0 && !X) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: X is NULL");
  if (dtype != RSKX_F64 && dtype != RSKX_F32) return fail(RSKX_ERR_ARG, "FLAM:rskx:dtype: unknown element type");
  if (F->cplx)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:dtype: the allocation-free solve is built for real factors");
  void* ws = nullptr;
  RSKX_TRY(scratch_address(&ws));
  if (F->prec == RSKX_PREC_FP64)
    return dtype == RSKX_F64 ? sweep_impl<double, double>(F, X, ldx, nrhs, t, true, 3, ws, kScratchBytes)
                             : sweep_impl<double, float>(F, X, ldx, nrhs, t, true, 3, ws, kScratchBytes);
  return dtype == RSKX_F64 ? sweep_impl<float, double>(F, X, ldx, nrhs, t, true, 3, ws, kScratchBytes)
                           : sweep_impl<float, float>(F, X, ldx, nrhs, t, true, 3, ws, kScratchBytes);
}

void rskx_free(rskx_factor* F) { delete F; }

int rskx_factor_info(const rskx_factor* F, rskx_info* info) {
  if (!F || !info) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  info->N = F->N;
  info->d = F->d;
  info->nlvl = F->nlvl;
  info->nlvl_done = F->nlvl_done;
  info->nfactor = (int64_t)F->recs.size();
  int es = (F->prec == RSKX_PREC_FP64 ? 8 : 4) * (F->cplx ? 2 : 1);
  info->bytes = F->fval_used * es + F->fidx_used * 4 + (int64_t)F->recs.size() * sizeof(FacRec) + F->N * 4;
  info->precision = F->prec;
  info->symm = F->symm;
  info->id_scratch = F->id_scratch;
  info->rank_fallbacks = F->rank_fallbacks;
  return RSKX_OK;
}

int rskx_factor_lvp(const rskx_factor* F, int64_t* lvp, int maxn) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  const int n = (int)F->lvp.size();
  for (int i = 0; i < n && i < maxn && lvp; ++i) lvp[i] = F->lvp[i];
  return n;
}

int rskx_level_table(const rskx_factor* F, int64_t* rows, int maxrows) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  int n = (int)F->table.size();
  for (int i = 0; i < n && i < maxrows && rows; ++i)
    for (int j = 0; j < 4; ++j) rows[4 * i + j] = F->table[i][j];
  return n;
}

int rskx_export_factor(const rskx_factor* F, int64_t i, rskx_factor_view* v) {
  if (!F || !v) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  if (i < 0 || i >= (int64_t)F->recs.size()) return fail(RSKX_ERR_ARG, "FLAM:rskx:index: factor index out of range");
  const FacRec& f = F->recs[i];
  v->k = f.k;
  v->r = f.r;
  if (!v->sk) return RSKX_OK;
  const int flv = F->level_of(i);
  const int64_t k = f.k, r = f.r;
  const bool lu = F->symm == 'n' || F->symm == 's';
  if ((int64_t)F->h_order.size() != F->N) {
    std::vector<int> o32(F->N);
    RSKX_HIP(hipMemcpy(o32.data(), F->order.p, F->N * sizeof(int), hipMemcpyDeviceToHost));
    F->h_order.assign(o32.begin(), o32.end());
  }
  if (F->cplx) {
    // complex factor: every matrix of the view holds interleaved (re, im)
    // pairs, twice the element counts of the real view
    const bool d64 = F->prec == RSKX_PREC_FP64;
    auto getz = [&](int64_t off, int64_t cnt, std::vector<double>& dst) -> int {
      dst.assign(2 * cnt, 0.0);
      if (!cnt) return RSKX_OK;
      if (d64) {
        RSKX_HIP(hipMemcpy(dst.data(), F->lval[flv].as<char>() + off * 16, cnt * 16, hipMemcpyDeviceToHost));
      } else {
        std::vector<float> raw(2 * cnt);
        RSKX_HIP(hipMemcpy(raw.data(), F->lval[flv].as<char>() + off * 8, cnt * 8, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < 2 * cnt; ++q) dst[q] = raw[q];
      }
      return RSKX_OK;
    };
    std::vector<double> T, LU, E, G;
    std::vector<int> sk(k), rd(r), pv(r);
    RSKX_TRY(getz(f.T, k * r, T));
    RSKX_TRY(getz(f.LU, r * r, LU));
    RSKX_TRY(getz(f.E, k * r, E));
    RSKX_TRY(getz(f.G, r * k, G));
    if (k) RSKX_HIP(hipMemcpy(sk.data(), F->lidx[flv].as<int>() + f.sk, k * 4, hipMemcpyDeviceToHost));
    if (r) RSKX_HIP(hipMemcpy(rd.data(), F->lidx[flv].as<int>() + f.rd, r * 4, hipMemcpyDeviceToHost));
    if (r) RSKX_HIP(hipMemcpy(pv.data(), F->lidx[flv].as<int>() + f.piv, r * 4, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < k; ++q) v->sk[q] = F->h_order[sk[q]];
    for (int64_t q = 0; q < r; ++q) v->rd[q] = F->h_order[rd[q]];
    if (v->T) std::copy(T.begin(), T.end(), v->T);
    if (v->E) std::copy(E.begin(), E.end(), v->E);
    if (v->F) std::copy(G.begin(), G.end(), v->F);
    for (int64_t c = 0; c < r; ++c)
      for (int64_t a = 0; a < r; ++a) {
        const int64_t e = 2 * (a + c * r);
        if (v->L) {
          v->L[e] = a > c ? LU[e] : (a == c ? 1.0 : 0.0);
          v->L[e + 1] = a > c ? LU[e + 1] : 0.0;
        }
        if (v->U) {
          v->U[e] = a <= c ? LU[e] : 0.0;
          v->U[e + 1] = a <= c ? LU[e + 1] : 0.0;
        }
      }
    if (v->p) for (int64_t q = 0; q < r; ++q) v->p[q] = pv[q];
    return RSKX_OK;
  }
  const int es = F->prec == RSKX_PREC_FP64 ? 8 : 4;
  auto getv = [&](int64_t off, int64_t cnt, std::vector<double>& dst) -> int {
    dst.resize(cnt);
    if (!cnt) return RSKX_OK;
    std::vector<char> raw(cnt * es);
    RSKX_HIP(hipMemcpy(raw.data(), F->lval[flv].as<char>() + off * es, cnt * es, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < cnt; ++q)
      dst[q] = es == 8 ? reinterpret_cast<double*>(raw.data())[q] : reinterpret_cast<float*>(raw.data())[q];
    return RSKX_OK;
  };
  auto geti = [&](int64_t off, int64_t cnt, std::vector<int>& dst) -> int {
    dst.resize(cnt);
    if (cnt) RSKX_HIP(hipMemcpy(dst.data(), F->lidx[flv].as<int>() + off, cnt * 4, hipMemcpyDeviceToHost));
    return RSKX_OK;
  };
  std::vector<double> T, LU, E, G, D;
  std::vector<int> sk, rd, pv;
  RSKX_TRY(getv(f.T, k * r, T));
  RSKX_TRY(getv(f.LU, r * r, LU));
  RSKX_TRY(getv(f.E, k * r, E));
  const bool ldl = F->symm == 'h';
  if (ldl) RSKX_TRY(getv(f.D, 2 * r, D));
  if (lu) RSKX_TRY(getv(f.G, r * k, G));
  RSKX_TRY(geti(f.sk, k, sk));
  RSKX_TRY(geti(f.rd, r, rd));
  if (lu) RSKX_TRY(geti(f.piv, r, pv));
  if ((int64_t)F->h_order.size() != F->N) {
    std::vector<int> o32(F->N);
    RSKX_HIP(hipMemcpy(o32.data(), F->order.p, F->N * sizeof(int), hipMemcpyDeviceToHost));
    F->h_order.assign(o32.begin(), o32.end());
  }
  for (int64_t q = 0; q < k; ++q) v->sk[q] = F->h_order[sk[q]];
  for (int64_t q = 0; q < r; ++q) v->rd[q] = F->h_order[rd[q]];
  if (v->T) std::copy(T.begin(), T.end(), v->T);
  if (v->E) std::copy(E.begin(), E.end(), v->E);
  if (v->L)
    for (int64_t c = 0; c < r; ++c)
      for (int64_t a = 0; a < r; ++a)
        v->L[a + c * r] = a > c ? LU[a + c * r] : (a == c ? ((lu || ldl) ? 1.0 : LU[a + c * r]) : 0.0);
  if (ldl && v->U)                            // U = D (rskelf.m:231: U = sparse(D)), 2 x 2 blocks
    for (int64_t c = 0; c < r; ++c)
      for (int64_t a = 0; a < r; ++a)
        v->U[a + c * r] = a == c ? D[a] : a == c + 1 ? D[r + c] : c == a + 1 ? D[r + a] : 0.0;
  if (lu) {
    if (v->U)
      for (int64_t c = 0; c < r; ++c)
        for (int64_t a = 0; a < r; ++a) v->U[a + c * r] = a <= c ? LU[a + c * r] : 0.0;
    if (v->F) std::copy(G.begin(), G.end(), v->F);
    if (v->p) for (int64_t q = 0; q < r; ++q) v->p[q] = pv[q];
  }
  return RSKX_OK;
}

// Import of a factor in the reference's layout (rskelf.m:156-158,249-257,263):
// the factors' values go to one value / index block per level, laid out as
// the device factorization lays them out; the inverted 32 x 32 diagonal
// blocks of L and U are formed on the host and the transposed copies by
// k_pack, so the sweeps run on an imported factor unchanged.  The tree
// order is the identity (points stay in the caller's order).
extern "C++" {
template <typename S>
static int import_impl(int64_t N, char symm, int nlvl, const int64_t* lvp,
                       const rskx_factor_view* fv, rskx_factor* F) {
  const bool lu = symm == 'n' || symm == 's', ldl = symm == 'h';
  const int64_t nf = lvp[nlvl];
  {
    std::vector<int> iota(N);
    for (int64_t i = 0; i < N; ++i) iota[i] = (int)i;
    RSKX_TRY(F->order.ensure(std::max<int64_t>(N, 1) * sizeof(int)));
    RSKX_HIP(hipMemcpy(F->order.p, iota.data(), N * sizeof(int), hipMemcpyHostToDevice));
  }
  int64_t nrem = N;
  F->lvp.assign(1, 0);
  for (int lv = 0; lv < nlvl; ++lv) {
    const int64_t f0 = lvp[lv], f1 = lvp[lv + 1];
    std::vector<S> val;
    std::vector<int> idx;
    int kmx = 0, rmx = 0;
    const int64_t nrem1 = nrem;
    for (int64_t f = f0; f < f1; ++f) {
      const rskx_factor_view& v = fv[f];
      const int64_t k = v.k, r = v.r, nbk = (r + kDiagBlock - 1) / kDiagBlock;
      FacRec fr{};
      fr.k = (int)k; fr.r = (int)r; fr.packed = 0; fr.D = -1;
      auto put = [&](const double* a, int64_t cnt) {
        const int64_t o = (int64_t)val.size();
        for (int64_t q = 0; q < cnt; ++q) val.push_back(a ? (S)a[q] : (S)0);
        return o;
      };
      fr.T = put(v.T, k * r);
      // packed LU: L strictly below the diagonal, U on and above ('n');
      // L with its diagonal ('p'); unit L ('h', D apart)
      std::vector<double> LU((size_t)r * r, 0.0);
      for (int64_t c = 0; c < r; ++c)
        for (int64_t a = 0; a < r; ++a) {
          const double l = v.L ? v.L[a + c * r] : 0.0, u = v.U ? v.U[a + c * r] : 0.0;
          LU[a + c * r] = lu ? (a > c ? l : u) : (a > c || (!ldl && a == c) ? l : (a == c ? 1.0 : 0.0));
        }
      fr.LU = put(LU.data(), r * r);
      fr.E = put(v.E, k * r);
      fr.G = lu ? put(v.F, r * k) : -1;
      // inverted diagonal blocks: unit lower for 'n' and 'h'
      auto inv_blocks = [&](bool upper) {
        std::vector<double> out((size_t)nbk * kDiagBlock * kDiagBlock, 0.0);
        for (int64_t b = 0; b < nbk; ++b) {
          const int64_t j0 = b * kDiagBlock, bs = std::min<int64_t>(kDiagBlock, r - j0);
          double* X = &out[(size_t)b * kDiagBlock * kDiagBlock];
          for (int64_t c = 0; c < bs; ++c) {
            if (!upper) {
              for (int64_t t = 0; t < bs; ++t) {
                double x = t == c ? 1.0 : 0.0;
                for (int64_t u = 0; u < t; ++u) x -= LU[(j0 + t) + (j0 + u) * r] * X[u + c * kDiagBlock];
                X[t + c * kDiagBlock] = (lu || ldl) ? x : x / LU[(j0 + t) + (j0 + t) * r];
              }
            } else {
              for (int64_t t = bs - 1; t >= 0; --t) {
                double x = t == c ? 1.0 : 0.0;
                for (int64_t u = t + 1; u < bs; ++u) x -= LU[(j0 + t) + (j0 + u) * r] * X[u + c * kDiagBlock];
                X[t + c * kDiagBlock] = x / LU[(j0 + t) + (j0 + t) * r];
              }
            }
          }
        }
        return put(out.data(), (int64_t)out.size());
      };
      fr.Li = inv_blocks(false);
      fr.Ui = lu ? inv_blocks(true) : -1;
      fr.Tt = put(nullptr, k * r);
      fr.Et = put(nullptr, k * r);
      fr.Gt = lu ? put(nullptr, r * k) : -1;
      fr.LUt = put(nullptr, r * r);
      fr.Lti = put(nullptr, nbk * kDiagBlock * kDiagBlock);
      fr.Uti = lu ? put(nullptr, nbk * kDiagBlock * kDiagBlock) : -1;
      if (ldl) {                       // U = D (rskelf.m:231): diagonal, then subdiagonal
        std::vector<double> D(2 * r, 0.0);
        for (int64_t a = 0; a < r; ++a) D[a] = v.U[a + a * r];
        for (int64_t a = 0; a + 1 < r; ++a) D[r + a] = v.U[(a + 1) + a * r];
        fr.D = put(D.data(), 2 * r);
      }
      fr.sk = (int64_t)idx.size();
      for (int64_t q = 0; q < k; ++q) idx.push_back((int)v.sk[q]);
      fr.rd = (int64_t)idx.size();
      for (int64_t q = 0; q < r; ++q) idx.push_back((int)v.rd[q]);
      fr.piv = (int64_t)idx.size();
      if (lu) for (int64_t q = 0; q < r; ++q) idx.push_back((int)v.p[q]);
      F->recs.push_back(fr);
      kmx = std::max<int>(kmx, (int)k);
      rmx = std::max<int>(rmx, (int)r);
      nrem -= r;
    }
    F->lval.emplace_back();
    F->lidx.emplace_back();
    RSKX_TRY(F->lval.back().ensure((val.size() + 1) * sizeof(S)));
    RSKX_TRY(F->lidx.back().ensure((idx.size() + 1) * sizeof(int)));
    if (!val.empty())
      RSKX_HIP(hipMemcpy(F->lval.back().p, val.data(), val.size() * sizeof(S), hipMemcpyHostToDevice));
    if (!idx.empty())
      RSKX_HIP(hipMemcpy(F->lidx.back().p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
    if (f1 > f0) {
      DevBuf recs;
      RSKX_TRY(recs.ensure((f1 - f0) * sizeof(FacRec)));
      RSKX_HIP(hipMemcpy(recs.p, F->recs.data() + f0, (f1 - f0) * sizeof(FacRec), hipMemcpyHostToDevice));
      int rmx = 0;
      for (int64_t f = f0; f < f1; ++f) rmx = std::max(rmx, F->recs[f].r);
      launch_pack<S>(lu, recs.as<FacRec>(), (int)(f1 - f0), rmx, F->lval.back().as<S>(), g_stream);
      RSKX_HIP(hipGetLastError());
      RSKX_TRY(sync_stream());
      recs.release();
    }
    F->fval_used += (int64_t)val.size();
    F->fidx_used += (int64_t)idx.size();
    F->lvp.push_back(f1);
    F->lev_kmax.push_back(kmx);
    F->lev_rmax.push_back(rmx);
    F->table.push_back({nlvl - lv, f1 - f0, nrem1, nrem});
    F->nlvl_done++;
  }
  for (FacRec& fr : F->recs) fr.packed = 1;
  RSKX_TRY(upload(F->d_recs, F->recs));
  RSKX_TRY(sync_stream());
  (void)nf;
  return RSKX_OK;
}
}  // extern "C++"

int rskx_import(int64_t N, char symm, int precision, int nlvl, const int64_t* lvp,
                const rskx_factor_view* factors, rskx_factor** out) {
  if (!out || !lvp || (!factors && nlvl > 0 && lvp[nlvl] > 0))
    return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  *out = nullptr;
  const char sy = (char)tolower(symm);
  if (sy != 'n' && sy != 's' && sy != 'h' && sy != 'p')
    return fail(RSKX_ERR_ARG, "FLAM:chksymm:invalidSymm: Symmetry parameter must be one of 'N', 'S', 'H', or 'P'.");
  if (precision != RSKX_PREC_FP64 && precision != RSKX_PREC_FP32)
    return fail(RSKX_ERR_ARG, "FLAM:rskx:precision: unknown precision");
  if (N < 1 || nlvl < 0 || lvp[0] != 0) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad N or lvp");
  for (int l = 0; l < nlvl; ++l)
    if (lvp[l + 1] < lvp[l]) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: lvp must be nondecreasing");
  for (int64_t f = 0; f < lvp[nlvl]; ++f) {
    const rskx_factor_view& v = factors[f];
    if (v.k < 0 || v.r < 1 || (v.k && !v.sk) || !v.rd || !v.L || (v.k && (!v.T || !v.E)))
      return fail(RSKX_ERR_ARG, "FLAM:rskx:import: factor " + std::to_string(f) + " is incomplete");
    if ((sy == 'n' || sy == 's') && (!v.U || !v.p || (v.k && !v.F)))
      return fail(RSKX_ERR_ARG, "FLAM:rskx:import: factor " + std::to_string(f) + " lacks U, p or F");
    if (sy == 'h') {                 // D: symmetric, 1 x 1 and 2 x 2 diagonal blocks
      if (!v.U) return fail(RSKX_ERR_ARG, "FLAM:rskx:import: symm 'h' needs D in U");
      const int64_t r = v.r;
      for (int64_t a = 0; a < r; ++a)
        for (int64_t c = 0; c < r; ++c) {
          const double u = v.U[a + c * r];
          const bool sub = a == c + 1, sup = c == a + 1;
          if (u != 0 && a != c && !sub && !sup)
            return fail(RSKX_ERR_ARG, "FLAM:rskx:import: D (U) is not block diagonal with 1 x 1 and 2 x 2 blocks");
          if (sub && (u != v.U[c + a * r] || (u != 0 && a + 1 < r && v.U[(a + 1) + a * r] != 0)))
            return fail(RSKX_ERR_ARG, "FLAM:rskx:import: D (U) is not block diagonal with 1 x 1 and 2 x 2 blocks");
        }
    }
    for (int64_t q = 0; q < v.k; ++q)
      if (v.sk[q] < 0 || v.sk[q] >= N) return fail(RSKX_ERR_ARG, "FLAM:rskx:import: sk out of range");
    for (int64_t q = 0; q < v.r; ++q)
      if (v.rd[q] < 0 || v.rd[q] >= N || ((sy == 'n' || sy == 's') && (v.p[q] < 0 || v.p[q] >= v.r)))
        return fail(RSKX_ERR_ARG, "FLAM:rskx:import: rd or p out of range");
  }
  rskx_factor* F = new rskx_factor();
  F->N = N;
  F->d = 0;
  F->symm = sy;
  F->prec = precision;
  F->nlvl = nlvl;
  const int rc = precision == RSKX_PREC_FP64 ? import_impl<double>(N, sy, nlvl, lvp, factors, F)
                                             : import_impl<float>(N, sy, nlvl, lvp, factors, F);
  if (rc != RSKX_OK) { delete F; return rc; }
  *out = F;
  return RSKX_OK;
}

int rskx_kernel_times(const rskx_factor* F, const char** names, double* ms, int64_t* launches,
                      int64_t* bytes, double* flops, int maxn) {
  if (!F) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: factor is NULL");
  int i = 0;
  for (auto& kv : F->prof.st) {
    if (i < maxn) {
      if (names) names[i] = kv.first.c_str();
      if (ms) ms[i] = kv.second.ms;
      if (launches) launches[i] = kv.second.launches;
      if (bytes) bytes[i] = kv.second.bytes;
      if (flops) flops[i] = kv.second.flops;
    }
    ++i;
  }
  return i;
}

int rskx_kernel_matvec(const double* x, int d, int64_t N, const rskx_kernel* A, const double* X,
                       int64_t ldx, double* Y, int64_t ldy, int nrhs) {
  RSKX_TRY(check_kernel(A, d));
  if (A->kind == RSKX_KERNEL_CALLBACK)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:kernel: no device product of a host-callback matrix");
  if (!x || !X || !Y) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL array");
  if (ldx < N || ldy < N || nrhs < 0) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad leading dimension");
  if (A->kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE)   // X, Y interleaved complex, ld in complex elements
    launch_zmatvec(x, N, A->nu, A->w, A->diag, A->wavenumber, A->eta, X, ldx, Y, ldy, nrhs, g_stream);
  else
    launch_matvec(A->kind, x, N, A->nu, A->w, A->diag, A->scale, X, ldx, Y, ldy, nrhs, nullptr, 1,
                  g_stream);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

int rskx_rank_model_nparams(void) { return mlp_nparams(); }

int rskx_rank_model_create(uint64_t seed, rskx_rank_model** out) {
  if (!out) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: output handle is NULL");
  *out = nullptr;
  rskx_rank_model* M = new rskx_rank_model();
  const int np = mlp_nparams();
  // Glorot-uniform weights (seeded), zero biases, in the parameter order
  // W1 (32 x 8), b1, W2 (32 x 32), b2, w3 (32), b3
  std::vector<double> p(np, 0.0);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto unif = [&]() {                       // splitmix64 -> [0, 1)
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const int F = RSKX_RANK_FEATURES, H = 32;
  auto fill = [&](int off, int cnt, int fin, int fout) {
    const double a = std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < cnt; ++i) p[off + i] = a * (2.0 * unif() - 1.0);
  };
  fill(0, H * F, F, H);
  fill(H * F + H, H * H, H, H);
  fill(H * F + 2 * H + H * H, H, H, 1);
  int rc = M->params.ensure(np * sizeof(double));
  if (rc == RSKX_OK) rc = M->m.ensure(np * sizeof(double));
  if (rc == RSKX_OK) rc = M->v.ensure(np * sizeof(double));
  if (rc == RSKX_OK) rc = M->g.ensure(np * sizeof(double));
  if (rc == RSKX_OK) rc = M->loss.ensure(sizeof(double));
  if (rc != RSKX_OK) { delete M; return rc; }
  std::vector<double> z(np, 0.0);
  if (hipMemcpy(M->params.p, p.data(), np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M->m.p, z.data(), np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M->v.p, z.data(), np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    delete M;
    return fail(RSKX_ERR_HIP, "FLAM:rskx:hip: rank model upload failed");
  }
  *out = M;
  return RSKX_OK;
}

void rskx_rank_model_free(rskx_rank_model* M) {
  if (M) (void)hipStreamSynchronize(g_stream);
  delete M;
}

int rskx_rank_model_params(rskx_rank_model* M, double* out, const double* in) {
  if (!M) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: model is NULL");
  const size_t bytes = mlp_nparams() * sizeof(double);
  RSKX_TRY(sync_stream());
  if (out) RSKX_HIP(hipMemcpy(out, M->params.p, bytes, hipMemcpyDeviceToHost));
  if (in) {
    RSKX_HIP(hipMemcpy(M->params.p, in, bytes, hipMemcpyHostToDevice));
    RSKX_HIP(hipMemset(M->m.p, 0, bytes));
    RSKX_HIP(hipMemset(M->v.p, 0, bytes));
    M->step = 0;
  }
  return RSKX_OK;
}

static int model_grad(rskx_rank_model* M, const double* X, const double* t, int64_t n, double* g) {
  const int nblk = mlp_grad_blocks(n);
  RSKX_TRY(M->part.ensure((size_t)nblk * mlp_nparams() * sizeof(double)));
  RSKX_TRY(M->lpart.ensure((size_t)nblk * sizeof(double)));
  launch_mlp_grad(M->params.as<double>(), X, t, n, M->part.as<double>(), M->lpart.as<double>(), g,
                  M->loss.as<double>(), g_stream);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

int rskx_rank_model_grad(rskx_rank_model* M, const double* X, const double* t, int64_t n, double* g,
                         double* loss) {
  if (!M || !X || !t || !g) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  if (n < 1) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: no samples");
  RSKX_TRY(model_grad(M, X, t, n, g));
  if (loss) {
    RSKX_TRY(download(loss, M->loss.p, sizeof(double)));
    RSKX_TRY(sync_stream());
  }
  return RSKX_OK;
}

int rskx_rank_model_train(rskx_rank_model* M, const double* X, const double* t, int64_t n,
                          int epochs, int64_t batch, double lr, double* loss) {
  if (!M || !X || !t) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  if (n < 1 || epochs < 0 || batch < 1 || !(lr > 0))
    return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad training parameters");
  for (int e = 0; e < epochs; ++e)
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min(batch, n - b0);
      RSKX_TRY(model_grad(M, X + b0 * RSKX_RANK_FEATURES, t + b0, nb, M->g.as<double>()));
      launch_mlp_adam(M->params.as<double>(), M->g.as<double>(), M->m.as<double>(),
                      M->v.as<double>(), lr, 0.9, 0.999, 1e-8, ++M->step, g_stream);
    }
  RSKX_HIP(hipGetLastError());
  if (loss) {
    RSKX_TRY(model_grad(M, X, t, n, M->g.as<double>()));
    RSKX_TRY(download(loss, M->loss.p, sizeof(double)));
  }
  RSKX_TRY(sync_stream());
  return RSKX_OK;
}

int rskx_rank_model_predict(const rskx_rank_model* M, const double* X, int64_t n, double* y) {
  if (!M || !X || !y) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  launch_mlp_forward(M->params.as<double>(), X, n, y, g_stream);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

int rskx_box_features(const rskx_factor* F, int64_t* nbox, double* X, int64_t* rank) {
  if (!F || !nbox) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL argument");
  int64_t c = 0;
  for (size_t b = 0; b < F->boxrank.size(); ++b) {
    if (F->boxrank[b] < 0) continue;
    if (X) std::copy(&F->boxfeat[8 * b], &F->boxfeat[8 * b] + 8, X + 8 * c);
    if (rank) rank[c] = F->boxrank[b];
    ++c;
  }
  *nbox = c;
  return RSKX_OK;
}

int rskx_cgs2(const double* V, int64_t ldv, int64_t ldr, int nv, double* w, int64_t ldw,
              int64_t N, int nrhs, double* hout) {
  if (!V || !w || !hout) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL array");
  if (nv < 1 || nv > cgs_max_vectors())
    return fail(RSKX_ERR_ARG, "FLAM:rskx:size: between 1 and " + std::to_string(cgs_max_vectors()) +
                                  " basis vectors per call");
  if (N < 1 || nrhs < 1 || ldv < N || ldw < N || ldr < (int64_t)nv * ldv)
    return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad sizes or leading dimensions");
  const int nblk = cgs_blocks(N);
  RSKX_TRY(g_ws.cgs.ensure(((size_t)nrhs * nblk * (nv + 1) + (size_t)nrhs * nv) * sizeof(double)));
  double* partial = g_ws.cgs.as<double>();
  double* normpart = partial + (size_t)nrhs * nblk * nv;
  launch_cgs(V, ldv, ldr, nv, w, ldw, N, nrhs, partial, normpart, normpart + (size_t)nrhs * nblk,
             hout, g_stream);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

int rskx_kernel_matvec_batch(int nprob, const int64_t* offsets, const double* x, int d,
                             const rskx_kernel* A, const double* X, int64_t ldx, double* Y,
                             int64_t ldy, int nrhs) {
  RSKX_TRY(check_kernel(A, d));
  if (A->kind == RSKX_KERNEL_HELMHOLTZ2D_CFIE || A->kind == RSKX_KERNEL_CALLBACK)
    return fail(RSKX_ERR_UNSUP, "FLAM:rskx:kernel: the batched product is built for the real built-in kernels");
  if (nprob < 1 || !offsets || offsets[0] != 0)
    return fail(RSKX_ERR_ARG, "FLAM:rskx:batch: need nprob >= 1 and offsets starting at 0");
  for (int j = 0; j < nprob; ++j)
    if (offsets[j + 1] <= offsets[j]) return fail(RSKX_ERR_ARG, "FLAM:rskx:batch: every problem needs at least one point");
  const int64_t N = offsets[nprob];
  if (!x || !X || !Y) return fail(RSKX_ERR_ARG, "FLAM:rskx:null: NULL array");
  if (ldx < N || ldy < N || nrhs < 0) return fail(RSKX_ERR_ARG, "FLAM:rskx:size: bad leading dimension");
  static DevBuf doff;                      // offsets on the device (kept for later calls)
  const std::vector<int64_t> hoff(offsets, offsets + nprob + 1);
  RSKX_TRY(upload(doff, hoff));
  launch_matvec(A->kind, x, N, A->nu, A->w, A->diag, A->scale, X, ldx, Y, ldy, nrhs,
                doff.as<int64_t>(), nprob, g_stream);
  RSKX_HIP(hipGetLastError());
  return RSKX_OK;
}

}  // extern "C"
