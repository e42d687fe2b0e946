// Host tree build; see tree.hpp.  Reference: core/hypoct.m:51-171.
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace rskx {

int HostTree::level_of(int b) const {
  // lvp is ascending; the level of b is the last lv with lvp[lv] <= b
  auto it = std::upper_bound(lvp.begin(), lvp.end(), (int64_t)b);
  return (int)(it - lvp.begin()) - 1;
}

static double matlab_round(double v) {
  // MATLAB round: halves away from zero
  return v < 0 ? -std::floor(-v + 0.5) : std::floor(v + 0.5);
}

// A small persistent pool of host threads for the per-level neighbour
// search: parallel_for(n, parts, f) runs f(part, lo, hi) on `parts`
// contiguous ranges covering [0, n) and returns when all have finished.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool pool;
    return pool;
  }
  int size() const { return (int)workers_.size() + 1; }
  template <typename F>
  void parallel_for(int64_t n, int parts, F&& f) {
    parts = std::max(1, std::min(parts, size()));
    if (parts == 1 || n < 2) { f(0, (int64_t)0, n); return; }
    std::function<void(int)> job = [&](int w) { f(w, n * w / parts, n * (w + 1) / parts); };
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &job;
      nparts_ = parts;
      next_ = 1;
      left_ = parts - 1;
      ++gen_;
    }
    cv_.notify_all();
    job(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return left_ == 0; });
    job_ = nullptr;
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }

 private:
  HostPool() {
    const int n = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency())) - 1;
    for (int i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::function<void(int)>* job;
      int w;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || (gen_ != seen && next_ < nparts_); });
        if (stop_) return;
        w = next_++;
        if (next_ >= nparts_) seen = gen_;
        job = job_;
      }
      (*job)(w);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::function<void(int)>* job_ = nullptr;
  int nparts_ = 0, next_ = 0, left_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

void append_level_neighbours(HostTree& t, int lv) {
  // neighbours of the boxes of level lv (hypoct.m:141-170): siblings,
  // adjacent non-split boxes among the parent's neighbours, and adjacent
  // children of the parent's neighbours.  Boxes are numbered level by
  // level and a box's list depends only on its parent's, so lists are
  // appended to the CSR arrays in box order as levels are completed; the
  // boxes of a large level are searched in parallel ranges.
  const int d = t.d;
  if (t.nbor_off.empty()) t.nbor_off.push_back(0);
  const int64_t b0 = t.lvp[lv], b1 = t.lvp[lv + 1], nb = b1 - b0;
  if ((int64_t)t.nbor_off.size() != b0 + 1) return;      // out of order: nothing to do
  if (lv == 0) {
    for (int64_t i = b0; i < b1; ++i) t.nbor_off.push_back((int)t.nbor.size());
    return;
  }
  auto box_list = [&](int64_t i, std::vector<int>& out) {
    const double* c = &t.ctr[(size_t)d * i];
    const double* lc = &t.bsize[(size_t)d * i];
    const int p = t.prnt[i];
    for (int j = t.cfirst[p]; j < t.cfirst[p] + t.ccnt[p]; ++j)
      if (j != i) out.push_back(j);
    const int pn0 = t.nbor_off[p], pn1 = t.nbor_off[p + 1];
    for (int e = pn0; e < pn1; ++e) {
      const int j = t.nbor[e];
      if (t.cnt[j] == 0) continue;            // only non-empty (leaf) boxes
      const double* jl = &t.bsize[(size_t)d * j];
      const double* jc = &t.ctr[(size_t)d * j];
      double dmax = -INFINITY;
      for (int q = 0; q < d; ++q)
        dmax = std::max(dmax, matlab_round((std::fabs(c[q] - jc[q]) - 0.5 * (lc[q] + jl[q])) / lc[q]));
      if (dmax <= 0) out.push_back(j);
    }
    for (int e = pn0; e < pn1; ++e) {
      const int j = t.nbor[e];
      for (int q2 = t.cfirst[j]; q2 < t.cfirst[j] + t.ccnt[j]; ++q2) {
        const double* qc = &t.ctr[(size_t)d * q2];
        double dmax = -INFINITY;
        for (int u = 0; u < d; ++u) dmax = std::max(dmax, matlab_round(std::fabs(c[u] - qc[u]) / lc[u]));
        if (dmax <= 1) out.push_back(q2);
      }
    }
  };
  HostPool& pool = HostPool::get();
  const int parts = (int)std::max<int64_t>(1, std::min<int64_t>(pool.size(), nb / 256));
  std::vector<std::vector<int>> lists(parts), counts(parts);
  pool.parallel_for(nb, parts, [&](int w, int64_t lo, int64_t hi) {
    std::vector<int>& out = lists[w];
    std::vector<int>& cnt = counts[w];
    out.reserve((size_t)(hi - lo) * (t.d == 3 ? 32 : 8));
    cnt.reserve((size_t)(hi - lo));
    for (int64_t q = lo; q < hi; ++q) {
      const size_t before = out.size();
      box_list(b0 + q, out);
      cnt.push_back((int)(out.size() - before));
    }
  });
  for (int w = 0; w < parts; ++w) {
    int base = (int)t.nbor.size();
    for (int c : counts[w]) {
      base += c;
      t.nbor_off.push_back(base);
    }
    t.nbor.insert(t.nbor.end(), lists[w].begin(), lists[w].end());
  }
}

void finish_tree(HostTree& t) {
  const int nbox = t.nbox();
  t.chld_off.assign(nbox + 1, 0);
  for (int b = 0; b < nbox; ++b) t.chld_off[b + 1] = t.chld_off[b] + t.ccnt[b];
  t.chld.resize(t.chld_off[nbox]);
  for (int b = 0; b < nbox; ++b)
    for (int c = 0; c < t.ccnt[b]; ++c) t.chld[t.chld_off[b] + c] = t.cfirst[b] + c;
  for (int lv = 0; lv < t.nlvl; ++lv) append_level_neighbours(t, lv);
}

}  // namespace rskx
