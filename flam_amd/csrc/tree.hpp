// Adaptive 2^d-tree over the points (host side).
//
// Follows core/hypoct.m:51-171: long sides (>= max side / sqrt 2) are halved
// each level, a box holding more than occ points is split, a point goes to
// the upper half of a divided dimension iff x > ctr strictly, and each box
// lists as neighbours its siblings, adjacent coarser leaves and adjacent
// children of its parent's neighbours.
//
// Points are kept in one index array that is stably partitioned as boxes
// split, so every leaf owns a contiguous range of it, in ascending original
// index order (the order hypoct's xi(idx == i) keeps).  That array is the
// "tree order" the device works in.
#pragma once
#include <cstdint>
#include <vector>

namespace rskx {

// A forest of such trees, one per independent problem, is stored the same
// way: level lv holds the level-lv boxes of every tree (trees aligned at
// their roots), and each box carries its own size, since every tree halves
// the sides of its own root box.
struct HostTree {
  int d = 0;
  int nlvl = 0;
  int nprob = 1;                  // trees (independent problems)
  std::vector<int64_t> lvp;       // box offsets per level, nlvl+1 entries
  std::vector<double> bsize;      // d x nbox box sizes
  std::vector<int> prob;          // tree (problem) of each box
  std::vector<double> ctr;        // d x nbox centres
  std::vector<int> prnt;          // parent box, -1 for the root
  std::vector<int> chld_off, chld;  // CSR children
  // while building: a split box's children are consecutive boxes
  // [cfirst, cfirst + ccnt) (ccnt 0: not split)
  std::vector<int> cfirst, ccnt;
  std::vector<int> nbor_off, nbor;  // CSR neighbours
  std::vector<int64_t> beg, cnt;  // leaf point range in `order` (cnt 0: not a leaf)
  std::vector<int64_t> order;     // tree order -> original point index

  int nbox() const { return (int)prnt.size(); }
  int level_of(int b) const;      // 0-based level (root = 0)
  bool leaf(int b) const { return chld_off[b + 1] == chld_off[b]; }
};

// Neighbour lists of the boxes of level lv (hypoct.m:141-170), appended to
// t.nbor / t.nbor_off; needs the lists of every earlier level, the final
// split status (cnt) of the levels above lv and the child ranges.
void append_level_neighbours(HostTree& t, int lv);

// Children in CSR form and the neighbour lists of the levels not yet
// appended, given the boxes (ctr, prnt, beg/cnt, lvp, bsize, nlvl) and the
// child ranges (cfirst, ccnt).
void finish_tree(HostTree& t);

}  // namespace rskx
