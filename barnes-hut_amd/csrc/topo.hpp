// Flag-and-scan topology of the adaptive Morton octree: every node of every
// level is derived from the sorted codes in one pass over the points.
//
// A node at level l >= 1 is a maximal run of sorted codes sharing their top
// l octal digits whose parent (level l - 1) holds more than leaf_size points
// and lies above the depth limit.  Point r of its set's range [sb, se)
// shares d(r) digits with its predecessor (d = 0 for the first point of a
// set), so it begins the cells of levels d + 1 .. D that contain it.  It
// begins NODES at levels d + 1 .. top(r) when its level-d cell is split, and
// none otherwise, with
//   W(r)   = digits shared by codes r and r + leaf_size (-1 past the set end):
//            the run from r holds more than leaf_size points at level l iff
//            W(r) >= l;
//   top(r) = min(D, max(d + 1, W(r) + 1)): the node at top(r) is a leaf and
//            the ones above it are split.
// Whether the level-d cell (begun before r) is split takes at most two
// binary searches of log2(leaf_size) steps around r.  The node index of
// (l, r) is level_start[l] + rank_l(r), rank_l(r) = the number of points
// before r that begin a level-l node, so nodes are numbered level by level
// in (set, Morton) order: the breadth-first order of the per-level builder.
//
// Query groups (the walk's units) are cut from the group cells: the maximal
// nodes holding at most kGroupCell points, and leaves holding more, each cut
// into consecutive runs of kGroupMax points.  The group cell that begins at
// r is one of the nodes r begins: the shallowest level whose node holds at
// most kGroupCell points, or the leaf at top(r) if every node r begins holds
// more (an over-full leaf at the depth limit, or a cell whose parent was
// already small: then r begins no group cell).  Groups are numbered in
// position order.
//
// Functions here are __host__ __device__ so that the same code can be run on
// the CPU against the breadth-first builder of the oracle.
#pragma once

#include <cstdint>

#include "tree.hpp"

namespace sbh {
namespace tree {
namespace topo {

constexpr int kLevels = 16;  // per-tile counter columns: levels 1..10, column 0 = groups

__host__ __device__ inline int clz32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __clz(static_cast<int>(x));
#else
  return x == 0 ? 32 : __builtin_clz(x);
#endif
}

// Leading octal digits two D-digit codes share (0..D).
__host__ __device__ inline int common_digits(uint32_t a, uint32_t b, int D) {
  const uint32_t x = a ^ b;
  if (x == 0) return D;
  const int h = 31 - clz32(x);
  return D - 1 - h / 3;
}

// First position e in [lo, hi) with common_digits(codes[r], codes[e]) < l
// (hi if none); the predicate is monotone over the sorted codes.
__host__ __device__ inline int64_t run_end(const uint32_t* codes, int64_t r, int64_t lo,
                                           int64_t hi, int l, int D) {
  const uint32_t cr = codes[r];
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (common_digits(cr, codes[mid], D) >= l)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First position k in [lo, r] with common_digits(codes[k], codes[r]) >= l.
__host__ __device__ inline int64_t run_begin(const uint32_t* codes, int64_t r, int64_t lo,
                                             int l, int D) {
  const uint32_t cr = codes[r];
  int64_t hi = r;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (common_digits(codes[mid], cr, D) >= l)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Whether the level-l cell holding r (which began before r when r > sb)
// holds more than `cap` points: the points of the cell before r (at most
// cap) plus those from r (at most cap + 1) reach cap + 1.
__host__ __device__ inline bool cell_exceeds(const uint32_t* codes, int64_t r, int64_t sb,
                                             int64_t se, int l, int D, int64_t cap) {
  // quick answer from two probes: with at most h points before r and at
  // most cap - h from r, the cell holds at most cap
  const int64_t h = cap / 2;
  const bool few_before = r - h - 1 < sb || common_digits(codes[r - h - 1], codes[r], D) < l;
  const bool few_after = r + cap - h >= se || common_digits(codes[r], codes[r + cap - h], D) < l;
  if (few_before && few_after) return false;
  const int64_t lo = r - cap > sb ? r - cap : sb;
  const int64_t before = r - run_begin(codes, r, lo, l, D);
  const int64_t hi = r + cap + 1 < se ? r + cap + 1 : se;
  const int64_t from = run_end(codes, r, r + 1, hi, l, D) - r;
  return before + from >= cap + 1;
}

struct PointNodes {
  int d;      // digits shared with the predecessor (0 at a set start)
  int w;      // W(r)
  int first;  // first level begun (d + 1), or 0 if r begins no node
  int top;    // deepest level begun (a leaf); first..top are r's nodes
};

__host__ __device__ inline PointNodes point_nodes(const uint32_t* codes, int64_t r, int64_t sb,
                                                  int64_t se, int D, int leaf_size) {
  PointNodes p;
  p.d = r == sb ? 0 : common_digits(codes[r - 1], codes[r], D);
  p.w = r + leaf_size < se ? common_digits(codes[r], codes[r + leaf_size], D) : -1;
  p.first = 0;
  p.top = -1;
  if (p.d >= D) return p;
  // the level-d cell holding r is split iff it holds more than leaf_size
  // points (d < D here); the window from r decides it when it fits inside
  const bool split = p.w >= p.d || cell_exceeds(codes, r, sb, se, p.d, D, leaf_size);
  if (!split) return p;
  p.first = p.d + 1;
  const int t = p.w + 1 > p.first ? p.w + 1 : p.first;
  p.top = t < D ? t : D;
  return p;
}

// Level of the group cell whose first point is r, or -1.  The cells with r
// as first point are r's nodes first..top, and the set's root (level 0)
// when r == sb.  The group cell is the shallowest of them holding at most
// kGroupCell points whose parent holds more (a root has no parent), or the
// deepest (a leaf) when even that holds more.  r's node at `first` has as
// parent the level-d cell, begun before r unless r == sb (then the root).
__host__ __device__ inline int group_cell_level(const uint32_t* codes, int64_t r, int64_t sb,
                                                int64_t se, int D, const PointNodes& p) {
  const int64_t g = kGroupCell;
  // r's cell at level l (one of its nodes, or the root) holds more than g
  // points iff the run from r does: iff l <= wg
  const int wg = r + g < se ? common_digits(codes[r], codes[r + g], D) : -1;
  if (r == sb) {
    if (wg < 0 || p.first == 0) return 0;  // root within g points, or an over-full leaf
  } else if (p.first == 0) {
    return -1;  // r begins no node
  }
  int l = wg + 1 > p.first ? wg + 1 : p.first;
  if (l > p.top) return p.top;  // every node of r holds more than g: over-full leaf
  if (l > p.first || r == sb) return l;  // parent: r's node at l - 1, or the root
  // l == first: the parent is the level-d cell begun before r
  return cell_exceeds(codes, r, sb, se, p.d, D, g) ? l : -1;
}

}  // namespace topo
}  // namespace tree
}  // namespace sbh
