// The reference's CPU trees Tree2D<1, 3> (sbh.Quadtree3, include/tree2d.h:
// 237-407) and Tree3D<1, 3> (sbh.Octree3, include/tree3d.h:238-413), built on
// the GPU with the same byte output.  python/jax_sbh_ffi.py:658-742 builds
// them on the host for the 2-D level-3 targets and for
// build_octree_buffers(backend="cpu").
//
// The reference builds breadth first: a node with more than one point and
// depth < 100 partitions its points into the 8^DIM cells of its box
// (AABB::quantize, cells visited in (z,) y, x order), and every non-empty
// cell becomes a child with the box min + k (max - min) / 8; children are
// reserved consecutively, so node ids are the breadth-first order with the
// children of a node contiguous in cell order.  Moments are aggregated
// bottom-up in child order.
//
// Here every level is one round of device work over the points that are
// still in internal nodes ("active" points, kept in position order):
//   1. key = (node - first node of the level) << cell_bits | cell,
//   2. a stable radix sort of (key, point),
//   3. runs of equal keys are the children (ids in key order = the
//      reference's breadth-first order); each point moves to its sorted slot
//      inside the union of the active ranges, which keeps every child
//      contiguous,
//   4. parents collect their children's cells into the bit mask,
//   5. the points of children that are not leaves stay active.
// One 8-byte read-back per level (child and active counts) sizes the next
// level.  The result equals the reference's leaf order wherever a leaf holds
// one point; inside a leaf at depth 100 (points that share a cell for 100
// levels) the order here is the input order, while std::partition leaves it
// unspecified.
#include <climits>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace sbh {
namespace {

constexpr int kB = 256;
constexpr int kMaxDepthNd = 100;  // PyQuadtree / PyOctree (include/py_tree.h:41, 69)

template <int DIM, int LOG>
struct NdT {
  using Node = NodeT<DIM, LOG>;
  using Leaf = LeafT<DIM>;
  static constexpr int kSplit = 1 << LOG;        // cells per axis
  static constexpr int kCellBits = DIM * LOG;
  static constexpr int kWords = Node::kWords;
};

// Node arrays of the tree under construction (structure of arrays; the
// record layout is written only by sbh_tree_emit).
struct NodeArrays {
  uint32_t* start;
  uint32_t* end;
  uint32_t* parent;
  uint32_t* first;  // first child id
  uint32_t* cell;   // cell index inside the parent
  int32_t* num;     // number of children
  int32_t* depth;
  uint32_t* leaf;   // is_leaf
  float* lo;        // DIM per node
  float* hi;
  uint64_t* mask;   // kWords per node
  float* cm;        // contrib: mass, position, normal, alpha
  float* cpos;      // DIM per node
  float* cnrm;
  float* calpha;
};

// AABB::quantize (include/aabb.h:36-42) as the reference's x86 build
// computes it: (int)(S (p - lo) / (hi - lo)), where the conversion of a NaN
// or out-of-range value gives INT_MIN, then min(., S - 1), max(., 0).
template <int S>
__device__ __forceinline__ int quantize_host(float p, float lo, float hi) {
#pragma clang fp contract(off)
  const float t = static_cast<float>(S) * (p - lo) / (hi - lo);
  int c = (t >= -2147483648.0f && t < 2147483648.0f) ? static_cast<int>(t) : INT_MIN;
  c = min(c, S - 1);
  return max(c, 0);
}

template <int DIM, int LOG>
__device__ __forceinline__ int point_cell(const float* __restrict__ pts, uint32_t i,
                                          const float* lo, const float* hi) {
  int c = 0;
#pragma unroll
  for (int a = DIM - 1; a >= 0; --a)
    c = (c << LOG) |
        quantize_host<(1 << LOG)>(pts[static_cast<int64_t>(DIM) * i + a], lo[a], hi[a]);
  return c;
}

template <int DIM, int LOG>
__global__ __launch_bounds__(kB) void k_nd_init(int64_t n, const float* __restrict__ lo,
                                                const float* __restrict__ hi,
                                                uint32_t* __restrict__ pidx,
                                                uint32_t* __restrict__ pnode,
                                                uint32_t* __restrict__ act, NodeArrays N) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (i < n) {
    pidx[i] = static_cast<uint32_t>(i);
    pnode[i] = 0u;
    act[i] = static_cast<uint32_t>(i);
  }
  if (i == 0) {
    N.start[0] = 0u;
    N.end[0] = static_cast<uint32_t>(n);
    N.parent[0] = 0u;
    N.first[0] = 0u;
    N.cell[0] = 0u;
    N.num[0] = 0;
    N.depth[0] = 0;
    N.leaf[0] = n <= 1 ? 1u : 0u;
    for (int a = 0; a < DIM; ++a) {
      N.lo[a] = lo[a];
      N.hi[a] = hi[a];
    }
    for (int w = 0; w < NdT<DIM, LOG>::kWords; ++w) N.mask[w] = 0ull;
  }
}

// Step 1: (node, cell) keys of the active points.
template <int DIM, int LOG>
__global__ __launch_bounds__(kB) void k_nd_keys(const float* __restrict__ pts,
                                                const uint32_t* __restrict__ act, int64_t na,
                                                const uint32_t* __restrict__ pidx,
                                                const uint32_t* __restrict__ pnode, uint32_t a,
                                                NodeArrays N, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k >= na) return;
  const uint32_t pos = act[k];
  const uint32_t i = pidx[pos];
  const uint32_t node = pnode[pos];
  const int cell = point_cell<DIM, LOG>(pts, i, N.lo + static_cast<int64_t>(DIM) * node,
                                   N.hi + static_cast<int64_t>(DIM) * node);
  keys[k] = ((node - a) << NdT<DIM, LOG>::kCellBits) | static_cast<uint32_t>(cell);
  vals[k] = i;
}

// Step 3a: run starts (flags[na] = 0 so that the exclusive scan's last
// entry is the run count).
__global__ __launch_bounds__(kB) void k_nd_run_flags(const uint32_t* __restrict__ keys,
                                                     int64_t na, uint32_t* __restrict__ flags) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k > na) return;
  flags[k] = (k < na && (k == 0 || keys[k] != keys[k - 1])) ? 1u : 0u;
}

// Step 3b: children records and the points' new slots and nodes.  Child
// bounds as the reference's partition loop: inc = (max - min) / S, lo = min
// + k inc, hi = lo + inc, in fp32 without contraction.
template <int DIM, int LOG>
__global__ __launch_bounds__(kB) void k_nd_children(const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ runs,
                                                    const uint32_t* __restrict__ act, int64_t na,
                                                    uint32_t a, uint32_t b, int32_t depth,
                                                    uint32_t* __restrict__ pidx,
                                                    uint32_t* __restrict__ pnode, NodeArrays N) {
#pragma clang fp contract(off)
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k >= na) return;
  constexpr int CB = NdT<DIM, LOG>::kCellBits;
  constexpr int S = NdT<DIM, LOG>::kSplit;
  const uint32_t key = keys[k];
  const bool first_of_run = k == 0 || keys[k - 1] != key;
  const bool last_of_run = k == na - 1 || keys[k + 1] != key;
  const uint32_t c = b + runs[k] - (first_of_run ? 0u : 1u);
  const uint32_t pos = act[k];
  pidx[pos] = vals[k];
  pnode[pos] = c;
  if (first_of_run) {
    const uint32_t parent = a + (key >> CB);
    const uint32_t cell = key & ((1u << CB) - 1u);
    N.start[c] = pos;
    N.parent[c] = parent;
    N.cell[c] = cell;
    N.depth[c] = depth + 1;
    N.num[c] = 0;
    N.first[c] = 0u;
    for (int w = 0; w < NdT<DIM, LOG>::kWords; ++w)
      N.mask[static_cast<int64_t>(NdT<DIM, LOG>::kWords) * c + w] = 0ull;
    for (int ax = 0; ax < DIM; ++ax) {
      const float plo = N.lo[static_cast<int64_t>(DIM) * parent + ax];
      const float phi = N.hi[static_cast<int64_t>(DIM) * parent + ax];
      const float inc = (phi - plo) / static_cast<float>(S);
      const uint32_t q = (cell >> (LOG * ax)) & static_cast<uint32_t>(S - 1);
      const float lo = plo + static_cast<float>(q) * inc;
      N.lo[static_cast<int64_t>(DIM) * c + ax] = lo;
      N.hi[static_cast<int64_t>(DIM) * c + ax] = lo + inc;
    }
    if (k == 0 || (keys[k - 1] >> CB) != (key >> CB)) N.first[parent] = c;
  }
  if (last_of_run) N.end[c] = pos + 1u;
}

// Step 3c: leaf status of the new children (LEAF_SIZE 1, max depth 100).
__global__ __launch_bounds__(kB) void k_nd_leaf_flags(const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ runs,
                                                      int64_t na, uint32_t b, NodeArrays N) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k >= na) return;
  if (!(k == 0 || keys[k - 1] != keys[k])) return;
  const uint32_t c = b + runs[k];
  N.leaf[c] = (N.depth[c] >= kMaxDepthNd || N.end[c] - N.start[c] <= 1u) ? 1u : 0u;
}

// Step 4: the parents' masks and child counts (children are contiguous from
// `first`, in cell order).
template <int DIM, int LOG>
__global__ __launch_bounds__(kB) void k_nd_parents(uint32_t a, uint32_t b,
                                                   const uint32_t* __restrict__ run_count,
                                                   NodeArrays N) {
  const uint32_t p = a + blockIdx.x * kB + threadIdx.x;
  if (p >= b || N.leaf[p]) return;
  const uint32_t end = b + *run_count;
  constexpr int kW = NdT<DIM, LOG>::kWords;
  uint64_t m[kW];
  for (int w = 0; w < kW; ++w) m[w] = 0ull;
  int cnt = 0;
  for (uint32_t c = N.first[p]; c < end && N.parent[c] == p; ++c) {
    const uint32_t cell = N.cell[c];
    m[cell >> 6] |= 1ull << (cell & 63u);
    ++cnt;
  }
  for (int w = 0; w < kW; ++w) N.mask[static_cast<int64_t>(kW) * p + w] = m[w];
  N.num[p] = cnt;
}

// Step 5: points of non-leaf children stay active (flags[na] = 0).
__global__ __launch_bounds__(kB) void k_nd_active_flags(const uint32_t* __restrict__ act,
                                                        int64_t na,
                                                        const uint32_t* __restrict__ pnode,
                                                        NodeArrays N,
                                                        uint32_t* __restrict__ flags) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k > na) return;
  flags[k] = (k < na && !N.leaf[pnode[act[k]]]) ? 1u : 0u;
}

__global__ __launch_bounds__(kB) void k_nd_compact(const uint32_t* __restrict__ act, int64_t na,
                                                   const uint32_t* __restrict__ pnode,
                                                   NodeArrays N,
                                                   const uint32_t* __restrict__ slot,
                                                   uint32_t* __restrict__ act_out) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k >= na) return;
  const uint32_t pos = act[k];
  if (!N.leaf[pnode[pos]]) act_out[slot[k]] = pos;
}

__global__ void k_nd_counts(const uint32_t* __restrict__ runs, const uint32_t* __restrict__ acts,
                            int64_t na, uint32_t* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = runs[na];
    out[1] = acts[na];
  }
}

// Moments of one level, bottom-up (Tree2DNode::build for leaves,
// aggregateChildInfo for internal nodes; include/tree2d.h:253-263,
// 304-321): sequential fp32 sums in point / child order, no contraction,
// then position and normal divided by the mass (0 / 0 for a massless node,
// as in the reference).
template <int DIM>
__global__ __launch_bounds__(kB) void k_nd_moments(uint32_t a, uint32_t b,
                                                   const float* __restrict__ pts,
                                                   const float* __restrict__ nrm,
                                                   const float* __restrict__ ms, float alpha,
                                                   const uint32_t* __restrict__ pidx,
                                                   NodeArrays N) {
#pragma clang fp contract(off)
  const uint32_t c = a + blockIdx.x * kB + threadIdx.x;
  if (c >= b) return;
  float m = 0.0f, al = 0.0f, p[DIM], q[DIM];
  for (int ax = 0; ax < DIM; ++ax) p[ax] = q[ax] = 0.0f;
  if (N.leaf[c]) {
    for (uint32_t j = N.start[c]; j < N.end[c]; ++j) {
      const int64_t i = pidx[j];
      const float mi = ms[i];
      al = alpha;
      m = m + mi;
      for (int ax = 0; ax < DIM; ++ax) {
        p[ax] = p[ax] + mi * pts[DIM * i + ax];
        q[ax] = q[ax] + mi * nrm[DIM * i + ax];
      }
    }
  } else {
    const uint32_t f = N.first[c];
    for (int32_t o = 0; o < N.num[c]; ++o) {
      const uint32_t ch = f + static_cast<uint32_t>(o);
      const float mc = N.cm[ch];
      al = N.calpha[ch];
      m = m + mc;
      for (int ax = 0; ax < DIM; ++ax) {
        p[ax] = p[ax] + mc * N.cpos[static_cast<int64_t>(DIM) * ch + ax];
        q[ax] = q[ax] + mc * N.cnrm[static_cast<int64_t>(DIM) * ch + ax];
      }
    }
  }
  N.cm[c] = m;
  N.calpha[c] = al;
  for (int ax = 0; ax < DIM; ++ax) {
    N.cpos[static_cast<int64_t>(DIM) * c + ax] = p[ax] / m;
    N.cnrm[static_cast<int64_t>(DIM) * c + ax] = q[ax] / m;
  }
}

// Point records in tree order (LeafData: position, normal, mass, alpha).
template <int DIM>
__global__ __launch_bounds__(kB) void k_nd_leaves(const float* __restrict__ pts,
                                                  const float* __restrict__ nrm,
                                                  const float* __restrict__ ms, float alpha,
                                                  int64_t n, const uint32_t* __restrict__ pidx,
                                                  float* __restrict__ leaf) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (j >= n) return;
  const int64_t i = pidx[j];
  float* o = leaf + (2 * DIM + 2) * j;
  for (int ax = 0; ax < DIM; ++ax) {
    o[ax] = pts[DIM * i + ax];
    o[DIM + ax] = nrm[DIM * i + ax];
  }
  o[2 * DIM] = ms[i];
  o[2 * DIM + 1] = alpha;
}

template <int DIM>
__global__ __launch_bounds__(kB) void k_nd_emit_contrib(int64_t nn, NodeArrays N,
                                                        float* __restrict__ out) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (c >= nn) return;
  float* o = out + (2 * DIM + 2) * c;
  for (int ax = 0; ax < DIM; ++ax) {
    o[ax] = N.cpos[DIM * c + ax];
    o[DIM + ax] = N.cnrm[DIM * c + ax];
  }
  o[2 * DIM] = N.cm[c];
  o[2 * DIM + 1] = N.calpha[c];
}

// One 32-bit word of a node record per thread (e.g. 154 words for
// Tree2DNode<1,3>, 1068 for Tree3DNode<1,3>), so that the stores of a wave
// are contiguous.
template <int DIM, int LOG>
__global__ __launch_bounds__(kB) void k_nd_emit_nodes(int64_t nn, NodeArrays N,
                                                      uint32_t* __restrict__ out) {
  using Node = typename NdT<DIM, LOG>::Node;
  constexpr int64_t kW = sizeof(Node) / 4;
  constexpr int kWords = NdT<DIM, LOG>::kWords;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (t >= nn * kW) return;
  const int64_t c = t / kW;
  const int w = static_cast<int>(t - c * kW);
  const int byte = 4 * w;
  uint32_t v = 0u;
  auto lo32 = [](uint64_t x) { return static_cast<uint32_t>(x); };
  auto hi32 = [](uint64_t x) { return static_cast<uint32_t>(x >> 32); };
  auto u64_word = [&](int off, uint64_t x) { return (byte - off) == 0 ? lo32(x) : hi32(x); };
  if (byte < 8) {
    v = u64_word(0, N.start[c]);
  } else if (byte < 16) {
    v = u64_word(8, N.end[c]);
  } else if (byte < 24) {
    v = u64_word(16, static_cast<uint64_t>(c));
  } else if (byte < 32) {
    v = u64_word(24, N.parent[c]);
  } else if (byte < static_cast<int>(offsetof(Node, bmax))) {
    v = __float_as_uint(N.lo[DIM * c + (byte - 32) / 4]);
  } else if (byte < static_cast<int>(offsetof(Node, diameter))) {
    v = __float_as_uint(N.hi[DIM * c + (byte - static_cast<int>(offsetof(Node, bmax))) / 4]);
  } else if (byte == static_cast<int>(offsetof(Node, diameter))) {
#pragma clang fp contract(off)
    // AABB<DIM>::diagonal().norm() (include/tree2d.h:248)
    float s = 0.0f;
    for (int ax = 0; ax < DIM; ++ax) {
      const float d = N.hi[DIM * c + ax] - N.lo[DIM * c + ax];
      s = ax == 0 ? d * d : s + d * d;
    }
    v = __float_as_uint(sqrtf(s));
  } else if (byte < static_cast<int>(offsetof(Node, contrib) + sizeof(Node::contrib))) {
    const int f = (byte - static_cast<int>(offsetof(Node, contrib))) / 4;
    float x;
    if (f < DIM)
      x = N.cpos[DIM * c + f];
    else if (f < 2 * DIM)
      x = N.cnrm[DIM * c + f - DIM];
    else if (f == 2 * DIM)
      x = N.cm[c];
    else
      x = N.calpha[c];
    v = __float_as_uint(x);
  } else if (byte < static_cast<int>(offsetof(Node, bitmask))) {
    v = 0u;  // padding
  } else if (byte < static_cast<int>(offsetof(Node, child))) {
    const int wd = (byte - static_cast<int>(offsetof(Node, bitmask))) / 8;
    v = u64_word(static_cast<int>(offsetof(Node, bitmask)) + 8 * wd, N.mask[kWords * c + wd]);
  } else if (byte < static_cast<int>(offsetof(Node, num_children))) {
    const int cell = (byte - static_cast<int>(offsetof(Node, child))) / 8;
    const uint64_t* mk = N.mask + kWords * c;
    uint64_t id = 0ull;
    if ((mk[cell >> 6] >> (cell & 63)) & 1ull) {
      int rank = 0;
      for (int wd = 0; wd < (cell >> 6); ++wd) rank += __popcll(mk[wd]);
      rank += __popcll(mk[cell >> 6] & ((1ull << (cell & 63)) - 1ull));
      id = static_cast<uint64_t>(N.first[c]) + static_cast<uint64_t>(rank);
    }
    v = u64_word(static_cast<int>(offsetof(Node, child)) + 8 * cell, id);
  } else if (byte == static_cast<int>(offsetof(Node, num_children))) {
    v = static_cast<uint32_t>(N.num[c]);
  } else if (byte == static_cast<int>(offsetof(Node, depth))) {
    v = static_cast<uint32_t>(N.depth[c]);
  } else if (byte == static_cast<int>(offsetof(Node, max_depth))) {
    v = static_cast<uint32_t>(kMaxDepthNd);
  } else if (byte == static_cast<int>(offsetof(Node, is_leaf))) {
    v = N.leaf[c] ? 1u : 0u;  // is_leaf byte, then three zero pad bytes
  }
  out[t] = v;
}

// ------------------------------------------------------------------ state
// Per (device, stream) build state: point arrays sized for the last n,
// node arrays grown (with copy) as levels are added.
struct NdState {
  int dim = 0;
  int level = 3;
  int64_t n = 0;
  int64_t num_nodes = 0;
  int64_t pt_cap = -1;
  int64_t node_cap = 0;
  int node_dim = 0;  // dimension and level the node arrays were carved for
  int node_level = 0;
  void* pt_block = nullptr;
  void* node_block = nullptr;
  uint32_t* host_counts = nullptr;  // pinned, 2 words
  uint32_t *pidx, *pnode, *act, *act2, *keys, *vals, *keys2, *vals2, *runs, *acts, *counts;
  float* leaf;
  void* sort_temp;
  void* scan_temp;
  NodeArrays N;
};

std::mutex g_nd_mutex;
std::map<std::tuple<int, hipStream_t>, NdState> g_nd;

size_t align256(size_t x) { return (x + 255) & ~size_t{255}; }

int mask_words(int dim, int level) {
  const int cells = 1 << (dim * level);
  return cells >= 64 ? cells / 64 : 1;
}

size_t node_bytes_per(int dim, int level) {
  const int words = mask_words(dim, level);
  return 4 * 8 + 4 * 2 * dim + 8 * words + 4 * (2 + 2 * dim);
}

// Carves the node arrays for `cap` nodes out of `base`.
NodeArrays carve_nodes(char* base, int64_t cap, int dim, int level) {
  NodeArrays N;
  const int words = mask_words(dim, level);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = base + o;
    o = align256(o + bytes);
    return p;
  };
  const size_t c = static_cast<size_t>(cap);
  N.start = reinterpret_cast<uint32_t*>(take(4 * c));
  N.end = reinterpret_cast<uint32_t*>(take(4 * c));
  N.parent = reinterpret_cast<uint32_t*>(take(4 * c));
  N.first = reinterpret_cast<uint32_t*>(take(4 * c));
  N.cell = reinterpret_cast<uint32_t*>(take(4 * c));
  N.num = reinterpret_cast<int32_t*>(take(4 * c));
  N.depth = reinterpret_cast<int32_t*>(take(4 * c));
  N.leaf = reinterpret_cast<uint32_t*>(take(4 * c));
  N.lo = reinterpret_cast<float*>(take(4 * dim * c));
  N.hi = reinterpret_cast<float*>(take(4 * dim * c));
  N.mask = reinterpret_cast<uint64_t*>(take(8 * words * c));
  N.cm = reinterpret_cast<float*>(take(4 * c));
  N.cpos = reinterpret_cast<float*>(take(4 * dim * c));
  N.cnrm = reinterpret_cast<float*>(take(4 * dim * c));
  N.calpha = reinterpret_cast<float*>(take(4 * c));
  return N;
}

size_t node_block_bytes(int64_t cap, int dim, int level) {
  return node_bytes_per(dim, level) * static_cast<size_t>(cap) + 16 * 256;
}

// Grows the node arrays to hold at least `need` nodes, keeping the first
// `keep` entries of every array.
int grow_nodes(NdState& S, hipStream_t s, int64_t need, int64_t keep) {
  if (S.node_block != nullptr && (S.node_dim != S.dim || S.node_level != S.level)) {
    SBH_HIP_TRY(hipStreamSynchronize(s));
    SBH_HIP_TRY(hipFree(S.node_block));
    S.node_block = nullptr;
    S.node_cap = 0;
    keep = 0;
  }
  if (need <= S.node_cap && S.node_block != nullptr) return SBH_OK;
  int64_t cap = S.node_cap > 0 ? S.node_cap : 1024;
  while (cap < need) cap *= 2;
  void* blk = nullptr;
  SBH_HIP_TRY(hipMalloc(&blk, node_block_bytes(cap, S.dim, S.level)));
  NodeArrays M = carve_nodes(static_cast<char*>(blk), cap, S.dim, S.level);
  if (S.node_block != nullptr && keep > 0) {
    const NodeArrays& O = S.N;
    const size_t k = static_cast<size_t>(keep);
    const int words = mask_words(S.dim, S.level);
    const std::pair<void*, const void*> arr[] = {
        {M.start, O.start}, {M.end, O.end},     {M.parent, O.parent}, {M.first, O.first},
        {M.cell, O.cell},   {M.num, O.num},     {M.depth, O.depth},   {M.leaf, O.leaf},
        {M.cm, O.cm},       {M.calpha, O.calpha}};
    for (const auto& pr : arr)
      SBH_HIP_TRY(hipMemcpyAsync(pr.first, pr.second, 4 * k, hipMemcpyDeviceToDevice, s));
    const std::pair<void*, const void*> arr_d[] = {
        {M.lo, O.lo}, {M.hi, O.hi}, {M.cpos, O.cpos}, {M.cnrm, O.cnrm}};
    for (const auto& pr : arr_d)
      SBH_HIP_TRY(hipMemcpyAsync(pr.first, pr.second, 4 * S.dim * k, hipMemcpyDeviceToDevice,
                                 s));
    SBH_HIP_TRY(hipMemcpyAsync(M.mask, O.mask, 8 * words * k, hipMemcpyDeviceToDevice, s));
  }
  if (S.node_block != nullptr) {
    SBH_HIP_TRY(hipStreamSynchronize(s));
    SBH_HIP_TRY(hipFree(S.node_block));
  }
  S.node_block = blk;
  S.node_cap = cap;
  S.node_dim = S.dim;
  S.node_level = S.level;
  S.N = M;
  return SBH_OK;
}

int prepare_points(NdState& S, hipStream_t s, int64_t n) {
  if (S.host_counts == nullptr)
    SBH_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&S.host_counts), 2 * sizeof(uint32_t),
                              hipHostMallocDefault));
  const int64_t nn = n > 0 ? n : 1;
  if (S.pt_block != nullptr && nn <= S.pt_cap) return SBH_OK;
  if (S.pt_block != nullptr) {
    SBH_HIP_TRY(hipStreamSynchronize(s));
    SBH_HIP_TRY(hipFree(S.pt_block));
    S.pt_block = nullptr;
  }
  const size_t u = 4 * static_cast<size_t>(nn);
  const size_t u1 = 4 * static_cast<size_t>(nn + 1);
  size_t o = 0;
  size_t off[13];
  off[0] = o; o = align256(o + u);   // pidx
  off[1] = o; o = align256(o + u);   // pnode
  off[2] = o; o = align256(o + u);   // act
  off[3] = o; o = align256(o + u);   // act2
  off[4] = o; o = align256(o + u);   // keys
  off[5] = o; o = align256(o + u);   // vals
  off[6] = o; o = align256(o + u);   // keys2
  off[7] = o; o = align256(o + u);   // vals2
  off[8] = o; o = align256(o + u1);  // runs
  off[9] = o; o = align256(o + u1);  // acts
  off[10] = o; o = align256(o + 32 * static_cast<size_t>(nn));  // leaf records (<= 32 B)
  off[11] = o; o = align256(o + radix_sort_temp_bytes(nn));
  off[12] = o; o = align256(o + exclusive_scan_temp_bytes(nn + 1));
  const size_t counts_off = o;
  o = align256(o + 16);
  void* blk = nullptr;
  SBH_HIP_TRY(hipMalloc(&blk, o));
  char* b = static_cast<char*>(blk);
  S.pt_block = blk;
  S.pt_cap = nn;
  S.pidx = reinterpret_cast<uint32_t*>(b + off[0]);
  S.pnode = reinterpret_cast<uint32_t*>(b + off[1]);
  S.act = reinterpret_cast<uint32_t*>(b + off[2]);
  S.act2 = reinterpret_cast<uint32_t*>(b + off[3]);
  S.keys = reinterpret_cast<uint32_t*>(b + off[4]);
  S.vals = reinterpret_cast<uint32_t*>(b + off[5]);
  S.keys2 = reinterpret_cast<uint32_t*>(b + off[6]);
  S.vals2 = reinterpret_cast<uint32_t*>(b + off[7]);
  S.runs = reinterpret_cast<uint32_t*>(b + off[8]);
  S.acts = reinterpret_cast<uint32_t*>(b + off[9]);
  S.leaf = reinterpret_cast<float*>(b + off[10]);
  S.sort_temp = b + off[11];
  S.scan_temp = b + off[12];
  S.counts = reinterpret_cast<uint32_t*>(b + counts_off);
  return SBH_OK;
}

int bits_for(uint64_t count) {  // bits of the largest index count - 1
  int b = 0;
  while (count > 1 && (1ull << b) < count) ++b;
  return b;
}

template <int DIM, int LOG>
int build_nd(NdState& S, hipStream_t s, const float* pts, const float* nrm, const float* ms,
             const float* lo, const float* hi, int64_t n, float alpha) {
  constexpr int CB = NdT<DIM, LOG>::kCellBits;
  int rc = prepare_points(S, s, n);
  if (rc != SBH_OK) return rc;
  rc = grow_nodes(S, s, 1, 0);
  if (rc != SBH_OK) return rc;
  std::vector<int64_t> level{0, 1};
  k_nd_init<DIM, LOG><<<grid_for(n, kB), kB, 0, s>>>(n, lo, hi, S.pidx, S.pnode, S.act, S.N);
  SBH_LAUNCH_CHECK();
  int64_t na = n > 1 ? n : 0;
  int32_t depth = 0;
  while (na > 0) {
    const int64_t a = level[depth], b = level[depth + 1];
    const int key_bits = CB + bits_for(static_cast<uint64_t>(b - a));
    if (key_bits > 32) return fail_invalid("tree level too wide for 32-bit keys");
    rc = grow_nodes(S, s, b + na, b);
    if (rc != SBH_OK) return rc;
    const unsigned g = grid_for(na, kB), g1 = grid_for(na + 1, kB);
    k_nd_keys<DIM, LOG><<<g, kB, 0, s>>>(pts, S.act, na, S.pidx, S.pnode,
                                         static_cast<uint32_t>(a), S.N, S.keys, S.vals);
    SBH_LAUNCH_CHECK();
    rc = radix_sort_pairs(s, S.keys, S.vals, S.keys2, S.vals2, na, key_bits, S.sort_temp);
    if (rc != SBH_OK) return rc;
    k_nd_run_flags<<<g1, kB, 0, s>>>(S.keys2, na, S.runs);
    SBH_LAUNCH_CHECK();
    rc = exclusive_scan(s, S.runs, na + 1, S.scan_temp);
    if (rc != SBH_OK) return rc;
    k_nd_children<DIM, LOG><<<g, kB, 0, s>>>(S.keys2, S.vals2, S.runs, S.act, na,
                                             static_cast<uint32_t>(a),
                                             static_cast<uint32_t>(b), depth, S.pidx,
                                             S.pnode, S.N);
    SBH_LAUNCH_CHECK();
    k_nd_leaf_flags<<<g, kB, 0, s>>>(S.keys2, S.runs, na, static_cast<uint32_t>(b), S.N);
    SBH_LAUNCH_CHECK();
    k_nd_parents<DIM, LOG><<<grid_for(b - a, kB), kB, 0, s>>>(
        static_cast<uint32_t>(a), static_cast<uint32_t>(b), S.runs + na, S.N);
    SBH_LAUNCH_CHECK();
    k_nd_active_flags<<<g1, kB, 0, s>>>(S.act, na, S.pnode, S.N, S.acts);
    SBH_LAUNCH_CHECK();
    rc = exclusive_scan(s, S.acts, na + 1, S.scan_temp);
    if (rc != SBH_OK) return rc;
    k_nd_compact<<<g, kB, 0, s>>>(S.act, na, S.pnode, S.N, S.acts, S.act2);
    SBH_LAUNCH_CHECK();
    k_nd_counts<<<1, 64, 0, s>>>(S.runs, S.acts, na, S.counts);
    SBH_LAUNCH_CHECK();
    SBH_HIP_TRY(hipMemcpyAsync(S.host_counts, S.counts, 2 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
    SBH_HIP_TRY(hipStreamSynchronize(s));
    const int64_t children = S.host_counts[0];
    const int64_t next_active = S.host_counts[1];
    level.push_back(b + children);
    std::swap(S.act, S.act2);
    na = next_active;
    ++depth;
    if (depth > kMaxDepthNd) return fail_internal_msg("tree build exceeded the maximum depth");
  }
  S.num_nodes = level.back();
  for (int l = static_cast<int>(level.size()) - 2; l >= 0; --l) {
    const int64_t a = level[l], b = level[l + 1];
    k_nd_moments<DIM><<<grid_for(b - a, kB), kB, 0, s>>>(
        static_cast<uint32_t>(a), static_cast<uint32_t>(b), pts, nrm, ms, alpha, S.pidx, S.N);
    SBH_LAUNCH_CHECK();
  }
  if (n > 0) {
    k_nd_leaves<DIM><<<grid_for(n, kB), kB, 0, s>>>(pts, nrm, ms, alpha, n, S.pidx, S.leaf);
    SBH_LAUNCH_CHECK();
  }
  S.n = n;
  S.dim = DIM;
  S.level = LOG;
  return SBH_OK;
}

int64_t node_size(int dim, int level) {
  if (dim == 2) return level == 1 ? sizeof(NodeT<2, 1>) : level == 2 ? sizeof(NodeT<2, 2>) : sizeof(NodeT<2, 3>);
  return level == 1 ? sizeof(NodeT<3, 1>) : level == 2 ? sizeof(NodeT<3, 2>) : sizeof(NodeT<3, 3>);
}

template <int DIM>
int build_dim(int level, NdState& S, hipStream_t s, const float* pts, const float* nrm,
              const float* ms, const float* lo, const float* hi, int64_t n, float alpha) {
  if (level == 1) return build_nd<DIM, 1>(S, s, pts, nrm, ms, lo, hi, n, alpha);
  if (level == 2) return build_nd<DIM, 2>(S, s, pts, nrm, ms, lo, hi, n, alpha);
  return build_nd<DIM, 3>(S, s, pts, nrm, ms, lo, hi, n, alpha);
}

template <int DIM, int LOG>
void emit_nodes(hipStream_t s, const NdState& S, void* out) {
  const int64_t words = S.num_nodes * static_cast<int64_t>(sizeof(NodeT<DIM, LOG>)) / 4;
  k_nd_emit_nodes<DIM, LOG><<<grid_for(words, kB), kB, 0, s>>>(S.num_nodes, S.N,
                                                              static_cast<uint32_t*>(out));
}

template <int DIM>
void emit_nodes_dim(hipStream_t s, const NdState& S, void* out) {
  if (S.level == 1)
    emit_nodes<DIM, 1>(s, S, out);
  else if (S.level == 2)
    emit_nodes<DIM, 2>(s, S, out);
  else
    emit_nodes<DIM, 3>(s, S, out);
}

NdState& state_for(hipStream_t s) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return g_nd[std::make_tuple(dev, s)];
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" size_t sbh_leaf2_size(void) { return sizeof(Leaf2); }
extern "C" size_t sbh_node2_size(void) { return sizeof(Node2); }

extern "C" int sbh_tree_build(void* stream, int32_t dim, int32_t level,
                              const SbhBuffer* points, const SbhBuffer* normals,
                              const SbhBuffer* masses, const SbhBuffer* min_corner,
                              const SbhBuffer* max_corner, float alpha, int64_t* num_nodes) {
  clear_error();
  if (dim != 2 && dim != 3) return fail_invalid("dim must be 2 or 3");
  if (level < 1 || level > 3) return fail_invalid("level must be 1, 2 or 3");
  if (num_nodes == nullptr) return fail_invalid("num_nodes must not be null");
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != dim)
    return fail_invalid(dim == 2 ? "points must have shape (N, 2)" : "points must have shape (N, 3)");
  if (!buffer_ok(normals, SBH_DTYPE_F32, 2) || normals->dims[0] != points->dims[0] ||
      normals->dims[1] != dim)
    return fail_invalid("normals shape mismatch");
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != points->dims[0])
    return fail_invalid("masses must have shape (N,)");
  if (!buffer_ok(min_corner, SBH_DTYPE_F32, 1) || min_corner->dims[0] != dim ||
      !buffer_ok(max_corner, SBH_DTYPE_F32, 1) || max_corner->dims[0] != dim)
    return fail_invalid("min_corner/max_corner must have shape (dim,)");
  const int64_t n = points->dims[0];
  // every level holds at most n nodes; their local index and the cell must
  // fit a 32-bit sort key
  if (n > (dim == 2 ? (int64_t{1} << 26) : (int64_t{1} << 23)))
    return fail_invalid("too many points for the tree builder");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (stream_capturing(s)) return fail_invalid("the tree builder cannot run under stream capture");
  std::lock_guard<std::mutex> lock(g_nd_mutex);
  NdState& S = state_for(s);
  S.dim = dim;
  S.level = level;
  S.num_nodes = 0;
  const auto* pts = static_cast<const float*>(points->data);
  const auto* nrm = static_cast<const float*>(normals->data);
  const auto* ms = static_cast<const float*>(masses->data);
  const auto* lo = static_cast<const float*>(min_corner->data);
  const auto* hi = static_cast<const float*>(max_corner->data);
  const int rc = dim == 2 ? build_dim<2>(level, S, s, pts, nrm, ms, lo, hi, n, alpha)
                          : build_dim<3>(level, S, s, pts, nrm, ms, lo, hi, n, alpha);
  if (rc != SBH_OK) {
    S.num_nodes = 0;
    return rc;
  }
  *num_nodes = S.num_nodes;
  return SBH_OK;
}

extern "C" int sbh_tree_emit(void* stream, SbhBuffer* leaf_out, SbhBuffer* node_out,
                             SbhBuffer* contrib_out) {
  clear_error();
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g_nd_mutex);
  NdState& S = state_for(s);
  if (S.num_nodes <= 0) return fail_invalid("no tree has been built on this stream");
  const int64_t leaf_sz = S.dim == 2 ? sizeof(Leaf2) : sizeof(Leaf3);
  const int64_t node_sz = node_size(S.dim, S.level);
  if (!buffer_ok(leaf_out, SBH_DTYPE_U8, -1) || byte_count(leaf_out) != S.n * leaf_sz)
    return fail_invalid("leaf_out size mismatch");
  if (!buffer_ok(node_out, SBH_DTYPE_U8, -1) || byte_count(node_out) != S.num_nodes * node_sz)
    return fail_invalid("node_out size mismatch");
  if (contrib_out != nullptr &&
      (!buffer_ok(contrib_out, SBH_DTYPE_U8, -1) ||
       byte_count(contrib_out) != S.num_nodes * leaf_sz))
    return fail_invalid("contrib_out size mismatch");
  if (S.n > 0)
    SBH_HIP_TRY(hipMemcpyAsync(leaf_out->data, S.leaf, static_cast<size_t>(S.n * leaf_sz),
                               hipMemcpyDeviceToDevice, s));
  if (S.dim == 2)
    emit_nodes_dim<2>(s, S, node_out->data);
  else
    emit_nodes_dim<3>(s, S, node_out->data);
  SBH_LAUNCH_CHECK();
  if (contrib_out != nullptr) {
    if (S.dim == 2)
      k_nd_emit_contrib<2><<<grid_for(S.num_nodes, kB), kB, 0, s>>>(
          S.num_nodes, S.N, static_cast<float*>(contrib_out->data));
    else
      k_nd_emit_contrib<3><<<grid_for(S.num_nodes, kB), kB, 0, s>>>(
          S.num_nodes, S.N, static_cast<float*>(contrib_out->data));
    SBH_LAUNCH_CHECK();
  }
  return SBH_OK;
}
