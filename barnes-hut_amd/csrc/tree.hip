// Adaptive Morton octree + wave-coherent group walk with full adjoint.
//
// Build (all on the device, no host synchronisation):
//   k_bbox_*      data bounds (min - 1e-3, max + 1e-3, as python/jax_sbh_ffi.py:879-884)
//   k_codes       D-level Morton codes with the reference quantiser (src/sbh_ffi.cu:162-179)
//   morton_sort   stable (set, code, row) order (onesweep.hip)
//   k_gather      sorted (x, y, z, m) and cotangent records
//   k_topo_*      every node and query group from the sorted codes: count ->
//                 scan -> emit (first and last points of each cell, topo.hpp)
//   k_topo_finalize  record counts, group sizes, leaf moments (mass, centre of
//                 mass, central second moment), one wavefront per node
//   per level:    k_moments_internal bottom-up over the internal nodes
// Walk (one wave = one group of 64 consecutive sorted points):
//   a node is used as a multipole for the whole group iff the distance from its
//   cell centre to the group's bounding box is >= beta * cell diagonal (the
//   reference's MAC, src/sbh_ffi.cu:1878-1893, made conservative for a group);
//   rejected leaves are summed exactly.  Lanes test 64 candidate nodes at once;
//   accepted nodes and near leaves are batched in LDS and evaluated with all 64
//   lanes sharing each node / source record, which is read once per wave
//   through the scalar cache (the records are read-only during the walk).
// Adjoint: query-side terms stay in registers; source-side terms of each
//   interaction batch are summed across the wave through an LDS slab and
//   added with one float atomic per value; node adjoints are pushed down to
//   the points in k_finalize.
#include "common.hpp"
#include "tree.hpp"
#include "groupwalk.hpp"
#include "topo.hpp"

namespace sbh {
namespace tree {
namespace {

using namespace gw;

// ------------------------------------------------------------------ bbox
__global__ __launch_bounds__(kB) void k_bbox_partial(const float* __restrict__ pts, int64_t n,
                                                     float* __restrict__ part) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    for (int a = 0; a < 3; ++a) {
      float v = pts[3 * i + a];
      lo[a] = fminf(lo[a], v);
      hi[a] = fmaxf(hi[a], v);
    }
  }
  __shared__ float sh[6][kB];
  for (int a = 0; a < 3; ++a) {
    sh[a][threadIdx.x] = lo[a];
    sh[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int off = kB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) {
      for (int a = 0; a < 3; ++a) {
        sh[a][threadIdx.x] = fminf(sh[a][threadIdx.x], sh[a][threadIdx.x + off]);
        sh[3 + a][threadIdx.x] = fmaxf(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + off]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// Reduces the per-block partials: bounds[0..3) = min - 1e-3, bounds[3..6) =
// max + 1e-3 (python/jax_sbh_ffi.py:879-884); an empty set gives +-1e-3.
__global__ __launch_bounds__(kB) void k_bbox_reduce(const float* __restrict__ part, int nparts,
                                                    int64_t n, float* __restrict__ bounds) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nparts; b += kB)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], part[6 * b + a]);
      hi[a] = fmaxf(hi[a], part[6 * b + 3 + a]);
    }
  __shared__ float sh[6][kB];
  for (int a = 0; a < 3; ++a) {
    sh[a][threadIdx.x] = lo[a];
    sh[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int off = kB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) {
      for (int a = 0; a < 3; ++a) {
        sh[a][threadIdx.x] = fminf(sh[a][threadIdx.x], sh[a][threadIdx.x + off]);
        sh[3 + a][threadIdx.x] = fmaxf(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + off]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float v = n == 0 ? 0.0f : sh[a][0];
    bounds[a] = a < 3 ? v - 1e-3f : v + 1e-3f;
  }
}

// Order-preserving map of floats to unsigned ints (for atomic min / max).
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ __launch_bounds__(kB) void k_setbox_init(uint32_t* __restrict__ box, int nsets) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < 6 * nsets; i += gridDim.x * kB)
    box[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
}

// Per-set data bounds of the input rows (sets are contiguous row ranges).
// Pass 1: one block per tile of kB consecutive points.  A tile
// inside one set stores its box in part[6 t ..]; a tile that straddles set
// boundaries (at most nsets - 1 of them) stores the neutral box and adds its
// points to the set boxes with ordered-integer atomics.
__global__ __launch_bounds__(kB) void k_set_bbox_tiles(const float* __restrict__ pts, int64_t n,
                                                       const int32_t* __restrict__ offs,
                                                       int nsets, uint32_t* __restrict__ box,
                                                       float* __restrict__ part) {
  __shared__ float sh[kB / 64][6];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t t = blockIdx.x;
  const int64_t b0 = t * kB;
  const int64_t i = b0 + threadIdx.x;
  const bool ok = i < n;
  const int64_t last = min<int64_t>(b0 + kB, n) - 1;
  const int s_first = set_of(offs, nsets, static_cast<int>(b0));
  const int s_last = set_of(offs, nsets, static_cast<int>(last));
  float v[6];
  for (int a = 0; a < 3; ++a) {
    const float x = ok ? pts[3 * i + a] : 0.0f;
    v[a] = ok ? x : INFINITY;
    v[3 + a] = ok ? x : -INFINITY;
  }
  if (s_first != s_last) {
    if (threadIdx.x < 6) part[6 * t + threadIdx.x] = threadIdx.x < 3 ? INFINITY : -INFINITY;
    if (ok) {
      const int my = set_of(offs, nsets, static_cast<int>(i));
      for (int a = 0; a < 3; ++a) {
        atomicMin(&box[6 * my + a], f2ord(v[a]));
        atomicMax(&box[6 * my + 3 + a], f2ord(v[3 + a]));
      }
    }
    return;
  }
  for (int d = 32; d > 0; d >>= 1)
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], __shfl_xor(v[a], d, 64));
      v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], d, 64));
    }
  if (lane == 0)
    for (int a = 0; a < 6; ++a) sh[w][a] = v[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float r = sh[0][a];
    for (int k = 1; k < kB / 64; ++k) r = a < 3 ? fminf(r, sh[k][a]) : fmaxf(r, sh[k][a]);
    part[6 * t + a] = r;
  }
}

// Pass 2: one block per set reduces the partial boxes of the tiles that
// overlap it (straddling tiles hold the neutral box) into the set's box.
__global__ __launch_bounds__(kB) void k_set_bbox_sets(const int32_t* __restrict__ offs,
                                                      const float* __restrict__ part,
                                                      uint32_t* __restrict__ box) {
  __shared__ float sh[kB / 64][6];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int k = blockIdx.x;
  const int o0 = offs[k], o1 = offs[k + 1];
  if (o1 <= o0) return;
  const int t0 = o0 / kB, t1 = (o1 - 1) / kB + 1;
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int t = t0 + static_cast<int>(threadIdx.x); t < t1; t += kB)
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], part[6 * static_cast<int64_t>(t) + a]);
      v[3 + a] = fmaxf(v[3 + a], part[6 * static_cast<int64_t>(t) + 3 + a]);
    }
  for (int d = 32; d > 0; d >>= 1)
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], __shfl_xor(v[a], d, 64));
      v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], d, 64));
    }
  if (lane == 0)
    for (int a = 0; a < 6; ++a) sh[w][a] = v[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float r = sh[0][a];
    for (int j = 1; j < kB / 64; ++j) r = a < 3 ? fminf(r, sh[j][a]) : fmaxf(r, sh[j][a]);
    if (a < 3)
      atomicMin(&box[6 * k + a], f2ord(r));
    else
      atomicMax(&box[6 * k + a], f2ord(r));
  }
}

// Initialises the forest: per-set cube geometry, one root per set (node k is
// the root of set k), the level table and the zero-mass sentinel node.  The
// bounds come either as six final floats (one set: min xyz, max xyz) or as
// the per-set atomics of k_set_bbox (expanded by 1e-3 here; an empty set
// gets the cube of the origin).
__global__ __launch_bounds__(kB) void k_forest_init(
    int64_t n, int nsets, const float* __restrict__ bounds6,
    const uint32_t* __restrict__ setbox, int32_t* __restrict__ offs, CubeGeom* __restrict__ geoms,
    Header* __restrict__ hdr, int4* __restrict__ topo, uint32_t* __restrict__ prefix,
    float4* __restrict__ geo, int32_t* __restrict__ parent, float4* __restrict__ mp,
    float4* __restrict__ qd, int sentinel) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  if (setbox == nullptr) {
    if (threadIdx.x == 0) {
      offs[0] = 0;
      offs[1] = static_cast<int32_t>(n);
    }
  } else {
    // offsets must run monotonically from 0 to n; otherwise every set is
    // left empty and the header reports overflow = 2
    for (int k = threadIdx.x; k < nsets; k += kB)
      if (offs[k] < 0 || offs[k] > offs[k + 1] || offs[k + 1] > n) atomicOr(&bad, 1);
    if (threadIdx.x == 0 && (offs[0] != 0 || offs[nsets] != n)) atomicOr(&bad, 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nsets; k += kB) {
    float lo[3], hi[3];
    int32_t s0, cnt;
    if (setbox == nullptr) {
      for (int a = 0; a < 3; ++a) {
        lo[a] = bounds6[a];
        hi[a] = bounds6[3 + a];
      }
      s0 = 0;
      cnt = static_cast<int32_t>(n);
    } else {
      s0 = bad ? 0 : offs[k];
      cnt = bad ? 0 : offs[k + 1] - s0;
      for (int a = 0; a < 3; ++a) {
        const float l = cnt > 0 ? ord2f(setbox[6 * k + a]) : 0.0f;
        const float h = cnt > 0 ? ord2f(setbox[6 * k + 3 + a]) : 0.0f;
        lo[a] = l - 1e-3f;
        hi[a] = h + 1e-3f;
      }
    }
    CubeGeom g;
    g.min_x = lo[0];
    g.min_y = lo[1];
    g.min_z = lo[2];
    g.span = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    g.cx = lo[0] + g.span * 0.5f;
    g.cy = lo[1] + g.span * 0.5f;
    g.cz = lo[2] + g.span * 0.5f;
    g.half_size = g.span * 0.5f;
    geoms[k] = g;
    topo[k] = make_int4(s0, cnt, -1, 0);
    prefix[k] = 0;
    parent[k] = -1;
    geo[k] = cell_geometry(g, 0u, 0);
    if (k == 0) hdr->geom = g;
  }
  if (threadIdx.x == 0) {
    for (int l = 0; l < kMaxLevels; ++l) {
      hdr->level_start[l] = 0;
      hdr->level_count[l] = 0;
    }
    hdr->level_count[0] = nsets;
    hdr->num_nodes = nsets;
    hdr->overflow = bad ? 2 : 0;
    for (int k = 0; k < 8; ++k) hdr->stats[k] = 0;
    mp[sentinel] = make_float4(0.f, 0.f, 0.f, 0.f);
    qd[2 * sentinel] = make_float4(0.f, 0.f, 0.f, 0.f);
    qd[2 * sentinel + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// ------------------------------------------------------------------ codes
__global__ __launch_bounds__(kB) void k_codes(const float* __restrict__ pts, int64_t n,
                                              const CubeGeom* __restrict__ geoms,
                                              const int32_t* __restrict__ offs, int nsets,
                                              int bits, uint32_t* __restrict__ codes_in) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (i >= n) return;
  const CubeGeom g = geoms[nsets > 1 ? set_of(offs, nsets, static_cast<int>(i)) : 0];
  uint32_t qx = quantize_axis(pts[3 * i + 0], g.min_x, g.span, bits);
  uint32_t qy = quantize_axis(pts[3 * i + 1], g.min_y, g.span, bits);
  uint32_t qz = quantize_axis(pts[3 * i + 2], g.min_z, g.span, bits);
  codes_in[i] = morton3(qx, qy, qz);
}

// Sorted records: row r of the sorted order is input row order[r] with code
// sorted[r]; writes the permutation and the codes where the sort did not
// leave them, (x, y, z, m) and the cotangents.
__global__ __launch_bounds__(kB) void k_gather(const float* __restrict__ pts,
                                               const float* __restrict__ ms,
                                               const float* __restrict__ cot, int64_t n,
                                               const uint32_t* order,
                                               const uint32_t* sorted, uint32_t* perm,
                                               uint32_t* codes, float4* __restrict__ sp,
                                               float4* __restrict__ sc) {
  int64_t r = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (r >= n) return;
  const uint32_t o = order[r];
  const int64_t i = o;
  if (perm != order) perm[r] = o;
  if (codes != sorted) codes[r] = sorted[r];
  sp[r] = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], ms != nullptr ? ms[i] : 0.0f);
  if (cot != nullptr) sc[r] = make_float4(cot[3 * i], cot[3 * i + 1], cot[3 * i + 2], 0.0f);
}

// ------------------------------------------------------------------ topology
// One pass over the sorted points (topo.hpp): point r begins the nodes of
// levels first..top and possibly a group cell.  The node index of (l, r) is
// level_start[l] + rank_l(r) (breadth-first, Morton order inside a level).
// A node's END is written by its last point e: e closes the cells of levels
// > dn(e) (the digits e shares with e + 1), and the node at level l <= lambda(e)
// holding e has index level_start[l] + rank_l(e) + begins(e, l) - 1, where
// lambda(e) = the deepest level of a node holding e = top of the last point
// <= e that begins a node (or 0 at a set whose root is a leaf).  The same
// point names the last child of each split node it closes.  Groups: a point
// starts a group iff it lies a multiple of kGroupMax after the first point of
// its group cell, the last group-cell first point <= r.  Kernels:
//   k_topo_count     per tile: level begin counts, group starts after the
//                    tile's first group cell, carries; a packed per-point record
//   k_topo_scan      one block: per-tile level offsets, group offsets and the
//                    carries (last group cell, last lambda) entering each tile;
//                    the level table and the totals
//   k_topo_emit      node records from their first points, node ends and last
//                    children from their last points, group starts
//   k_topo_finalize  counts and child counts of the records, group sizes, and
//                    the moments of every leaf (a row of 16 lanes per node);
//                    k_moments_internal then does the internal levels, one
//                    launch per level, deepest first
constexpr int kTopoTile = 1024;  // one point per thread, one tile per block
constexpr int kTopoWaves = kTopoTile / 64;
constexpr int kTopoCols = topo::kLevels;
// columns of the per-tile counters (1..10: level begins)
constexpr int kColGroups = 0;   // count: group starts after the first group cell; scan: offset
constexpr int kColFirstG = 11;  // count: first group-cell first point in the tile, or -1
constexpr int kColLastG = 12;   // count: last one, or -1; scan: the carry entering the tile
constexpr int kColLamPos = 13;  // count: last lambda source (a node-beginning point or a set start)
constexpr int kColLamTop = 14;  // count: its top; scan: lambda entering the tile (-1: none)

// Per-point record of k_topo_count read back by k_topo_emit:
// first | (top + 1) << 4 | (group level + 1) << 8.
__device__ __forceinline__ uint32_t pack_point(const topo::PointNodes& p, int gl) {
  return static_cast<uint32_t>(p.first) | static_cast<uint32_t>(p.top + 1) << 4 |
         static_cast<uint32_t>(gl + 1) << 8;
}

struct SetRange {
  int set;
  int64_t sb, se;
};

__device__ __forceinline__ SetRange set_range(const int32_t* __restrict__ offs, int nsets,
                                              int64_t r) {
  SetRange q;
  q.set = nsets > 1 ? set_of(offs, nsets, static_cast<int>(r)) : 0;
  q.sb = offs[q.set];
  q.se = offs[q.set + 1];
  return q;
}

// Value of the last thread of the block at or before this one with `has`
// set, else `fallback`.  All threads of the block must call it.
__device__ __forceinline__ int block_last(bool has, int v, int fallback, int* s_w, int lane,
                                          int w) {
  const uint64_t b = __ballot(has);
  const uint64_t le = b & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  // every lane takes part in both shuffles (sources are lanes with `has`)
  const int src = le != 0ull ? 63 - __clzll(static_cast<long long>(le)) : lane;
  const int got = __shfl(v, src, 64);
  int mine = le != 0ull ? got : fallback;
  const int wsrc = b != 0ull ? 63 - __clzll(static_cast<long long>(b)) : 0;
  const int wlast = __shfl(v, wsrc, 64);
  if (lane == 0) s_w[2 * w] = b != 0ull ? 1 : 0;
  if (lane == 0) s_w[2 * w + 1] = wlast;
  __syncthreads();
  if (le == 0ull) {
    for (int k = w - 1; k >= 0; --k)
      if (s_w[2 * k]) {
        mine = s_w[2 * k + 1];
        break;
      }
  }
  __syncthreads();  // s_w is reused by the next call
  return mine;
}

__global__ __launch_bounds__(kTopoTile) void k_topo_count(const Header* __restrict__ hdr,
                                                          int64_t n,
                                                   const uint32_t* __restrict__ codes,
                                                   const int32_t* __restrict__ offs, int nsets,
                                                   int D, int leaf_size,
                                                   int32_t* __restrict__ tiles,
                                                   uint32_t* __restrict__ prec) {
  __shared__ int s_cnt[kTopoWaves][kTopoCols];
  __shared__ int s_w[2 * kTopoWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kTopoTile + threadIdx.x;
  const bool ok = r < n && (hdr->overflow & 2) == 0;
  topo::PointNodes p{0, -1, 0, -1};
  int gl = -1;
  bool sb = false;
  if (ok) {
    const SetRange q = set_range(offs, nsets, r);
    p = topo::point_nodes(codes, r, q.sb, q.se, D, leaf_size);
    gl = topo::group_cell_level(codes, r, q.sb, q.se, D, p);
    sb = r == q.sb;
    prec[r] = pack_point(p, gl);
  }
  const int ri = static_cast<int>(r);
  const bool has_g = ok && gl >= 0;
  const int gs = block_last(has_g, ri, -1, s_w, lane, w);
  const bool gstart = ok && gs >= 0 && ((ri - gs) % kGroupMax) == 0;
  const bool lam_src = ok && (p.first > 0 || sb);
  // tile summary: first / last group cell, last lambda source (thread 255 sees all)
  const int lastg = block_last(has_g, ri, -1, s_w, lane, w);
  const int lampos = block_last(lam_src, ri, -1, s_w, lane, w);
  const int lamtop = block_last(lam_src, p.first > 0 ? p.top : 0, -1, s_w, lane, w);
  const uint64_t bg = __ballot(has_g);
  const int fsrc = bg != 0ull ? __ffsll(static_cast<long long>(bg)) - 1 : 0;
  const int fval = __shfl(ri, fsrc, 64);
  if (lane == 0) s_cnt[w][kColFirstG] = bg != 0ull ? fval : -1;
  const int cg = __popcll(__ballot(gstart));
  if (lane == 0) s_cnt[w][kColGroups] = cg;
  for (int l = 1; l <= 10; ++l) {
    const int c = __popcll(__ballot(ok && l <= D && p.first > 0 && l >= p.first && l <= p.top));
    if (lane == 0) s_cnt[w][l] = c;
  }
  __syncthreads();
  int32_t* out = tiles + static_cast<int64_t>(blockIdx.x) * kTopoCols;
  if (threadIdx.x <= 10) {
    int c = 0;
    for (int k = 0; k < kTopoWaves; ++k) c += s_cnt[k][threadIdx.x];
    out[threadIdx.x] = c;
  }
  if (threadIdx.x == 0) {
    int f = -1;
    for (int k = 0; k < kTopoWaves && f < 0; ++k) f = s_cnt[k][kColFirstG];
    out[kColFirstG] = f;
  }
  if (threadIdx.x == kTopoTile - 1) {
    out[kColLastG] = lastg;
    out[kColLamPos] = lampos;
    out[kColLamTop] = lamtop;
    out[15] = 0;
  }
}

// Group starts in [t0, t1) of a group cell whose first point is g (-1: none).
__device__ __forceinline__ int64_t lead_groups(int64_t g, int64_t t0, int64_t t1) {
  if (g < 0 || t1 <= t0) return 0;
  auto le = [&](int64_t x) { return x < g ? int64_t{0} : (x - g) / kGroupMax + 1; };
  return le(t1 - 1) - le(t0 - 1);
}

// Exclusive block scan (sum, or max with identity -1) of one value per
// thread of a 1024-thread block; the block total (maximum) through *all.
template <typename T, bool SUM>
__device__ __forceinline__ T block_scan1(T v, T* s_w, T* all) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T t = __shfl_up(incl, d, 64);
    if (lane >= d) incl = SUM ? incl + t : (incl > t ? incl : t);
  }
  T ex = __shfl_up(incl, 1, 64);
  if (lane == 0) ex = SUM ? T(0) : T(-1);
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  T before = SUM ? T(0) : T(-1), tot = SUM ? T(0) : T(-1);
  for (int j = 0; j < 16; ++j) {
    const T x = s_w[j];
    if (SUM) {
      if (j < w) before += x;
      tot += x;
    } else {
      if (j < w) before = before > x ? before : x;
      tot = tot > x ? tot : x;
    }
  }
  __syncthreads();
  *all = tot;
  return SUM ? before + ex : (before > ex ? before : ex);
}

__global__ __launch_bounds__(1024) void k_topo_scan(Header* __restrict__ hdr,
                                                    int32_t* __restrict__ tiles, int64_t ntiles,
                                                    int64_t n, int nsets, int D, int64_t cap) {
  __shared__ int32_t s_w32[16];
  __shared__ int64_t s_w64[16];
  __shared__ int32_t s_tot[11];
  __shared__ int32_t s_lvl[10][1024];  // exclusive level offsets entering each chunk
  const int tid = threadIdx.x;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t t0 = tid * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
#pragma unroll 1
  for (int l = 0; l < 10; ++l) {
    int32_t sum = 0;
    for (int64_t t = t0; t < t1; ++t) sum += tiles[t * kTopoCols + 1 + l];
    int32_t tot;
    s_lvl[l][tid] = block_scan1<int32_t, true>(sum, s_w32, &tot);
    if (tid == 0) s_tot[l] = tot;
  }
  // carries entering the chunk: the last group cell (positions rise, so a
  // max) and the last lambda source as (position << 4 | top)
  int64_t cg = -1, cl = -1;
  for (int64_t t = t0; t < t1; ++t) {
    const int32_t* c = tiles + t * kTopoCols;
    if (c[kColLastG] >= 0) cg = c[kColLastG];
    if (c[kColLamPos] >= 0) cl = (static_cast<int64_t>(c[kColLamPos]) << 4) | c[kColLamTop];
  }
  int64_t dummy;
  const int64_t g_in = block_scan1<int64_t, false>(cg, s_w64, &dummy);
  const int64_t l_in = block_scan1<int64_t, false>(cl, s_w64, &dummy);
  // group starts per chunk (the leading ones need the carried cell)
  int32_t gsum = 0;
  {
    int64_t g = g_in;
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t* c = tiles + t * kTopoCols;
      const int64_t x = t * kTopoTile, e = x + kTopoTile < n ? x + kTopoTile : n;
      gsum += static_cast<int32_t>(lead_groups(g, x, c[kColFirstG] >= 0 ? c[kColFirstG] : e)) +
              c[kColGroups];
      if (c[kColLastG] >= 0) g = c[kColLastG];
    }
  }
  int32_t gtot;
  int32_t goff = block_scan1<int32_t, true>(gsum, s_w32, &gtot);
  // per-tile outputs
  {
    int64_t g = g_in;
    int32_t lam = l_in >= 0 ? static_cast<int32_t>(l_in & 15) : -1;
    for (int64_t t = t0; t < t1; ++t) {
      int32_t* c = tiles + t * kTopoCols;
      const int64_t x = t * kTopoTile, e = x + kTopoTile < n ? x + kTopoTile : n;
      const int32_t first_g = c[kColFirstG], last_g = c[kColLastG];
      const int32_t lam_pos = c[kColLamPos], lam_top = c[kColLamTop];
      const int32_t here =
          static_cast<int32_t>(lead_groups(g, x, first_g >= 0 ? first_g : e)) + c[kColGroups];
      c[kColGroups] = goff;
      goff += here;
      for (int l = 0; l < 10; ++l) {
        const int32_t v = c[1 + l];
        c[1 + l] = s_lvl[l][tid];
        s_lvl[l][tid] += v;
      }
      c[kColLastG] = static_cast<int32_t>(g);
      c[kColLamTop] = lam;
      if (last_g >= 0) g = last_g;
      if (lam_pos >= 0) lam = lam_top;
    }
  }
  if (tid == 0) {
    int64_t start = nsets;
    for (int l = 1; l < kMaxLevels; ++l) {
      const int64_t cnt = l <= D && l <= 10 ? s_tot[l - 1] : 0;
      hdr->level_start[l] = static_cast<int32_t>(start);
      hdr->level_count[l] = static_cast<int32_t>(cnt);
      start += cnt;
    }
    if (start > cap) {
      for (int l = 1; l < kMaxLevels; ++l) {
        hdr->level_start[l] = nsets;
        hdr->level_count[l] = 0;
      }
      hdr->overflow |= 1;
      start = nsets;
    }
    hdr->num_nodes = static_cast<int32_t>(start);
    hdr->num_groups = gtot;
    hdr->queue = 0;
    for (int x = 0; x < 8; ++x) hdr->xq[x] = 0;
  }
}

__global__ __launch_bounds__(kTopoTile) void k_topo_emit(
    const Header* __restrict__ hdr, int64_t n, const uint32_t* __restrict__ codes,
    const int32_t* __restrict__ offs, int nsets, int D, const int32_t* __restrict__ tiles,
    const uint32_t* __restrict__ prec, const CubeGeom* __restrict__ geoms,
    int4* __restrict__ topo, uint32_t* __restrict__ prefix, float4* __restrict__ geo,
    int32_t* __restrict__ parent, int2* __restrict__ ends, int2* __restrict__ groups) {
  __shared__ int s_cnt[kTopoWaves][kTopoCols];
  __shared__ int s_w[2 * kTopoWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kTopoTile + threadIdx.x;
  const bool nodes_ok = (hdr->overflow & 3) == 0;
  const bool ok = r < n && (hdr->overflow & 2) == 0;
  const int32_t* base = tiles + static_cast<int64_t>(blockIdx.x) * kTopoCols;
  topo::PointNodes p{0, -1, 0, -1};
  int gl = -1;
  SetRange q{0, 0, 0};
  if (ok) {
    const uint32_t v = prec[r];
    p.first = static_cast<int>(v & 15u);
    p.top = static_cast<int>((v >> 4) & 15u) - 1;
    gl = static_cast<int>((v >> 8) & 15u) - 1;
    q = set_range(offs, nsets, r);
  }
  const int ri = static_cast<int>(r);
  const bool sb = ok && r == q.sb;
  const int gs = block_last(ok && gl >= 0, ri, base[kColLastG], s_w, lane, w);
  const int lam = block_last(ok && (p.first > 0 || sb), p.first > 0 ? p.top : 0,
                             base[kColLamTop], s_w, lane, w);
  const bool gstart = ok && gs >= 0 && ((ri - gs) % kGroupMax) == 0;
  int rank[kTopoCols];
  {
    const uint64_t b = __ballot(gstart);
    rank[0] = __popcll(b & lt);
    if (lane == 0) s_cnt[w][0] = __popcll(b);
  }
  for (int l = 1; l <= 10; ++l) {
    const uint64_t b = __ballot(ok && l <= D && p.first > 0 && l >= p.first && l <= p.top);
    rank[l] = __popcll(b & lt);
    if (lane == 0) s_cnt[w][l] = __popcll(b);
  }
  __syncthreads();
  if (!ok) return;
  auto rank_of = [&](int l) {
    int v = base[l] + rank[l];
    for (int k = 0; k < w; ++k) v += s_cnt[k][l];
    return v;
  };
  auto begins = [&](int l) { return p.first > 0 && l >= p.first && l <= p.top ? 1 : 0; };
  if (gstart) groups[rank_of(0)] = make_int2(ri, 0);
  if (!nodes_ok) return;
  // records of the nodes r begins
  if (sb && p.first == 1) topo[q.set].z = hdr->level_start[1] + rank_of(1);
  if (p.first > 0) {
    const CubeGeom gm = geoms[q.set];
    for (int l = p.first; l <= p.top; ++l) {
      const int idx = hdr->level_start[l] + rank_of(l);
      const uint32_t pre = codes[r] >> (3 * (D - l));
      const int fc = l < p.top ? hdr->level_start[l + 1] + rank_of(l + 1) : -1;
      topo[idx] = make_int4(ri, 0, fc, 0);
      prefix[idx] = pre;
      parent[idx] = l == 1 ? q.set
                           : hdr->level_start[l - 1] + rank_of(l - 1) - (l - 1 >= p.first ? 0 : 1);
      geo[idx] = cell_geometry(gm, pre, l);
    }
  }
  // ends and last children of the nodes r closes
  const int dn = r + 1 < q.se ? topo::common_digits(codes[r], codes[r + 1], D) : -1;
  for (int l = dn + 1 > 1 ? dn + 1 : 1; l <= lam; ++l) {
    const int idx = hdr->level_start[l] + rank_of(l) + begins(l) - 1;
    ends[idx].x = ri + 1;
    if (l < lam) ends[idx].y = hdr->level_start[l + 1] + rank_of(l + 1) + begins(l + 1) - 1;
  }
  if (r + 1 == q.se && lam >= 1) ends[q.set].y = hdr->level_start[1] + rank_of(1) + begins(1) - 1;
}

// Sum over a row of 16 lanes (every lane of the row active), by DPP moves:
// partners l ^ 8 (row rotate by 8), 7 - l within each half row, then the
// quad permutations l ^ 2 and l ^ 1; each stage adds a disjoint set of lanes.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_sum16(float v) {
  v += dpp_move<0x128>(v);  // row_ror:8
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  return v;
}

// Moments of node k by a row of 16 lanes (sl = lane within the row): a leaf
// sums its points 16 at a time (mass and mass-weighted position, then the
// central second moment about the centre of mass: two passes, as the FP64
// oracle), an internal node its <= 8 children (parallel-axis shift of their
// second moments).  The row sums in a fixed order: deterministic.
__device__ __forceinline__ void node_moments(int k, int4 t, int sl,
                                             const float4* __restrict__ geo,
                                             const float4* __restrict__ sp,
                                             float4* __restrict__ mp, float4* __restrict__ qd,
                                             const float4* __restrict__ sv,
                                             float4* __restrict__ wd) {
  if (wd != nullptr) {
    // dipole sum m v of the node (winding numbers: areas times unit normals)
    float vx = 0.f, vy = 0.f, vz = 0.f;
    if (t.w == 0) {
      for (int j = t.x + sl; j < t.x + t.y; j += 16) {
        const float a = sp[j].w;
        const float4 nv = sv[j];
        vx = fmaf(a, nv.x, vx);
        vy = fmaf(a, nv.y, vy);
        vz = fmaf(a, nv.z, vz);
      }
    } else if (sl < t.w) {
      const float4 dc = wd[t.z + sl];
      vx = dc.x;
      vy = dc.y;
      vz = dc.z;
    }
    vx = row_sum16(vx);
    vy = row_sum16(vy);
    vz = row_sum16(vz);
    if (sl == 0) wd[k] = make_float4(vx, vy, vz, 0.f);
  }
  float M = 0.f, X = 0.f, Y = 0.f, Z = 0.f;
  if (t.w == 0) {
    for (int j = t.x + sl; j < t.x + t.y; j += 16) {
      const float4 p = sp[j];
      M += p.w;
      X = fmaf(p.w, p.x, X);
      Y = fmaf(p.w, p.y, Y);
      Z = fmaf(p.w, p.z, Z);
    }
  } else if (sl < t.w) {
    const float4 m = mp[t.z + sl];
    M = m.w;
    X = m.w * m.x;
    Y = m.w * m.y;
    Z = m.w * m.z;
  }
  M = row_sum16(M);
  X = row_sum16(X);
  Y = row_sum16(Y);
  Z = row_sum16(Z);
  if (M != 0.f) {
    X /= M;
    Y /= M;
    Z /= M;
  } else {
    const float4 gg = geo[k];
    X = gg.x;
    Y = gg.y;
    Z = gg.z;
  }
  float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
  if (t.w == 0) {
    for (int j = t.x + sl; j < t.x + t.y; j += 16) {
      const float4 p = sp[j];
      const float dx = p.x - X, dy = p.y - Y, dz = p.z - Z;
      const float wx = p.w * dx, wy = p.w * dy;
      sxx = fmaf(wx, dx, sxx);
      syy = fmaf(wy, dy, syy);
      szz = fmaf(p.w * dz, dz, szz);
      sxy = fmaf(wx, dy, sxy);
      sxz = fmaf(wx, dz, sxz);
      syz = fmaf(wy, dz, syz);
    }
  } else if (sl < t.w) {
    const int c = t.z + sl;
    const float4 m = mp[c];
    const float4 q0 = qd[2 * c], q1 = qd[2 * c + 1];
    const float dx = m.x - X, dy = m.y - Y, dz = m.z - Z;
    const float wx = m.w * dx, wy = m.w * dy;
    sxx = fmaf(wx, dx, q0.x);
    syy = fmaf(wy, dy, q0.y);
    szz = fmaf(m.w * dz, dz, q0.z);
    sxy = fmaf(wx, dy, q0.w);
    sxz = fmaf(wx, dz, q1.x);
    syz = fmaf(wy, dz, q1.y);
  }
  sxx = row_sum16(sxx);
  syy = row_sum16(syy);
  szz = row_sum16(szz);
  sxy = row_sum16(sxy);
  sxz = row_sum16(sxz);
  syz = row_sum16(syz);
  if (M == 0.f) sxx = syy = szz = sxy = sxz = syz = 0.f;
  if (sl == 0) {
    mp[k] = make_float4(X, Y, Z, M);
    qd[2 * k] = make_float4(sxx, syy, szz, sxy);
    // + 1.5 tr(S) and -2.5 tr(S), constants of the walk's far-field terms
    const float tr = sxx + syy + szz;
    qd[2 * k + 1] = make_float4(sxz, syz, 1.5f * tr, -2.5f * tr);
  }
}

// Counts and child counts of every record, group sizes (groups tile the
// sorted points), and with `moments` the moments of every leaf (a row of 16
// lanes per node).
__global__ __launch_bounds__(kB) void k_topo_finalize(const Header* __restrict__ hdr, int64_t n,
                                                      int nsets, int moments,
                                                      int4* __restrict__ topo,
                                                      const int2* __restrict__ ends,
                                                      int2* __restrict__ groups,
                                                      const float4* __restrict__ geo,
                                                      const float4* __restrict__ sp,
                                                      float4* __restrict__ mp,
                                                      float4* __restrict__ qd,
                                                      const float4* __restrict__ sv,
                                                      float4* __restrict__ wd, int sentinel) {
  if (wd != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
    wd[sentinel] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int ng = hdr->num_groups;
  for (int g = blockIdx.x * kB + threadIdx.x; g < ng; g += gridDim.x * kB) {
    const int s0 = groups[g].x;
    const int s1 = g + 1 < ng ? groups[g + 1].x : static_cast<int>(n);
    groups[g].y = s1 - s0;
  }
  const int sl = threadIdx.x & 15;
  const int row = (blockIdx.x * kB + threadIdx.x) >> 4;
  const int nrows = (gridDim.x * kB) >> 4;
  const int nn = hdr->num_nodes;
  for (int k = row; k < nn; k += nrows) {
    int4 t = topo[k];
    const int2 e = ends[k];
    if (k >= nsets) t.y = e.x - t.x;
    t.w = t.z >= 0 ? e.y - t.z + 1 : 0;
    if (sl == 0) {
      topo[k].y = t.y;
      topo[k].w = t.w;
    }
    if (moments && t.w == 0) node_moments(k, t, sl, geo, sp, mp, qd, sv, wd);
  }
}

// Moments of the internal nodes of one level (children done).
__global__ __launch_bounds__(kB) void k_moments_internal(const Header* __restrict__ hdr,
                                                         int level, const int4* __restrict__ topo,
                                                         const float4* __restrict__ geo,
                                                         const float4* __restrict__ sp,
                                                         float4* __restrict__ mp,
                                                         float4* __restrict__ qd,
                                                         const float4* __restrict__ sv,
                                                         float4* __restrict__ wd) {
  const int sl = threadIdx.x & 15;
  const int row = (blockIdx.x * kB + threadIdx.x) >> 4;
  const int nrows = (gridDim.x * kB) >> 4;
  const int start = hdr->level_start[level];
  const int count = hdr->level_count[level];
  for (int u = row; u < count; u += nrows) {
    const int k = start + u;
    const int4 t = topo[k];
    if (t.w > 0) node_moments(k, t, sl, geo, sp, mp, qd, sv, wd);
  }
}

// ------------------------------------------------------------------ wave helpers
// Partner value across lanes by a DPP move (no LDS crossbar):
//   kRowRor8     lane l <- (l + 8) mod 16 within its row of 16, i.e. l ^ 8
//   kHalfMirror  lane l <- 7 - l within its half-row of 8 (partner bit 2 flipped)
//   kQuadXor2 / kQuadXor1  quad permutations l ^ 2 and l ^ 1
constexpr int kRowRor8 = 0x128;
constexpr int kHalfMirror = 0x141;
constexpr int kQuadXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kQuadXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]

template <int CTRL>
__device__ __forceinline__ float dpp_partner(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Wave sums through LDS, used by the adjoint batches.  The V values of each
// lane are written to the wave's slab as V rows of 64 (one float per lane,
// rows kRedRow floats apart); 64 / V lanes per value then each add a run of
// V consecutive floats of that value's row, and DPP adds combine the runs.
// Lane l ends with the sum of value l / (64 / V) (all lanes of its run hold
// it).  A value costs about one VALU add; a transposed butterfly in
// registers (permlane swaps, selects and DPP adds) takes about two VALU
// operations per value, and the walk is VALU bound, so the LDS pipe does
// the transposition.  The slab is private to the wave, and the LDS
// operations of one wave complete in issue order, so passes may reuse it.
constexpr int kRedRow = 68;  // 64 lanes + 4 floats: staggers the runs over the banks
constexpr int kRedMax = 16;  // values per pass (the slab holds kRedMax rows)

template <int V>
__device__ __forceinline__ float lds_wave_sum(const float* v, float* red, int lane) {
  static_assert(V == 4 || V == 8 || V == 16, "unsupported width");
  constexpr int kLanes = 64 / V;  // lanes per value
#pragma unroll
  for (int k = 0; k < V; ++k) red[k * kRedRow + lane] = v[k];
  const float4* p =
      reinterpret_cast<const float4*>(red + (lane / kLanes) * kRedRow + (lane % kLanes) * V);
  float r[V / 4];
#pragma unroll
  for (int k = 0; k < V / 4; ++k) {
    const float4 q = p[k];
    r[k] = (q.x + q.y) + (q.z + q.w);
  }
#pragma unroll
  for (int h = V / 8; h >= 1; h /= 2)
#pragma unroll
    for (int k = 0; k < h; ++k) r[k] += r[k + h];
  float sum = r[0];
  sum += dpp_partner<kQuadXor1>(sum);
  sum += dpp_partner<kQuadXor2>(sum);
  if constexpr (kLanes >= 8) sum += dpp_partner<kHalfMirror>(sum);
  if constexpr (kLanes >= 16) sum += dpp_partner<kRowRor8>(sum);
  return sum;
}

struct WalkArgs {
  Header* hdr;
  const int2* groups;
  const int4* topo;
  const float4* geo;
  const float4* mp;
  const float4* qd;
  const float4* sp;
  const float4* sc;
  const uint32_t* perm;
  const int32_t* offs;  // set offsets (sorted order), nsets + 1 entries
  int nsets;
  int n;
  int sentinel;  // zero-mass node used to pad far batches
  float beta;
  float e2;  // eps^2 + 1e-12
  float* forces;
  float4* qgrad;
  float* sgrad;  // 4 floats per sorted point (atomics)
  float* adj;    // 4 floats per node: dL/dX (3), dL/dM
  float* adjq;   // 8 floats per node: dL/dS (6 used)
  unsigned long long* stats;
  const int* order;  // claim order of the groups (k_group_order)
  float* gcost;      // measured cost of each group, for the next walk's order
};

struct Lane {
  float x, y, z;     // query position
  float cx, cy, cz;  // cotangent
  float c3x, c3y, c3z;  // 3 x cotangent (adjoint): the 3 of 3 s^2 (c.D) comes free
  float* red;      // the wave's LDS slab for the adjoint sums (kRedMax rows)
};

// Adds the wave-summed batch partials v[0..V) (PER values per entry, the
// first USED of them meaningful; negated when NEG) to dst: value j goes to
// dst[stride * e + j % PER] with e = entry[j / PER] (or base + j / PER when
// entry is null); e == skip and e >= limit are dropped.  THIRD: value 3 of
// each entry arrives tripled and is scaled by 1/3.  The sums run through the
// wave's slab in passes of up to kRedMax values; one lane per value issues
// the atomic.
template <int V, int PER, int USED, bool NEG, bool THIRD = false>
__device__ __forceinline__ void scatter_batch(const float (&v)[V], float* red, int lane,
                                              float* dst, int stride, const int* entry, int skip,
                                              int base, int limit) {
  constexpr int P = V < kRedMax ? V : kRedMax;  // values per pass
  constexpr int kLanes = 64 / P;
#pragma unroll
  for (int c = 0; c < V; c += P) {
    const float sum = lds_wave_sum<P>(v + c, red, lane);
    const int j = c + lane / kLanes;
    const float scale = (THIRD && j % PER == 3 ? 1.0f / 3.0f : 1.0f) * (NEG ? -1.0f : 1.0f);
    if (lane % kLanes == 0 && (j % PER) < USED) {
      const int e = entry != nullptr ? entry[j / PER] : base + j / PER;
      if (e != skip && e < limit) atomicAdd(&dst[stride * e + (j % PER)], scale * sum);
    }
  }
}

// Far-field batch of C nodes (sentinel-padded): forces (if FORCES),
// query-side gradient and the node adjoint partials (if VJP).  ORDER: 1
// monopole, 2 monopole + quadrupole, 3 (adjoint only) monopole terms plus
// the quadrupole term's dL/dS, which does not depend on S: pushed down to
// the sources it adds the first-order term of their position gradients and
// the second-order term of their mass gradients.
template <bool VJP, int ORDER, bool FORCES, int C>
__device__ __forceinline__ void far_batch(const WalkArgs& A, const Lane& L, const int* far,
                                          int lane, float& fx, float& fy, float& fz, float& gx,
                                          float& gy, float& gz) {
  constexpr bool QUAD = ORDER == 2;
  constexpr bool SMOM = VJP && ORDER == 3;  // second-moment adjoint with monopole terms
  int idx[C];
#pragma unroll
  for (int t = 0; t < C; ++t) idx[t] = __builtin_amdgcn_readfirstlane(far[t]);
  float part[4 * C];   // minus the node-side partials (negated in the scatter)
  // quadrupole adjoint, 10 values per node packed as 8 + 2: per node
  // (dL/dS xx yy zz xy xz yz, minus dL/dM, minus dL/dX x) and (minus dL/dX y, z)
  float q8[8 * C];
  float q2[2 * C];
  float mw[C];  // node masses (second-moment adjoint)
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const float4 m = sload4(A.mp, idx[t]);
    const float Dx = L.x - m.x, Dy = L.y - m.y, Dz = L.z - m.z;
    const float s = rsq(fmaf(Dz, Dz, fmaf(Dy, Dy, fmaf(Dx, Dx, A.e2))));
    const float s2 = s * s, s3 = s2 * s;
    if constexpr (SMOM) {
      // d(c.F)/dD = M s^3 e with e = 3 s^2 u D - c; the lanes' s^3 e are
      // summed and scaled by the node's M in the scatter.  dL/dS as in the
      // quadrupole branch below, doubled: with p = s^5 (c - 5 e), whose half
      // is P2 there, the diagonal is p_a D_a + 3 s^5 u and the off-diagonal
      // p_a D_b + D_a p_b (halved in the scatter).
      if constexpr (FORCES) {
        const float a = -m.w * s3;
        fx = fmaf(a, Dx, fx);
        fy = fmaf(a, Dy, fy);
        fz = fmaf(a, Dz, fz);
      }
      mw[t] = m.w;
      const float u3 = L.c3x * Dx + L.c3y * Dy + L.c3z * Dz;  // 3 u
      const float h = s2 * u3;
      const float ex = fmaf(h, Dx, -L.cx), ey = fmaf(h, Dy, -L.cy), ez = fmaf(h, Dz, -L.cz);
      const float sx = s3 * ex, sy = s3 * ey, sz = s3 * ez;
      gx = fmaf(m.w, sx, gx);
      gy = fmaf(m.w, sy, gy);
      gz = fmaf(m.w, sz, gz);
      const float s5 = s3 * s2, w5 = s3 * h;  // s^5, 3 s^5 u
      const float px = s5 * fmaf(-5.0f, ex, L.cx);
      const float py = s5 * fmaf(-5.0f, ey, L.cy);
      const float pz = s5 * fmaf(-5.0f, ez, L.cz);
      q8[8 * t + 0] = fmaf(px, Dx, w5);
      q8[8 * t + 1] = fmaf(py, Dy, w5);
      q8[8 * t + 2] = fmaf(pz, Dz, w5);
      q8[8 * t + 3] = fmaf(px, Dy, Dx * py);
      q8[8 * t + 4] = fmaf(px, Dz, Dx * pz);
      q8[8 * t + 5] = fmaf(py, Dz, Dy * pz);
      q8[8 * t + 6] = s3 * u3;  // scaled by 1/3 in the scatter
      q8[8 * t + 7] = sx;
      q2[2 * t + 0] = sy;
      q2[2 * t + 1] = sz;
    } else if constexpr (!QUAD) {
      const float a = -m.w * s3;
      if constexpr (FORCES) {
        fx = fmaf(a, Dx, fx);
        fy = fmaf(a, Dy, fy);
        fz = fmaf(a, Dz, fz);
      }
      if constexpr (VJP) {
        const float u3 = L.c3x * Dx + L.c3y * Dy + L.c3z * Dz;  // 3 u
        const float kD = -a * s2 * u3;  // 3 M s^5 u
        const float dDx = kD * Dx + a * L.cx;
        const float dDy = kD * Dy + a * L.cy;
        const float dDz = kD * Dz + a * L.cz;
        gx += dDx;
        gy += dDy;
        gz += dDz;
        part[4 * t + 0] = dDx;
        part[4 * t + 1] = dDy;
        part[4 * t + 2] = dDz;
        part[4 * t + 3] = s3 * u3;  // scaled by 1/3 in the scatter
      }
    } else {
      const float4 q0 = sload4(A.qd, 2 * idx[t]);
      const float4 q1 = sload4(A.qd, 2 * idx[t] + 1);
      // powers of s through w = s^2; second moment (xx, yy, zz, xy), (xz, yz)
      const float w = s2, s5 = s3 * w;
      const float SDx = q0.x * Dx + q0.w * Dy + q1.x * Dz;
      const float SDy = q0.w * Dx + q0.y * Dy + q1.y * Dz;
      const float SDz = q1.x * Dx + q1.y * Dy + q0.z * Dz;
      const float tq = Dx * SDx + Dy * SDy + Dz * SDz;
      const float wtq = w * tq;
      // a = -M s^3 + 1.5 tr(S) s^5 - 7.5 s^7 t,  b = 3 s^5
      const float a = s3 * fmaf(w, fmaf(-7.5f, wtq, q1.z), -m.w);
      const float b = 3.0f * s5;
      if constexpr (FORCES) {
        fx = fmaf(a, Dx, fmaf(b, SDx, fx));
        fy = fmaf(a, Dy, fmaf(b, SDy, fy));
        fz = fmaf(a, Dz, fmaf(b, SDz, fz));
      }
      if constexpr (VJP) {
        const float u = L.cx * Dx + L.cy * Dy + L.cz * Dz;
        const float Scx = q0.x * L.cx + q0.w * L.cy + q1.x * L.cz;
        const float Scy = q0.w * L.cx + q0.y * L.cy + q1.y * L.cz;
        const float Scz = q1.x * L.cx + q1.y * L.cy + q0.z * L.cz;
        const float cw = Scx * Dx + Scy * Dy + Scz * Dz;  // c . S D
        // d(c.F)/dD = kD D + a c + b S c + b15 S D with
        //   kD = u (3 M s^5 - 5 (1.5 tr S) s^7 + 52.5 s^9 t) - 15 s^7 (c.S D),
        //   b15 = -15 s^7 u; the bracket is b (M + w (17.5 w t - 2.5 tr S))
        const float A1 = b * fmaf(w, fmaf(17.5f, wtq, q1.w), m.w);
        const float m15s7 = -15.0f * (s5 * w);
        const float kD = fmaf(m15s7, cw, u * A1);
        const float b15 = m15s7 * u;
        const float dDx = kD * Dx + a * L.cx + b * Scx + b15 * SDx;
        const float dDy = kD * Dy + a * L.cy + b * Scy + b15 * SDy;
        const float dDz = kD * Dz + a * L.cz + b * Scz + b15 * SDz;
        gx += dDx;
        gy += dDy;
        gz += dDz;
        q8[8 * t + 6] = s3 * u;
        q8[8 * t + 7] = dDx;
        q2[2 * t + 0] = dDy;
        q2[2 * t + 1] = dDz;
        // dL/dS (symmetric) = P D^T + D P^T + 1.5 s^5 u I with
        //   P = 1.5 s^5 c - 3.75 s^7 u D = (b c + (b15 / 2) D) / 2 =: P2 / 2;
        //   diagonal: P2_a D_a + 1.5 s^5 u, off-diagonal stored doubled:
        //   P2_a D_b + D_a P2_b (halved in k_finalize)
        const float hb15 = 0.5f * b15, hu = (0.5f * b) * u;
        const float Px = fmaf(hb15, Dx, b * L.cx);
        const float Py = fmaf(hb15, Dy, b * L.cy);
        const float Pz = fmaf(hb15, Dz, b * L.cz);
        q8[8 * t + 0] = fmaf(Px, Dx, hu);
        q8[8 * t + 1] = fmaf(Py, Dy, hu);
        q8[8 * t + 2] = fmaf(Pz, Dz, hu);
        q8[8 * t + 3] = fmaf(Px, Dy, Dx * Py);
        q8[8 * t + 4] = fmaf(Px, Dz, Dx * Pz);
        q8[8 * t + 5] = fmaf(Py, Dz, Dy * Pz);
      }
    }
  }
  if constexpr (VJP && !QUAD && !SMOM) {
    // node adjoints (dL/dX, dL/dM)
    scatter_batch<4 * C, 4, 4, true, true>(part, L.red, lane, A.adj, 4, far, A.sentinel, 0,
                                           INT_MAX);
  }
  if constexpr (VJP && (QUAD || SMOM)) {
    static_assert(C == 4, "the packed quadrupole adjoint assumes 4 nodes per batch");
    // the second-moment adjoint's dL/dX partials are scaled by the node's
    // mass, its dL/dS partials (doubled) by 1/2 and its dL/dM partials
    // (tripled) by 1/3
    constexpr float sq = SMOM ? 0.5f : 1.0f;
    constexpr float s6 = SMOM ? -1.0f / 3.0f : -1.0f;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      // q8 of nodes 2c, 2c + 1: value j = lane >> 2 of the pass is slot j & 7
      // of node 2c + (j >> 3)
      const float sum = lds_wave_sum<16>(q8 + 16 * c, L.red, lane);
      const int j = lane >> 2, slot = j & 7;
      const float mx = SMOM ? -(lane < 32 ? mw[2 * c] : mw[2 * c + 1]) : -1.0f;
      if ((lane & 3) == 0) {
        const int e = far[2 * c + (j >> 3)];
        if (e != A.sentinel) {
          float* dst = slot < 6 ? A.adjq + 8 * e + slot : A.adj + 4 * e + (slot == 6 ? 3 : 0);
          atomicAdd(dst, sum * (slot < 6 ? sq : slot == 6 ? s6 : mx));
        }
      }
    }
    {
      // q2: value j = lane >> 3 is component 1 + (j & 1) of node j >> 1's dL/dX
      const float sum = lds_wave_sum<8>(q2, L.red, lane);
      const int j = lane >> 3;
      float mx = -1.0f;
      if constexpr (SMOM)
        mx = -(lane < 32 ? (lane < 16 ? mw[0] : mw[1]) : (lane < 48 ? mw[2] : mw[3]));
      if ((lane & 7) == 0) {
        const int e = far[j >> 1];
        if (e != A.sentinel) atomicAdd(A.adj + 4 * e + 1 + (j & 1), mx * sum);
      }
    }
  }
}

// Near-field batch: CH consecutive sources starting at j0 (entries at or
// beyond s1 carry zero weight).
// SELF: the batch may hold the lane's own point. Batches of a range that
// lies wholly outside the group's points pass false and skip that test.
// PAD: the batch may run past s1; batches that fit read CH consecutive
// records with no clamp.
template <bool VJP, bool FORCES, int CH, bool SELF = true, bool PAD = false>
__device__ __forceinline__ void near_batch(const WalkArgs& A, const Lane& L, int i, int j0,
                                           int s1, int lane, float& fx, float& fy, float& fz,
                                           float& gx, float& gy, float& gz) {
  float part[4 * CH];  // minus the source-side partials (negated in the scatter)
  cfloat* run = run_base(A.sp, j0);
#pragma unroll
  for (int u = 0; u < CH; ++u) {
    const int j = j0 + u;
    // the adjoint reads runs of 4 records with one wide load (its batches
    // are long enough to hide the wait); the forward walk's own schedule of
    // single-record loads overlaps them better with its shorter batches
    const float4 src = PAD    ? sload4(A.sp, min(j, s1 - 1))
                       : VJP ? sload4_run(run, u)
                             : sload4(A.sp, j);
    // zero weight past the range end; j, s1 and the record are wave-uniform,
    // so the mask stays on the scalar unit
    const float w =
        PAD ? __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(src.w) &
                                                             (j < s1 ? ~0u : 0u)))
            : src.w;
    const float dx = L.x - src.x, dy = L.y - src.y, dz = L.z - src.z;
    const float s = rsq(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, A.e2))));
    const float s2 = s * s, s3 = s * s2;
    float ws3 = w * s3;  // w s^3
    // the self pair: its force term vanishes with d = 0, its adjoint terms do not
    if constexpr (VJP && SELF) ws3 = (j == i) ? 0.0f : ws3;
    if constexpr (FORCES) {
      fx = fmaf(-ws3, dx, fx);
      fy = fmaf(-ws3, dy, fy);
      fz = fmaf(-ws3, dz, fz);
    }
    if constexpr (VJP) {
      const float cd3 = L.c3x * dx + L.c3y * dy + L.c3z * dz;  // 3 (c.d)
      // v = 3 w s^5 (c.d) d - w s^3 c = w s^3 (3 s^2 (c.d) d - c)
      const float k = s2 * cd3;
      const float vx = ws3 * fmaf(k, dx, -L.cx);
      const float vy = ws3 * fmaf(k, dy, -L.cy);
      const float vz = ws3 * fmaf(k, dz, -L.cz);
      gx += vx;
      gy += vy;
      gz += vz;
      part[4 * u + 0] = vx;
      part[4 * u + 1] = vy;
      part[4 * u + 2] = vz;
      part[4 * u + 3] = s3 * cd3;  // 3 s^3 (c.d); exactly 0 for the self pair (d = 0)
    }
  }
  if constexpr (VJP)
    scatter_batch<4 * CH, 4, 4, true, true>(part, L.red, lane, A.sgrad, 4, nullptr, -1, j0, s1);
}

// One query group (sorted points [base, base + cnt)) processed by one wave.
template <bool VJP, int ORDER, bool FORCES>
__device__ __forceinline__ void walk_group(const WalkArgs& A, int base, int cnt, int lane,
                                           int* stack, int* far, int* near, float* red,
                                           unsigned long long& n_far,
                                           unsigned long long& n_near) {
  const int i = base + lane;
  const bool valid = lane < cnt;
  Lane L;
  L.red = red;
  {
    const float4 me = A.sp[valid ? i : base];
    L.x = me.x;
    L.y = me.y;
    L.z = me.z;
    L.cx = L.cy = L.cz = 0.f;
    if constexpr (VJP) {
      if (valid) {
        const float4 c = A.sc[i];
        L.cx = c.x;
        L.cy = c.y;
        L.cz = c.z;
      }
    }
    L.c3x = 3.0f * L.cx;
    L.c3y = 3.0f * L.cy;
    L.c3z = 3.0f * L.cz;
  }
  const Box box = wave_box(L.x, L.y, L.z);
  const int nvalid = cnt;

  float fx = 0.f, fy = 0.f, fz = 0.f;
  float gx = 0.f, gy = 0.f, gz = 0.f;

  // The quadrupole and second-moment adjoints hold 10 partials per node:
  // batches of 4 nodes, not 8 (8 spill).
  constexpr int C = (VJP && ORDER != 1) ? kChunk / 2 : kChunk;
  // The monopole forward walk reads 16 nodes / sources per batch while the
  // loads of a batch are in flight together (its registers allow it); the
  // rest of a list or range is done in batches of 8 as the others.
  constexpr bool kWide = !VJP && ORDER == 1;
  auto flush_far = [&](int nfar) {
    n_far += static_cast<unsigned long long>(nfar) * nvalid;
    const int padded = (nfar + C - 1) & ~(C - 1);
    if (lane < padded - nfar) far[nfar + lane] = A.sentinel;
    __builtin_amdgcn_wave_barrier();
    int f0 = 0;
    if constexpr (kWide) {
      for (; f0 + 2 * C <= padded; f0 += 2 * C)
        far_batch<VJP, ORDER, FORCES, 2 * C>(A, L, far + f0, lane, fx, fy, fz, gx, gy, gz);
    }
    for (; f0 < padded; f0 += C)
      far_batch<VJP, ORDER, FORCES, C>(A, L, far + f0, lane, fx, fy, fz, gx, gy, gz);
    __builtin_amdgcn_wave_barrier();
  };

  // Near leaves whose points directly follow each other in sorted order
  // (sibling leaves, appended in order by the traversal) are merged into one
  // source range, so only a range end pads its last batch of kChunk.
  auto flush_near = [&](int nnear) {
    int b = 0;
    while (b < nnear) {
      const int4 t = sload4(A.topo, __builtin_amdgcn_readfirstlane(near[b++]));
      const int s0 = t.x;
      int s1 = t.x + t.y;
      while (b < nnear) {
        const int4 tn = sload4(A.topo, __builtin_amdgcn_readfirstlane(near[b]));
        if (tn.x != s1) break;
        s1 += tn.y;
        ++b;
      }
      s1 = __builtin_amdgcn_readfirstlane(s1);  // wave-uniform: keeps source loads scalar
      // self pairs are skipped
      n_near += static_cast<unsigned long long>(s1 - s0) * nvalid -
                static_cast<unsigned long long>(max(0, min(s1, base + nvalid) - max(s0, base)));
      // batches of 8, then the rest exactly (4, 2, 1); the quadrupole adjoint
      // and the adjoint that also returns forces pad their last batch of 8
      // instead (tail variants push them past their register budget and cost
      // more than the padding saves)
      if constexpr (VJP && (ORDER == 2 || FORCES)) {
        for (int j0 = s0; j0 < s1; j0 += kChunk)
          near_batch<VJP, FORCES, kChunk, true, true>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy,
                                                      gz);
      } else {
        int j0 = s0;
        if constexpr (VJP) {
          // ranges away from the group's own points skip the self-pair test
          if (s1 <= base || s0 >= base + nvalid) {
            for (; j0 + kChunk <= s1; j0 += kChunk)
              near_batch<VJP, FORCES, kChunk, false>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy,
                                                     gz);
          }
        }
        if constexpr (kWide) {
          for (; j0 + 2 * kChunk <= s1; j0 += 2 * kChunk)
            near_batch<VJP, FORCES, 2 * kChunk>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy, gz);
        }
        for (; j0 + kChunk <= s1; j0 += kChunk)
          near_batch<VJP, FORCES, kChunk>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy, gz);
        if (j0 + 4 <= s1) {
          near_batch<VJP, FORCES, 4>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy, gz);
          j0 += 4;
        }
        if (j0 + 2 <= s1) {
          near_batch<VJP, FORCES, 2>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy, gz);
          j0 += 2;
        }
        if (j0 < s1) near_batch<VJP, FORCES, 1>(A, L, i, j0, s1, lane, fx, fy, fz, gx, gy, gz);
      }
    }
  };

  const int root = A.nsets > 1 ? set_of(A.offs, A.nsets, base) : 0;
  traverse<true>(A.hdr, &A.hdr->overflow, A.topo, A.geo, A.mp, A.beta, box, lane, root, stack,
                 far, near, flush_far, flush_near);

  if (valid) {
    const int64_t o = static_cast<int64_t>(A.perm[i]);
    if constexpr (FORCES) {
      A.forces[3 * o + 0] = fx;
      A.forces[3 * o + 1] = fy;
      A.forces[3 * o + 2] = fz;
    }
    if constexpr (VJP) A.qgrad[i] = make_float4(gx, gy, gz, 0.f);
  }
}

// Claim order of the next walk: within each eighth of the groups (one block
// each), heaviest first by the costs that the previous walk of the same kind
// (forward or adjoint) over this workspace measured, when its tag says it
// walked as many groups; else by the other kind's costs under the same test;
// else the groups' own order.  The persistent waves drain the queue unevenly
// at the end, and dealing the cheap groups last shortens that tail.  Costs
// are binned (256 bins, heaviest first) and the order inside a bin is that
// of the atomics.  Only the schedule depends on the costs, never a result.
constexpr int kOrderBins = 256;
__global__ __launch_bounds__(1024) void k_group_order(const Header* __restrict__ hdr,
                                                      const float* __restrict__ cost_own,
                                                      const float* __restrict__ cost_other,
                                                      int* __restrict__ gorder, Sched* sched_own,
                                                      const Sched* sched_other) {
  __shared__ float s_max[16];
  __shared__ int s_bin[kOrderBins];
  const int ng = hdr->num_groups;
  const int2 r = gw::eighth_range(ng, static_cast<int>(blockIdx.x));
  // block 0 rewrites the own tag below; a block that reads the new tag early
  // orders by stale costs, which changes only the schedule
  const bool own = sched_own->magic == kSchedMagic && sched_own->groups == ng;
  const bool other = sched_other->magic == kSchedMagic && sched_other->groups == ng;
  const float* gcost = own ? cost_own : cost_other;
  if (!own && !other) {
    for (int t = r.x + static_cast<int>(threadIdx.x); t < r.y; t += 1024) gorder[t] = t;
  } else {
    auto clean = [](float c) { return c > 0.0f && c < 3.0e38f ? c : 0.0f; };  // NaN -> 0
    float m = 0.0f;
    for (int t = r.x + static_cast<int>(threadIdx.x); t < r.y; t += 1024)
      m = fmaxf(m, clean(gcost[t]));
    m = gw::wave_max(m);
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    for (int b = threadIdx.x; b < kOrderBins; b += 1024) s_bin[b] = 0;
    __syncthreads();
    m = 0.0f;
    for (int k = 0; k < 16; ++k) m = fmaxf(m, s_max[k]);
    const float scale = m > 0.0f ? (kOrderBins - 0.5f) / m : 0.0f;
    auto bin_of = [&](int t) {
      const int b = static_cast<int>(clean(gcost[t]) * scale);
      return kOrderBins - 1 - min(max(b, 0), kOrderBins - 1);  // heaviest in bin 0
    };
    for (int t = r.x + static_cast<int>(threadIdx.x); t < r.y; t += 1024)
      atomicAdd(&s_bin[bin_of(t)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan of the bins (256 entries)
      int run = r.x;
      for (int b = 0; b < kOrderBins; ++b) {
        const int c = s_bin[b];
        s_bin[b] = run;
        run += c;
      }
    }
    __syncthreads();
    for (int t = r.x + static_cast<int>(threadIdx.x); t < r.y; t += 1024)
      gorder[atomicAdd(&s_bin[bin_of(t)], 1)] = t;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sched_own->magic = kSchedMagic;  // the walk after this order measures every group
    sched_own->groups = ng;
  }
}

// Persistent walk: each wave takes query groups from a device work queue
// until none are left (heavy groups do not hold up a static schedule).
// Occupancy targets: the adjoint walks at 4 waves per SIMD (at most 128
// VGPRs; their 40 KB of LDS per 4-wave block allow no more), the forward
// walks at 7 (at most 72 VGPRs; 22 KB of LDS per block, 7 blocks per CU;
// 6 waves were 5% slower).  WPB waves per block (the block size is a tuning
// dimension; every wave walks on its own).
template <bool VJP, int ORDER, bool FORCES, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(VJP ? 4 : 7))) void k_walk(
    WalkArgs A) {
  __shared__ int s_stack[WPB][kRunStackWords];
  __shared__ int s_far[WPB][kFarCap];
  __shared__ int s_near[WPB][kNearCap];
  // the adjoint's wave sums (the forward walk needs none)
  __shared__ __align__(16) float s_red[WPB][VJP ? kRedMax * kRedRow : 4];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int num_groups = A.hdr->num_groups;
  unsigned long long n_far = 0, n_near = 0, n_groups = 0;
  int tries = 0;
  while (true) {
    const int g = next_group(A.hdr, num_groups, lane, tries, A.order);
    if (g < 0) break;  // every wave reaches this exit
    const int2 grp = A.groups[g];
    const int base = __builtin_amdgcn_readfirstlane(grp.x);
    const int cnt = __builtin_amdgcn_readfirstlane(grp.y);
    if (cnt < 1 || cnt > 64 || base < 0 || base + cnt > A.n) continue;  // corrupt entry
    const uint64_t t0 = __builtin_amdgcn_s_memtime();
    walk_group<VJP, ORDER, FORCES>(A, base, cnt, lane, s_stack[w], s_far[w], s_near[w],
                                   s_red[w], n_far, n_near);
    const uint64_t t1 = __builtin_amdgcn_s_memtime();
    if (lane == 0) A.gcost[g] = static_cast<float>(t1 - t0);
    ++n_groups;
  }
  if (A.stats != nullptr && lane == 0) {
    atomicAdd(&A.stats[0], n_far);
    atomicAdd(&A.stats[1], n_near);
    atomicAdd(&A.stats[2], n_groups);
  }
}

// Zero the adjoint accumulators (fill kernel; graph-capture friendly).
__global__ __launch_bounds__(kB) void k_zero_adjoint(const Header* __restrict__ hdr, int64_t n,
                                                     float* __restrict__ sgrad,
                                                     float* __restrict__ adj,
                                                     float* __restrict__ adjq) {
  const int64_t nn = hdr->num_nodes;
  const int64_t total = 4 * n + 4 * nn + 8 * nn;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < total;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    if (i < 4 * n)
      sgrad[i] = 0.f;
    else if (i < 4 * n + 4 * nn)
      adj[i - 4 * n] = 0.f;
    else
      adjq[i - 4 * n - 4 * nn] = 0.f;
  }
}

// Push node adjoints down to the points and write gradients in input order.
template <bool QUAD>
__global__ __launch_bounds__(kB) void k_finalize(const Header* __restrict__ hdr, int64_t n,
                                                 const int32_t* __restrict__ offs, int nsets,
                                                 const int4* __restrict__ topo,
                                                 const float4* __restrict__ mp,
                                                 const float4* __restrict__ sp,
                                                 const uint32_t* __restrict__ perm,
                                                 const float4* __restrict__ qgrad,
                                                 const float* __restrict__ sgrad,
                                                 const float* __restrict__ adj,
                                                 const float* __restrict__ adjq,
                                                 float* __restrict__ gx_out,
                                                 float* __restrict__ gm_out) {
  int64_t r = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (r >= n) return;
  const float4 p = sp[r];
  const float4 q = qgrad[r];
  float ax = q.x + sgrad[4 * r + 0];
  float ay = q.y + sgrad[4 * r + 1];
  float az = q.z + sgrad[4 * r + 2];
  float am = sgrad[4 * r + 3];
  const int rr = static_cast<int>(r);
  int node = nsets > 1 ? set_of(offs, nsets, rr) : 0;  // the root of the point's set
  for (int depth = 0; depth < kMaxLevels; ++depth) {
    const float4 m = mp[node];
    if (m.w != 0.0f) {
      const float Ax = adj[4 * node], Ay = adj[4 * node + 1], Az = adj[4 * node + 2];
      const float B = adj[4 * node + 3];
      const float dx = p.x - m.x, dy = p.y - m.y, dz = p.z - m.z;
      const float wm = p.w / m.w;
      ax += wm * Ax;
      ay += wm * Ay;
      az += wm * Az;
      am += B + (Ax * dx + Ay * dy + Az * dz) / m.w;
      if constexpr (QUAD) {
        // xx yy zz, then xy xz yz stored doubled by the walk
        const float* G = adjq + 8 * node;
        const float Gxy = 0.5f * G[3], Gxz = 0.5f * G[4], Gyz = 0.5f * G[5];
        const float Gdx = G[0] * dx + Gxy * dy + Gxz * dz;
        const float Gdy = Gxy * dx + G[1] * dy + Gyz * dz;
        const float Gdz = Gxz * dx + Gyz * dy + G[2] * dz;
        ax += 2.0f * p.w * Gdx;
        ay += 2.0f * p.w * Gdy;
        az += 2.0f * p.w * Gdz;
        am += dx * Gdx + dy * Gdy + dz * Gdz;
      }
    }
    const int4 t = topo[node];
    if (t.w == 0) break;
    int ch = t.z;
    for (int c = 1; c < t.w; ++c)
      if (topo[t.z + c].x <= rr) ch = t.z + c;
    node = ch;
  }
  const int64_t o = perm[r];
  gx_out[3 * o + 0] = ax;
  gx_out[3 * o + 1] = ay;
  gx_out[3 * o + 2] = az;
  gm_out[o] = am;
}

__global__ void k_copy_stats(const Header* __restrict__ hdr, int64_t* __restrict__ out) {
  if (threadIdx.x != 0) return;
  out[0] = static_cast<int64_t>(hdr->stats[0]);
  out[1] = static_cast<int64_t>(hdr->stats[1]);
  out[2] = static_cast<int64_t>(hdr->stats[2]);
  out[3] = hdr->num_nodes;
  int nl = 0;
  for (int l = 0; l < kMaxLevels; ++l)
    if (hdr->level_count[l] > 0) nl = l + 1;
  out[4] = nl;
  out[5] = hdr->overflow;
  out[6] = 0;
  out[7] = 0;
}

}  // namespace

int64_t node_capacity(int64_t n, int max_depth, int leaf_size, int nsets) {
  // Level 0 holds the roots; every deeper level holds at most 8 children per
  // split node, nodes of one level are disjoint, and a split node holds more
  // than leaf_size points.
  int64_t cap = nsets;
  int64_t internal_max = n / (static_cast<int64_t>(leaf_size) + 1);
  for (int l = 1; l <= max_depth; ++l) {
    int64_t a = (l < 18) ? static_cast<int64_t>(nsets) << (3 * l) : INT64_MAX;
    int64_t b = 8 * internal_max;
    int64_t c = n;
    cap += std::min(a, std::min(b, c));
  }
  return cap;
}

Layout layout(int64_t n, int max_depth, int leaf_size, int nsets) {
  Layout L;
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  size_t nn = static_cast<size_t>(n > 0 ? n : 1);
  L.nsets = nsets;
  L.cap = node_capacity(n, max_depth, leaf_size, nsets);
  size_t cap = static_cast<size_t>(L.cap) + 1;  // + the sentinel node
  size_t o = 0;
  L.header = o; o = align(o + sizeof(Header));
  L.codes = o; o = align(o + 4 * nn);
  L.perm = o; o = align(o + 4 * nn);
  L.keys_tmp = o; o = align(o + 4 * nn);
  L.vals_tmp = o; o = align(o + 4 * nn);
  // radix sort temporaries, reused by the scan of the group counts
  L.sort_tmp = o;
  o = align(o + morton_sort_temp_bytes(n, nsets));
  L.bbox_tmp = o; o = align(o + 6 * 4 * (kBboxBlocks + 1));
  L.sp = o; o = align(o + 16 * nn);
  L.sc = o; o = align(o + 16 * nn);
  L.topo = o; o = align(o + 16 * cap);
  L.prefix = o; o = align(o + 4 * cap);
  L.parent = o; o = align(o + 4 * cap);
  L.geo = o; o = align(o + 16 * cap);
  L.mp = o; o = align(o + 16 * cap);
  L.qd = o; o = align(o + 32 * cap);
  L.adj = o; o = align(o + 16 * cap);
  L.adjq = o; o = align(o + 32 * cap);
  L.qgrad = o; o = align(o + 16 * nn);
  L.sgrad = o; o = align(o + 16 * nn);
  L.groups = o; o = align(o + 8 * nn);
  L.gcost = o; o = align(o + 8 * nn);  // forward, then adjoint walks
  L.gorder = o; o = align(o + 4 * nn);
  L.sched = o; o = align(o + 2 * sizeof(Sched));
  L.tiles = o;
  o = align(o + 4 * static_cast<size_t>(kTopoCols) * ((nn + kTopoTile - 1) / kTopoTile));
  L.geoms = o; o = align(o + sizeof(CubeGeom) * static_cast<size_t>(nsets));
  L.offs = o; o = align(o + 4 * (static_cast<size_t>(nsets) + 1));
  L.setbox = o; o = align(o + 24 * static_cast<size_t>(nsets));
  L.codes_in = o; o = align(o + 4 * nn);
  L.total = o;
  return L;
}

int data_bounds(hipStream_t s, const float* pts, int64_t n, float* tmp, float** bounds) {
  const unsigned nb = static_cast<unsigned>(
      std::min<int64_t>(kBboxBlocks, std::max<int64_t>(1, (n + kB - 1) / kB)));
  k_bbox_partial<<<nb, kB, 0, s>>>(pts, n, tmp);
  SBH_LAUNCH_CHECK();
  *bounds = tmp + 6 * kBboxBlocks;
  k_bbox_reduce<<<1, kB, 0, s>>>(tmp, static_cast<int>(nb), n, *bounds);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

namespace {

// Everything after the roots exist: codes, sort, gather, levels, moments, groups.
int build_rest(hipStream_t s, char* ws, const Layout& L, const float* pts, const float* ms,
               const float* cot, int64_t n, const int32_t* offs, int max_depth,
               int leaf_size, int parts, float4* dipoles) {
  auto* hdr = reinterpret_cast<Header*>(ws + L.header);
  auto* codes = reinterpret_cast<uint32_t*>(ws + L.codes);
  auto* codes_in = reinterpret_cast<uint32_t*>(ws + L.codes_in);
  auto* perm = reinterpret_cast<uint32_t*>(ws + L.perm);
  auto* keys = reinterpret_cast<uint32_t*>(ws + L.keys_tmp);
  auto* vals = reinterpret_cast<uint32_t*>(ws + L.vals_tmp);
  auto* topo = reinterpret_cast<int4*>(ws + L.topo);
  auto* prefix = reinterpret_cast<uint32_t*>(ws + L.prefix);
  auto* parent = reinterpret_cast<int32_t*>(ws + L.parent);
  auto* geo = reinterpret_cast<float4*>(ws + L.geo);
  auto* mp = reinterpret_cast<float4*>(ws + L.mp);
  auto* qd = reinterpret_cast<float4*>(ws + L.qd);
  auto* sp = reinterpret_cast<float4*>(ws + L.sp);
  auto* sc = reinterpret_cast<float4*>(ws + L.sc);
  auto* geoms = reinterpret_cast<CubeGeom*>(ws + L.geoms);
  const int nsets = L.nsets;
  if (n > 0) {
    k_codes<<<grid_for(n, kB), kB, 0, s>>>(pts, n, geoms, offs, nsets, max_depth, codes_in);
    SBH_LAUNCH_CHECK();
    // stable order by (set, code, input row); the permutation lands in perm
    // or in vals (then k_gather copies it)
    const uint32_t* order = nullptr;
    const uint32_t* sorted = nullptr;
    int rc = morton_sort(s, codes_in, n, 3 * max_depth, offs, nsets, keys, vals, codes, perm,
                         ws + L.sort_tmp, &hdr->overflow, &sorted, &order);
    if (rc != SBH_OK) return rc;
    k_gather<<<grid_for(n, kB), kB, 0, s>>>(pts, ms, cot, n, order, sorted, perm, codes, sp, sc);
    SBH_LAUNCH_CHECK();
  }
  // topology, query groups and moments from the sorted codes
  auto* tiles = reinterpret_cast<int32_t*>(ws + L.tiles);
  auto* prec = reinterpret_cast<uint32_t*>(ws + L.qgrad);  // per-point scratch of the build
  auto* ends = reinterpret_cast<int2*>(ws + L.adj);        // per-node scratch of the build
  auto* groups = reinterpret_cast<int2*>(ws + L.groups);
  const int64_t ntiles = (n + kTopoTile - 1) / kTopoTile;
  if (ntiles > 0) {
    k_topo_count<<<static_cast<unsigned>(ntiles), kTopoTile, 0, s>>>(hdr, n, codes, offs, nsets,
                                                              max_depth, leaf_size, tiles, prec);
    SBH_LAUNCH_CHECK();
  }
  k_topo_scan<<<1, 1024, 0, s>>>(hdr, tiles, ntiles, n, nsets, max_depth, L.cap);
  SBH_LAUNCH_CHECK();
  if (ntiles > 0) {
    k_topo_emit<<<static_cast<unsigned>(ntiles), kTopoTile, 0, s>>>(hdr, n, codes, offs, nsets,
                                                             max_depth, tiles, prec, geoms, topo,
                                                             prefix, geo, parent, ends, groups);
    SBH_LAUNCH_CHECK();
  }
  const bool moments = (parts & kBuildMoments) != 0;
  float4* wd = moments ? dipoles : nullptr;
  k_topo_finalize<<<grid_for(L.cap, kB / 16, 2048), kB, 0, s>>>(
      hdr, n, nsets, moments ? 1 : 0, topo, ends, groups, geo, sp, mp, qd, sc, wd,
      static_cast<int>(L.cap));
  SBH_LAUNCH_CHECK();
  if (moments) {
    int64_t internal_max = n / (static_cast<int64_t>(leaf_size) + 1);
    for (int l = max_depth - 1; l >= 0; --l) {
      // internal nodes of level l: at most nsets 8^l and at most internal_max
      const int64_t cap_l = std::min<int64_t>(
          l < 18 ? static_cast<int64_t>(nsets) << (3 * l) : INT64_MAX,
          std::max<int64_t>(nsets, internal_max));
      k_moments_internal<<<grid_for(cap_l, kB / 16, 1024), kB, 0, s>>>(hdr, l, topo, geo, sp,
                                                                        mp, qd, sc, wd);
      SBH_LAUNCH_CHECK();
    }
  }
  return SBH_OK;
}

int init_roots(hipStream_t s, char* ws, const Layout& L, int64_t n, const float* bounds6,
               const uint32_t* setbox) {
  k_forest_init<<<1, kB, 0, s>>>(
      n, L.nsets, bounds6, setbox, reinterpret_cast<int32_t*>(ws + L.offs),
      reinterpret_cast<CubeGeom*>(ws + L.geoms), reinterpret_cast<Header*>(ws + L.header),
      reinterpret_cast<int4*>(ws + L.topo), reinterpret_cast<uint32_t*>(ws + L.prefix),
      reinterpret_cast<float4*>(ws + L.geo), reinterpret_cast<int32_t*>(ws + L.parent),
      reinterpret_cast<float4*>(ws + L.mp), reinterpret_cast<float4*>(ws + L.qd),
      static_cast<int>(L.cap));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

}  // namespace

int build(hipStream_t s, char* ws, const Layout& L, const float* pts, const float* ms,
          const float* cot, int64_t n, int max_depth, int leaf_size,
          const float* user_bounds, int parts, float4* dipoles) {
  if (L.nsets != 1) return fail_internal_msg("build: layout is a forest");
  if (user_bounds == nullptr) {
    float* b = nullptr;
    int rc = data_bounds(s, pts, n, reinterpret_cast<float*>(ws + L.bbox_tmp), &b);
    if (rc != SBH_OK) return rc;
    user_bounds = b;
  }
  int rc = init_roots(s, ws, L, n, user_bounds, nullptr);
  if (rc != SBH_OK) return rc;
  return build_rest(s, ws, L, pts, ms, cot, n, reinterpret_cast<const int32_t*>(ws + L.offs),
                    max_depth, leaf_size, parts, dipoles);
}

int build_forest(hipStream_t s, char* ws, const Layout& L, const float* pts, const float* ms,
                 const float* cot, int64_t n, const int32_t* offs, int max_depth,
                 int leaf_size, int parts) {
  auto* box = reinterpret_cast<uint32_t*>(ws + L.setbox);
  auto* ofs = reinterpret_cast<int32_t*>(ws + L.offs);
  SBH_HIP_TRY(hipMemcpyAsync(ofs, offs, 4 * (static_cast<size_t>(L.nsets) + 1),
                             hipMemcpyDeviceToDevice, s));
  k_setbox_init<<<grid_for(6 * L.nsets, kB, 64), kB, 0, s>>>(box, L.nsets);
  SBH_LAUNCH_CHECK();
  if (n > 0) {
    // tile partials in the sort keys' scratch (free until the codes are made)
    auto* part = reinterpret_cast<float*>(ws + L.keys_tmp);
    const int64_t tiles = (n + kB - 1) / kB;
    k_set_bbox_tiles<<<static_cast<unsigned>(tiles), kB, 0, s>>>(pts, n, ofs, L.nsets, box,
                                                                part);
    SBH_LAUNCH_CHECK();
    k_set_bbox_sets<<<static_cast<unsigned>(L.nsets), kB, 0, s>>>(ofs, part, box);
    SBH_LAUNCH_CHECK();
  }
  int rc = init_roots(s, ws, L, n, nullptr, box);
  if (rc != SBH_OK) return rc;
  return build_rest(s, ws, L, pts, ms, cot, n, ofs, max_depth, leaf_size, parts, nullptr);
}

namespace {
template <int WPB>
void launch_walk(hipStream_t s, unsigned grid, bool vjp, int order, bool f, const WalkArgs& A) {
  const unsigned block = 64 * WPB;
  if (vjp && order == 2 && f)
    k_walk<true, 2, true, WPB><<<grid, block, 0, s>>>(A);
  else if (vjp && order == 2)
    k_walk<true, 2, false, WPB><<<grid, block, 0, s>>>(A);
  else if (vjp && order == 3 && f)
    k_walk<true, 3, true, WPB><<<grid, block, 0, s>>>(A);
  else if (vjp && order == 3)
    k_walk<true, 3, false, WPB><<<grid, block, 0, s>>>(A);
  else if (vjp && f)
    k_walk<true, 1, true, WPB><<<grid, block, 0, s>>>(A);
  else if (vjp)
    k_walk<true, 1, false, WPB><<<grid, block, 0, s>>>(A);
  else if (order == 2)
    k_walk<false, 2, true, WPB><<<grid, block, 0, s>>>(A);
  else
    k_walk<false, 1, true, WPB><<<grid, block, 0, s>>>(A);
}
}  // namespace

int schedule_groups(hipStream_t s, char* ws, const Layout& L, int64_t n, int kind,
                    const int** order, float** cost) {
  auto* costs = reinterpret_cast<float*>(ws + L.gcost);
  auto* tags = reinterpret_cast<Sched*>(ws + L.sched);
  const int64_t nn = n > 0 ? n : 1, own = kind != 0 ? 1 : 0;
  *order = reinterpret_cast<const int*>(ws + L.gorder);
  *cost = costs + own * nn;
  k_group_order<<<8, 1024, 0, s>>>(reinterpret_cast<const Header*>(ws + L.header), *cost,
                                   costs + (1 - own) * nn, reinterpret_cast<int*>(ws + L.gorder),
                                   tags + own, tags + (1 - own));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

int walk(hipStream_t s, char* ws, const Layout& L, int64_t n, float beta, float softening,
         bool vjp, int order, float* forces, float* grad_points, float* grad_masses,
         bool want_stats, hipEvent_t* ev_walk, int walk_blocks, int walk_threads) {
  if (n == 0) return SBH_OK;
  auto* hdr = reinterpret_cast<Header*>(ws + L.header);
  WalkArgs A;
  A.hdr = hdr;
  A.groups = reinterpret_cast<const int2*>(ws + L.groups);
  A.topo = reinterpret_cast<const int4*>(ws + L.topo);
  A.geo = reinterpret_cast<const float4*>(ws + L.geo);
  A.mp = reinterpret_cast<const float4*>(ws + L.mp);
  A.qd = reinterpret_cast<const float4*>(ws + L.qd);
  A.sp = reinterpret_cast<const float4*>(ws + L.sp);
  A.sc = reinterpret_cast<const float4*>(ws + L.sc);
  A.perm = reinterpret_cast<const uint32_t*>(ws + L.perm);
  A.offs = reinterpret_cast<const int32_t*>(ws + L.offs);
  A.nsets = L.nsets;
  A.n = static_cast<int>(n);
  A.sentinel = static_cast<int>(L.cap);
  A.beta = beta;
  float eps = fmaxf(softening, 0.0f);
  A.e2 = eps * eps + 1e-12f;
  A.forces = forces;
  A.qgrad = reinterpret_cast<float4*>(ws + L.qgrad);
  A.sgrad = reinterpret_cast<float*>(ws + L.sgrad);
  A.adj = reinterpret_cast<float*>(ws + L.adj);
  A.adjq = reinterpret_cast<float*>(ws + L.adjq);
  A.stats = want_stats ? hdr->stats : nullptr;
  {
    const int* order = nullptr;
    float* cost = nullptr;
    const int rc = schedule_groups(s, ws, L, n, vjp ? 1 : 0, &order, &cost);
    if (rc != SBH_OK) return rc;
    A.order = order;
    A.gcost = cost;
  }
  SBH_LAUNCH_CHECK();
  if (vjp) {
    k_zero_adjoint<<<grid_for(4 * n + 12 * L.cap, kB, 8192), kB, 0, s>>>(hdr, n, A.sgrad, A.adj,
                                                                        A.adjq);
    SBH_LAUNCH_CHECK();
  }
  // Persistent waves drain the group queue: 8192 waves by default (8 per
  // SIMD, twice what fits at once, so that every CU stays busy to the end),
  // never more than one wave per point.
  const int threads = walk_threads > 0 ? walk_threads : kB;
  const int wpb = threads / 64;
  const int64_t waves_cap = walk_blocks > 0 ? static_cast<int64_t>(walk_blocks) * wpb : 8192;
  unsigned grid = grid_for(n, 64 * wpb, static_cast<unsigned>((waves_cap + wpb - 1) / wpb));
  if (ev_walk != nullptr) SBH_HIP_TRY(hipEventRecord(ev_walk[0], s));
  const bool f = forces != nullptr;
  switch (wpb) {
    case 1: launch_walk<1>(s, grid, vjp, order, f, A); break;
    case 2: launch_walk<2>(s, grid, vjp, order, f, A); break;
    case 8: launch_walk<8>(s, grid, vjp, order, f, A); break;
    default: launch_walk<4>(s, grid, vjp, order, f, A); break;
  }
  SBH_LAUNCH_CHECK();
  if (ev_walk != nullptr) SBH_HIP_TRY(hipEventRecord(ev_walk[1], s));
  if (vjp) {
    if (order != 1)  // the node adjoints carry dL/dS
      k_finalize<true><<<grid_for(n, kB), kB, 0, s>>>(hdr, n, A.offs, A.nsets, A.topo, A.mp,
                                                       A.sp, A.perm, A.qgrad, A.sgrad, A.adj,
                                                       A.adjq, grad_points, grad_masses);
    else
      k_finalize<false><<<grid_for(n, kB), kB, 0, s>>>(hdr, n, A.offs, A.nsets, A.topo, A.mp,
                                                        A.sp, A.perm, A.qgrad, A.sgrad, A.adj,
                                                        A.adjq, grad_points, grad_masses);
    SBH_LAUNCH_CHECK();
  }
  return SBH_OK;
}

int copy_stats(hipStream_t s, char* ws, const Layout& L, int64_t* out) {
  k_copy_stats<<<1, 64, 0, s>>>(reinterpret_cast<const Header*>(ws + L.header), out);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

// ------------------------------------------------------------------ query order
namespace {

// 30-bit Morton key of each query in the queries' own bounding box.
__global__ __launch_bounds__(kB) void k_query_keys(const float* __restrict__ q, int64_t nq,
                                                   const float* __restrict__ bounds,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
  const float lx = bounds[0], ly = bounds[1], lz = bounds[2];
  const float span = fmaxf(fmaxf(bounds[3] - lx, bounds[4] - ly), fmaxf(bounds[5] - lz, 1e-30f));
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < nq;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    keys[i] = morton3(quantize_axis(q[3 * i], lx, span, 10), quantize_axis(q[3 * i + 1], ly, span, 10),
                      quantize_axis(q[3 * i + 2], lz, span, 10));
    vals[i] = static_cast<uint32_t>(i);
  }
}

struct QueryOrderLayout {
  size_t tmp, keys, vals, keys_out, perm, sort, total;
};

QueryOrderLayout query_order_layout(int64_t nq) {
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  const size_t n4 = 4 * static_cast<size_t>(nq > 0 ? nq : 1);
  QueryOrderLayout L;
  size_t o = 0;
  L.tmp = o; o = align(o + 6 * 4 * (kBboxBlocks + 1));
  L.keys = o; o = align(o + n4);
  L.vals = o; o = align(o + n4);
  L.keys_out = o; o = align(o + n4);
  L.perm = o; o = align(o + n4);
  L.sort = o; o = align(o + radix_sort_temp_bytes(nq));
  L.total = o;
  return L;
}

}  // namespace

size_t query_order_bytes(int64_t nq) { return query_order_layout(nq).total; }

int query_order(hipStream_t s, const float* q, int64_t nq, char* ws, const uint32_t** perm) {
  const QueryOrderLayout L = query_order_layout(nq);
  float* bounds = nullptr;
  int rc = data_bounds(s, q, nq, reinterpret_cast<float*>(ws + L.tmp), &bounds);
  if (rc != SBH_OK) return rc;
  auto* keys = reinterpret_cast<uint32_t*>(ws + L.keys);
  auto* vals = reinterpret_cast<uint32_t*>(ws + L.vals);
  auto* out = reinterpret_cast<uint32_t*>(ws + L.perm);
  k_query_keys<<<grid_for(nq, kB, 4096), kB, 0, s>>>(q, nq, bounds, keys, vals);
  SBH_LAUNCH_CHECK();
  rc = radix_sort_pairs(s, keys, vals, reinterpret_cast<uint32_t*>(ws + L.keys_out), out, nq, 30,
                        ws + L.sort);
  if (rc != SBH_OK) return rc;
  *perm = out;
  return SBH_OK;
}

}  // namespace tree
}  // namespace sbh
