// Adaptive Morton octree: device-side header, workspace layout and the build
// / walk drivers shared by the C-ABI entry points.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace sbh {
namespace tree {

constexpr int kMaxLevels = 24;
constexpr int kBboxBlocks = 1024;

struct CubeGeom {
  float min_x, min_y, min_z, span;
  float cx, cy, cz, half_size;
};

// Lives at offset 0 of the workspace; written only by device kernels.
struct Header {
  CubeGeom geom;
  int32_t level_start[kMaxLevels];
  int32_t level_count[kMaxLevels];
  int32_t num_nodes;
  int32_t overflow;
  int32_t num_groups;  // query groups of the walk
  int32_t queue;       // unused by the walks (kept zero)
  int32_t xq[8];       // per-XCD work queues: next group of each eighth of the groups
  unsigned long long stats[8];  // far interactions, near pairs, groups
};

// Tag of the walk schedule: the group costs in the workspace were measured
// by a walk over `groups` groups (the tag is checked, never trusted: any
// cost values give a valid claim order).
constexpr uint32_t kSchedMagic = 0x5c4ed01eu;
struct Sched {
  uint32_t magic;
  int32_t groups;
};

constexpr int kGroupMax = 64;     // one wavefront per query group
constexpr int kGroupCell = 2048;  // groups are runs inside cells of at most this many points
// Query trees (queries walked against a separate source tree): topology and
// groups only, depth 10, leaves of at most one group.
constexpr int kQueryTreeDepth = 10;
constexpr int kQueryTreeLeaf = kGroupMax;

// Byte offsets of every array inside the workspace.  Node arrays hold `cap`
// entries (an upper bound on the node count, see node_capacity()).  A
// workspace holds a forest of `nsets` independent trees (one per point set;
// node k < nsets is the root of set k), one tree in the plain case.
struct Layout {
  int64_t cap;
  int nsets;
  size_t geoms, offs, setbox, codes_in;  // per-set cubes, set offsets, bbox atomics
  size_t header, codes, perm, keys_tmp, vals_tmp, sort_tmp, bbox_tmp;
  size_t sp, sc;                       // sorted (x,y,z,m) and cotangent records
  size_t topo, prefix, parent, geo;    // int4 {start,count,first_child,nchild}, cell prefix
  size_t mp, qd;                       // float4 {com, mass}, 2 x float4 second moment
  size_t adj, adjq, qgrad, sgrad;      // adjoint accumulators
  size_t groups, tiles;                // int2 {start, count} per group; per-tile counters
  size_t gcost, gorder, sched;         // walk schedule: group costs (forward, adjoint), order, tags
  size_t total;
};

int64_t node_capacity(int64_t n, int max_depth, int leaf_size, int nsets = 1);
Layout layout(int64_t n, int max_depth, int leaf_size, int nsets = 1);

// Set of sorted (or input) point i given the set offsets offs[0..nsets]:
// the largest s with offs[s] <= i.
__device__ __forceinline__ int set_of(const int32_t* __restrict__ offs, int nsets, int i) {
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Claim order of the next walk over the query groups of the tree in ws
// (n points): the groups' own order, or heaviest first within each eighth by
// the costs the previous walk of the same kind (0: forward-type walks, 1:
// adjoint walks; else the other kind) measured in this workspace.  *order is
// the permutation for gw::next_group, *cost where the walk stores each
// group's measured cost.  Only the schedule depends on them.
int schedule_groups(hipStream_t s, char* ws, const Layout& L, int64_t n, int kind,
                    const int** order, float** cost);

// Data bounds of n points on the device (no host synchronisation): min - 1e-3
// (3 floats) then max + 1e-3 (3 floats), written to tmp + 6 * kBboxBlocks and
// returned in *bounds; tmp holds 6 * (kBboxBlocks + 1) floats.
int data_bounds(hipStream_t s, const float* pts, int64_t n, float* tmp, float** bounds);

// Morton order of nq query points (30-bit keys in their own bounding box):
// *perm (inside ws, nq entries) lists the query indices in that order.  Used
// to schedule per-query walks so that the lanes of a wave are neighbours;
// ws holds query_order_bytes(nq).
size_t query_order_bytes(int64_t nq);
int query_order(hipStream_t s, const float* q, int64_t nq, char* ws, const uint32_t** perm);

// Parts of build() beyond the topology (sorted records, nodes, geometry).
constexpr int kBuildMoments = 1;  // mass, centre of mass, second moment per node
constexpr int kBuildGroups = 2;   // query groups of the walk (cells of <= kGroupMax points)
constexpr int kBuildAll = kBuildMoments | kBuildGroups;

// Build topology, moments and groups.  `cot` may be null (forces only).
// `user_bounds` (device, 6 floats: min xyz, max xyz) may be null to use the
// data bounds expanded by 1e-3.  Without kBuildMoments the node masses
// stay unset (zero only for the sentinel).
// dipoles (may be null): with kBuildMoments, also the per-node sums of the
// masses times the gathered vectors (winding numbers: areas times unit
// normals), L.cap + 1 float4 (the sentinel's zero included).
int build(hipStream_t s, char* ws, const Layout& L, const float* pts, const float* ms,
          const float* cot, int64_t n, int max_depth, int leaf_size,
          const float* user_bounds, int parts = kBuildAll, float4* dipoles = nullptr);

// Forest build: points are the concatenation of L.nsets sets, set k holding
// input rows [offs[k], offs[k+1]) (device int32 array of nsets + 1 entries,
// offs[0] = 0, offs[nsets] = n).  Every set gets its own cube (its data bounds
// expanded by 1e-3), its own root and its own groups; no node mixes sets.
// With one set this is build() without user bounds.
int build_forest(hipStream_t s, char* ws, const Layout& L, const float* pts, const float* ms,
                 const float* cot, int64_t n, const int32_t* offs, int max_depth,
                 int leaf_size, int parts = kBuildAll);

// Walk; forces (input order, N x 3) may be null; with vjp, writes grad_points
// (N x 3) and grad_masses (N) in input order.  order: 1 monopole, 2 monopole
// + quadrupole, 3 (with vjp) monopole terms with the quadrupole term's
// second-moment adjoint (forces, if requested, are the monopole's).
// ev_walk (may be null): two events recorded immediately before and after the
// walk kernel on stream s.
// Works on forests: each group walks the tree of its own set.  walk_blocks
// (0: default) caps the persistent grid; walk_threads (0: 256, else 64, 128,
// 256 or 512) sets its block size.
int walk(hipStream_t s, char* ws, const Layout& L, int64_t n, float beta, float softening,
         bool vjp, int order, float* forces, float* grad_points, float* grad_masses,
         bool want_stats, hipEvent_t* ev_walk = nullptr, int walk_blocks = 0,
         int walk_threads = 0);

// Copies {far, near, groups, nodes, levels, overflow, 0, 0} into a device int64[8].
int copy_stats(hipStream_t s, char* ws, const Layout& L, int64_t* out);

// Cell geometry from a level-l prefix (same expression as the octree8 node
// centres, src/sbh_ffi.cu:1828-1834): {cx, cy, cz, diagonal}.
__device__ __forceinline__ float4 cell_geometry(const CubeGeom& g, uint32_t prefix, int level) {
#pragma clang fp contract(off)
  int ix = 0, iy = 0, iz = 0;
  for (int b = 0; b < level; ++b) {
    ix |= static_cast<int>((prefix >> (3 * b + 0)) & 1u) << b;
    iy |= static_cast<int>((prefix >> (3 * b + 1)) & 1u) << b;
    iz |= static_cast<int>((prefix >> (3 * b + 2)) & 1u) << b;
  }
  float step = g.half_size * 2.0f / static_cast<float>(1 << level);
  float4 r;
  r.x = fmaf(static_cast<float>(ix) + 0.5f, step, g.cx - g.half_size);
  r.y = fmaf(static_cast<float>(iy) + 0.5f, step, g.cy - g.half_size);
  r.z = fmaf(static_cast<float>(iz) + 0.5f, step, g.cz - g.half_size);
  r.w = step * 1.7320508075688772f;
  return r;
}

}  // namespace tree
}  // namespace sbh
