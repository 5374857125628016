// Level-3 targets: the single-split 8 x 8 x 8 tree in the Tree3DNode<1, 3>
// layout (src/sbh_ffi.cu:181-397, 736-1037) and the walks over
// Tree3DNode<1, 3> and Tree2DNode<1, 3> buffers (src/sbh_ffi.cu:1422-1680 via
// include/barnes_hut_gpu.h:19-72 and include/tree_sampler_gpu.h:25-140).
//
// Build: per-axis cell index of every point (quantize_axis_level3), a stable
// radix sort on the 9-bit cell id (the reference scatters with an atomic
// cursor, so its order inside a cell is a race; here it is (cell, index)),
// cell offsets by binary search, per-cell sums by one block per cell in a
// fixed order, and one 512-thread block writing the root and the child
// records (ranks of present cells by an LDS scan, the root's bit mask by one
// wave ballot per 64 cells).  No float atomics: the build is deterministic.
// Multi-level buffers in the same layouts (the reference's CPU trees) come
// from ndtree.hip.
//
// Classic walk: the reference's warp-voted traversal: a node is used as a far
// field only if every active query of a 32-query warp accepts it, so all
// queries of that warp take the same path.  For level-3 buffers (every child
// of the root a leaf) a warp either uses the root monopole or sums every
// point exactly.
//
// Stochastic walk: per subdomain (child of the root) its exact sum or
// monopole plus the control-variate corrections along roulette paths to a
// sample point; level-3 buffers have no paths (their subdomains are leaves).
#include <cfloat>
#include "common.hpp"
#include "o8walk.hpp"

namespace sbh {
namespace {

constexpr int kSplit = 8;
constexpr int kCells = kSplit * kSplit * kSplit;  // 512
constexpr int kMaxNodes3 = 1 + kCells;
constexpr int kB = 256;

// quantize_axis_level3 (src/sbh_ffi.cu:181-190), IEEE fp32.
__device__ __forceinline__ int quantize3(float x, float lo, float hi) {
#pragma clang fp contract(off)
  float span = hi - lo;
  if (span <= 1e-12f) return 0;
  float t = (x - lo) / span;
  t = fminf(fmaxf(t, 0.0f), 0.999999f);
  return static_cast<int>(t * static_cast<float>(kSplit));
}

__global__ __launch_bounds__(kB) void k_l3_cells(const float* __restrict__ pts, int64_t n,
                                                 const float* __restrict__ lo,
                                                 const float* __restrict__ hi,
                                                 uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ idx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (i >= n) return;
  const int qx = quantize3(pts[3 * i + 0], lo[0], hi[0]);
  const int qy = quantize3(pts[3 * i + 1], lo[1], hi[1]);
  const int qz = quantize3(pts[3 * i + 2], lo[2], hi[2]);
  keys[i] = static_cast<uint32_t>(qx + kSplit * (qy + kSplit * qz));
  idx[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kB) void k_l3_gather(const float* __restrict__ pts,
                                                  const float* __restrict__ nrm,
                                                  const float* __restrict__ ms, int64_t n,
                                                  const uint32_t* __restrict__ perm,
                                                  Leaf3* __restrict__ leaf) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (r >= n) return;
  const int64_t i = perm[r];
  Leaf3 L;
  for (int a = 0; a < 3; ++a) {
    L.pos[a] = pts[3 * i + a];
    L.normal[a] = nrm[3 * i + a];
  }
  L.mass = ms[i];
  L.alpha = 1.0f;
  leaf[r] = L;
}

// offs[c] = first sorted point of cell c (c = 0..512; offs[512] = n).
__global__ __launch_bounds__(kB) void k_l3_offsets(const uint32_t* __restrict__ cells, int64_t n,
                                                   int32_t* __restrict__ offs) {
  const int c = blockIdx.x * kB + threadIdx.x;
  if (c > kCells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (cells[mid] < static_cast<uint32_t>(c))
      lo = mid + 1;
    else
      hi = mid;
  }
  offs[c] = static_cast<int32_t>(lo);
}

// Per-cell sums {m, m x, m y, m z, m nx, m ny, m nz} in a fixed order (one
// block per cell, strided partial sums then an LDS tree).
__global__ __launch_bounds__(kB) void k_l3_cell_sums(const Leaf3* __restrict__ leaf,
                                                     const int32_t* __restrict__ offs,
                                                     float* __restrict__ sums) {
  __shared__ float sh[7][kB];
  const int c = blockIdx.x;
  const int s0 = offs[c], s1 = offs[c + 1];
  float v[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int j = s0 + static_cast<int>(threadIdx.x); j < s1; j += kB) {
    const Leaf3 L = leaf[j];
    v[0] += L.mass;
    for (int a = 0; a < 3; ++a) {
      v[1 + a] += L.mass * L.pos[a];
      v[4 + a] += L.mass * L.normal[a];
    }
  }
  for (int a = 0; a < 7; ++a) sh[a][threadIdx.x] = v[a];
  __syncthreads();
  for (int off = kB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off))
      for (int a = 0; a < 7; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 7) sums[8 * c + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kB) void k_fill_zero(uint32_t* __restrict__ p, int64_t words) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < words;
       i += static_cast<int64_t>(gridDim.x) * kB)
    p[i] = 0u;
}

__device__ __forceinline__ void set_mean(Leaf3& L, float m, const float* s) {
  L.alpha = 1.0f;
  L.mass = m;
  for (int a = 0; a < 3; ++a) {
    L.pos[a] = m > 0.0f ? s[a] / m : 0.0f;
    L.normal[a] = m > 0.0f ? s[3 + a] / m : 0.0f;
  }
}

// One block of kCells threads: thread c owns cell c.  Node and contrib
// buffers are zero-filled beforehand (absent cells leave zero records).
__global__ __launch_bounds__(kCells) void k_l3_nodes(const int32_t* __restrict__ offs,
                                                     const float* __restrict__ sums, int64_t n,
                                                     const float* __restrict__ lo,
                                                     const float* __restrict__ hi,
                                                     Node3* __restrict__ nodes,
                                                     Leaf3* __restrict__ contrib) {
  __shared__ int rank[kCells];
  __shared__ float tot[7][kCells];
  __shared__ uint64_t mask[kCells / 64];
  const int c = threadIdx.x;
  const int count = offs[c + 1] - offs[c];
  const int present = count > 0 ? 1 : 0;
  rank[c] = present;
  for (int a = 0; a < 7; ++a) tot[a][c] = sums[8 * c + a];
  const uint64_t bal = __ballot(present);
  if ((c & 63) == 0) mask[c >> 6] = bal;
  __syncthreads();
  // inclusive scan of the present flags, then the totals (fixed tree order)
  for (int off = 1; off < kCells; off <<= 1) {
    const int t = c >= off ? rank[c - off] : 0;
    __syncthreads();
    rank[c] += t;
    __syncthreads();
  }
  for (int off = kCells / 2; off > 0; off >>= 1) {
    if (c < off)
      for (int a = 0; a < 7; ++a) tot[a][c] += tot[a][c + off];
    __syncthreads();
  }
  const int child_idx = rank[c];  // 1 + exclusive rank
  float lo3[3], hi3[3], d3[3];
  for (int a = 0; a < 3; ++a) {
    lo3[a] = lo[a];
    hi3[a] = hi[a];
    d3[a] = (hi3[a] - lo3[a]) / static_cast<float>(kSplit);
  }
  if (present) {
    Node3* ch = &nodes[child_idx];
    const int q[3] = {c & (kSplit - 1), (c >> 3) & (kSplit - 1), (c >> 6) & (kSplit - 1)};
    ch->start = static_cast<uint64_t>(offs[c]);
    ch->end = static_cast<uint64_t>(offs[c] + count);
    ch->buffer_idx = static_cast<uint64_t>(child_idx);
    ch->parent_idx = 0;
    for (int a = 0; a < 3; ++a) {
      // child_min = min + d * q (contracted, as nvcc does by default)
      const float mn = fmaf(d3[a], static_cast<float>(q[a]), lo3[a]);
      ch->bmin[a] = mn;
      ch->bmax[a] = mn + d3[a];
    }
    {
#pragma clang fp contract(off)
      ch->diameter = sqrtf(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
    }
    Leaf3 L;
    set_mean(L, sums[8 * c], sums + 8 * c + 1);
    ch->contrib = L;
    ch->num_children = 0;
    ch->depth = 1;
    ch->max_depth = 1;
    ch->is_leaf = 1;
    contrib[child_idx] = L;
    nodes[0].child[c] = static_cast<uint64_t>(child_idx);
  }
  if (c == 0) {
    Node3* r = &nodes[0];
    r->start = 0;
    r->end = static_cast<uint64_t>(n);
    r->buffer_idx = 0;
    r->parent_idx = 0;
    float dd[3];
    for (int a = 0; a < 3; ++a) {
      r->bmin[a] = lo3[a];
      r->bmax[a] = hi3[a];
      dd[a] = hi3[a] - lo3[a];
    }
    {
#pragma clang fp contract(off)
      r->diameter = sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
    }
    const float s[6] = {tot[1][0], tot[2][0], tot[3][0], tot[4][0], tot[5][0], tot[6][0]};
    Leaf3 L;
    set_mean(L, tot[0][0], s);
    r->contrib = L;
    for (int w = 0; w < kCells / 64; ++w) r->bitmask[w] = mask[w];
    r->num_children = rank[kCells - 1];
    r->depth = 0;
    r->max_depth = 1;
    r->is_leaf = n <= 1 ? 1 : 0;
    contrib[0] = L;
  }
}

// ------------------------------------------------------------------ walks
// The walks read Tree3DNode<1, 3> records (level-3 buffers, or CPU-layout
// octrees) and Tree2DNode<1, 3> records (quadtrees) alike.
enum L3Kind { kPotential = 0, kForce = 1, kSoftForce3 = 2, kWn = 3, kSmooth = 4 };

// Kinds with one value per query.
template <int KIND>
constexpr bool kScalarL3 = KIND == kPotential || KIND == kWn || KIND == kSmooth;
// Kinds whose terms also read a record's normal and alpha (the wide rope).
template <int KIND>
constexpr bool kFullRope = KIND == kWn || KIND == kSmooth;

constexpr float kPiF = 3.14159265358979323846f;

template <int DIM, int LOG>
struct L3T {
  using Node = NodeT<DIM, LOG>;
  using Leaf = LeafT<DIM>;
  static constexpr int kSplitN = 1 << LOG;  // cells per axis
  static constexpr int kCellsN = Node::kCells;
  static constexpr int kWords = Node::kWords;
};

struct L3Params {
  o8::WalkParams P;
  float beta;
  int threads;  // the caller's block size: reference warps are 32 queries of a block
};

// The reference builds its GPU kernels with --use_fast_math
// (CMakeLists.txt:34-39): square roots and quotients in the kernel
// functions below are the hardware approximations (1 ulp), not the
// correctly rounded sequences.  Decisions (the MAC, cell quantisation) keep
// IEEE arithmetic so that they equal the oracle's.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Contribution of one record to a query (gravityPotential, gravityGrad,
// gravityGradSoftCutoff<N>: include/kernel.h:55-60, 89-94; src/sbh_ffi.cu:142-160).
template <int KIND>
__device__ __forceinline__ void l3_term(const Leaf3& L, const float* q,
                                        const o8::WalkParams& P, float* acc) {
  const float qx = q[0], qy = q[1], qz = q[2];
  if constexpr (KIND == kPotential || KIND == kWn || KIND == kSmooth) {
    const float dx = min_image_floor(L.pos[0] - qx, P.periodic, P.box, P.inv_box);
    const float dy = min_image_floor(L.pos[1] - qy, P.periodic, P.box, P.inv_box);
    const float dz = min_image_floor(L.pos[2] - qz, P.periodic, P.box, P.inv_box);
    const float dist = hw_sqrt(dx * dx + dy * dy + dz * dz);
    if constexpr (KIND == kPotential) {
      acc[0] += hw_div(-L.mass, dist + 1e-1f);
    } else if constexpr (KIND == kWn) {
      // signedSolidAngle<3> (include/kernel.h:73-80)
      const float qn = L.normal[0] * dx + L.normal[1] * dy + L.normal[2] * dz;
      acc[0] += hw_div(L.mass * qn, 4.0f * kPiF * (dist * dist * dist + 1e-5f));
    } else {
      // smoothDistExp (include/kernel.h:82-87)
      acc[0] += expf(-L.alpha * dist);
    }
  } else if constexpr (KIND == kForce) {
    o8::contribution<o8::kGravGrad>(L, qx, qy, qz, 0.f, 0.f, 0.f, P, acc[0], acc[1], acc[2]);
  } else {
    o8::contribution<o8::kSoftForce>(L, qx, qy, qz, 0.f, 0.f, 0.f, P, acc[0], acc[1], acc[2]);
  }
}

template <int KIND>
__device__ __forceinline__ void l3_term(const Leaf2& L, const float* q,
                                        const o8::WalkParams& P, float* acc) {
  if constexpr (KIND == kPotential || KIND == kWn || KIND == kSmooth) {
    // d = p - q
    const float dx = min_image_floor(L.pos[0] - q[0], P.periodic, P.box, P.inv_box);
    const float dy = min_image_floor(L.pos[1] - q[1], P.periodic, P.box, P.inv_box);
    if constexpr (KIND == kPotential) {
      // gravityPotential<2>
      acc[0] += hw_div(-L.mass, hw_sqrt(dx * dx + dy * dy) + 1e-1f);
    } else if constexpr (KIND == kWn) {
      // signedSolidAngle<2> (include/kernel.h:65-71)
      const float qn = L.normal[0] * dx + L.normal[1] * dy;
      acc[0] += hw_div(L.mass * qn, 2.0f * kPiF * (dx * dx + dy * dy + 1e-5f));
    } else {
      acc[0] += expf(-L.alpha * hw_sqrt(dx * dx + dy * dy));
    }
  } else {
    // gravityGrad<2> and gravityGradSoftCutoff<2>: d = q - p
    const float dx = min_image_floor(q[0] - L.pos[0], P.periodic, P.box, P.inv_box);
    const float dy = min_image_floor(q[1] - L.pos[1], P.periodic, P.box, P.inv_box);
    float c;
    if constexpr (KIND == kForce) {
      const float t = hw_sqrt(dx * dx + dy * dy) + 1e-1f;
      c = hw_div(-L.mass, t * t * t);
    } else {
      const float r2 = dx * dx + dy * dy;
      const float inv_r = rsq_softened(r2 + P.eps2 + 1e-12f);
      float att = 1.0f;
      if (P.cutoff) {
        float e, dist;
        att = attenuation(r2, P.rc, &e, &dist);
      }
      c = -L.mass * att * (inv_r * inv_r * inv_r);
    }
    acc[0] += c * dx;
    acc[1] += c * dy;
  }
}

// farFieldRatio (include/tree3d.h:134-141, include/tree2d.h:132-138):
// |q - centre of mass| / diameter, infinite for a leaf.
template <int DIM, int LOG>
__device__ __forceinline__ float far_ratio(const typename L3T<DIM, LOG>::Node* nd, const float* q) {
#pragma clang fp contract(off)
  if (nd->is_leaf) return INFINITY;
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const float d = q[a] - nd->contrib.pos[a];
    s = a == 0 ? d * d : s + d * d;
  }
  return sqrtf(s) / nd->diameter;
}

// Cell of point p in `parent` (AABB::quantize, include/aabb.h:36-42), or -1
// when p is outside parent's bounds widened by 1e-4 (AABB::contains).
template <int DIM, int LOG>
__device__ __forceinline__ int cell_in(const typename L3T<DIM, LOG>::Node* parent, const float* p) {
#pragma clang fp contract(off)
  constexpr int kS = L3T<DIM, LOG>::kSplitN;
  int q[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if (!(parent->bmin[a] - 1e-4f <= p[a] && p[a] <= parent->bmax[a] + 1e-4f)) return -1;
    const float t =
        static_cast<float>(kS) * (p[a] - parent->bmin[a]) / (parent->bmax[a] - parent->bmin[a]);
    q[a] = min(max(static_cast<int>(t), 0), kS - 1);
  }
  int c = 0;
#pragma unroll
  for (int a = DIM - 1; a >= 0; --a) c = c * kS + q[a];
  return c;
}

// First set bit of the node's mask at or after `from` (kCellsN if none).
template <int DIM, int LOG>
__device__ __forceinline__ int next_bit(const typename L3T<DIM, LOG>::Node* nd, int from) {
  constexpr int kN = L3T<DIM, LOG>::kCellsN;
  for (int w = from >> 6; w < L3T<DIM, LOG>::kWords; ++w) {
    uint64_t m = nd->bitmask[w];
    if (w == (from >> 6)) m &= ~0ull << (from & 63);
    if (m) return 64 * w + __builtin_ctzll(m);
  }
  return kN;
}

// Exact sum over the point records [j0, j1) in index order; records are
// fetched 8 at a time so that their loads overlap.
template <int KIND, class Leaf, class Params>
__device__ __forceinline__ void l3_leaf_sum(const Leaf* __restrict__ leaf, uint64_t j0,
                                            uint64_t j1, const float* q, const Params& P,
                                            float* acc) {
  uint64_t j = j0;
  for (; j + 8 <= j1; j += 8) {
    Leaf r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = leaf[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) l3_term<KIND>(r[u], q, P, acc);
  }
  for (; j < j1; ++j) l3_term<KIND>(leaf[j], q, P, acc);
}

template <int DIM, int KIND>
__device__ __forceinline__ void store_out(float* __restrict__ out, int q, const float* acc) {
  if constexpr (kScalarL3<KIND>) {
    out[q] = acc[0];
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) out[static_cast<int64_t>(DIM) * q + a] = acc[a];
  }
}

// Vote walk (include/barnes_hut_gpu.h:19-72): stackless traversal in child
// order; a node is used as a far field only if every active query of its
// 32-query reference warp accepts it, so a warp's queries share one path.
// Launched with the caller's block size, every reference warp is one half
// of a 64-wide wavefront here, and the vote is a ballot restricted to that
// half.
// PER: periodicity as a template argument (no minimum-image arithmetic in
// the non-periodic walk; same bits either way).
template <int DIM, int LOG, int KIND, bool PER>
__global__ void k_l3_walk(const typename L3T<DIM, LOG>::Leaf* __restrict__ leaf,
                          const typename L3T<DIM, LOG>::Node* __restrict__ nodes,
                          const float* __restrict__ queries, int nq, float* __restrict__ out,
                          L3Params A, uint32_t num_nodes, uint32_t num_leaves) {
  A.P.periodic = PER ? 1 : 0;
  using Node = typename L3T<DIM, LOG>::Node;
  constexpr int kN = L3T<DIM, LOG>::kCellsN;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = q < nq;
  const int lane = __lane_id();
  const uint64_t half = (lane < 32) ? 0x00000000ffffffffull : 0xffffffff00000000ull;
  const uint64_t members = __ballot(active) & half;
  if (!active) return;
  float qv[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) qv[a] = queries[static_cast<int64_t>(DIM) * q + a];
  float acc[DIM] = {};
  uint32_t node = 0;
  bool from_child = false;
  // every node is entered once (from its parent or its previous sibling) and
  // left once upwards, so 2 num_nodes + 1 steps end any walk of a valid tree
  const uint64_t guard_max = 4ull * num_nodes + 64ull;
  for (uint64_t guard = 0; guard < guard_max; ++guard) {
    const Node* nd = &nodes[node];
    if (!from_child) {
      const bool far = far_ratio<DIM, LOG>(nd, qv) >= A.beta;
      const bool all = (__ballot(far) & members) == members;
      if (all) {
        if (nd->is_leaf) {
          const uint64_t je = nd->end < num_leaves ? nd->end : num_leaves;
          l3_leaf_sum<KIND>(leaf, nd->start, je, qv, A.P, acc);
        } else {
          l3_term<KIND>(nd->contrib, qv, A.P, acc);
        }
      } else {
        const int b = next_bit<DIM, LOG>(nd, 0);
        if (b < kN && nd->child[b] < num_nodes) {
          node = static_cast<uint32_t>(nd->child[b]);
          continue;
        }
      }
    }
    if (node == 0) break;
    // next sibling: the parent's next present cell after this node's cell
    const Node* parent = &nodes[nd->parent_idx < num_nodes ? nd->parent_idx : 0];
    float mid[DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a) mid[a] = (nd->bmin[a] + nd->bmax[a]) / 2.0f;
    const int cell = cell_in<DIM, LOG>(parent, mid);
    const int b = cell < 0 ? kN : next_bit<DIM, LOG>(parent, cell + 1);
    if (b >= kN || parent->child[b] >= num_nodes) {
      node = static_cast<uint32_t>(nd->parent_idx < num_nodes ? nd->parent_idx : 0);
      from_child = true;
    } else {
      node = static_cast<uint32_t>(parent->child[b]);
      from_child = false;
    }
  }
  store_out<DIM, KIND>(out, q, acc);
}

// ------------------------------------------------------------- rope walk
// For trees of many nodes the vote walk runs over a compact copy of the
// node records ("rope"): centre of mass, mass, diameter, point range with
// the leaf flag, first child, and the escape link = the node the reference's
// stackless traversal moves to after finishing this node (its next sibling by
// AABB::quantize of the midpoint in the parent, else the parent's escape;
// none past the root).  Visit order, votes and sums are those of k_l3_walk;
// each step is one dependent load instead of the node, its parent and the
// child tables.  A leaf's centre of mass is never read (leaves always
// accept), so a leaf of one point carries that point's record (position,
// mass, normal, alpha) in those fields and needs no second load.  The walk
// is bound by vector instruction issue (eight waves per SIMD hide the load
// latency), so its loop keeps the exact ratio off the common path.  Measured
// slower: fetching both possible successors ahead of the vote (the selects
// cost more than the latency they hide), and votes kept wholly in scalar
// masks (the scalar unit, shared by a CU's four SIMDs, becomes the bound).
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kSumBit = 0x40000000u;
constexpr uint32_t kIndexMask = 0x3fffffffu;
constexpr uint32_t kRopeMin = 2048;  // fewer nodes: the plain walk
// Relative margins of the far-field test's sure answers (k_l3_walk_rope).
constexpr float kMarginHi = 1.0f + 0x1p-14f;
constexpr float kMarginLo = 1.0f - 0x1p-13f;  // applied to the sure-far bound

// FULL records also carry the normal and alpha (winding number, smooth
// distance); the others only what the gravity kernels read.  Per call (the
// records are rebuilt for every walk): start = first point | kLeafBit for a
// leaf | kSumBit for a leaf whose points need the leaf sum (any count but
// one); aux = end of the point range for a leaf, else the bits of the sure-
// far bound (beta diameter)^2 (1 + 2^-14), NaN where that bound is not
// usable; open = the node to visit when this one is opened (its first
// child, else its escape).
template <int DIM, bool FULL>
struct Rope;
template <>
struct Rope<2, false> {
  float c[2];
  float m, diam;
  uint32_t start, aux, open, escape;
};
template <>
struct Rope<3, false> {
  float c[3];
  float m, diam;
  uint32_t start, aux, open, escape, pad[3];
};
template <>
struct Rope<2, true> {
  float c[2];
  float m, diam;
  float n[2];
  float alpha;
  uint32_t start, aux, open, escape, pad;
};
template <>
struct Rope<3, true> {
  float c[3];
  float m, diam;
  float n[3];
  float alpha;
  uint32_t start, aux, open, escape, pad[3];
};
static_assert(sizeof(Rope<2, false>) == 32 && sizeof(Rope<3, false>) == 48 &&
                  sizeof(Rope<2, true>) == 48 && sizeof(Rope<3, true>) == 64,
              "rope records");

template <int DIM, int LOG>
__device__ __forceinline__ uint32_t next_sibling(const typename L3T<DIM, LOG>::Node* __restrict__ nodes,
                                                 uint32_t c, uint32_t num_nodes) {
  using Node = typename L3T<DIM, LOG>::Node;
  constexpr int kN = L3T<DIM, LOG>::kCellsN;
  const Node* nd = &nodes[c];
  const Node* parent = &nodes[nd->parent_idx < num_nodes ? nd->parent_idx : 0];
  float mid[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) mid[a] = (nd->bmin[a] + nd->bmax[a]) / 2.0f;
  const int cell = cell_in<DIM, LOG>(parent, mid);
  const int b = cell < 0 ? kN : next_bit<DIM, LOG>(parent, cell + 1);
  if (b >= kN || parent->child[b] >= num_nodes) return kNone;
  return static_cast<uint32_t>(parent->child[b]);
}

template <int DIM, int LOG, bool FULL>
__global__ __launch_bounds__(kB) void k_l3_rope(const typename L3T<DIM, LOG>::Node* __restrict__ nodes,
                                                const typename L3T<DIM, LOG>::Leaf* __restrict__ leaf,
                                                uint32_t num_nodes, uint32_t num_leaves,
                                                float beta, Rope<DIM, FULL>* __restrict__ rope) {
  using Node = typename L3T<DIM, LOG>::Node;
  constexpr int kN = L3T<DIM, LOG>::kCellsN;
  const uint32_t c = blockIdx.x * kB + threadIdx.x;
  if (c >= num_nodes) return;
  const Node* nd = &nodes[c];
  Rope<DIM, FULL> r = {};
#pragma unroll
  for (int a = 0; a < DIM; ++a) r.c[a] = nd->contrib.pos[a];
  r.m = nd->contrib.mass;
  if constexpr (FULL) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) r.n[a] = nd->contrib.normal[a];
    r.alpha = nd->contrib.alpha;
  }
  r.diam = nd->diameter;
  const uint64_t je = nd->end < num_leaves ? nd->end : num_leaves;
  const uint64_t js = nd->start < je ? nd->start : je;
  r.start = static_cast<uint32_t>(js) |
            (nd->is_leaf ? (je - js == 1 ? kLeafBit : kLeafBit | kSumBit) : 0u);
  if (nd->is_leaf) {
    r.aux = static_cast<uint32_t>(je);
  } else {
    // the sure-far bound of the walk's far-field test
#pragma clang fp contract(off)
    const float t = beta * nd->diameter;
    const float t2 = t * t;
    const float hi = t2 * kMarginHi;
    // a non-positive diameter (node bytes not written by this library's
    // builders) has no usable bound: NaN sends every vote to the exact ratio
    r.aux = __float_as_uint(beta > 0.0f && nd->diameter > 0.0f && t2 >= FLT_MIN &&
                                    hi <= FLT_MAX
                                ? hi
                                : NAN);
  }
  if (nd->is_leaf && je - js == 1) {
    // the leaf's only point, as the leaf sum would read it
    const auto& L = leaf[js];
#pragma unroll
    for (int a = 0; a < DIM; ++a) r.c[a] = L.pos[a];
    r.m = L.mass;
    if constexpr (FULL) {
#pragma unroll
      for (int a = 0; a < DIM; ++a) r.n[a] = L.normal[a];
      r.alpha = L.alpha;
    }
  }
  const int b = next_bit<DIM, LOG>(nd, 0);
  const uint32_t first =
      (b < kN && nd->child[b] < num_nodes) ? static_cast<uint32_t>(nd->child[b]) : kNone;
  // finishing node c: next sibling, else climb (the reference's FromChild
  // state) until a node with a next sibling; the root ends the walk
  uint32_t esc = kNone, cur = c;
  for (int hop = 0; hop < 256 && cur != 0; ++hop) {
    const uint32_t s = next_sibling<DIM, LOG>(nodes, cur, num_nodes);
    if (s != kNone) {
      esc = s;
      break;
    }
    const uint64_t pi = nodes[cur].parent_idx;
    cur = static_cast<uint32_t>(pi < num_nodes ? pi : 0);
  }
  r.escape = esc;
  r.open = first != kNone ? first : esc;
  rope[c] = r;
}

// farFieldRatio >= beta (far_ratio(): correctly rounded square root and
// quotient) is decided from the squared distance s alone where s is clear of
// (beta diameter)^2 by a relative 2^-14: the roundings of the exact
// expression and of (beta diameter)^2 move it by under 2^-21, so both give
// the same answer there.  Everything else (a margin hit, beta <= 0, zero,
// huge, infinite or NaN operands) takes the exact expression.

template <int DIM, bool FULL>
__device__ __forceinline__ Rope<DIM, FULL> rope_at(const char* __restrict__ base, uint32_t c) {
  // record offsets fit 32 bits (the launcher checks): a scalar base and a
  // 32-bit vector offset per load
  return *reinterpret_cast<const Rope<DIM, FULL>*>(base + c * static_cast<uint32_t>(sizeof(Rope<DIM, FULL>)));
}

template <int DIM, int LOG, int KIND, bool PER>
__global__ void k_l3_walk_rope(const typename L3T<DIM, LOG>::Leaf* __restrict__ leaf,
                               const Rope<DIM, kFullRope<KIND>>* __restrict__ rope,
                               const float* __restrict__ queries, int nq,
                               float* __restrict__ out, L3Params A, uint32_t num_nodes) {
  using Leaf = typename L3T<DIM, LOG>::Leaf;
  using R = Rope<DIM, kFullRope<KIND>>;
  // periodicity is a template argument here: the non-periodic walk carries
  // no minimum-image arithmetic
  o8::WalkParams P = A.P;
  P.periodic = PER ? 1 : 0;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = q < nq;
  const int lane = __lane_id();
  const uint64_t half = (lane < 32) ? 0x00000000ffffffffull : 0xffffffff00000000ull;
  const uint64_t members = __ballot(active) & half;
  if (!active) return;
  float qv[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) qv[a] = queries[static_cast<int64_t>(DIM) * q + a];
  float acc[DIM] = {};
  const char* __restrict__ base = reinterpret_cast<const char*>(rope);
  // a leaf's ratio is infinite: far for any beta but NaN
  const bool leaf_far = INFINITY >= A.beta;
  const uint32_t guard_max = num_nodes > 0x3ffffff0u ? 0xffffffffu : 4u * num_nodes + 64u;
  uint32_t c = 0;
  for (uint32_t guard = 0; guard < guard_max && c < num_nodes; ++guard) {
    const R r = rope_at<DIM, kFullRope<KIND>>(base, c);
    const bool is_leaf = (r.start & kLeafBit) != 0u;
    uint64_t far_mask;
    {
#pragma clang fp contract(off)
      // The difference takes the sign opposite to the kind's pair kernel
      // (p - q for the scalar kinds, q - p for the forces): the same s, bit
      // for bit, but no product the compiler could share with the pair
      // kernel of a one-point leaf (whose record holds the point), so the
      // pair kernel's contraction does not depend on the walk around it.
      float s = 0.0f;
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        const float d = kScalarL3<KIND> ? qv[a] - r.c[a] : r.c[a] - qv[a];
        s = a == 0 ? d * d : s + d * d;
      }
      // hi = (beta diam)^2 (1 + 2^-14), lo = hi (1 - 2^-13) ~ (beta diam)^2
      // (1 - 2^-14); NaN (no sure answer) where the bound is not usable
      const float hi = __uint_as_float(r.aux);
      const bool sure_far = is_leaf ? leaf_far : s > hi;
      const bool sure_near = is_leaf ? !leaf_far : s < hi * kMarginLo;
      far_mask = __ballot(sure_far) & members;
      // the exact ratio only for a vote that the sure answers leave open
      if (far_mask != members && (__ballot(sure_near) & members) == 0)
        far_mask |= __ballot(!sure_far && sqrtf(s) / r.diam >= A.beta) & members;
    }
    const bool all = far_mask == members;
    if (all) {
      if (r.start & kSumBit) {
        l3_leaf_sum<KIND>(leaf, r.start & kIndexMask, r.aux, qv, P, acc);
      } else {
        // an internal node's monopole, or a one-point leaf's point
        Leaf L = {};
#pragma unroll
        for (int a = 0; a < DIM; ++a) L.pos[a] = r.c[a];
        L.mass = r.m;
        if constexpr (kFullRope<KIND>) {
#pragma unroll
          for (int a = 0; a < DIM; ++a) L.normal[a] = r.n[a];
          L.alpha = r.alpha;
        }
        l3_term<KIND>(L, qv, P, acc);
      }
    }
    c = all ? r.escape : r.open;
  }
  store_out<DIM, KIND>(out, q, acc);
}

// Node of the child cell of `node` that holds p, skipping children that have
// exactly one child (Tree2DNode / Tree3DNode::containedInChild(p),
// include/tree2d.h:168-179); -1 when p is outside a box on the way (the
// reference returns nullptr there).  An absent cell reads child id 0, as the
// reference's zero-filled child table does.
template <int DIM, int LOG>
__device__ __forceinline__ int64_t contained_child(const typename L3T<DIM, LOG>::Node* __restrict__ nodes,
                                                   uint64_t node, const float* p,
                                                   uint32_t num_nodes) {
  for (int hop = 0; hop < 128; ++hop) {
    const int cell = cell_in<DIM, LOG>(&nodes[node], p);
    if (cell < 0) return -1;
    const uint64_t ch = nodes[node].child[cell];
    if (ch >= num_nodes) return -1;
    if (nodes[ch].num_children != 1) return static_cast<int64_t>(ch);
    node = ch;
  }
  return -1;
}

// Multi-level prefix control-variate walk (include/tree_sampler_gpu.h:25-140).
// Subdomains are the root's children: each adds its exact sum (leaf) or its
// monopole, plus, per sample, a sample point's roulette path whose steps add
// n / (P count(prev)) (sum over prev's children - prev's monopole).
// Random numbers: Philox4x32-10 in place of cuRAND XORWOW.  As in the
// reference, the sample point is shared by all queries of a block (its
// stream depends only on the block index) and the roulette stream on
// seed + query.
template <int DIM, int LOG, int KIND>
__global__ void k_l3_stoch(const typename L3T<DIM, LOG>::Leaf* __restrict__ leaf,
                           const typename L3T<DIM, LOG>::Node* __restrict__ nodes,
                           const typename L3T<DIM, LOG>::Leaf* __restrict__ contrib,
                           const float* __restrict__ queries, int nq, float* __restrict__ out,
                           o8::WalkParams P, int samples, int seed, uint32_t seed_per_query,
                           uint32_t num_nodes,
                           uint32_t num_leaves, uint32_t num_contrib) {
  using Node = typename L3T<DIM, LOG>::Node;
  constexpr int kN = L3T<DIM, LOG>::kCellsN;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  float qv[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) qv[a] = queries[static_cast<int64_t>(DIM) * q + a];
  float acc[DIM] = {};
  float mlcv[DIM] = {};
  // the FFI targets seed the roulette with seed + query; the pybind11 module
  // gives every query seed 0 (src/sbh_gpu.cu:143-164)
  const uint32_t rseed = static_cast<uint32_t>(seed) + seed_per_query * static_cast<uint32_t>(q);
  const Node* root = &nodes[0];
  for (int b = next_bit<DIM, LOG>(root, 0); b < kN; b = next_bit<DIM, LOG>(root, b + 1)) {
    const uint64_t ci = root->child[b];
    if (ci >= num_nodes) continue;
    const Node* sub = &nodes[ci];
    if (sub->is_leaf) {
      const uint64_t je = sub->end < num_leaves ? sub->end : num_leaves;
      l3_leaf_sum<KIND>(leaf, sub->start, je, qv, P, acc);
    } else {
      l3_term<KIND>(sub->contrib, qv, P, acc);
    }
    const uint64_t count = sub->end - sub->start;
    float total[DIM] = {};
    for (int ss = 0; ss < samples; ++ss) {
      const float u2 = sbh_uniform(blockIdx.x, 1u, static_cast<uint32_t>(ci),
                                   static_cast<uint32_t>(ss), 0u);
      uint64_t idx = static_cast<uint64_t>(
          fmaf(static_cast<float>(count), u2, static_cast<float>(sub->start)));
      if (idx > sub->end - 1) idx = sub->end - 1;
      if (count == 0 || idx >= num_leaves) continue;
      float sp[DIM];
#pragma unroll
      for (int a = 0; a < DIM; ++a) sp[a] = leaf[idx].pos[a];
      uint64_t cur = ci;
      float event_prob = 1.0f;
      float wsum[DIM] = {};
      float cur_beta = fmaxf(1.0f, far_ratio<DIM, LOG>(sub, qv));
      // a path ends at a leaf; depth <= 100 in trees from the builders
      for (uint32_t step = 0; step < 1024u && !nodes[cur].is_leaf; ++step) {
        const Node* cn = &nodes[cur];
        const float ffr = far_ratio<DIM, LOG>(cn, qv);
        const float tp = fminf(cur_beta / ffr, 1.0f);
        const float pu = sbh_uniform(rseed, 0u, static_cast<uint32_t>(ci),
                                     static_cast<uint32_t>(ss), step);
        if (pu > tp) break;
        event_prob *= tp;
        const int64_t nxt = contained_child<DIM, LOG>(nodes, cur, sp, num_nodes);
        cur_beta = fmaxf(cur_beta, ffr);
        float prev_res[DIM] = {}, cur_res[DIM] = {};
        l3_term<KIND>(cn->contrib, qv, P, prev_res);
        const int fb = next_bit<DIM, LOG>(cn, 0);
        const uint64_t first = fb < kN ? cn->child[fb] : 0ull;
        for (int32_t o = 0; o < cn->num_children; ++o) {
          const uint64_t c = first + static_cast<uint64_t>(o);
          if (c >= num_nodes) break;
          const Node* ch = &nodes[c];
          if (ch->is_leaf) {
            const uint64_t je = ch->end < num_leaves ? ch->end : num_leaves;
            l3_leaf_sum<KIND>(leaf, ch->start, je, qv, P, cur_res);
          } else {
            l3_term<KIND>(c < num_contrib ? contrib[c] : ch->contrib, qv, P, cur_res);
          }
        }
        const float w = static_cast<float>(count) /
                        (event_prob * static_cast<float>(cn->end - cn->start));
#pragma unroll
        for (int a = 0; a < DIM; ++a) wsum[a] += w * (cur_res[a] - prev_res[a]);
        if (nxt < 0) break;  // the reference would dereference nullptr here
        cur = static_cast<uint64_t>(nxt);
      }
#pragma unroll
      for (int a = 0; a < DIM; ++a) total[a] += wsum[a];
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) mlcv[a] += total[a] / static_cast<float>(samples);
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) acc[a] += mlcv[a];
  store_out<DIM, KIND>(out, q, acc);
}

template <int DIM, int LOG>
int check_l3_walk(const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                  const SbhBuffer* contrib_bytes, const SbhBuffer* queries,
                  const SbhBuffer* out, int threads, bool vec_out) {
  using Node = typename L3T<DIM, LOG>::Node;
  using Leaf = typename L3T<DIM, LOG>::Leaf;
  if (!buffer_ok(leaf_bytes, SBH_DTYPE_U8, -1) || !buffer_ok(node_bytes, SBH_DTYPE_U8, -1))
    return fail_invalid("leaf_bytes/node_bytes must be uint8 device buffers");
  if (byte_count(leaf_bytes) % static_cast<int64_t>(sizeof(Leaf)) != 0)
    return fail_invalid("leaf_bytes size mismatch");
  if (byte_count(node_bytes) % static_cast<int64_t>(sizeof(Node)) != 0 ||
      byte_count(node_bytes) == 0)
    return fail_invalid("node_bytes size mismatch");
  if (byte_count(node_bytes) / static_cast<int64_t>(sizeof(Node)) > INT32_MAX)
    return fail_invalid("too many nodes");
  if (contrib_bytes != nullptr &&
      (!buffer_ok(contrib_bytes, SBH_DTYPE_U8, -1) ||
       byte_count(contrib_bytes) % static_cast<int64_t>(sizeof(Leaf)) != 0))
    return fail_invalid("contrib_bytes size mismatch");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2))
    return fail_invalid(DIM == 2 ? "queries must be (M, 2) float32" : "queries must be (M, 3) float32");
  if (queries->dims[1] != DIM)
    return fail_invalid(DIM == 2 ? "queries dim must be 2" : "queries dim must be 3");
  if (queries->dims[0] > INT32_MAX) return fail_invalid("too many queries");
  if (vec_out ? (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != queries->dims[0] ||
                 out->dims[1] != DIM)
              : (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != queries->dims[0]))
    return fail_invalid("out shape mismatch");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  return SBH_OK;
}

template <int DIM, int LOG, int KIND>
int launch_l3_walk(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                   const SbhBuffer* queries, SbhBuffer* out, float beta, float softening,
                   float cutoff, int periodic, float box, int threads) {
  using Node = typename L3T<DIM, LOG>::Node;
  using Leaf = typename L3T<DIM, LOG>::Leaf;
  int rc = check_l3_walk<DIM, LOG>(leaf_bytes, node_bytes, nullptr, queries, out, threads,
                              !kScalarL3<KIND>);
  if (rc != SBH_OK) return rc;
  const int nq = static_cast<int>(queries->dims[0]);
  if (nq == 0) return SBH_OK;
  L3Params A;
  A.P = o8::make_walk_params(beta, softening, cutoff, 0, 0.f, periodic, box);
  // gravityGradSoftCutoff applies the attenuation for any cutoff > 0; at
  // +inf it is exactly 1 (u = 0), so it is skipped there
  A.P.cutoff = (cutoff > 0.0f && std::isfinite(cutoff)) ? 1 : 0;
  A.beta = beta;
  A.threads = threads;
  const unsigned blocks = static_cast<unsigned>((nq + threads - 1) / threads);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t num_nodes = static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node));
  const uint32_t num_leaves = static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf));
  const auto* nodes = static_cast<const Node*>(node_bytes->data);
  // the rope needs a workspace (not under stream capture: no allocation
  // there), record offsets within 32 bits and point indices within 30
  using R = Rope<DIM, kFullRope<KIND>>;
  R* rope = nullptr;
  if (num_nodes >= kRopeMin && static_cast<uint64_t>(num_nodes) * sizeof(R) <= (1ull << 32) &&
      num_leaves <= kIndexMask && !stream_capturing(s))
    rope = static_cast<R*>(workspace(s, kWsRope, static_cast<size_t>(num_nodes) * sizeof(R)));
  if (rope != nullptr) {
    k_l3_rope<DIM, LOG, kFullRope<KIND>><<<grid_for(num_nodes, kB), kB, 0, s>>>(
        nodes, static_cast<const Leaf*>(leaf_bytes->data), num_nodes, num_leaves, beta, rope);
    SBH_LAUNCH_CHECK();
    auto walk = A.P.periodic ? k_l3_walk_rope<DIM, LOG, KIND, true>
                             : k_l3_walk_rope<DIM, LOG, KIND, false>;
    walk<<<blocks, threads, 0, s>>>(static_cast<const Leaf*>(leaf_bytes->data), rope,
                                    static_cast<const float*>(queries->data), nq,
                                    static_cast<float*>(out->data), A, num_nodes);
  } else {
    auto plain = A.P.periodic ? k_l3_walk<DIM, LOG, KIND, true> : k_l3_walk<DIM, LOG, KIND, false>;
    plain<<<blocks, threads, 0, s>>>(
        static_cast<const Leaf*>(leaf_bytes->data), nodes,
        static_cast<const float*>(queries->data), nq, static_cast<float*>(out->data), A,
        num_nodes, num_leaves);
  }
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

template <int DIM, int LOG, int KIND>
int launch_l3_stoch(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                    const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
                    int samples, int seed, float softening, float cutoff, int periodic,
                    float box, int threads, uint32_t seed_per_query = 1u) {
  using Node = typename L3T<DIM, LOG>::Node;
  using Leaf = typename L3T<DIM, LOG>::Leaf;
  int rc = check_l3_walk<DIM, LOG>(leaf_bytes, node_bytes, contrib_bytes, queries, out, threads,
                              !kScalarL3<KIND>);
  if (rc != SBH_OK) return rc;
  if (contrib_bytes == nullptr) return fail_invalid("contrib_bytes must be given");
  const int nq = static_cast<int>(queries->dims[0]);
  if (nq == 0) return SBH_OK;
  o8::WalkParams P = o8::make_walk_params(1.0f, softening, cutoff, 0, 0.f, periodic, box);
  P.cutoff = (cutoff > 0.0f && std::isfinite(cutoff)) ? 1 : 0;
  const unsigned blocks = static_cast<unsigned>((nq + threads - 1) / threads);
  k_l3_stoch<DIM, LOG, KIND><<<blocks, threads, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const Leaf*>(leaf_bytes->data), static_cast<const Node*>(node_bytes->data),
      static_cast<const Leaf*>(contrib_bytes->data), static_cast<const float*>(queries->data),
      nq, static_cast<float*>(out->data), P, samples, seed, seed_per_query,
      static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node)),
      static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf)),
      static_cast<uint32_t>(byte_count(contrib_bytes) / sizeof(Leaf)));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

// --------------------------------------------------- pybind11 GPU module
// The reference's GPU module (src/sbh_gpu.cu) evaluates gravityPotential,
// signedSolidAngle<N> and smoothDistExp over Tree2D / Tree3D<1, L> buffers:
// brute force (compensated sums over the leaf records in tree order,
// include/brute_force_gpu.h:18-66), the vote walk and the control-variate
// sampler, all with 32-thread blocks; the sampler passes seed 0 to every
// query (src/sbh_gpu.cu:143-211, 436-550).  The sampler keeps 32-thread
// blocks (its sample point is drawn per block).  The walk's vote groups are
// 32 consecutive queries for any block size that is a multiple of 32 (each
// group is one half of a wave), so it runs in 256-thread blocks: the same
// votes and sums, with 64 lanes per wave instead of 32.
constexpr int kModuleThreads = 32;
constexpr int kModuleWalkThreads = 256;

template <int DIM, int KIND>
__global__ __launch_bounds__(kB) void k_tree_brute(const LeafT<DIM>* __restrict__ leaf,
                                                   int64_t n, const float* __restrict__ queries,
                                                   int nq, float* __restrict__ out,
                                                   o8::WalkParams P) {
  // the module's fields are never periodic: no minimum-image arithmetic
  P.periodic = 0;
  __shared__ LeafT<DIM> tile[kB];
  const int q = blockIdx.x * kB + threadIdx.x;
  float qv[DIM] = {};
  if (q < nq)
#pragma unroll
    for (int a = 0; a < DIM; ++a) qv[a] = queries[static_cast<int64_t>(DIM) * q + a];
  float total = 0.0f, comp = 0.0f;
  for (int64_t i = 0; i < n; i += kB) {
    if (i + threadIdx.x < n) tile[threadIdx.x] = leaf[i + threadIdx.x];
    __syncthreads();
    if (q < nq) {
      const int cnt = static_cast<int>(n - i < kB ? n - i : kB);
      for (int b = 0; b < cnt; ++b) {
        float v[DIM] = {};
        l3_term<KIND>(tile[b], qv, P, v);
        {
#pragma clang fp contract(off)
          const float y = v[0] - comp;
          const float t = total + y;
          comp = (t - total) - y;
          total = t;
        }
      }
    }
    __syncthreads();
  }
  if (q < nq) out[q] = total;
}

template <int DIM, int LOG>
int module_eval(void* stream, int kind, const SbhBuffer* leaf, const SbhBuffer* node,
                const SbhBuffer* contrib, const SbhBuffer* q, SbhBuffer* out, float beta,
                int samples) {
  constexpr int T = kModuleThreads;
  constexpr int W = kModuleWalkThreads;
  if (samples < 0) {
    if (kind == 0)
      return launch_l3_walk<DIM, LOG, kPotential>(stream, leaf, node, q, out, beta, 0.f, 0.f, 0,
                                                  1.f, W);
    if (kind == 1)
      return launch_l3_walk<DIM, LOG, kWn>(stream, leaf, node, q, out, beta, 0.f, 0.f, 0, 1.f,
                                           W);
    return launch_l3_walk<DIM, LOG, kSmooth>(stream, leaf, node, q, out, beta, 0.f, 0.f, 0,
                                             1.f, W);
  }
  if (kind == 0)
    return launch_l3_stoch<DIM, LOG, kPotential>(stream, leaf, node, contrib, q, out, samples,
                                                 0, 0.f, 0.f, 0, 1.f, T, 0u);
  if (kind == 1)
    return launch_l3_stoch<DIM, LOG, kWn>(stream, leaf, node, contrib, q, out, samples, 0, 0.f,
                                          0.f, 0, 1.f, T, 0u);
  return launch_l3_stoch<DIM, LOG, kSmooth>(stream, leaf, node, contrib, q, out, samples, 0,
                                            0.f, 0.f, 0, 1.f, T, 0u);
}

template <int DIM>
int module_eval_level(void* stream, int level, int kind, const SbhBuffer* leaf,
                      const SbhBuffer* node, const SbhBuffer* contrib, const SbhBuffer* q,
                      SbhBuffer* out, float beta, int samples) {
  if (level == 1) return module_eval<DIM, 1>(stream, kind, leaf, node, contrib, q, out, beta, samples);
  if (level == 2) return module_eval<DIM, 2>(stream, kind, leaf, node, contrib, q, out, beta, samples);
  return module_eval<DIM, 3>(stream, kind, leaf, node, contrib, q, out, beta, samples);
}

template <int DIM, int KIND>
int launch_brute(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* queries,
                 SbhBuffer* out) {
  const int nq = static_cast<int>(queries->dims[0]);
  if (nq == 0) return SBH_OK;
  const int64_t n = byte_count(leaf_bytes) / static_cast<int64_t>(sizeof(LeafT<DIM>));
  const o8::WalkParams P = o8::make_walk_params(1.0f, 0.f, 0.f, 0, 0.f, 0, 1.f);
  k_tree_brute<DIM, KIND><<<grid_for(nq, kB), kB, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const LeafT<DIM>*>(leaf_bytes->data), n,
      static_cast<const float*>(queries->data), nq, static_cast<float*>(out->data), P);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" int sbh_build_octree_level3_buffers(void* stream, const SbhBuffer* points,
                                               const SbhBuffer* normals, const SbhBuffer* masses,
                                               const SbhBuffer* min_corner,
                                               const SbhBuffer* max_corner, SbhBuffer* leaf_out,
                                               SbhBuffer* node_out, SbhBuffer* contrib_out) {
  clear_error();
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  if (!buffer_ok(normals, SBH_DTYPE_F32, 2) || normals->dims[1] != 3)
    return fail_invalid("normals must have shape (N, 3)");
  if (points->dims[0] != normals->dims[0]) return fail_invalid("points/normals row mismatch");
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != points->dims[0])
    return fail_invalid("masses must have shape (N,)");
  if (!buffer_ok(min_corner, SBH_DTYPE_F32, 1) || min_corner->dims[0] != 3 ||
      !buffer_ok(max_corner, SBH_DTYPE_F32, 1) || max_corner->dims[0] != 3)
    return fail_invalid("min_corner/max_corner must have shape (3,)");
  const int64_t n = points->dims[0];
  if (n > INT32_MAX) return fail_invalid("too many points");
  if (!buffer_ok(leaf_out, SBH_DTYPE_U8, -1) ||
      byte_count(leaf_out) != n * static_cast<int64_t>(sizeof(Leaf3)))
    return fail_invalid("leaf_out size mismatch");
  if (!buffer_ok(node_out, SBH_DTYPE_U8, -1) ||
      byte_count(node_out) != kMaxNodes3 * static_cast<int64_t>(sizeof(Node3)))
    return fail_invalid("node_out size mismatch");
  if (!buffer_ok(contrib_out, SBH_DTYPE_U8, -1) ||
      byte_count(contrib_out) != kMaxNodes3 * static_cast<int64_t>(sizeof(Leaf3)))
    return fail_invalid("contrib_out size mismatch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nn = static_cast<size_t>(n > 0 ? n : 1);
  // workspace: keys, vals, sorted cells, perm, sort temp, offsets, cell sums
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  size_t o = 0;
  const size_t o_keys = o; o = align(o + 4 * nn);
  const size_t o_vals = o; o = align(o + 4 * nn);
  const size_t o_cells = o; o = align(o + 4 * nn);
  const size_t o_perm = o; o = align(o + 4 * nn);
  const size_t o_sort = o; o = align(o + radix_sort_temp_bytes(n));
  const size_t o_offs = o; o = align(o + 4 * (kCells + 1));
  const size_t o_sums = o; o = align(o + 4 * 8 * kCells);
  char* ws = nullptr;
  if (int rcw = cached_workspace(s, kWsBuild, o, &ws); rcw != SBH_OK) return rcw;
  auto* keys = reinterpret_cast<uint32_t*>(ws + o_keys);
  auto* vals = reinterpret_cast<uint32_t*>(ws + o_vals);
  auto* cells = reinterpret_cast<uint32_t*>(ws + o_cells);
  auto* perm = reinterpret_cast<uint32_t*>(ws + o_perm);
  auto* offs = reinterpret_cast<int32_t*>(ws + o_offs);
  auto* sums = reinterpret_cast<float*>(ws + o_sums);
  const auto* pts = static_cast<const float*>(points->data);
  const auto* lo = static_cast<const float*>(min_corner->data);
  const auto* hi = static_cast<const float*>(max_corner->data);
  auto* leaf = static_cast<Leaf3*>(leaf_out->data);
  auto* nodes = static_cast<Node3*>(node_out->data);
  auto* contrib = static_cast<Leaf3*>(contrib_out->data);
  k_fill_zero<<<256, kB, 0, s>>>(reinterpret_cast<uint32_t*>(nodes),
                                 byte_count(node_out) / 4);
  SBH_LAUNCH_CHECK();
  k_fill_zero<<<16, kB, 0, s>>>(reinterpret_cast<uint32_t*>(contrib),
                                byte_count(contrib_out) / 4);
  SBH_LAUNCH_CHECK();
  if (n > 0) {
    k_l3_cells<<<grid_for(n, kB), kB, 0, s>>>(pts, n, lo, hi, keys, vals);
    SBH_LAUNCH_CHECK();
    int rc = radix_sort_pairs(s, keys, vals, cells, perm, n, 9, ws + o_sort);
    if (rc != SBH_OK) return rc;
    k_l3_gather<<<grid_for(n, kB), kB, 0, s>>>(pts, static_cast<const float*>(normals->data),
                                               static_cast<const float*>(masses->data), n, perm,
                                               leaf);
    SBH_LAUNCH_CHECK();
  }
  k_l3_offsets<<<grid_for(kCells + 1, kB), kB, 0, s>>>(cells, n, offs);
  SBH_LAUNCH_CHECK();
  k_l3_cell_sums<<<kCells, kB, 0, s>>>(leaf, offs, sums);
  SBH_LAUNCH_CHECK();
  k_l3_nodes<<<1, kCells, 0, s>>>(offs, sums, n, lo, hi, nodes, contrib);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_barnes_hut_gravity_3d(void* stream, const SbhBuffer* leaf_bytes,
                                         const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                         SbhBuffer* out, float beta, int32_t periodic,
                                         float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<3, 3, kPotential>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f,
                                       0.f, periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                                       const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                       SbhBuffer* out, float beta, int32_t periodic,
                                       float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<3, 3, kForce>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f,
                                   0.f, periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force_soft_3d(void* stream, const SbhBuffer* leaf_bytes,
                                            const SbhBuffer* node_bytes,
                                            const SbhBuffer* queries, SbhBuffer* out,
                                            float beta, float softening_scale,
                                            float cutoff_scale, int32_t periodic,
                                            float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<3, 3, kSoftForce3>(stream, leaf_bytes, node_bytes, queries, out, beta,
                                        softening_scale, cutoff_scale, periodic, box_length,
                                        threads);
}

extern "C" int sbh_stochastic_bh_gravity_3d(void* stream, const SbhBuffer* leaf_bytes,
                                            const SbhBuffer* node_bytes,
                                            const SbhBuffer* contrib_bytes,
                                            const SbhBuffer* queries, SbhBuffer* out,
                                            int32_t samples_per_subdomain, int32_t seed,
                                            int32_t periodic, float box_length,
                                            int32_t threads) {
  clear_error();
  return launch_l3_stoch<3, 3, kPotential>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                        out, samples_per_subdomain, seed, 0.f, 0.f, periodic,
                                        box_length, threads);
}

extern "C" int sbh_stochastic_bh_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                                          const SbhBuffer* node_bytes,
                                          const SbhBuffer* contrib_bytes,
                                          const SbhBuffer* queries, SbhBuffer* out,
                                          int32_t samples_per_subdomain, int32_t seed,
                                          int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  return launch_l3_stoch<3, 3, kForce>(stream, leaf_bytes, node_bytes, contrib_bytes, queries, out,
                                    samples_per_subdomain, seed, 0.f, 0.f, periodic, box_length,
                                    threads);
}

extern "C" int sbh_stochastic_bh_force_soft_3d(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
    int32_t samples_per_subdomain, int32_t seed, float softening_scale, float cutoff_scale,
    int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  return launch_l3_stoch<3, 3, kSoftForce3>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                         out, samples_per_subdomain, seed, softening_scale,
                                         cutoff_scale, periodic, box_length, threads);
}

// 2-D targets over Tree2DNode<1, 3> buffers (src/sbh_ffi.cu:2444-2669,
// 2779-3064): the same walks with two coordinates.
extern "C" int sbh_barnes_hut_gravity(void* stream, const SbhBuffer* leaf_bytes,
                                      const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                      SbhBuffer* out, float beta, int32_t periodic,
                                      float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<2, 3, kPotential>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f,
                                       0.f, periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force(void* stream, const SbhBuffer* leaf_bytes,
                                    const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                    SbhBuffer* out, float beta, int32_t periodic,
                                    float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<2, 3, kForce>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f,
                                   0.f, periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force_soft(void* stream, const SbhBuffer* leaf_bytes,
                                         const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                         SbhBuffer* out, float beta, float softening_scale,
                                         float cutoff_scale, int32_t periodic,
                                         float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<2, 3, kSoftForce3>(stream, leaf_bytes, node_bytes, queries, out, beta,
                                        softening_scale, cutoff_scale, periodic, box_length,
                                        threads);
}

extern "C" int sbh_stochastic_bh_gravity(void* stream, const SbhBuffer* leaf_bytes,
                                         const SbhBuffer* node_bytes,
                                         const SbhBuffer* contrib_bytes,
                                         const SbhBuffer* queries, SbhBuffer* out,
                                         int32_t samples_per_subdomain, int32_t seed,
                                         int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  return launch_l3_stoch<2, 3, kPotential>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                        out, samples_per_subdomain, seed, 0.f, 0.f, periodic,
                                        box_length, threads);
}

extern "C" int sbh_stochastic_bh_force(void* stream, const SbhBuffer* leaf_bytes,
                                       const SbhBuffer* node_bytes,
                                       const SbhBuffer* contrib_bytes, const SbhBuffer* queries,
                                       SbhBuffer* out, int32_t samples_per_subdomain,
                                       int32_t seed, int32_t periodic, float box_length,
                                       int32_t threads) {
  clear_error();
  return launch_l3_stoch<2, 3, kForce>(stream, leaf_bytes, node_bytes, contrib_bytes, queries, out,
                                    samples_per_subdomain, seed, 0.f, 0.f, periodic, box_length,
                                    threads);
}

extern "C" int sbh_stochastic_bh_force_soft(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
    int32_t samples_per_subdomain, int32_t seed, float softening_scale, float cutoff_scale,
    int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  return launch_l3_stoch<2, 3, kSoftForce3>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                         out, samples_per_subdomain, seed, softening_scale,
                                         cutoff_scale, periodic, box_length, threads);
}

// The pybind11 GPU module's evaluations (src/sbh_gpu.cu:324-550): kind 0
// potential, 1 winding number, 2 smooth distance; samples < 0: vote walk with
// `beta`, samples >= 0: control-variate sampler.
extern "C" int sbh_tree_eval(void* stream, int32_t dim, int32_t level, int32_t kind,
                             const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                             const SbhBuffer* contrib_bytes, const SbhBuffer* queries,
                             SbhBuffer* out, float beta, int32_t samples_per_subdomain) {
  clear_error();
  if (dim != 2 && dim != 3) return fail_invalid("dim must be 2 or 3");
  if (level < 1 || level > 3) return fail_invalid("level must be 1, 2 or 3");
  if (kind < 0 || kind > 2) return fail_invalid("kind must be 0 (gravity), 1 (wn) or 2 (smoothdist)");
  const SbhBuffer* c = samples_per_subdomain >= 0 ? contrib_bytes : nullptr;
  if (samples_per_subdomain >= 0 && contrib_bytes == nullptr)
    return fail_invalid("contrib_bytes must be given for the sampler");
  return dim == 2 ? module_eval_level<2>(stream, level, kind, leaf_bytes, node_bytes, c, queries,
                                         out, beta, samples_per_subdomain)
                  : module_eval_level<3>(stream, level, kind, leaf_bytes, node_bytes, c, queries,
                                         out, beta, samples_per_subdomain);
}

// evalBruteForce_{gravity,wn,smoothdist} (src/sbh_gpu.cu:39-77, 213-322).
extern "C" int sbh_tree_bruteforce(void* stream, int32_t dim, int32_t kind,
                                   const SbhBuffer* leaf_bytes, const SbhBuffer* queries,
                                   SbhBuffer* out) {
  clear_error();
  if (dim != 2 && dim != 3) return fail_invalid("dim must be 2 or 3");
  if (kind < 0 || kind > 2) return fail_invalid("kind must be 0 (gravity), 1 (wn) or 2 (smoothdist)");
  const int64_t lsz = dim == 2 ? sizeof(Leaf2) : sizeof(Leaf3);
  if (!buffer_ok(leaf_bytes, SBH_DTYPE_U8, -1) || byte_count(leaf_bytes) % lsz != 0)
    return fail_invalid("leaf_bytes size mismatch");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2) || queries->dims[1] != dim)
    return fail_invalid(dim == 2 ? "queries dim must be 2" : "queries dim must be 3");
  if (queries->dims[0] > INT32_MAX) return fail_invalid("too many queries");
  if (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != queries->dims[0])
    return fail_invalid("out shape mismatch");
  if (dim == 2) {
    if (kind == 0) return launch_brute<2, kPotential>(stream, leaf_bytes, queries, out);
    if (kind == 1) return launch_brute<2, kWn>(stream, leaf_bytes, queries, out);
    return launch_brute<2, kSmooth>(stream, leaf_bytes, queries, out);
  }
  if (kind == 0) return launch_brute<3, kPotential>(stream, leaf_bytes, queries, out);
  if (kind == 1) return launch_brute<3, kWn>(stream, leaf_bytes, queries, out);
  return launch_brute<3, kSmooth>(stream, leaf_bytes, queries, out);
}
