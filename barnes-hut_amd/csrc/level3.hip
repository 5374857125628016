// Level-3 targets: the single-split 8 x 8 x 8 tree in the Tree3DNode<1, 3>
// layout (src/sbh_ffi.cu:181-397, 736-1037) and the walks over it
// (src/sbh_ffi.cu:1441-1680 via include/barnes_hut_gpu.h:19-72 and
// include/tree_sampler_gpu.h:25-140).
//
// Build: per-axis cell index of every point (quantize_axis_level3), a stable
// radix sort on the 9-bit cell id (the reference scatters with an atomic
// cursor, so its order inside a cell is a race; here it is (cell, index)),
// cell offsets by binary search, per-cell sums by one block per cell in a
// fixed order, and one 512-thread block writing the root and the child
// records (ranks of present cells by an LDS scan, the root's bit mask by one
// wave ballot per 64 cells).  No float atomics: the build is deterministic.
//
// Classic walk: the reference's warp-voted traversal: a node is used as a far
// field only if every active query of a 32-query warp accepts it, so all
// queries of that warp take the same path.  Launched with the caller's block
// size, every reference warp is one half of a 64-wide wavefront here, and the
// vote is a ballot restricted to that half.  Leaves always accept, so for
// level-3 buffers a warp either uses the root monopole (all its queries far
// from the whole set) or sums every point exactly.
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
enum L3Kind { kPotential = 0, kForce = 1, kSoftForce3 = 2 };

struct L3Params {
  o8::WalkParams P;
  float beta;
  int threads;  // the caller's block size: reference warps are 32 queries of a block
};

// Contribution of one record to a query (gravityPotential, gravityGrad,
// gravityGradSoftCutoff<3>: include/kernel.h:55-60, 89-94; src/sbh_ffi.cu:142-160).
template <int KIND>
__device__ __forceinline__ void l3_term(const Leaf3& L, float qx, float qy, float qz,
                                        const o8::WalkParams& P, float* acc) {
  if constexpr (KIND == kPotential) {
    const float dx = min_image_floor(L.pos[0] - qx, P.periodic, P.box, P.inv_box);
    const float dy = min_image_floor(L.pos[1] - qy, P.periodic, P.box, P.inv_box);
    const float dz = min_image_floor(L.pos[2] - qz, P.periodic, P.box, P.inv_box);
    acc[0] += -L.mass / (sqrtf(dx * dx + dy * dy + dz * dz) + 1e-1f);
  } else if constexpr (KIND == kForce) {
    o8::contribution<o8::kGravGrad>(L, qx, qy, qz, 0.f, 0.f, 0.f, P, acc[0], acc[1], acc[2]);
  } else {
    o8::contribution<o8::kSoftForce>(L, qx, qy, qz, 0.f, 0.f, 0.f, P, acc[0], acc[1], acc[2]);
  }
}

// farFieldRatio (include/tree3d.h:134-141): |q - centre of mass| / diameter,
// infinite for a leaf.
__device__ __forceinline__ float far_ratio(const Node3* nd, float qx, float qy, float qz) {
#pragma clang fp contract(off)
  if (nd->is_leaf) return INFINITY;
  const float dx = qx - nd->contrib.pos[0];
  const float dy = qy - nd->contrib.pos[1];
  const float dz = qz - nd->contrib.pos[2];
  return sqrtf(dx * dx + dy * dy + dz * dz) / nd->diameter;
}

// Cell of point p in `parent` (AABB::quantize, include/aabb.h:36-42), or -1
// when p is outside parent's bounds widened by 1e-4 (AABB::contains).
__device__ __forceinline__ int cell_in(const Node3* parent, const float* p) {
#pragma clang fp contract(off)
  int q[3];
  for (int a = 0; a < 3; ++a) {
    if (!(parent->bmin[a] - 1e-4f <= p[a] && p[a] <= parent->bmax[a] + 1e-4f)) return -1;
    const float t =
        static_cast<float>(kSplit) * (p[a] - parent->bmin[a]) / (parent->bmax[a] - parent->bmin[a]);
    q[a] = min(max(static_cast<int>(t), 0), kSplit - 1);
  }
  return q[0] + kSplit * (q[1] + kSplit * q[2]);
}

// First set bit of the node's mask at or after `from` (kCells if none).
__device__ __forceinline__ int next_bit(const Node3* nd, int from) {
  for (int w = from >> 6; w < kCells / 64; ++w) {
    uint64_t m = nd->bitmask[w];
    if (w == (from >> 6)) m &= ~0ull << (from & 63);
    if (m) return 64 * w + __builtin_ctzll(m);
  }
  return kCells;
}

// Exact sum over the point records [j0, j1) in index order; records are
// fetched 8 at a time so that their loads overlap.
template <int KIND, class Params>
__device__ __forceinline__ void l3_leaf_sum(const Leaf3* __restrict__ leaf, uint64_t j0,
                                            uint64_t j1, float qx, float qy, float qz,
                                            const Params& P, float* acc) {
  uint64_t j = j0;
  for (; j + 8 <= j1; j += 8) {
    Leaf3 r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = leaf[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) l3_term<KIND>(r[u], qx, qy, qz, P, acc);
  }
  for (; j < j1; ++j) l3_term<KIND>(leaf[j], qx, qy, qz, P, acc);
}

template <int KIND>
__global__ void k_l3_walk(const Leaf3* __restrict__ leaf, const Node3* __restrict__ nodes,
                          const float* __restrict__ queries, int nq, float* __restrict__ out,
                          L3Params A, uint32_t num_nodes, uint32_t num_leaves) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = q < nq;
  const int lane = __lane_id();
  // the reference warp of this query: lanes [32 h, 32 h + 32) of the wave
  const uint64_t half = (lane < 32) ? 0x00000000ffffffffull : 0xffffffff00000000ull;
  const uint64_t members = __ballot(active) & half;
  if (!active) return;
  const float qx = queries[3 * static_cast<int64_t>(q)];
  const float qy = queries[3 * static_cast<int64_t>(q) + 1];
  const float qz = queries[3 * static_cast<int64_t>(q) + 2];
  float acc[3] = {0.f, 0.f, 0.f};
  // stackless traversal (include/barnes_hut_gpu.h:19-72)
  uint32_t node = 0;
  bool from_child = false;
  for (int guard = 0; guard < 4 * kMaxNodes3 * 64; ++guard) {
    const Node3* nd = &nodes[node];
    if (!from_child) {
      const bool far = far_ratio(nd, qx, qy, qz) >= A.beta;
      const bool all = (__ballot(far) & members) == members;
      if (all) {
        if (nd->is_leaf) {
          const uint64_t je = nd->end < num_leaves ? nd->end : num_leaves;
          l3_leaf_sum<KIND>(leaf, nd->start, je, qx, qy, qz, A.P, acc);
        } else {
          l3_term<KIND>(nd->contrib, qx, qy, qz, A.P, acc);
        }
      } else {
        const int b = next_bit(nd, 0);
        if (b < kCells && nd->child[b] < num_nodes) {
          node = static_cast<uint32_t>(nd->child[b]);
          continue;
        }
      }
    }
    if (node == 0) break;
    // next sibling: the parent's next present cell after this node's cell
    const Node3* parent = &nodes[nd->parent_idx < num_nodes ? nd->parent_idx : 0];
    float mid[3];
    for (int a = 0; a < 3; ++a) mid[a] = (nd->bmin[a] + nd->bmax[a]) / 2.0f;
    const int cell = cell_in(parent, mid);
    const int b = cell < 0 ? kCells : next_bit(parent, cell + 1);
    if (b >= kCells || parent->child[b] >= num_nodes) {
      node = static_cast<uint32_t>(nd->parent_idx < num_nodes ? nd->parent_idx : 0);
      from_child = true;
    } else {
      node = static_cast<uint32_t>(parent->child[b]);
      from_child = false;
    }
  }
  if constexpr (KIND == kPotential) {
    out[q] = acc[0];
  } else {
    out[3 * static_cast<int64_t>(q) + 0] = acc[0];
    out[3 * static_cast<int64_t>(q) + 1] = acc[1];
    out[3 * static_cast<int64_t>(q) + 2] = acc[2];
  }
}

// Stochastic walk over level-3 buffers (include/tree_sampler_gpu.h:25-140):
// the subdomains are the root's children, which are all leaves in buffers
// from the level-3 builder, so every subdomain is summed exactly and no
// sample path has a step; the estimate is divided by samples_per_subdomain
// as in the reference (0 samples give 0 / 0).  A non-leaf subdomain
// contributes its monopole only.
template <int KIND>
__global__ __launch_bounds__(kB) void k_l3_stoch(const Leaf3* __restrict__ leaf,
                                                 const Node3* __restrict__ nodes,
                                                 const Leaf3* __restrict__ contrib,
                                                 const float* __restrict__ queries, int nq,
                                                 float* __restrict__ out, o8::WalkParams P,
                                                 int samples, uint32_t num_nodes,
                                                 uint32_t num_leaves, uint32_t num_contrib) {
  const int q = blockIdx.x * kB + threadIdx.x;
  if (q >= nq) return;
  const float qx = queries[3 * static_cast<int64_t>(q)];
  const float qy = queries[3 * static_cast<int64_t>(q) + 1];
  const float qz = queries[3 * static_cast<int64_t>(q) + 2];
  float acc[3] = {0.f, 0.f, 0.f};
  float mlcv[3] = {0.f, 0.f, 0.f};
  const Node3* root = &nodes[0];
  for (int b = next_bit(root, 0); b < kCells; b = next_bit(root, b + 1)) {
    const uint64_t ci = root->child[b];
    if (ci >= num_nodes) continue;
    const Node3* sub = &nodes[ci];
    if (sub->is_leaf) {
      const uint64_t je = sub->end < num_leaves ? sub->end : num_leaves;
      l3_leaf_sum<KIND>(leaf, sub->start, je, qx, qy, qz, P, acc);
    } else {
      l3_term<KIND>(ci < num_contrib ? contrib[ci] : sub->contrib, qx, qy, qz, P, acc);
    }
    const float inv_s = 1.0f / static_cast<float>(samples);
    for (int a = 0; a < 3; ++a) mlcv[a] += 0.0f * inv_s;  // 0 / samples, as the reference
  }
  if constexpr (KIND == kPotential) {
    out[q] = acc[0] + mlcv[0];
  } else {
    for (int a = 0; a < 3; ++a) out[3 * static_cast<int64_t>(q) + a] = acc[a] + mlcv[a];
  }
}

int check_l3_walk(const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                  const SbhBuffer* contrib_bytes, const SbhBuffer* queries,
                  const SbhBuffer* out, int threads, bool vec_out) {
  if (!buffer_ok(leaf_bytes, SBH_DTYPE_U8, -1) || !buffer_ok(node_bytes, SBH_DTYPE_U8, -1))
    return fail_invalid("leaf_bytes/node_bytes must be uint8 device buffers");
  if (byte_count(leaf_bytes) % static_cast<int64_t>(sizeof(Leaf3)) != 0)
    return fail_invalid("leaf_bytes size mismatch");
  if (byte_count(node_bytes) % static_cast<int64_t>(sizeof(Node3)) != 0 ||
      byte_count(node_bytes) == 0)
    return fail_invalid("node_bytes size mismatch");
  if (contrib_bytes != nullptr &&
      (!buffer_ok(contrib_bytes, SBH_DTYPE_U8, -1) ||
       byte_count(contrib_bytes) % static_cast<int64_t>(sizeof(Leaf3)) != 0))
    return fail_invalid("contrib_bytes size mismatch");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2)) return fail_invalid("queries must be (M, 3) float32");
  if (queries->dims[1] != 3) return fail_invalid("queries dim must be 3");
  if (queries->dims[0] > INT32_MAX) return fail_invalid("too many queries");
  if (vec_out ? (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != queries->dims[0] ||
                 out->dims[1] != 3)
              : (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != queries->dims[0]))
    return fail_invalid("out shape mismatch");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  return SBH_OK;
}

template <int KIND>
int launch_l3_walk(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                   const SbhBuffer* queries, SbhBuffer* out, float beta, float softening,
                   float cutoff, int periodic, float box, int threads) {
  int rc = check_l3_walk(leaf_bytes, node_bytes, nullptr, queries, out, threads,
                         KIND != kPotential);
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
  k_l3_walk<KIND><<<blocks, threads, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const Leaf3*>(leaf_bytes->data), static_cast<const Node3*>(node_bytes->data),
      static_cast<const float*>(queries->data), nq, static_cast<float*>(out->data), A,
      static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node3)),
      static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf3)));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

template <int KIND>
int launch_l3_stoch(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                    const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
                    int samples, float softening, float cutoff, int periodic, float box,
                    int threads) {
  int rc = check_l3_walk(leaf_bytes, node_bytes, contrib_bytes, queries, out, threads,
                         KIND != kPotential);
  if (rc != SBH_OK) return rc;
  const int nq = static_cast<int>(queries->dims[0]);
  if (nq == 0) return SBH_OK;
  o8::WalkParams P = o8::make_walk_params(1.0f, softening, cutoff, 0, 0.f, periodic, box);
  P.cutoff = cutoff > 0.0f ? 1 : 0;
  k_l3_stoch<KIND><<<grid_for(nq, kB), kB, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const Leaf3*>(leaf_bytes->data), static_cast<const Node3*>(node_bytes->data),
      static_cast<const Leaf3*>(contrib_bytes->data), static_cast<const float*>(queries->data),
      nq, static_cast<float*>(out->data), P, samples,
      static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node3)),
      static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf3)),
      static_cast<uint32_t>(byte_count(contrib_bytes) / sizeof(Leaf3)));
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
  char* ws = static_cast<char*>(workspace(s, kWsBuild, o));
  if (ws == nullptr) return fail_internal_msg("workspace allocation failed");
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
  return launch_l3_walk<kPotential>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f,
                                    0.f, periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                                       const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                       SbhBuffer* out, float beta, int32_t periodic,
                                       float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<kForce>(stream, leaf_bytes, node_bytes, queries, out, beta, 0.f, 0.f,
                                periodic, box_length, threads);
}

extern "C" int sbh_barnes_hut_force_soft_3d(void* stream, const SbhBuffer* leaf_bytes,
                                            const SbhBuffer* node_bytes,
                                            const SbhBuffer* queries, SbhBuffer* out,
                                            float beta, float softening_scale,
                                            float cutoff_scale, int32_t periodic,
                                            float box_length, int32_t threads) {
  clear_error();
  return launch_l3_walk<kSoftForce3>(stream, leaf_bytes, node_bytes, queries, out, beta,
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
  (void)seed;
  return launch_l3_stoch<kPotential>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                     out, samples_per_subdomain, 0.f, 0.f, periodic, box_length,
                                     threads);
}

extern "C" int sbh_stochastic_bh_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                                          const SbhBuffer* node_bytes,
                                          const SbhBuffer* contrib_bytes,
                                          const SbhBuffer* queries, SbhBuffer* out,
                                          int32_t samples_per_subdomain, int32_t seed,
                                          int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  (void)seed;
  return launch_l3_stoch<kForce>(stream, leaf_bytes, node_bytes, contrib_bytes, queries, out,
                                 samples_per_subdomain, 0.f, 0.f, periodic, box_length,
                                 threads);
}

extern "C" int sbh_stochastic_bh_force_soft_3d(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
    int32_t samples_per_subdomain, int32_t seed, float softening_scale, float cutoff_scale,
    int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  (void)seed;
  return launch_l3_stoch<kSoftForce3>(stream, leaf_bytes, node_bytes, contrib_bytes, queries,
                                      out, samples_per_subdomain, softening_scale, cutoff_scale,
                                      periodic, box_length, threads);
}
