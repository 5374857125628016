// Dense octree8 (drop-in mode): builder and classic walks over the
// reference's byte layout (OctreeNode8, src/sbh_ffi.cu:47-65).
//
// Builder (replaces src/sbh_ffi.cu:1039-1420):
//   codes -> stable radix sort (no atomics: intra-bin order = input order)
//   -> leaf records in sorted order -> bin offsets by binary search
//   -> leaf-level sums in sorted order -> bottom-up 8-child sums
//   -> one kernel per level writing complete node records.
// Walks (replace src/sbh_ffi.cu:1950-2144): one thread per query, depth-first
// with the reference's push order, parameters passed by value.
#include "common.hpp"
#include "o8walk.hpp"
#include "tree.hpp"

namespace sbh {
namespace {

constexpr int kB = 256;

struct BuildGeom {
  float min_x, min_y, min_z;
  float span;
  float cx, cy, cz, half_size;
};

__global__ __launch_bounds__(kB) void k_codes(const float* __restrict__ pts, int64_t n,
                                              BuildGeom g, int bits,
                                              uint32_t* __restrict__ codes,
                                              uint32_t* __restrict__ idx) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (i >= n) return;
  uint32_t qx = quantize_axis(pts[3 * i + 0], g.min_x, g.span, bits);
  uint32_t qy = quantize_axis(pts[3 * i + 1], g.min_y, g.span, bits);
  uint32_t qz = quantize_axis(pts[3 * i + 2], g.min_z, g.span, bits);
  codes[i] = morton3(qx, qy, qz);
  idx[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kB) void k_leaf_gather(const float* __restrict__ pts,
                                                    const float* __restrict__ nrm,
                                                    const float* __restrict__ ms, int64_t n,
                                                    const uint32_t* __restrict__ perm,
                                                    Leaf3* __restrict__ leaf) {
  int64_t r = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (r >= n) return;
  int64_t i = perm[r];
  Leaf3 o;
  o.pos[0] = pts[3 * i + 0];
  o.pos[1] = pts[3 * i + 1];
  o.pos[2] = pts[3 * i + 2];
  o.normal[0] = nrm[3 * i + 0];
  o.normal[1] = nrm[3 * i + 1];
  o.normal[2] = nrm[3 * i + 2];
  o.mass = ms[i];
  o.alpha = 1.0f;
  leaf[r] = o;
}

// offsets[b] = number of sorted codes < b, for b in [0, nbins].
__global__ __launch_bounds__(kB) void k_bin_offsets(const uint32_t* __restrict__ codes,
                                                    int64_t n, int64_t nbins,
                                                    uint32_t* __restrict__ offsets) {
  int64_t b = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (b > nbins) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (static_cast<int64_t>(codes[mid]) < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  offsets[b] = static_cast<uint32_t>(lo);
}

// Level-D sums: count, mass, sum m*x (3), sum m*n (3), in sorted order.
__global__ __launch_bounds__(kB) void k_leaf_sums(const Leaf3* __restrict__ leaf,
                                                  const uint32_t* __restrict__ offsets,
                                                  int64_t nbins, uint32_t* __restrict__ cnt,
                                                  float* __restrict__ sums) {
  int64_t b = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (b >= nbins) return;
  uint32_t s = offsets[b], e = offsets[b + 1];
  float a[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (uint32_t j = s; j < e; ++j) {
    const Leaf3 L = leaf[j];
    float m = L.mass;
    a[0] += m;
    a[1] += m * L.pos[0];
    a[2] += m * L.pos[1];
    a[3] += m * L.pos[2];
    a[4] += m * L.normal[0];
    a[5] += m * L.normal[1];
    a[6] += m * L.normal[2];
  }
  cnt[b] = e - s;
  for (int k = 0; k < 7; ++k) sums[7 * b + k] = a[k];
}

// Parent sums from 8 children (src/sbh_ffi.cu:1745-1786).
__global__ __launch_bounds__(kB) void k_reduce_level(const uint32_t* __restrict__ ccnt,
                                                     const float* __restrict__ csum,
                                                     uint32_t* __restrict__ pcnt,
                                                     float* __restrict__ psum,
                                                     int64_t parents) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (p >= parents) return;
  uint32_t count = 0;
  float a[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int c = 0; c < 8; ++c) {
    int64_t ch = p * 8 + c;
    uint32_t cc = ccnt[ch];
    if (cc == 0) continue;
    count += cc;
    for (int k = 0; k < 7; ++k) a[k] += csum[7 * ch + k];
  }
  pcnt[p] = count;
  for (int k = 0; k < 7; ++k) psum[7 * p + k] = a[k];
}

// Node records for one level (src/sbh_ffi.cu:1788-1876).
__global__ __launch_bounds__(kB) void k_build_nodes(
    Node8* __restrict__ nodes, Leaf3* __restrict__ contrib_out,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ cnt,
    const float* __restrict__ sums, const uint32_t* __restrict__ child_cnt, int level,
    int max_level, uint32_t level_offset, uint32_t parent_level_offset,
    uint32_t child_level_offset, BuildGeom g) {
  int64_t idx64 = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  int64_t level_nodes = int64_t{1} << (3 * level);
  if (idx64 >= level_nodes) return;
  uint32_t idx = static_cast<uint32_t>(idx64);
  uint32_t node_idx = level_offset + idx;
  Node8 node;
  node.depth = static_cast<uint8_t>(level);
  node.parent = (level == 0) ? node_idx : (parent_level_offset + (idx >> 3));
  node.is_leaf = (level == max_level) ? 1 : 0;
  node.child_mask = 0;
  node.pad0 = 0;
  node.first_child = 0xFFFFFFFFu;
  int shift = 3 * (max_level - level);
  uint32_t first_leaf = idx << shift;
  node.start = offsets[first_leaf];
  node.end = node.start + cnt[idx];
  int ix = 0, iy = 0, iz = 0;
  for (int b = 0; b < level; ++b) {
    ix |= static_cast<int>((idx >> (3 * b + 0)) & 1u) << b;
    iy |= static_cast<int>((idx >> (3 * b + 1)) & 1u) << b;
    iz |= static_cast<int>((idx >> (3 * b + 2)) & 1u) << b;
  }
  {
#pragma clang fp contract(off)
    float step = g.half_size * 2.0f / static_cast<float>(1 << level);
    // (c - h) + (i + 0.5) * step with the a*b+c fused, as the reference's
    // compiler contracts it by default.
    node.cx = fmaf(static_cast<float>(ix) + 0.5f, step, g.cx - g.half_size);
    node.cy = fmaf(static_cast<float>(iy) + 0.5f, step, g.cy - g.half_size);
    node.cz = fmaf(static_cast<float>(iz) + 0.5f, step, g.cz - g.half_size);
    node.half_size = step * 0.5f;
    node.diameter = step * 1.7320508075688772f;
  }
  if (level < max_level) {
    uint32_t child_base = child_level_offset + (idx << 3);
    uint8_t mask = 0;
    for (uint32_t c = 0; c < 8; ++c) {
      node.child[c] = child_base + c;
      if (child_cnt[(idx << 3) + c] > 0) {
        mask |= static_cast<uint8_t>(1u << c);
        if (node.first_child == 0xFFFFFFFFu) node.first_child = child_base + c;
      }
    }
    node.child_mask = mask;
    if (mask == 0) node.is_leaf = 1;
  } else {
    for (int c = 0; c < 8; ++c) node.child[c] = 0xFFFFFFFFu;
  }
  const float* s = sums + 7 * static_cast<int64_t>(idx);
  float mass = s[0];
  node.contrib.alpha = 1.0f;
  node.contrib.mass = mass;
  if (mass > 0.0f) {
    node.contrib.pos[0] = s[1] / mass;
    node.contrib.pos[1] = s[2] / mass;
    node.contrib.pos[2] = s[3] / mass;
    node.contrib.normal[0] = s[4] / mass;
    node.contrib.normal[1] = s[5] / mass;
    node.contrib.normal[2] = s[6] / mass;
  } else {
    for (int a = 0; a < 3; ++a) node.contrib.pos[a] = node.contrib.normal[a] = 0.0f;
  }
  nodes[node_idx] = node;
  contrib_out[node_idx] = node.contrib;
}

// ------------------------------------------------------------------ walks
using o8::WalkParams;
using o8::use_far_field;
using o8::should_prune;
using o8::contribution;
using o8::kGravGrad;
using o8::kSoftForce;
using o8::kSoftVjp;
using o8::kWinding;


// Node record read once per wave through the scalar data cache (the node
// buffer is read-only during a walk; records are 4-byte aligned).
typedef __attribute__((address_space(4))) const uint32_t cu32;

__device__ __forceinline__ Node8 sload_node8(const Node8* nodes, uint32_t id) {
  static_assert(sizeof(Node8) == 26 * 4, "Node8 is 26 dwords");
  cu32* src = (cu32*)(nodes) + 26 * static_cast<uint64_t>(id);
  Node8 n;
  uint32_t* d = reinterpret_cast<uint32_t*>(&n);
#pragma unroll
  for (int k = 0; k < 26; ++k) d[k] = src[k];
  return n;
}

__device__ __forceinline__ Leaf3 sload_leaf3(const Leaf3* leaf, uint32_t j) {
  cu32* src = (cu32*)(leaf) + 8 * static_cast<uint64_t>(j);
  Leaf3 r;
  uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (int k = 0; k < 8; ++k) d[k] = src[k];
  return r;
}

constexpr int kWStack = 96;  // >= 7 * 10 + 1 entries for max_depth 10

// Wave-coherent form of the reference's per-query DFS
// (src/sbh_ffi.cu:1950-2144): the 64 queries of a wave share one stack of
// (node, lanes) entries.  A popped node is evaluated by the lanes listed with
// it; each of them skips it (empty, pruned), uses it (far field; a leaf is
// summed exactly) or opens it, and the children are pushed, in the
// reference's order, with the set of lanes that opened it.  Each lane thus
// visits exactly the nodes of its own DFS and in the same order (a DFS
// restricted to an ancestor-closed subset keeps its order), so its result is
// bit-identical to the per-query walk, while every node and point record is
// read once per wave through the scalar cache.  Queries run in Morton order
// (order[t]) so that the lanes of a wave share most of their paths.
// PER: periodicity as a template argument (the non-periodic walk carries no
// minimum-image arithmetic; same bits either way).
template <int KIND, bool PER>
__global__ __launch_bounds__(kB) void k_walk8(const Leaf3* __restrict__ leaf,
                                              const Node8* __restrict__ nodes,
                                              const float* __restrict__ queries,
                                              const float* __restrict__ cot, int64_t nq,
                                              const uint32_t* __restrict__ order,
                                              float* __restrict__ out, WalkParams P,
                                              uint32_t num_nodes, uint32_t num_leaves) {
  P.periodic = PER ? 1 : 0;
  __shared__ uint32_t s_node[kB / 64][kWStack];
  __shared__ uint64_t s_lanes[kB / 64][kWStack];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  const bool valid = t < nq;
  const int64_t q = !valid ? 0 : (order != nullptr ? static_cast<int64_t>(order[t]) : t);
  float qx = 0.f, qy = 0.f, qz = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  if (valid) {
    qx = queries[3 * q + 0];
    qy = queries[3 * q + 1];
    qz = queries[3 * q + 2];
    if constexpr (KIND == kSoftVjp) {
      gx = cot[3 * q + 0];
      gy = cot[3 * q + 1];
      gz = cot[3 * q + 2];
    }
  }
  const uint64_t all = __ballot(valid);
  if (all == 0) return;  // the whole wave is past the end
  float ax = 0.f, ay = 0.f, az = 0.f;
  uint32_t* stk_node = s_node[w];
  uint64_t* stk_lanes = s_lanes[w];
  if (lane == 0) {
    stk_node[0] = 0;
    stk_lanes[0] = all;
  }
  int top = 1;
  __builtin_amdgcn_wave_barrier();
  while (top > 0) {
    --top;
    const uint32_t id = __builtin_amdgcn_readfirstlane(stk_node[top]);
    const uint64_t lanes = stk_lanes[top];
    __builtin_amdgcn_wave_barrier();
    if (id >= num_nodes) continue;
    const Node8 nd = sload_node8(nodes, id);
    if (nd.contrib.mass == 0.0f) continue;
    const bool mine = (lanes >> lane) & 1ull;
    bool skip = !mine, far = false;
    if (mine) {
      if constexpr (o8::kPrunedKind<KIND>) skip = should_prune(&nd, qx, qy, qz, P);
      if (!skip) far = use_far_field(&nd, qx, qy, qz, P);
    }
    if (nd.is_leaf) {
      // a leaf always takes the far branch: the exact sum of its points
      if (far) {
        const uint32_t je = nd.end < num_leaves ? nd.end : num_leaves;
        for (uint32_t j = nd.start; j < je; ++j)
          contribution<KIND>(sload_leaf3(leaf, j), qx, qy, qz, gx, gy, gz, P, ax, ay, az);
      }
      continue;
    }
    if (far) contribution<KIND>(nd.contrib, qx, qy, qz, gx, gy, gz, P, ax, ay, az);
    const uint64_t open = __ballot(!skip && !far);
    if (open == 0) continue;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if ((nd.child_mask & (1u << c)) && top < kWStack && nd.child[c] < num_nodes) {
        if (lane == 0) {
          stk_node[top] = nd.child[c];
          stk_lanes[top] = open;
        }
        ++top;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (!valid) return;
  if constexpr (o8::kScalarKind<KIND>) {
    out[q] = ax;
  } else {
    out[3 * q + 0] = ax;
    out[3 * q + 1] = ay;
    out[3 * q + 2] = az;
  }
}

uint64_t total_nodes8(int max_depth) {
  uint64_t total = 0, level = 1;
  for (int l = 0; l <= max_depth; ++l) {
    total += level;
    level <<= 3;
  }
  return total;
}

using o8::check_walk_common;
using o8::make_walk_params;

template <int KIND>
int launch_walk(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                const SbhBuffer* queries, const SbhBuffer* cot, SbhBuffer* out,
                const WalkParams& P) {
  int64_t nq = queries->dims[0];
  if (nq == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t* order = nullptr;
  // (not while the stream is captured: the ordering scratch may need an
  // allocation, and the unordered walk gives the same results)
  if (nq >= o8::kOrderMin && !stream_capturing(s)) {
    char* ws = static_cast<char*>(workspace(s, kWsWalk, tree::query_order_bytes(nq)));
    if (ws == nullptr) return fail_internal_msg("workspace allocation failed");
    int rc = tree::query_order(s, static_cast<const float*>(queries->data), nq, ws, &order);
    if (rc != SBH_OK) return rc;
  }
  (P.periodic ? k_walk8<KIND, true> : k_walk8<KIND, false>)<<<grid_for(nq, kB), kB, 0, s>>>(
      static_cast<const Leaf3*>(leaf_bytes->data), static_cast<const Node8*>(node_bytes->data),
      static_cast<const float*>(queries->data),
      cot ? static_cast<const float*>(cot->data) : nullptr, nq, order,
      static_cast<float*>(out->data),
      P, static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node8)),
      static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf3)));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" int sbh_build_octree8_buffers(void* stream, const SbhBuffer* points,
                                         const SbhBuffer* normals, const SbhBuffer* masses,
                                         const SbhBuffer* min_corner,
                                         const SbhBuffer* max_corner, SbhBuffer* leaf_out,
                                         SbhBuffer* node_out, SbhBuffer* contrib_out,
                                         int32_t max_depth) {
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
  if (max_depth < 1 || max_depth > 10) return fail_invalid("max_depth must be in [1, 10]");
  int64_t n = points->dims[0];
  uint64_t total = total_nodes8(max_depth);
  int64_t nbins = int64_t{1} << (3 * max_depth);
  if (!buffer_ok(leaf_out, SBH_DTYPE_U8, -1) ||
      byte_count(leaf_out) != n * static_cast<int64_t>(sizeof(Leaf3)))
    return fail_invalid("leaf_out size mismatch");
  if (!buffer_ok(node_out, SBH_DTYPE_U8, -1) ||
      byte_count(node_out) != static_cast<int64_t>(total * sizeof(Node8)))
    return fail_invalid("node_out size mismatch");
  if (!buffer_ok(contrib_out, SBH_DTYPE_U8, -1) ||
      byte_count(contrib_out) != static_cast<int64_t>(total * sizeof(Leaf3)))
    return fail_invalid("contrib_out size mismatch");
  if (n > INT32_MAX) return fail_invalid("too many points");

  hipStream_t s = static_cast<hipStream_t>(stream);
  // Bounds are read back like the reference does (src/sbh_ffi.cu:1111-1126)
  // so that an empty span is reported as an argument error.
  float h[6];
  SBH_HIP_TRY(hipMemcpyAsync(h, min_corner->data, 3 * sizeof(float), hipMemcpyDeviceToHost, s));
  SBH_HIP_TRY(hipMemcpyAsync(h + 3, max_corner->data, 3 * sizeof(float), hipMemcpyDeviceToHost, s));
  SBH_HIP_TRY(hipStreamSynchronize(s));
  BuildGeom g;
  g.min_x = h[0];
  g.min_y = h[1];
  g.min_z = h[2];
  g.span = fmaxf(fmaxf(h[3] - h[0], h[4] - h[1]), h[5] - h[2]);
  if (!(g.span > 0.0f)) return fail_invalid("invalid bounds span");
  g.cx = h[0] + g.span * 0.5f;
  g.cy = h[1] + g.span * 0.5f;
  g.cz = h[2] + g.span * 0.5f;
  g.half_size = g.span * 0.5f;

  // Workspace: keys/vals x2, sort temp, offsets, per-level counts and sums.
  size_t nn = static_cast<size_t>(n > 0 ? n : 1);
  size_t off_codes = 0;
  size_t off_idx = off_codes + nn * 4;
  size_t off_codes2 = off_idx + nn * 4;
  size_t off_perm = off_codes2 + nn * 4;
  size_t off_sort = off_perm + nn * 4;
  size_t sort_bytes = (radix_sort_temp_bytes(n) + 255) & ~size_t{255};
  size_t off_offsets = off_sort + sort_bytes;
  size_t off_levels = off_offsets + ((static_cast<size_t>(nbins) + 1) * 4 + 255) / 256 * 256;
  size_t level_bytes = 0;
  for (int l = 0; l <= max_depth; ++l) level_bytes += (size_t{1} << (3 * l)) * 32;
  size_t ws_bytes = off_levels + level_bytes;
  char* ws = nullptr;
  if (int rcw = cached_workspace(s, kWsBuild, ws_bytes, &ws); rcw != SBH_OK) return rcw;
  auto* codes = reinterpret_cast<uint32_t*>(ws + off_codes);
  auto* idx = reinterpret_cast<uint32_t*>(ws + off_idx);
  auto* codes_sorted = reinterpret_cast<uint32_t*>(ws + off_codes2);
  auto* perm = reinterpret_cast<uint32_t*>(ws + off_perm);
  auto* offsets = reinterpret_cast<uint32_t*>(ws + off_offsets);
  uint32_t* lcnt[11];
  float* lsum[11];
  {
    size_t o = off_levels;
    for (int l = 0; l <= max_depth; ++l) {
      size_t c = size_t{1} << (3 * l);
      lcnt[l] = reinterpret_cast<uint32_t*>(ws + o);
      lsum[l] = reinterpret_cast<float*>(ws + o + c * 4);
      o += c * 32;
    }
  }
  auto* pp = static_cast<const float*>(points->data);
  auto* leaf = static_cast<Leaf3*>(leaf_out->data);
  if (n > 0) {
    k_codes<<<grid_for(n, kB), kB, 0, s>>>(pp, n, g, max_depth, codes, idx);
    SBH_LAUNCH_CHECK();
    int rc = radix_sort_pairs(s, codes, idx, codes_sorted, perm, n, 3 * max_depth, ws + off_sort);
    if (rc != SBH_OK) return rc;
    k_leaf_gather<<<grid_for(n, kB), kB, 0, s>>>(pp, static_cast<const float*>(normals->data),
                                                 static_cast<const float*>(masses->data), n,
                                                 perm, leaf);
    SBH_LAUNCH_CHECK();
  }
  k_bin_offsets<<<grid_for(nbins + 1, kB), kB, 0, s>>>(codes_sorted, n, nbins, offsets);
  SBH_LAUNCH_CHECK();
  k_leaf_sums<<<grid_for(nbins, kB), kB, 0, s>>>(leaf, offsets, nbins, lcnt[max_depth],
                                                 lsum[max_depth]);
  SBH_LAUNCH_CHECK();
  for (int l = max_depth - 1; l >= 0; --l) {
    int64_t parents = int64_t{1} << (3 * l);
    k_reduce_level<<<grid_for(parents, kB), kB, 0, s>>>(lcnt[l + 1], lsum[l + 1], lcnt[l],
                                                        lsum[l], parents);
    SBH_LAUNCH_CHECK();
  }
  uint32_t level_offset = 0, parent_offset = 0;
  auto* nodes = static_cast<Node8*>(node_out->data);
  auto* contrib = static_cast<Leaf3*>(contrib_out->data);
  for (int l = 0; l <= max_depth; ++l) {
    uint32_t level_nodes = 1u << (3 * l);
    uint32_t child_offset = level_offset + level_nodes;
    k_build_nodes<<<grid_for(level_nodes, kB), kB, 0, s>>>(
        nodes, contrib, offsets, lcnt[l], lsum[l], l < max_depth ? lcnt[l + 1] : nullptr, l,
        max_depth, level_offset, parent_offset, l < max_depth ? child_offset : 0u, g);
    SBH_LAUNCH_CHECK();
    parent_offset = level_offset;
    level_offset = child_offset;
  }
  return SBH_OK;
}

extern "C" int sbh_barnes_hut_force_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                            const SbhBuffer* node_bytes,
                                            const SbhBuffer* queries, SbhBuffer* out,
                                            float beta, int32_t periodic, float box_length,
                                            int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK) return rc;
  WalkParams P = make_walk_params(beta, 0.0f, 0.0f, 0, 0.0f, periodic, box_length);
  return launch_walk<kGravGrad>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P);
}

// Winding numbers over octree8 records built from oriented area elements
// (points = element centroids, normals = unit normals, masses = areas): the
// classic vote walk of src/sbh_gpu.cu:100-120 with signedSolidAngle<3>,
// include/kernel.h:73-80.  out: (M,).
extern "C" int sbh_barnes_hut_wn_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                         const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                         SbhBuffer* out, float beta, int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  WalkParams P = make_walk_params(beta, 0.0f, 0.0f, 0, 0.0f, 0, 1.0f);
  return launch_walk<o8::kWinding>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P);
}

// Potential -m / (r + 0.1) (gravityPotential, include/kernel.h:55-60) over
// octree8 records: the classic walk of src/sbh_gpu.cu:79-98, 324-361
// (evalBarnesHut_gravity).  out: (M,).
extern "C" int sbh_barnes_hut_potential_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                                const SbhBuffer* node_bytes,
                                                const SbhBuffer* queries, SbhBuffer* out,
                                                float beta, int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  WalkParams P = make_walk_params(beta, 0.0f, 0.0f, 0, 0.0f, 0, 1.0f);
  return launch_walk<o8::kPotential>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P);
}

// Smooth distance field sum exp(-alpha r) (smoothDistExp, include/kernel.h:
// 82-87) over octree8 records (alpha = 1 in every record, as the builder of
// src/sbh_ffi.cu:1741, 1860 writes): src/sbh_gpu.cu:122-141, 400-434
// (evalBarnesHut_smoothdist).  out: (M,).
extern "C" int sbh_barnes_hut_smoothdist_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                                 const SbhBuffer* node_bytes,
                                                 const SbhBuffer* queries, SbhBuffer* out,
                                                 float beta, int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  WalkParams P = make_walk_params(beta, 0.0f, 0.0f, 0, 0.0f, 0, 1.0f);
  return launch_walk<o8::kSmoothDist>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P);
}

extern "C" int sbh_barnes_hut_force_octree8_soft(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, SbhBuffer* out, float beta, float softening_scale,
    float cutoff_scale, int32_t prune_enabled, float prune_r_cut_mult, int32_t periodic,
    float box_length, int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK) return rc;
  if (prune_enabled != 0 && prune_r_cut_mult <= 0.0f)
    return fail_invalid("prune_r_cut_mult must be > 0 when prune_enabled is set");
  WalkParams P = make_walk_params(beta, softening_scale, cutoff_scale, prune_enabled,
                                  prune_r_cut_mult, periodic, box_length);
  return launch_walk<kSoftForce>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P);
}

extern "C" int sbh_barnes_hut_force_octree8_soft_vjp(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, const SbhBuffer* cotangent, SbhBuffer* out, float beta,
    float softening_scale, float cutoff_scale, int32_t prune_enabled, float prune_r_cut_mult,
    int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  int rc = check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK) return rc;
  if (!buffer_ok(cotangent, SBH_DTYPE_F32, 2) || cotangent->dims[0] != queries->dims[0] ||
      cotangent->dims[1] != 3)
    return fail_invalid("cotangent shape mismatch");
  if (prune_enabled != 0 && prune_r_cut_mult <= 0.0f)
    return fail_invalid("prune_r_cut_mult must be > 0 when prune_enabled is set");
  WalkParams P = make_walk_params(beta, softening_scale, cutoff_scale, prune_enabled,
                                  prune_r_cut_mult, periodic, box_length);
  return launch_walk<kSoftVjp>(stream, leaf_bytes, node_bytes, queries, cotangent, out, P);
}
