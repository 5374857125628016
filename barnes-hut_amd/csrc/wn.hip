// Winding numbers of an oriented-area (dipole) cloud on the adaptive octree.
//
// Sources are area elements (centroid p, unit normal n, area a); the value at
// q is sum a n.(p - q) / (4 pi (|p - q|^3 + 1e-5)) (signedSolidAngle<3>,
// include/kernel.h:73-80).  The tree is built over the centroids with the
// areas as weights (tree::build: area-weighted centroid c and area A per
// node), then per node the dipole D = sum a n, so an accepted node counts as
// the dipole D at c -- the reference's node record (A, D / A) at c
// (src/sbh_gpu.cu:100-120 with src/sbh_ffi.cu:1840-1868).  A first-moment
// (sum a n (p - c)^T) correction was tried and rejected: for smooth surface
// patches at the MAC distance the next Taylor term is larger and of opposite
// sign, so the truncated series is less accurate than the dipole alone.
//
// Queries are a separate cloud with a tree of their own (topology only, in
// their own cube); its cell groups of <= 64 queries (one wavefront each) are
// walked against the element tree with the group MAC and traversal of the
// gravity walk.  (Plain runs of 64 Morton-sorted queries are not used: runs
// that cross a coarse cell boundary have boxes spanning much of the domain
// and accept almost nothing.)
#include "common.hpp"
#include "profile.hpp"
#include "groupwalk.hpp"
#include "tree.hpp"

namespace sbh {
namespace tree {
namespace {

using namespace gw;

constexpr float kInvFourPi = 0.07957747154594767f;
constexpr int kQueryDepth = kQueryTreeDepth;
constexpr int kQueryLeaf = kQueryTreeLeaf;

struct WnLayout {
  Layout t;      // element tree, at the workspace base
  size_t wd;     // per element-tree node: dipole (float4)
  size_t qbase;  // query tree, at workspace + qbase
  Layout qt;
  size_t total;
};

WnLayout wn_layout(int64_t n, int64_t m, int max_depth, int leaf_size) {
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  WnLayout W;
  W.t = layout(n, max_depth, leaf_size);
  const size_t cap = static_cast<size_t>(W.t.cap) + 1;
  size_t o = W.t.total;
  W.wd = o; o = align(o + 16 * cap);
  W.qbase = o;
  W.qt = layout(m, kQueryDepth, kQueryLeaf);
  W.total = W.qbase + W.qt.total;
  return W;
}

// The dipoles are computed by the tree build's moment pass (tree::build with
// a dipole array), which also clears the sentinel node's entry.
// ------------------------------------------------------------------ walk
struct WnArgs {
  const Header* hdr;  // element tree
  Header* qhdr;       // query tree: group count, work queue, statistics
  const int2* groups; // query groups (start, count) in sorted query order
  const int4* topo;
  const float4* geo;
  const float4* mp;   // centroid, area
  const float4* wd;   // dipole
  const float4* sp;   // sorted centroid, area
  const float4* sn;   // sorted normal
  const float4* qs;   // sorted queries
  const uint32_t* qperm;
  int m;
  int sentinel;
  float beta;
  float* out;
  bool stats;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// C far nodes (sentinel-padded) against the lane's query; like the near batch
// below it returns 4 pi times the contribution (the caller scales once).
template <int C>
__device__ __forceinline__ float wn_far_batch(const WnArgs& A, float qx, float qy, float qz,
                                              const int* far) {
  float w = 0.f;
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const int k = __builtin_amdgcn_readfirstlane(far[t]);
    const float4 c = sload4(A.mp, k);
    const float4 d = sload4(A.wd, k);
    const float Xx = c.x - qx, Xy = c.y - qy, Xz = c.z - qz;
    const float R2 = Xx * Xx + Xy * Xy + Xz * Xz;
    const float R = sqrt_hw(R2);
    w += (d.x * Xx + d.y * Xy + d.z * Xz) * rcp(fmaf(R, R2, 1e-5f));
  }
  return w;
}

// kChunk consecutive elements of one leaf from j0 (with PAD, entries at or
// beyond s1 carry zero area; without, the batch lies inside the range and
// reads its records as runs, which the scalar loads fetch 4 at a time).
template <bool PAD>
__device__ __forceinline__ float wn_near_batch(const WnArgs& A, float qx, float qy, float qz,
                                               int j0, int s1) {
  float w = 0.f;
  cfloat* rp = run_base(A.sp, j0);
  cfloat* rn = run_base(A.sn, j0);
#pragma unroll
  for (int u = 0; u < kChunk; ++u) {
    const int j = j0 + u;
    const int jj = min(j, s1 - 1);
    const float4 p = PAD ? sload4(A.sp, jj) : sload4_run(rp, u);
    const float4 nv = PAD ? sload4(A.sn, jj) : sload4_run(rn, u);
    const float a = (!PAD || j < s1) ? p.w : 0.0f;
    const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float r = sqrt_hw(r2);
    w += a * (nv.x * dx + nv.y * dy + nv.z * dz) * rcp(fmaf(r, r2, 1e-5f));
  }
  return w;
}

__global__ __launch_bounds__(kB) void k_wn_walk(WnArgs A) {
  __shared__ int s_stack[kWavesPerBlock][kStackCap];
  __shared__ int s_far[kWavesPerBlock][kFarCap];
  __shared__ int s_near[kWavesPerBlock][kNearCap];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  int* stack = s_stack[wv];
  int* far = s_far[wv];
  int* near = s_near[wv];
  const int num_groups = A.qhdr->num_groups;
  unsigned long long n_far = 0, n_near = 0, n_groups = 0;
  int tries = 0;
  while (true) {
    const int g = next_group(A.qhdr, num_groups, lane, tries);
    if (g < 0) break;  // every wave reaches this exit
    const int2 grp = A.groups[g];
    const int base = __builtin_amdgcn_readfirstlane(grp.x);
    const int cnt = __builtin_amdgcn_readfirstlane(grp.y);
    if (cnt < 1 || cnt > 64 || base < 0 || base + cnt > A.m) continue;  // corrupt entry
    const bool valid = lane < cnt;
    const int i = base + (valid ? lane : 0);
    const float4 q = A.qs[i];
    const Box box = wave_box(q.x, q.y, q.z);
    float w = 0.f;
    constexpr int C = kChunk;
    auto flush_far = [&](int nfar) {
      n_far += static_cast<unsigned long long>(nfar) * cnt;
      const int padded = (nfar + C - 1) & ~(C - 1);
      if (lane < padded - nfar) far[nfar + lane] = A.sentinel;
      __builtin_amdgcn_wave_barrier();
      for (int f0 = 0; f0 < padded; f0 += C) w += wn_far_batch<C>(A, q.x, q.y, q.z, far + f0);
      __builtin_amdgcn_wave_barrier();
    };
    // consecutive near leaves with adjacent element ranges are merged
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
        s1 = __builtin_amdgcn_readfirstlane(s1);  // wave-uniform: keeps element loads scalar
        n_near += static_cast<unsigned long long>(s1 - s0) * cnt;
        int j0 = s0;
        for (; j0 + kChunk <= s1; j0 += kChunk)
          w += wn_near_batch<false>(A, q.x, q.y, q.z, j0, s1);
        if (j0 < s1) w += wn_near_batch<true>(A, q.x, q.y, q.z, j0, s1);
      }
    };
    traverse<false>(A.hdr, &A.qhdr->overflow, A.topo, A.geo, A.mp, A.beta, box, lane, 0,
                    stack, far, near, flush_far, flush_near);
    if (valid) A.out[A.qperm[i]] = w * kInvFourPi;
    ++n_groups;
  }
  if (A.stats && lane == 0) {
    atomicAdd(&A.qhdr->stats[0], n_far);
    atomicAdd(&A.qhdr->stats[1], n_near);
    atomicAdd(&A.qhdr->stats[2], n_groups);
  }
}

__global__ void k_wn_copy_stats(const Header* __restrict__ hdr, const Header* __restrict__ qhdr,
                                int64_t* __restrict__ out) {
  if (threadIdx.x != 0) return;
  out[0] = static_cast<int64_t>(qhdr->stats[0]);
  out[1] = static_cast<int64_t>(qhdr->stats[1]);
  out[2] = static_cast<int64_t>(qhdr->stats[2]);
  out[3] = hdr->num_nodes;
  int levels = 0;
  for (int l = 0; l < kMaxLevels; ++l)
    if (hdr->level_count[l] > 0) levels = l + 1;
  out[4] = levels;
  out[5] = hdr->overflow | qhdr->overflow;
  out[6] = 0;
  out[7] = 0;
}

int check_wn(const SbhBuffer* points, const SbhBuffer* normals, const SbhBuffer* areas,
             const SbhBuffer* queries, const SbhBuffer* out, float beta, int max_depth,
             int leaf_size) {
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  const int64_t n = points->dims[0];
  if (!buffer_ok(normals, SBH_DTYPE_F32, 2) || normals->dims[0] != n || normals->dims[1] != 3)
    return fail_invalid("normals must have shape (N, 3)");
  if (!buffer_ok(areas, SBH_DTYPE_F32, 1) || areas->dims[0] != n)
    return fail_invalid("areas must have shape (N,)");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2) || queries->dims[1] != 3)
    return fail_invalid("queries must have shape (M, 3)");
  if (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != queries->dims[0])
    return fail_invalid("out must have shape (M,)");
  if (n > (int64_t{1} << 30) || queries->dims[0] > (int64_t{1} << 30))
    return fail_invalid("too many points");
  if (!(beta >= 0.0f)) return fail_invalid("beta must be >= 0");
  if (max_depth < 1 || max_depth > 10) return fail_invalid("max_depth must be in [1, 10]");
  if (leaf_size < 1) return fail_invalid("leaf_size must be >= 1");
  return SBH_OK;
}

}  // namespace

}  // namespace tree
}  // namespace sbh

using namespace sbh;

extern "C" int64_t sbh_winding_number_workspace_bytes(int64_t num_elements, int64_t num_queries,
                                                      int32_t max_depth, int32_t leaf_size) {
  if (num_elements < 0 || num_queries < 0 || max_depth < 1 || max_depth > 10 || leaf_size < 1)
    return -1;
  return static_cast<int64_t>(
      tree::wn_layout(num_elements, num_queries, max_depth, leaf_size).total);
}

extern "C" int sbh_winding_number(void* stream, const SbhBuffer* points, const SbhBuffer* normals,
                                  const SbhBuffer* areas, const SbhBuffer* queries,
                                  SbhBuffer* out, float beta, int32_t max_depth,
                                  int32_t leaf_size, SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = tree::check_wn(points, normals, areas, queries, out, beta, max_depth, leaf_size);
  if (rc != SBH_OK) return rc;
  if (stats != nullptr &&
      (!buffer_ok(stats, SBH_DTYPE_S64, -1) || elem_count(stats) < 8))
    return fail_invalid("stats must be an int64 buffer of at least 8 elements");
  const int64_t n = points->dims[0], m = queries->dims[0];
  hipStream_t s = static_cast<hipStream_t>(stream);
  const tree::WnLayout W = tree::wn_layout(n, m, max_depth, leaf_size);
  char* ws = nullptr;
  if (workspace != nullptr) {
    if (!buffer_ok(workspace, SBH_DTYPE_U8, -1) ||
        byte_count(workspace) < static_cast<int64_t>(W.total))
      return fail_invalid("workspace too small");
    ws = static_cast<char*>(workspace->data);
  } else {
    if ((rc = cached_workspace(s, kWsPipeline, W.total, &ws)) != SBH_OK) return rc;
  }
  const auto* pts = static_cast<const float*>(points->data);
  const auto* q = static_cast<const float*>(queries->data);
  prof::Scope pr(s, 2);
  const bool profiled = pr.on() && m > 0;
  // Source tree over the centroids with areas as weights; the normals ride
  // along as the gathered per-point vector.
  auto* wd = reinterpret_cast<float4*>(ws + W.wd);
  rc = tree::build(s, ws, W.t, pts, static_cast<const float*>(areas->data),
                   static_cast<const float*>(normals->data), n, max_depth, leaf_size, nullptr,
                   tree::kBuildAll, wd);
  if (rc != SBH_OK) return rc;
  auto* hdr = reinterpret_cast<tree::Header*>(ws + W.t.header);
  const auto* topo = reinterpret_cast<const int4*>(ws + W.t.topo);
  const auto* mp = reinterpret_cast<const float4*>(ws + W.t.mp);
  const auto* sp = reinterpret_cast<const float4*>(ws + W.t.sp);
  const auto* sn = reinterpret_cast<const float4*>(ws + W.t.sc);
  // Query tree (topology and groups only).
  char* qws = ws + W.qbase;
  rc = tree::build(s, qws, W.qt, q, nullptr, nullptr, m, tree::kQueryDepth, tree::kQueryLeaf,
                   nullptr, tree::kBuildGroups);
  if (rc != SBH_OK) return rc;
  auto* qhdr = reinterpret_cast<tree::Header*>(qws + W.qt.header);
  if (m > 0) {
    tree::WnArgs A;
    A.hdr = hdr;
    A.qhdr = qhdr;
    A.groups = reinterpret_cast<const int2*>(qws + W.qt.groups);
    A.topo = topo;
    A.geo = reinterpret_cast<const float4*>(ws + W.t.geo);
    A.mp = mp;
    A.wd = wd;
    A.sp = sp;
    A.sn = sn;
    A.qs = reinterpret_cast<const float4*>(qws + W.qt.sp);
    A.qperm = reinterpret_cast<const uint32_t*>(qws + W.qt.perm);
    A.m = static_cast<int>(m);
    A.sentinel = static_cast<int>(W.t.cap);
    A.beta = beta;
    A.out = static_cast<float*>(out->data);
    A.stats = stats != nullptr;
    const unsigned grid = grid_for(m, 64 * tree::gw::kWavesPerBlock, 256 * 8);
    if (profiled) {
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[2], s));
    }
    tree::k_wn_walk<<<grid, tree::gw::kB, 0, s>>>(A);
    SBH_LAUNCH_CHECK();
    if (profiled) {
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[3], s));
      pr.commit();
    }
  }
  if (stats != nullptr) {
    tree::k_wn_copy_stats<<<1, 64, 0, s>>>(hdr, qhdr, static_cast<int64_t*>(stats->data));
    SBH_LAUNCH_CHECK();
  }
  return SBH_OK;
}
