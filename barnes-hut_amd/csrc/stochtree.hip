// Stochastic Barnes-Hut over the adaptive octree: the multi-level prefix
// control variate of src/sbh_ffi.cu:2146-2291 (forces), 2293-2442 (their
// query-side VJP) and, with the dipole kernel, the winding-number sampler of
// src/sbh_gpu.cu:166-187.
//
// Estimator, per query: every control subdomain (a node at level 3, or a
// leaf above it) contributes its node term (or its exact leaf sum); then
// each of S samples draws a source of the subdomain uniformly and walks from
// the subdomain towards it with Russian roulette -- at node N continue with
// probability t = min(beta_cur / ratio(N), 1), ratio = |q - cell centre| /
// cell diagonal, beta_cur starting at max(beta, ratio(subdomain)) and rising
// to the largest ratio passed -- adding n_sub / (P * count(N)) * (sum of N's
// children - N), P the product of the continuation probabilities so far.
// Children that are leaves are summed exactly.  Its expectation is the exact
// sum over the subdomain (the differences telescope).
//
// Wave64 design: one wavefront per query group (<= 64 queries of a query
// tree cell, the groups of the classic walk).  The group shares every draw
// -- the sampled source and the roulette uniforms, Philox4x32-10 keyed by
// (seed, first sorted query of the group, subdomain node, sample, step) --
// so all lanes walk ONE path per (subdomain, sample): the nodes, children
// and source records of a step are read once per wave through the scalar
// cache and evaluated by every lane, and each lane keeps its own
// continuation probability, P and survival (a lane whose draw exceeds its t
// stops adding; the wave stops when no lane is left).  Each query's
// estimate has the reference's distribution (a uniform source, independent
// uniforms along its path); queries of one group share randomness.
#include "common.hpp"
#include "groupwalk.hpp"
#include "profile.hpp"
#include "tree.hpp"

namespace sbh {
namespace tree {
namespace {

using namespace gw;

constexpr int kControlLevel = 3;      // src/sbh_ffi.cu:2167
constexpr int kMaxSubdomains = 512;   // disjoint cells of level <= 3
constexpr float kInvFourPi = 0.07957747154594767f;

enum { kSForce = 0, kSVjp = 1, kSWn = 2 };

struct StochLayout {
  Layout t;      // source tree at the workspace base
  size_t wd;     // per source node: dipole (winding numbers)
  size_t subs;   // control subdomains: count, then node indices
  size_t qbase;  // query tree at workspace + qbase
  Layout qt;
  size_t total;
};

StochLayout stoch_layout(int64_t n, int64_t m, int max_depth, int leaf_size) {
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  StochLayout W;
  W.t = layout(n, max_depth, leaf_size);
  size_t o = W.t.total;
  W.wd = o; o = align(o + 16 * (static_cast<size_t>(W.t.cap) + 1));
  W.subs = o; o = align(o + 4 * (kMaxSubdomains + 1));
  W.qbase = o;
  W.qt = layout(m, kQueryTreeDepth, kQueryTreeLeaf);
  W.total = W.qbase + W.qt.total;
  return W;
}

// Control subdomains in node order: nodes at level 3 and leaves at levels
// 0-2, non-empty (mass or area != 0).  One wave; subs[0] = count, the nodes
// follow.  More than kMaxSubdomains (impossible for one tree: the cells are
// disjoint cells of the level-3 grid) sets bit 8 of the overflow word.
__global__ __launch_bounds__(64) void k_subdomains(Header* __restrict__ hdr,
                                                   const int4* __restrict__ topo,
                                                   const float4* __restrict__ mp,
                                                   int32_t* __restrict__ subs) {
  const int lane = threadIdx.x;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int ns = 0;
  for (int l = 0; l <= kControlLevel; ++l) {
    const int cnt = hdr->level_count[l];
    if (cnt <= 0) break;
    const int st = hdr->level_start[l];
    for (int b = 0; b < cnt; b += 64) {
      const int k = st + b + lane;
      bool take = false;
      if (b + lane < cnt) {
        const int4 t = topo[k];
        take = (l == kControlLevel || t.w == 0) && mp[k].w != 0.0f;
      }
      const uint64_t bal = __ballot(take);
      const int pos = ns + __popcll(bal & lt);
      if (take && pos < kMaxSubdomains) subs[1 + pos] = k;
      ns += __popcll(bal);
    }
  }
  if (lane == 0) {
    if (ns > kMaxSubdomains) atomicOr(&hdr->overflow, 8);
    subs[0] = min(ns, kMaxSubdomains);
  }
}

struct StochArgs {
  Header* qhdr;           // query tree: groups, work queues, statistics
  const int2* groups;     // query groups (start, count), sorted query order
  const int4* topo;       // source tree
  const float4* geo;      // cell centre, cell diagonal
  const float4* mp;       // centre of mass (centroid), mass (area)
  const float4* qd;       // second moments (gravity, quadrupole)
  const float4* wd;       // dipoles (winding numbers)
  const float4* sp;       // sorted sources (x, y, z, mass or area)
  const float4* sn;       // sorted unit normals (winding numbers)
  const float4* qs;       // sorted queries
  const float4* qc;       // sorted cotangents (VJP)
  const uint32_t* qperm;  // sorted query -> input row
  const int32_t* subs;    // count, then subdomain nodes
  int m;
  float beta;
  float e2;  // eps^2 + 1e-12
  int samples;
  uint32_t seed;
  float* out;
  bool stats;
};

struct Q {
  float x, y, z;     // query
  float cx, cy, cz;  // cotangent (VJP)
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Node term of the lane's query, added to (ax, ay, az) (winding numbers: ax).
template <int KIND, bool QUAD>
__device__ __forceinline__ void node_term(const StochArgs& A, int k, const Q& L, float& ax,
                                          float& ay, float& az) {
  const float4 m = sload4(A.mp, k);
  if constexpr (KIND == kSWn) {
    const float4 d = sload4(A.wd, k);
    const float Xx = m.x - L.x, Xy = m.y - L.y, Xz = m.z - L.z;
    const float R2 = Xx * Xx + Xy * Xy + Xz * Xz;
    const float R = sqrt_hw(R2);
    ax += (d.x * Xx + d.y * Xy + d.z * Xz) * kInvFourPi * rcp(fmaf(R, R2, 1e-5f));
  } else {
    const float Dx = L.x - m.x, Dy = L.y - m.y, Dz = L.z - m.z;
    const float s = rsq(fmaf(Dz, Dz, fmaf(Dy, Dy, fmaf(Dx, Dx, A.e2))));
    const float s2 = s * s, s3 = s2 * s;
    if constexpr (!QUAD) {
      const float a = -m.w * s3;
      if constexpr (KIND == kSForce) {
        ax = fmaf(a, Dx, ax);
        ay = fmaf(a, Dy, ay);
        az = fmaf(a, Dz, az);
      } else {
        // J^T c = a c + 3 M s^5 (c . D) D
        const float u = L.cx * Dx + L.cy * Dy + L.cz * Dz;
        const float kD = -3.0f * a * s2 * u;
        ax += kD * Dx + a * L.cx;
        ay += kD * Dy + a * L.cy;
        az += kD * Dz + a * L.cz;
      }
    } else {
      const float4 q0 = sload4(A.qd, 2 * k);
      const float4 q1 = sload4(A.qd, 2 * k + 1);
      // F = a D + b S D, a = -M s^3 + 1.5 tr(S) s^5 - 7.5 s^7 (D.S.D), b = 3 s^5
      const float SDx = q0.x * Dx + q0.w * Dy + q1.x * Dz;
      const float SDy = q0.w * Dx + q0.y * Dy + q1.y * Dz;
      const float SDz = q1.x * Dx + q1.y * Dy + q0.z * Dz;
      const float tq = Dx * SDx + Dy * SDy + Dz * SDz;
      const float wtq = s2 * tq;
      const float a = s3 * fmaf(s2, fmaf(-7.5f, wtq, q1.z), -m.w);
      const float b = 3.0f * s3 * s2;
      if constexpr (KIND == kSForce) {
        ax = fmaf(a, Dx, fmaf(b, SDx, ax));
        ay = fmaf(a, Dy, fmaf(b, SDy, ay));
        az = fmaf(a, Dz, fmaf(b, SDz, az));
      } else {
        // d(c.F)/dD = kD D + a c + b S c - 15 s^7 (c.D) S D,
        // kD = (c.D) b (M + s^2 (17.5 s^2 t - 2.5 tr S)) - 15 s^7 (c.S D)
        const float u = L.cx * Dx + L.cy * Dy + L.cz * Dz;
        const float Scx = q0.x * L.cx + q0.w * L.cy + q1.x * L.cz;
        const float Scy = q0.w * L.cx + q0.y * L.cy + q1.y * L.cz;
        const float Scz = q1.x * L.cx + q1.y * L.cy + q0.z * L.cz;
        const float cw = Scx * Dx + Scy * Dy + Scz * Dz;
        const float A1 = b * fmaf(s2, fmaf(17.5f, wtq, q1.w), m.w);
        const float m15s7 = -15.0f * (s3 * s2 * s2);
        const float kD = fmaf(m15s7, cw, u * A1);
        const float b15 = m15s7 * u;
        ax += kD * Dx + a * L.cx + b * Scx + b15 * SDx;
        ay += kD * Dy + a * L.cy + b * Scy + b15 * SDy;
        az += kD * Dz + a * L.cz + b * Scz + b15 * SDz;
      }
    }
  }
}

// kChunk sources from b for leaf_sum: with PAD, entries past j1 carry zero
// weight; without, the batch lies inside [b, j1) and reads its records as
// runs, which the scalar loads fetch 4 at a time.
template <int KIND, bool PAD>
__device__ __forceinline__ void leaf_batch(const StochArgs& A, int b, int j1, const Q& L,
                                           float& ax, float& ay, float& az) {
  cfloat* rp = run_base(A.sp, b);
  cfloat* rn = run_base(A.sn, b);
#pragma unroll
  for (int u = 0; u < kChunk; ++u) {
    const int j = b + u;
    const int jj = min(j, j1 - 1);
    const float4 p = PAD ? sload4(A.sp, jj) : sload4_run(rp, u);
    const float w = PAD ? __uint_as_float(__builtin_amdgcn_readfirstlane(
                              __float_as_uint(p.w) & (j < j1 ? ~0u : 0u)))
                        : p.w;
    if constexpr (KIND == kSWn) {
      const float4 nv = PAD ? sload4(A.sn, jj) : sload4_run(rn, u);
      const float dx = p.x - L.x, dy = p.y - L.y, dz = p.z - L.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float r = sqrt_hw(r2);
      ax += w * (nv.x * dx + nv.y * dy + nv.z * dz) * kInvFourPi * rcp(fmaf(r, r2, 1e-5f));
    } else {
      const float dx = L.x - p.x, dy = L.y - p.y, dz = L.z - p.z;
      const float s = rsq(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, A.e2))));
      const float s2 = s * s, ws3 = w * s2 * s;
      if constexpr (KIND == kSForce) {
        ax = fmaf(-ws3, dx, ax);
        ay = fmaf(-ws3, dy, ay);
        az = fmaf(-ws3, dz, az);
      } else {
        // w s^3 (3 s^2 (c.d) d - c)
        const float k3 = (3.0f * s2) * (L.cx * dx + L.cy * dy + L.cz * dz);
        ax += ws3 * fmaf(k3, dx, -L.cx);
        ay += ws3 * fmaf(k3, dy, -L.cy);
        az += ws3 * fmaf(k3, dz, -L.cz);
      }
    }
  }
}

// Exact sum of the lane's query over the sorted sources [j0, j1), in batches
// of kChunk (the last one padded).  The query-side VJP keeps padded batches
// throughout: with record runs it measured slower (10.9 -> 11.6 ms per 2^20
// queries) while the force went 9.2 -> 8.1 ms.
template <int KIND>
__device__ __forceinline__ void leaf_sum(const StochArgs& A, int j0, int j1, const Q& L,
                                         float& ax, float& ay, float& az) {
  int b = j0;
  if constexpr (KIND != kSVjp)
    for (; b + kChunk <= j1; b += kChunk) leaf_batch<KIND, false>(A, b, j1, L, ax, ay, az);
  for (; b < j1; b += kChunk) leaf_batch<KIND, true>(A, b, j1, L, ax, ay, az);
}

// |q - cell centre| / cell diagonal in IEEE fp32 (the oracle's expression).
__device__ __forceinline__ float centre_ratio(const Q& L, float4 g) {
#pragma clang fp contract(off)
  const float dx = L.x - g.x, dy = L.y - g.y, dz = L.z - g.z;
  return sqrtf(dx * dx + dy * dy + dz * dz) / g.w;
}

// Persistent waves over the query groups (XCD-aware queue of groupwalk.hpp).
template <int KIND, bool QUAD>
__global__ __launch_bounds__(kB) void k_stoch_walk(StochArgs A) {
  constexpr bool kVec = KIND != kSWn;
  const int lane = threadIdx.x & 63;
  const int num_groups = A.qhdr->num_groups;
  const int nsub = __builtin_amdgcn_readfirstlane(A.subs[0]);
  const float inv_s = A.samples > 0 ? 1.0f / static_cast<float>(A.samples) : 0.0f;
  unsigned long long n_node = 0, n_pair = 0, n_groups = 0;
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
    Q L;
    {
      const float4 q = A.qs[i];
      L.x = q.x;
      L.y = q.y;
      L.z = q.z;
      L.cx = L.cy = L.cz = 0.f;
      if constexpr (KIND == kSVjp) {
        const float4 c = A.qc[i];
        L.cx = c.x;
        L.cy = c.y;
        L.cz = c.z;
      }
    }
    const uint32_t key = static_cast<uint32_t>(base);
    float tx = 0.f, ty = 0.f, tz = 0.f;
    unsigned long long w_node = 0, w_pair = 0;  // terms this wave evaluates
    for (int sd = 0; sd < nsub; ++sd) {
      const int k = __builtin_amdgcn_readfirstlane(A.subs[1 + sd]);
      const int4 t = sload4(A.topo, k);
      if (t.w == 0) {
        leaf_sum<KIND>(A, t.x, t.x + t.y, L, tx, ty, tz);
        w_pair += static_cast<unsigned>(t.y);
        continue;
      }
      node_term<KIND, QUAD>(A, k, L, tx, ty, tz);
      ++w_node;
      if (A.samples <= 0) continue;
      const float4 g0 = sload4(A.geo, k);
      const uint32_t n_sub = static_cast<uint32_t>(t.y);
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (int s = 0; s < A.samples; ++s) {
        const float u0 = sbh_uniform(A.seed, key, static_cast<uint32_t>(k),
                                     static_cast<uint32_t>(s), 0u);
        const uint32_t sidx = __builtin_amdgcn_readfirstlane(
            min(static_cast<uint32_t>(static_cast<float>(n_sub) * u0) +
                    static_cast<uint32_t>(t.x),
                static_cast<uint32_t>(t.x + t.y - 1)));
        int cur = k;
        int4 tc = t;
        float4 gc = g0;
        float ep = 1.0f;
        bool alive = valid;
        float cur_beta = fmaxf(A.beta, centre_ratio(L, gc));
        float wx = 0.f, wy = 0.f, wz = 0.f;
        uint32_t step = 1;
        while (tc.w > 0 && step < 64) {
          const float ratio = centre_ratio(L, gc);
          float tp;
          {
#pragma clang fp contract(off)
            tp = fminf(cur_beta / ratio, 1.0f);
          }
          const float pu = sbh_uniform(A.seed, key, static_cast<uint32_t>(k),
                                       static_cast<uint32_t>(s), step++);
          alive = alive && !(pu > tp);
          if (__ballot(alive) == 0ull) break;  // every lane has stopped
          if (alive) ep *= tp;
          const int prev = cur;
          const int fc = __builtin_amdgcn_readfirstlane(tc.z);
          const int nch = __builtin_amdgcn_readfirstlane(tc.w);
          const int cnt_prev = __builtin_amdgcn_readfirstlane(tc.y);
          // the child holding the sampled source (children partition the range)
          int nxt = fc;
          for (int c = 1; c < nch; ++c)
            if (static_cast<uint32_t>(sload4(A.topo, fc + c).x) <= sidx) nxt = fc + c;
          cur = __builtin_amdgcn_readfirstlane(nxt);
          cur_beta = fmaxf(cur_beta, ratio);
          // sum of prev's children minus prev's own term
          float cx = 0.f, cy = 0.f, cz = 0.f;
          for (int c = 0; c < nch; ++c) {
            const int4 tk = sload4(A.topo, fc + c);
            if (tk.w == 0) {
              leaf_sum<KIND>(A, tk.x, tk.x + tk.y, L, cx, cy, cz);
              w_pair += static_cast<unsigned>(tk.y);
            } else {
              node_term<KIND, QUAD>(A, fc + c, L, cx, cy, cz);
              ++w_node;
            }
          }
          float px = 0.f, py = 0.f, pz = 0.f;
          node_term<KIND, QUAD>(A, prev, L, px, py, pz);
          ++w_node;
          const float w =
              alive ? static_cast<float>(n_sub) / (ep * static_cast<float>(cnt_prev)) : 0.0f;
          wx += w * (cx - px);
          if constexpr (kVec) {
            wy += w * (cy - py);
            wz += w * (cz - pz);
          }
          tc = sload4(A.topo, cur);
          gc = sload4(A.geo, cur);
        }
        sx += wx;
        sy += wy;
        sz += wz;
      }
      tx += sx * inv_s;
      ty += sy * inv_s;
      tz += sz * inv_s;
    }
    if (valid) {
      const int64_t o = static_cast<int64_t>(A.qperm[i]);
      if constexpr (kVec) {
        A.out[3 * o + 0] = tx;
        A.out[3 * o + 1] = ty;
        A.out[3 * o + 2] = tz;
      } else {
        A.out[o] = tx;
      }
    }
    n_node += w_node * static_cast<unsigned>(cnt);
    n_pair += w_pair * static_cast<unsigned>(cnt);
    ++n_groups;
  }
  if (A.stats && lane == 0) {
    atomicAdd(&A.qhdr->stats[0], n_node);
    atomicAdd(&A.qhdr->stats[1], n_pair);
    atomicAdd(&A.qhdr->stats[2], n_groups);
  }
}

__global__ void k_stoch_stats(const Header* __restrict__ hdr, const Header* __restrict__ qhdr,
                              const int32_t* __restrict__ subs, int64_t* __restrict__ out) {
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
  out[6] = subs[0];
  out[7] = 0;
}

int check_counts(const SbhBuffer* points, const SbhBuffer* queries, float beta, int max_depth,
                 int leaf_size, int samples, const SbhBuffer* stats) {
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2) || queries->dims[1] != 3)
    return fail_invalid("queries must have shape (M, 3)");
  if (points->dims[0] > (int64_t{1} << 30) || queries->dims[0] > (int64_t{1} << 30))
    return fail_invalid("too many points");
  if (!(beta >= 0.0f)) return fail_invalid("beta must be >= 0");
  if (max_depth < 1 || max_depth > 10) return fail_invalid("max_depth must be in [1, 10]");
  if (leaf_size < 1) return fail_invalid("leaf_size must be >= 1");
  if (samples < 0) return fail_invalid("samples_per_subdomain must be >= 0");
  if (stats != nullptr && (!buffer_ok(stats, SBH_DTYPE_S64, -1) || elem_count(stats) < 8))
    return fail_invalid("stats must be an int64 buffer of at least 8 elements");
  return SBH_OK;
}

// Builds the source tree (moments; dipoles for winding numbers), the
// subdomain list and the query tree, then runs the walk.  `vec` (per source:
// cotangent-free) is the normals for winding numbers; `cot` the query
// cotangents for the VJP.
int run_stoch(int kind, bool quad, hipStream_t s, const float* pts, const float* ms,
              const float* normals, const float* q, const float* cot, int64_t n, int64_t m,
              float* out, float beta, float softening, int max_depth, int leaf_size, int samples,
              int seed, SbhBuffer* workspace, SbhBuffer* stats) {
  const StochLayout W = stoch_layout(n, m, max_depth, leaf_size);
  char* ws = nullptr;
  if (workspace != nullptr) {
    if (!buffer_ok(workspace, SBH_DTYPE_U8, -1) ||
        byte_count(workspace) < static_cast<int64_t>(W.total))
      return fail_invalid("workspace too small");
    ws = static_cast<char*>(workspace->data);
  } else {
    int rc = cached_workspace(s, kWsPipeline, W.total, &ws);
    if (rc != SBH_OK) return rc;
  }
  prof::Scope pr(s, 3);
  auto* wd = reinterpret_cast<float4*>(ws + W.wd);
  int rc = build(s, ws, W.t, pts, ms, normals, n, max_depth, leaf_size, nullptr, kBuildMoments,
                 kind == kSWn ? wd : nullptr);
  if (rc != SBH_OK) return rc;
  auto* hdr = reinterpret_cast<Header*>(ws + W.t.header);
  auto* subs = reinterpret_cast<int32_t*>(ws + W.subs);
  k_subdomains<<<1, 64, 0, s>>>(hdr, reinterpret_cast<const int4*>(ws + W.t.topo),
                                reinterpret_cast<const float4*>(ws + W.t.mp), subs);
  SBH_LAUNCH_CHECK();
  char* qws = ws + W.qbase;
  rc = build(s, qws, W.qt, q, nullptr, cot, m, kQueryTreeDepth, kQueryTreeLeaf, nullptr,
             kBuildGroups);
  if (rc != SBH_OK) return rc;
  auto* qhdr = reinterpret_cast<Header*>(qws + W.qt.header);
  if (m > 0) {
    StochArgs A;
    A.qhdr = qhdr;
    A.groups = reinterpret_cast<const int2*>(qws + W.qt.groups);
    A.topo = reinterpret_cast<const int4*>(ws + W.t.topo);
    A.geo = reinterpret_cast<const float4*>(ws + W.t.geo);
    A.mp = reinterpret_cast<const float4*>(ws + W.t.mp);
    A.qd = reinterpret_cast<const float4*>(ws + W.t.qd);
    A.wd = wd;
    A.sp = reinterpret_cast<const float4*>(ws + W.t.sp);
    A.sn = reinterpret_cast<const float4*>(ws + W.t.sc);
    A.qs = reinterpret_cast<const float4*>(qws + W.qt.sp);
    A.qc = reinterpret_cast<const float4*>(qws + W.qt.sc);
    A.qperm = reinterpret_cast<const uint32_t*>(qws + W.qt.perm);
    A.subs = subs;
    A.m = static_cast<int>(m);
    A.beta = beta;
    const float eps = fmaxf(softening, 0.0f);
    A.e2 = eps * eps + 1e-12f;
    A.samples = samples;
    A.seed = static_cast<uint32_t>(seed);
    A.out = out;
    A.stats = stats != nullptr;
    const unsigned grid = grid_for(m, 64 * kWavesPerBlock, 256 * 8);
    if (pr.on()) {
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[2], s));
    }
    if (kind == kSForce && quad)
      k_stoch_walk<kSForce, true><<<grid, kB, 0, s>>>(A);
    else if (kind == kSForce)
      k_stoch_walk<kSForce, false><<<grid, kB, 0, s>>>(A);
    else if (kind == kSVjp && quad)
      k_stoch_walk<kSVjp, true><<<grid, kB, 0, s>>>(A);
    else if (kind == kSVjp)
      k_stoch_walk<kSVjp, false><<<grid, kB, 0, s>>>(A);
    else
      k_stoch_walk<kSWn, false><<<grid, kB, 0, s>>>(A);
    SBH_LAUNCH_CHECK();
    if (pr.on()) {
      SBH_HIP_TRY(hipEventRecord(pr.rec().ev[3], s));
      pr.commit();
    }
  }
  if (stats != nullptr) {
    k_stoch_stats<<<1, 64, 0, s>>>(hdr, qhdr, subs, static_cast<int64_t*>(stats->data));
    SBH_LAUNCH_CHECK();
  }
  return SBH_OK;
}

}  // namespace
}  // namespace tree
}  // namespace sbh

using namespace sbh;

extern "C" int64_t sbh_stochastic_workspace_bytes(int64_t num_sources, int64_t num_queries,
                                                  int32_t max_depth, int32_t leaf_size) {
  if (num_sources < 0 || num_queries < 0 || max_depth < 1 || max_depth > 10 || leaf_size < 1)
    return -1;
  return static_cast<int64_t>(
      tree::stoch_layout(num_sources, num_queries, max_depth, leaf_size).total);
}

extern "C" int sbh_stochastic_gravity_force(void* stream, const SbhBuffer* points,
                                            const SbhBuffer* masses, const SbhBuffer* queries,
                                            SbhBuffer* out, float beta, float softening_scale,
                                            int32_t max_depth, int32_t leaf_size,
                                            int32_t multipole_order,
                                            int32_t samples_per_subdomain, int32_t seed,
                                            SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = tree::check_counts(points, queries, beta, max_depth, leaf_size,
                              samples_per_subdomain, stats);
  if (rc != SBH_OK) return rc;
  const int64_t n = points->dims[0], m = queries->dims[0];
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != n)
    return fail_invalid("masses must have shape (N,)");
  if (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != m || out->dims[1] != 3)
    return fail_invalid("out must have shape (M, 3)");
  if (softening_scale < 0.0f) return fail_invalid("softening_scale must be >= 0");
  if (multipole_order != 1 && multipole_order != 2)
    return fail_invalid("multipole_order must be 1 or 2");
  return tree::run_stoch(tree::kSForce, multipole_order == 2, static_cast<hipStream_t>(stream),
                         static_cast<const float*>(points->data),
                         static_cast<const float*>(masses->data), nullptr,
                         static_cast<const float*>(queries->data), nullptr, n, m,
                         static_cast<float*>(out->data), beta, softening_scale, max_depth,
                         leaf_size, samples_per_subdomain, seed, workspace, stats);
}

extern "C" int sbh_stochastic_gravity_force_vjp(
    void* stream, const SbhBuffer* points, const SbhBuffer* masses, const SbhBuffer* queries,
    const SbhBuffer* cotangent, SbhBuffer* out, float beta, float softening_scale,
    int32_t max_depth, int32_t leaf_size, int32_t multipole_order,
    int32_t samples_per_subdomain, int32_t seed, SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = tree::check_counts(points, queries, beta, max_depth, leaf_size,
                              samples_per_subdomain, stats);
  if (rc != SBH_OK) return rc;
  const int64_t n = points->dims[0], m = queries->dims[0];
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != n)
    return fail_invalid("masses must have shape (N,)");
  if (!buffer_ok(cotangent, SBH_DTYPE_F32, 2) || cotangent->dims[0] != m ||
      cotangent->dims[1] != 3)
    return fail_invalid("cotangent must have shape (M, 3)");
  if (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != m || out->dims[1] != 3)
    return fail_invalid("out must have shape (M, 3)");
  if (softening_scale < 0.0f) return fail_invalid("softening_scale must be >= 0");
  if (multipole_order != 1 && multipole_order != 2)
    return fail_invalid("multipole_order must be 1 or 2");
  return tree::run_stoch(tree::kSVjp, multipole_order == 2, static_cast<hipStream_t>(stream),
                         static_cast<const float*>(points->data),
                         static_cast<const float*>(masses->data), nullptr,
                         static_cast<const float*>(queries->data),
                         static_cast<const float*>(cotangent->data), n, m,
                         static_cast<float*>(out->data), beta, softening_scale, max_depth,
                         leaf_size, samples_per_subdomain, seed, workspace, stats);
}

extern "C" int sbh_stochastic_winding_number(void* stream, const SbhBuffer* points,
                                             const SbhBuffer* normals, const SbhBuffer* areas,
                                             const SbhBuffer* queries, SbhBuffer* out,
                                             float beta, int32_t max_depth, int32_t leaf_size,
                                             int32_t samples_per_subdomain, int32_t seed,
                                             SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = tree::check_counts(points, queries, beta, max_depth, leaf_size,
                              samples_per_subdomain, stats);
  if (rc != SBH_OK) return rc;
  const int64_t n = points->dims[0], m = queries->dims[0];
  if (!buffer_ok(normals, SBH_DTYPE_F32, 2) || normals->dims[0] != n || normals->dims[1] != 3)
    return fail_invalid("normals must have shape (N, 3)");
  if (!buffer_ok(areas, SBH_DTYPE_F32, 1) || areas->dims[0] != n)
    return fail_invalid("areas must have shape (N,)");
  if (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != m)
    return fail_invalid("out must have shape (M,)");
  return tree::run_stoch(tree::kSWn, false, static_cast<hipStream_t>(stream),
                         static_cast<const float*>(points->data),
                         static_cast<const float*>(areas->data),
                         static_cast<const float*>(normals->data),
                         static_cast<const float*>(queries->data), nullptr, n, m,
                         static_cast<float*>(out->data), beta, 0.0f, max_depth, leaf_size,
                         samples_per_subdomain, seed, workspace, stats);
}
