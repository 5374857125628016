// Device helpers shared by the classic and stochastic walks over the dense
// octree8 records: walk parameters (passed by value, never __constant__),
// the far-field test, the AABB prune and the four pair kernels.
#pragma once

#include "common.hpp"

namespace sbh {
namespace o8 {

struct WalkParams {
  float beta;
  float eps2;
  float rc;
  int cutoff;  // attenuation active
  int periodic;
  float box;
  float inv_box;
  int prune;
  float r_cut_sq;
};

// Distance from q to the cell centre with the kernel.h min-image
// (src/sbh_ffi.cu:1885-1891), IEEE fp32, not contracted.
__device__ __forceinline__ float centre_dist(const Node8* __restrict__ nd, float qx, float qy,
                                             float qz, const WalkParams& P) {
#pragma clang fp contract(off)
  float dx = min_image_floor(qx - nd->cx, P.periodic, P.box, P.inv_box);
  float dy = min_image_floor(qy - nd->cy, P.periodic, P.box, P.inv_box);
  float dz = min_image_floor(qz - nd->cz, P.periodic, P.box, P.inv_box);
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// Far-field test (src/sbh_ffi.cu:1878-1893).
__device__ __forceinline__ bool use_far_field(const Node8* __restrict__ nd, float qx, float qy,
                                              float qz, const WalkParams& P) {
#pragma clang fp contract(off)
  if (nd->is_leaf) return true;
  return centre_dist(nd, qx, qy, qz, P) / nd->diameter >= P.beta;
}

__device__ __forceinline__ float axis_dist(float q, float lo, float hi) {
  if (q < lo) return lo - q;
  if (q > hi) return q - hi;
  return 0.0f;
}

// AABB prune (src/sbh_ffi.cu:1903-1948).
__device__ __forceinline__ bool should_prune(const Node8* __restrict__ nd, float qx, float qy,
                                             float qz, const WalkParams& P) {
#pragma clang fp contract(off)
  if (!P.prune) return false;
  if (P.periodic) {
    qx -= P.box * rintf((qx - nd->cx) * P.inv_box);
    qy -= P.box * rintf((qy - nd->cy) * P.inv_box);
    qz -= P.box * rintf((qz - nd->cz) * P.inv_box);
  }
  float h = nd->half_size;
  float dx = axis_dist(qx, nd->cx - h, nd->cx + h);
  float dy = axis_dist(qy, nd->cy - h, nd->cy + h);
  float dz = axis_dist(qz, nd->cz - h, nd->cz + h);
  return dx * dx + dy * dy + dz * dz > P.r_cut_sq;
}

// Child octant of point p by comparison with the cell centre
// (src/sbh_ffi.cu:1895-1901).
__device__ __forceinline__ uint32_t child_slot(const Node8* __restrict__ nd, float px, float py,
                                               float pz) {
  uint32_t ix = px >= nd->cx ? 1u : 0u;
  uint32_t iy = py >= nd->cy ? 1u : 0u;
  uint32_t iz = pz >= nd->cz ? 1u : 0u;
  return ix | (iy << 1) | (iz << 2);
}

enum WalkKind {
  kGravGrad = 0,
  kSoftForce = 1,
  kSoftVjp = 2,
  kGravVjp = 3,
  kWinding = 4,
  kPotential = 5,   // gravityPotential (include/kernel.h:55-60)
  kSmoothDist = 6,  // smoothDistExp (include/kernel.h:82-87)
};

// Kinds with one value per query (the others give a 3-vector).
template <int KIND>
constexpr bool kScalarKind = KIND == kWinding || KIND == kPotential || KIND == kSmoothDist;
// Kinds that use the softened cutoff force and hence the AABB prune.
template <int KIND>
constexpr bool kPrunedKind = KIND == kSoftForce || KIND == kSoftVjp;

constexpr float kInvFourPi = 0.07957747154594767f;

// Query count from which the per-query walks run in Morton order of the
// queries (below it the ordering costs more than it saves).
constexpr int64_t kOrderMin = 4096;

// signedSolidAngle<3> (include/kernel.h:73-80): a n.(p - q) / (4 pi (|p - q|^3 + 1e-5)).
__device__ __forceinline__ float wn_pair(const Leaf3& L, float qx, float qy, float qz,
                                         const WalkParams& P) {
  float dx = min_image_floor(L.pos[0] - qx, P.periodic, P.box, P.inv_box);
  float dy = min_image_floor(L.pos[1] - qy, P.periodic, P.box, P.inv_box);
  float dz = min_image_floor(L.pos[2] - qz, P.periodic, P.box, P.inv_box);
  float r = sqrtf(dx * dx + dy * dy + dz * dz);
  float nd = L.normal[0] * dx + L.normal[1] * dy + L.normal[2] * dz;
  return L.mass * nd * kInvFourPi / (r * r * r + 1e-5f);
}

// gravityPotential (include/kernel.h:55-60): -m / (|p - q| + 0.1).
__device__ __forceinline__ float potential_pair(const Leaf3& L, float qx, float qy, float qz,
                                                const WalkParams& P) {
  float dx = min_image_floor(L.pos[0] - qx, P.periodic, P.box, P.inv_box);
  float dy = min_image_floor(L.pos[1] - qy, P.periodic, P.box, P.inv_box);
  float dz = min_image_floor(L.pos[2] - qz, P.periodic, P.box, P.inv_box);
  return -L.mass / (sqrtf(dx * dx + dy * dy + dz * dz) + 1e-1f);
}

// smoothDistExp (include/kernel.h:82-87): exp(-alpha |p - q|).
__device__ __forceinline__ float smooth_pair(const Leaf3& L, float qx, float qy, float qz,
                                             const WalkParams& P) {
  float dx = min_image_floor(L.pos[0] - qx, P.periodic, P.box, P.inv_box);
  float dy = min_image_floor(L.pos[1] - qy, P.periodic, P.box, P.inv_box);
  float dz = min_image_floor(L.pos[2] - qz, P.periodic, P.box, P.inv_box);
  return expf(-L.alpha * sqrtf(dx * dx + dy * dy + dz * dz));
}

// Pair/node contribution to the accumulator.
template <int KIND>
__device__ __forceinline__ void contribution(float px, float py, float pz, float m, float qx,
                                             float qy, float qz, float gx, float gy, float gz,
                                             const WalkParams& P, float& ax, float& ay,
                                             float& az) {
  float dx = min_image_floor(qx - px, P.periodic, P.box, P.inv_box);
  float dy = min_image_floor(qy - py, P.periodic, P.box, P.inv_box);
  float dz = min_image_floor(qz - pz, P.periodic, P.box, P.inv_box);
  if constexpr (KIND == kGravGrad) {
    // gravityGrad (include/kernel.h:89-94)
    float t = sqrtf(dx * dx + dy * dy + dz * dz) + 1e-1f;
    float c = -m / (t * t * t);
    ax += c * dx;
    ay += c * dy;
    az += c * dz;
  } else if constexpr (KIND == kGravVjp) {
    // gravityGradVjp (src/sbh_ffi.cu:2044-2060)
    float r2 = dx * dx + dy * dy + dz * dz;
    float dist = sqrtf(r2 + 1e-12f);
    float den = dist + 1e-1f;
    float i3 = 1.0f / (den * den * den);
    float i4 = i3 / den;
    float rg = dx * gx + dy * gy + dz * gz;
    float k = 3.0f * rg * (1.0f / (dist + 1e-12f)) * i4;
    ax += -m * (i3 * gx - k * dx);
    ay += -m * (i3 * gy - k * dy);
    az += -m * (i3 * gz - k * dz);
  } else {
    float r2 = dx * dx + dy * dy + dz * dz;
    float inv_r = rsq_softened(r2 + P.eps2 + 1e-12f);
    float s3 = inv_r * inv_r * inv_r;
    float att = 1.0f, cc = 0.0f;
    if (P.cutoff) {
      float e, dist;
      att = attenuation(r2, P.rc, &e, &dist);
      if constexpr (KIND == kSoftVjp) cc = s3 * dist * e / (kTwoSqrtPi * P.rc * P.rc * P.rc);
    }
    if constexpr (KIND == kSoftForce) {
      // gravityGradSoftCutoff (src/sbh_ffi.cu:142-160)
      float c = -m * att * s3;
      ax += c * dx;
      ay += c * dy;
      az += c * dz;
    } else {
      // gravityGradSoftCutoffVjp (src/sbh_ffi.cu:2062-2090)
      float s5 = s3 * inv_r * inv_r;
      float rg = dx * gx + dy * gy + dz * gz;
      float lin = att * s3;
      float rad = rg * (3.0f * att * s5 + cc);
      ax += -m * (lin * gx - rad * dx);
      ay += -m * (lin * gy - rad * dy);
      az += -m * (lin * gz - rad * dz);
    }
  }
}

template <int KIND>
__device__ __forceinline__ void contribution(const Leaf3& L, float qx, float qy, float qz,
                                             float gx, float gy, float gz, const WalkParams& P,
                                             float& ax, float& ay, float& az) {
  if constexpr (KIND == kWinding) {
    ax += wn_pair(L, qx, qy, qz, P);
  } else if constexpr (KIND == kPotential) {
    ax += potential_pair(L, qx, qy, qz, P);
  } else if constexpr (KIND == kSmoothDist) {
    ax += smooth_pair(L, qx, qy, qz, P);
  } else {
    contribution<KIND>(L.pos[0], L.pos[1], L.pos[2], L.mass, qx, qy, qz, gx, gy, gz, P, ax, ay,
                       az);
  }
}

// Exact sum of the point records [j0, j1) in index order; records are
// fetched 8 at a time so that their loads overlap.
template <int KIND>
__device__ __forceinline__ void leaf_sum(const Leaf3* __restrict__ leaf, uint32_t j0,
                                         uint32_t j1, float qx, float qy, float qz, float gx,
                                         float gy, float gz, const WalkParams& P, float& ax,
                                         float& ay, float& az) {
  uint32_t j = j0;
  for (; j + 8 <= j1; j += 8) {
    Leaf3 r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = leaf[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) contribution<KIND>(r[u], qx, qy, qz, gx, gy, gz, P, ax, ay, az);
  }
  for (; j < j1; ++j) contribution<KIND>(leaf[j], qx, qy, qz, gx, gy, gz, P, ax, ay, az);
}

// Host: walk parameters from the handler attributes (prune radius as
// src/sbh_ffi.cu:3233-3237: pruning only with a positive cutoff).
inline WalkParams make_walk_params(float beta, float softening, float cutoff, int prune_enabled,
                                   float prune_mult, int periodic, float box) {
  WalkParams P;
  P.beta = beta;
  float eps = fmaxf(softening, 0.0f);
  P.eps2 = eps * eps;
  P.rc = cutoff;
  P.cutoff = (cutoff > 0.0f && std::isfinite(cutoff)) ? 1 : 0;
  P.periodic = (periodic != 0 && box > 0.0f) ? 1 : 0;
  P.box = box;
  P.inv_box = box > 0.0f ? 1.0f / box : 0.0f;
  P.prune = (prune_enabled != 0 && cutoff > 0.0f) ? 1 : 0;
  float r_cut = P.prune ? prune_mult * cutoff : 0.0f;
  P.r_cut_sq = r_cut * r_cut;
  return P;
}

// Host: shared validation of the walk handlers (src/sbh_ffi.cu:3142-3159).
// `scalar_out`: the result is one value per query, shape (M,), instead of (M, 3).
inline int check_walk_common(const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                             const SbhBuffer* queries, const SbhBuffer* out, int threads,
                             bool scalar_out = false) {
  if (!buffer_ok(leaf_bytes, SBH_DTYPE_U8, -1) || !buffer_ok(node_bytes, SBH_DTYPE_U8, -1))
    return fail_invalid("leaf_bytes/node_bytes must be uint8 device buffers");
  if (byte_count(leaf_bytes) % static_cast<int64_t>(sizeof(Leaf3)) != 0)
    return fail_invalid("leaf_bytes size mismatch");
  if (byte_count(node_bytes) % static_cast<int64_t>(sizeof(Node8)) != 0 ||
      byte_count(node_bytes) == 0)
    return fail_invalid("node_bytes size mismatch");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2)) return fail_invalid("queries must be (M, 3) float32");
  if (queries->dims[1] != 3) return fail_invalid("queries dim must be 3");
  if (scalar_out ? (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != queries->dims[0])
                 : (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != queries->dims[0] ||
                    out->dims[1] != 3))
    return fail_invalid("out shape mismatch");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  return SBH_OK;
}

}  // namespace o8
}  // namespace sbh
