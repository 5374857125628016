// Shared host/device definitions for libsbh_amd (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "sbh_amd.h"

namespace sbh {

// LeafData<N> (include/leaf_data.h:8-16): position, normal, mass, alpha;
// 24 bytes in 2-D, 32 in 3-D.
template <int DIM>
struct LeafT {
  float pos[DIM];
  float normal[DIM];
  float mass;
  float alpha;
};
using Leaf2 = LeafT<2>;
using Leaf3 = LeafT<3>;

// OctreeNode8 (src/sbh_ffi.cu:47-65): 104 bytes, same field order.
struct Node8 {
  uint32_t start, end, parent, first_child;
  uint32_t child[8];
  uint8_t child_mask, depth, is_leaf, pad0;
  float cx, cy, cz, half_size, diameter;
  Leaf3 contrib;
};

// Tree2DNode<1, LOG> / Tree3DNode<1, LOG> (include/tree2d.h:212-225,
// include/tree3d.h:212-225; for LOG = 3 also the reference's Node3Layout,
// src/sbh_ffi.cu:123-140): the same members in the same order, hence the same
// size and padding.  2^LOG cells per axis; cell b <-> x + S y (+ S^2 z),
// S = 2^LOG; the bit mask (BitMask<LOG, DIM>, include/bitmask.h:20-27) has
// max(1, S^DIM / 64) words.
template <int DIM, int LOG>
struct NodeT {
  static constexpr int kCells = 1 << (DIM * LOG);
  static constexpr int kWords = kCells >= 64 ? kCells / 64 : 1;
  uint64_t start, end, buffer_idx, parent_idx;
  float bmin[DIM], bmax[DIM];  // AABB<DIM>
  float diameter;
  LeafT<DIM> contrib;
  uint64_t bitmask[kWords];
  uint64_t child[kCells];      // buffer index of the node of each present cell
  int32_t num_children, depth, max_depth;
  uint8_t is_leaf, pad[3];
};
using Node2 = NodeT<2, 3>;  // 616 bytes
using Node3 = NodeT<3, 3>;  // 4272 bytes

static_assert(sizeof(Leaf2) == 24, "Leaf2 layout");
static_assert(sizeof(Node2) == 616, "Tree2DNode<1,3> layout");
static_assert(offsetof(Node2, bmin) == 32 && offsetof(Node2, diameter) == 48 &&
                  offsetof(Node2, contrib) == 52 && offsetof(Node2, bitmask) == 80 &&
                  offsetof(Node2, child) == 88 && offsetof(Node2, num_children) == 600 &&
                  offsetof(Node2, is_leaf) == 612,
              "Tree2DNode<1,3> layout");
static_assert(sizeof(Leaf3) == 32, "Leaf3 layout");
static_assert(sizeof(Node8) == 104, "OctreeNode8 layout");
static_assert(offsetof(Node8, child_mask) == 48, "OctreeNode8 layout");
static_assert(offsetof(Node8, cx) == 52, "OctreeNode8 layout");
static_assert(offsetof(Node8, contrib) == 72, "OctreeNode8 layout");
static_assert(sizeof(Node3) == 4272, "Tree3DNode<1,3> layout");
static_assert(sizeof(NodeT<2, 1>) == 136 && sizeof(NodeT<2, 2>) == 232 &&
                  sizeof(NodeT<3, 1>) == 184 && sizeof(NodeT<3, 2>) == 632,
              "Tree2DNode<1,1..2> / Tree3DNode<1,1..2> layouts");
static_assert(offsetof(Node3, bmin) == 32 && offsetof(Node3, diameter) == 56 &&
                  offsetof(Node3, contrib) == 60 && offsetof(Node3, bitmask) == 96 &&
                  offsetof(Node3, child) == 160 && offsetof(Node3, num_children) == 4256 &&
                  offsetof(Node3, is_leaf) == 4268,
              "Tree3DNode<1,3> layout");

// ---------------------------------------------------------------- errors
int fail_invalid(const char* msg);      // returns SBH_INVALID_ARGUMENT
int fail_internal(hipError_t err);      // returns SBH_INTERNAL
int fail_internal_msg(const char* msg); // returns SBH_INTERNAL
void clear_error();

#define SBH_HIP_TRY(expr)                         \
  do {                                            \
    hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return ::sbh::fail_internal(_e); \
  } while (0)

#define SBH_LAUNCH_CHECK() SBH_HIP_TRY(hipGetLastError())

// ---------------------------------------------------------------- buffers
inline bool has_dtype(const SbhBuffer* b, int32_t dtype) {
  return b != nullptr && b->dtype == dtype;
}
inline int64_t elem_count(const SbhBuffer* b) {
  int64_t c = 1;
  for (int i = 0; i < b->rank; i++) c *= b->dims[i];
  return c;
}
inline size_t elem_size(int32_t dtype) {
  switch (dtype) {
    case SBH_DTYPE_U8: return 1;
    case SBH_DTYPE_S32: return 4;
    case SBH_DTYPE_U32: return 4;
    case SBH_DTYPE_F32: return 4;
    case SBH_DTYPE_S64: return 8;
    default: return 0;
  }
}
inline int64_t byte_count(const SbhBuffer* b) {
  return elem_count(b) * static_cast<int64_t>(elem_size(b->dtype));
}
// Validates pointer/dtype/rank (rank < 0: any).  Returns false on mismatch.
inline bool buffer_ok(const SbhBuffer* b, int32_t dtype, int rank) {
  if (b == nullptr) return false;
  if (b->dtype != dtype) return false;
  if (rank >= 0 && b->rank != rank) return false;
  if (b->rank < 0 || b->rank > 4) return false;
  if (elem_count(b) > 0 && b->data == nullptr) return false;
  return true;
}

// ---------------------------------------------------------------- workspace
// Grow-only device scratch cached per (device, stream, slot).  Allocation
// happens only when a call needs more than any earlier call on that stream.
void* workspace(hipStream_t stream, int slot, size_t bytes);
// True while `stream` is being captured into a graph (no allocation allowed).
bool stream_capturing(hipStream_t stream);
// workspace() for entry points whose caller passed no buffer: refuses
// (SBH_INVALID_ARGUMENT) while `stream` is being captured, SBH_INTERNAL when
// the allocation fails.
int cached_workspace(hipStream_t stream, int slot, size_t bytes, char** out);
constexpr int kWsBuild = 0;
constexpr int kWsSort = 1;
constexpr int kWsWalk = 2;
constexpr int kWsPipeline = 3;
constexpr int kWsRope = 4;
constexpr int kWsSlots = 5;

inline unsigned grid_for(int64_t n, int block, unsigned cap = 0x7fffffffu) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// ---------------------------------------------------------------- radix sort
// Stable LSD sort of (key, value) pairs on the low `key_bits` bits of key.
// `temp` must hold radix_sort_temp_bytes(n).  Result lands in keys_out/vals_out
// (keys_in/vals_in are clobbered).
size_t radix_sort_temp_bytes(int64_t n);
// In-place exclusive prefix sum of n uint32 values (three launches: tile
// scans, a scan of the tile sums, the fix-up); `temp` holds
// exclusive_scan_temp_bytes(n).
size_t exclusive_scan_temp_bytes(int64_t n);
int exclusive_scan(hipStream_t stream, uint32_t* data, int64_t n, void* temp);
int radix_sort_pairs(hipStream_t stream, uint32_t* keys_in, uint32_t* vals_in,
                     uint32_t* keys_out, uint32_t* vals_out, int64_t n,
                     int key_bits, void* temp);
// Stable order of n rows by (set, code, row) in 2 + ceil(key_bits / 8) +
// set-byte launches (onesweep.hip): codes (n, input order, key_bits wide),
// set_offs the nsets + 1 row offsets of the sets (contiguous row ranges).
// keys_a/vals_a and keys_b/vals_b are ping-pong buffers of n entries; *keys
// and *perm point at the ones holding the sorted codes and rows.  temp holds
// morton_sort_temp_bytes(n, nsets).  A look-back that finds a missing
// predecessor (it cannot happen) sets bit 16 of *flag.
size_t morton_sort_temp_bytes(int64_t n, int nsets);
int morton_sort(hipStream_t stream, const uint32_t* codes, int64_t n, int key_bits,
                const int32_t* set_offs, int nsets, uint32_t* keys_a, uint32_t* vals_a,
                uint32_t* keys_b, uint32_t* vals_b, void* temp, int32_t* flag,
                const uint32_t** keys, const uint32_t** perm);

// ---------------------------------------------------------------- device math
// Min-image as include/kernel.h:27-39 (floor(x + 0.5)); not contracted so
// that MAC decisions that depend on it equal the IEEE restatement.
__device__ __forceinline__ float min_image_floor(float d, int periodic, float box,
                                                 float inv_box) {
#pragma clang fp contract(off)
  return periodic ? d - box * floorf(d * inv_box + 0.5f) : d;
}
// Min-image as the direct-sum pair functions (src/sbh_ffi.cu:440-445).
__device__ __forceinline__ float min_image_rint(float d, int periodic, float box,
                                                float inv_box) {
#pragma clang fp contract(off)
  return periodic ? d - box * rintf(d * inv_box) : d;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based RNG used
// by the stochastic walks in place of cuRAND XORWOW.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform in (0, 1] indexed by (query, node, sample, step): 24 random bits,
// so P(u > t) = 1 - t exactly for t on the 2^-24 grid (the reference draws
// curand_uniform, also in (0, 1], src/sbh_ffi.cu:2213, 2235).
__device__ __forceinline__ float sbh_uniform(uint32_t seed, uint32_t q, uint32_t node,
                                             uint32_t sample, uint32_t step) {
  const uint4 r = philox4x32_10(make_uint4(q, node, sample, step), seed, 0x5BD1E995u);
  return static_cast<float>((r.x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Morton quantiser (src/sbh_ffi.cu:173-179), IEEE fp32 without contraction.
__device__ __forceinline__ uint32_t quantize_axis(float v, float min_v, float span, int bits) {
#pragma clang fp contract(off)
  float t = (v - min_v) / span;
  t = fminf(fmaxf(t, 0.0f), 0.999999f);
  uint32_t max_bin = (1u << bits) - 1u;
  return static_cast<uint32_t>(t * static_cast<float>(max_bin + 1u));
}

// Bit interleave (src/sbh_ffi.cu:162-171): x -> bit 3i, y -> 3i+1, z -> 3i+2,
// for coordinates of up to 10 bits.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

constexpr float kTwoOverSqrtPi = 2.0f / 1.772453850905516f;
constexpr float kTwoSqrtPi = 2.0f * 1.772453850905516f;

// 1/sqrt(x) by the hardware instruction (about 1 ulp).  Callers pass
// r^2 + eps^2 + 1e-12 >= 1e-12, far from the denormal range that rsqrtf's
// extra compare-and-scale (three more VALU operations) handles.
__device__ __forceinline__ float rsq_softened(float x) { return __builtin_amdgcn_rsqf(x); }

// Short-range attenuation erfc(u) + 2u/sqrt(pi) exp(-u^2), u = r / (2 r_c)
// (src/sbh_ffi.cu:151-157).  Also returns exp(-u^2) and r.
__device__ __forceinline__ float attenuation(float r2, float rc, float* exp_u2,
                                             float* dist) {
  float r = sqrtf(r2 + 1e-12f);
  float u = r / (2.0f * rc);
  float e = expf(-(u * u));
  *exp_u2 = e;
  *dist = r;
  return erfcf(u) + kTwoOverSqrtPi * u * e;
}

}  // namespace sbh
