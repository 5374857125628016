// Wave-coherent group traversal of the adaptive octree, shared by the
// gravity walk (tree.hip) and the winding-number walk (wn.hip).
//
// One wavefront owns a group of up to 64 queries with bounding box
// [lo, hi].  A node is accepted as a multipole for the whole group iff the
// distance from its cell centre to the box is >= beta * cell diagonal (the
// reference's MAC, src/sbh_ffi.cu:1878-1893, made conservative for a group);
// rejected leaves are summed exactly.  Lanes test up to 64 stack entries at
// once; accepted nodes and near leaves are appended to per-wave LDS lists and
// handed to the caller's batch evaluators.
#pragma once

#include "common.hpp"
#include "tree.hpp"

namespace sbh {
namespace tree {
namespace gw {

constexpr int kB = 256;
constexpr int kWavesPerBlock = kB / 64;
// Run stack (traverse<true>): one entry per opened node, holding the run of
// its children.  An iteration pops whole entries from the top (at most 64
// nodes) and pushes at most 64 entries, each one level below the node that
// made it.  The entries left from one push lie below every later push, and
// each push is at least one level deeper than the one beneath it, so a tree
// of depth <= 10 holds at most 11 pushes of 64 entries (704) on the stack.
constexpr int kRunStackCap = 768;
// first-child words, the wave's 64 spread slots, then one count byte per entry
constexpr int kRunStackWords = kRunStackCap + 64 + kRunStackCap / 4;
// Node stack (traverse<false>): one entry per node; a pop takes at most
// (kStackSoft - top) / 7 nodes near the soft limit, so a push of <= 8 children
// each stays under kStackCap.
constexpr int kStackSoft = 1024;
constexpr int kStackCap = kStackSoft + 8 * 24;
constexpr int kFarCap = 128;
constexpr int kNearFlush = 16;
constexpr int kNearCap = kNearFlush + 64;
constexpr int kChunk = 8;  // nodes / sources evaluated per batch

// Wave-uniform reads through the scalar data cache: the arrays are not
// written while a walk runs, so they may be read as constant memory.
typedef __attribute__((address_space(4))) const float cfloat;
typedef __attribute__((address_space(4))) const int cint;

__device__ __forceinline__ float4 sload4(const float4* p, int i) {
  cfloat* q = (cfloat*)(p) + 4 * static_cast<int64_t>(i);  // generic -> constant space
  return make_float4(q[0], q[1], q[2], q[3]);
}
// Record u of a run starting at q (constant offsets from one base, so the
// loads of consecutive records can merge into wider scalar loads).
__device__ __forceinline__ float4 sload4_run(cfloat* q, int u) {
  return make_float4(q[4 * u], q[4 * u + 1], q[4 * u + 2], q[4 * u + 3]);
}
__device__ __forceinline__ cfloat* run_base(const float4* p, int i0) {
  return (cfloat*)(p) + 4 * static_cast<int64_t>(i0);
}
__device__ __forceinline__ int4 sload4(const int4* p, int i) {
  cint* q = (cint*)(p) + 4 * static_cast<int64_t>(i);
  return make_int4(q[0], q[1], q[2], q[3]);
}

// 1/sqrt(x) with the hardware instruction; x >= 1e-12 wherever it is used,
// far from the denormal range that rsqrtf's extra scaling handles.
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
// Square root by the hardware instruction (about 1 ulp); sqrtf wraps it in a
// correctly rounding sequence of about a dozen more VALU operations.
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ float wave_min(float v) {
  for (int d = 32; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}
// Inclusive prefix sum over the 64 lanes by DPP moves (no LDS round trips):
// Hillis-Steele inside each row of 16 (row_shr 1, 2, 4, 8; lanes whose
// source falls outside the row add 0), then the last lane of row 0 added to
// rows 1 and 3 (row_bcast:15) and the last lane of row 1 to rows 2 and 3
// (row_bcast:31).  The sums run in fp32, where the compiler fuses each move
// into its add; they are exact for non-negative totals below 2^24 (the
// traversal's child counts sum to at most 512).  All lanes must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_shr(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_bcast(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ int wave_incl_scan(int v) {
  float x = static_cast<float>(v);
  x += dpp_shr<0x111>(x);        // row_shr:1
  x += dpp_shr<0x112>(x);        // row_shr:2
  x += dpp_shr<0x114>(x);        // row_shr:4
  x += dpp_shr<0x118>(x);        // row_shr:8
  x += dpp_bcast<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_bcast<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return static_cast<int>(x);
}

// Number of lanes below this one whose bit is set in the wave mask m.
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(
      static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
}

// XCD-aware work queue of a persistent walk.  The groups (in Morton order)
// are split into 8 contiguous eighths, one per XCD; workgroups are dispatched
// to the 8 XCDs round-robin, so block b starts on its home eighth b mod 8 and
// each XCD's private L2 serves a compact region of the tree (measured: the
// forward walk 3% faster than with one global queue; chunks of 16 groups
// dealt round-robin to the XCDs gained half of that).  A wave that finds its
// current eighth exhausted moves on to the next one (work stealing); after 8
// exhausted eighths it gets -1.  `tries` is the calling wave's state (used by
// lane 0 only), starting at 0.  Each wave makes at most 8 failing claims, so
// every wave reaches the -1 exit.  With `order` (a permutation of the groups
// that keeps each eighth's groups inside it) claim t of an eighth takes
// group order[t] instead of group t.
__device__ __forceinline__ int next_group(Header* hdr, int num_groups, int lane, int& tries,
                                          const int* order = nullptr) {
  int g = -1;
  if (lane == 0) {
    const int home = static_cast<int>(blockIdx.x & 7u);
    while (tries < 8) {
      const int q = (home + tries) & 7;
      const int lo = static_cast<int>(static_cast<int64_t>(num_groups) * q / 8);
      const int hi = static_cast<int>(static_cast<int64_t>(num_groups) * (q + 1) / 8);
      const int t = lo + atomicAdd(&hdr->xq[q], 1);
      if (t < hi) {
        g = order != nullptr ? order[t] : t;
        break;
      }
      ++tries;
    }
  }
  return __builtin_amdgcn_readfirstlane(g);
}

// Eighth q of `num_groups` groups: [lo, hi), as next_group deals them.
__device__ __forceinline__ int2 eighth_range(int num_groups, int q) {
  return make_int2(static_cast<int>(static_cast<int64_t>(num_groups) * q / 8),
                   static_cast<int>(static_cast<int64_t>(num_groups) * (q + 1) / 8));
}

// Group MAC, fp32 without contraction (identical expression in the oracle).
__device__ __forceinline__ bool group_accept(float lx, float ly, float lz, float hx, float hy,
                                             float hz, float4 g, float beta) {
#pragma clang fp contract(off)
  float dx = fmaxf(fmaxf(lx - g.x, g.x - hx), 0.0f);
  float dy = fmaxf(fmaxf(ly - g.y, g.y - hy), 0.0f);
  float dz = fmaxf(fmaxf(lz - g.z, g.z - hz), 0.0f);
  float d2 = dx * dx + dy * dy + dz * dz;
  float t = beta * g.w;
  return d2 >= t * t;
}

struct Box {
  float lx, ly, lz, hx, hy, hz;
};

// Wave-wide bounding box of the lanes' positions (every lane must hold a
// position of the group; idle lanes repeat a valid one).
__device__ __forceinline__ Box wave_box(float x, float y, float z) {
  return Box{wave_min(x), wave_min(y), wave_min(z), wave_max(x), wave_max(y), wave_max(z)};
}

// Depth-first group traversal from `root`.  far[0..nfar) holds accepted
// non-empty nodes when flush_far(nfar) is called (at 64 entries and at the
// end); near[0..nnear) holds rejected non-empty leaves when
// flush_near(nnear) is called (at kNearFlush entries and at the end).  Both
// callbacks run on the whole wave.  A traversal whose stack would exceed
// its capacity stops early and sets bit 4 of *overflow (a header's overflow
// word, reported by the walk's stats), so a truncated sum never passes
// unnoticed.  RUNS selects the run stack (kRunStackWords ints per wave) or the
// node stack (kStackCap ints per wave); both visit the same nodes.
template <bool RUNS, class FarFn, class NearFn>
__device__ __forceinline__ void traverse(const Header* hdr, int* overflow, const int4* topo,
                                         const float4* geo, const float4* mp, float beta,
                                         const Box& b, int lane,
                                         int root, int* stack, int* far, int* near,
                                         FarFn&& flush_far, NearFn&& flush_near) {
  // run stack entry e: nodes stack[e] .. stack[e] + count[e] - 1 (count 1 to
  // 8, one byte); spread[l] marks the lane where a popped entry's first node
  // lands
  int* spread = stack + kRunStackCap;
  unsigned char* count = reinterpret_cast<unsigned char*>(spread + 64);
  const int cap = RUNS ? kRunStackCap : kStackCap;
  int top = 1, nfar = 0, nnear = 0;
  if (lane == 0) {
    stack[0] = root;
    if (RUNS) count[0] = 1;
  }
  if (RUNS) spread[lane] = 0;
  __builtin_amdgcn_wave_barrier();
  // Every node is pushed at most once per group and each iteration pops at
  // least one, so num_nodes iterations bound the loop even for a corrupt tree.
  const int num_nodes = hdr->num_nodes;
  int iters = 0;
  while (top > 0 && iters++ <= num_nodes) {
    int node;
    if constexpr (RUNS) {
      // lane r holds the r-th entry from the top; the whole entries whose
      // nodes fit in 64 lanes are popped (at least one: an entry holds <= 8)
      const int ne = min(64, top);
      const int ent = lane < ne ? stack[top - 1 - lane] : 0;
      const int cnt = lane < ne ? count[top - 1 - lane] : 0;
      const int incl = wave_incl_scan(cnt);
      const int npop = __popcll(__ballot(lane < ne && incl <= 64));
      if (npop == 0) {  // an entry of more than 64 nodes: not a tree of ours
        if (lane == 0) atomicOr(overflow, 4);
        break;
      }
      const int nodes = __builtin_amdgcn_readlane(incl, npop - 1);
      top -= npop;
      // lane l < nodes takes node l: entry r's nodes occupy lanes
      // [incl_r - cnt_r, incl_r); mark each popped entry's first lane
      if (lane < npop) spread[incl - cnt] = lane + 1;
      __builtin_amdgcn_wave_barrier();
      const bool first = spread[lane] != 0;
      spread[lane] = 0;
      const int r = lanes_below(__ballot(first)) + (first ? 1 : 0) - 1;
      const int er = __shfl(ent, max(r, 0), 64);
      const int sr = __shfl(incl - cnt, max(r, 0), 64);
      node = lane < nodes ? er + (lane - sr) : -1;
    } else {
      int k = min(64, top);
      k = min(k, max(1, (kStackSoft - top) / 7));
      const int from = top - k;
      node = lane < k ? stack[from + lane] : -1;
      top = from;
    }
    __builtin_amdgcn_wave_barrier();
    const bool have = node >= 0 && node < num_nodes;
    const int nd = have ? node : 0;
    const float4 m = mp[nd];
    const float4 g = geo[nd];
    const int4 t = topo[nd];
    int kind = 0, nch = 0, fch = 0;
    if (have && m.w != 0.0f) {
      if (group_accept(b.lx, b.ly, b.lz, b.hx, b.hy, b.hz, g, beta)) {
        kind = 1;
      } else if (t.w == 0) {
        kind = 2;
      } else {
        kind = 3;
        nch = t.w;
        fch = t.z;
      }
    }
    const uint64_t bf = __ballot(kind == 1);
    if (kind == 1) far[nfar + lanes_below(bf)] = node;
    nfar += __popcll(bf);
    const uint64_t bn = __ballot(kind == 2);
    if (kind == 2) near[nnear + lanes_below(bn)] = node;
    nnear += __popcll(bn);
    // children fch .. fch + nch - 1: one run entry per opened node, or one
    // entry per child
    int total, incl = 0;
    uint64_t bo = 0;
    if constexpr (RUNS) {
      bo = __ballot(kind == 3);
      total = __popcll(bo);
    } else {
      incl = wave_incl_scan(nch);
      total = __builtin_amdgcn_readlane(incl, 63);
    }
    if (top + total > cap) {  // unreachable for depth <= 10 (see the capacities)
      if (lane == 0) atomicOr(overflow, 4);
      break;
    }
    if constexpr (RUNS) {
      if (kind == 3) {
        const int e = top + lanes_below(bo);
        stack[e] = fch;
        count[e] = static_cast<unsigned char>(nch);
      }
    } else {
      int* dst = stack + (top + incl - nch);  // two per store (ds_write2)
      for (int c = 0; c < nch; c += 2) {
        dst[c] = fch + c;
        if (c + 1 < nch) dst[c + 1] = fch + c + 1;
      }
    }
    top += total;
    __builtin_amdgcn_wave_barrier();
    if (nfar >= 64) {
      flush_far(nfar);
      nfar = 0;
    }
    if (nnear >= kNearFlush) {
      flush_near(nnear);
      nnear = 0;
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (nfar > 0) flush_far(nfar);
  if (nnear > 0) flush_near(nnear);
}

}  // namespace gw
}  // namespace tree
}  // namespace sbh
