// Allocation-free variant of sbh_gravity_force_vjp: the tree workspace lives
// in a file-scope device array instead of a caller buffer or the library's
// cached hipMalloc, for callers that cannot allocate (code under stream
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array is `scratch`, kNoAllocScratchBytes bytes per device, held by the
// loaded code object until the process ends.  Host code never passes the
// symbol itself to a kernel: its device address is resolved once per device
// with hipGetSymbolAddress (with that device current) and kept in g_ptr.
// Every call on a device shares that device's array, so calls that overlap
// on one device (different streams or host threads) race; calls queued on
// one stream are ordered.  The pipeline clears what must start at zero (the
// tree header, the adjoint accumulators) with fill kernels on every call.
#include <mutex>

#include "common.hpp"
#include "tree.hpp"

using namespace sbh;

namespace {

constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (resolved on first use).
int scratch_ptr(char** out) {
  *out = nullptr;
  int dev = -1;
  SBH_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail_internal_msg("device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_internal(e);
    if (p == nullptr) return fail_internal_msg("scratch symbol has no device address");
    g_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_ptr[dev];
  return SBH_OK;
}

}  // namespace

extern "C" int64_t sbh_gravity_noalloc_scratch_bytes(void) {
  return static_cast<int64_t>(kNoAllocScratchBytes);
}

extern "C" int sbh_gravity_force_vjp_noalloc(void* stream, const SbhBuffer* points,
                                             const SbhBuffer* masses, const SbhBuffer* cotangent,
                                             SbhBuffer* forces, SbhBuffer* grad_points,
                                             SbhBuffer* grad_masses, float beta,
                                             float softening_scale, int32_t max_depth,
                                             int32_t leaf_size, int32_t multipole_order,
                                             SbhBuffer* stats) {
  clear_error();
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != points->dims[0])
    return fail_invalid("masses must have shape (N,)");
  if (!(beta >= 0.0f)) return fail_invalid("beta must be >= 0");
  if (softening_scale < 0.0f) return fail_invalid("softening_scale must be >= 0");
  if (max_depth < 1 || max_depth > 10) return fail_invalid("max_depth must be in [1, 10]");
  if (leaf_size < 1) return fail_invalid("leaf_size must be >= 1");
  if (multipole_order < 1 || multipole_order > 3)
    return fail_invalid("multipole_order must be 1, 2 or 3");
  const int64_t n = points->dims[0];
  if (n > (int64_t{1} << 30)) return fail_invalid("too many points");
  if (!buffer_ok(cotangent, SBH_DTYPE_F32, 2) || cotangent->dims[0] != n ||
      cotangent->dims[1] != 3)
    return fail_invalid("cotangent shape mismatch");
  if (forces != nullptr &&
      (!buffer_ok(forces, SBH_DTYPE_F32, 2) || forces->dims[0] != n || forces->dims[1] != 3))
    return fail_invalid("forces shape mismatch");
  if (!buffer_ok(grad_points, SBH_DTYPE_F32, 2) || grad_points->dims[0] != n ||
      grad_points->dims[1] != 3)
    return fail_invalid("grad_points shape mismatch");
  if (!buffer_ok(grad_masses, SBH_DTYPE_F32, -1) || elem_count(grad_masses) != n)
    return fail_invalid("grad_masses shape mismatch");
  if (stats != nullptr && (!buffer_ok(stats, SBH_DTYPE_S64, -1) || elem_count(stats) < 8))
    return fail_invalid("stats must be an int64 buffer of at least 8 elements");
  const tree::Layout L = tree::layout(n, max_depth, leaf_size);
  if (L.total > kNoAllocScratchBytes)
    return fail_invalid(
        "problem too large for the allocation-free scratch (sbh_gravity_workspace_bytes() "
        "must not exceed sbh_gravity_noalloc_scratch_bytes())");
  char* ws = nullptr;
  int rc = scratch_ptr(&ws);
  if (rc != SBH_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  rc = tree::build(s, ws, L, static_cast<const float*>(points->data),
                   static_cast<const float*>(masses->data),
                   static_cast<const float*>(cotangent->data), n, max_depth, leaf_size, nullptr);
  if (rc != SBH_OK) return rc;
  rc = tree::walk(s, ws, L, n, beta, softening_scale, true, multipole_order,
                  forces ? static_cast<float*>(forces->data) : nullptr,
                  static_cast<float*>(grad_points->data), static_cast<float*>(grad_masses->data),
                  stats != nullptr);
  if (rc != SBH_OK) return rc;
  if (stats != nullptr) return tree::copy_stats(s, ws, L, static_cast<int64_t*>(stats->data));
  return SBH_OK;
}
