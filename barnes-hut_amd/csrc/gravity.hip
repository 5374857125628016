// C-ABI entry points of the adaptive-tree gravity pipeline (new targets: the
// reference has no adaptive tree, no leaf-size parameter and no full adjoint).
#include <mutex>
#include <vector>

#include "common.hpp"
#include "profile.hpp"
#include "tree.hpp"

using namespace sbh;

namespace sbh {
namespace prof {
namespace {
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
}  // namespace

bool begin(hipStream_t s, Rec* r, int kind) {
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_on) return false;
  }
  for (int k = 0; k < 4; ++k) r->ev[k] = nullptr;
  r->kind = kind;
  for (int k = 0; k < 4; ++k) {
    if (hipEventCreate(&r->ev[k]) != hipSuccess) {
      r->ev[k] = nullptr;
      destroy(*r);
      return false;
    }
  }
  if (hipEventRecord(r->ev[0], s) != hipSuccess) {
    destroy(*r);
    return false;
  }
  return true;
}

void end(const Rec& r) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_recs.push_back(r);
}

void destroy(Rec& r) {
  for (int k = 0; k < 4; ++k) {
    if (r.ev[k] != nullptr) (void)hipEventDestroy(r.ev[k]);
    r.ev[k] = nullptr;
  }
}

}  // namespace prof
}  // namespace sbh

namespace {

// order: 1 or 2; the VJP entry points also take 3 (monopole terms with the
// quadrupole term's second-moment adjoint), passed as max_order.
int check_common(const SbhBuffer* points, const SbhBuffer* masses, float beta, float softening,
                 int max_depth, int leaf_size, int order, int max_order = 2) {
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  if (!buffer_ok(masses, SBH_DTYPE_F32, 1) || masses->dims[0] != points->dims[0])
    return fail_invalid("masses must have shape (N,)");
  if (points->dims[0] > (int64_t{1} << 30)) return fail_invalid("too many points");
  if (!(beta >= 0.0f)) return fail_invalid("beta must be >= 0");
  if (softening < 0.0f) return fail_invalid("softening_scale must be >= 0");
  if (max_depth < 1 || max_depth > 10) return fail_invalid("max_depth must be in [1, 10]");
  if (leaf_size < 1) return fail_invalid("leaf_size must be >= 1");
  if (order < 1 || order > max_order)
    return fail_invalid(max_order == 2 ? "multipole_order must be 1 or 2"
                                       : "multipole_order must be 1, 2 or 3");
  return SBH_OK;
}

int resolve_workspace(hipStream_t s, SbhBuffer* workspace, const tree::Layout& L, char** ws) {
  if (workspace != nullptr) {
    if (!buffer_ok(workspace, SBH_DTYPE_U8, -1) ||
        byte_count(workspace) < static_cast<int64_t>(L.total))
      return fail_invalid("workspace too small");
    *ws = static_cast<char*>(workspace->data);
    return SBH_OK;
  }
  return cached_workspace(s, kWsPipeline, L.total, ws);
}

int check_stats(const SbhBuffer* stats) {
  if (stats == nullptr) return SBH_OK;
  if (!buffer_ok(stats, SBH_DTYPE_S64, -1) || elem_count(stats) < 8)
    return fail_invalid("stats must be an int64 buffer of at least 8 elements");
  return SBH_OK;
}

}  // namespace

extern "C" int64_t sbh_gravity_workspace_bytes(int64_t n, int32_t max_depth, int32_t leaf_size) {
  if (n < 0 || max_depth < 1 || max_depth > 10 || leaf_size < 1) return -1;
  return static_cast<int64_t>(tree::layout(n, max_depth, leaf_size).total);
}

extern "C" int sbh_gravity_workspace_layout(int64_t n, int32_t max_depth, int32_t leaf_size,
                                            int64_t* offsets, int32_t count) {
  clear_error();
  if (n < 0 || max_depth < 1 || max_depth > 10 || leaf_size < 1 || offsets == nullptr)
    return fail_invalid("invalid layout query");
  tree::Layout L = tree::layout(n, max_depth, leaf_size);
  const int64_t v[] = {static_cast<int64_t>(L.header), static_cast<int64_t>(L.codes),
                       static_cast<int64_t>(L.perm),   static_cast<int64_t>(L.sp),
                       static_cast<int64_t>(L.sc),     static_cast<int64_t>(L.topo),
                       static_cast<int64_t>(L.prefix), static_cast<int64_t>(L.parent),
                       static_cast<int64_t>(L.geo),    static_cast<int64_t>(L.mp),
                       static_cast<int64_t>(L.qd),     static_cast<int64_t>(L.adj),
                       static_cast<int64_t>(L.adjq),   static_cast<int64_t>(L.qgrad),
                       static_cast<int64_t>(L.sgrad),  static_cast<int64_t>(L.total),
                       L.cap};
  const int nv = static_cast<int>(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < count && i < nv; ++i) offsets[i] = v[i];
  return SBH_OK;
}

extern "C" int sbh_gravity_force(void* stream, const SbhBuffer* points, const SbhBuffer* masses,
                                 SbhBuffer* forces, float beta, float softening_scale,
                                 int32_t max_depth, int32_t leaf_size, int32_t multipole_order,
                                 SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = check_common(points, masses, beta, softening_scale, max_depth, leaf_size,
                        multipole_order);
  if (rc != SBH_OK) return rc;
  int64_t n = points->dims[0];
  if (!buffer_ok(forces, SBH_DTYPE_F32, 2) || forces->dims[0] != n || forces->dims[1] != 3)
    return fail_invalid("forces shape mismatch");
  if ((rc = check_stats(stats)) != SBH_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  tree::Layout L = tree::layout(n, max_depth, leaf_size);
  char* ws = nullptr;
  if ((rc = resolve_workspace(s, workspace, L, &ws)) != SBH_OK) return rc;
  prof::Scope pr(s, 0);
  rc = tree::build(s, ws, L, static_cast<const float*>(points->data),
                   static_cast<const float*>(masses->data), nullptr, n, max_depth, leaf_size,
                   nullptr);
  if (rc != SBH_OK) return rc;
  if (pr.on()) SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
  rc = tree::walk(s, ws, L, n, beta, softening_scale, false, multipole_order,
                  static_cast<float*>(forces->data), nullptr, nullptr, stats != nullptr,
                  pr.on() && n > 0 ? pr.rec().ev + 2 : nullptr);
  if (rc != SBH_OK) return rc;
  if (n > 0) pr.commit();
  if (stats != nullptr) return tree::copy_stats(s, ws, L, static_cast<int64_t*>(stats->data));
  return SBH_OK;
}

extern "C" int sbh_gravity_force_vjp(void* stream, const SbhBuffer* points,
                                     const SbhBuffer* masses, const SbhBuffer* cotangent,
                                     SbhBuffer* forces, SbhBuffer* grad_points,
                                     SbhBuffer* grad_masses, float beta, float softening_scale,
                                     int32_t max_depth, int32_t leaf_size,
                                     int32_t multipole_order, SbhBuffer* workspace,
                                     SbhBuffer* stats) {
  clear_error();
  int rc = check_common(points, masses, beta, softening_scale, max_depth, leaf_size,
                        multipole_order, 3);
  if (rc != SBH_OK) return rc;
  int64_t n = points->dims[0];
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
  if ((rc = check_stats(stats)) != SBH_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  tree::Layout L = tree::layout(n, max_depth, leaf_size);
  char* ws = nullptr;
  if ((rc = resolve_workspace(s, workspace, L, &ws)) != SBH_OK) return rc;
  prof::Scope pr(s, 1);
  rc = tree::build(s, ws, L, static_cast<const float*>(points->data),
                   static_cast<const float*>(masses->data),
                   static_cast<const float*>(cotangent->data), n, max_depth, leaf_size, nullptr);
  if (rc != SBH_OK) return rc;
  if (pr.on()) SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
  rc = tree::walk(s, ws, L, n, beta, softening_scale, true, multipole_order,
                  forces ? static_cast<float*>(forces->data) : nullptr,
                  static_cast<float*>(grad_points->data), static_cast<float*>(grad_masses->data),
                  stats != nullptr, pr.on() && n > 0 ? pr.rec().ev + 2 : nullptr);
  if (rc != SBH_OK) return rc;
  if (n > 0) pr.commit();
  if (stats != nullptr) return tree::copy_stats(s, ws, L, static_cast<int64_t*>(stats->data));
  return SBH_OK;
}

extern "C" void sbh_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lock(prof::g_mu);
  prof::g_on = on != 0;
}

// out[3 k .. 3 k + 2] = {calls, build ms, walk-kernel ms} of kind k (0
// gravity forward, 1 gravity adjoint, 2 winding numbers, 3 stochastic walks)
// summed over the calls recorded since the last collect.  Waits for the
// recorded events.
extern "C" int sbh_profile_collect(double* out, int32_t count) {
  clear_error();
  std::vector<prof::Rec> recs;
  {
    std::lock_guard<std::mutex> lock(prof::g_mu);
    recs.swap(prof::g_recs);
  }
  double acc[3 * prof::kKinds] = {};
  int rc = SBH_OK;
  for (auto& r : recs) {
    float b = 0.f, w = 0.f;
    if (hipEventSynchronize(r.ev[3]) != hipSuccess ||
        hipEventElapsedTime(&b, r.ev[0], r.ev[1]) != hipSuccess ||
        hipEventElapsedTime(&w, r.ev[2], r.ev[3]) != hipSuccess)
      rc = fail_internal_msg("profile event query failed");
    const int o = 3 * r.kind;
    acc[o] += 1.0;
    acc[o + 1] += b;
    acc[o + 2] += w;
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(r.ev[k]);
  }
  for (int i = 0; i < count && i < 3 * prof::kKinds; ++i) out[i] = acc[i];
  return rc;
}

// ---------------------------------------------------------------- batches
namespace {

bool walk_threads_ok(int t) { return t == 0 || t == 64 || t == 128 || t == 256 || t == 512; }

int check_batch(const SbhBuffer* points, const SbhBuffer* set_offsets, int* nsets) {
  if (!buffer_ok(set_offsets, SBH_DTYPE_S32, 1) || set_offsets->dims[0] < 2)
    return fail_invalid("set_offsets must be an int32 vector of num_sets + 1 entries");
  const int64_t s = set_offsets->dims[0] - 1;
  if (s > (int64_t{1} << 20)) return fail_invalid("too many sets");
  if (points->dims[0] > INT32_MAX) return fail_invalid("too many points");
  *nsets = static_cast<int>(s);
  return SBH_OK;
}

}  // namespace

extern "C" int64_t sbh_gravity_batch_workspace_bytes(int64_t total_points, int32_t num_sets,
                                                     int32_t max_depth, int32_t leaf_size) {
  if (total_points < 0 || num_sets < 1 || max_depth < 1 || max_depth > 10 || leaf_size < 1)
    return -1;
  return static_cast<int64_t>(tree::layout(total_points, max_depth, leaf_size, num_sets).total);
}

extern "C" int sbh_gravity_batch_force(void* stream, const SbhBuffer* points,
                                       const SbhBuffer* masses, const SbhBuffer* set_offsets,
                                       SbhBuffer* forces, float beta, float softening_scale,
                                       int32_t max_depth, int32_t leaf_size,
                                       int32_t multipole_order, int32_t walk_blocks,
                                       int32_t walk_threads, SbhBuffer* workspace,
                                       SbhBuffer* stats) {
  clear_error();
  int rc = check_common(points, masses, beta, softening_scale, max_depth, leaf_size,
                        multipole_order);
  if (rc != SBH_OK) return rc;
  int nsets = 0;
  if ((rc = check_batch(points, set_offsets, &nsets)) != SBH_OK) return rc;
  const int64_t n = points->dims[0];
  if (!buffer_ok(forces, SBH_DTYPE_F32, 2) || forces->dims[0] != n || forces->dims[1] != 3)
    return fail_invalid("forces shape mismatch");
  if ((rc = check_stats(stats)) != SBH_OK) return rc;
  if (walk_blocks < 0) return fail_invalid("walk_blocks must be >= 0");
  if (!walk_threads_ok(walk_threads))
    return fail_invalid("walk_threads must be 0, 64, 128, 256 or 512");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tree::Layout L = tree::layout(n, max_depth, leaf_size, nsets);
  char* ws = nullptr;
  if ((rc = resolve_workspace(s, workspace, L, &ws)) != SBH_OK) return rc;
  prof::Scope pr(s, 0);
  rc = tree::build_forest(s, ws, L, static_cast<const float*>(points->data),
                          static_cast<const float*>(masses->data), nullptr, n,
                          static_cast<const int32_t*>(set_offsets->data), max_depth, leaf_size);
  if (rc != SBH_OK) return rc;
  if (pr.on()) SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
  rc = tree::walk(s, ws, L, n, beta, softening_scale, false, multipole_order,
                  static_cast<float*>(forces->data), nullptr, nullptr, stats != nullptr,
                  pr.on() && n > 0 ? pr.rec().ev + 2 : nullptr, walk_blocks, walk_threads);
  if (rc != SBH_OK) return rc;
  if (n > 0) pr.commit();
  if (stats != nullptr) return tree::copy_stats(s, ws, L, static_cast<int64_t*>(stats->data));
  return SBH_OK;
}

extern "C" int sbh_gravity_batch_force_vjp(void* stream, const SbhBuffer* points,
                                           const SbhBuffer* masses, const SbhBuffer* cotangent,
                                           const SbhBuffer* set_offsets, SbhBuffer* forces,
                                           SbhBuffer* grad_points, SbhBuffer* grad_masses,
                                           float beta, float softening_scale, int32_t max_depth,
                                           int32_t leaf_size, int32_t multipole_order,
                                           int32_t walk_blocks, int32_t walk_threads,
                                           SbhBuffer* workspace, SbhBuffer* stats) {
  clear_error();
  int rc = check_common(points, masses, beta, softening_scale, max_depth, leaf_size,
                        multipole_order, 3);
  if (rc != SBH_OK) return rc;
  int nsets = 0;
  if ((rc = check_batch(points, set_offsets, &nsets)) != SBH_OK) return rc;
  const int64_t n = points->dims[0];
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
  if ((rc = check_stats(stats)) != SBH_OK) return rc;
  if (walk_blocks < 0) return fail_invalid("walk_blocks must be >= 0");
  if (!walk_threads_ok(walk_threads))
    return fail_invalid("walk_threads must be 0, 64, 128, 256 or 512");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tree::Layout L = tree::layout(n, max_depth, leaf_size, nsets);
  char* ws = nullptr;
  if ((rc = resolve_workspace(s, workspace, L, &ws)) != SBH_OK) return rc;
  prof::Scope pr(s, 1);
  rc = tree::build_forest(s, ws, L, static_cast<const float*>(points->data),
                          static_cast<const float*>(masses->data),
                          static_cast<const float*>(cotangent->data), n,
                          static_cast<const int32_t*>(set_offsets->data), max_depth, leaf_size);
  if (rc != SBH_OK) return rc;
  if (pr.on()) SBH_HIP_TRY(hipEventRecord(pr.rec().ev[1], s));
  rc = tree::walk(s, ws, L, n, beta, softening_scale, true, multipole_order,
                  forces ? static_cast<float*>(forces->data) : nullptr,
                  static_cast<float*>(grad_points->data), static_cast<float*>(grad_masses->data),
                  stats != nullptr, pr.on() && n > 0 ? pr.rec().ev + 2 : nullptr, walk_blocks, walk_threads);
  if (rc != SBH_OK) return rc;
  if (n > 0) pr.commit();
  if (stats != nullptr) return tree::copy_stats(s, ws, L, static_cast<int64_t*>(stats->data));
  return SBH_OK;
}

// ---------------------------------------------------------------- graphs
// Stream capture of any sequence of library calls into a HIP graph.  Calls
// captured this way must use caller workspaces (no library-owned buffer may
// need to grow while capturing).
extern "C" int sbh_graph_begin(void* stream) {
  clear_error();
  SBH_HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                    hipStreamCaptureModeThreadLocal));
  return SBH_OK;
}

extern "C" int sbh_graph_end(void* stream, void** graph_exec) {
  clear_error();
  if (graph_exec == nullptr) return fail_invalid("graph_exec must not be NULL");
  hipGraph_t g = nullptr;
  SBH_HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail_internal(e);
  *graph_exec = ex;
  return SBH_OK;
}

extern "C" int sbh_graph_launch(void* graph_exec, void* stream) {
  clear_error();
  if (graph_exec == nullptr) return fail_invalid("graph_exec must not be NULL");
  SBH_HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                             static_cast<hipStream_t>(stream)));
  return SBH_OK;
}

extern "C" int sbh_graph_destroy(void* graph_exec) {
  clear_error();
  if (graph_exec != nullptr) SBH_HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return SBH_OK;
}
