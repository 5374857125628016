// Stochastic Barnes-Hut (multi-level prefix control variate) over the dense
// octree8 records: forward and query-side VJP, unsoftened and softened.
//
// Estimator (src/sbh_ffi.cu:2146-2442): every subdomain (a node at depth 3,
// or a leaf above it) contributes its monopole (or its exact leaf sum); then
// each of S samples picks a point of the subdomain uniformly, walks down
// towards it with Russian roulette (continue with probability
// min(beta_cur / ratio, 1)), and adds (sum of children - parent monopole)
// weighted by n_sub / (P(path) * count(parent)).  Random numbers come from
// Philox4x32-10 indexed by (query, subdomain node, sample, step), so every
// sample path is reproducible independently of thread scheduling.
#include "common.hpp"
#include "o8walk.hpp"
#include "tree.hpp"

namespace sbh {
namespace {

using namespace o8;

constexpr int kB = 128;
constexpr int kControlDepth = 3;  // src/sbh_ffi.cu:2167

// PER: periodicity as a template argument (no minimum-image arithmetic in
// the non-periodic walk; same bits either way).
template <int KIND, bool PER>
__global__ __launch_bounds__(kB) void k_stoch8(const Leaf3* __restrict__ leaf,
                                               const Node8* __restrict__ nodes,
                                               const float* __restrict__ queries,
                                               const float* __restrict__ cot, int64_t nq,
                                               const uint32_t* __restrict__ order,
                                               float* __restrict__ out, WalkParams P,
                                               int samples, uint32_t seed, uint32_t num_nodes,
                                               uint32_t num_leaves) {
  P.periodic = PER ? 1 : 0;
  constexpr bool kSoft = (KIND == kSoftForce || KIND == kSoftVjp);
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (t >= nq) return;
  // thread t estimates query order[t] (Morton order); the random numbers are
  // keyed by the query index q, so the estimate does not depend on the thread
  const int64_t q = order != nullptr ? static_cast<int64_t>(order[t]) : t;
  const float qx = queries[3 * q], qy = queries[3 * q + 1], qz = queries[3 * q + 2];
  float gx = 0.f, gy = 0.f, gz = 0.f;
  if constexpr (KIND == kSoftVjp || KIND == kGravVjp) {
    gx = cot[3 * q];
    gy = cot[3 * q + 1];
    gz = cot[3 * q + 2];
  }
  const uint32_t qi = static_cast<uint32_t>(q);
  float tx = 0.f, ty = 0.f, tz = 0.f;
  // subdomain discovery: depth <= 3, at most 7 * 3 + 1 pending entries
  int stack[32];
  int top = 0;
  stack[top++] = 0;
  while (top > 0) {
    const int idx = stack[--top];
    const Node8* nd = &nodes[idx];
    if (nd->contrib.mass == 0.0f) continue;
    if (!(nd->is_leaf || nd->depth >= kControlDepth)) {
      for (int c = 0; c < 8; ++c)
        if ((nd->child_mask & (1u << c)) && nd->child[c] < num_nodes && top < 32)
          stack[top++] = static_cast<int>(nd->child[c]);
      continue;
    }
    const Node8* sub = nd;
    if (kSoft && should_prune(sub, qx, qy, qz, P)) continue;
    if (sub->is_leaf) {
      const uint32_t je = min(sub->end, num_leaves);
      for (uint32_t j = sub->start; j < je; ++j)
        contribution<KIND>(leaf[j], qx, qy, qz, gx, gy, gz, P, tx, ty, tz);
    } else {
      contribution<KIND>(sub->contrib, qx, qy, qz, gx, gy, gz, P, tx, ty, tz);
    }
    const uint32_t num_el = sub->end - sub->start;
    if (num_el == 0 || sub->end > num_leaves) continue;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int ss = 0; ss < samples; ++ss) {
      const float u = sbh_uniform(seed, qi, static_cast<uint32_t>(idx), ss, 0);
      const uint32_t sidx =
          min(static_cast<uint32_t>(static_cast<float>(num_el) * u) + sub->start, sub->end - 1);
      const float px = leaf[sidx].pos[0], py = leaf[sidx].pos[1], pz = leaf[sidx].pos[2];
      const Node8* prev = sub;
      const Node8* cur = sub;
      float event_prob = 1.0f;
      float wx = 0.f, wy = 0.f, wz = 0.f;
      float cur_beta;
      {
#pragma clang fp contract(off)
        cur_beta = fmaxf(P.beta, centre_dist(cur, qx, qy, qz, P) / cur->diameter);
      }
      uint32_t step = 1;
      while (!cur->is_leaf && step < 64) {
        if (kSoft && should_prune(cur, qx, qy, qz, P)) break;
        float ratio, tp;
        {
#pragma clang fp contract(off)
          ratio = centre_dist(cur, qx, qy, qz, P) / cur->diameter;
          tp = fminf(cur_beta / ratio, 1.0f);
        }
        const float pu = sbh_uniform(seed, qi, static_cast<uint32_t>(idx), ss, step++);
        if (pu > tp) break;
        event_prob *= tp;
        prev = cur;
        const uint32_t slot = child_slot(cur, px, py, pz);
        if (!(cur->child_mask & (1u << slot)) || cur->child[slot] >= num_nodes) break;
        cur = &nodes[cur->child[slot]];
        cur_beta = fmaxf(cur_beta, ratio);
        float prx = 0.f, pry = 0.f, prz = 0.f;
        contribution<KIND>(prev->contrib, qx, qy, qz, gx, gy, gz, P, prx, pry, prz);
        float crx = 0.f, cry = 0.f, crz = 0.f;
        for (int cc = 0; cc < 8; ++cc) {
          if (!(prev->child_mask & (1u << cc)) || prev->child[cc] >= num_nodes) continue;
          const Node8* ch = &nodes[prev->child[cc]];
          if (kSoft && should_prune(ch, qx, qy, qz, P)) continue;
          if (ch->is_leaf) {
            const uint32_t je = min(ch->end, num_leaves);
            for (uint32_t j = ch->start; j < je; ++j)
              contribution<KIND>(leaf[j], qx, qy, qz, gx, gy, gz, P, crx, cry, crz);
          } else {
            contribution<KIND>(ch->contrib, qx, qy, qz, gx, gy, gz, P, crx, cry, crz);
          }
        }
        const float prob_inv = static_cast<float>(num_el) /
                               (event_prob * static_cast<float>(prev->end - prev->start));
        wx += prob_inv * (crx - prx);
        wy += prob_inv * (cry - pry);
        wz += prob_inv * (crz - prz);
      }
      sx += wx;
      sy += wy;
      sz += wz;
    }
    if (samples > 0) {
      const float inv = 1.0f / static_cast<float>(samples);
      tx += sx * inv;
      ty += sy * inv;
      tz += sz * inv;
    }
  }
  if constexpr (kScalarKind<KIND>) {
    out[q] = tx;
  } else {
    out[3 * q + 0] = tx;
    out[3 * q + 1] = ty;
    out[3 * q + 2] = tz;
  }
}

template <int KIND>
int launch(void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
           const SbhBuffer* queries, const SbhBuffer* cot, SbhBuffer* out, const WalkParams& P,
           int samples, int seed) {
  const int64_t nq = queries->dims[0];
  if (nq == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t* order = nullptr;
  // (not while the stream is captured: the ordering scratch may need an
  // allocation, and the unordered walk gives the same results)
  if (nq >= kOrderMin && !stream_capturing(s)) {
    char* ws = static_cast<char*>(workspace(s, kWsWalk, tree::query_order_bytes(nq)));
    if (ws == nullptr) return fail_internal_msg("workspace allocation failed");
    int rc = tree::query_order(s, static_cast<const float*>(queries->data), nq, ws, &order);
    if (rc != SBH_OK) return rc;
  }
  (P.periodic ? k_stoch8<KIND, true> : k_stoch8<KIND, false>)<<<grid_for(nq, kB), kB, 0, s>>>(
      static_cast<const Leaf3*>(leaf_bytes->data), static_cast<const Node8*>(node_bytes->data),
      static_cast<const float*>(queries->data),
      cot ? static_cast<const float*>(cot->data) : nullptr, nq, order,
      static_cast<float*>(out->data),
      P, samples, static_cast<uint32_t>(seed),
      static_cast<uint32_t>(byte_count(node_bytes) / sizeof(Node8)),
      static_cast<uint32_t>(byte_count(leaf_bytes) / sizeof(Leaf3)));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

int check_cot(const SbhBuffer* queries, const SbhBuffer* cot) {
  if (!buffer_ok(cot, SBH_DTYPE_F32, 2) || cot->dims[0] != queries->dims[0] ||
      cot->dims[1] != 3)
    return fail_invalid("cotangent shape mismatch");
  return SBH_OK;
}

int check_prune(int prune_enabled, float mult) {
  if (prune_enabled != 0 && mult <= 0.0f)
    return fail_invalid("prune_r_cut_mult must be > 0 when prune_enabled is set");
  return SBH_OK;
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" int sbh_stochastic_bh_force_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                               const SbhBuffer* node_bytes,
                                               const SbhBuffer* queries, SbhBuffer* out,
                                               int32_t samples_per_subdomain, int32_t seed,
                                               float beta, int32_t periodic, float box_length,
                                               int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, 0.f, 0.f, 0, 0.f, periodic, box_length);
  return launch<o8::kGravGrad>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P,
                               samples_per_subdomain, seed);
}

extern "C" int sbh_stochastic_bh_force_octree8_soft(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, SbhBuffer* out, int32_t samples_per_subdomain, int32_t seed,
    float beta, float softening_scale, float cutoff_scale, int32_t prune_enabled,
    float prune_r_cut_mult, int32_t periodic, float box_length, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK || (rc = check_prune(prune_enabled, prune_r_cut_mult)) != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, softening_scale, cutoff_scale, prune_enabled,
                                          prune_r_cut_mult, periodic, box_length);
  return launch<o8::kSoftForce>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P,
                                samples_per_subdomain, seed);
}

extern "C" int sbh_stochastic_bh_force_octree8_vjp(void* stream, const SbhBuffer* leaf_bytes,
                                                   const SbhBuffer* node_bytes,
                                                   const SbhBuffer* queries,
                                                   const SbhBuffer* cotangent, SbhBuffer* out,
                                                   int32_t samples_per_subdomain, int32_t seed,
                                                   float beta, int32_t periodic,
                                                   float box_length, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK || (rc = check_cot(queries, cotangent)) != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, 0.f, 0.f, 0, 0.f, periodic, box_length);
  return launch<o8::kGravVjp>(stream, leaf_bytes, node_bytes, queries, cotangent, out, P,
                              samples_per_subdomain, seed);
}

extern "C" int sbh_stochastic_bh_force_octree8_soft_vjp(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, const SbhBuffer* cotangent, SbhBuffer* out,
    int32_t samples_per_subdomain, int32_t seed, float beta, float softening_scale,
    float cutoff_scale, int32_t prune_enabled, float prune_r_cut_mult, int32_t periodic,
    float box_length, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads);
  if (rc != SBH_OK || (rc = check_cot(queries, cotangent)) != SBH_OK ||
      (rc = check_prune(prune_enabled, prune_r_cut_mult)) != SBH_OK)
    return rc;
  o8::WalkParams P = o8::make_walk_params(beta, softening_scale, cutoff_scale, prune_enabled,
                                          prune_r_cut_mult, periodic, box_length);
  return launch<o8::kSoftVjp>(stream, leaf_bytes, node_bytes, queries, cotangent, out, P,
                              samples_per_subdomain, seed);
}

// Stochastic winding numbers (MLPCV of src/sbh_gpu.cu:166-187 with
// signedSolidAngle<3>) over octree8 records of oriented area elements.
extern "C" int sbh_stochastic_bh_wn_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                            const SbhBuffer* node_bytes,
                                            const SbhBuffer* queries, SbhBuffer* out,
                                            int32_t samples_per_subdomain, int32_t seed,
                                            float beta, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, 0.f, 0.f, 0, 0.f, 0, 1.0f);
  return launch<o8::kWinding>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P,
                              samples_per_subdomain, seed);
}

// Stochastic potential and smooth-distance sums (MLPCV of src/sbh_gpu.cu:
// 143-164, 436-472 and 189-211, 512-550 with gravityPotential and
// smoothDistExp) over octree8 records.  out: (M,).
extern "C" int sbh_stochastic_bh_potential_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                                   const SbhBuffer* node_bytes,
                                                   const SbhBuffer* queries, SbhBuffer* out,
                                                   int32_t samples_per_subdomain, int32_t seed,
                                                   float beta, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, 0.f, 0.f, 0, 0.f, 0, 1.0f);
  return launch<o8::kPotential>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P,
                                samples_per_subdomain, seed);
}

extern "C" int sbh_stochastic_bh_smoothdist_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                                    const SbhBuffer* node_bytes,
                                                    const SbhBuffer* queries, SbhBuffer* out,
                                                    int32_t samples_per_subdomain, int32_t seed,
                                                    float beta, int32_t threads) {
  clear_error();
  int rc = o8::check_walk_common(leaf_bytes, node_bytes, queries, out, threads, true);
  if (rc != SBH_OK) return rc;
  o8::WalkParams P = o8::make_walk_params(beta, 0.f, 0.f, 0, 0.f, 0, 1.0f);
  return launch<o8::kSmoothDist>(stream, leaf_bytes, node_bytes, queries, nullptr, out, P,
                                 samples_per_subdomain, seed);
}
