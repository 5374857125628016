/*
 * sbh_oracle.c -- CPU restatement of the reference Barnes-Hut path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker (or, in bench.py, as the timed CPU baseline).
 *
 * Every function cites the reference lines it restates.  Reference files are
 * cited relative to the reference repository root.
 *
 * Arithmetic conventions
 *   - Morton quantisation, bounds, node geometry and MAC decisions are
 *     evaluated in IEEE fp32 with the reference's operation order (this file is
 *     compiled with -ffp-contract=off; the one contraction the reference's
 *     compiler performs by default -- the node-centre a+b*c -- is written as an
 *     explicit fmaf).
 *   - Sums (moments, forces, gradients) are accumulated in fp64: the reference
 *     accumulates in fp32 with atomics in an undefined order, so only a
 *     tolerance comparison is meaningful there.
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py)
 *   orc_quantize_axis / orc_morton3d      hand-computed codes + integer brute force
 *   orc_build_octree8                     invariants + independent numpy build (tiny N)
 *   orc_direct_*                          closed forms, Newton's third law,
 *                                         potential gradient, central differences
 *   orc_bh_octree8_*                      beta->inf equals direct sum, beta=0 equals
 *                                         root monopole, prune equivalence
 *   orc_tree_*                            topology invariants, beta->inf == direct,
 *                                         quadrupole error order, FD of the adjoint
 *   orc_philox4x32                        Random123 known-answer vectors
 *   orc_tree_stoch                        zero samples == sum of subdomain terms;
 *                                         every path forced to the leaves (beta
 *                                         huge) == exact sum; seed mean == exact
 *                                         sum (z-test); huge beta VJP == FD of
 *                                         the forward on the same paths
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef __cplusplus
#error "compile as C"
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Byte layouts (include/leaf_data.h:8-16; src/sbh_ffi.cu:47-65)              */
/* ------------------------------------------------------------------------- */
typedef struct {
  float pos[3];
  float normal[3];
  float mass;
  float alpha;
} OrcLeaf3; /* 32 bytes */

typedef struct {
  uint32_t start, end, parent, first_child;
  uint32_t child[8];
  uint8_t child_mask, depth, is_leaf, pad0;
  float cx, cy, cz, half_size, diameter;
  OrcLeaf3 contrib;
} OrcNode8; /* 104 bytes */

_Static_assert(sizeof(OrcLeaf3) == 32, "Leaf3 layout");
_Static_assert(sizeof(OrcNode8) == 104, "OctreeNode8 layout");

ORC_EXPORT size_t orc_leaf3_size(void) { return sizeof(OrcLeaf3); }
ORC_EXPORT size_t orc_node8_size(void) { return sizeof(OrcNode8); }

/* ------------------------------------------------------------------------- */
/* B1: quantisation and Morton interleave (src/sbh_ffi.cu:162-179)            */
/* ------------------------------------------------------------------------- */
ORC_EXPORT uint32_t orc_quantize_axis(float v, float min_v, float span, int bits) {
  float t = (v - min_v) / span;
  t = fminf(fmaxf(t, 0.0f), 0.999999f);
  uint32_t max_bin = (1u << bits) - 1u;
  return (uint32_t)(t * (float)(max_bin + 1u));
}

ORC_EXPORT uint32_t orc_morton3d(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t code = 0;
  for (int i = 0; i < bits; i++) {
    code |= ((x >> i) & 1u) << (3 * i);
    code |= ((y >> i) & 1u) << (3 * i + 1);
    code |= ((z >> i) & 1u) << (3 * i + 2);
  }
  return code;
}

ORC_EXPORT void orc_morton_codes(const float* pts, int64_t n, const float* min_corner,
                                 float span, int bits, uint32_t* codes) {
  for (int64_t i = 0; i < n; i++) {
    uint32_t qx = orc_quantize_axis(pts[3 * i + 0], min_corner[0], span, bits);
    uint32_t qy = orc_quantize_axis(pts[3 * i + 1], min_corner[1], span, bits);
    uint32_t qz = orc_quantize_axis(pts[3 * i + 2], min_corner[2], span, bits);
    codes[i] = orc_morton3d(qx, qy, qz, bits);
  }
}

/* B2: default bounds (python/jax_sbh_ffi.py:879-884: min-1e-3 / max+1e-3 in
 * fp32) and the cube span (src/sbh_ffi.cu:1128-1130). */
ORC_EXPORT void orc_default_bounds(const float* pts, int64_t n, float* min_corner,
                                   float* max_corner) {
  for (int a = 0; a < 3; a++) {
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = 0; i < n; i++) {
      float v = pts[3 * i + a];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    min_corner[a] = lo - 1e-3f;
    max_corner[a] = hi + 1e-3f;
  }
}

ORC_EXPORT float orc_span(const float* min_corner, const float* max_corner) {
  return fmaxf(fmaxf(max_corner[0] - min_corner[0], max_corner[1] - min_corner[1]),
               max_corner[2] - min_corner[2]);
}

/* Stable refinement of the reference's counting-sort scatter
 * (src/sbh_ffi.cu:1723-1743): points ordered by (code, original index).
 * The reference's intra-bin order is an atomic race; this is the
 * deterministic order the product builder must reproduce bit-exactly. */
ORC_EXPORT void orc_stable_order(const uint32_t* codes, int64_t n, int bits3,
                                 uint32_t* perm) {
  /* LSD radix sort, 8-bit digits: an independent algorithm from the GPU's. */
  uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) a[i] = (uint32_t)i;
  for (int shift = 0; shift < bits3; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; i++) cnt[((codes[a[i]] >> shift) & 255u) + 1]++;
    for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; i++) b[cnt[(codes[a[i]] >> shift) & 255u]++] = a[i];
    uint32_t* t = a;
    a = b;
    b = t;
  }
  memcpy(perm, a, sizeof(uint32_t) * (size_t)n);
  free(a);
  free(b);
}

/* ------------------------------------------------------------------------- */
/* B3-B8: dense octree8 builder (src/sbh_ffi.cu:1039-1420, 1682-1876)         */
/* ------------------------------------------------------------------------- */
ORC_EXPORT uint64_t orc_octree8_total_nodes(int max_depth) {
  uint64_t total = 0, level = 1;
  for (int l = 0; l <= max_depth; l++) {
    total += level;
    level <<= 3;
  }
  return total;
}

/* Returns 0 on success, 1 for the reference's InvalidArgument branches. */
ORC_EXPORT int orc_build_octree8(const float* pts, const float* normals,
                                 const float* masses, int64_t n,
                                 const float* min_corner, const float* max_corner,
                                 int max_depth, uint8_t* leaf_out, uint8_t* node_out,
                                 uint8_t* contrib_out) {
  if (max_depth < 1 || max_depth > 10) return 1;
  float span = orc_span(min_corner, max_corner);
  if (!(span > 0.0f)) return 1; /* "invalid bounds span" (sbh_ffi.cu:1131-1133) */
  const uint32_t leaf_count = 1u << (3 * max_depth);
  const uint64_t total_nodes = orc_octree8_total_nodes(max_depth);

  OrcLeaf3* leaf = (OrcLeaf3*)leaf_out;
  OrcNode8* nodes = (OrcNode8*)node_out;
  OrcLeaf3* contrib = (OrcLeaf3*)contrib_out;
  memset(leaf_out, 0, sizeof(OrcLeaf3) * (size_t)n);
  memset(node_out, 0, sizeof(OrcNode8) * (size_t)total_nodes);
  memset(contrib_out, 0, sizeof(OrcLeaf3) * (size_t)total_nodes);

  uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  orc_morton_codes(pts, n, min_corner, span, max_depth, codes);
  orc_stable_order(codes, n, 3 * max_depth, perm);

  /* Per-level sums, level D first (histogram, sbh_ffi.cu:1706-1713). */
  uint32_t** cnt = (uint32_t**)calloc((size_t)max_depth + 1, sizeof(uint32_t*));
  double** sums = (double**)calloc((size_t)max_depth + 1, sizeof(double*)); /* 7/bin */
  for (int l = 0; l <= max_depth; l++) {
    size_t c = (size_t)1 << (3 * l);
    cnt[l] = (uint32_t*)calloc(c, sizeof(uint32_t));
    sums[l] = (double*)calloc(c * 7, sizeof(double));
  }
  for (int64_t i = 0; i < n; i++) {
    uint32_t c = codes[i];
    float m = masses[i];
    double* s = sums[max_depth] + 7 * (size_t)c;
    cnt[max_depth][c]++;
    s[0] += m;
    s[1] += (double)(m * pts[3 * i + 0]);
    s[2] += (double)(m * pts[3 * i + 1]);
    s[3] += (double)(m * pts[3 * i + 2]);
    s[4] += (double)(m * normals[3 * i + 0]);
    s[5] += (double)(m * normals[3 * i + 1]);
    s[6] += (double)(m * normals[3 * i + 2]);
  }
  /* Exclusive scan of bin counts (sbh_ffi.cu:1303-1321). */
  uint32_t* offsets = (uint32_t*)calloc((size_t)leaf_count + 1, sizeof(uint32_t));
  for (uint32_t b = 0; b < leaf_count; b++) offsets[b + 1] = offsets[b] + cnt[max_depth][b];

  /* Scatter in stable order (sbh_ffi.cu:1723-1743). */
  for (int64_t r = 0; r < n; r++) {
    uint32_t i = perm[r];
    OrcLeaf3* o = &leaf[r];
    o->pos[0] = pts[3 * i + 0];
    o->pos[1] = pts[3 * i + 1];
    o->pos[2] = pts[3 * i + 2];
    o->normal[0] = normals[3 * i + 0];
    o->normal[1] = normals[3 * i + 1];
    o->normal[2] = normals[3 * i + 2];
    o->mass = masses[i];
    o->alpha = 1.0f;
  }

  /* Bottom-up reduction over 8 children (sbh_ffi.cu:1745-1786). */
  for (int l = max_depth - 1; l >= 0; l--) {
    size_t c = (size_t)1 << (3 * l);
    for (size_t p = 0; p < c; p++) {
      for (int k = 0; k < 8; k++) {
        size_t ch = p * 8 + (size_t)k;
        if (cnt[l + 1][ch] == 0) continue;
        cnt[l][p] += cnt[l + 1][ch];
        for (int q = 0; q < 7; q++) sums[l][7 * p + q] += sums[l + 1][7 * ch + q];
      }
    }
  }

  /* Node records (sbh_ffi.cu:1390-1416, 1788-1876). */
  float cx = min_corner[0] + span * 0.5f;
  float cy = min_corner[1] + span * 0.5f;
  float cz = min_corner[2] + span * 0.5f;
  float half_size = span * 0.5f;
  uint32_t level_offset = 0, parent_level_offset = 0;
  for (int l = 0; l <= max_depth; l++) {
    uint32_t level_nodes = 1u << (3 * l);
    uint32_t child_level_offset = level_offset + level_nodes;
    for (uint32_t idx = 0; idx < level_nodes; idx++) {
      uint32_t node_idx = level_offset + idx;
      OrcNode8 node;
      memset(&node, 0, sizeof(node));
      node.depth = (uint8_t)l;
      node.parent = (l == 0) ? node_idx : (parent_level_offset + (idx >> 3));
      node.is_leaf = (l == max_depth) ? 1 : 0;
      node.child_mask = 0;
      node.first_child = 0xFFFFFFFFu;
      int shift = 3 * (max_depth - l);
      uint32_t first_leaf = idx << shift;
      node.start = offsets[first_leaf];
      node.end = node.start + cnt[l][idx];
      int grid = 1 << l;
      int ix = 0, iy = 0, iz = 0;
      for (int b = 0; b < l; b++) {
        ix |= (int)((idx >> (3 * b + 0)) & 1u) << b;
        iy |= (int)((idx >> (3 * b + 1)) & 1u) << b;
        iz |= (int)((idx >> (3 * b + 2)) & 1u) << b;
      }
      float step = half_size * 2.0f / (float)grid;
      node.cx = fmaf((float)ix + 0.5f, step, cx - half_size);
      node.cy = fmaf((float)iy + 0.5f, step, cy - half_size);
      node.cz = fmaf((float)iz + 0.5f, step, cz - half_size);
      node.half_size = step * 0.5f;
      node.diameter = step * 1.7320508075688772f;
      if (l < max_depth) {
        uint32_t child_base = child_level_offset + (idx << 3);
        uint8_t mask = 0;
        for (uint32_t c = 0; c < 8; c++) {
          node.child[c] = child_base + c;
          if (cnt[l + 1][(idx << 3) + c] > 0) {
            mask |= (uint8_t)(1u << c);
            if (node.first_child == 0xFFFFFFFFu) node.first_child = child_base + c;
          }
        }
        node.child_mask = mask;
        if (mask == 0) node.is_leaf = 1;
      } else {
        for (int c = 0; c < 8; c++) node.child[c] = 0xFFFFFFFFu;
      }
      const double* s = sums[l] + 7 * (size_t)idx;
      float mass = (float)s[0];
      node.contrib.alpha = 1.0f;
      node.contrib.mass = mass;
      if (mass > 0.0f) {
        for (int a = 0; a < 3; a++) {
          node.contrib.pos[a] = (float)(s[1 + a] / s[0]);
          node.contrib.normal[a] = (float)(s[4 + a] / s[0]);
        }
      }
      nodes[node_idx] = node;
      contrib[node_idx] = node.contrib;
    }
    parent_level_offset = level_offset;
    level_offset = child_level_offset;
  }

  for (int l = 0; l <= max_depth; l++) {
    free(cnt[l]);
    free(sums[l]);
  }
  free(cnt);
  free(sums);
  free(offsets);
  free(codes);
  free(perm);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Pair kernels in fp64 (src/sbh_ffi.cu:142-160, 431-501, 2044-2090;          */
/* include/kernel.h:27-39, 55-60, 73-80, 89-94)                              */
/* ------------------------------------------------------------------------- */
typedef struct {
  double eps;       /* max(softening_scale, 0) */
  double rc;        /* cutoff_scale; <= 0 or inf => attenuation 1 */
  int periodic;     /* min-image active */
  double box;       /* box length */
  int nearbyint_mi; /* 1: direct-sum min-image (nearbyint), 0: kernel.h floor(x+0.5) */
} OrcParams;

static inline double orc_mi(double d, const OrcParams* P) {
  if (!P->periodic || !(P->box > 0.0)) return d;
  double inv_l = 1.0 / P->box;
  if (P->nearbyint_mi) return d - P->box * nearbyint(d * inv_l);
  return d - P->box * floor(d * inv_l + 0.5);
}

static inline int orc_has_cutoff(const OrcParams* P) {
  return P->rc > 0.0 && isfinite(P->rc);
}

/* Softened+cutoff force of a unit... source (p, m) on q: -m*att*s^3*d. */
static inline void orc_soft_force(const double* q, const double* p, double m,
                                  const OrcParams* P, double* f) {
  double d[3];
  for (int a = 0; a < 3; a++) d[a] = orc_mi(q[a] - p[a], P);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double s = 1.0 / sqrt(r2 + P->eps * P->eps + 1e-12);
  double s3 = s * s * s;
  double att = 1.0;
  if (orc_has_cutoff(P)) {
    double dist = sqrt(r2 + 1e-12);
    double u = dist / (2.0 * P->rc);
    att = erfc(u) + (2.0 / sqrt(M_PI)) * u * exp(-u * u);
  }
  double c = -m * att * s3;
  for (int a = 0; a < 3; a++) f[a] += c * d[a];
}

/* J^T g of the softened force w.r.t. the query position. */
static inline void orc_soft_force_vjp(const double* q, const double* p, double m,
                                      const double* g, const OrcParams* P,
                                      double* out) {
  double d[3];
  for (int a = 0; a < 3; a++) d[a] = orc_mi(q[a] - p[a], P);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double s = 1.0 / sqrt(r2 + P->eps * P->eps + 1e-12);
  double s3 = s * s * s;
  double s5 = s3 * s * s;
  double att = 1.0, cc = 0.0;
  if (orc_has_cutoff(P)) {
    double dist = sqrt(r2 + 1e-12);
    double u = dist / (2.0 * P->rc);
    double e = exp(-u * u);
    att = erfc(u) + (2.0 / sqrt(M_PI)) * u * e;
    cc = s3 * dist * e / (2.0 * sqrt(M_PI) * P->rc * P->rc * P->rc);
  }
  double rg = d[0] * g[0] + d[1] * g[1] + d[2] * g[2];
  double lin = att * s3;
  double rad = rg * (3.0 * att * s5 + cc);
  for (int a = 0; a < 3; a++) out[a] += -m * (lin * g[a] - rad * d[a]);
}

/* gravityGrad (kernel.h:89-94): -m*d/(|d|+0.1)^3 with d = minimage(q-p). */
static inline void orc_grav_grad(const double* q, const double* p, double m,
                                 const OrcParams* P, double* f) {
  double d[3];
  for (int a = 0; a < 3; a++) d[a] = orc_mi(q[a] - p[a], P);
  double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double den = dist + 0.1;
  double c = -m / (den * den * den);
  for (int a = 0; a < 3; a++) f[a] += c * d[a];
}

/* gravityGradVjp (sbh_ffi.cu:2044-2060). */
static inline void orc_grav_grad_vjp(const double* q, const double* p, double m,
                                     const double* g, const OrcParams* P, double* out) {
  double d[3];
  for (int a = 0; a < 3; a++) d[a] = orc_mi(q[a] - p[a], P);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double dist = sqrt(r2 + 1e-12);
  double den = dist + 0.1;
  double i3 = 1.0 / (den * den * den);
  double i4 = i3 / den;
  double rg = d[0] * g[0] + d[1] * g[1] + d[2] * g[2];
  double dinv = 1.0 / (dist + 1e-12);
  for (int a = 0; a < 3; a++) out[a] += -m * (i3 * g[a] - 3.0 * rg * dinv * i4 * d[a]);
}

static void orc_params(OrcParams* P, double eps, double rc, int periodic, double box,
                       int nearbyint_mi) {
  P->eps = eps > 0.0 ? eps : 0.0;
  P->rc = rc;
  P->periodic = periodic != 0 && box > 0.0;
  P->box = box;
  P->nearbyint_mi = nearbyint_mi;
}

/* D1 (sbh_ffi.cu:431-460, 503-522; python/unified_force_benchmark.py:208-230). */
ORC_EXPORT void orc_direct_force_soft(const float* pts, const float* masses, int64_t n,
                                      const float* queries, int64_t nq, double eps,
                                      double rc, int periodic, double box, double* out) {
  OrcParams P;
  orc_params(&P, eps, rc, periodic, box, 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double f[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      double p[3] = {pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]};
      orc_soft_force(q, p, masses[j], &P, f);
    }
    out[3 * i] = f[0];
    out[3 * i + 1] = f[1];
    out[3 * i + 2] = f[2];
  }
}

/* D2 (sbh_ffi.cu:462-501, 524-546): query-side J^T g. */
ORC_EXPORT void orc_direct_force_soft_vjp(const float* pts, const float* masses,
                                          int64_t n, const float* queries,
                                          const float* cot, int64_t nq, double eps,
                                          double rc, int periodic, double box,
                                          double* out) {
  OrcParams P;
  orc_params(&P, eps, rc, periodic, box, 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double g[3] = {cot[3 * i], cot[3 * i + 1], cot[3 * i + 2]};
    double o[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      double p[3] = {pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]};
      orc_soft_force_vjp(q, p, masses[j], g, &P, o);
    }
    out[3 * i] = o[0];
    out[3 * i + 1] = o[1];
    out[3 * i + 2] = o[2];
  }
}

/* Full adjoint of L = sum_i c_i . F_i for self-gravity (queries == sources),
 * no cutoff.  Not in the reference (its custom_vjp returns the query side
 * only and zero mass gradients, python/jax_sbh_ffi.py:576-588).
 *   grad_x[k] = sum_{i!=k} v_ik - sum_{j!=k} v_kj ... written per pair below.
 *   grad_m[j] = sum_{i!=j} -s^3 (c_i . d_ij) * att-free form.            */
ORC_EXPORT void orc_direct_full_vjp(const float* pts, const float* masses, int64_t n,
                                    const float* cot, double eps, int periodic,
                                    double box, double* gx, double* gm) {
  OrcParams P;
  orc_params(&P, eps, 0.0, periodic, box, 1);
  double e2 = P.eps * P.eps + 1e-12;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; k++) {
    double xk[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
    double ck[3] = {cot[3 * k], cot[3 * k + 1], cot[3 * k + 2]};
    double g[3] = {0, 0, 0};
    double gmk = 0.0;
    for (int64_t j = 0; j < n; j++) {
      if (j == k) continue;
      double xj[3] = {pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]};
      double cj[3] = {cot[3 * j], cot[3 * j + 1], cot[3 * j + 2]};
      double d[3];
      for (int a = 0; a < 3; a++) d[a] = orc_mi(xk[a] - xj[a], &P); /* d_kj */
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double s = 1.0 / sqrt(r2 + e2);
      double s3 = s * s * s, s5 = s3 * s * s;
      /* k as query, j as source: v_kj = m_j (3 s5 (c_k.d) d - s3 c_k) */
      double ckd = ck[0] * d[0] + ck[1] * d[1] + ck[2] * d[2];
      /* j as query, k as source: d_jk = -d; contribution to x_k is -v_jk with
       * v_jk = m_k (3 s5 (c_j.d_jk) d_jk - s3 c_j) */
      double cjd = cj[0] * d[0] + cj[1] * d[1] + cj[2] * d[2];
      double mj = masses[j], mk = masses[k];
      for (int a = 0; a < 3; a++) {
        double v_kj = mj * (3.0 * s5 * ckd * d[a] - s3 * ck[a]);
        double v_jk = mk * (3.0 * s5 * cjd * d[a] - s3 * cj[a]); /* (-d)(-cjd) = d*cjd */
        g[a] += v_kj - v_jk;
      }
      /* dL/dm_k from query j: c_j . (-s3 d_jk) = s3 (c_j . d) */
      gmk += s3 * cjd;
    }
    gx[3 * k] = g[0];
    gx[3 * k + 1] = g[1];
    gx[3 * k + 2] = g[2];
    gm[k] = gmk;
  }
}

/* Full adjoint with the short-range cutoff, built from the pinned pair
 * functions: grad_x[k] = sum_j [vjp(q=x_k,p=x_j,m_j,c_k) - vjp(q=x_j,p=x_k,m_k,c_j)],
 * grad_m[k] = sum_j c_j . F_unit(source x_k on query x_j).  The no-cutoff case
 * is pinned by central differences (orc_direct_full_vjp). */
ORC_EXPORT void orc_direct_full_vjp_cut(const float* pts, const float* masses, int64_t n,
                                        const float* cot, double eps, double rc,
                                        int periodic, double box, double* gx, double* gm) {
  OrcParams P;
  orc_params(&P, eps, rc, periodic, box, 1);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; k++) {
    double xk[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
    double ck[3] = {cot[3 * k], cot[3 * k + 1], cot[3 * k + 2]};
    double g[3] = {0, 0, 0}, neg[3] = {0, 0, 0};
    double gmk = 0.0;
    for (int64_t j = 0; j < n; j++) {
      if (j == k) continue;
      double xj[3] = {pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]};
      double cj[3] = {cot[3 * j], cot[3 * j + 1], cot[3 * j + 2]};
      orc_soft_force_vjp(xk, xj, masses[j], ck, &P, g);
      orc_soft_force_vjp(xj, xk, masses[k], cj, &P, neg);
      double f[3] = {0, 0, 0};
      orc_soft_force(xj, xk, 1.0, &P, f);
      gmk += cj[0] * f[0] + cj[1] * f[1] + cj[2] * f[2];
    }
    for (int a = 0; a < 3; a++) gx[3 * k + a] = g[a] - neg[a];
    gm[k] = gmk;
  }
}

/* D3 (sbh_ffi.cu:399-429; kernel.h:55-60): potential sum -m/(|d|+0.1). */
ORC_EXPORT void orc_bruteforce_gravity(const float* pts, const float* masses, int64_t n,
                                       int dim, const float* queries, int64_t nq,
                                       int periodic, double box, double* out) {
  OrcParams P;
  orc_params(&P, 0.0, 0.0, periodic, box, 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; i++) {
    double acc = 0.0;
    for (int64_t j = 0; j < n; j++) {
      double r2 = 0.0;
      for (int a = 0; a < dim; a++) {
        double d = orc_mi((double)pts[dim * j + a] - (double)queries[dim * i + a], &P);
        r2 += d * d;
      }
      acc += -masses[j] / (sqrt(r2) + 0.1);
    }
    out[i] = acc;
  }
}

/* ------------------------------------------------------------------------- */
/* W1-W4: classic walks over the dense octree8 (sbh_ffi.cu:1878-2144)         */
/* ------------------------------------------------------------------------- */
/* MAC (sbh_ffi.cu:1878-1893) in fp32; min-image as kernel.h:27-39. */
static inline float orc_mi_f(float d, int periodic, float box) {
  if (periodic && box > 0.0f) {
    float inv_l = 1.0f / box;
    return d - box * floorf(d * inv_l + 0.5f);
  }
  return d;
}

static inline int orc_use_far_field8(const OrcNode8* nd, const float* q, float beta,
                                     int periodic, float box) {
  if (nd->is_leaf) return 1;
  float dx = orc_mi_f(q[0] - nd->cx, periodic, box);
  float dy = orc_mi_f(q[1] - nd->cy, periodic, box);
  float dz = orc_mi_f(q[2] - nd->cz, periodic, box);
  float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  return dist / nd->diameter >= beta;
}

static inline float orc_axis_dist(float q, float lo, float hi) {
  if (q < lo) return lo - q;
  if (q > hi) return q - hi;
  return 0.0f;
}

/* Prune test (sbh_ffi.cu:1903-1948). */
static inline int orc_should_prune8(const OrcNode8* nd, const float* q, int prune_on,
                                    float r_cut_sq, int periodic, float box) {
  if (!prune_on) return 0;
  float qx = q[0], qy = q[1], qz = q[2];
  if (periodic && box > 0.0f) {
    float inv_l = 1.0f / box;
    qx -= box * nearbyintf((qx - nd->cx) * inv_l);
    qy -= box * nearbyintf((qy - nd->cy) * inv_l);
    qz -= box * nearbyintf((qz - nd->cz) * inv_l);
  }
  float dx = orc_axis_dist(qx, nd->cx - nd->half_size, nd->cx + nd->half_size);
  float dy = orc_axis_dist(qy, nd->cy - nd->half_size, nd->cy + nd->half_size);
  float dz = orc_axis_dist(qz, nd->cz - nd->half_size, nd->cz + nd->half_size);
  return dx * dx + dy * dy + dz * dz > r_cut_sq;
}

/* signedSolidAngle<3> (include/kernel.h:73-80) in fp64:
 * a n.(p - q) / (4 pi (|p - q|^3 + 1e-5)), added to acc[0]. */
static inline void orc_wn_pair(const double* q, const OrcLeaf3* src, const OrcParams* P,
                               double* acc) {
  double d0 = orc_mi((double)src->pos[0] - q[0], P);
  double d1 = orc_mi((double)src->pos[1] - q[1], P);
  double d2 = orc_mi((double)src->pos[2] - q[2], P);
  double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double nd = src->normal[0] * d0 + src->normal[1] * d1 + src->normal[2] * d2;
  acc[0] += (double)src->mass * nd / (4.0 * M_PI * (r * r * r + 1e-5));
}

/* gravityPotential (include/kernel.h:55-60): -m / (|p - q| + 0.1). */
static inline void orc_potential_pair(const double* q, const OrcLeaf3* src, const OrcParams* P,
                                      double* acc) {
  double d0 = orc_mi((double)src->pos[0] - q[0], P);
  double d1 = orc_mi((double)src->pos[1] - q[1], P);
  double d2 = orc_mi((double)src->pos[2] - q[2], P);
  acc[0] += -(double)src->mass / (sqrt(d0 * d0 + d1 * d1 + d2 * d2) + 0.1);
}

/* smoothDistExp (include/kernel.h:82-87): exp(-alpha |p - q|). */
static inline void orc_smooth_pair(const double* q, const OrcLeaf3* src, const OrcParams* P,
                                   double* acc) {
  double d0 = orc_mi((double)src->pos[0] - q[0], P);
  double d1 = orc_mi((double)src->pos[1] - q[1], P);
  double d2 = orc_mi((double)src->pos[2] - q[2], P);
  acc[0] += exp(-(double)src->alpha * sqrt(d0 * d0 + d1 * d1 + d2 * d2));
}

/* Exact smooth-distance sum (src/sbh_gpu.cu:66-77, 287-322), fp64. */
ORC_EXPORT void orc_bruteforce_smoothdist(const float* pts, const float* alpha, int64_t n,
                                          const float* queries, int64_t nq, double* out) {
  OrcParams P;
  orc_params(&P, 0.0, 0.0, 0, 1.0f, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double acc[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      OrcLeaf3 s;
      s.pos[0] = pts[3 * j]; s.pos[1] = pts[3 * j + 1]; s.pos[2] = pts[3 * j + 2];
      s.normal[0] = s.normal[1] = s.normal[2] = 0.0f;
      s.mass = 0.0f;
      s.alpha = alpha[j];
      orc_smooth_pair(q, &s, &P, acc);
    }
    out[i] = acc[0];
  }
}

/* Exact winding-number sum of dipoles (the brute_force_gpu.h:18-66 summation
 * with signedSolidAngle, in fp64). */
ORC_EXPORT void orc_bruteforce_wn(const float* pts, const float* nrm, const float* area, int64_t n,
                                  const float* queries, int64_t nq, double* out) {
  OrcParams P;
  orc_params(&P, 0.0, 0.0, 0, 1.0f, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double acc[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      OrcLeaf3 s;
      s.pos[0] = pts[3 * j]; s.pos[1] = pts[3 * j + 1]; s.pos[2] = pts[3 * j + 2];
      s.normal[0] = nrm[3 * j]; s.normal[1] = nrm[3 * j + 1]; s.normal[2] = nrm[3 * j + 2];
      s.mass = area[j];
      s.alpha = 1.0f;
      orc_wn_pair(q, &s, &P, acc);
    }
    out[i] = acc[0];
  }
}

/* kind: 0 = gravityGrad (unsoftened, W3), 1 = softened/cutoff force (W2),
 *       2 = softened/cutoff VJP (W4, needs cot), 3 = gravityGrad VJP,
 *       4 = winding number (signedSolidAngle, scalar in out[3 i]),
 *       5 = gravityPotential, 6 = smoothDistExp (scalars in out[3 i]). */
ORC_EXPORT void orc_bh_octree8(const uint8_t* leaf_bytes, const uint8_t* node_bytes,
                               const float* queries, const float* cot, int64_t nq,
                               int kind, float beta, double eps, double rc,
                               int prune_enabled, float prune_r_cut_mult, int periodic,
                               float box, double* out, int64_t* interactions) {
  const OrcLeaf3* leaf = (const OrcLeaf3*)leaf_bytes;
  const OrcNode8* nodes = (const OrcNode8*)node_bytes;
  OrcParams P;
  orc_params(&P, eps, rc, periodic, box, 0);
  int per = periodic != 0 && box > 0.0f;
  /* Prune radius (sbh_ffi.cu:3233-3237). */
  int prune_on = (prune_enabled != 0 && rc > 0.0) ? 1 : 0;
  float r_cut = prune_on ? prune_r_cut_mult * (float)rc : 0.0f;
  float r_cut_sq = r_cut * r_cut;
  int64_t total_inter = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total_inter)
  for (int64_t i = 0; i < nq; i++) {
    float qf[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double q[3] = {qf[0], qf[1], qf[2]};
    double g[3] = {0, 0, 0};
    if (cot) {
      g[0] = cot[3 * i];
      g[1] = cot[3 * i + 1];
      g[2] = cot[3 * i + 2];
    }
    double acc[3] = {0, 0, 0};
    int stack[128];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
      const OrcNode8* nd = &nodes[stack[--top]];
      if (nd->contrib.mass == 0.0f) continue;
      if ((kind == 1 || kind == 2) &&
          orc_should_prune8(nd, qf, prune_on, r_cut_sq, per, box))
        continue;
      if (orc_use_far_field8(nd, qf, beta, per, box)) {
        uint32_t s0 = nd->is_leaf ? nd->start : 0, s1 = nd->is_leaf ? nd->end : 1;
        for (uint32_t j = s0; j < s1; j++) {
          const OrcLeaf3* src = nd->is_leaf ? &leaf[j] : &nd->contrib;
          double p[3] = {src->pos[0], src->pos[1], src->pos[2]};
          switch (kind) {
            case 0: orc_grav_grad(q, p, src->mass, &P, acc); break;
            case 1: orc_soft_force(q, p, src->mass, &P, acc); break;
            case 2: orc_soft_force_vjp(q, p, src->mass, g, &P, acc); break;
            case 4: orc_wn_pair(q, src, &P, acc); break;
            case 5: orc_potential_pair(q, src, &P, acc); break;
            case 6: orc_smooth_pair(q, src, &P, acc); break;
            default: orc_grav_grad_vjp(q, p, src->mass, g, &P, acc); break;
          }
          total_inter++;
        }
      } else {
        for (int c = 0; c < 8; c++)
          if (nd->child_mask & (1u << c)) stack[top++] = (int)nd->child[c];
      }
    }
    out[3 * i] = acc[0];
    out[3 * i + 1] = acc[1];
    out[3 * i + 2] = acc[2];
  }
  if (interactions) *interactions = total_inter;
}

/* ------------------------------------------------------------------------- */
/* L1/L2: level-3 tree (src/sbh_ffi.cu:181-397, 736-1037) and its walks       */
/* (src/sbh_ffi.cu:1441-1680, include/barnes_hut_gpu.h:19-72,                 */
/* include/tree_sampler_gpu.h:25-140)                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint64_t start, end, buffer_idx, parent_idx;
  float bmin[3], bmax[3];
  float diameter;
  OrcLeaf3 contrib;
  uint64_t bitmask[8];
  uint64_t child[512];
  int32_t num_children, depth, max_depth;
  uint8_t is_leaf, pad[3];
} OrcNode3; /* 4272 bytes: Tree3DNode<1, 3> (include/tree3d.h:212-225) */

_Static_assert(sizeof(OrcNode3) == 4272, "Tree3DNode<1,3> layout");

ORC_EXPORT size_t orc_node3_size(void) { return sizeof(OrcNode3); }

/* quantize_axis_level3 (src/sbh_ffi.cu:181-190). */
static inline int orc_quantize3(float x, float lo, float hi) {
  float span = hi - lo;
  if (span <= 1e-12f) return 0;
  float t = (x - lo) / span;
  t = fminf(fmaxf(t, 0.0f), 0.999999f);
  return (int)(t * 8.0f);
}

/* Level-3 builder: cells by the per-axis quantiser, points in stable (cell,
 * index) order, fp64 cell sums, child records at 1 + rank in cell order.
 * The child minimum corner is fmaf(d, q, min) (nvcc contracts
 * min + d * q); diameters are not contracted. */
ORC_EXPORT void orc_build_level3(const float* pts, const float* normals, const float* masses,
                                 int64_t n, const float* lo, const float* hi,
                                 uint8_t* leaf_out, uint8_t* node_out, uint8_t* contrib_out) {
  OrcLeaf3* leaf = (OrcLeaf3*)leaf_out;
  OrcNode3* nodes = (OrcNode3*)node_out;
  OrcLeaf3* contrib = (OrcLeaf3*)contrib_out;
  memset(node_out, 0, sizeof(OrcNode3) * 513);
  memset(contrib_out, 0, sizeof(OrcLeaf3) * 513);
  uint32_t* cell = (uint32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) {
    int qx = orc_quantize3(pts[3 * i], lo[0], hi[0]);
    int qy = orc_quantize3(pts[3 * i + 1], lo[1], hi[1]);
    int qz = orc_quantize3(pts[3 * i + 2], lo[2], hi[2]);
    cell[i] = (uint32_t)(qx + 8 * (qy + 8 * qz));
  }
  orc_stable_order(cell, n, 9, perm);
  int64_t cnt[512], off[513];
  double sum[512][7], tot[7] = {0, 0, 0, 0, 0, 0, 0};
  memset(cnt, 0, sizeof(cnt));
  memset(sum, 0, sizeof(sum));
  for (int64_t r = 0; r < n; r++) {
    uint32_t i = perm[r];
    OrcLeaf3* o = &leaf[r];
    for (int a = 0; a < 3; a++) {
      o->pos[a] = pts[3 * i + a];
      o->normal[a] = normals[3 * i + a];
    }
    o->mass = masses[i];
    o->alpha = 1.0f;
    uint32_t c = cell[i];
    cnt[c]++;
    double m = masses[i];
    double v[7] = {m, m * pts[3 * i], m * pts[3 * i + 1], m * pts[3 * i + 2],
                   m * normals[3 * i], m * normals[3 * i + 1], m * normals[3 * i + 2]};
    for (int a = 0; a < 7; a++) {
      sum[c][a] += v[a];
      tot[a] += v[a];
    }
  }
  off[0] = 0;
  for (int c = 0; c < 512; c++) off[c + 1] = off[c] + cnt[c];
  float d[3];
  for (int a = 0; a < 3; a++) d[a] = (hi[a] - lo[a]) / 8.0f;
  OrcNode3* root = &nodes[0];
  int rank = 0;
  for (int c = 0; c < 512; c++) {
    if (cnt[c] == 0) continue;
    int idx = 1 + rank++;
    OrcNode3* ch = &nodes[idx];
    int q[3] = {c & 7, (c >> 3) & 7, (c >> 6) & 7};
    ch->start = (uint64_t)off[c];
    ch->end = (uint64_t)(off[c] + cnt[c]);
    ch->buffer_idx = (uint64_t)idx;
    ch->parent_idx = 0;
    for (int a = 0; a < 3; a++) {
      ch->bmin[a] = fmaf(d[a], (float)q[a], lo[a]);
      ch->bmax[a] = ch->bmin[a] + d[a];
    }
    ch->diameter = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float m = (float)sum[c][0];
    ch->contrib.mass = m;
    ch->contrib.alpha = 1.0f;
    if (m > 0.0f)
      for (int a = 0; a < 3; a++) {
        ch->contrib.pos[a] = (float)(sum[c][1 + a] / sum[c][0]);
        ch->contrib.normal[a] = (float)(sum[c][4 + a] / sum[c][0]);
      }
    ch->depth = 1;
    ch->max_depth = 1;
    ch->is_leaf = 1;
    contrib[idx] = ch->contrib;
    root->bitmask[c >> 6] |= 1ull << (c & 63);
    root->child[c] = (uint64_t)idx;
  }
  root->start = 0;
  root->end = (uint64_t)n;
  float dd[3];
  for (int a = 0; a < 3; a++) {
    root->bmin[a] = lo[a];
    root->bmax[a] = hi[a];
    dd[a] = hi[a] - lo[a];
  }
  root->diameter = sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
  float M = (float)tot[0];
  root->contrib.mass = M;
  root->contrib.alpha = 1.0f;
  if (M > 0.0f)
    for (int a = 0; a < 3; a++) {
      root->contrib.pos[a] = (float)(tot[1 + a] / tot[0]);
      root->contrib.normal[a] = (float)(tot[4 + a] / tot[0]);
    }
  root->num_children = rank;
  root->max_depth = 1;
  root->is_leaf = n <= 1 ? 1 : 0;
  contrib[0] = root->contrib;
  free(cell);
  free(perm);
}

/* kind: 0 = gravityPotential (scalar in out[3 i]), 1 = gravityGrad,
 * 2 = gravityGradSoftCutoff<3> (eps, rc; attenuation for any rc > 0). */
static void orc_l3_term(int kind, const double* q, const OrcLeaf3* src, const OrcParams* P,
                        double* acc) {
  double p[3] = {src->pos[0], src->pos[1], src->pos[2]};
  if (kind == 0) {
    double d0 = orc_mi(p[0] - q[0], P), d1 = orc_mi(p[1] - q[1], P), d2 = orc_mi(p[2] - q[2], P);
    acc[0] += -(double)src->mass / (sqrt(d0 * d0 + d1 * d1 + d2 * d2) + 0.1);
  } else if (kind == 1) {
    orc_grav_grad(q, p, src->mass, P, acc);
  } else {
    orc_soft_force(q, p, src->mass, P, acc);
  }
}

/* Recursive vote walk for one reference warp (queries g0..g1-1): a node is a
 * far field only if every query of the warp has |q - COM| / diameter >= beta
 * (leaves: always); the path is common to the warp. */
static void orc_l3_vote(const OrcLeaf3* leaf, const OrcNode3* nodes, int64_t nn, int64_t nl,
                        int64_t node, const float* qf, int64_t g0, int64_t g1, int kind,
                        float beta, const OrcParams* P, double* out) {
  const OrcNode3* nd = &nodes[node];
  int all = 1;
  for (int64_t i = g0; i < g1 && all; i++) {
    if (nd->is_leaf) break;
    float dx = qf[3 * i] - nd->contrib.pos[0];
    float dy = qf[3 * i + 1] - nd->contrib.pos[1];
    float dz = qf[3 * i + 2] - nd->contrib.pos[2];
    float ratio = sqrtf(dx * dx + dy * dy + dz * dz) / nd->diameter;
    if (!(ratio >= beta)) all = 0;
  }
  if (all) {
    for (int64_t i = g0; i < g1; i++) {
      double q[3] = {qf[3 * i], qf[3 * i + 1], qf[3 * i + 2]};
      if (nd->is_leaf) {
        uint64_t je = nd->end < (uint64_t)nl ? nd->end : (uint64_t)nl;
        for (uint64_t j = nd->start; j < je; j++) orc_l3_term(kind, q, &leaf[j], P, out + 3 * i);
      } else {
        orc_l3_term(kind, q, &nd->contrib, P, out + 3 * i);
      }
    }
    return;
  }
  for (int c = 0; c < 512; c++)
    if ((nd->bitmask[c >> 6] >> (c & 63)) & 1ull)
      if (nd->child[c] < (uint64_t)nn)
        orc_l3_vote(leaf, nodes, nn, nl, (int64_t)nd->child[c], qf, g0, g1, kind, beta, P, out);
}

/* Classic level-3 walk (samples < 0) or the stochastic one (samples >= 0:
 * every subdomain is a leaf in level-3 buffers; 0 samples give 0 / 0).
 * Reference warps: 32 consecutive queries of a block of `threads`. */
ORC_EXPORT void orc_l3_eval(const uint8_t* leaf_bytes, int64_t nl, const uint8_t* node_bytes,
                            int64_t nn, const float* queries, int64_t nq, int kind, float beta,
                            int threads, int samples, double eps, double rc, int periodic,
                            float box, double* out) {
  const OrcLeaf3* leaf = (const OrcLeaf3*)leaf_bytes;
  const OrcNode3* nodes = (const OrcNode3*)node_bytes;
  OrcParams P;
  orc_params(&P, eps, rc > 0.0 ? rc : 0.0, periodic, box, 0);
  memset(out, 0, sizeof(double) * 3 * (size_t)nq);
  if (samples < 0) {
    for (int64_t b0 = 0; b0 < nq; b0 += threads)
      for (int64_t w0 = b0; w0 < b0 + threads && w0 < nq; w0 += 32) {
        int64_t w1 = w0 + 32;
        if (w1 > b0 + threads) w1 = b0 + threads;
        if (w1 > nq) w1 = nq;
        orc_l3_vote(leaf, nodes, nn, nl, 0, queries, w0, w1, kind, beta, &P, out);
      }
    return;
  }
  const OrcNode3* root = &nodes[0];
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double extra = 0.0;
    for (int c = 0; c < 512; c++) {
      if (!((root->bitmask[c >> 6] >> (c & 63)) & 1ull) || root->child[c] >= (uint64_t)nn)
        continue;
      const OrcNode3* sub = &nodes[root->child[c]];
      if (sub->is_leaf) {
        uint64_t je = sub->end < (uint64_t)nl ? sub->end : (uint64_t)nl;
        for (uint64_t j = sub->start; j < je; j++) orc_l3_term(kind, q, &leaf[j], &P, out + 3 * i);
      } else {
        orc_l3_term(kind, q, &sub->contrib, &P, out + 3 * i);
      }
      extra += 0.0 / (double)samples;
    }
    for (int a = 0; a < 3; a++) out[3 * i + a] += extra;
  }
}

/* ------------------------------------------------------------------------- */
/* S1/S2: stochastic multi-level control-variate walk (src/sbh_ffi.cu:2146-2442) */
/* ------------------------------------------------------------------------- */
ORC_EXPORT void orc_philox4x32(const uint32_t* ctr_in, const uint32_t* key_in,
                               uint32_t* out);

/* Uniform in (0, 1] from one Philox draw indexed by (query, node, sample,
 * step); replaces curand_uniform on the XORWOW streams seeded with seed + q
 * (src/sbh_ffi.cu:2160-2163).  P(u > t) = 1 - t for t a multiple of 2^-24. */
static inline float orc_sbh_uniform(uint32_t seed, uint32_t q, uint32_t node, uint32_t sample,
                                    uint32_t step) {
  uint32_t ctr[4] = {q, node, sample, step};
  uint32_t key[2] = {seed, 0x5BD1E995u};
  uint32_t out[4];
  orc_philox4x32(ctr, key, out);
  return (float)((out[0] >> 8) + 1u) * (1.0f / 16777216.0f);
}

static inline uint32_t orc_child_slot8(const OrcNode8* nd, const float* p) {
  uint32_t ix = p[0] >= nd->cx ? 1u : 0u;
  uint32_t iy = p[1] >= nd->cy ? 1u : 0u;
  uint32_t iz = p[2] >= nd->cz ? 1u : 0u;
  return ix | (iy << 1) | (iz << 2);
}

/* Distance from q to the cell centre (min-image, kernel.h:27-39), fp32. */
static inline float orc_centre_dist8(const OrcNode8* nd, const float* q, int per, float box) {
  float dx = orc_mi_f(q[0] - nd->cx, per, box);
  float dy = orc_mi_f(q[1] - nd->cy, per, box);
  float dz = orc_mi_f(q[2] - nd->cz, per, box);
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* kind as orc_bh_octree8 (0 grav, 1 soft, 2 soft vjp, 3 grav vjp, 4 wn); the
 * softened kinds apply the prune test like the reference (softened = true).
 * Subdomains: nodes at depth 3 or leaves above it (kControlDepth = 3,
 * src/sbh_ffi.cu:2167-2192), in the reference's push/pop order. */
ORC_EXPORT void orc_sbh_octree8(const uint8_t* leaf_bytes, const uint8_t* node_bytes,
                                const float* queries, const float* cot, int64_t nq, int kind,
                                int samples, int seed, float beta, double eps, double rc,
                                int prune_enabled, float prune_r_cut_mult, int periodic,
                                float box, double* out) {
  const OrcLeaf3* leaf = (const OrcLeaf3*)leaf_bytes;
  const OrcNode8* nodes = (const OrcNode8*)node_bytes;
  OrcParams P;
  orc_params(&P, eps, rc, periodic, box, 0);
  int per = periodic != 0 && box > 0.0f;
  int softened = (kind == 1 || kind == 2);
  int prune_on = (softened && prune_enabled != 0 && rc > 0.0) ? 1 : 0;
  float r_cut = prune_on ? prune_r_cut_mult * (float)rc : 0.0f;
  float r_cut_sq = r_cut * r_cut;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < nq; i++) {
    float qf[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
    double q[3] = {qf[0], qf[1], qf[2]};
    double g[3] = {0, 0, 0};
    if (cot) {
      g[0] = cot[3 * i];
      g[1] = cot[3 * i + 1];
      g[2] = cot[3 * i + 2];
    }
#define ORC_TERM(src, acc)                                                 \
  do {                                                                     \
    double p_[3] = {(src)->pos[0], (src)->pos[1], (src)->pos[2]};          \
    switch (kind) {                                                        \
      case 0: orc_grav_grad(q, p_, (src)->mass, &P, acc); break;           \
      case 1: orc_soft_force(q, p_, (src)->mass, &P, acc); break;          \
      case 2: orc_soft_force_vjp(q, p_, (src)->mass, g, &P, acc); break;   \
      case 4: orc_wn_pair(q, (src), &P, acc); break;                       \
      case 5: orc_potential_pair(q, (src), &P, acc); break;                \
      case 6: orc_smooth_pair(q, (src), &P, acc); break;                   \
      default: orc_grav_grad_vjp(q, p_, (src)->mass, g, &P, acc); break;   \
    }                                                                      \
  } while (0)
    double total[3] = {0, 0, 0};
    int stack[64];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
      int idx = stack[--top];
      const OrcNode8* nd = &nodes[idx];
      if (nd->contrib.mass == 0.0f) continue;
      if (!(nd->is_leaf || nd->depth >= 3)) {
        for (int c = 0; c < 8; c++)
          if (nd->child_mask & (1u << c)) stack[top++] = (int)nd->child[c];
        continue;
      }
      const OrcNode8* sub = nd;
      if (prune_on && orc_should_prune8(sub, qf, 1, r_cut_sq, per, box)) continue;
      if (sub->is_leaf) {
        for (uint32_t j = sub->start; j < sub->end; j++) ORC_TERM(&leaf[j], total);
      } else {
        ORC_TERM(&sub->contrib, total);
      }
      uint32_t num_el = sub->end - sub->start;
      double sub_total[3] = {0, 0, 0};
      for (int ss = 0; ss < samples; ss++) {
        float u = orc_sbh_uniform((uint32_t)seed, (uint32_t)i, (uint32_t)idx, (uint32_t)ss, 0);
        uint32_t sidx = (uint32_t)((float)num_el * u) + sub->start;
        if (sidx > sub->end - 1) sidx = sub->end - 1;
        const float* spos = leaf[sidx].pos;
        const OrcNode8* prev = sub;
        const OrcNode8* cur = sub;
        float event_prob = 1.0f;
        double wsum[3] = {0, 0, 0};
        float cur_beta = fmaxf(beta, orc_centre_dist8(cur, qf, per, box) / cur->diameter);
        uint32_t step = 1;
        while (!cur->is_leaf) {
          if (prune_on && orc_should_prune8(cur, qf, 1, r_cut_sq, per, box)) break;
          float ratio = orc_centre_dist8(cur, qf, per, box) / cur->diameter;
          float tp = fminf(cur_beta / ratio, 1.0f);
          float pu = orc_sbh_uniform((uint32_t)seed, (uint32_t)i, (uint32_t)idx, (uint32_t)ss,
                                     step++);
          if (pu > tp) break;
          event_prob *= tp;
          prev = cur;
          uint32_t slot = orc_child_slot8(cur, spos);
          if (!(cur->child_mask & (1u << slot))) break;
          cur = &nodes[cur->child[slot]];
          cur_beta = fmaxf(cur_beta, ratio);
          double prev_r[3] = {0, 0, 0}, cur_r[3] = {0, 0, 0};
          ORC_TERM(&prev->contrib, prev_r);
          for (int cc = 0; cc < 8; cc++) {
            if (!(prev->child_mask & (1u << cc))) continue;
            const OrcNode8* ch = &nodes[prev->child[cc]];
            if (prune_on && orc_should_prune8(ch, qf, 1, r_cut_sq, per, box)) continue;
            if (ch->is_leaf) {
              for (uint32_t j = ch->start; j < ch->end; j++) ORC_TERM(&leaf[j], cur_r);
            } else {
              ORC_TERM(&ch->contrib, cur_r);
            }
          }
          float prob_inv = (float)num_el / (event_prob * (float)(prev->end - prev->start));
          for (int a = 0; a < 3; a++) wsum[a] += (double)prob_inv * (cur_r[a] - prev_r[a]);
        }
        for (int a = 0; a < 3; a++) sub_total[a] += wsum[a];
      }
      if (samples > 0)
        for (int a = 0; a < 3; a++) total[a] += sub_total[a] / (double)samples;
    }
#undef ORC_TERM
    out[3 * i] = total[0];
    out[3 * i + 1] = total[1];
    out[3 * i + 2] = total[2];
  }
}

/* ------------------------------------------------------------------------- */
/* Adaptive Morton octree (new; no reference counterpart beyond the octree8    */
/* quantiser and geometry it shares).                                         */
/*   - codes: orc_morton_codes at depth D, stable order by (code, index)      */
/*   - node = non-empty cell whose parent holds more than leaf_size points    */
/*   - leaf  = count <= leaf_size or level == D                               */
/*   - nodes in breadth-first order; children of a node are contiguous        */
/* ------------------------------------------------------------------------- */
static inline void orc_cell_geometry(uint32_t prefix, int level, const float* min_corner,
                                     float span, float* geo) {
  float cx = min_corner[0] + span * 0.5f;
  float cy = min_corner[1] + span * 0.5f;
  float cz = min_corner[2] + span * 0.5f;
  float half_size = span * 0.5f;
  int ix = 0, iy = 0, iz = 0;
  for (int b = 0; b < level; b++) {
    ix |= (int)((prefix >> (3 * b + 0)) & 1u) << b;
    iy |= (int)((prefix >> (3 * b + 1)) & 1u) << b;
    iz |= (int)((prefix >> (3 * b + 2)) & 1u) << b;
  }
  float step = half_size * 2.0f / (float)(1 << level);
  geo[0] = fmaf((float)ix + 0.5f, step, cx - half_size);
  geo[1] = fmaf((float)iy + 0.5f, step, cy - half_size);
  geo[2] = fmaf((float)iz + 0.5f, step, cz - half_size);
  geo[3] = step * 1.7320508075688772f;
}

/* Returns the node count, or -1 if cap is too small. */
ORC_EXPORT int64_t orc_tree_build(const uint32_t* codes_sorted, int64_t n, int max_depth,
                                  int leaf_size, const float* min_corner, float span,
                                  int64_t cap, int32_t* start, int32_t* count,
                                  int32_t* first_child, int32_t* nchild, int32_t* parent,
                                  int32_t* level, float* geo) {
  if (cap < 1) return -1;
  int64_t num = 1;
  start[0] = 0;
  count[0] = (int32_t)n;
  parent[0] = -1;
  level[0] = 0;
  orc_cell_geometry(0u, 0, min_corner, span, geo);
  int64_t head = 0; /* BFS */
  while (head < num) {
    int64_t nd = head++;
    int l = level[nd];
    first_child[nd] = -1;
    nchild[nd] = 0;
    if (l >= max_depth || count[nd] <= leaf_size || count[nd] == 0) continue;
    int shift = 3 * (max_depth - l - 1);
    int32_t s = start[nd], e = start[nd] + count[nd];
    first_child[nd] = (int32_t)num;
    int32_t i = s;
    while (i < e) {
      uint32_t pre = codes_sorted[i] >> shift;
      int32_t j = i;
      while (j < e && (codes_sorted[j] >> shift) == pre) j++;
      if (num >= cap) return -1;
      start[num] = i;
      count[num] = j - i;
      parent[num] = (int32_t)nd;
      level[num] = l + 1;
      orc_cell_geometry(pre, l + 1, min_corner, span, geo + 4 * num);
      num++;
      nchild[nd]++;
      i = j;
    }
  }
  return num;
}

/* Moments in fp64 over sorted sources: mass, centre of mass, and the central
 * second moment S = sum m (x-X)(x-X)^T stored as (xx, yy, zz, xy, xz, yz). */
ORC_EXPORT void orc_tree_moments(const int32_t* start, const int32_t* count, int64_t nn,
                                 const double* xs, const double* ms, double* mass,
                                 double* com, double* quad) {
  for (int64_t k = 0; k < nn; k++) {
    double M = 0, X[3] = {0, 0, 0};
    for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
      M += ms[j];
      for (int a = 0; a < 3; a++) X[a] += ms[j] * xs[3 * j + a];
    }
    double Q[6] = {0, 0, 0, 0, 0, 0};
    if (M != 0.0) {
      for (int a = 0; a < 3; a++) X[a] /= M;
      for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
        double d0 = xs[3 * j] - X[0], d1 = xs[3 * j + 1] - X[1], d2 = xs[3 * j + 2] - X[2];
        double m = ms[j];
        Q[0] += m * d0 * d0;
        Q[1] += m * d1 * d1;
        Q[2] += m * d2 * d2;
        Q[3] += m * d0 * d1;
        Q[4] += m * d0 * d2;
        Q[5] += m * d1 * d2;
      }
    }
    mass[k] = M;
    for (int a = 0; a < 3; a++) com[3 * k + a] = X[a];
    for (int a = 0; a < 6; a++) quad[6 * k + a] = Q[a];
  }
}

/* Query groups of the walk: a partition of the sorted points into contiguous
 * ranges of at most gmax points that stay inside one cell.  Every maximal
 * node holding at most cell_max points (its parent holds more, or it has no
 * parent) and every leaf holding more than cell_max points (possible at the
 * depth limit) is cut into consecutive runs of gmax points, the last run
 * shorter.  Order: by node index.  Returns the number of groups. */
ORC_EXPORT int64_t orc_tree_groups(int64_t nn, const int32_t* start, const int32_t* count,
                                   const int32_t* first_child, const int32_t* nchild,
                                   const int32_t* parent, int gmax, int cell_max,
                                   int32_t* g_start, int32_t* g_count) {
  (void)first_child;
  int64_t ng = 0;
  for (int64_t k = 0; k < nn; k++) {
    if (count[k] == 0) continue;
    int32_t p = parent[k];
    int maximal = count[k] <= cell_max && (p < 0 || count[p] > cell_max);
    int overfull_leaf = count[k] > cell_max && nchild[k] == 0;
    if (!maximal && !overfull_leaf) continue;
    for (int32_t s = start[k]; s < start[k] + count[k]; s += gmax) {
      g_start[ng] = s;
      g_count[ng] = (start[k] + count[k] - s) < gmax ? (start[k] + count[k] - s) : gmax;
      ng++;
    }
  }
  return ng;
}

/* Group boxes (fp32 min / max of the group's sorted query positions). */
ORC_EXPORT void orc_group_boxes(const float* q, int64_t ng, const int32_t* g_start,
                                const int32_t* g_count, float* boxes) {
  for (int64_t g = 0; g < ng; g++) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = g_start[g]; i < g_start[g] + g_count[g]; i++)
      for (int a = 0; a < 3; a++) {
        lo[a] = fminf(lo[a], q[3 * i + a]);
        hi[a] = fmaxf(hi[a], q[3 * i + a]);
      }
    for (int a = 0; a < 3; a++) {
      boxes[6 * g + a] = lo[a];
      boxes[6 * g + 3 + a] = hi[a];
    }
  }
}

/* Group MAC: accept node for every query of the group iff the distance from
 * the cell centre to the group box is at least beta * diagonal.  fp32, no
 * contraction; the product kernel evaluates the identical expression. */
static inline int orc_group_accept(const float* box, const float* geo, float beta) {
  float dx = fmaxf(fmaxf(box[0] - geo[0], geo[0] - box[3]), 0.0f);
  float dy = fmaxf(fmaxf(box[1] - geo[1], geo[1] - box[4]), 0.0f);
  float dz = fmaxf(fmaxf(box[2] - geo[2], geo[2] - box[5]), 0.0f);
  float d2 = dx * dx + dy * dy + dz * dz;
  float t = beta * geo[3];
  return d2 >= t * t;
}

/* Far-field (monopole [+ quadrupole]) force and its adjoint for one query.
 * F = -M s^3 D + 3 s^5 (S D) + 1.5 tr(S) s^5 D - 7.5 s^7 (D.S.D) D,
 * s = (|D|^2 + eps^2 + 1e-12)^(-1/2), D = q - X.
 * use_quad: 0 monopole; 1 monopole + quadrupole; 2 monopole terms (force,
 * query-side gradient, dL/dX, dL/dM) plus the quadrupole term's dL/dS, which
 * does not depend on S: pushed down to the sources it adds the first-order
 * term of their position gradients and the second-order term of their mass
 * gradients (the Taylor expansion of the exact source-side gradient about
 * the centre of mass). */
static void orc_far(const double* q, double M, const double* X, const double* Q,
                    int use_quad, double e2, const double* c, double* F, double* gq,
                    double* adjX, double* adjM, double* adjS) {
  double D[3] = {q[0] - X[0], q[1] - X[1], q[2] - X[2]};
  double r2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  double s = 1.0 / sqrt(r2 + e2);
  double s2 = s * s, s3 = s2 * s, s5 = s3 * s2, s7 = s5 * s2, s9 = s7 * s2;
  double SD[3] = {0, 0, 0}, t = 0.0, T = 0.0;
  if (use_quad == 1) {
    SD[0] = Q[0] * D[0] + Q[3] * D[1] + Q[4] * D[2];
    SD[1] = Q[3] * D[0] + Q[1] * D[1] + Q[5] * D[2];
    SD[2] = Q[4] * D[0] + Q[5] * D[1] + Q[2] * D[2];
    t = D[0] * SD[0] + D[1] * SD[1] + D[2] * SD[2];
    T = Q[0] + Q[1] + Q[2];
  }
  for (int a = 0; a < 3; a++)
    F[a] += -M * s3 * D[a] + 3.0 * s5 * SD[a] + 1.5 * T * s5 * D[a] - 7.5 * s7 * t * D[a];
  if (!c) return;
  double u = c[0] * D[0] + c[1] * D[1] + c[2] * D[2];
  double Sc[3] = {0, 0, 0}, cw = 0.0;
  if (use_quad == 1) {
    Sc[0] = Q[0] * c[0] + Q[3] * c[1] + Q[4] * c[2];
    Sc[1] = Q[3] * c[0] + Q[1] * c[1] + Q[5] * c[2];
    Sc[2] = Q[4] * c[0] + Q[5] * c[1] + Q[2] * c[2];
    cw = c[0] * SD[0] + c[1] * SD[1] + c[2] * SD[2];
  }
  double kD = 3.0 * M * s5 * u - 15.0 * s7 * cw - 7.5 * T * s7 * u + 52.5 * s9 * t * u;
  double kc = -M * s3 + 1.5 * T * s5 - 7.5 * s7 * t;
  double dD[3];
  for (int a = 0; a < 3; a++)
    dD[a] = kD * D[a] + kc * c[a] + 3.0 * s5 * Sc[a] - 15.0 * s7 * u * SD[a];
  for (int a = 0; a < 3; a++) {
    gq[a] += dD[a];
    adjX[a] -= dD[a];
  }
  *adjM += -s3 * u;
  if (use_quad) {
    /* symmetric gradient w.r.t. S, stored (xx, yy, zz, xy, xz, yz) as the
     * matrix entries G_jl (off-diagonal entries counted once). */
    int jj[6] = {0, 1, 2, 0, 0, 1}, ll[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; k++) {
      int j = jj[k], l = ll[k];
      double G = 1.5 * s5 * (c[j] * D[l] + c[l] * D[j]) - 7.5 * s7 * u * D[j] * D[l];
      if (j == l) G += 1.5 * s5 * u;
      adjS[k] += G;
    }
  }
}

/* Evaluate the adaptive-tree walk over sorted queries in groups (ranges
 * g_start/g_count, bounding boxes `boxes`, see orc_tree_groups).
 *   self_mode: queries are the sorted sources (query i == source i); the
 *              i == j pair is skipped.
 *   cot == NULL: forces only.  Otherwise also
 *     gq   (nq*3): query-side gradient,
 *     gsx  (n*3), gsm (n): source-side position and mass gradients
 *                 (pair terms plus the push-down of node adjoints).
 *   counts[0] = far (node) interactions, counts[1] = near pairs.          */
ORC_EXPORT void orc_tree_eval_range(int64_t nn, const int32_t* start, const int32_t* count,
                                    const int32_t* first_child, const int32_t* nchild,
                                    const float* geo, const double* mass, const double* com,
                                    const double* quad, const double* xs, const double* ms,
                                    int64_t n, const float* boxes, const double* qx,
                                    int64_t nq, const int32_t* g_start,
                                    const int32_t* g_count, float beta, int use_quad,
                                    double eps, int self_mode, const double* cot,
                                    double* force, double* gq, double* gsx, double* gsm,
                                    int64_t* counts, int64_t g_begin, int64_t g_end);

ORC_EXPORT void orc_tree_eval(int64_t nn, const int32_t* start, const int32_t* count,
                              const int32_t* first_child, const int32_t* nchild,
                              const float* geo, const double* mass, const double* com,
                              const double* quad, const double* xs, const double* ms,
                              int64_t n, const float* boxes, const double* qx,
                              int64_t nq, int64_t ng, const int32_t* g_start,
                              const int32_t* g_count, float beta, int use_quad,
                              double eps, int self_mode, const double* cot,
                              double* force, double* gq, double* gsx, double* gsm,
                              int64_t* counts) {
  orc_tree_eval_range(nn, start, count, first_child, nchild, geo, mass, com, quad, xs, ms, n,
                      boxes, qx, nq, g_start, g_count, beta, use_quad, eps, self_mode, cot,
                      force, gq, gsx, gsm, counts, 0, ng);
}

/* Same walk restricted to groups [g_begin, g_end) (used to time a bounded
 * sample of the full evaluation). */
ORC_EXPORT void orc_tree_eval_range(int64_t nn, const int32_t* start, const int32_t* count,
                                    const int32_t* first_child, const int32_t* nchild,
                                    const float* geo, const double* mass, const double* com,
                                    const double* quad, const double* xs, const double* ms,
                                    int64_t n, const float* boxes, const double* qx,
                                    int64_t nq, const int32_t* g_start,
                                    const int32_t* g_count, float beta, int use_quad,
                                    double eps, int self_mode, const double* cot,
                                    double* force, double* gq, double* gsx, double* gsm,
                                    int64_t* counts, int64_t g_begin, int64_t g_end) {
  double e2 = (eps > 0 ? eps * eps : 0.0) + 1e-12;
  /* Groups run on OpenMP threads; source-side and node adjoints are summed in
   * per-thread buffers and reduced at the end (fp64, order-independent up to
   * rounding). */
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  double* adjX = NULL;
  double* adjM = NULL;
  double* adjS = NULL;
  double* t_adj = NULL; /* per thread: 10 per node, then 4 per source */
  const size_t per_thread = (size_t)nn * 10 + (size_t)n * 4;
  if (cot) {
    adjX = (double*)calloc((size_t)nn * 3, sizeof(double));
    adjM = (double*)calloc((size_t)nn, sizeof(double));
    adjS = (double*)calloc((size_t)nn * 6, sizeof(double));
    t_adj = (double*)calloc(per_thread * (size_t)nthreads, sizeof(double));
    memset(gq, 0, sizeof(double) * 3 * (size_t)nq);
    memset(gsx, 0, sizeof(double) * 3 * (size_t)n);
    memset(gsm, 0, sizeof(double) * (size_t)n);
  }
  memset(force, 0, sizeof(double) * 3 * (size_t)nq);
  int64_t nfar = 0, nnear = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nfar, nnear)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* my = cot ? t_adj + per_thread * (size_t)tid : NULL;
    double* myX = my;                        /* 3 per node */
    double* myM = my ? my + 3 * nn : NULL;   /* 1 per node */
    double* myS = my ? my + 4 * nn : NULL;   /* 6 per node */
    double* myGx = my ? my + 10 * nn : NULL; /* 3 per source */
    double* myGm = my ? my + 10 * nn + 3 * n : NULL;
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 8));
    int32_t* far = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
    int32_t* near = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
#pragma omp for schedule(dynamic, 8)
    for (int64_t g = g_begin; g < g_end; g++) {
      const float* box = boxes + 6 * g;
      int64_t nf = 0, nl = 0, top = 0;
      stack[top++] = 0;
      while (top > 0) {
        int32_t k = stack[--top];
        if (mass[k] == 0.0) continue;
        if (orc_group_accept(box, geo + 4 * k, beta)) {
          far[nf++] = k;
        } else if (nchild[k] == 0) {
          near[nl++] = k;
        } else {
          for (int32_t c = 0; c < nchild[k]; c++) stack[top++] = first_child[k] + c;
        }
      }
      for (int64_t i = g_start[g]; i < g_start[g] + g_count[g] && i < nq; i++) {
        const double* q = qx + 3 * i;
        const double* c = cot ? cot + 3 * i : NULL;
        double F[3] = {0, 0, 0}, G[3] = {0, 0, 0};
        for (int64_t f = 0; f < nf; f++) {
          int32_t k = far[f];
          double dummyM = 0.0;
          orc_far(q, mass[k], com + 3 * k, quad + 6 * k, use_quad, e2, c, F, G,
                  cot ? myX + 3 * k : NULL, cot ? myM + k : &dummyM, cot ? myS + 6 * k : NULL);
        }
        nfar += nf;
        for (int64_t b = 0; b < nl; b++) {
          int32_t k = near[b];
          for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
            if (self_mode && j == i) continue;
            const double* p = xs + 3 * j;
            double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double s = 1.0 / sqrt(r2 + e2);
            double s3 = s * s * s, s5 = s3 * s * s;
            double m = ms[j];
            for (int a = 0; a < 3; a++) F[a] += -m * s3 * d[a];
            nnear++;
            if (c) {
              double cd = c[0] * d[0] + c[1] * d[1] + c[2] * d[2];
              for (int a = 0; a < 3; a++) {
                double v = m * (3.0 * s5 * cd * d[a] - s3 * c[a]);
                G[a] += v;
                myGx[3 * j + a] -= v;
              }
              myGm[j] += -s3 * cd;
            }
          }
        }
        for (int a = 0; a < 3; a++) force[3 * i + a] = F[a];
        if (c)
          for (int a = 0; a < 3; a++) gq[3 * i + a] = G[a];
      }
    }
    free(stack);
    free(far);
    free(near);
  }
  if (cot) {
    for (int t = 0; t < nthreads; t++) {
      const double* my = t_adj + per_thread * (size_t)t;
      for (int64_t k = 0; k < nn; k++) {
        for (int a = 0; a < 3; a++) adjX[3 * k + a] += my[3 * k + a];
        adjM[k] += my[3 * nn + k];
        for (int a = 0; a < 6; a++) adjS[6 * k + a] += my[4 * nn + 6 * k + a];
      }
      for (int64_t j = 0; j < n; j++) {
        for (int a = 0; a < 3; a++) gsx[3 * j + a] += my[10 * nn + 3 * j + a];
        gsm[j] += my[10 * nn + 3 * n + j];
      }
    }
    /* Push node adjoints down to the sources of each node. */
    for (int64_t k = 0; k < nn; k++) {
      double M = mass[k];
      if (M == 0.0) continue;
      const double* X = com + 3 * k;
      const double* A = adjX + 3 * k;
      const double* S = adjS + 6 * k;
      double B = adjM[k];
      for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
        double dl[3] = {xs[3 * j] - X[0], xs[3 * j + 1] - X[1], xs[3 * j + 2] - X[2]};
        double m = ms[j];
        double Gd[3] = {S[0] * dl[0] + S[3] * dl[1] + S[4] * dl[2],
                        S[3] * dl[0] + S[1] * dl[1] + S[5] * dl[2],
                        S[4] * dl[0] + S[5] * dl[1] + S[2] * dl[2]};
        for (int a = 0; a < 3; a++) gsx[3 * j + a] += (m / M) * A[a] + 2.0 * m * Gd[a];
        gsm[j] += B + (A[0] * dl[0] + A[1] * dl[1] + A[2] * dl[2]) / M +
                  (dl[0] * Gd[0] + dl[1] * Gd[1] + dl[2] * Gd[2]);
      }
    }
    free(adjX);
    free(adjM);
    free(adjS);
    free(t_adj);
  }
  if (counts) {
    counts[0] = nfar;
    counts[1] = nnear;
  }
}

/* ------------------------------------------------------------------------- */
/* K1 on the adaptive tree: winding numbers of a dipole (oriented-area) cloud */
/* ------------------------------------------------------------------------- */
/* Per node (sorted sources xs, unit normals ns, areas as): area A,
 * area-weighted centroid c = sum a p / A and dipole D = sum a n -- the
 * reference's node record (A, D / A) at c (src/sbh_ffi.cu:1840-1868 with
 * normals), in fp64. */
ORC_EXPORT void orc_wn_moments(const int32_t* start, const int32_t* count, int64_t nn,
                               const double* xs, const double* ns, const double* as,
                               double* area, double* cen, double* dip) {
  for (int64_t k = 0; k < nn; k++) {
    double A = 0, C[3] = {0, 0, 0}, D[3] = {0, 0, 0};
    for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
      A += as[j];
      for (int a = 0; a < 3; a++) {
        C[a] += as[j] * xs[3 * j + a];
        D[a] += as[j] * ns[3 * j + a];
      }
    }
    if (A != 0.0)
      for (int a = 0; a < 3; a++) C[a] /= A;
    area[k] = A;
    for (int a = 0; a < 3; a++) {
      cen[3 * k + a] = C[a];
      dip[3 * k + a] = D[a];
    }
  }
}

/* Far field of one node at query q: the dipole D at the centroid c,
 * (D.(c - q)) / (4 pi (|c - q|^3 + 1e-5)). */
static inline double orc_wn_far(const double* q, const double* c, const double* D) {
  double X[3] = {c[0] - q[0], c[1] - q[1], c[2] - q[2]};
  double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  return (D[0] * X[0] + D[1] * X[1] + D[2] * X[2]) / (4.0 * M_PI * (R * R * R + 1e-5));
}

/* Group walk for winding numbers: queries (sorted, fp64 copies of fp32
 * values) in groups g_start/g_count with boxes; same MAC and traversal as
 * orc_tree_eval_range.  out[i] in sorted query order; counts as there. */
ORC_EXPORT void orc_wn_eval(int64_t nn, const int32_t* start, const int32_t* count,
                            const int32_t* first_child, const int32_t* nchild, const float* geo,
                            const double* area, const double* cen, const double* dip,
                            const double* xs, const double* ns, const double* as,
                            const float* boxes, const double* qx, int64_t nq, int64_t ng,
                            const int32_t* g_start, const int32_t* g_count, float beta,
                            double* out, int64_t* counts) {
  int64_t nfar = 0, nnear = 0;
  memset(out, 0, sizeof(double) * (size_t)nq);
#pragma omp parallel reduction(+ : nfar, nnear)
  {
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 8));
    int32_t* far = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
    int32_t* near = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
#pragma omp for schedule(dynamic, 4)
    for (int64_t g = 0; g < ng; g++) {
      const float* box = boxes + 6 * g;
      int64_t nf = 0, nl = 0, top = 0;
      stack[top++] = 0;
      while (top > 0) {
        int32_t k = stack[--top];
        if (area[k] == 0.0) continue;
        if (orc_group_accept(box, geo + 4 * k, beta)) {
          far[nf++] = k;
        } else if (nchild[k] == 0) {
          near[nl++] = k;
        } else {
          for (int32_t c = 0; c < nchild[k]; c++) stack[top++] = first_child[k] + c;
        }
      }
      for (int64_t i = g_start[g]; i < g_start[g] + g_count[g] && i < nq; i++) {
        const double* q = qx + 3 * i;
        double w = 0.0;
        for (int64_t f = 0; f < nf; f++) {
          int32_t k = far[f];
          w += orc_wn_far(q, cen + 3 * k, dip + 3 * k);
        }
        nfar += nf;
        for (int64_t b = 0; b < nl; b++) {
          int32_t k = near[b];
          for (int32_t j = start[k]; j < start[k] + count[k]; j++) {
            const double* p = xs + 3 * j;
            const double* nv = ns + 3 * j;
            double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            w += as[j] * (nv[0] * d[0] + nv[1] * d[1] + nv[2] * d[2]) /
                 (4.0 * M_PI * (r * r * r + 1e-5));
            nnear++;
          }
        }
        out[i] = w;
      }
    }
    free(stack);
    free(far);
    free(near);
  }
  if (counts) {
    counts[0] = nfar;
    counts[1] = nnear;
  }
}

/* ------------------------------------------------------------------------- */
/* Stochastic BH over the adaptive tree: the multi-level prefix control        */
/* variate of src/sbh_ffi.cu:2146-2442 (forward 2146-2291, query-side VJP      */
/* 2293-2442) on the adaptive octree's nodes, with query groups sharing draws. */
/* ------------------------------------------------------------------------- */

/* Node term of one query: the node's expansion at its centre of mass
 * (gravity: monopole, + quadrupole when use_quad; kind 1 its query-side VJP)
 * or its dipole at the area-weighted centroid (kind 2, winding number). */
static void orc_stoch_node(int kind, const double* q, const double* c, int64_t k,
                           const double* mass, const double* com, const double* quad,
                           const double* dip, int use_quad, double e2, double* acc) {
  if (kind == 2) {
    acc[0] += orc_wn_far(q, com + 3 * k, dip + 3 * k);
    return;
  }
  double F[3] = {0, 0, 0}, G[3] = {0, 0, 0}, aX[3] = {0, 0, 0}, aM = 0.0;
  double aS[6] = {0, 0, 0, 0, 0, 0};
  orc_far(q, mass[k], com + 3 * k, quad + 6 * k, use_quad, e2, kind == 1 ? c : NULL, F, G, aX,
          &aM, aS);
  for (int a = 0; a < 3; a++) acc[a] += kind == 1 ? G[a] : F[a];
}

/* Exact sum of one query over the sorted sources [j0, j1). */
static void orc_stoch_leaf(int kind, const double* q, const double* c, int64_t j0, int64_t j1,
                           const double* xs, const double* ms, const double* nrm, double e2,
                           double* acc) {
  for (int64_t j = j0; j < j1; j++) {
    const double* p = xs + 3 * j;
    if (kind == 2) {
      const double* nv = nrm + 3 * j;
      double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      acc[0] += ms[j] * (nv[0] * d[0] + nv[1] * d[1] + nv[2] * d[2]) /
                (4.0 * M_PI * (r * r * r + 1e-5));
      continue;
    }
    double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double s = 1.0 / sqrt(r2 + e2);
    double s3 = s * s * s, s5 = s3 * s * s;
    if (kind == 0) {
      for (int a = 0; a < 3; a++) acc[a] += -ms[j] * s3 * d[a];
    } else {
      double cd = c[0] * d[0] + c[1] * d[1] + c[2] * d[2];
      for (int a = 0; a < 3; a++) acc[a] += ms[j] * (3.0 * s5 * cd * d[a] - s3 * c[a]);
    }
  }
}

/* Children sum minus the node's own term, and the terms evaluated for it:
 * nodes[0] node terms, nodes[1] source pairs (per query). */
static void orc_stoch_diff(int kind, const double* q, const double* c, int32_t prev,
                           const int32_t* start, const int32_t* count,
                           const int32_t* first_child, const int32_t* nchild,
                           const double* mass, const double* com, const double* quad,
                           const double* dip, const double* xs, const double* ms,
                           const double* nrm, int use_quad, double e2, double* diff,
                           int64_t* terms) {
  double ch[3] = {0, 0, 0}, pr[3] = {0, 0, 0};
  for (int32_t cc = 0; cc < nchild[prev]; cc++) {
    int32_t k = first_child[prev] + cc;
    if (nchild[k] == 0) {
      orc_stoch_leaf(kind, q, c, start[k], start[k] + count[k], xs, ms, nrm, e2, ch);
      terms[1] += count[k];
    } else {
      orc_stoch_node(kind, q, c, k, mass, com, quad, dip, use_quad, e2, ch);
      terms[0] += 1;
    }
  }
  orc_stoch_node(kind, q, c, prev, mass, com, quad, dip, use_quad, e2, pr);
  terms[0] += 1;
  for (int a = 0; a < 3; a++) diff[a] = ch[a] - pr[a];
}

/* Ratio |q - cell centre| / cell diagonal of node k, IEEE fp32. */
static inline float orc_stoch_ratio(const float* q, const float* g) {
  float dx = q[0] - g[0], dy = q[1] - g[1], dz = q[2] - g[2];
  return sqrtf(dx * dx + dy * dy + dz * dz) / g[3];
}

/* Subdomains: nodes at level 3 and leaves above it (kControlDepth = 3,
 * src/sbh_ffi.cu:2167-2192) with non-zero mass, in node order; returns their
 * count (at most nn) in subs. */
ORC_EXPORT int64_t orc_tree_subdomains(int64_t nn, const int32_t* nchild, const int32_t* level,
                                       const double* mass, int32_t* subs) {
  int64_t ns = 0;
  for (int64_t k = 0; k < nn; k++)
    if ((level[k] == 3 || (level[k] < 3 && nchild[k] == 0)) && mass[k] != 0.0)
      subs[ns++] = (int32_t)k;
  return ns;
}

/* Estimator per query q of a group (sorted queries g_start[g] + 0..g_count-1):
 *   sum over subdomains of [its node term, or its exact leaf sum]
 *   + (1/S) sum over S samples of sum over path steps of
 *       n_sub / (P * count(prev)) * (sum of children(prev) - term(prev)),
 * where the path starts at the subdomain with beta_cur = max(beta, ratio),
 * continues at a node with probability t = min(beta_cur / ratio, 1) (ratio =
 * |q - cell centre| / cell diagonal; stop when the draw exceeds t), P is the
 * product of the continuation probabilities so far, the next node is the
 * child holding the sampled source, and beta_cur = max(beta_cur, ratio).
 * The group shares every draw: the sampled source (uniform over the
 * subdomain's sorted range) and the roulette draws are
 * sbh_uniform(seed, g_start[g], subdomain node, sample, step 0 / 1, 2, ...);
 * each query keeps its own t and P.  Leaves are summed exactly (children
 * that are leaves too).  kind 0: softened force (out nq x 3), 1: its
 * query-side VJP with cot (sorted, nq x 3), 2: winding number (out nq; mass =
 * areas, com = centroids, dip = sum a n per node, nrm = unit normals).
 * counts[0] / counts[1]: node terms / source pairs evaluated by a group for
 * every one of its queries (a step is evaluated while any query of the group
 * continues), summed over the groups and multiplied by the group sizes. */
ORC_EXPORT void orc_tree_stoch(int64_t nn, const int32_t* start, const int32_t* count,
                               const int32_t* first_child, const int32_t* nchild,
                               const int32_t* level, const float* geo, const double* mass,
                               const double* com, const double* quad, const double* dip,
                               const double* xs, const double* ms, const double* nrm,
                               const float* qf, int64_t nq, int64_t ng, const int32_t* g_start,
                               const int32_t* g_count, int kind, int use_quad, float beta,
                               double eps, int samples, uint32_t seed, const double* cot,
                               double* out, int64_t* counts) {
  const double e2 = (eps > 0 ? eps * eps : 0.0) + 1e-12;
  const int width = kind == 2 ? 1 : 3;
  int32_t* subs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn + 1));
  const int64_t nsub = orc_tree_subdomains(nn, nchild, level, mass, subs);
  memset(out, 0, sizeof(double) * (size_t)nq * (size_t)width);
  int64_t c_node = 0, c_pair = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : c_node, c_pair)
  for (int64_t g = 0; g < ng; g++) {
    const int64_t i0 = g_start[g], i1 = g_start[g] + g_count[g];
    const uint32_t key = (uint32_t)g_start[g];
    int64_t wave_node = 0, wave_pair = 0; /* terms the group evaluates */
    for (int64_t sd = 0; sd < nsub; sd++) {
      const int32_t k = subs[sd];
      const int leaf = nchild[k] == 0;
      if (leaf)
        wave_pair += count[k];
      else
        wave_node += 1;
      for (int64_t i = i0; i < i1 && i < nq; i++) {
        double qd[3] = {qf[3 * i], qf[3 * i + 1], qf[3 * i + 2]};
        const double* c = cot ? cot + 3 * i : NULL;
        double acc[3] = {0, 0, 0};
        if (leaf)
          orc_stoch_leaf(kind, qd, c, start[k], start[k] + count[k], xs, ms, nrm, e2, acc);
        else
          orc_stoch_node(kind, qd, c, k, mass, com, quad, dip, use_quad, e2, acc);
        for (int a = 0; a < width; a++) out[width * i + a] += acc[a];
      }
      if (leaf || samples <= 0) continue;
      const uint32_t n_sub = (uint32_t)count[k];
      for (int s = 0; s < samples; s++) {
        const float u0 = orc_sbh_uniform(seed, key, (uint32_t)k, (uint32_t)s, 0);
        uint32_t sidx = (uint32_t)((float)n_sub * u0) + (uint32_t)start[k];
        if (sidx > (uint32_t)(start[k] + count[k] - 1)) sidx = (uint32_t)(start[k] + count[k] - 1);
        int64_t steps_max = 0;
        int64_t step_node[64], step_pair[64];
        for (int64_t i = i0; i < i1 && i < nq; i++) {
          const float* qi = qf + 3 * i;
          double qd[3] = {qi[0], qi[1], qi[2]};
          const double* c = cot ? cot + 3 * i : NULL;
          double wsum[3] = {0, 0, 0};
          int32_t cur = k;
          float ep = 1.0f;
          float cur_beta = fmaxf(beta, orc_stoch_ratio(qi, geo + 4 * (int64_t)cur));
          uint32_t step = 1;
          int64_t taken = 0;
          while (nchild[cur] > 0 && step < 64) {
            const float ratio = orc_stoch_ratio(qi, geo + 4 * (int64_t)cur);
            const float tp = fminf(cur_beta / ratio, 1.0f);
            const float pu = orc_sbh_uniform(seed, key, (uint32_t)k, (uint32_t)s, step++);
            if (pu > tp) break;
            ep *= tp;
            const int32_t prev = cur;
            int32_t nxt = first_child[prev];
            for (int32_t cc = 1; cc < nchild[prev]; cc++)
              if ((uint32_t)start[first_child[prev] + cc] <= sidx) nxt = first_child[prev] + cc;
            cur = nxt;
            cur_beta = fmaxf(cur_beta, ratio);
            double diff[3];
            int64_t terms[2] = {0, 0};
            orc_stoch_diff(kind, qd, c, prev, start, count, first_child, nchild, mass, com, quad,
                           dip, xs, ms, nrm, use_quad, e2, diff, terms);
            step_node[taken] = terms[0];
            step_pair[taken] = terms[1];
            const float w = (float)n_sub / (ep * (float)count[prev]);
            for (int a = 0; a < width; a++) wsum[a] += (double)w * diff[a];
            taken++;
          }
          if (taken > steps_max) {
            for (int64_t t = steps_max; t < taken; t++) {
              wave_node += step_node[t];
              wave_pair += step_pair[t];
            }
            steps_max = taken;
          }
          for (int a = 0; a < width; a++) out[width * i + a] += wsum[a] / (double)samples;
        }
      }
    }
    const int64_t nv = (i1 < nq ? i1 : nq) - i0;
    c_node += wave_node * nv;
    c_pair += wave_pair * nv;
  }
  free(subs);
  if (counts) {
    counts[0] = c_node;
    counts[1] = c_pair;
  }
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based RNG   */
/* the stochastic walks use in place of cuRAND XORWOW.                        */
/* ------------------------------------------------------------------------- */
ORC_EXPORT void orc_philox4x32(const uint32_t* ctr_in, const uint32_t* key_in,
                               uint32_t* out) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
