/* TEST INFRASTRUCTURE ONLY -- part of the oracle: a C restatement of the
 * reference's CPU-built trees and of the GPU walks over them, used as the
 * checker for the library's tree builder and its 2-D / multi-level walks.
 * Nothing in the product links or calls it.
 *
 * Restated:
 *   Tree2D<1, L> / Tree3D<1, L>, L = 1, 2, 3 (include/tree2d.h:237-407,
 *     include/tree3d.h:238-413; include/py_tree.h:20-73 sets max_depth 100
 *     and every point's alpha): breadth-first build from a queue; a node is a
 *     leaf when depth >= 100 or it holds <= 1 point; otherwise its points are
 *     partitioned by AABB::quantize (include/aabb.h:36-42) into the S^dim
 *     cells (S = 2^L), visited in (z,) y, x order, and each non-empty cell
 *     becomes a child with box lo + k inc, lo + k inc + inc, inc = (hi - lo) / S.
 *     Moments: sequential fp32 sums over points (leaves) or children in cell
 *     order, position and normal divided by the mass.
 *   evaluateBarnesHutGPUVote (include/barnes_hut_gpu.h:19-72): recursive form
 *     of the stackless walk for one 32-query warp.
 *   multiLevelPrefixControlVariateSampleGPUQuery
 *     (include/tree_sampler_gpu.h:25-140).
 *   Kernels (include/kernel.h:55-94): gravityPotential, gravityGrad,
 *     signedSolidAngle<N>, smoothDistExp; gravityGradSoftCutoff
 *     (src/sbh_ffi.cu:142-160).
 * Stated differences: the partition here is stable (the reference's
 * std::partition leaves the order inside a cell unspecified; it matters only
 * inside leaves at depth 100).  cuRAND XORWOW is replaced by Philox4x32-10
 * streams: the sample point from key = block index, counter (1, subdomain,
 * sample, 0); the roulette from key = seed + query (the FFI targets) or seed
 * (the pybind11 module, which passes seed 0 to every query), counter (0,
 * subdomain, sample, step).  Tree arithmetic and walk decisions are fp32 in the
 * reference's operation order (this file is built with -ffp-contract=off;
 * the sample index uses an explicit fmaf, as nvcc contracts it); the sums of
 * the walks are fp64.
 *
 * Pinning (tests/test_oracle_ndtree.py): tree invariants (cells, bounds,
 * breadth-first order, contiguity, leaf rule) against an independent numpy
 * partition, moments against fp64 sums; huge beta = the fp64 direct sum and
 * beta = 0 = the root monopole for the vote walk; for the sampler, one-level
 * trees give the exact sum and the mean over blocks (independent sample
 * points and paths) is the direct sum within its standard error; the module's
 * fixed seed equals the FFI seeding at seed = -query; the brute force is the
 * fp64 direct sum over the leaf records. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

ORC_EXPORT void orc_philox4x32(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out);

#define ND_MAX_DEPTH 100

/* Byte layout of Tree2DNode<1, L> / Tree3DNode<1, L> (include/tree2d.h:
 * 212-225, include/tree3d.h:212-225) under the Itanium C++ ABI: members in
 * declaration order, each at its natural alignment, size rounded to 8. */
typedef struct {
  int dim, log, split, cells, words;
  size_t o_bmin, o_bmax, o_diam, o_contrib, o_mask, o_child, o_num, o_depth, o_maxd, o_leaf;
  size_t size, leaf_size;
} NdLayout;

static size_t nd_align(size_t x, size_t a) { return (x + a - 1) / a * a; }

static NdLayout nd_layout(int dim, int log) {
  NdLayout L;
  L.dim = dim;
  L.log = log;
  L.split = 1 << log;
  L.cells = 1 << (dim * log);
  L.words = L.cells >= 64 ? L.cells / 64 : 1;
  L.leaf_size = (size_t)(2 * dim + 2) * 4;
  size_t o = 32; /* start, end, buffer_idx, parent_idx */
  L.o_bmin = o;
  o += 4 * (size_t)dim;
  L.o_bmax = o;
  o += 4 * (size_t)dim;
  L.o_diam = o;
  o += 4;
  L.o_contrib = o;
  o += L.leaf_size;
  o = nd_align(o, 8);
  L.o_mask = o;
  o += 8 * (size_t)L.words;
  L.o_child = o;
  o += 8 * (size_t)L.cells;
  L.o_num = o;
  L.o_depth = o + 4;
  L.o_maxd = o + 8;
  L.o_leaf = o + 12;
  L.size = nd_align(o + 13, 8);
  return L;
}

ORC_EXPORT size_t orc_node2_size(void) { return nd_layout(2, 3).size; }
ORC_EXPORT size_t orc_nd_node_size(int dim, int log) { return nd_layout(dim, log).size; }

/* Generic in-memory node (converted to the byte layouts on output). */
typedef struct {
  int64_t start, end, parent, first;
  int32_t depth, num, leaf;
  float lo[3], hi[3];
  uint64_t mask[8];
  float m, pos[3], nrm[3], alpha;
} NdNode;

typedef struct {
  int dim, log;
  int64_t n, nn, cap;
  float* pts; /* tree order, dim per point */
  float* nrm;
  float* ms;
  float alpha;
  NdNode* nodes;
} NdTree;

/* AABB::quantize as the x86 build evaluates it: (int) of S (p - lo) / (hi -
 * lo) (INT_MIN for NaN or out of range), then min(., S - 1), max(., 0). */
static int nd_quantize(float p, float lo, float hi, int S) {
  float t = (float)S * (p - lo) / (hi - lo);
  int c = (t >= -2147483648.0f && t < 2147483648.0f) ? (int)t : INT32_MIN;
  if (c > S - 1) c = S - 1;
  if (c < 0) c = 0;
  return c;
}

static int nd_cell(const float* p, const float* lo, const float* hi, int dim, int S) {
  int c = 0;
  for (int a = dim - 1; a >= 0; a--) c = c * S + nd_quantize(p[a], lo[a], hi[a], S);
  return c;
}

static int64_t nd_new_node(NdTree* t) {
  if (t->nn == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 256;
    t->nodes = (NdNode*)realloc(t->nodes, sizeof(NdNode) * (size_t)t->cap);
  }
  memset(&t->nodes[t->nn], 0, sizeof(NdNode));
  return t->nn++;
}

ORC_EXPORT void orc_nd_free(NdTree* t) {
  if (!t) return;
  free(t->pts);
  free(t->nrm);
  free(t->ms);
  free(t->nodes);
  free(t);
}

ORC_EXPORT NdTree* orc_nd_build(int dim, int log, const float* pts, const float* nrm,
                                const float* ms, int64_t n, const float* lo, const float* hi,
                                float alpha) {
  NdTree* t = (NdTree*)calloc(1, sizeof(NdTree));
  const int S = 1 << log;
  const int cells = 1 << (dim * log);
  t->dim = dim;
  t->log = log;
  t->n = n;
  t->alpha = alpha;
  size_t np = (size_t)(n > 0 ? n : 1);
  t->pts = (float*)malloc(sizeof(float) * dim * np);
  t->nrm = (float*)malloc(sizeof(float) * dim * np);
  t->ms = (float*)malloc(sizeof(float) * np);
  memcpy(t->pts, pts, sizeof(float) * dim * (size_t)n);
  memcpy(t->nrm, nrm, sizeof(float) * dim * (size_t)n);
  memcpy(t->ms, ms, sizeof(float) * (size_t)n);
  float* tp = (float*)malloc(sizeof(float) * dim * np);
  float* tn = (float*)malloc(sizeof(float) * dim * np);
  float* tm = (float*)malloc(sizeof(float) * np);
  int* cell = (int*)malloc(sizeof(int) * np);
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (cells + 1));
  int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * cells);
  int64_t r = nd_new_node(t);
  NdNode* root = &t->nodes[r];
  root->start = 0;
  root->end = n;
  root->parent = 0;
  for (int a = 0; a < dim; a++) {
    root->lo[a] = lo[a];
    root->hi[a] = hi[a];
  }
  /* nodes are created in breadth-first order, so a sweep over ids is the
   * reference's queue */
  for (int64_t id = 0; id < t->nn; id++) {
    NdNode* nd = &t->nodes[id];
    nd->leaf = (nd->depth >= ND_MAX_DEPTH || nd->end - nd->start <= 1) ? 1 : 0;
    if (nd->leaf) continue;
    const int64_t s0 = nd->start, s1 = nd->end;
    memset(cnt, 0, sizeof(int64_t) * (cells + 1));
    for (int64_t j = s0; j < s1; j++) {
      cell[j - s0] = nd_cell(&t->pts[dim * j], nd->lo, nd->hi, dim, S);
      cnt[cell[j - s0] + 1]++;
    }
    for (int c = 0; c < cells; c++) cnt[c + 1] += cnt[c];
    /* stable scatter by cell */
    for (int c = 0; c < cells; c++) cur[c] = cnt[c];
    for (int64_t j = s0; j < s1; j++) {
      int64_t d = s0 + cur[cell[j - s0]]++;
      for (int a = 0; a < dim; a++) {
        tp[dim * d + a] = t->pts[dim * j + a];
        tn[dim * d + a] = t->nrm[dim * j + a];
      }
      tm[d] = t->ms[j];
    }
    memcpy(&t->pts[dim * s0], &tp[dim * s0], sizeof(float) * dim * (size_t)(s1 - s0));
    memcpy(&t->nrm[dim * s0], &tn[dim * s0], sizeof(float) * dim * (size_t)(s1 - s0));
    memcpy(&t->ms[s0], &tm[s0], sizeof(float) * (size_t)(s1 - s0));
    float inc[3], plo[3];
    int32_t depth = nd->depth;
    for (int a = 0; a < dim; a++) {
      inc[a] = (nd->hi[a] - nd->lo[a]) / (float)S;
      plo[a] = nd->lo[a];
    }
    int64_t first = -1;
    int num = 0;
    uint64_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < cells; c++) {
      if (cnt[c + 1] == cnt[c]) continue;
      int64_t ch = nd_new_node(t); /* may move t->nodes */
      NdNode* cn = &t->nodes[ch];
      cn->start = s0 + cnt[c];
      cn->end = s0 + cnt[c + 1];
      cn->parent = id;
      cn->depth = depth + 1;
      int rem = c;
      for (int a = 0; a < dim; a++) {
        int k = rem % S;
        rem /= S;
        cn->lo[a] = plo[a] + (float)k * inc[a];
        cn->hi[a] = cn->lo[a] + inc[a];
      }
      if (first < 0) first = ch;
      num++;
      mask[c >> 6] |= 1ull << (c & 63);
    }
    nd = &t->nodes[id];
    nd->first = first;
    nd->num = num;
    memcpy(nd->mask, mask, sizeof(mask));
  }
  /* moments, bottom-up (children have larger ids) */
  for (int64_t id = t->nn - 1; id >= 0; id--) {
    NdNode* nd = &t->nodes[id];
    float m = 0.0f, al = 0.0f, p[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    if (nd->leaf) {
      for (int64_t j = nd->start; j < nd->end; j++) {
        float mj = t->ms[j];
        al = alpha;
        m = m + mj;
        for (int a = 0; a < dim; a++) {
          p[a] = p[a] + mj * t->pts[dim * j + a];
          q[a] = q[a] + mj * t->nrm[dim * j + a];
        }
      }
    } else {
      for (int o = 0; o < nd->num; o++) {
        const NdNode* cn = &t->nodes[nd->first + o];
        al = cn->alpha;
        m = m + cn->m;
        for (int a = 0; a < dim; a++) {
          p[a] = p[a] + cn->m * cn->pos[a];
          q[a] = q[a] + cn->m * cn->nrm[a];
        }
      }
    }
    nd->m = m;
    nd->alpha = al;
    for (int a = 0; a < dim; a++) {
      nd->pos[a] = p[a] / m;
      nd->nrm[a] = q[a] / m;
    }
  }
  free(tp);
  free(tn);
  free(tm);
  free(cell);
  free(cnt);
  free(cur);
  return t;
}

ORC_EXPORT int64_t orc_nd_num_nodes(const NdTree* t) { return t->nn; }

static uint64_t nd_child_id(const NdNode* nd, int c) {
  if (!((nd->mask[c >> 6] >> (c & 63)) & 1ull)) return 0;
  int rank = 0;
  for (int w = 0; w < (c >> 6); w++) rank += __builtin_popcountll(nd->mask[w]);
  rank += __builtin_popcountll(nd->mask[c >> 6] & ((1ull << (c & 63)) - 1ull));
  return (uint64_t)(nd->first + rank);
}

static float nd_diameter(const NdNode* nd, int dim) {
  float s = 0.0f;
  for (int a = 0; a < dim; a++) {
    float d = nd->hi[a] - nd->lo[a];
    s = a == 0 ? d * d : s + d * d;
  }
  return sqrtf(s);
}

static void nd_put_leaf(uint8_t* dst, int dim, const float* pos, const float* nrm, float m,
                        float alpha) {
  float v[8];
  for (int a = 0; a < dim; a++) {
    v[a] = pos[a];
    v[dim + a] = nrm[a];
  }
  v[2 * dim] = m;
  v[2 * dim + 1] = alpha;
  memcpy(dst, v, sizeof(float) * (size_t)(2 * dim + 2));
}

/* Byte output as the reference's leaf_data_bytes / node_buffer_bytes /
 * contrib_data_bytes (src/sbh.cpp:227-235); padding bytes are zero. */
ORC_EXPORT void orc_nd_emit(const NdTree* t, uint8_t* leaf_out, uint8_t* node_out,
                            uint8_t* contrib_out) {
  const int dim = t->dim;
  const NdLayout L = nd_layout(dim, t->log);
  for (int64_t j = 0; j < t->n; j++)
    nd_put_leaf(leaf_out + L.leaf_size * (size_t)j, dim, &t->pts[dim * j], &t->nrm[dim * j],
                t->ms[j], t->alpha);
  memset(node_out, 0, L.size * (size_t)t->nn);
  for (int64_t id = 0; id < t->nn; id++) {
    const NdNode* nd = &t->nodes[id];
    uint8_t* o = node_out + L.size * (size_t)id;
    uint64_t u[4] = {(uint64_t)nd->start, (uint64_t)nd->end, (uint64_t)id, (uint64_t)nd->parent};
    memcpy(o, u, sizeof(u));
    memcpy(o + L.o_bmin, nd->lo, 4 * (size_t)dim);
    memcpy(o + L.o_bmax, nd->hi, 4 * (size_t)dim);
    float diam = nd_diameter(nd, dim);
    memcpy(o + L.o_diam, &diam, 4);
    nd_put_leaf(o + L.o_contrib, dim, nd->pos, nd->nrm, nd->m, nd->alpha);
    memcpy(o + L.o_mask, nd->mask, 8 * (size_t)L.words);
    for (int c = 0; c < L.cells; c++) {
      uint64_t ch = nd_child_id(nd, c);
      memcpy(o + L.o_child + 8 * (size_t)c, &ch, 8);
    }
    int32_t iv[3] = {nd->num, nd->depth, ND_MAX_DEPTH};
    memcpy(o + L.o_num, iv, sizeof(iv));
    o[L.o_leaf] = (uint8_t)nd->leaf;
    if (contrib_out)
      memcpy(contrib_out + L.leaf_size * (size_t)id, o + L.o_contrib, L.leaf_size);
  }
}

/* ------------------------------------------------------------------ walks */
typedef struct {
  const uint8_t* base;
  NdLayout L;
} NdView;

typedef struct {
  uint64_t start, end, parent;
  float bmin[3], bmax[3], pos[3], nrm[3];
  float diameter, mass, alpha;
  const uint8_t* mask;
  const uint8_t* child;
  int32_t num;
  int leaf;
} NdRec;

static uint64_t nd_u64(const uint8_t* p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}

static NdRec nd_rec(const NdView* v, uint64_t id) {
  const NdLayout* L = &v->L;
  const uint8_t* o = v->base + L->size * id;
  NdRec r;
  float c[8];
  r.start = nd_u64(o);
  r.end = nd_u64(o + 8);
  r.parent = nd_u64(o + 24);
  memcpy(r.bmin, o + L->o_bmin, 4 * (size_t)L->dim);
  memcpy(r.bmax, o + L->o_bmax, 4 * (size_t)L->dim);
  memcpy(&r.diameter, o + L->o_diam, 4);
  memcpy(c, o + L->o_contrib, L->leaf_size);
  for (int a = 0; a < L->dim; a++) {
    r.pos[a] = c[a];
    r.nrm[a] = c[L->dim + a];
  }
  r.mass = c[2 * L->dim];
  r.alpha = c[2 * L->dim + 1];
  r.mask = o + L->o_mask;
  r.child = o + L->o_child;
  memcpy(&r.num, o + L->o_num, 4);
  r.leaf = o[L->o_leaf];
  return r;
}

static int nd_has_cell(const NdRec* r, int c) {
  return (int)((nd_u64(r->mask + 8 * (size_t)(c >> 6)) >> (c & 63)) & 1ull);
}
static uint64_t nd_child_at(const NdRec* r, int c) { return nd_u64(r->child + 8 * (size_t)c); }

typedef struct {
  int kind; /* 0 gravityPotential, 1 gravityGrad, 2 gravityGradSoftCutoff, 3 signedSolidAngle, 4 smoothDistExp */
  double eps, rc;
  int cutoff, periodic;
  double box;
} NdParams;

static double nd_mi(double d, const NdParams* P) {
  if (!P->periodic || !(P->box > 0.0)) return d;
  return d - P->box * floor(d / P->box + 0.5);
}

/* Kernels (include/kernel.h:55-94; src/sbh_ffi.cu:142-160) in fp64 for a
 * record at p with normal nv, mass m and alpha al. */
static void nd_term(const NdParams* P, int dim, const double* q, const float* p,
                    const float* nv, float m, float al, double* acc) {
  double d[3], r2 = 0.0;
  if (P->kind == 0 || P->kind == 3 || P->kind == 4) {
    double dn = 0.0;
    for (int a = 0; a < dim; a++) {
      d[a] = nd_mi((double)p[a] - q[a], P);
      r2 += d[a] * d[a];
      dn += (double)nv[a] * d[a];
    }
    if (P->kind == 0) {
      acc[0] += -(double)m / (sqrt(r2) + 0.1);
    } else if (P->kind == 3) {
      if (dim == 2)
        acc[0] += (double)m * dn / (2.0 * M_PI * (r2 + 1e-5));
      else
        acc[0] += (double)m * dn / (4.0 * M_PI * (r2 * sqrt(r2) + 1e-5));
    } else {
      acc[0] += exp(-(double)al * sqrt(r2));
    }
    return;
  }
  for (int a = 0; a < dim; a++) {
    d[a] = nd_mi(q[a] - (double)p[a], P);
    r2 += d[a] * d[a];
  }
  double c;
  if (P->kind == 1) {
    double t = sqrt(r2) + 0.1;
    c = -(double)m / (t * t * t);
  } else {
    double s = 1.0 / sqrt(r2 + P->eps * P->eps + 1e-12);
    double att = 1.0;
    if (P->cutoff) {
      double u = sqrt(r2 + 1e-12) / (2.0 * P->rc);
      att = erfc(u) + (2.0 / sqrt(M_PI)) * u * exp(-u * u);
    }
    c = -(double)m * att * s * s * s;
  }
  for (int a = 0; a < dim; a++) acc[a] += c * d[a];
}

static void nd_leaf_term(const NdParams* P, const NdLayout* L, const uint8_t* leaf, uint64_t j,
                         const double* q, double* acc) {
  float c[8];
  memcpy(c, leaf + L->leaf_size * j, L->leaf_size);
  nd_term(P, L->dim, q, c, c + L->dim, c[2 * L->dim], c[2 * L->dim + 1], acc);
}

static void nd_leaf_sum(const uint8_t* leaf, uint64_t nl, const NdLayout* L, const NdRec* r,
                        const NdParams* P, const double* q, double* acc) {
  uint64_t je = r->end < nl ? r->end : nl;
  for (uint64_t j = r->start; j < je; j++) nd_leaf_term(P, L, leaf, j, q, acc);
}

static void nd_rec_term(const NdParams* P, int dim, const NdRec* r, const double* q,
                        double* acc) {
  nd_term(P, dim, q, r->pos, r->nrm, r->mass, r->alpha, acc);
}

/* farFieldRatio (include/tree2d.h:132-138), fp32. */
static float nd_ratio(const NdRec* r, const float* q, int dim) {
  if (r->leaf) return INFINITY;
  float s = 0.0f;
  for (int a = 0; a < dim; a++) {
    float d = q[a] - r->pos[a];
    s = a == 0 ? d * d : s + d * d;
  }
  return sqrtf(s) / r->diameter;
}

static void nd_vote(const uint8_t* leaf, uint64_t nl, const NdView* v, uint64_t nn, uint64_t id,
                    const float* qf, int64_t g0, int64_t g1, float beta, const NdParams* P,
                    double* out, int vec) {
  const int dim = v->L.dim;
  NdRec r = nd_rec(v, id);
  int all = 1;
  for (int64_t i = g0; i < g1 && all; i++)
    if (!(nd_ratio(&r, &qf[dim * i], dim) >= beta)) all = 0;
  if (all) {
    for (int64_t i = g0; i < g1; i++) {
      double q[3];
      for (int a = 0; a < dim; a++) q[a] = qf[dim * i + a];
      double* o = out + (vec ? dim * i : i);
      if (r.leaf)
        nd_leaf_sum(leaf, nl, &v->L, &r, P, q, o);
      else
        nd_rec_term(P, dim, &r, q, o);
    }
    return;
  }
  for (int c = 0; c < v->L.cells; c++)
    if (nd_has_cell(&r, c)) {
      uint64_t ch = nd_child_at(&r, c);
      if (ch < nn) nd_vote(leaf, nl, v, nn, ch, qf, g0, g1, beta, P, out, vec);
    }
}

static inline float nd_uniform(uint32_t key, uint32_t tag, uint32_t node, uint32_t sample,
                               uint32_t step) {
  uint32_t ctr[4] = {tag, node, sample, step};
  uint32_t k[2] = {key, 0x5BD1E995u};
  uint32_t o[4];
  orc_philox4x32(ctr, k, o);
  return (float)((o[0] >> 8) + 1u) * (1.0f / 16777216.0f);
}

/* AABB::contains with the 1e-4 margin, then AABB::quantize as nvcc's
 * saturating conversion; -1 outside. */
static int nd_cell_in(const NdRec* r, const float* p, int dim, int S) {
  int c = 0;
  int qa[3];
  for (int a = 0; a < dim; a++) {
    if (!(r->bmin[a] - 1e-4f <= p[a] && p[a] <= r->bmax[a] + 1e-4f)) return -1;
    float t = (float)S * (p[a] - r->bmin[a]) / (r->bmax[a] - r->bmin[a]);
    int k = isnan(t) ? 0
                     : (t >= 2147483647.0f ? 2147483647
                                           : (t <= -2147483648.0f ? INT32_MIN : (int)t));
    if (k < 0) k = 0;
    if (k > S - 1) k = S - 1;
    qa[a] = k;
  }
  for (int a = dim - 1; a >= 0; a--) c = c * S + qa[a];
  return c;
}

static int64_t nd_contained(const NdView* v, uint64_t nn, uint64_t id, const float* p) {
  for (int hop = 0; hop < 128; hop++) {
    NdRec r = nd_rec(v, id);
    int c = nd_cell_in(&r, p, v->L.dim, v->L.split);
    if (c < 0) return -1;
    uint64_t ch = nd_child_at(&r, c);
    if (ch >= nn) return -1;
    NdRec cr = nd_rec(v, ch);
    if (cr.num != 1) return (int64_t)ch;
    id = ch;
  }
  return -1;
}

/* One query of the sampler; `block` is the query's block index. */
static void nd_mlpcv(const uint8_t* leaf, uint64_t nl, const NdView* v, uint64_t nn,
                     const uint8_t* contrib, uint64_t ncontrib, const float* qf, int64_t qi,
                     uint32_t block, int samples, int seed, int seed_per_query,
                     const NdParams* P, double* out) {
  const NdLayout* L = &v->L;
  const int dim = L->dim;
  double q[3] = {0, 0, 0}, bh[3] = {0, 0, 0}, ml[3] = {0, 0, 0};
  for (int a = 0; a < dim; a++) q[a] = qf[a];
  const uint32_t rseed = (uint32_t)seed + (seed_per_query ? (uint32_t)qi : 0u);
  NdRec root = nd_rec(v, 0);
  for (int b = 0; b < L->cells; b++) {
    if (!nd_has_cell(&root, b)) continue;
    uint64_t ci = nd_child_at(&root, b);
    if (ci >= nn) continue;
    NdRec sub = nd_rec(v, ci);
    if (sub.leaf)
      nd_leaf_sum(leaf, nl, L, &sub, P, q, bh);
    else
      nd_rec_term(P, dim, &sub, q, bh);
    uint64_t count = sub.end - sub.start;
    double total[3] = {0, 0, 0};
    for (int ss = 0; ss < samples; ss++) {
      float u2 = nd_uniform(block, 1u, (uint32_t)ci, (uint32_t)ss, 0u);
      uint64_t idx = (uint64_t)fmaf((float)count, u2, (float)sub.start);
      if (idx > sub.end - 1) idx = sub.end - 1;
      if (count == 0 || idx >= nl) continue;
      float sp[8];
      memcpy(sp, leaf + L->leaf_size * idx, L->leaf_size);
      uint64_t cur = ci;
      float event_prob = 1.0f;
      double w_all[3] = {0, 0, 0};
      float cur_beta = fmaxf(1.0f, nd_ratio(&sub, qf, dim));
      for (uint32_t step = 0; step < 1024u; step++) {
        NdRec cn = nd_rec(v, cur);
        if (cn.leaf) break;
        float ffr = nd_ratio(&cn, qf, dim);
        float tp = fminf(cur_beta / ffr, 1.0f);
        float pu = nd_uniform(rseed, 0u, (uint32_t)ci, (uint32_t)ss, step);
        if (pu > tp) break;
        event_prob *= tp;
        int64_t nxt = nd_contained(v, nn, cur, sp);
        cur_beta = fmaxf(cur_beta, ffr);
        double prev_res[3] = {0, 0, 0}, cur_res[3] = {0, 0, 0};
        nd_rec_term(P, dim, &cn, q, prev_res);
        uint64_t first = 0;
        for (int c = 0; c < L->cells; c++)
          if (nd_has_cell(&cn, c)) {
            first = nd_child_at(&cn, c);
            break;
          }
        for (int32_t o = 0; o < cn.num; o++) {
          uint64_t c = first + (uint64_t)o;
          if (c >= nn) break;
          NdRec ch = nd_rec(v, c);
          if (ch.leaf)
            nd_leaf_sum(leaf, nl, L, &ch, P, q, cur_res);
          else if (c < ncontrib)
            nd_leaf_term(P, L, contrib, c, q, cur_res);
          else
            nd_rec_term(P, dim, &ch, q, cur_res);
        }
        double w = (double)((float)count / (event_prob * (float)(cn.end - cn.start)));
        for (int a = 0; a < dim; a++) w_all[a] += w * (cur_res[a] - prev_res[a]);
        if (nxt < 0) break;
        cur = (uint64_t)nxt;
      }
      for (int a = 0; a < dim; a++) total[a] += w_all[a];
    }
    for (int a = 0; a < dim; a++) ml[a] += total[a] / (double)samples;
  }
  for (int a = 0; a < dim; a++) out[a] = bh[a] + ml[a];
}

/* Classic vote walk (samples < 0) or the sampler (samples >= 0) over node
 * records Tree{2,3}DNode<1, log>; kinds 0, 3, 4 write out[i], kinds 1, 2
 * out[dim i + a].  Reference warps: 32 consecutive queries of a block of
 * `threads`. */
ORC_EXPORT void orc_nd_eval(int dim, int log, const uint8_t* leaf_bytes, int64_t nl,
                            const uint8_t* node_bytes, int64_t nn, const uint8_t* contrib_bytes,
                            int64_t ncontrib, const float* queries, int64_t nq, int kind,
                            float beta, int threads, int samples, int seed, int seed_per_query,
                            double eps, double rc, int periodic, double box, double* out) {
  NdView v = {node_bytes, nd_layout(dim, log)};
  NdParams P;
  P.kind = kind;
  P.eps = eps > 0.0 ? eps : 0.0;
  P.rc = rc;
  P.cutoff = rc > 0.0 && isfinite(rc);
  P.periodic = periodic != 0 && box > 0.0;
  P.box = box;
  const int vec = kind == 1 || kind == 2;
  memset(out, 0, sizeof(double) * (size_t)(vec ? dim * nq : nq));
  if (nn <= 0) return;
  if (samples < 0) {
    for (int64_t b0 = 0; b0 < nq; b0 += threads)
      for (int64_t w0 = b0; w0 < b0 + threads && w0 < nq; w0 += 32) {
        int64_t w1 = w0 + 32;
        if (w1 > b0 + threads) w1 = b0 + threads;
        if (w1 > nq) w1 = nq;
        nd_vote(leaf_bytes, (uint64_t)nl, &v, (uint64_t)nn, 0, queries, w0, w1, beta, &P, out,
                vec);
      }
    return;
  }
  for (int64_t i = 0; i < nq; i++) {
    double o[3] = {0, 0, 0};
    nd_mlpcv(leaf_bytes, (uint64_t)nl, &v, (uint64_t)nn, contrib_bytes, (uint64_t)ncontrib,
             &queries[dim * i], i, (uint32_t)(i / threads), samples, seed, seed_per_query, &P,
             o);
    if (vec)
      for (int a = 0; a < dim; a++) out[dim * i + a] = o[a];
    else
      out[i] = o[0];
  }
}

/* The pybind11 module's brute force (src/sbh_gpu.cu:39-77): the sum of a
 * scalar kernel (0 potential, 3 winding number, 4 smooth distance) over the
 * leaf records, here in fp64. */
ORC_EXPORT void orc_nd_brute(int dim, const uint8_t* leaf_bytes, int64_t nl,
                             const float* queries, int64_t nq, int kind, double* out) {
  NdLayout L = nd_layout(dim, 3);
  NdParams P;
  memset(&P, 0, sizeof(P));
  P.kind = kind;
  for (int64_t i = 0; i < nq; i++) {
    double q[3] = {0, 0, 0}, acc[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) q[a] = queries[dim * i + a];
    for (int64_t j = 0; j < nl; j++) nd_leaf_term(&P, &L, leaf_bytes, (uint64_t)j, q, acc);
    out[i] = acc[0];
  }
}
