/* TEST INFRASTRUCTURE ONLY -- part of the oracle: a C restatement of the
 * reference's CPU-built trees and of the GPU walks over them, used as the
 * checker for the library's tree builder and its 2-D / multi-level walks.
 * Nothing in the product links or calls it.
 *
 * Restated:
 *   Tree2D<1, 3> / Tree3D<1, 3> (include/tree2d.h:237-407,
 *     include/tree3d.h:238-413; include/py_tree.h:20-73 sets max_depth 100
 *     and every point's alpha): breadth-first build from a queue; a node is a
 *     leaf when depth >= 100 or it holds <= 1 point; otherwise its points are
 *     partitioned by AABB::quantize (include/aabb.h:36-42) into the 8^dim
 *     cells, visited in (z,) y, x order, and each non-empty cell becomes a
 *     child with box lo + k inc, lo + k inc + inc, inc = (hi - lo) / 8.
 *     Moments: sequential fp32 sums over points (leaves) or children in cell
 *     order, position and normal divided by the mass.
 *   evaluateBarnesHutGPUVote (include/barnes_hut_gpu.h:19-72): recursive form
 *     of the stackless walk for one 32-query warp.
 *   multiLevelPrefixControlVariateSampleGPUQuery
 *     (include/tree_sampler_gpu.h:25-140).
 * Stated differences: the partition here is stable (the reference's
 * std::partition leaves the order inside a cell unspecified; it matters only
 * inside leaves at depth 100).  cuRAND XORWOW is replaced by Philox4x32-10
 * streams: the sample point from key = block index, counter (1, subdomain,
 * sample, 0); the roulette from key = seed + query, counter (0, subdomain,
 * sample, step).  Tree arithmetic and walk decisions are fp32 in the
 * reference's operation order (this file is built with -ffp-contract=off;
 * the sample index uses an explicit fmaf, as nvcc contracts it); the sums of
 * the walks are fp64.
 *
 * Pinning (tests/test_oracle_ndtree.py): tree invariants (cells, bounds,
 * breadth-first order, contiguity, leaf rule) against an independent numpy
 * partition, moments against fp64 sums; huge beta = the fp64 direct sum and
 * beta = 0 = the root monopole for the vote walk; for the sampler, one-level
 * trees give the exact sum and the mean over blocks (independent sample
 * points and paths) is the direct sum within its standard error. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

ORC_EXPORT void orc_philox4x32(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out);

typedef struct {
  float pos[2], normal[2], mass, alpha;
} NdLeaf2; /* LeafData<2>: 24 bytes */

typedef struct {
  float pos[3], normal[3], mass, alpha;
} NdLeaf3; /* LeafData<3>: 32 bytes */

typedef struct {
  uint64_t start, end, buffer_idx, parent_idx;
  float bmin[2], bmax[2];
  float diameter;
  NdLeaf2 contrib;
  uint64_t bitmask[1];
  uint64_t child[64];
  int32_t num_children, depth, max_depth;
  uint8_t is_leaf, pad[3];
} NdNode2; /* Tree2DNode<1, 3>: 616 bytes */

typedef struct {
  uint64_t start, end, buffer_idx, parent_idx;
  float bmin[3], bmax[3];
  float diameter;
  NdLeaf3 contrib;
  uint64_t bitmask[8];
  uint64_t child[512];
  int32_t num_children, depth, max_depth;
  uint8_t is_leaf, pad[3];
} NdNode3; /* Tree3DNode<1, 3>: 4272 bytes */

_Static_assert(sizeof(NdLeaf2) == 24, "LeafData<2>");
_Static_assert(sizeof(NdNode2) == 616, "Tree2DNode<1,3>");
_Static_assert(sizeof(NdNode3) == 4272, "Tree3DNode<1,3>");

ORC_EXPORT size_t orc_node2_size(void) { return sizeof(NdNode2); }

#define ND_MAX_DEPTH 100

/* Generic in-memory node (converted to the byte layouts on output). */
typedef struct {
  int64_t start, end, parent, first;
  int32_t depth, num, leaf;
  float lo[3], hi[3];
  uint64_t mask[8];
  float m, pos[3], nrm[3], alpha;
} NdNode;

typedef struct {
  int dim;
  int64_t n, nn, cap;
  float* pts; /* tree order, dim per point */
  float* nrm;
  float* ms;
  float alpha;
  NdNode* nodes;
} NdTree;

/* AABB::quantize as the x86 build evaluates it: (int) of 8 (p - lo) / (hi -
 * lo) (INT_MIN for NaN or out of range), then min(., 7), max(., 0). */
static int nd_quantize(float p, float lo, float hi) {
  float t = 8.0f * (p - lo) / (hi - lo);
  int c = (t >= -2147483648.0f && t < 2147483648.0f) ? (int)t : INT32_MIN;
  if (c > 7) c = 7;
  if (c < 0) c = 0;
  return c;
}

static int nd_cell(const float* p, const float* lo, const float* hi, int dim) {
  int c = 0;
  for (int a = dim - 1; a >= 0; a--) c = c * 8 + nd_quantize(p[a], lo[a], hi[a]);
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

ORC_EXPORT NdTree* orc_nd_build(int dim, const float* pts, const float* nrm, const float* ms,
                                int64_t n, const float* lo, const float* hi, float alpha) {
  NdTree* t = (NdTree*)calloc(1, sizeof(NdTree));
  const int cells = dim == 2 ? 64 : 512;
  t->dim = dim;
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
      cell[j - s0] = nd_cell(&t->pts[dim * j], nd->lo, nd->hi, dim);
      cnt[cell[j - s0] + 1]++;
    }
    for (int c = 0; c < cells; c++) cnt[c + 1] += cnt[c];
    /* stable scatter by cell */
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * cells);
    for (int c = 0; c < cells; c++) cur[c] = cnt[c];
    for (int64_t j = s0; j < s1; j++) {
      int64_t d = s0 + cur[cell[j - s0]]++;
      for (int a = 0; a < dim; a++) {
        tp[dim * d + a] = t->pts[dim * j + a];
        tn[dim * d + a] = t->nrm[dim * j + a];
      }
      tm[d] = t->ms[j];
    }
    free(cur);
    memcpy(&t->pts[dim * s0], &tp[dim * s0], sizeof(float) * dim * (size_t)(s1 - s0));
    memcpy(&t->nrm[dim * s0], &tn[dim * s0], sizeof(float) * dim * (size_t)(s1 - s0));
    memcpy(&t->ms[s0], &tm[s0], sizeof(float) * (size_t)(s1 - s0));
    float inc[3], plo[3];
    int32_t depth = nd->depth;
    for (int a = 0; a < dim; a++) {
      inc[a] = (nd->hi[a] - nd->lo[a]) / 8.0f;
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
      for (int a = 0; a < dim; a++) {
        int k = (c >> (3 * a)) & 7;
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

/* Byte output as the reference's leaf_data_bytes / node_buffer_bytes /
 * contrib_data_bytes (src/sbh.cpp:227-235); padding bytes are zero. */
ORC_EXPORT void orc_nd_emit(const NdTree* t, uint8_t* leaf_out, uint8_t* node_out,
                            uint8_t* contrib_out) {
  const int dim = t->dim;
  for (int64_t j = 0; j < t->n; j++) {
    if (dim == 2) {
      NdLeaf2* L = (NdLeaf2*)leaf_out + j;
      for (int a = 0; a < 2; a++) {
        L->pos[a] = t->pts[2 * j + a];
        L->normal[a] = t->nrm[2 * j + a];
      }
      L->mass = t->ms[j];
      L->alpha = t->alpha;
    } else {
      NdLeaf3* L = (NdLeaf3*)leaf_out + j;
      for (int a = 0; a < 3; a++) {
        L->pos[a] = t->pts[3 * j + a];
        L->normal[a] = t->nrm[3 * j + a];
      }
      L->mass = t->ms[j];
      L->alpha = t->alpha;
    }
  }
  const size_t nsz = dim == 2 ? sizeof(NdNode2) : sizeof(NdNode3);
  memset(node_out, 0, nsz * (size_t)t->nn);
  for (int64_t id = 0; id < t->nn; id++) {
    const NdNode* nd = &t->nodes[id];
    if (dim == 2) {
      NdNode2* o = (NdNode2*)node_out + id;
      NdLeaf2* c = contrib_out ? (NdLeaf2*)contrib_out + id : NULL;
      o->start = (uint64_t)nd->start;
      o->end = (uint64_t)nd->end;
      o->buffer_idx = (uint64_t)id;
      o->parent_idx = (uint64_t)nd->parent;
      for (int a = 0; a < 2; a++) {
        o->bmin[a] = nd->lo[a];
        o->bmax[a] = nd->hi[a];
        o->contrib.pos[a] = nd->pos[a];
        o->contrib.normal[a] = nd->nrm[a];
      }
      o->diameter = nd_diameter(nd, 2);
      o->contrib.mass = nd->m;
      o->contrib.alpha = nd->alpha;
      o->bitmask[0] = nd->mask[0];
      for (int cl = 0; cl < 64; cl++) o->child[cl] = nd_child_id(nd, cl);
      o->num_children = nd->num;
      o->depth = nd->depth;
      o->max_depth = ND_MAX_DEPTH;
      o->is_leaf = (uint8_t)nd->leaf;
      if (c) *c = o->contrib;
    } else {
      NdNode3* o = (NdNode3*)node_out + id;
      NdLeaf3* c = contrib_out ? (NdLeaf3*)contrib_out + id : NULL;
      o->start = (uint64_t)nd->start;
      o->end = (uint64_t)nd->end;
      o->buffer_idx = (uint64_t)id;
      o->parent_idx = (uint64_t)nd->parent;
      for (int a = 0; a < 3; a++) {
        o->bmin[a] = nd->lo[a];
        o->bmax[a] = nd->hi[a];
        o->contrib.pos[a] = nd->pos[a];
        o->contrib.normal[a] = nd->nrm[a];
      }
      o->diameter = nd_diameter(nd, 3);
      o->contrib.mass = nd->m;
      o->contrib.alpha = nd->alpha;
      for (int w = 0; w < 8; w++) o->bitmask[w] = nd->mask[w];
      for (int cl = 0; cl < 512; cl++) o->child[cl] = nd_child_id(nd, cl);
      o->num_children = nd->num;
      o->depth = nd->depth;
      o->max_depth = ND_MAX_DEPTH;
      o->is_leaf = (uint8_t)nd->leaf;
      if (c) *c = o->contrib;
    }
  }
}

/* ------------------------------------------------------------------ walks */
/* Uniform view of a node record of either layout. */
typedef struct {
  const uint8_t* base;
  size_t size;
  int dim;
} NdView;

typedef struct {
  uint64_t start, end, parent;
  const float *bmin, *bmax, *pos;
  float diameter, mass;
  const uint64_t* mask;
  const uint64_t* child;
  int32_t num;
  int leaf;
} NdRec;

static NdRec nd_rec(const NdView* v, uint64_t id) {
  NdRec r;
  if (v->dim == 2) {
    const NdNode2* o = (const NdNode2*)(v->base) + id;
    r.start = o->start;
    r.end = o->end;
    r.parent = o->parent_idx;
    r.bmin = o->bmin;
    r.bmax = o->bmax;
    r.pos = o->contrib.pos;
    r.diameter = o->diameter;
    r.mass = o->contrib.mass;
    r.mask = o->bitmask;
    r.child = o->child;
    r.num = o->num_children;
    r.leaf = o->is_leaf;
  } else {
    const NdNode3* o = (const NdNode3*)(v->base) + id;
    r.start = o->start;
    r.end = o->end;
    r.parent = o->parent_idx;
    r.bmin = o->bmin;
    r.bmax = o->bmax;
    r.pos = o->contrib.pos;
    r.diameter = o->diameter;
    r.mass = o->contrib.mass;
    r.mask = o->bitmask;
    r.child = o->child;
    r.num = o->num_children;
    r.leaf = o->is_leaf;
  }
  return r;
}

typedef struct {
  int kind; /* 0 gravityPotential, 1 gravityGrad, 2 gravityGradSoftCutoff */
  double eps, rc;
  int cutoff, periodic;
  double box;
} NdParams;

static double nd_mi(double d, const NdParams* P) {
  if (!P->periodic || !(P->box > 0.0)) return d;
  return d - P->box * floor(d / P->box + 0.5);
}

/* Pair kernels (include/kernel.h:55-60, 89-94; src/sbh_ffi.cu:142-160) in
 * fp64 for a source at p with mass m. */
static void nd_term(const NdParams* P, int dim, const double* q, const float* p, float m,
                    double* acc) {
  double d[3], r2 = 0.0;
  if (P->kind == 0) {
    for (int a = 0; a < dim; a++) {
      d[a] = nd_mi((double)p[a] - q[a], P);
      r2 += d[a] * d[a];
    }
    acc[0] += -(double)m / (sqrt(r2) + 0.1);
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

static const float* nd_leaf_pos(const uint8_t* leaf, int dim, uint64_t j, float* mass) {
  if (dim == 2) {
    const NdLeaf2* L = (const NdLeaf2*)leaf + j;
    *mass = L->mass;
    return L->pos;
  }
  const NdLeaf3* L = (const NdLeaf3*)leaf + j;
  *mass = L->mass;
  return L->pos;
}

static void nd_leaf_sum(const uint8_t* leaf, uint64_t nl, int dim, const NdRec* r,
                        const NdParams* P, const double* q, double* acc) {
  uint64_t je = r->end < nl ? r->end : nl;
  for (uint64_t j = r->start; j < je; j++) {
    float m;
    const float* p = nd_leaf_pos(leaf, dim, j, &m);
    nd_term(P, dim, q, p, m, acc);
  }
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
  const int dim = v->dim;
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
        nd_leaf_sum(leaf, nl, dim, &r, P, q, o);
      else
        nd_term(P, dim, q, r.pos, r.mass, o);
    }
    return;
  }
  const int cells = dim == 2 ? 64 : 512;
  for (int c = 0; c < cells; c++)
    if ((r.mask[c >> 6] >> (c & 63)) & 1ull)
      if (r.child[c] < nn) nd_vote(leaf, nl, v, nn, r.child[c], qf, g0, g1, beta, P, out, vec);
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
static int nd_cell_in(const NdRec* r, const float* p, int dim) {
  int c = 0;
  int qa[3];
  for (int a = 0; a < dim; a++) {
    if (!(r->bmin[a] - 1e-4f <= p[a] && p[a] <= r->bmax[a] + 1e-4f)) return -1;
    float t = 8.0f * (p[a] - r->bmin[a]) / (r->bmax[a] - r->bmin[a]);
    int k = isnan(t) ? 0 : (t >= 2147483647.0f ? 2147483647 : (t <= -2147483648.0f ? INT32_MIN : (int)t));
    if (k < 0) k = 0;
    if (k > 7) k = 7;
    qa[a] = k;
  }
  for (int a = dim - 1; a >= 0; a--) c = c * 8 + qa[a];
  return c;
}

static int64_t nd_contained(const NdView* v, uint64_t nn, uint64_t id, const float* p) {
  for (int hop = 0; hop < 128; hop++) {
    NdRec r = nd_rec(v, id);
    int c = nd_cell_in(&r, p, v->dim);
    if (c < 0) return -1;
    uint64_t ch = r.child[c];
    if (ch >= nn) return -1;
    NdRec cr = nd_rec(v, ch);
    if (cr.num != 1) return (int64_t)ch;
    id = ch;
  }
  return -1;
}

static void nd_contrib_term(const uint8_t* contrib, uint64_t ncontrib, const NdView* v,
                            uint64_t c, const NdParams* P, const double* q, double* acc) {
  const int dim = v->dim;
  if (c < ncontrib) {
    float m;
    const float* p = nd_leaf_pos(contrib, dim, c, &m);
    nd_term(P, dim, q, p, m, acc);
  } else {
    NdRec r = nd_rec(v, c);
    nd_term(P, dim, q, r.pos, r.mass, acc);
  }
}

/* One query of the sampler; `block` is the query's block index. */
static void nd_mlpcv(const uint8_t* leaf, uint64_t nl, const NdView* v, uint64_t nn,
                     const uint8_t* contrib, uint64_t ncontrib, const float* qf, int64_t qi,
                     uint32_t block, int samples, int seed, const NdParams* P, double* out) {
  const int dim = v->dim;
  const int cells = dim == 2 ? 64 : 512;
  double q[3] = {0, 0, 0}, bh[3] = {0, 0, 0}, ml[3] = {0, 0, 0};
  for (int a = 0; a < dim; a++) q[a] = qf[a];
  const uint32_t rseed = (uint32_t)seed + (uint32_t)qi;
  NdRec root = nd_rec(v, 0);
  for (int b = 0; b < cells; b++) {
    if (!((root.mask[b >> 6] >> (b & 63)) & 1ull)) continue;
    uint64_t ci = root.child[b];
    if (ci >= nn) continue;
    NdRec sub = nd_rec(v, ci);
    if (sub.leaf)
      nd_leaf_sum(leaf, nl, dim, &sub, P, q, bh);
    else
      nd_term(P, dim, q, sub.pos, sub.mass, bh);
    uint64_t count = sub.end - sub.start;
    double total[3] = {0, 0, 0};
    for (int ss = 0; ss < samples; ss++) {
      float u2 = nd_uniform(block, 1u, (uint32_t)ci, (uint32_t)ss, 0u);
      uint64_t idx = (uint64_t)fmaf((float)count, u2, (float)sub.start);
      if (idx > sub.end - 1) idx = sub.end - 1;
      if (count == 0 || idx >= nl) continue;
      float mm;
      const float* sp0 = nd_leaf_pos(leaf, dim, idx, &mm);
      float sp[3] = {0, 0, 0};
      for (int a = 0; a < dim; a++) sp[a] = sp0[a];
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
        nd_term(P, dim, q, cn.pos, cn.mass, prev_res);
        uint64_t first = 0;
        for (int c = 0; c < cells; c++)
          if ((cn.mask[c >> 6] >> (c & 63)) & 1ull) {
            first = cn.child[c];
            break;
          }
        for (int32_t o = 0; o < cn.num; o++) {
          uint64_t c = first + (uint64_t)o;
          if (c >= nn) break;
          NdRec ch = nd_rec(v, c);
          if (ch.leaf)
            nd_leaf_sum(leaf, nl, dim, &ch, P, q, cur_res);
          else
            nd_contrib_term(contrib, ncontrib, v, c, P, q, cur_res);
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
 * records of either layout; kind 0 writes out[i], kinds 1, 2 out[dim i + a].
 * Reference warps: 32 consecutive queries of a block of `threads`. */
ORC_EXPORT void orc_nd_eval(int dim, const uint8_t* leaf_bytes, int64_t nl,
                            const uint8_t* node_bytes, int64_t nn, const uint8_t* contrib_bytes,
                            int64_t ncontrib, const float* queries, int64_t nq, int kind,
                            float beta, int threads, int samples, int seed, double eps,
                            double rc, int periodic, double box, double* out) {
  NdView v = {node_bytes, dim == 2 ? sizeof(NdNode2) : sizeof(NdNode3), dim};
  NdParams P;
  P.kind = kind;
  P.eps = eps > 0.0 ? eps : 0.0;
  P.rc = rc;
  P.cutoff = rc > 0.0 && isfinite(rc);
  P.periodic = periodic != 0 && box > 0.0;
  P.box = box;
  const int vec = kind != 0;
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
             &queries[dim * i], i, (uint32_t)(i / threads), samples, seed, &P, o);
    if (vec)
      for (int a = 0; a < dim; a++) out[dim * i + a] = o[a];
    else
      out[i] = o[0];
  }
}
