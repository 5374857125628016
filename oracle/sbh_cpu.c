/* TEST INFRASTRUCTURE ONLY -- part of the oracle: a C restatement of the
 * reference's CPU Barnes-Hut path, used as a checker and as the CPU baseline
 * beside the GPU numbers.  Nothing in the product links or calls it.
 *
 * Restated (src/sbh.cpp:43-153, include/tree3d.h:30-206 and 238-371,
 * include/barnes_hut.h:24-78, include/tree_sampler.h:21-113):
 *   Tree3D<LEAF_SIZE, DIM_SPLIT_LOG>: breadth-first build; a node is a leaf
 *     when depth >= max_depth or it holds <= LEAF_SIZE points; otherwise its
 *     points are partitioned into S^3 cells (S = 2^DIM_SPLIT_LOG) of its box
 *     by AABB::quantize, in (z, y, x) loop order, and every non-empty cell
 *     becomes a child with the cell's box; children are reserved
 *     consecutively, so they are contiguous and in ascending cell order.
 *     Leaf moments are mass-weighted means of their points, internal ones of
 *     their children (bottom-up).
 *   MAC: |q - centre of mass| / |box diagonal| >= beta (leaves always).
 *   evaluateBarnesHut_stackless: depth-first in child order.
 *   multiLevelPrefixControlVariateSample_stackless: subdomains are the
 *     root's children.
 * Differences, stated: std::partition is not stable, so the reference's order
 * of points inside a cell is unspecified; here the partition is stable.  The
 * reference shares one std::mt19937 between threads (a data race); here the
 * stochastic draws are Philox4x32-10 numbers indexed by (query, subdomain,
 * sample, step).  Moments are summed in double and stored as float.
 * Queries run on OpenMP threads with the reference's strided assignment
 * (query i on thread i mod T). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

ORC_EXPORT void orc_philox4x32(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out);

typedef struct {
  float pos[3], normal[3], mass, alpha;
} CpuLeaf;

typedef struct {
  int64_t n, nodes, cap;
  int split;       /* S = 2^DIM_SPLIT_LOG cells per axis */
  CpuLeaf* leaf;   /* points in tree order */
  int64_t* start;  /* per node: point range [start, end) */
  int64_t* end;
  int64_t* first_child; /* children contiguous, ascending cell */
  int32_t* nchild;
  int32_t* cell;        /* cell index of the node inside its parent */
  int64_t* parent;
  int32_t* depth;
  uint8_t* is_leaf;
  float* bmin; /* 3 per node */
  float* bmax;
  float* diam;
  float* com;  /* 3 per node */
  float* mass;
  float* nrm;  /* 3 per node: mass-weighted mean normal */
} CpuTree;

static void cpu_grow(CpuTree* t, int64_t need) {
  if (need <= t->cap) return;
  int64_t cap = t->cap ? t->cap : 64;
  while (cap < need) cap *= 2;
#define GROW(f, T, k) t->f = (T*)realloc(t->f, sizeof(T) * (size_t)(k) * (size_t)cap)
  GROW(start, int64_t, 1);
  GROW(end, int64_t, 1);
  GROW(first_child, int64_t, 1);
  GROW(nchild, int32_t, 1);
  GROW(cell, int32_t, 1);
  GROW(parent, int64_t, 1);
  GROW(depth, int32_t, 1);
  GROW(is_leaf, uint8_t, 1);
  GROW(bmin, float, 3);
  GROW(bmax, float, 3);
  GROW(diam, float, 1);
  GROW(com, float, 3);
  GROW(mass, float, 1);
  GROW(nrm, float, 3);
#undef GROW
  t->cap = cap;
}

/* AABB::quantize (include/aabb.h:36-42). */
static inline int cpu_quantize(float p, float lo, float hi, int s) {
  int q = (int)((float)s * (p - lo) / (hi - lo));
  if (q > s - 1) q = s - 1;
  if (q < 0) q = 0;
  return q;
}

ORC_EXPORT void orc_cpu_tree_free(CpuTree* t) {
  if (!t) return;
  free(t->leaf);
  free(t->start);
  free(t->end);
  free(t->first_child);
  free(t->nchild);
  free(t->cell);
  free(t->parent);
  free(t->depth);
  free(t->is_leaf);
  free(t->bmin);
  free(t->bmax);
  free(t->diam);
  free(t->com);
  free(t->mass);
  free(t->nrm);
  free(t);
}

ORC_EXPORT CpuTree* orc_cpu_tree_build(const float* pts, const float* normals,
                                       const float* masses, int64_t n, const float* lo,
                                       const float* hi, int leaf_size, int dim_split_log,
                                       int max_depth) {
  CpuTree* t = (CpuTree*)calloc(1, sizeof(CpuTree));
  t->n = n;
  t->split = 1 << dim_split_log;
  const int s = t->split, s3 = s * s * s;
  t->leaf = (CpuLeaf*)malloc(sizeof(CpuLeaf) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      t->leaf[i].pos[a] = pts[3 * i + a];
      t->leaf[i].normal[a] = normals ? normals[3 * i + a] : 0.0f;
    }
    t->leaf[i].mass = masses[i];
    t->leaf[i].alpha = 1.0f;
  }
  CpuLeaf* tmp = (CpuLeaf*)malloc(sizeof(CpuLeaf) * (size_t)(n > 0 ? n : 1));
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s3 + 1));
  int* cellof = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  cpu_grow(t, 1);
  t->nodes = 1;
  t->start[0] = 0;
  t->end[0] = n;
  t->parent[0] = 0;
  t->cell[0] = 0;
  t->depth[0] = 0;
  for (int a = 0; a < 3; a++) {
    t->bmin[a] = lo[a];
    t->bmax[a] = hi[a];
  }
  /* breadth-first: nodes are processed in index order, children appended */
  for (int64_t k = 0; k < t->nodes; k++) {
    const int64_t s0 = t->start[k], s1 = t->end[k];
    float d[3];
    for (int a = 0; a < 3; a++) d[a] = t->bmax[3 * k + a] - t->bmin[3 * k + a];
    t->diam[k] = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    t->first_child[k] = -1;
    t->nchild[k] = 0;
    t->is_leaf[k] = (t->depth[k] >= max_depth || s1 - s0 <= leaf_size) ? 1 : 0;
    if (t->is_leaf[k]) continue;
    memset(cnt, 0, sizeof(int64_t) * (size_t)(s3 + 1));
    for (int64_t i = s0; i < s1; i++) {
      int qx = cpu_quantize(t->leaf[i].pos[0], t->bmin[3 * k], t->bmax[3 * k], s);
      int qy = cpu_quantize(t->leaf[i].pos[1], t->bmin[3 * k + 1], t->bmax[3 * k + 1], s);
      int qz = cpu_quantize(t->leaf[i].pos[2], t->bmin[3 * k + 2], t->bmax[3 * k + 2], s);
      int c = (qz * s + qy) * s + qx;
      cellof[i] = c;
      cnt[c + 1]++;
    }
    for (int c = 0; c < s3; c++) cnt[c + 1] += cnt[c];
    for (int64_t i = s0; i < s1; i++) tmp[s0 + cnt[cellof[i]]++] = t->leaf[i];
    memcpy(t->leaf + s0, tmp + s0, sizeof(CpuLeaf) * (size_t)(s1 - s0));
    /* cnt[c] is now the end offset of cell c; children in ascending cell */
    float inc[3];
    for (int a = 0; a < 3; a++) inc[a] = d[a] / (float)s;
    int64_t prev = 0;
    for (int c = 0; c < s3; c++) {
      const int64_t e = cnt[c];
      if (e > prev) {
        cpu_grow(t, t->nodes + 1);
        const int64_t ch = t->nodes++;
        if (t->first_child[k] < 0) t->first_child[k] = ch;
        t->nchild[k]++;
        const int x = c % s, y = (c / s) % s, z = c / (s * s);
        const int q[3] = {x, y, z};
        for (int a = 0; a < 3; a++) {
          float m0 = t->bmin[3 * k + a] + (float)q[a] * inc[a];
          t->bmin[3 * ch + a] = m0;
          t->bmax[3 * ch + a] = m0 + inc[a];
        }
        t->start[ch] = s0 + prev;
        t->end[ch] = s0 + e;
        t->parent[ch] = k;
        t->cell[ch] = c;
        t->depth[ch] = t->depth[k] + 1;
      }
      prev = e;
    }
  }
  /* moments bottom-up (children have larger indices than their parent) */
  for (int64_t k = t->nodes - 1; k >= 0; k--) {
    double M = 0, X[3] = {0, 0, 0}, N[3] = {0, 0, 0};
    if (t->is_leaf[k]) {
      for (int64_t i = t->start[k]; i < t->end[k]; i++) {
        const CpuLeaf* L = &t->leaf[i];
        M += L->mass;
        for (int a = 0; a < 3; a++) {
          X[a] += (double)L->mass * L->pos[a];
          N[a] += (double)L->mass * L->normal[a];
        }
      }
    } else {
      for (int64_t c = t->first_child[k]; c < t->first_child[k] + t->nchild[k]; c++) {
        M += t->mass[c];
        for (int a = 0; a < 3; a++) {
          X[a] += (double)t->mass[c] * t->com[3 * c + a];
          N[a] += (double)t->mass[c] * t->nrm[3 * c + a];
        }
      }
    }
    t->mass[k] = (float)M;
    for (int a = 0; a < 3; a++) {
      t->com[3 * k + a] = M != 0.0 ? (float)(X[a] / M) : 0.0f;
      t->nrm[3 * k + a] = M != 0.0 ? (float)(N[a] / M) : 0.0f;
    }
  }
  free(tmp);
  free(cnt);
  free(cellof);
  return t;
}

ORC_EXPORT int64_t orc_cpu_tree_nodes(const CpuTree* t) { return t->nodes; }

/* Copies per-node arrays out: start, end, first_child, nchild, is_leaf, mass, com. */
ORC_EXPORT void orc_cpu_tree_arrays(const CpuTree* t, int64_t* start, int64_t* end,
                                    int64_t* first_child, int32_t* nchild, uint8_t* is_leaf,
                                    float* mass, float* com, float* leaf_pos) {
  for (int64_t k = 0; k < t->nodes; k++) {
    start[k] = t->start[k];
    end[k] = t->end[k];
    first_child[k] = t->first_child[k];
    nchild[k] = t->nchild[k];
    is_leaf[k] = t->is_leaf[k];
    mass[k] = t->mass[k];
    for (int a = 0; a < 3; a++) com[3 * k + a] = t->com[3 * k + a];
  }
  for (int64_t i = 0; i < t->n; i++)
    for (int a = 0; a < 3; a++) leaf_pos[3 * i + a] = t->leaf[i].pos[a];
}

/* kind: 0 gravityPotential -m/(|p-q|+0.1); 1 signedSolidAngle<3>;
 *       2 smoothDistExp exp(-alpha |p-q|); 3 softened force (eps) -m s^3 (q-p). */
static inline void cpu_term(int kind, const float* q, const float* p, const float* nv, float m,
                            float alpha, double e2, double* acc) {
  double d[3] = {(double)p[0] - q[0], (double)p[1] - q[1], (double)p[2] - q[2]};
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (kind == 0) {
    acc[0] += -(double)m / (sqrt(r2) + 0.1);
  } else if (kind == 1) {
    double r = sqrt(r2);
    acc[0] += (double)m * (nv[0] * d[0] + nv[1] * d[1] + nv[2] * d[2]) /
              (4.0 * M_PI * (r * r * r + 1e-5));
  } else if (kind == 2) {
    acc[0] += exp(-(double)alpha * sqrt(r2));
  } else {
    double s = 1.0 / sqrt(r2 + e2);
    double c = (double)m * s * s * s; /* force on q: -m s^3 (q - p) = m s^3 d */
    acc[0] += c * d[0];
    acc[1] += c * d[1];
    acc[2] += c * d[2];
  }
}

static inline float cpu_ratio(const CpuTree* t, int64_t k, const float* q) {
  if (t->is_leaf[k]) return INFINITY;
  float dx = q[0] - t->com[3 * k], dy = q[1] - t->com[3 * k + 1], dz = q[2] - t->com[3 * k + 2];
  return sqrtf(dx * dx + dy * dy + dz * dz) / t->diam[k];
}

static inline void cpu_node_term(const CpuTree* t, int kind, int64_t k, const float* q,
                                 double e2, double* acc, int64_t* nint) {
  if (t->is_leaf[k]) {
    for (int64_t i = t->start[k]; i < t->end[k]; i++) {
      const CpuLeaf* L = &t->leaf[i];
      cpu_term(kind, q, L->pos, L->normal, L->mass, L->alpha, e2, acc);
      (*nint)++;
    }
  } else {
    cpu_term(kind, q, t->com + 3 * k, t->nrm + 3 * k, t->mass[k], 1.0f, e2, acc);
    (*nint)++;
  }
}

/* evaluateBarnesHut_stackless, one query (the stackless walk visits the
 * children of an opened node in order and returns to its parent; a
 * depth-first recursion in child order visits the same nodes in the same
 * order). */
static void cpu_walk(const CpuTree* t, int kind, int64_t k, const float* q, float beta,
                     double e2, double* acc, int64_t* nint) {
  if (cpu_ratio(t, k, q) >= beta) {
    cpu_node_term(t, kind, k, q, e2, acc, nint);
    return;
  }
  for (int64_t c = t->first_child[k]; c >= 0 && c < t->first_child[k] + t->nchild[k]; c++)
    cpu_walk(t, kind, c, q, beta, e2, acc, nint);
}

/* All queries, `threads` OpenMP threads with the strided assignment of
 * src/sbh.cpp:71-86 (0: all available).  out: 3 doubles per query (scalar
 * kinds use out[3 i]).  Returns the total number of interactions. */
ORC_EXPORT int64_t orc_cpu_bh_eval(const CpuTree* t, const float* queries, int64_t nq, int kind,
                                   float beta, double eps, int threads, double* out) {
  int64_t total = 0;
  const double e2 = eps * eps;
#ifdef _OPENMP
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(+ : total)
#endif
  for (int64_t i = 0; i < nq; i++) {
    double acc[3] = {0, 0, 0};
    int64_t nint = 0;
    if (t->nodes > 0 && t->n > 0) cpu_walk(t, kind, 0, queries + 3 * i, beta, e2, acc, &nint);
    out[3 * i] = acc[0];
    out[3 * i + 1] = acc[1];
    out[3 * i + 2] = acc[2];
    total += nint;
  }
  return total;
}

/* Node of `p` below node k: the child whose cell holds p (AABB::quantize in
 * k's box), skipping single-child chains as Tree3DNode::containedInChild
 * does (include/tree3d.h:172-183); -1 when p is outside k's box (+- 1e-4). */
static int64_t cpu_contained_child(const CpuTree* t, int64_t k, const float* p) {
  for (int a = 0; a < 3; a++)
    if (!(t->bmin[3 * k + a] - 1e-4f <= p[a] && p[a] <= t->bmax[3 * k + a] + 1e-4f)) return -1;
  const int s = t->split;
  int qx = cpu_quantize(p[0], t->bmin[3 * k], t->bmax[3 * k], s);
  int qy = cpu_quantize(p[1], t->bmin[3 * k + 1], t->bmax[3 * k + 1], s);
  int qz = cpu_quantize(p[2], t->bmin[3 * k + 2], t->bmax[3 * k + 2], s);
  int c = (qz * s + qy) * s + qx;
  for (int64_t ch = t->first_child[k]; ch >= 0 && ch < t->first_child[k] + t->nchild[k]; ch++) {
    if (t->cell[ch] == c) {
      if (t->nchild[ch] == 1) {
        int64_t deeper = cpu_contained_child(t, ch, p);
        return deeper >= 0 ? deeper : ch;
      }
      return ch;
    }
  }
  return -1;
}

static inline double cpu_uniform(uint32_t seed, uint32_t q, uint32_t node, uint32_t sample,
                                 uint32_t step) {
  uint32_t ctr[4] = {q, node, sample, step};
  uint32_t key[2] = {seed, 0x0C1CCAFEu};
  uint32_t o[4];
  orc_philox4x32(ctr, key, o);
  return ((double)(o[0] >> 8) + 0.5) * (1.0 / 16777216.0); /* (0, 1) */
}

/* multiLevelPrefixControlVariateSample_stackless (include/tree_sampler.h:21-113)
 * for scalar kinds 0..2 and the force kind 3. */
ORC_EXPORT void orc_cpu_mlpcv(const CpuTree* t, const float* queries, int64_t nq, int kind,
                              int samples, uint32_t seed, double eps, int threads, double* out) {
  const double e2 = eps * eps;
#ifdef _OPENMP
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static, 1)
#endif
  for (int64_t i = 0; i < nq; i++) {
    const float* q = queries + 3 * i;
    double bh[3] = {0, 0, 0}, mlcv[3] = {0, 0, 0};
    int64_t nint = 0;
    const int64_t root = 0;
    for (int64_t sub = t->first_child[root];
         t->nodes > 0 && sub >= 0 && sub < t->first_child[root] + t->nchild[root]; sub++) {
      cpu_node_term(t, kind, sub, q, e2, bh, &nint);
      const int64_t nel = t->end[sub] - t->start[sub];
      double sub_total[3] = {0, 0, 0};
      for (int ss = 0; ss < samples; ss++) {
        int64_t idx = (int64_t)(cpu_uniform(seed, (uint32_t)i, (uint32_t)sub, (uint32_t)ss, 0) *
                                (double)nel);
        if (idx >= nel) idx = nel - 1;
        const float* sp = t->leaf[t->start[sub] + idx].pos;
        double event_prob = 1.0;
        int64_t prev = sub, cur = sub;
        double wsum[3] = {0, 0, 0};
        double cur_beta = fmax(1.0, (double)cpu_ratio(t, cur, q));
        uint32_t step = 1;
        while (!t->is_leaf[cur]) {
          double ratio = cpu_ratio(t, cur, q);
          double tp = fmin(cur_beta / ratio, 1.0);
          double p = cpu_uniform(seed, (uint32_t)i, (uint32_t)sub, (uint32_t)ss, step++);
          if (p > tp) {
            event_prob *= 1.0 - tp;
            break;
          }
          event_prob *= tp;
          prev = cur;
          cur = cpu_contained_child(t, cur, sp);
          if (cur < 0) break;
          cur_beta = fmax(cur_beta, ratio);
          double pr[3] = {0, 0, 0}, cr[3] = {0, 0, 0};
          cpu_term(kind, q, t->com + 3 * prev, t->nrm + 3 * prev, t->mass[prev], 1.0f, e2, pr);
          for (int64_t c = t->first_child[prev]; c < t->first_child[prev] + t->nchild[prev]; c++)
            cpu_node_term(t, kind, c, q, e2, cr, &nint);
          double w = (double)nel / (event_prob * (double)(t->end[prev] - t->start[prev]));
          for (int a = 0; a < 3; a++) wsum[a] += w * (cr[a] - pr[a]);
        }
        for (int a = 0; a < 3; a++) sub_total[a] += wsum[a];
      }
      for (int a = 0; a < 3; a++) mlcv[a] += sub_total[a] / (double)samples;
    }
    for (int a = 0; a < 3; a++) out[3 * i + a] = bh[a] + mlcv[a];
  }
}

/* Threads the OpenMP loops of the oracle use (OMP_NUM_THREADS or all). */
ORC_EXPORT int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
