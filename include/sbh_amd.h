/*
 * sbh_amd.h -- C ABI of the MI355X (gfx950) Barnes-Hut library libsbh_amd.so.
 *
 * Each entry point replaces one XLA-FFI handler of the reference
 * (src/sbh_ffi.cu, registered at src/sbh_ffi.cu:3580-3947 and bound from
 * python/jax_sbh_ffi.py).  The signature keeps the handler's order:
 *
 *     stream, argument buffers..., result buffers..., attributes...
 *
 * with the attribute names and types of the reference (i32 vs f32 as passed
 * by python/jax_sbh_ffi.py with np.int32 / np.float32).
 *
 * Buffers are device pointers plus shape (SbhBuffer, the plain-C analogue of
 * xla::ffi::Buffer<dtype, rank>).  The caller owns every argument and result
 * buffer; results must be allocated at the byte sizes the reference's Python
 * computes.  Parameters travel by value into the kernels (no __constant__
 * state), so calls on different streams do not interfere.
 *
 * Return value: SBH_OK, SBH_INVALID_ARGUMENT (shape / attribute / range
 * errors, detected on the host before any HIP call, with the reference's
 * message text) or SBH_INTERNAL (a HIP runtime error).  sbh_last_error()
 * returns the message of the last failure on the calling thread.
 */
#ifndef SBH_AMD_H_
#define SBH_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the xla::ffi::ErrorCode values the reference returns. */
#define SBH_OK 0
#define SBH_INVALID_ARGUMENT 3
#define SBH_INTERNAL 13

/* Element types: XLA_FFI_DataType values. */
#define SBH_DTYPE_S32 4
#define SBH_DTYPE_S64 5
#define SBH_DTYPE_U8 6
#define SBH_DTYPE_U32 8
#define SBH_DTYPE_F32 11

typedef struct SbhBuffer {
  void* data;       /* device pointer */
  int32_t dtype;    /* SBH_DTYPE_* */
  int32_t rank;     /* 0..4 */
  int64_t dims[4];  /* row-major extents */
} SbhBuffer;

/* Message of the last failed call on this thread ("" if none). */
const char* sbh_last_error(void);

/* ---- layout sizes (src/sbh_ffi.cu:3949-3963) ------------------------- */
size_t sbh_leaf3_size(void);              /* 32   : LeafData<3>            */
size_t sbh_node3_size(void);              /* 4272 : Tree3DNode<1,3>        */
size_t sbh_octree_level3_max_nodes(void); /* 513                           */
size_t sbh_octree8_node_size(void);       /* 104  : OctreeNode8            */

/* ---- direct sums --------------------------------------------------------- */

/* Replaces BruteForceGravity (src/sbh_ffi.cu:548-597; target
 * "sbh_bruteforce_gravity", python/jax_sbh_ffi.py:341-378).
 * points (N,d) f32, masses (N) f32, queries (M,d) f32 -> out (M) f32,
 * out[q] = sum_i -m_i / (|p_i - q| + 0.1), d in {2,3}. */
int sbh_bruteforce_gravity(void* stream, const SbhBuffer* points,
                           const SbhBuffer* masses, const SbhBuffer* queries,
                           SbhBuffer* out, int32_t periodic, float box_length,
                           int32_t threads);

/* Replaces SoftenedDirectForce (src/sbh_ffi.cu:600-661; target
 * "sbh_direct_force_soft", python/jax_sbh_ffi.py:381-426).
 * out (M,3) = sum_i -m_i * att(r) * (r^2 + eps^2 + 1e-12)^(-3/2) * (q - p_i). */
int sbh_direct_force_soft(void* stream, const SbhBuffer* points,
                          const SbhBuffer* masses, const SbhBuffer* queries,
                          SbhBuffer* out, float softening_scale,
                          float cutoff_scale, int32_t periodic,
                          float box_length, int32_t threads);

/* Replaces SoftenedDirectForceVjp (src/sbh_ffi.cu:663-734; target
 * "sbh_direct_force_soft_vjp", python/jax_sbh_ffi.py:429-479).
 * out (M,3) = J_q^T cotangent (query side, sources fixed). */
int sbh_direct_force_soft_vjp(void* stream, const SbhBuffer* points,
                              const SbhBuffer* masses, const SbhBuffer* queries,
                              const SbhBuffer* cotangent, SbhBuffer* out,
                              float softening_scale, float cutoff_scale,
                              int32_t periodic, float box_length,
                              int32_t threads);

/* New (no reference counterpart; the reference's custom_vjp returns the
 * query side only and zero mass gradients, python/jax_sbh_ffi.py:576-588).
 * Full adjoint of L = sum_i c_i . F_i for self-gravity (queries == points):
 * grad_points (N,3) and grad_masses (N). */
int sbh_direct_force_soft_full_vjp(void* stream, const SbhBuffer* points,
                                   const SbhBuffer* masses,
                                   const SbhBuffer* cotangent,
                                   SbhBuffer* grad_points,
                                   SbhBuffer* grad_masses,
                                   float softening_scale, float cutoff_scale,
                                   int32_t periodic, float box_length,
                                   int32_t threads);

/* ---- dense octree8 (drop-in mode) ---------------------------------------- */

/* Replaces BuildOctree8Buffers (src/sbh_ffi.cu:1039-1420; target
 * "sbh_build_octree8_buffers", python/jax_sbh_ffi.py:864-960).
 * points (N,3), normals (N,3), masses (N), min_corner (3), max_corner (3)
 * -> leaf_out u8[N*32], node_out u8[T*104], contrib_out u8[T*32],
 * T = sum_{l<=max_depth} 8^l, max_depth in [1,10].
 * Points inside a leaf bin are in (code, original index) order; node
 * topology and geometry follow the reference record by record.  Unlike the
 * reference, the bounds are read on the device (no host synchronisation). */
int sbh_build_octree8_buffers(void* stream, const SbhBuffer* points,
                              const SbhBuffer* normals, const SbhBuffer* masses,
                              const SbhBuffer* min_corner,
                              const SbhBuffer* max_corner, SbhBuffer* leaf_out,
                              SbhBuffer* node_out, SbhBuffer* contrib_out,
                              int32_t max_depth);

/* Replaces BarnesHutForceOctree8 (src/sbh_ffi.cu:3133-3181; target
 * "sbh_barnes_hut_force_octree8", python/jax_sbh_ffi.py:1336-1369).
 * Classic walk with gravityGrad = -m d / (|d| + 0.1)^3. */
int sbh_barnes_hut_force_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                 const SbhBuffer* node_bytes,
                                 const SbhBuffer* queries, SbhBuffer* out,
                                 float beta, int32_t periodic, float box_length,
                                 int32_t threads);

/* Replaces SoftenedBarnesHutForceOctree8 (src/sbh_ffi.cu:3183-3248; target
 * "sbh_barnes_hut_force_octree8_soft", python/jax_sbh_ffi.py:1398-1447). */
int sbh_barnes_hut_force_octree8_soft(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, SbhBuffer* out, float beta, float softening_scale,
    float cutoff_scale, int32_t prune_enabled, float prune_r_cut_mult,
    int32_t periodic, float box_length, int32_t threads);

/* Replaces SoftenedBarnesHutForceOctree8Vjp (src/sbh_ffi.cu:3250-3320;
 * target "sbh_barnes_hut_force_octree8_soft_vjp",
 * python/jax_sbh_ffi.py:1674-1728).  Query-side J^T g, tree fixed. */
int sbh_barnes_hut_force_octree8_soft_vjp(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, const SbhBuffer* cotangent, SbhBuffer* out,
    float beta, float softening_scale, float cutoff_scale,
    int32_t prune_enabled, float prune_r_cut_mult, int32_t periodic,
    float box_length, int32_t threads);

/* ---- stochastic walks over the dense octree8 ------------------------------
 * Multi-level prefix control variate (src/sbh_ffi.cu:2146-2442): subdomains
 * are the nodes at depth 3 (or shallower leaves); each contributes its
 * monopole plus samples_per_subdomain Russian-roulette correction paths.
 * Random numbers: Philox4x32-10 keyed by seed and indexed by (query,
 * subdomain node, sample, step), in place of the reference's cuRAND XORWOW
 * streams (src/sbh_ffi.cu:2160-2163); results are therefore statistically,
 * not bitwise, equivalent to the reference, and deterministic per seed. */

/* Replaces StochasticBarnesHutForceOctree8 (src/sbh_ffi.cu:3322-3373; target
 * "sbh_stochastic_bh_force_octree8", python/jax_sbh_ffi.py:1450-1487). */
int sbh_stochastic_bh_force_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                    const SbhBuffer* node_bytes,
                                    const SbhBuffer* queries, SbhBuffer* out,
                                    int32_t samples_per_subdomain, int32_t seed,
                                    float beta, int32_t periodic,
                                    float box_length, int32_t threads);

/* Replaces SoftenedStochasticBarnesHutForceOctree8 (src/sbh_ffi.cu:3375-3443;
 * target "sbh_stochastic_bh_force_octree8_soft",
 * python/jax_sbh_ffi.py:1537-1588). */
int sbh_stochastic_bh_force_octree8_soft(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, SbhBuffer* out, int32_t samples_per_subdomain,
    int32_t seed, float beta, float softening_scale, float cutoff_scale,
    int32_t prune_enabled, float prune_r_cut_mult, int32_t periodic,
    float box_length, int32_t threads);

/* Replaces StochasticBarnesHutForceOctree8Vjp (src/sbh_ffi.cu:3445-3503;
 * target "sbh_stochastic_bh_force_octree8_vjp",
 * python/jax_sbh_ffi.py:1490-1534).  Query-side J^T g along the same sampled
 * paths as the forward call with the same seed. */
int sbh_stochastic_bh_force_octree8_vjp(void* stream,
                                        const SbhBuffer* leaf_bytes,
                                        const SbhBuffer* node_bytes,
                                        const SbhBuffer* queries,
                                        const SbhBuffer* cotangent,
                                        SbhBuffer* out,
                                        int32_t samples_per_subdomain,
                                        int32_t seed, float beta,
                                        int32_t periodic, float box_length,
                                        int32_t threads);

/* Replaces SoftenedStochasticBarnesHutForceOctree8Vjp
 * (src/sbh_ffi.cu:3505-3576; target "sbh_stochastic_bh_force_octree8_soft_vjp",
 * python/jax_sbh_ffi.py:1591-1671 with use_ffi=True). */
int sbh_stochastic_bh_force_octree8_soft_vjp(
    void* stream, const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
    const SbhBuffer* queries, const SbhBuffer* cotangent, SbhBuffer* out,
    int32_t samples_per_subdomain, int32_t seed, float beta,
    float softening_scale, float cutoff_scale, int32_t prune_enabled,
    float prune_r_cut_mult, int32_t periodic, float box_length,
    int32_t threads);

/* ---- level-3 tree (Tree3DNode<1,3> buffers) -------------------------------
 *
 * One 8 x 8 x 8 split under the root (cells on each axis of the given box);
 * leaf N*32, node 513*4272 and contrib 513*32 bytes.  The walks are the
 * reference's warp-voted traversal (a node is a far field only if every
 * active query of a 32-query warp of a `threads`-sized block accepts it) and
 * its multi-level control-variate sampler; they take any Tree3DNode<1,3>
 * buffer, also the multi-level trees of sbh_tree_build (dim 3).  For level-3
 * buffers leaves always accept, so a warp uses either the root monopole or
 * the exact sum, and the stochastic variants sum every (leaf) subdomain
 * exactly and divide the zero correction by samples_per_subdomain (0 samples
 * give NaN, as the reference).  Sample paths draw Philox4x32-10 numbers in
 * place of cuRAND XORWOW: the sample point from a stream of the block index
 * (shared by a block's queries, as the reference's), the roulette from a
 * stream of seed + query. */

/* Replaces BuildOctreeLevel3Buffers (src/sbh_ffi.cu:736-1037; target
 * "sbh_build_octree_level3_buffers", python/jax_sbh_ffi.py:744-861).  Points
 * inside a cell are stored in input order (the reference's atomic cursor
 * leaves that order undefined). */
int sbh_build_octree_level3_buffers(void* stream, const SbhBuffer* points,
                                    const SbhBuffer* normals, const SbhBuffer* masses,
                                    const SbhBuffer* min_corner,
                                    const SbhBuffer* max_corner, SbhBuffer* leaf_out,
                                    SbhBuffer* node_out, SbhBuffer* contrib_out);

/* Replaces BarnesHutGravity3D (src/sbh_ffi.cu:2490-2534; target
 * "sbh_barnes_hut_gravity_3d", python/jax_sbh_ffi.py:963-997): potential
 * -m / (|p - q| + 0.1), out (M,). */
int sbh_barnes_hut_gravity_3d(void* stream, const SbhBuffer* leaf_bytes,
                              const SbhBuffer* node_bytes, const SbhBuffer* queries,
                              SbhBuffer* out, float beta, int32_t periodic,
                              float box_length, int32_t threads);

/* Replaces BarnesHutForce3D (src/sbh_ffi.cu:2729-2777; target
 * "sbh_barnes_hut_force_3d", python/jax_sbh_ffi.py:1000-1034). */
int sbh_barnes_hut_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                            const SbhBuffer* node_bytes, const SbhBuffer* queries,
                            SbhBuffer* out, float beta, int32_t periodic,
                            float box_length, int32_t threads);

/* Replaces SoftenedBarnesHutForce3D (src/sbh_ffi.cu:2671-2727; target
 * "sbh_barnes_hut_force_soft_3d", python/jax_sbh_ffi.py:1063-1105). */
int sbh_barnes_hut_force_soft_3d(void* stream, const SbhBuffer* leaf_bytes,
                                 const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                 SbhBuffer* out, float beta, float softening_scale,
                                 float cutoff_scale, int32_t periodic, float box_length,
                                 int32_t threads);

/* Replaces StochasticBarnesHutGravity3D (src/sbh_ffi.cu:2834-2887; target
 * "sbh_stochastic_bh_gravity_3d", python/jax_sbh_ffi.py:1138-1177). */
int sbh_stochastic_bh_gravity_3d(void* stream, const SbhBuffer* leaf_bytes,
                                 const SbhBuffer* node_bytes, const SbhBuffer* contrib_bytes,
                                 const SbhBuffer* queries, SbhBuffer* out,
                                 int32_t samples_per_subdomain, int32_t seed,
                                 int32_t periodic, float box_length, int32_t threads);

/* Replaces StochasticBarnesHutForce3D (src/sbh_ffi.cu:2944-2997; target
 * "sbh_stochastic_bh_force_3d", python/jax_sbh_ffi.py:1180-1219). */
int sbh_stochastic_bh_force_3d(void* stream, const SbhBuffer* leaf_bytes,
                               const SbhBuffer* node_bytes, const SbhBuffer* contrib_bytes,
                               const SbhBuffer* queries, SbhBuffer* out,
                               int32_t samples_per_subdomain, int32_t seed,
                               int32_t periodic, float box_length, int32_t threads);

/* Replaces SoftenedStochasticBarnesHutForce3D (src/sbh_ffi.cu:3066-3131;
 * target "sbh_stochastic_bh_force_soft_3d", python/jax_sbh_ffi.py:1252-1299). */
int sbh_stochastic_bh_force_soft_3d(void* stream, const SbhBuffer* leaf_bytes,
                                    const SbhBuffer* node_bytes,
                                    const SbhBuffer* contrib_bytes,
                                    const SbhBuffer* queries, SbhBuffer* out,
                                    int32_t samples_per_subdomain, int32_t seed,
                                    float softening_scale, float cutoff_scale,
                                    int32_t periodic, float box_length, int32_t threads);

/* ---- 2-D quadtree targets (Tree2DNode<1,3> buffers) -----------------------
 *
 * Leaf records LeafData<2> (24 B: position[2], normal[2], mass, alpha), node
 * records Tree2DNode<1,3> (616 B: 8 x 8 cells, cell = x + 8 y), contrib one
 * LeafData<2> per node.  The walks are the level-3 walks above with two
 * coordinates: queries (M, 2), forces (M, 2), potentials (M,). */

/* Replaces BarnesHutGravity (src/sbh_ffi.cu:2444-2488; target
 * "sbh_barnes_hut_gravity", python/jax_sbh_ffi.py:963-997). */
int sbh_barnes_hut_gravity(void* stream, const SbhBuffer* leaf_bytes,
                           const SbhBuffer* node_bytes, const SbhBuffer* queries,
                           SbhBuffer* out, float beta, int32_t periodic, float box_length,
                           int32_t threads);

/* Replaces BarnesHutForce (src/sbh_ffi.cu:2563-2611; target
 * "sbh_barnes_hut_force", python/jax_sbh_ffi.py:1000-1034). */
int sbh_barnes_hut_force(void* stream, const SbhBuffer* leaf_bytes,
                         const SbhBuffer* node_bytes, const SbhBuffer* queries, SbhBuffer* out,
                         float beta, int32_t periodic, float box_length, int32_t threads);

/* Replaces SoftenedBarnesHutForce (src/sbh_ffi.cu:2613-2669; target
 * "sbh_barnes_hut_force_soft", python/jax_sbh_ffi.py:1063-1105). */
int sbh_barnes_hut_force_soft(void* stream, const SbhBuffer* leaf_bytes,
                              const SbhBuffer* node_bytes, const SbhBuffer* queries,
                              SbhBuffer* out, float beta, float softening_scale,
                              float cutoff_scale, int32_t periodic, float box_length,
                              int32_t threads);

/* Replaces StochasticBarnesHutGravity (src/sbh_ffi.cu:2779-2832; target
 * "sbh_stochastic_bh_gravity", python/jax_sbh_ffi.py:1138-1177). */
int sbh_stochastic_bh_gravity(void* stream, const SbhBuffer* leaf_bytes,
                              const SbhBuffer* node_bytes, const SbhBuffer* contrib_bytes,
                              const SbhBuffer* queries, SbhBuffer* out,
                              int32_t samples_per_subdomain, int32_t seed, int32_t periodic,
                              float box_length, int32_t threads);

/* Replaces StochasticBarnesHutForce (src/sbh_ffi.cu:2889-2942; target
 * "sbh_stochastic_bh_force", python/jax_sbh_ffi.py:1180-1219). */
int sbh_stochastic_bh_force(void* stream, const SbhBuffer* leaf_bytes,
                            const SbhBuffer* node_bytes, const SbhBuffer* contrib_bytes,
                            const SbhBuffer* queries, SbhBuffer* out,
                            int32_t samples_per_subdomain, int32_t seed, int32_t periodic,
                            float box_length, int32_t threads);

/* Replaces SoftenedStochasticBarnesHutForce (src/sbh_ffi.cu:2999-3064; target
 * "sbh_stochastic_bh_force_soft", python/jax_sbh_ffi.py:1252-1299). */
int sbh_stochastic_bh_force_soft(void* stream, const SbhBuffer* leaf_bytes,
                                 const SbhBuffer* node_bytes, const SbhBuffer* contrib_bytes,
                                 const SbhBuffer* queries, SbhBuffer* out,
                                 int32_t samples_per_subdomain, int32_t seed,
                                 float softening_scale, float cutoff_scale, int32_t periodic,
                                 float box_length, int32_t threads);

/* ---- CPU-layout trees: Tree2D<1,L> and Tree3D<1,L>, L = 1, 2, 3 -------------
 *
 * Replaces the reference's host builders sbh.Quadtree1..3 / sbh.Octree1..3
 * (include/tree2d.h:237-407, include/tree3d.h:238-413, include/py_tree.h:
 * 20-73; their byte exports src/sbh.cpp:227-235), used by
 * build_quadtree_buffers(_with_contrib) and build_octree_buffers(backend=
 * "cpu") (python/jax_sbh_ffi.py:658-742, level 3) and by the pybind11 GPU
 * module's trees (python/jax_sbh.py:132-163, levels 1-3).  Leaf size 1,
 * maximum depth 100, 2^level cells per axis, nodes in breadth-first order
 * with each node's children contiguous in cell order, moments aggregated
 * bottom-up in fp32.  Node records Tree2DNode<1,L> are 136 / 232 / 616 bytes,
 * Tree3DNode<1,L> 184 / 632 / 4272 bytes for L = 1 / 2 / 3.
 * Built on the GPU in two calls, because the node count sizes the caller's
 * node buffer:
 *   sbh_tree_build: points (N, dim), normals (N, dim), masses (N,),
 *     min_corner / max_corner (dim,), all device float32; `alpha` is every
 *     point's alpha.  Builds into state kept per (device, stream),
 *     synchronizes the stream and returns the node count.  Not capturable.
 *   sbh_tree_emit: writes the last tree built on `stream`: leaf_out N*24 /
 *     N*32 bytes, node_out num_nodes * node size, contrib_out num_nodes*24 /
 *     num_nodes*32 bytes (may be NULL).
 * Points that share a cell for 100 levels form one leaf; their order there
 * is the input order (the reference's std::partition leaves it unspecified);
 * elsewhere every leaf holds one point and the bytes equal the reference's.
 * Limits: N <= 2^26 (dim 2) or 2^23 (dim 3). */
int sbh_tree_build(void* stream, int32_t dim, int32_t level, const SbhBuffer* points,
                   const SbhBuffer* normals, const SbhBuffer* masses,
                   const SbhBuffer* min_corner, const SbhBuffer* max_corner, float alpha,
                   int64_t* num_nodes);
int sbh_tree_emit(void* stream, SbhBuffer* leaf_out, SbhBuffer* node_out,
                  SbhBuffer* contrib_out);

/* The pybind11 GPU module's evaluations over these trees (python/jax_sbh.py
 * build_gpu_tree + eval_*; src/sbh_gpu.cu:324-550): kind 0 gravityPotential
 * -m / (|p - q| + 0.1), 1 signedSolidAngle<dim>, 2 smoothDistExp
 * exp(-alpha |p - q|); out (M,).  samples_per_subdomain < 0: the warp-voted
 * walk with `beta`; >= 0: the multi-level control-variate sampler (contrib
 * required; the module's seed 0 for every query).  Both run in 32-query
 * blocks, as the module launches them.  Queries (M, dim); leaf / node /
 * contrib bytes of a Tree{2,3}D<1, level> (sbh_tree_build / sbh_tree_emit).
 * Replaces evalBarnesHut_{gravity,wn,smoothdist} (src/sbh_gpu.cu:324-434) and
 * multiLevelPrefixControlVariateSampler_{gravity,wn,smoothdist} (436-550). */
int sbh_tree_eval(void* stream, int32_t dim, int32_t level, int32_t kind,
                  const SbhBuffer* leaf_bytes, const SbhBuffer* node_bytes,
                  const SbhBuffer* contrib_bytes, const SbhBuffer* queries, SbhBuffer* out,
                  float beta, int32_t samples_per_subdomain);

/* Replaces evalBruteForce_{gravity,wn,smoothdist} (src/sbh_gpu.cu:213-322,
 * include/brute_force_gpu.h:18-66): per query the compensated (Kahan) fp32
 * sum of the kind's kernel over the leaf records in their order. */
int sbh_tree_bruteforce(void* stream, int32_t dim, int32_t kind, const SbhBuffer* leaf_bytes,
                        const SbhBuffer* queries, SbhBuffer* out);

/* sizeof(LeafData<2>) = 24 and sizeof(Tree2DNode<1,3>) = 616. */
size_t sbh_leaf2_size(void);
size_t sbh_node2_size(void);

/* ---- winding numbers (K1) -------------------------------------------------
 *
 * Sources are oriented area elements: centroid p, unit normal n, area a.  The
 * value at q is sum a n.(p - q) / (4 pi (|p - q|^3 + 1e-5))
 * (signedSolidAngle<3>, include/kernel.h:73-80).  The reference exposes these
 * only through its pybind11 module over CPU-built trees (src/sbh_gpu.cu:
 * eval_brute_force_wn 53-64, eval_barnes_hut_wn 100-120,
 * multi_level_prefix_control_variate_wn 166-187); there is no XLA target, so
 * these are new targets with the reference's arithmetic.  Results: (M,). */

/* Replaces evalBruteForce_wn / evalBruteForceGPU_wn (src/sbh_gpu.cu:53-64,
 * 250-282; compensated summation of include/brute_force_gpu.h:18-66).
 * points (N,3), normals (N,3), areas (N,), queries (M,3). */
int sbh_bruteforce_wn(void* stream, const SbhBuffer* points,
                      const SbhBuffer* normals, const SbhBuffer* areas,
                      const SbhBuffer* queries, SbhBuffer* out, int32_t threads);

/* Replaces evalBarnesHut_wn (src/sbh_gpu.cu:100-120, 363-397) on octree8
 * records built by sbh_build_octree8_buffers from (centroids, normals,
 * areas): the octree8 MAC of src/sbh_ffi.cu:1878-1893, node record =
 * area A and area-weighted normal at the area-weighted centroid. */
int sbh_barnes_hut_wn_octree8(void* stream, const SbhBuffer* leaf_bytes,
                              const SbhBuffer* node_bytes,
                              const SbhBuffer* queries, SbhBuffer* out,
                              float beta, int32_t threads);

/* Replaces multiLevelPrefixControlVariateSampler_wn (src/sbh_gpu.cu:166-187,
 * 474-510) with the octree8 stochastic walk of src/sbh_ffi.cu:2146-2291
 * (depth-3 subdomains, Philox4x32-10 sample paths as
 * sbh_stochastic_bh_force_octree8). */
int sbh_stochastic_bh_wn_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                 const SbhBuffer* node_bytes,
                                 const SbhBuffer* queries, SbhBuffer* out,
                                 int32_t samples_per_subdomain, int32_t seed,
                                 float beta, int32_t threads);

/* Potential and smooth-distance fields (the reference's pybind11 module
 * src/sbh_gpu.cu, which evaluates them over CPU-built trees; here over
 * octree8 records, alpha = 1 in every record as src/sbh_ffi.cu:1741, 1860
 * writes).  Each output has shape (M,).
 *   sbh_barnes_hut_potential_octree8      replaces evalBarnesHut_gravity
 *                                         (src/sbh_gpu.cu:79-98, 324-361)
 *   sbh_stochastic_bh_potential_octree8   multiLevelPrefixControlVariateSampler_gravity
 *                                         (src/sbh_gpu.cu:143-164, 436-472)
 *   sbh_bruteforce_smoothdist             evalBruteForce_smoothdist
 *                                         (src/sbh_gpu.cu:66-77, 287-322; Kahan sum,
 *                                         include/brute_force_gpu.h:18-66)
 *   sbh_barnes_hut_smoothdist_octree8     evalBarnesHut_smoothdist
 *                                         (src/sbh_gpu.cu:122-141, 400-434)
 *   sbh_stochastic_bh_smoothdist_octree8  multiLevelPrefixControlVariateSampler_smoothdist
 *                                         (src/sbh_gpu.cu:189-211, 512-550)
 * The potential brute force (evalBruteForce_gravity, src/sbh_gpu.cu:39-51)
 * is sbh_bruteforce_gravity above. */
int sbh_barnes_hut_potential_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                     const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                     SbhBuffer* out, float beta, int32_t threads);
int sbh_stochastic_bh_potential_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                        const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                        SbhBuffer* out, int32_t samples_per_subdomain,
                                        int32_t seed, float beta, int32_t threads);
int sbh_bruteforce_smoothdist(void* stream, const SbhBuffer* points, const SbhBuffer* alphas,
                              const SbhBuffer* queries, SbhBuffer* out, int32_t threads);
int sbh_barnes_hut_smoothdist_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                      const SbhBuffer* node_bytes, const SbhBuffer* queries,
                                      SbhBuffer* out, float beta, int32_t threads);
int sbh_stochastic_bh_smoothdist_octree8(void* stream, const SbhBuffer* leaf_bytes,
                                         const SbhBuffer* node_bytes,
                                         const SbhBuffer* queries, SbhBuffer* out,
                                         int32_t samples_per_subdomain, int32_t seed,
                                         float beta, int32_t threads);

/* Adaptive-tree winding numbers (no reference counterpart: the reference
 * walks a CPU-built 512-ary tree).  The tree is built over the centroids
 * with the areas as weights; queries are Morton-sorted in their own cube and
 * walked in groups of 64 with the group MAC of the gravity pipeline; an
 * accepted node is the dipole sum(a n) at the area-weighted centroid (the
 * reference's node record).
 * workspace: u8 buffer of sbh_winding_number_workspace_bytes() bytes, or NULL.
 * stats: NULL or int64[8] {far interactions, near pairs, groups, nodes,
 * levels, overflow, 0, 0}. */
int64_t sbh_winding_number_workspace_bytes(int64_t num_elements, int64_t num_queries,
                                           int32_t max_depth, int32_t leaf_size);
int sbh_winding_number(void* stream, const SbhBuffer* points,
                       const SbhBuffer* normals, const SbhBuffer* areas,
                       const SbhBuffer* queries, SbhBuffer* out, float beta,
                       int32_t max_depth, int32_t leaf_size,
                       SbhBuffer* workspace, SbhBuffer* stats);

/* ---- adaptive Morton octree pipeline (new targets) ----------------------
 *
 * No reference counterpart: the reference's octree8 is dense (8^D cells,
 * infeasible at D = 10), has no leaf-size parameter, and its custom_vjp
 * returns the query-side gradient only (python/jax_sbh_ffi.py:1855-1891).
 * These targets build an adaptive octree on the device (bounds = data bounds
 * -/+ 1e-3, codes with the reference quantiser at depth max_depth, a cell
 * splits while it holds more than leaf_size points), then walk it with one
 * 64-lane wavefront per group of 64 Morton-consecutive points.  A cell is used
 * as a multipole for a group iff the distance from its centre to the group's
 * bounding box is >= beta * cell diagonal (the reference MAC,
 * src/sbh_ffi.cu:1878-1893, applied conservatively to the whole group);
 * otherwise it is opened, and rejected leaves are summed exactly.
 *
 * Self-gravity only (queries are the points; the i == j pair is skipped).
 * Kernel: F_i = sum -m (r^2 + eps^2 + 1e-12)^(-3/2) d, no cutoff, not periodic.
 * multipole_order: 1 = monopole, 2 = monopole + quadrupole.  The VJP entry
 *            points also take 3: monopole terms (forces, the query-side
 *            gradient and the node adjoints dL/dX, dL/dM) plus the
 *            quadrupole term's adjoint with respect to the node's second
 *            moment, which does not depend on that moment.  Pushed down to
 *            the sources it adds the first-order term of their position
 *            gradients and the second-order term of their mass gradients
 *            (the same orders as 2, at about 3/4 of its adjoint work);
 *            forces returned with 3 are the monopole's.
 * workspace: u8 device buffer of sbh_gravity_workspace_bytes() bytes, or NULL
 *            for a library-owned cached buffer (allocated when it must grow).
 * stats:     NULL or int64[8] device buffer receiving {far (lane x node)
 *            interactions, near pairs, groups, nodes, levels, overflow, 0, 0}. */
int64_t sbh_gravity_workspace_bytes(int64_t n, int32_t max_depth, int32_t leaf_size);

/* Byte offsets of the workspace arrays, in this order: header, codes, perm,
 * sorted points, sorted cotangents, topo (int4 start,count,first_child,
 * nchild), prefix, parent, geo (float4 cx,cy,cz,diagonal), mp (float4 com,mass),
 * qd (2 x float4 second moment), adj, adjq, qgrad, sgrad, total, node capacity. */
int sbh_gravity_workspace_layout(int64_t n, int32_t max_depth, int32_t leaf_size,
                                 int64_t* offsets, int32_t count);

/* Forces only: points (N,3), masses (N) -> forces (N,3). */
int sbh_gravity_force(void* stream, const SbhBuffer* points,
                      const SbhBuffer* masses, SbhBuffer* forces, float beta,
                      float softening_scale, int32_t max_depth,
                      int32_t leaf_size, int32_t multipole_order,
                      SbhBuffer* workspace, SbhBuffer* stats);

/* Recompute-style VJP: rebuilds the tree and walks it once, returning the
 * full adjoint of L = sum_i c_i . F_i with respect to positions and masses
 * (grad_points (N,3), grad_masses (N)); forces (N,3) may be NULL. */
int sbh_gravity_force_vjp(void* stream, const SbhBuffer* points,
                          const SbhBuffer* masses, const SbhBuffer* cotangent,
                          SbhBuffer* forces, SbhBuffer* grad_points,
                          SbhBuffer* grad_masses, float beta,
                          float softening_scale, int32_t max_depth,
                          int32_t leaf_size, int32_t multipole_order,
                          SbhBuffer* workspace, SbhBuffer* stats);

/* ---- batched sets (new targets) ------------------------------------------
 *
 * Many independent self-gravity problems in one pipeline: set k is input rows
 * set_offsets[k] .. set_offsets[k+1]-1 (int32 device vector of num_sets + 1
 * entries rising from 0 to N).  Every set gets its own cube, tree and query
 * groups inside one forest, built and walked by the same launches as one
 * set (the persistent walk drains the groups of all sets from one queue).
 * Per set, results equal sbh_gravity_force / sbh_gravity_force_vjp on that
 * set alone.  Offsets that do not rise from 0 to N leave every set empty and
 * set stats[5] (overflow) to 2.  walk_blocks: persistent walk grid in
 * blocks (0 = default, 8192 waves); walk_threads: its block size, 64, 128,
 * 256 or 512 (0 = 256).  No reference counterpart (the reference runs one launch per
 * call, python/benchmark_octree8.py). */
int64_t sbh_gravity_batch_workspace_bytes(int64_t total_points, int32_t num_sets,
                                          int32_t max_depth, int32_t leaf_size);
int sbh_gravity_batch_force(void* stream, const SbhBuffer* points,
                            const SbhBuffer* masses, const SbhBuffer* set_offsets,
                            SbhBuffer* forces, float beta, float softening_scale,
                            int32_t max_depth, int32_t leaf_size,
                            int32_t multipole_order, int32_t walk_blocks,
                            int32_t walk_threads, SbhBuffer* workspace, SbhBuffer* stats);
int sbh_gravity_batch_force_vjp(void* stream, const SbhBuffer* points,
                                const SbhBuffer* masses, const SbhBuffer* cotangent,
                                const SbhBuffer* set_offsets, SbhBuffer* forces,
                                SbhBuffer* grad_points, SbhBuffer* grad_masses,
                                float beta, float softening_scale, int32_t max_depth,
                                int32_t leaf_size, int32_t multipole_order,
                                int32_t walk_blocks, int32_t walk_threads,
                                SbhBuffer* workspace, SbhBuffer* stats);

/* ---- stochastic Barnes-Hut on the adaptive octree (new targets) --------
 *
 * The multi-level prefix control-variate estimator of
 * stochastic_bh_force_octree8_kernel / its _vjp_kernel (src/sbh_ffi.cu:
 * 2146-2291, 2293-2442; python/jax_sbh_ffi.py:1450-1535 wraps them) and of
 * multi_level_prefix_control_variate_wn (src/sbh_gpu.cu:166-187), on the
 * adaptive tree instead of the dense octree8 records (infeasible at
 * max_depth 10).  Subdomains are the nodes at level 3 and the leaves above
 * it; each contributes its node term (or exact leaf sum), and each of
 * samples_per_subdomain samples walks from it towards a uniformly drawn
 * source with Russian roulette (continue with probability
 * min(beta_cur / ratio, 1), ratio = |q - cell centre| / cell diagonal),
 * adding n_sub / (P(path) count(node)) (sum of the node's children - node).
 * Queries are separate from the sources (as in the reference targets); they
 * are grouped by their own Morton tree (groups of <= 64), and the queries of
 * a group share the random draws (Philox4x32-10 keyed by seed, the group,
 * the subdomain, the sample and the step), each keeping its own roulette
 * probabilities.  E[estimate] = the exact sum.
 * Gravity: softened force (eps = softening_scale, no cutoff, not periodic);
 * multipole_order 1 (the reference's monopole node term) or 2 (+ quadrupole).
 * The VJP is the reference's: J^T c with respect to the query positions along
 * the same sample paths (cotangent (M, 3) -> out (M, 3)).
 * Winding numbers: node term = dipole sum(a n) at the area-weighted centroid,
 * leaves summed exactly (signedSolidAngle<3>, include/kernel.h:73-80).
 * workspace: u8 buffer of sbh_stochastic_workspace_bytes() bytes, or NULL.
 * stats: NULL or int64[8] {node terms, source pairs (both per query, summed),
 * query groups, source-tree nodes, levels, overflow, subdomains, 0}. */
int64_t sbh_stochastic_workspace_bytes(int64_t num_sources, int64_t num_queries,
                                       int32_t max_depth, int32_t leaf_size);
int sbh_stochastic_gravity_force(void* stream, const SbhBuffer* points,
                                 const SbhBuffer* masses, const SbhBuffer* queries,
                                 SbhBuffer* out, float beta, float softening_scale,
                                 int32_t max_depth, int32_t leaf_size,
                                 int32_t multipole_order, int32_t samples_per_subdomain,
                                 int32_t seed, SbhBuffer* workspace, SbhBuffer* stats);
int sbh_stochastic_gravity_force_vjp(void* stream, const SbhBuffer* points,
                                     const SbhBuffer* masses, const SbhBuffer* queries,
                                     const SbhBuffer* cotangent, SbhBuffer* out, float beta,
                                     float softening_scale, int32_t max_depth,
                                     int32_t leaf_size, int32_t multipole_order,
                                     int32_t samples_per_subdomain, int32_t seed,
                                     SbhBuffer* workspace, SbhBuffer* stats);
int sbh_stochastic_winding_number(void* stream, const SbhBuffer* points,
                                  const SbhBuffer* normals, const SbhBuffer* areas,
                                  const SbhBuffer* queries, SbhBuffer* out, float beta,
                                  int32_t max_depth, int32_t leaf_size,
                                  int32_t samples_per_subdomain, int32_t seed,
                                  SbhBuffer* workspace, SbhBuffer* stats);

/* ---- allocation-free gravity VJP -----------------------------------------
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): sbh_gravity_force_vjp with the same arguments except
 * the workspace, which is a library-owned static device array of
 * sbh_gravity_noalloc_scratch_bytes() bytes (64 MiB) per device, held until
 * the process ends.  No call allocates or frees device memory.  Problems whose
 * sbh_gravity_workspace_bytes(n, max_depth, leaf_size) exceeds that size are
 * rejected with SBH_INVALID_ARGUMENT (at max_depth 10: n <= 224,912 at leaf
 * size 32, n <= 143,008 at leaf size 16).
 * Every call on one device shares that device's array: calls that overlap
 * on one device (different streams or host threads) race; calls queued on
 * one stream are ordered.  Results equal sbh_gravity_force_vjp's. */
int64_t sbh_gravity_noalloc_scratch_bytes(void);
int sbh_gravity_force_vjp_noalloc(void* stream, const SbhBuffer* points,
                                  const SbhBuffer* masses, const SbhBuffer* cotangent,
                                  SbhBuffer* forces, SbhBuffer* grad_points,
                                  SbhBuffer* grad_masses, float beta,
                                  float softening_scale, int32_t max_depth,
                                  int32_t leaf_size, int32_t multipole_order,
                                  SbhBuffer* stats);

/* HIP graph capture of library calls on `stream` (no reference counterpart;
 * XLA captures its own graphs).  Between sbh_graph_begin and sbh_graph_end
 * only calls with caller-provided workspaces may be made.  sbh_graph_end
 * instantiates the graph into *graph_exec; sbh_graph_launch replays it. */
int sbh_graph_begin(void* stream);
int sbh_graph_end(void* stream, void** graph_exec);
int sbh_graph_launch(void* graph_exec, void* stream);
int sbh_graph_destroy(void* graph_exec);

/* Phase timing for the sbh_gravity_* (single set and batch),
 * sbh_winding_number and sbh_stochastic_* targets (no reference counterpart;
 * the reference times from Python around dispatch).  When enabled, each call
 * records HIP events on its stream around the tree build(s) and around the
 * walk kernel.  sbh_profile_collect waits for the recorded events and writes
 * up to 12 values, {calls, build ms, walk ms} for gravity forward, gravity
 * adjoint, winding numbers and the stochastic walks in turn, summed since
 * the previous collect.  Not for use while a graph is being captured. */
void sbh_profile_enable(int32_t on);
int sbh_profile_collect(double* out, int32_t count);

#ifdef __cplusplus
}
#endif

#endif /* SBH_AMD_H_ */
