"""ctypes front-end of the CPU oracle (oracle/sbh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LEAF3_DTYPE = np.dtype(
    [("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("mass", "<f4"), ("alpha", "<f4")]
)
NODE8_DTYPE = np.dtype(
    [
        ("start", "<u4"),
        ("end", "<u4"),
        ("parent", "<u4"),
        ("first_child", "<u4"),
        ("child", "<u4", (8,)),
        ("child_mask", "u1"),
        ("depth", "u1"),
        ("is_leaf", "u1"),
        ("pad0", "u1"),
        ("cx", "<f4"),
        ("cy", "<f4"),
        ("cz", "<f4"),
        ("half_size", "<f4"),
        ("diameter", "<f4"),
        ("contrib", LEAF3_DTYPE),
    ]
)
assert LEAF3_DTYPE.itemsize == 32 and NODE8_DTYPE.itemsize == 104


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def _declare(L: ctypes.CDLL) -> None:
    L.orc_quantize_axis.restype = ctypes.c_uint32
    L.orc_quantize_axis.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    L.orc_morton3d.restype = ctypes.c_uint32
    L.orc_morton3d.argtypes = [ctypes.c_uint32] * 3 + [ctypes.c_int]
    L.orc_morton_codes.argtypes = [_f32p, _i64, _f32p, ctypes.c_float, ctypes.c_int, _u32p]
    L.orc_default_bounds.argtypes = [_f32p, _i64, _f32p, _f32p]
    L.orc_span.restype = ctypes.c_float
    L.orc_span.argtypes = [_f32p, _f32p]
    L.orc_stable_order.argtypes = [_u32p, _i64, ctypes.c_int, _u32p]
    L.orc_octree8_total_nodes.restype = ctypes.c_uint64
    L.orc_octree8_total_nodes.argtypes = [ctypes.c_int]
    L.orc_build_octree8.restype = ctypes.c_int
    L.orc_build_octree8.argtypes = [_f32p, _f32p, _f32p, _i64, _f32p, _f32p, ctypes.c_int,
                                    _u8p, _u8p, _u8p]
    L.orc_direct_force_soft.argtypes = [_f32p, _f32p, _i64, _f32p, _i64, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_double, _f64p]
    L.orc_direct_force_soft_vjp.argtypes = [_f32p, _f32p, _i64, _f32p, _f32p, _i64,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, _f64p]
    L.orc_direct_full_vjp.argtypes = [_f32p, _f32p, _i64, _f32p, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double, _f64p, _f64p]
    L.orc_direct_full_vjp_cut.argtypes = [_f32p, _f32p, _i64, _f32p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                          _f64p, _f64p]
    L.orc_bruteforce_gravity.argtypes = [_f32p, _f32p, _i64, ctypes.c_int, _f32p, _i64,
                                         ctypes.c_int, ctypes.c_double, _f64p]
    L.orc_bh_octree8.argtypes = [_u8p, _u8p, _f32p, _vp, _i64, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_int, ctypes.c_float, _f64p, ctypes.POINTER(_i64)]
    L.orc_sbh_octree8.argtypes = [_u8p, _u8p, _f32p, _vp, _i64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_float, _f64p]
    L.orc_tree_build.restype = _i64
    L.orc_tree_build.argtypes = [_u32p, _i64, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_float,
                                 _i64, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _f32p]
    L.orc_tree_moments.argtypes = [_i32p, _i32p, _i64, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_tree_groups.restype = _i64
    L.orc_tree_groups.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _i32p, ctypes.c_int,
                                  ctypes.c_int, _i32p, _i32p]
    L.orc_group_boxes.argtypes = [_f32p, _i64, _i32p, _i32p, _f32p]
    L.orc_tree_eval.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _f32p, _f64p, _f64p, _f64p,
                                _f64p, _f64p, _i64, _f32p, _f64p, _i64, _i64, _i32p, _i32p,
                                ctypes.c_float, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                _vp, _f64p, _vp, _vp, _vp, ctypes.POINTER(_i64)]
    L.orc_tree_eval_range.argtypes = (list(L.orc_tree_eval.argtypes[:15]) +
                                      list(L.orc_tree_eval.argtypes[16:]) + [_i64, _i64])
    L.orc_philox4x32.argtypes = [_u32p, _u32p, _u32p]
    L.orc_node3_size.restype = ctypes.c_size_t
    L.orc_build_level3.argtypes = [_f32p, _f32p, _f32p, _i64, _f32p, _f32p, _u8p, _u8p, _u8p]
    L.orc_l3_eval.argtypes = [_u8p, _i64, _u8p, _i64, _f32p, _i64, ctypes.c_int, ctypes.c_float,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, ctypes.c_float, _f64p]
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_node2_size.restype = ctypes.c_size_t
    L.orc_nd_node_size.restype = ctypes.c_size_t
    L.orc_nd_node_size.argtypes = [ctypes.c_int, ctypes.c_int]
    L.orc_nd_build.restype = ctypes.c_void_p
    L.orc_nd_build.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p, _i64, _f32p,
                               _f32p, ctypes.c_float]
    L.orc_nd_num_nodes.restype = _i64
    L.orc_nd_num_nodes.argtypes = [_vp]
    L.orc_nd_emit.argtypes = [_vp, _u8p, _u8p, _u8p]
    L.orc_nd_free.argtypes = [_vp]
    L.orc_nd_eval.argtypes = [ctypes.c_int, ctypes.c_int, _u8p, _i64, _u8p, _i64, _u8p, _i64,
                              _f32p, _i64, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.c_double, _f64p]
    L.orc_nd_brute.argtypes = [ctypes.c_int, _u8p, _i64, _f32p, _i64, ctypes.c_int, _f64p]
    L.orc_cpu_tree_build.restype = ctypes.c_void_p
    L.orc_cpu_tree_build.argtypes = [_f32p, _vp, _f32p, _i64, _f32p, _f32p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
    L.orc_cpu_tree_free.argtypes = [_vp]
    L.orc_cpu_tree_nodes.restype = _i64
    L.orc_cpu_tree_nodes.argtypes = [_vp]
    L.orc_cpu_tree_arrays.argtypes = [_vp, _i64p, _i64p, _i64p, _i32p,
                                      np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"),
                                      _f32p, _f32p, _f32p]
    L.orc_cpu_bh_eval.restype = _i64
    L.orc_cpu_bh_eval.argtypes = [_vp, _f32p, _i64, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_double, ctypes.c_int, _f64p]
    L.orc_cpu_mlpcv.argtypes = [_vp, _f32p, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                ctypes.c_double, ctypes.c_int, _f64p]
    L.orc_bruteforce_wn.argtypes = [_f32p, _f32p, _f32p, _i64, _f32p, _i64, _f64p]
    L.orc_bruteforce_smoothdist.argtypes = [_f32p, _f32p, _i64, _f32p, _i64, _f64p]
    L.orc_wn_moments.argtypes = [_i32p, _i32p, _i64, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_wn_eval.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _f32p, _f64p, _f64p, _f64p,
                              _f64p, _f64p, _f64p, _f32p, _f64p, _i64, _i64, _i32p, _i32p,
                              ctypes.c_float, _f64p, ctypes.POINTER(_i64)]
    L.orc_tree_subdomains.restype = _i64
    L.orc_tree_subdomains.argtypes = [_i64, _i32p, _i32p, _f64p, _i32p]
    L.orc_tree_stoch.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _i32p, _f32p, _f64p, _f64p,
                                 _f64p, _f64p, _f64p, _f64p, _f64p, _f32p, _i64, _i64, _i32p,
                                 _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _vp, _f64p,
                                 ctypes.POINTER(_i64)]


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- Morton / bounds
def quantize_axis(v, min_v, span, bits):
    return int(lib().orc_quantize_axis(float(np.float32(v)), float(np.float32(min_v)),
                                       float(np.float32(span)), int(bits)))


def morton3d(x, y, z, bits):
    return int(lib().orc_morton3d(int(x), int(y), int(z), int(bits)))


def default_bounds(points):
    p = _c(points, np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().orc_default_bounds(p, p.shape[0], lo, hi)
    return lo, hi


def span(min_corner, max_corner):
    return np.float32(lib().orc_span(_c(min_corner, np.float32), _c(max_corner, np.float32)))


def morton_codes(points, min_corner, span_, bits):
    p = _c(points, np.float32)
    out = np.zeros(p.shape[0], np.uint32)
    lib().orc_morton_codes(p, p.shape[0], _c(min_corner, np.float32), float(span_), bits, out)
    return out


def stable_order(codes, bits3):
    c = _c(codes, np.uint32)
    out = np.zeros(c.shape[0], np.uint32)
    lib().orc_stable_order(c, c.shape[0], int(bits3), out)
    return out


# ---------------------------------------------------------------- octree8
def octree8_total_nodes(max_depth):
    return int(lib().orc_octree8_total_nodes(int(max_depth)))


def build_octree8(points, normals, masses, min_corner, max_corner, max_depth):
    p = _c(points, np.float32)
    nrm = _c(normals, np.float32)
    m = _c(masses, np.float32)
    n = p.shape[0]
    if not 1 <= int(max_depth) <= 10:
        raise ValueError("max_depth must be in [1, 10]")
    t = octree8_total_nodes(max_depth)
    leaf = np.zeros(n * 32, np.uint8)
    node = np.zeros(t * 104, np.uint8)
    contrib = np.zeros(t * 32, np.uint8)
    rc = lib().orc_build_octree8(p, nrm, m, n, _c(min_corner, np.float32),
                                 _c(max_corner, np.float32), int(max_depth), leaf, node, contrib)
    if rc != 0:
        raise ValueError("invalid octree8 arguments")
    return leaf, node, contrib


def decode_nodes8(node_bytes):
    return np.frombuffer(np.ascontiguousarray(node_bytes, dtype=np.uint8).tobytes(), NODE8_DTYPE)


def decode_leaf3(leaf_bytes):
    return np.frombuffer(np.ascontiguousarray(leaf_bytes, dtype=np.uint8).tobytes(), LEAF3_DTYPE)


# ---------------------------------------------------------------- direct sums
def direct_force_soft(points, masses, queries, eps, rc, periodic=False, box=1.0):
    p, m, q = _c(points, np.float32), _c(masses, np.float32), _c(queries, np.float32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_direct_force_soft(p, m, p.shape[0], q, q.shape[0], float(eps), float(rc),
                                int(bool(periodic)), float(box), out)
    return out


def direct_force_soft_vjp(points, masses, queries, cot, eps, rc, periodic=False, box=1.0):
    p, m, q = _c(points, np.float32), _c(masses, np.float32), _c(queries, np.float32)
    g = _c(cot, np.float32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_direct_force_soft_vjp(p, m, p.shape[0], q, g, q.shape[0], float(eps), float(rc),
                                    int(bool(periodic)), float(box), out)
    return out


def direct_full_vjp(points, masses, cot, eps, periodic=False, box=1.0):
    p, m, g = _c(points, np.float32), _c(masses, np.float32), _c(cot, np.float32)
    gx = np.zeros((p.shape[0], 3), np.float64)
    gm = np.zeros(p.shape[0], np.float64)
    lib().orc_direct_full_vjp(p, m, p.shape[0], g, float(eps), int(bool(periodic)), float(box),
                              gx, gm)
    return gx, gm


def direct_full_vjp_cut(points, masses, cot, eps, rc, periodic=False, box=1.0):
    p, m, g = _c(points, np.float32), _c(masses, np.float32), _c(cot, np.float32)
    gx = np.zeros((p.shape[0], 3), np.float64)
    gm = np.zeros(p.shape[0], np.float64)
    lib().orc_direct_full_vjp_cut(p, m, p.shape[0], g, float(eps), float(rc),
                                  int(bool(periodic)), float(box), gx, gm)
    return gx, gm


def bruteforce_gravity(points, masses, queries, periodic=False, box=1.0):
    p, m, q = _c(points, np.float32), _c(masses, np.float32), _c(queries, np.float32)
    out = np.zeros(q.shape[0], np.float64)
    lib().orc_bruteforce_gravity(p, m, p.shape[0], p.shape[1], q, q.shape[0],
                                 int(bool(periodic)), float(box), out)
    return out


# ---------------------------------------------------------------- octree8 walks
KIND_GRAV, KIND_SOFT, KIND_SOFT_VJP, KIND_GRAV_VJP, KIND_WN = 0, 1, 2, 3, 4
KIND_POTENTIAL, KIND_SMOOTH = 5, 6
_SCALAR_KINDS = (KIND_WN, KIND_POTENTIAL, KIND_SMOOTH)


def bh_octree8(leaf_bytes, node_bytes, queries, kind, beta, eps=0.0, rc=0.0, cot=None,
               prune_enabled=False, prune_r_cut_mult=4.0, periodic=False, box=1.0):
    leaf = _c(leaf_bytes, np.uint8)
    node = _c(node_bytes, np.uint8)
    q = _c(queries, np.float32)
    g = None if cot is None else _c(cot, np.float32)
    out = np.zeros((q.shape[0], 3), np.float64)
    cnt = _i64(0)
    lib().orc_bh_octree8(leaf, node, q, _ptr(g), q.shape[0], int(kind), float(beta),
                         float(eps), float(rc), int(bool(prune_enabled)),
                         float(prune_r_cut_mult), int(bool(periodic)), float(box), out,
                         ctypes.byref(cnt))
    if kind in _SCALAR_KINDS:
        return out[:, 0].copy(), int(cnt.value)
    return out, int(cnt.value)


def sbh_octree8(leaf_bytes, node_bytes, queries, kind, samples, seed, beta, eps=0.0, rc=0.0,
                cot=None, prune_enabled=False, prune_r_cut_mult=4.0, periodic=False, box=1.0):
    """Stochastic walk (same Philox sample paths as the GPU kernels)."""
    leaf = _c(leaf_bytes, np.uint8)
    node = _c(node_bytes, np.uint8)
    q = _c(queries, np.float32)
    g = None if cot is None else _c(cot, np.float32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_sbh_octree8(leaf, node, q, _ptr(g), q.shape[0], int(kind), int(samples), int(seed),
                          float(beta), float(eps), float(rc), int(bool(prune_enabled)),
                          float(prune_r_cut_mult), int(bool(periodic)), float(box), out)
    if kind in _SCALAR_KINDS:
        return out[:, 0].copy()
    return out


NODE3_DTYPE = np.dtype({
    "names": ["start", "end", "buffer_idx", "parent_idx", "bmin", "bmax", "diameter", "contrib",
              "bitmask", "child", "num_children", "depth", "max_depth", "is_leaf"],
    "formats": ["<u8", "<u8", "<u8", "<u8", ("<f4", 3), ("<f4", 3), "<f4", LEAF3_DTYPE,
                ("<u8", 8), ("<u8", 512), "<i4", "<i4", "<i4", "u1"],
    "offsets": [0, 8, 16, 24, 32, 44, 56, 60, 96, 160, 4256, 4260, 4264, 4268],
    "itemsize": 4272})
L3_MAX_NODES = 513
L3_POTENTIAL, L3_FORCE, L3_SOFT = 0, 1, 2


def build_level3(points, normals, masses, min_corner, max_corner):
    """Level-3 buffers (leaf N*32, node 513*4272, contrib 513*32 bytes)."""
    p = _c(points, np.float32)
    nv = _c(normals, np.float32)
    m = _c(masses, np.float32)
    n = p.shape[0]
    leaf = np.zeros(n * 32, np.uint8)
    node = np.zeros(L3_MAX_NODES * 4272, np.uint8)
    contrib = np.zeros(L3_MAX_NODES * 32, np.uint8)
    lib().orc_build_level3(p, nv, m, n, _c(min_corner, np.float32), _c(max_corner, np.float32),
                           leaf, node, contrib)
    return leaf, node, contrib


def decode_nodes3(node_bytes):
    return np.frombuffer(np.ascontiguousarray(node_bytes, np.uint8).tobytes(), NODE3_DTYPE)


def l3_eval(leaf_bytes, node_bytes, queries, kind, *, beta=2.0, threads=256, samples=-1,
            eps=0.0, rc=0.0, periodic=False, box=1.0):
    """Level-3 walks: classic vote walk (samples < 0) or stochastic."""
    leaf = _c(leaf_bytes, np.uint8)
    node = _c(node_bytes, np.uint8)
    q = _c(queries, np.float32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_l3_eval(leaf, leaf.size // 32, node, node.size // 4272, q, q.shape[0], int(kind),
                      float(beta), int(threads), int(samples), float(eps), float(rc),
                      int(bool(periodic)), float(box), out)
    return out[:, 0].copy() if kind == L3_POTENTIAL else out


LEAF2_DTYPE = np.dtype(
    [("pos", "<f4", (2,)), ("normal", "<f4", (2,)), ("mass", "<f4"), ("alpha", "<f4")]
)
NODE2_DTYPE = np.dtype({
    "names": ["start", "end", "buffer_idx", "parent_idx", "bmin", "bmax", "diameter", "contrib",
              "bitmask", "child", "num_children", "depth", "max_depth", "is_leaf"],
    "formats": ["<u8", "<u8", "<u8", "<u8", ("<f4", 2), ("<f4", 2), "<f4", LEAF2_DTYPE,
                ("<u8", 1), ("<u8", 64), "<i4", "<i4", "<i4", "u1"],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 52, 80, 88, 600, 604, 608, 612],
    "itemsize": 616})
assert LEAF2_DTYPE.itemsize == 24


ND_WN, ND_SMOOTH = 3, 4  # signedSolidAngle<N>, smoothDistExp (kinds 0-2 as L3_*)


def nd_node_dtype(dim, level):
    """Tree2DNode<1, level> / Tree3DNode<1, level> record dtype."""
    size = int(lib().orc_nd_node_size(dim, level))
    cells = 1 << (dim * level)
    words = max(1, cells // 64)
    leaf = LEAF2_DTYPE if dim == 2 else LEAF3_DTYPE
    o_contrib = 36 + 8 * dim
    o_mask = (o_contrib + leaf.itemsize + 7) // 8 * 8
    o_child = o_mask + 8 * words
    o_num = o_child + 8 * cells
    return np.dtype({
        "names": ["start", "end", "buffer_idx", "parent_idx", "bmin", "bmax", "diameter",
                  "contrib", "bitmask", "child", "num_children", "depth", "max_depth",
                  "is_leaf"],
        "formats": ["<u8", "<u8", "<u8", "<u8", ("<f4", dim), ("<f4", dim), "<f4", leaf,
                    ("<u8", words), ("<u8", cells), "<i4", "<i4", "<i4", "u1"],
        "offsets": [0, 8, 16, 24, 32, 32 + 4 * dim, 32 + 8 * dim, o_contrib, o_mask, o_child,
                    o_num, o_num + 4, o_num + 8, o_num + 12],
        "itemsize": size})


def nd_build(points, normals, masses, min_corner, max_corner, alpha=1.0, level=3):
    """Tree2D<1, level> (points (N, 2)) or Tree3D<1, level> (points (N, 3))
    bytes: (leaf, node, contrib) uint8 arrays, as the reference's
    sbh.Quadtree<level> / sbh.Octree<level> leaf_data_bytes /
    node_buffer_bytes / contrib_data_bytes."""
    p = _c(points, np.float32)
    dim = p.shape[1]
    nv = _c(normals, np.float32)
    m = _c(masses, np.float32)
    n = p.shape[0]
    L = lib()
    t = L.orc_nd_build(dim, int(level), p, nv, m, n, _c(min_corner, np.float32),
                       _c(max_corner, np.float32), float(alpha))
    try:
        nn = int(L.orc_nd_num_nodes(t))
        lsz = 24 if dim == 2 else 32
        nsz = int(L.orc_nd_node_size(dim, int(level)))
        leaf = np.zeros(max(n, 1) * lsz, np.uint8)
        node = np.zeros(nn * nsz, np.uint8)
        contrib = np.zeros(nn * lsz, np.uint8)
        L.orc_nd_emit(t, leaf, node, contrib)
    finally:
        L.orc_nd_free(t)
    return leaf[: n * lsz].copy(), node, contrib


def decode_nodes2(node_bytes):
    return np.frombuffer(np.ascontiguousarray(node_bytes, np.uint8).tobytes(), NODE2_DTYPE)


def decode_nodes(node_bytes, dim, level):
    return np.frombuffer(np.ascontiguousarray(node_bytes, np.uint8).tobytes(),
                         nd_node_dtype(dim, level))


def nd_eval(leaf_bytes, node_bytes, contrib_bytes, queries, kind, *, beta=2.0, threads=256,
            samples=-1, seed=0, eps=0.0, rc=0.0, periodic=False, box=1.0, level=3,
            seed_per_query=True):
    """Vote walk (samples < 0) or control-variate sampler (samples >= 0) over
    Tree{2,3}DNode<1, level> buffers; the layout follows queries' width.
    seed_per_query: roulette stream of seed + query (the FFI targets) or of
    seed alone (the pybind11 module)."""
    q = _c(queries, np.float32)
    dim = q.shape[1]
    lsz = 24 if dim == 2 else 32
    nsz = int(lib().orc_nd_node_size(dim, int(level)))
    leaf = _c(leaf_bytes, np.uint8)
    node = _c(node_bytes, np.uint8)
    contrib = _c(contrib_bytes if contrib_bytes is not None else np.zeros(0, np.uint8), np.uint8)
    out = np.zeros((q.shape[0], dim), np.float64)
    lib().orc_nd_eval(dim, int(level), leaf if leaf.size else np.zeros(1, np.uint8),
                      leaf.size // lsz, node, node.size // nsz,
                      contrib if contrib.size else np.zeros(1, np.uint8), contrib.size // lsz,
                      q, q.shape[0], int(kind), float(beta), int(threads), int(samples),
                      int(seed), int(bool(seed_per_query)), float(eps), float(rc),
                      int(bool(periodic)), float(box), out)
    if kind in (L3_POTENTIAL, ND_WN, ND_SMOOTH):  # written as a flat (M,) array
        return out.reshape(-1)[: q.shape[0]].copy()
    return out


def nd_brute(leaf_bytes, queries, kind):
    """fp64 sum of a scalar kernel (L3_POTENTIAL, ND_WN, ND_SMOOTH) over the
    leaf records."""
    q = _c(queries, np.float32)
    dim = q.shape[1]
    leaf = _c(leaf_bytes, np.uint8)
    out = np.zeros(q.shape[0], np.float64)
    lib().orc_nd_brute(dim, leaf if leaf.size else np.zeros(1, np.uint8),
                       leaf.size // (24 if dim == 2 else 32), q, q.shape[0], int(kind), out)
    return out


CPU_POTENTIAL, CPU_WN, CPU_SMOOTHDIST, CPU_SOFT_FORCE = 0, 1, 2, 3


def num_threads():
    """Threads the oracle's OpenMP loops use."""
    return int(lib().orc_num_threads())


class CpuTree:
    """Restatement of the reference's CPU Tree3D<LEAF_SIZE, DIM_SPLIT_LOG>
    (see oracle/sbh_cpu.c): dim_split_log 1 is a true octree, 3 the
    512-ary Octree3."""

    def __init__(self, points, masses, normals=None, min_corner=None, max_corner=None,
                 leaf_size=1, dim_split_log=3, max_depth=100):
        p = _c(points, np.float32)
        self.n = p.shape[0]
        if min_corner is None or max_corner is None:
            lo, hi = default_bounds(p)
            min_corner = lo if min_corner is None else min_corner
            max_corner = hi if max_corner is None else max_corner
        nv = None if normals is None else _c(normals, np.float32)
        self._nv = nv
        self._h = lib().orc_cpu_tree_build(p, _ptr(nv), _c(masses, np.float32), self.n,
                                           _c(min_corner, np.float32),
                                           _c(max_corner, np.float32), int(leaf_size),
                                           int(dim_split_log), int(max_depth))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_cpu_tree_free(h)
            self._h = None

    @property
    def num_nodes(self):
        return int(lib().orc_cpu_tree_nodes(self._h))

    def arrays(self):
        nn = self.num_nodes
        a = dict(start=np.zeros(nn, np.int64), end=np.zeros(nn, np.int64),
                 first_child=np.zeros(nn, np.int64), nchild=np.zeros(nn, np.int32),
                 is_leaf=np.zeros(nn, np.uint8), mass=np.zeros(nn, np.float32),
                 com=np.zeros((nn, 3), np.float32), leaf_pos=np.zeros((self.n, 3), np.float32))
        lib().orc_cpu_tree_arrays(self._h, a["start"], a["end"], a["first_child"], a["nchild"],
                                  a["is_leaf"], a["mass"], a["com"], a["leaf_pos"])
        return a

    def eval(self, queries, kind=CPU_POTENTIAL, beta=2.0, eps=0.0, threads=0):
        q = _c(queries, np.float32)
        out = np.zeros((q.shape[0], 3), np.float64)
        n_int = lib().orc_cpu_bh_eval(self._h, q, q.shape[0], int(kind), float(beta),
                                      float(eps), int(threads), out)
        res = out if kind == CPU_SOFT_FORCE else out[:, 0].copy()
        return res, int(n_int)

    def mlpcv(self, queries, kind=CPU_POTENTIAL, samples=1, seed=0, eps=0.0, threads=0):
        q = _c(queries, np.float32)
        out = np.zeros((q.shape[0], 3), np.float64)
        lib().orc_cpu_mlpcv(self._h, q, q.shape[0], int(kind), int(samples), int(seed),
                            float(eps), int(threads), out)
        return out if kind == CPU_SOFT_FORCE else out[:, 0].copy()


def bruteforce_smoothdist(points, alphas, queries):
    """Exact fp64 sum of exp(-alpha_j |p_j - q|) (smoothDistExp)."""
    p = _c(points, np.float32)
    a = _c(alphas, np.float32)
    q = _c(queries, np.float32)
    out = np.zeros(q.shape[0], np.float64)
    lib().orc_bruteforce_smoothdist(p, a, p.shape[0], q, q.shape[0], out)
    return out


def bruteforce_wn(points, normals, areas, queries):
    """Exact fp64 winding-number sum of dipoles (signedSolidAngle<3>)."""
    p = _c(points, np.float32)
    nv = _c(normals, np.float32)
    a = _c(areas, np.float32)
    q = _c(queries, np.float32)
    out = np.zeros(q.shape[0], np.float64)
    lib().orc_bruteforce_wn(p, nv, a, p.shape[0], q, q.shape[0], out)
    return out


def triangle_wn(verts, faces, queries, chunk=256):
    """Exact winding number of a closed triangle mesh at each query: the sum
    of triangle solid angles (Van Oosterom & Strackee, IEEE TBME 1983:
    tan(omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|))
    over 4 pi, in fp64."""
    V = np.asarray(verts, np.float64)
    F = np.asarray(faces, np.int64)
    Q = np.asarray(queries, np.float64)
    A0, B0, C0 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    out = np.zeros(Q.shape[0], np.float64)
    for s0 in range(0, Q.shape[0], chunk):
        q = Q[s0:s0 + chunk, None, :]
        a, b, c = A0[None] - q, B0[None] - q, C0[None] - q
        la, lb, lc = (np.linalg.norm(x, axis=2) for x in (a, b, c))
        det = np.einsum("qfk,qfk->qf", a, np.cross(b, c))
        den = (la * lb * lc + np.einsum("qfk,qfk->qf", a, b) * lc +
               np.einsum("qfk,qfk->qf", a, c) * lb + np.einsum("qfk,qfk->qf", b, c) * la)
        out[s0:s0 + chunk] = (2.0 * np.arctan2(det, den)).sum(axis=1) / (4.0 * np.pi)
    return out


# ---------------------------------------------------------------- adaptive tree
GROUP_CELL = 2048  # query groups are runs of <= 64 points inside cells of <= 2048 points


class Tree:
    """Adaptive Morton octree built by the oracle (topology in numpy arrays)."""

    def __init__(self, points, masses, max_depth=10, leaf_size=32, min_corner=None,
                 max_corner=None):
        p = _c(points, np.float32)
        self.n = p.shape[0]
        if min_corner is None or max_corner is None:
            lo, hi = default_bounds(p)
            min_corner = lo if min_corner is None else min_corner
            max_corner = hi if max_corner is None else max_corner
        self.min_corner = _c(min_corner, np.float32)
        self.max_corner = _c(max_corner, np.float32)
        self.span = span(self.min_corner, self.max_corner)
        self.max_depth = int(max_depth)
        self.leaf_size = int(leaf_size)
        codes = morton_codes(p, self.min_corner, self.span, self.max_depth)
        self.perm = stable_order(codes, 3 * self.max_depth)
        self.codes = codes[self.perm]
        self.points_sorted = p[self.perm]
        self.masses_sorted = _c(masses, np.float32)[self.perm]
        cap = 1 + 8 * self.max_depth * (self.n // (self.leaf_size + 1) + 1)
        cap = max(cap, 16)
        arrs = [np.zeros(cap, np.int32) for _ in range(6)]
        geo = np.zeros(cap * 4, np.float32)
        nn = lib().orc_tree_build(self.codes, self.n, self.max_depth, self.leaf_size,
                                  self.min_corner, float(self.span), cap, *arrs, geo)
        if nn < 0:
            raise RuntimeError("node capacity exceeded")
        self.num_nodes = int(nn)
        (self.start, self.count, self.first_child, self.nchild, self.parent,
         self.level) = [a[:nn].copy() for a in arrs]
        self.geo = geo[: 4 * nn].reshape(nn, 4).copy()

    def moments(self, xs=None, ms=None):
        xs = self.points_sorted.astype(np.float64) if xs is None else _c(xs, np.float64)
        ms = self.masses_sorted.astype(np.float64) if ms is None else _c(ms, np.float64)
        nn = self.num_nodes
        mass = np.zeros(nn, np.float64)
        com = np.zeros((nn, 3), np.float64)
        quad = np.zeros((nn, 6), np.float64)
        lib().orc_tree_moments(self.start, self.count, nn, _c(xs, np.float64),
                               _c(ms, np.float64), mass, com, quad)
        return mass, com, quad

    def groups(self, gmax=64, cell_max=GROUP_CELL):
        """Query groups of the walk as (start, count) arrays (orc_tree_groups)."""
        gs = np.zeros(max(self.n, 1), np.int32)
        gc = np.zeros(max(self.n, 1), np.int32)
        ng = lib().orc_tree_groups(self.num_nodes, self.start, self.count, self.first_child,
                                   self.nchild, self.parent, int(gmax), int(cell_max), gs, gc)
        return gs[:ng].copy(), gc[:ng].copy()

    def group_boxes(self, q_sorted_f32, gmax=64):
        q = _c(q_sorted_f32, np.float32)
        gs, gc = self.groups(gmax)
        boxes = np.zeros(max(len(gs), 1) * 6, np.float32)
        lib().orc_group_boxes(q, len(gs), gs, gc, boxes)
        return boxes

    def eval(self, *, beta=2.0, eps=0.0, use_quad=True, group_size=64, cot_sorted=None,
             xs=None, ms=None, boxes=None):
        """Self-gravity walk over the sorted sources (queries == sources).

        use_quad: False / 0 monopole, True / 1 monopole + quadrupole, 2 monopole
        terms with the quadrupole's second-moment adjoint (orc_far).
        group_size is the largest query group (see orc_tree_groups).  xs/ms
        override the source positions/masses (fp64, sorted) while the
        topology and the group boxes stay frozen -- used for finite differences.
        Returns dict(force, gq, gsx, gsm, far, near, groups) in sorted order.
        """
        xs = self.points_sorted.astype(np.float64) if xs is None else _c(xs, np.float64)
        ms = self.masses_sorted.astype(np.float64) if ms is None else _c(ms, np.float64)
        gs, gc = self.groups(group_size)
        if boxes is None:
            boxes = self.group_boxes(self.points_sorted, group_size)
        mass, com, quad = self.moments(xs, ms)
        n = self.n
        force = np.zeros((n, 3), np.float64)
        cot = None if cot_sorted is None else _c(cot_sorted, np.float64)
        gq = np.zeros((n, 3), np.float64) if cot is not None else None
        gsx = np.zeros((n, 3), np.float64) if cot is not None else None
        gsm = np.zeros(n, np.float64) if cot is not None else None
        counts = (_i64 * 2)()
        lib().orc_tree_eval(self.num_nodes, self.start, self.count, self.first_child,
                            self.nchild, _c(self.geo, np.float32).ravel(), mass,
                            com.ravel(), quad.ravel(), xs, ms, n, boxes, xs, n, len(gs),
                            gs, gc, float(beta), int(use_quad), float(eps), 1,
                            _ptr(cot), force, _ptr(gq), _ptr(gsx), _ptr(gsm), counts)
        return dict(force=force, gq=gq, gsx=gsx, gsm=gsm, far=int(counts[0]),
                    near=int(counts[1]), groups=len(gs))

    def eval_groups(self, g_begin, g_end, *, beta=2.0, eps=0.0, use_quad=True, group_size=64,
                    cot_sorted=None):
        """Walk only groups [g_begin, g_end) (the node adjoints are still
        pushed down over the whole tree).  Used to time a bounded sample."""
        xs = self.points_sorted.astype(np.float64)
        ms = self.masses_sorted.astype(np.float64)
        gs, gc = self.groups(group_size)
        boxes = self.group_boxes(self.points_sorted, group_size)
        mass, com, quad = self.moments(xs, ms)
        n = self.n
        force = np.zeros((n, 3), np.float64)
        cot = None if cot_sorted is None else _c(cot_sorted, np.float64)
        gq = np.zeros((n, 3), np.float64) if cot is not None else None
        gsx = np.zeros((n, 3), np.float64) if cot is not None else None
        gsm = np.zeros(n, np.float64) if cot is not None else None
        counts = (_i64 * 2)()
        lib().orc_tree_eval_range(self.num_nodes, self.start, self.count, self.first_child,
                                  self.nchild, _c(self.geo, np.float32).ravel(), mass,
                                  com.ravel(), quad.ravel(), xs, ms, n, boxes, xs, n, gs, gc,
                                  float(beta), int(use_quad), float(eps), 1, _ptr(cot),
                                  force, _ptr(gq), _ptr(gsx), _ptr(gsm), counts,
                                  int(g_begin), int(g_end))
        return dict(force=force, far=int(counts[0]), near=int(counts[1]), groups=len(gs))


    def wn_moments(self, normals_sorted):
        nn = self.num_nodes
        area = np.zeros(nn, np.float64)
        cen = np.zeros((nn, 3), np.float64)
        dip = np.zeros((nn, 3), np.float64)
        lib().orc_wn_moments(self.start, self.count, nn,
                             self.points_sorted.astype(np.float64),
                             _c(normals_sorted, np.float64),
                             self.masses_sorted.astype(np.float64), area, cen, dip)
        return area, cen, dip

    def wn_eval(self, normals, queries, *, beta=2.0, group_size=64):
        """Winding numbers at `queries` (input order) with the group walk: the
        tree's points are dipole positions, its masses the areas, `normals`
        (input order) the unit normals.  The queries get a tree of their own
        (depth 10, leaf size group_size, own cube) whose cell groups
        (orc_tree_groups) are walked against this tree."""
        ns = _c(normals, np.float64)[self.perm]
        area, cen, dip = self.wn_moments(ns)
        q = _c(queries, np.float32)
        nq = q.shape[0]
        out = np.zeros(nq, np.float64)
        if nq == 0:
            return dict(wn=out, far=0, near=0, groups=0)
        QT = Tree(q, np.zeros(nq, np.float32), max_depth=10, leaf_size=group_size)
        gs, gc = QT.groups(group_size)
        boxes = QT.group_boxes(QT.points_sorted, group_size)
        ng = len(gs)
        counts = (_i64 * 2)()
        lib().orc_wn_eval(self.num_nodes, self.start, self.count, self.first_child, self.nchild,
                          _c(self.geo, np.float32).ravel(), area, cen.ravel(), dip.ravel(),
                          self.points_sorted.astype(np.float64), ns,
                          self.masses_sorted.astype(np.float64), boxes,
                          QT.points_sorted.astype(np.float64), nq, ng, gs, gc, float(beta), out,
                          counts)
        res = np.zeros(nq, np.float64)
        res[QT.perm] = out
        return dict(wn=res, far=int(counts[0]), near=int(counts[1]), groups=int(ng))


    # stochastic walk kinds (orc_tree_stoch)
    STOCH_FORCE, STOCH_VJP, STOCH_WN = 0, 1, 2

    def subdomains(self, masses_or_areas=None):
        """Control subdomains of the stochastic walk: nodes at level 3 and
        leaves above it with non-zero mass, in node order."""
        mass = self.moments()[0] if masses_or_areas is None else masses_or_areas
        subs = np.zeros(max(self.num_nodes, 1), np.int32)
        ns = lib().orc_tree_subdomains(self.num_nodes, self.nchild, self.level,
                                       _c(mass, np.float64), subs)
        return subs[:ns].copy()

    def query_groups(self, queries, group_size=64):
        """The query tree of a stochastic or winding-number walk: (tree,
        group starts, group counts) over the queries' own Morton order."""
        q = _c(queries, np.float32)
        QT = Tree(q, np.zeros(q.shape[0], np.float32), max_depth=10, leaf_size=group_size)
        gs, gc = QT.groups(group_size)
        return QT, gs, gc

    def stoch_eval_sorted(self, qs, gs, gc, *, kind=0, beta=2.0, eps=0.0, samples=1, seed=0,
                          use_quad=False, cot_sorted=None, normals=None):
        """orc_tree_stoch over queries already in group order (qs float32,
        sorted; groups gs / gc; cotangents in the same order).  Returns
        dict(out (sorted order), node, pair)."""
        qs = _c(qs, np.float32)
        nq = qs.shape[0]
        nn = self.num_nodes
        if kind == self.STOCH_WN:
            ns = _c(normals, np.float64)[self.perm]
            mass, com, dip = self.wn_moments(ns)
            quad = np.zeros(6 * max(nn, 1), np.float64)
        else:
            mass, com, quad = self.moments()
            dip = np.zeros(3, np.float64)
            ns = np.zeros(3, np.float64)
        width = 1 if kind == self.STOCH_WN else 3
        out = np.zeros(max(nq, 1) * width, np.float64)
        cs = None if cot_sorted is None else _c(cot_sorted, np.float64)
        counts = (_i64 * 2)()
        lib().orc_tree_stoch(nn, self.start, self.count, self.first_child, self.nchild,
                             self.level, _c(self.geo, np.float32).ravel(), _c(mass, np.float64),
                             _c(com, np.float64).ravel(), _c(quad, np.float64).ravel(),
                             _c(dip, np.float64).ravel(), self.points_sorted.astype(np.float64),
                             self.masses_sorted.astype(np.float64), _c(ns, np.float64).ravel(),
                             qs.reshape(-1) if nq else np.zeros(3, np.float32), nq, len(gs),
                             _c(gs, np.int32), _c(gc, np.int32), int(kind), int(bool(use_quad)),
                             float(beta), float(eps), int(samples), int(seed) & 0xFFFFFFFF,
                             _ptr(cs), out, counts)
        out = out[:nq * width].reshape(nq, width) if width == 3 else out[:nq]
        return dict(out=out, node=int(counts[0]), pair=int(counts[1]))

    def stoch_eval(self, queries, *, kind=0, beta=2.0, eps=0.0, samples=1, seed=0,
                   use_quad=False, cot=None, normals=None, group_size=64):
        """Stochastic BH estimate at `queries` (input order): forces (kind 0),
        their query-side VJP with `cot` (kind 1) or winding numbers (kind 2,
        this tree's points are element centroids, its masses the areas,
        `normals` the unit normals in input order).  The queries get a tree
        of their own (depth 10, leaf group_size) whose groups share draws."""
        QT, gs, gc = self.query_groups(queries, group_size)
        cs = None if cot is None else _c(cot, np.float64)[QT.perm]
        r = self.stoch_eval_sorted(QT.points_sorted, gs, gc, kind=kind, beta=beta, eps=eps,
                                   samples=samples, seed=seed, use_quad=use_quad,
                                   cot_sorted=cs, normals=normals)
        res = np.zeros_like(r["out"])
        res[QT.perm] = r["out"]
        return dict(out=res, node=r["node"], pair=r["pair"], groups=len(gs))


# ---------------------------------------------------------------- Philox
def philox4x32(ctr, key):
    c = _c(ctr, np.uint32)
    k = _c(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(c, k, out)
    return out


def rel_l2(pred, ref):
    pred = np.asarray(pred, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm((pred - ref).ravel()) / max(np.linalg.norm(ref.ravel()), 1e-12))


def outlier_frac(pred, ref, thr=0.05):
    pred = np.asarray(pred, np.float64)
    ref = np.asarray(ref, np.float64)
    e = np.linalg.norm(pred - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-12)
    return float(np.mean(e > thr))
