"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, reports the reference's layout sizes, and rejects bad
arguments on the host (the reference's InvalidArgument branches) before any
HIP call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sbh_amd.h")


def _lib():
    from sbh_amd import _lib as L

    return L.load_library()


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sbh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib()
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name


def test_layout_sizes_match_reference():
    lib = _lib()
    # src/sbh_ffi.cu:3949-3963
    assert lib.sbh_leaf3_size() == 32
    assert lib.sbh_node3_size() == 4272
    assert lib.sbh_octree_level3_max_nodes() == 513
    assert lib.sbh_octree8_node_size() == 104
    # LeafData<2>, Tree2DNode<1,3> (include/leaf_data.h:8-16, include/tree2d.h:212-225)
    assert lib.sbh_leaf2_size() == 24
    assert lib.sbh_node2_size() == 616


def _buf(dtype, *dims, ptr=None):
    from sbh_amd._lib import SbhBuffer

    b = SbhBuffer()
    b.data = ptr
    b.dtype = dtype
    b.rank = len(dims)
    for i, d in enumerate(dims):
        b.dims[i] = d
    return b


F32, U8 = 11, 6
FAKE = 0x1000  # never dereferenced: validation fails first


def _call(name, *args):
    lib = _lib()
    fn = getattr(lib, name)
    fn.restype = ctypes.c_int
    rc = fn(*args)
    return rc, lib.sbh_last_error().decode()


def _f(x):
    return ctypes.c_float(x)


def _i(x):
    return ctypes.c_int32(x)


@pytest.mark.parametrize(
    "pts,ms,qs,out,soft,cut,threads,msg",
    [
        ((F32, 5), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 256, "must be rank-2"),
        ((F32, 5, 2), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 256, "must have dim=3"),
        ((F32, 5, 3), (F32, 6), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 256, "masses length mismatch"),
        ((F32, 5, 3), (F32, 5), (F32, 4, 3), (F32, 3, 3), 0.01, 1.0, 256, "out shape mismatch"),
        ((F32, 5, 3), (F32, 5), (F32, 4, 3), (F32, 4, 3), -1.0, 1.0, 256, "softening_scale must be >= 0"),
        ((F32, 5, 3), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 0.0, 256, "cutoff_scale must be > 0"),
        ((F32, 5, 3), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 0, "threads must be in [1, 1024]"),
        ((F32, 5, 3), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 2048, "threads must be in [1, 1024]"),
        ((U8, 5, 3), (F32, 5), (F32, 4, 3), (F32, 4, 3), 0.01, 1.0, 256, "float32"),
    ],
)
def test_direct_force_rejects(pts, ms, qs, out, soft, cut, threads, msg):
    bp = _buf(*pts, ptr=FAKE)
    bm = _buf(*ms, ptr=FAKE)
    bq = _buf(*qs, ptr=FAKE)
    bo = _buf(*out, ptr=FAKE)
    rc, err = _call("sbh_direct_force_soft", None, ctypes.byref(bp), ctypes.byref(bm),
                    ctypes.byref(bq), ctypes.byref(bo), _f(soft), _f(cut), _i(0), _f(1.0),
                    _i(threads))
    assert rc == 3 and msg in err


def test_direct_vjp_rejects_cotangent_mismatch():
    bp, bm = _buf(F32, 5, 3, ptr=FAKE), _buf(F32, 5, ptr=FAKE)
    bq, bc, bo = _buf(F32, 4, 3, ptr=FAKE), _buf(F32, 3, 3, ptr=FAKE), _buf(F32, 4, 3, ptr=FAKE)
    rc, err = _call("sbh_direct_force_soft_vjp", None, *map(ctypes.byref, (bp, bm, bq, bc, bo)),
                    _f(0.01), _f(1.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "cotangent shape mismatch" in err


def test_bruteforce_rejects_dim():
    bp, bm = _buf(F32, 5, 4, ptr=FAKE), _buf(F32, 5, ptr=FAKE)
    bq, bo = _buf(F32, 4, 4, ptr=FAKE), _buf(F32, 4, ptr=FAKE)
    rc, err = _call("sbh_bruteforce_gravity", None, *map(ctypes.byref, (bp, bm, bq, bo)), _i(0),
                    _f(1.0), _i(64))
    assert rc == 3 and "only dim=2 or dim=3 supported" in err
    bq = _buf(F32, 4, 2, ptr=FAKE)
    rc, err = _call("sbh_bruteforce_gravity", None, *map(ctypes.byref, (bp, bm, bq, bo)), _i(0),
                    _f(1.0), _i(64))
    assert rc == 3 and "queries dim mismatch" in err


@pytest.mark.parametrize("depth,leaf_n,node_n,msg", [
    (0, 10, 1, "max_depth must be in [1, 10]"),
    (11, 10, 1, "max_depth must be in [1, 10]"),
    (2, 9, 73, "leaf_out size mismatch"),
    (2, 10, 72, "node_out size mismatch"),
])
def test_octree8_builder_rejects(depth, leaf_n, node_n, msg):
    bp, bn, bm = _buf(F32, 10, 3, ptr=FAKE), _buf(F32, 10, 3, ptr=FAKE), _buf(F32, 10, ptr=FAKE)
    lo, hi = _buf(F32, 3, ptr=FAKE), _buf(F32, 3, ptr=FAKE)
    leaf = _buf(U8, leaf_n * 32, ptr=FAKE)
    node = _buf(U8, node_n * 104, ptr=FAKE)
    contrib = _buf(U8, node_n * 32, ptr=FAKE)
    rc, err = _call("sbh_build_octree8_buffers", None,
                    *map(ctypes.byref, (bp, bn, bm, lo, hi, leaf, node, contrib)), _i(depth))
    assert rc == 3 and msg in err


def test_octree8_builder_rejects_shapes():
    bp, bn, bm = _buf(F32, 10, 3, ptr=FAKE), _buf(F32, 9, 3, ptr=FAKE), _buf(F32, 10, ptr=FAKE)
    lo, hi = _buf(F32, 3, ptr=FAKE), _buf(F32, 2, ptr=FAKE)
    leaf, node, contrib = _buf(U8, 320, ptr=FAKE), _buf(U8, 9 * 104, ptr=FAKE), _buf(U8, 9 * 32, ptr=FAKE)
    rc, err = _call("sbh_build_octree8_buffers", None,
                    *map(ctypes.byref, (bp, bn, bm, lo, hi, leaf, node, contrib)), _i(1))
    assert rc == 3 and "points/normals row mismatch" in err


def test_walk_rejects():
    leaf, node = _buf(U8, 33, ptr=FAKE), _buf(U8, 104, ptr=FAKE)
    q, o = _buf(F32, 4, 3, ptr=FAKE), _buf(F32, 4, 3, ptr=FAKE)
    rc, err = _call("sbh_barnes_hut_force_octree8", None, *map(ctypes.byref, (leaf, node, q, o)),
                    _f(2.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "leaf_bytes size mismatch" in err
    leaf = _buf(U8, 64, ptr=FAKE)
    q2 = _buf(F32, 4, 2, ptr=FAKE)
    rc, err = _call("sbh_barnes_hut_force_octree8", None, *map(ctypes.byref, (leaf, node, q2, o)),
                    _f(2.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "queries dim must be 3" in err
    rc, err = _call("sbh_barnes_hut_force_octree8_soft", None,
                    *map(ctypes.byref, (leaf, node, q, o)), _f(2.0), _f(0.01), _f(1.0), _i(1),
                    _f(0.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "prune_r_cut_mult must be > 0" in err


def test_python_mirror_validates_like_reference():
    import numpy as np

    from sbh_amd import ffi

    p = np.zeros((4, 3), np.float32)
    with pytest.raises((ValueError, RuntimeError)):
        ffi.softened_direct_force(p, np.ones(3, np.float32), p)


@pytest.mark.parametrize("dim,pts,nrm,ms,lo,msg", [
    (4, (F32, 5, 4), (F32, 5, 4), (F32, 5), (F32, 4), "dim must be 2 or 3"),
    (2, (F32, 5, 3), (F32, 5, 3), (F32, 5), (F32, 2), "points must have shape (N, 2)"),
    (3, (F32, 5, 3), (F32, 4, 3), (F32, 5), (F32, 3), "normals shape mismatch"),
    (2, (F32, 5, 2), (F32, 5, 2), (F32, 6), (F32, 2), "masses must have shape (N,)"),
    (2, (F32, 5, 2), (F32, 5, 2), (F32, 5), (F32, 3), "min_corner/max_corner"),
])
def test_tree_builder_rejects(dim, pts, nrm, ms, lo, msg):
    bufs = [_buf(*x, ptr=FAKE) for x in (pts, nrm, ms, lo, lo)]
    nodes = ctypes.c_int64(-1)
    rc, err = _call("sbh_tree_build", None, _i(dim), _i(3), *map(ctypes.byref, bufs), _f(1.0),
                    ctypes.byref(nodes))
    assert rc == 3 and msg in err and nodes.value == -1
    if dim == 2 and msg.startswith("masses"):
        bufs = [_buf(*x, ptr=FAKE) for x in ((F32, 5, 2), (F32, 5, 2), (F32, 5), lo, lo)]
        rc, err = _call("sbh_tree_build", None, _i(2), _i(4), *map(ctypes.byref, bufs),
                        _f(1.0), ctypes.byref(nodes))
        assert rc == 3 and "level must be 1, 2 or 3" in err


def test_quadtree_walks_reject_wrong_widths():
    # 2-D targets: LeafData<2> / Tree2DNode<1,3> byte multiples, (M, 2) queries
    leaf, node = _buf(U8, 24 * 5, ptr=FAKE), _buf(U8, 616 * 3, ptr=FAKE)
    q3, out = _buf(F32, 4, 3, ptr=FAKE), _buf(F32, 4, 2, ptr=FAKE)
    rc, err = _call("sbh_barnes_hut_force", None, *map(ctypes.byref, (leaf, node, q3, out)),
                    _f(2.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "queries dim must be 2" in err
    bad_node = _buf(U8, 616 * 3 + 8, ptr=FAKE)
    q2 = _buf(F32, 4, 2, ptr=FAKE)
    rc, err = _call("sbh_barnes_hut_force", None, *map(ctypes.byref, (leaf, bad_node, q2, out)),
                    _f(2.0), _i(0), _f(1.0), _i(256))
    assert rc == 3 and "node_bytes size mismatch" in err
    bad_leaf = _buf(U8, 32 * 5, ptr=FAKE)  # 160 is not a multiple of 24
    rc, err = _call("sbh_barnes_hut_gravity", None,
                    *map(ctypes.byref, (bad_leaf, node, q2, _buf(F32, 4, ptr=FAKE))),
                    _f(2.0), _i(0), _f(1.0), _i(64))
    assert rc == 3 and "leaf_bytes size mismatch" in err


def test_tree_module_entries_reject():
    # sbh_tree_eval / sbh_tree_bruteforce: the pybind11 module's evaluations
    leaf, node = _buf(U8, 24 * 5, ptr=FAKE), _buf(U8, 136 * 3, ptr=FAKE)
    q2, out = _buf(F32, 4, 2, ptr=FAKE), _buf(F32, 4, ptr=FAKE)
    args = (leaf, node, node, q2, out)
    rc, err = _call("sbh_tree_eval", None, _i(4), _i(1), _i(0), *map(ctypes.byref, args),
                    _f(2.0), _i(-1))
    assert rc == 3 and "dim must be 2 or 3" in err
    rc, err = _call("sbh_tree_eval", None, _i(2), _i(0), _i(0), *map(ctypes.byref, args),
                    _f(2.0), _i(-1))
    assert rc == 3 and "level must be 1, 2 or 3" in err
    rc, err = _call("sbh_tree_eval", None, _i(2), _i(1), _i(5), *map(ctypes.byref, args),
                    _f(2.0), _i(-1))
    assert rc == 3 and "kind must be" in err
    # level 2 records are 232 bytes: 3 * 136 is not a multiple
    rc, err = _call("sbh_tree_eval", None, _i(2), _i(2), _i(0), *map(ctypes.byref, args),
                    _f(2.0), _i(-1))
    assert rc == 3 and "node_bytes size mismatch" in err
    rc, err = _call("sbh_tree_eval", None, _i(2), _i(1), _i(0), ctypes.byref(leaf),
                    ctypes.byref(node), None, ctypes.byref(q2), ctypes.byref(out), _f(2.0), _i(1))
    assert rc == 3 and "contrib_bytes must be given" in err
    rc, err = _call("sbh_tree_bruteforce", None, _i(3), _i(0), ctypes.byref(leaf),
                    ctypes.byref(q2), ctypes.byref(out))
    assert rc == 3 and "leaf_bytes size mismatch" in err  # 120 bytes are not LeafData<3>s
    rc, err = _call("sbh_tree_bruteforce", None, _i(2), _i(0), ctypes.byref(leaf),
                    ctypes.byref(_buf(F32, 4, 3, ptr=FAKE)), ctypes.byref(out))
    assert rc == 3 and "queries dim must be 2" in err


def test_stochastic_adaptive_entries_reject():
    # argument checks of the sbh_stochastic_* targets run before any HIP call
    p, m = _buf(F32, 5, 3, ptr=FAKE), _buf(F32, 5, ptr=FAKE)
    q, o = _buf(F32, 4, 3, ptr=FAKE), _buf(F32, 4, 3, ptr=FAKE)
    args = (_f(1.0), _f(0.01), _i(10), _i(32), _i(1))
    rc, err = _call("sbh_stochastic_gravity_force", None, *map(ctypes.byref, (p, m, q, o)),
                    *args, _i(-1), _i(0), None, None)
    assert rc == 3 and "samples_per_subdomain must be >= 0" in err
    rc, err = _call("sbh_stochastic_gravity_force", None, *map(ctypes.byref, (p, m, q, o)),
                    _f(1.0), _f(0.01), _i(11), _i(32), _i(1), _i(1), _i(0), None, None)
    assert rc == 3 and "max_depth must be in [1, 10]" in err
    rc, err = _call("sbh_stochastic_gravity_force", None, *map(ctypes.byref, (p, m, q, o)),
                    _f(1.0), _f(0.01), _i(10), _i(32), _i(3), _i(1), _i(0), None, None)
    assert rc == 3 and "multipole_order must be 1 or 2" in err
    o2 = _buf(F32, 5, 3, ptr=FAKE)
    rc, err = _call("sbh_stochastic_gravity_force", None, *map(ctypes.byref, (p, m, q, o2)),
                    *args, _i(1), _i(0), None, None)
    assert rc == 3 and "out must have shape (M, 3)" in err
    c_bad = _buf(F32, 5, 3, ptr=FAKE)
    rc, err = _call("sbh_stochastic_gravity_force_vjp", None,
                    *map(ctypes.byref, (p, m, q, c_bad, o)), *args, _i(1), _i(0), None, None)
    assert rc == 3 and "cotangent must have shape (M, 3)" in err
    n_bad, a = _buf(F32, 5, 2, ptr=FAKE), _buf(F32, 5, ptr=FAKE)
    w = _buf(F32, 4, ptr=FAKE)
    rc, err = _call("sbh_stochastic_winding_number", None,
                    *map(ctypes.byref, (p, n_bad, a, q, w)), _f(2.0), _i(10), _i(16), _i(1),
                    _i(0), None, None)
    assert rc == 3 and "normals must have shape (N, 3)" in err
    lib = _lib()
    lib.sbh_stochastic_workspace_bytes.restype = ctypes.c_int64
    assert lib.sbh_stochastic_workspace_bytes(ctypes.c_int64(-1), ctypes.c_int64(4), _i(10),
                                              _i(32)) == -1
    assert lib.sbh_stochastic_workspace_bytes(ctypes.c_int64(1000), ctypes.c_int64(10), _i(10),
                                              _i(32)) > 0


def test_gravity_expansion_orders_reject():
    # forces take orders 1 and 2; the VJP entry points also take 3 (the
    # second-moment adjoint); checks run before any HIP call
    p, m, f = _buf(F32, 5, 3, ptr=FAKE), _buf(F32, 5, ptr=FAKE), _buf(F32, 5, 3, ptr=FAKE)
    c, gx, gm = _buf(F32, 5, 3, ptr=FAKE), _buf(F32, 5, 3, ptr=FAKE), _buf(F32, 5, ptr=FAKE)
    rc, err = _call("sbh_gravity_force", None, *map(ctypes.byref, (p, m, f)), _f(2.0),
                    _f(0.01), _i(10), _i(32), _i(3), None, None)
    assert rc == 3 and "multipole_order must be 1 or 2" in err
    for name, extra in (("sbh_gravity_force_vjp", (None,)), ("sbh_gravity_force_vjp_noalloc", ())):
        rc, err = _call(name, None, *map(ctypes.byref, (p, m, c)), None,
                        *map(ctypes.byref, (gx, gm)), _f(2.0), _f(0.01), _i(10), _i(32), _i(4),
                        *extra, None)
        assert rc == 3 and "multipole_order must be 1, 2 or 3" in err, (name, err)
    bad_gm = _buf(F32, 4, ptr=FAKE)
    rc, err = _call("sbh_gravity_force_vjp", None, *map(ctypes.byref, (p, m, c)), None,
                    *map(ctypes.byref, (gx, bad_gm)), _f(2.0), _f(0.01), _i(10), _i(32), _i(3),
                    None, None)
    assert rc == 3 and "grad_masses shape mismatch" in err
