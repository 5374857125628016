"""GPU parity: dense octree8 builder and classic walks (drop-in mode).

Builder: leaf records, node topology and node geometry must equal the
oracle's restatement bit for bit; moments (fp32 sums vs fp64) within a stated
tolerance.  Walks: identical MAC decisions, so forces differ from the oracle
only by fp32 accumulation.  The reference's own tests
(python/test_jax_ffi_tree_builder.py:206-412,
python/test_jax_custom_vjp_recompute.py:105-151) are replicated at the end.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOPO_FIELDS = ["start", "end", "parent", "first_child", "child", "child_mask", "depth",
               "is_leaf", "pad0"]
GEO_FIELDS = ["cx", "cy", "cz", "half_size", "diameter"]


def _ffi():
    from sbh_amd import ffi

    return ffi


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _points(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    if kind == "clump":
        p = rng.uniform(-1, 1, (n, 3))
        p[: n // 2] = 0.3 + rng.normal(0, 1e-3, (n // 2, 3))
        return p.astype(np.float32)
    if kind == "plummer":
        xmax = 20.0 ** 3 / (20.0 ** 2 + 1.0) ** 1.5
        X = rng.uniform(0, xmax, n)
        r = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return (r[:, None] * v).astype(np.float32)
    raise ValueError(kind)


@pytest.mark.parametrize("kind,n,depth", [
    ("uniform", 5000, 5), ("clump", 3000, 6), ("plummer", 20000, 7), ("uniform", 1, 3),
    ("uniform", 0, 2), ("uniform", 777, 1),
])
def test_builder_matches_oracle(gpu, kind, n, depth):
    p = _points(kind, n, 100 + n)
    rng = np.random.default_rng(n)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    m = rng.uniform(0.1, 1.0, n).astype(np.float32)
    if n > 0:
        lo, hi = O.default_bounds(p)
    else:
        lo, hi = -np.ones(3, np.float32), np.ones(3, np.float32)
    leaf, node, contrib = _ffi().build_octree8_buffers_with_contrib(p, nrm, m, lo, hi,
                                                                    max_depth=depth)
    leaf, node, contrib = _np(leaf), _np(node), _np(contrib)
    rleaf, rnode, rcontrib = O.build_octree8(p, nrm, m, lo, hi, depth)
    # leaf records: the stable (code, index) order, byte for byte
    np.testing.assert_array_equal(leaf, rleaf)
    g, r = O.decode_nodes8(node), O.decode_nodes8(rnode)
    for f in TOPO_FIELDS + GEO_FIELDS:
        np.testing.assert_array_equal(g[f], r[f], err_msg=f)
    # moments: fp32 sums in sorted order vs fp64
    gm, rm = g["contrib"], r["contrib"]
    np.testing.assert_array_equal(gm["alpha"], 1.0)
    np.testing.assert_allclose(gm["mass"], rm["mass"], rtol=2e-5, atol=0)
    big = rm["mass"] > 0
    for f in ("pos", "normal"):
        scale = np.abs(rm[f][big]).max() if big.any() else 1.0
        np.testing.assert_allclose(gm[f][big], rm[f][big], rtol=0, atol=2e-5 * max(scale, 1.0))
        assert np.all(gm[f][~big] == 0)
    np.testing.assert_array_equal(contrib.reshape(-1, 32), node.reshape(-1, 104)[:, 72:])


def test_builder_default_bounds_and_determinism(gpu):
    ffi = _ffi()
    p = _points("plummer", 30000, 5)
    m = np.full(30000, 1 / 30000, np.float32)
    z = np.zeros_like(p)
    a = [_np(x) for x in ffi.build_octree8_buffers(p, z, m, max_depth=6)]
    b = [_np(x) for x in ffi.build_octree8_buffers(p, z, m, max_depth=6)]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    lo, hi = O.default_bounds(p)
    rleaf, rnode, _ = O.build_octree8(p, z, m, lo, hi, 6)
    np.testing.assert_array_equal(a[0], rleaf)
    g, r = O.decode_nodes8(a[1]), O.decode_nodes8(rnode)
    for f in TOPO_FIELDS + GEO_FIELDS:
        np.testing.assert_array_equal(g[f], r[f], err_msg=f)


def _tree(n=4000, depth=5, seed=3, kind="uniform"):
    p = _points(kind, n, seed)
    m = np.random.default_rng(seed).uniform(0.2, 1.0, n).astype(np.float32)
    lo, hi = O.default_bounds(p)
    leaf, node, _ = O.build_octree8(p, np.zeros_like(p), m, lo, hi, depth)
    return p, m, leaf, node


@pytest.mark.parametrize("kind", ["uniform", "plummer"])
@pytest.mark.parametrize("beta", [0.8, 2.0])
def test_walks_match_oracle(gpu, kind, beta):
    ffi = _ffi()
    p, m, leaf, node = _tree(kind=kind)
    q = p[::3].copy()
    c = np.random.default_rng(1).normal(size=q.shape).astype(np.float32)
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    out = _np(ffi.barnes_hut_force_octree8(lt, nt, q, beta=beta))
    ref, _ = O.bh_octree8(leaf, node, q, O.KIND_GRAV, beta)
    assert O.rel_l2(out, ref) < 1e-5
    np.testing.assert_allclose(out, ref, rtol=1e-3, atol=1e-5 * np.abs(ref).max())
    for rc, prune in ((np.inf, False), (0.5, True), (0.5, False)):
        kw = dict(beta=beta, softening_scale=0.02, cutoff_scale=float(rc), prune_enabled=prune,
                  prune_r_cut_mult=2.0)
        out = _np(ffi.softened_barnes_hut_force_octree8(lt, nt, q, **kw))
        ref, _ = O.bh_octree8(leaf, node, q, O.KIND_SOFT, beta, eps=0.02, rc=rc,
                              prune_enabled=prune, prune_r_cut_mult=2.0)
        assert O.rel_l2(out, ref) < 1e-5, (rc, prune)
        out = _np(ffi.softened_barnes_hut_force_octree8_vjp(lt, nt, q, c, **kw))
        ref, _ = O.bh_octree8(leaf, node, q, O.KIND_SOFT_VJP, beta, eps=0.02, rc=rc, cot=c,
                              prune_enabled=prune, prune_r_cut_mult=2.0)
        assert O.rel_l2(out, ref) < 1e-5, (rc, prune)


def test_walk_periodic_matches_oracle(gpu):
    ffi = _ffi()
    rng = np.random.default_rng(13)
    L = 3.0
    p = rng.uniform(0, L, (2000, 3)).astype(np.float32)
    m = np.ones(2000, np.float32)
    lo, hi = np.zeros(3, np.float32), np.full(3, L, np.float32)
    leaf, node, _ = O.build_octree8(p, np.zeros_like(p), m, lo, hi, 5)
    q = p[:300]
    kw = dict(beta=1.9, softening_scale=4e-2, cutoff_scale=0.7, prune_enabled=True,
              prune_r_cut_mult=3.0, periodic=True, box_length=L)
    out = _np(ffi.softened_barnes_hut_force_octree8(leaf, node, q, **kw))
    ref, _ = O.bh_octree8(leaf, node, q, O.KIND_SOFT, 1.9, eps=4e-2, rc=0.7, prune_enabled=True,
                          prune_r_cut_mult=3.0, periodic=True, box=L)
    assert O.rel_l2(out, ref) < 1e-5


def test_walk_huge_beta_equals_direct_sum(gpu):
    ffi = _ffi()
    p, m, leaf, node = _tree(n=3000, depth=4, seed=9)
    q = p[:500]
    out = _np(ffi.softened_barnes_hut_force_octree8(leaf, node, q, beta=1e30,
                                                    softening_scale=0.03,
                                                    cutoff_scale=float("inf")))
    ref = O.direct_force_soft(p, m, q, 0.03, np.inf)
    assert O.rel_l2(out, ref) < 1e-5


# ---------------------------------------------------------- reference test replicas
def test_reference_octree8_chain(gpu):
    # python/test_jax_ffi_tree_builder.py:206-261 (shapes, dtypes, finiteness)
    ffi = _ffi()
    p = np.random.default_rng(1).uniform(-1, 1, (256, 3)).astype(np.float32)
    leaf, node, contrib = ffi.build_octree8_buffers_with_contrib(
        p, np.zeros_like(p), np.ones(256, np.float32), max_depth=5)
    assert leaf.dtype == torch.uint8 and node.dtype == torch.uint8
    assert contrib.dtype == torch.uint8
    f = ffi.barnes_hut_force_octree8(leaf, node, p, beta=2.0)
    fs = ffi.softened_barnes_hut_force_octree8(leaf, node, p, beta=2.0, softening_scale=5e-2,
                                               cutoff_scale=1.5)
    assert f.shape == (256, 3) and fs.shape == (256, 3)
    assert np.isfinite(_np(f)).all() and np.isfinite(_np(fs)).all()


def test_reference_periodic_shift_invariance(gpu):
    # python/test_jax_ffi_tree_builder.py:264-305
    ffi = _ffi()
    L = 2.0
    p = np.random.default_rng(3).uniform(0, L, (192, 3)).astype(np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), np.ones(192, np.float32),
                                           min_corner=np.zeros(3, np.float32),
                                           max_corner=np.full(3, L, np.float32), max_depth=5)
    q = p[:32]
    f0 = ffi.barnes_hut_force_octree8(leaf, node, q, beta=1.8, periodic=True, box_length=L)
    f1 = ffi.barnes_hut_force_octree8(leaf, node, q + np.array([L, 0, 0], np.float32), beta=1.8,
                                      periodic=True, box_length=L)
    np.testing.assert_allclose(_np(f0), _np(f1), rtol=1e-2, atol=5e-3)


def test_reference_prune_huge_cutoff_matches_disabled(gpu):
    # python/test_jax_ffi_tree_builder.py:308-360
    ffi = _ffi()
    L = 2.0
    p = np.random.default_rng(11).uniform(0, L, (320, 3)).astype(np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), np.ones(320, np.float32),
                                           min_corner=np.zeros(3, np.float32),
                                           max_corner=np.full(3, L, np.float32), max_depth=6)
    q = p[:96]
    kw = dict(beta=1.8, softening_scale=5e-2, cutoff_scale=0.8, periodic=True, box_length=L)
    a = ffi.softened_barnes_hut_force_octree8(leaf, node, q, prune_enabled=False, **kw)
    b = ffi.softened_barnes_hut_force_octree8(leaf, node, q, prune_enabled=True,
                                              prune_r_cut_mult=1e6, **kw)
    np.testing.assert_allclose(_np(a), _np(b), rtol=1e-6, atol=1e-6)


def test_reference_prune_shift_invariance(gpu):
    # python/test_jax_ffi_tree_builder.py:363-412
    ffi = _ffi()
    L = 3.0
    p = np.random.default_rng(13).uniform(0, L, (256, 3)).astype(np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), np.ones(256, np.float32),
                                           min_corner=np.zeros(3, np.float32),
                                           max_corner=np.full(3, L, np.float32), max_depth=6)
    q = p[:48]
    kw = dict(beta=1.9, softening_scale=4e-2, cutoff_scale=0.7, prune_enabled=True,
              prune_r_cut_mult=3.0, periodic=True, box_length=L)
    f0 = ffi.softened_barnes_hut_force_octree8(leaf, node, q, **kw)
    f1 = ffi.softened_barnes_hut_force_octree8(leaf, node, q + np.array([L, 0, 0], np.float32),
                                               **kw)
    np.testing.assert_allclose(_np(f0), _np(f1), rtol=1e-2, atol=5e-3)


def test_reference_recompute_custom_vjp_matches_fixed_tree_vjp(gpu):
    # python/test_jax_custom_vjp_recompute.py:105-151
    ffi = _ffi()
    rng = np.random.default_rng(123)
    n = 104
    p = rng.uniform(-0.7, 0.7, (n, 3)).astype(np.float32)
    m = rng.uniform(0.3, 1.2, n).astype(np.float32)
    c = torch.tensor(rng.normal(size=(n, 3)).astype(np.float32), device="cuda")
    kw = dict(beta=1.6, softening_scale=0.05, cutoff_scale=1.4)
    x = torch.tensor(p, device="cuda", requires_grad=True)
    (ffi.softened_barnes_hut_force_octree8_recompute_vjp(x, m, max_depth=6, **kw) * c).sum().backward()
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), m, max_depth=6)
    ref = ffi.softened_barnes_hut_force_octree8_vjp(leaf, node, p, c, **kw)
    np.testing.assert_allclose(_np(x.grad), _np(ref), rtol=1e-5, atol=1e-6)


def test_morton_scheduled_walks_equal_unordered_chunks(gpu):
    # From 4096 queries the walks run in Morton order of the queries; each
    # query's result must not change: compare with chunked calls below the
    # threshold (input order), bit for bit, and with the oracle.
    ffi = _ffi()
    p, m, leaf, node = _tree(n=20000, depth=6, seed=5, kind="plummer")
    rng = np.random.default_rng(6)
    q = np.concatenate([p[rng.permutation(p.shape[0])[:9000]],
                        rng.normal(0, 3, (1000, 3)).astype(np.float32)])
    c = rng.normal(size=q.shape).astype(np.float32)
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    qt, ct = torch.tensor(q, device="cuda"), torch.tensor(c, device="cuda")
    kw = dict(beta=1.5, softening_scale=0.02, cutoff_scale=0.5, prune_enabled=True,
              prune_r_cut_mult=2.0)
    calls = [
        lambda a, b: ffi.barnes_hut_force_octree8(lt, nt, a, beta=1.5),
        lambda a, b: ffi.softened_barnes_hut_force_octree8(lt, nt, a, **kw),
        lambda a, b: ffi.softened_barnes_hut_force_octree8_vjp(lt, nt, a, b, **kw),
    ]
    for f in calls:
        whole = _np(f(qt, ct))
        parts = np.concatenate([_np(f(qt[k:k + 2500].contiguous(), ct[k:k + 2500].contiguous()))
                                for k in range(0, q.shape[0], 2500)])
        np.testing.assert_array_equal(whole, parts)
    out = _np(ffi.softened_barnes_hut_force_octree8(lt, nt, qt, **kw))
    ref, _ = O.bh_octree8(leaf, node, q, O.KIND_SOFT, 1.5, eps=0.02, rc=0.5,
                          prune_enabled=True, prune_r_cut_mult=2.0)
    assert O.rel_l2(out, ref) < 1e-5


def test_walk_captured_in_graph_equals_eager(gpu):
    # under stream capture the walk skips the query ordering (no scratch
    # allocation while capturing); results are the same bit for bit
    from sbh_amd.graph import Graph

    ffi = _ffi()
    p, m, leaf, node = _tree(n=12000, depth=5, seed=8, kind="plummer")
    q = p[np.random.default_rng(9).permutation(p.shape[0])[:6000]].copy()
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    qt = torch.tensor(q, device="cuda")
    kw = dict(beta=1.5, softening_scale=0.02, cutoff_scale=0.5, prune_enabled=True,
              prune_r_cut_mult=2.0)
    eager = _np(ffi.softened_barnes_hut_force_octree8(lt, nt, qt, **kw))
    from sbh_amd._lib import buf, call, f32, i32, stream_handle

    out = torch.empty((q.shape[0], 3), dtype=torch.float32, device="cuda")

    def step():
        # the C-ABI call with a preallocated output (nothing allocates)
        call("sbh_barnes_hut_force_octree8_soft", stream_handle(), buf(lt), buf(nt), buf(qt),
             buf(out), f32(1.5), f32(0.02), f32(0.5), i32(1), f32(2.0), i32(0), f32(1.0),
             i32(256))

    g = Graph()
    g.capture(step)
    for _ in range(2):
        out.zero_()
        g.replay()
        np.testing.assert_array_equal(_np(out), eager)
    g.close()
