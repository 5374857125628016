"""GPU parity: the reference's CPU-layout trees (Tree2D<1,3> quadtrees for the
2-D targets, Tree3D<1,3> octrees of build_octree_buffers(backend="cpu")),
built on the device, and the vote / control-variate walks over them.

Builder: leaf, node and contrib bytes equal the oracle's (oracle/sbh_ndtree.c,
fp32 in the reference's operation order) word for word; the only words
allowed to differ are NaN moments of massless nodes (the NaN's sign bit is
the hardware's).  Walks: the same warp votes and the same sample paths as the
oracle, so results differ only by fp32 against fp64 accumulation.  The
reference's 2-D examples (python/example_jax_bh_compare.py,
python/sbh_jax_examples/stochastic_2d.py) close the file.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ffi():
    from sbh_amd import ffi

    return ffi


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _case(n, dim, seed, kind="uniform"):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (n, dim))
    if kind == "clump":
        p[: n // 2] = 0.3 + rng.normal(0, 2e-3, (n // 2, dim))
    if kind == "dups":
        p[: n // 4] = p[0]
    p = p.astype(np.float32)
    nv = rng.normal(size=(n, dim)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32)
    return p, nv, m


def _corners(p, dim):
    if p.shape[0] == 0:
        return np.zeros(dim, np.float32), np.ones(dim, np.float32)
    return ((p.min(0) - np.float32(1e-3)).astype(np.float32),
            (p.max(0) + np.float32(1e-3)).astype(np.float32))


def _same_words(a, b):
    a = np.ascontiguousarray(a, np.uint8).view(np.uint32)
    b = np.ascontiguousarray(b, np.uint8).view(np.uint32)
    assert a.shape == b.shape
    diff = np.nonzero(a != b)[0]
    if diff.size:
        fa, fb = a[diff].view(np.float32), b[diff].view(np.float32)
        assert (np.isnan(fa) & np.isnan(fb)).all(), (diff[:8], a[diff[:8]], b[diff[:8]])


def _build(ffi, dim, p, nv, m, lo, hi, alpha=1.0):
    if dim == 2:
        return ffi.build_quadtree_buffers_with_contrib(p, nv, m, lo, hi, alpha=alpha)
    return ffi.build_octree_buffers_with_contrib(p, nv, m, lo, hi, backend="cpu")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n,seed,kind", [(3000, 1, "uniform"), (2000, 2, "clump"),
                                         (500, 3, "dups"), (1, 4, "uniform"),
                                         (0, 5, "uniform"), (2, 6, "uniform"),
                                         (40000, 7, "clump")])
def test_tree_builder_bytes_match_oracle(gpu, dim, n, seed, kind):
    ffi = _ffi()
    p, nv, m = _case(n, dim, seed, kind)
    lo, hi = _corners(p, dim)
    alpha = 0.5 if dim == 2 else 1.0
    leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi, alpha)
    ol, on, oc = O.nd_build(p, nv, m, lo, hi, alpha=alpha)
    assert node.numel() == on.size and contrib.numel() == oc.size
    _same_words(_np(leaf), ol)
    _same_words(_np(node), on)
    _same_words(_np(contrib), oc)


def test_tree_builder_alternating_dims_and_growth(gpu):
    ffi = _ffi()
    for dim, n, seed in [(2, 60000, 11), (3, 500, 12), (3, 30000, 13), (2, 700, 14)]:
        p, nv, m = _case(n, dim, seed, "clump")
        lo, hi = _corners(p, dim)
        leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi)
        ol, on, oc = O.nd_build(p, nv, m, lo, hi)
        _same_words(_np(node), on)
        _same_words(_np(leaf), ol)


def test_tree_builder_points_outside_corners(gpu):
    # quantize clamps points outside the box into the edge cells at every level
    ffi = _ffi()
    p, nv, m = _case(2000, 2, 21)
    lo, hi = np.array([-0.5, -0.5], np.float32), np.array([0.5, 0.5], np.float32)
    leaf, node, contrib = _build(ffi, 2, p, nv, m, lo, hi)
    ol, on, oc = O.nd_build(p, nv, m, lo, hi)
    _same_words(_np(node), on)
    _same_words(_np(leaf), ol)


def _walk(ffi, name, leaf, node, contrib, q, **kw):
    fn = getattr(ffi, name)
    if "stochastic" in name:
        return _np(fn(leaf, node, contrib, q, **kw))
    return _np(fn(leaf, node, q, **kw))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("threads", [32, 48, 96, 256])  # 48: a 16-query warp per block
def test_tree_vote_walks_match_oracle(gpu, dim, threads):
    ffi = _ffi()
    p, nv, m = _case(20000, dim, 31, "clump")
    lo, hi = _corners(p, dim)
    leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi)
    lb, nb, cb = _np(leaf), _np(node), _np(contrib)
    rng = np.random.default_rng(3)
    q = np.concatenate([p[:3000], rng.uniform(-1.5, 1.5, (1000, dim))]).astype(np.float32)
    cases = [
        ("barnes_hut_gravity", O.L3_POTENTIAL, dict(beta=2.0), {}),
        ("barnes_hut_force", O.L3_FORCE, dict(beta=2.0), {}),
        ("softened_barnes_hut_force", O.L3_SOFT,
         dict(beta=1.5, softening_scale=0.02, cutoff_scale=0.5), dict(eps=0.02, rc=0.5)),
        ("softened_barnes_hut_force", O.L3_SOFT,
         dict(beta=2.0, softening_scale=0.01, cutoff_scale=float("inf")), dict(eps=0.01)),
        ("barnes_hut_force", O.L3_FORCE, dict(beta=2.0, periodic=True, box_length=2.0),
         dict(periodic=True, box=2.0)),
    ]
    for name, kind, kw, okw in cases:
        out = _walk(ffi, name, leaf, node, contrib, q, threads=threads, **kw)
        ref = O.nd_eval(lb, nb, cb, q, kind, beta=kw["beta"], threads=threads, **okw)
        # fp32 sums of ~1e4 terms of both signs (forces up to 1e7 in the clump)
        assert O.rel_l2(out, ref) < 1e-5, (name, kw)


@pytest.mark.parametrize("dim", [2, 3])
def test_tree_stochastic_walks_match_oracle(gpu, dim):
    ffi = _ffi()
    p, nv, m = _case(20000, dim, 41, "clump")
    lo, hi = _corners(p, dim)
    leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi)
    lb, nb, cb = _np(leaf), _np(node), _np(contrib)
    q = p[:2048]
    for threads in (64, 256):
        for name, kind, kw, okw in [
            ("stochastic_barnes_hut_gravity", O.L3_POTENTIAL, {}, {}),
            ("stochastic_barnes_hut_force", O.L3_FORCE, {}, {}),
            ("softened_stochastic_barnes_hut_force", O.L3_SOFT,
             dict(softening_scale=0.02, cutoff_scale=0.5), dict(eps=0.02, rc=0.5)),
            ("stochastic_barnes_hut_force", O.L3_FORCE, dict(periodic=True, box_length=2.0),
             dict(periodic=True, box=2.0)),
        ]:
            out = _walk(ffi, name, leaf, node, contrib, q, samples_per_subdomain=2, seed=7,
                        threads=threads, **kw)
            ref = O.nd_eval(lb, nb, cb, q, kind, samples=2, seed=7, threads=threads, **okw)
            assert O.rel_l2(out, ref) < 1e-4, (name, threads)
    a = _walk(ffi, "stochastic_barnes_hut_force", leaf, node, contrib, q, seed=3)
    b = _walk(ffi, "stochastic_barnes_hut_force", leaf, node, contrib, q, seed=3)
    np.testing.assert_array_equal(a, b)


def test_tree_stochastic_unbiased_over_blocks(gpu):
    # One query repeated, one query per block: independent sample points and
    # paths.  The oracle's estimates (same paths, fp64 sums) average to the
    # direct sum within their standard error.  The device's estimates follow
    # the same paths in fp32: their mean also carries the fp32 rounding of
    # every node's (children - parent) difference, which the estimator sums
    # over all nodes (about 1e-4 of the force here), as the reference's fp32
    # arithmetic does.
    ffi = _ffi()
    p, nv, m = _case(3000, 2, 51, "clump")
    lo, hi = _corners(p, 2)
    leaf, node, contrib = _build(ffi, 2, p, nv, m, lo, hi)
    K = 20000
    q = np.full((K, 2), 0.1, np.float32)
    est = _walk(ffi, "stochastic_barnes_hut_force", leaf, node, contrib, q, seed=9, threads=1)
    ref = O.nd_eval(_np(leaf), _np(node), _np(contrib), q, O.L3_FORCE, samples=1, seed=9,
                    threads=1)
    d = (q[:1, None] - p[None]).astype(np.float64)
    exact = (-m[None, :, None] * d / (np.linalg.norm(d, axis=2)[..., None] + 0.1) ** 3).sum(1)[0]
    se = ref.std(0, ddof=1) / np.sqrt(K)
    z = np.abs(ref.mean(0) - exact) / se
    assert (se > 0).all() and (z < 4.5).all(), (ref.mean(0), exact, se)
    assert O.rel_l2(est, ref) < 1e-3
    mean = est.astype(np.float64).mean(0)
    assert (np.abs(mean - ref.mean(0)) < 5e-4 * np.abs(exact)).all(), (mean, ref.mean(0))


def test_quadtree_reference_examples(gpu):
    # python/example_jax_bh_compare.py:20-60 and
    # python/sbh_jax_examples/stochastic_2d.py:19-45
    ffi = _ffi()
    rng = np.random.default_rng(0)
    n = 300
    points = rng.uniform(-1.0, 1.0, (n, 2)).astype(np.float32)
    normals = np.zeros_like(points)
    masses = np.full(n, 1.0 / n, np.float32)
    leaf, node = ffi.build_quadtree_buffers(points, normals, masses, level=3)
    assert leaf.dtype == node.dtype == torch.uint8 and leaf.numel() == 24 * n
    f = _np(ffi.barnes_hut_force(leaf, node, points, beta=2.0))
    d = (points[:, None] - points[None]).astype(np.float64)
    exact = (-masses[None, :, None] * d /
             (np.linalg.norm(d, axis=2)[..., None] + 0.1) ** 3).sum(1)
    assert f.shape == (n, 2) and O.rel_l2(f, exact) < 1e-3
    leaf, node, contrib = ffi.build_quadtree_buffers_with_contrib(points, normals, masses,
                                                                  level=3)
    s = _np(ffi.stochastic_barnes_hut_gravity(leaf, node, contrib, points,
                                              samples_per_subdomain=2, seed=123))
    bf = (-masses[None] / (np.linalg.norm(d, axis=2) + 0.1)).sum(1)
    assert s.shape == (n,) and np.isfinite(s).all() and O.rel_l2(s, bf) < 0.05
    g = _np(ffi.barnes_hut_gravity(leaf, node, points, beta=2.0))
    assert O.rel_l2(g, bf) < 1e-3


def test_quadtree_rejects(gpu):
    ffi = _ffi()
    p, nv, m = _case(100, 2, 61)
    with pytest.raises(ValueError):
        ffi.build_quadtree_buffers(p, nv, m, level=2)
    with pytest.raises(ValueError):
        ffi.build_quadtree_buffers(np.zeros((5, 3), np.float32), np.zeros((5, 3), np.float32),
                                   np.ones(5, np.float32))
    leaf, node, contrib = ffi.build_quadtree_buffers_with_contrib(p, nv, m)
    q3 = np.zeros((4, 3), np.float32)
    with pytest.raises(ValueError):
        ffi.barnes_hut_force(leaf, node, q3)  # 2-D buffers read as 3-D records
    with pytest.raises(ValueError):
        ffi.stochastic_barnes_hut_force(leaf, node[:-3], contrib, p)
    with pytest.raises(ValueError):
        ffi.build_octree_buffers(np.zeros((5, 3), np.float32), np.zeros((5, 3), np.float32),
                                 np.ones(5, np.float32), backend="gpu")


@pytest.mark.parametrize("dim", [2, 3])
def test_tree_vote_walk_rope_equals_plain_walk(gpu, dim):
    # Trees of >= 2048 nodes are walked over the compact rope records; under
    # stream capture (no workspace allocation) the plain walk over the node
    # records runs instead.  Same visit order and votes: identical bits.
    ffi = _ffi()
    p, nv, m = _case(20000, dim, 71, "clump")
    lo, hi = _corners(p, dim)
    leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi)
    rng = np.random.default_rng(4)
    Q = torch.tensor(np.concatenate([p[:2000], rng.uniform(-1.5, 1.5, (1000, dim))]),
                     dtype=torch.float32, device="cuda")
    eager = {
        "f": _np(ffi.barnes_hut_force(leaf, node, Q, beta=2.0, threads=96)),
        "g": _np(ffi.barnes_hut_gravity(leaf, node, Q, beta=1.5)),
        "s": _np(ffi.softened_barnes_hut_force(leaf, node, Q, beta=2.0, softening_scale=0.01,
                                               cutoff_scale=0.5)),
    }
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            ffi.barnes_hut_force(leaf, node, Q, beta=2.0, threads=96)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        f = ffi.barnes_hut_force(leaf, node, Q, beta=2.0, threads=96)
        g = ffi.barnes_hut_gravity(leaf, node, Q, beta=1.5)
        s = ffi.softened_barnes_hut_force(leaf, node, Q, beta=2.0, softening_scale=0.01,
                                          cutoff_scale=0.5)
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_np(f), eager["f"])
        np.testing.assert_array_equal(_np(g), eager["g"])
        np.testing.assert_array_equal(_np(s), eager["s"])


def test_tree_builder_full_size_2d(gpu):
    # 2^20 points (the quadtree workload's size): bytes equal the oracle's
    ffi = _ffi()
    rng = np.random.default_rng(7000)
    p = rng.uniform(-1.0, 1.0, (1 << 20, 2)).astype(np.float32)
    nv = np.zeros_like(p)
    m = np.full(1 << 20, 1.0 / (1 << 20), np.float32)
    lo, hi = _corners(p, 2)
    leaf, node, contrib = _build(ffi, 2, p, nv, m, lo, hi)
    ol, on, oc = O.nd_build(p, nv, m, lo, hi)
    assert node.numel() == on.size
    assert np.array_equal(_np(leaf), ol)
    assert np.array_equal(_np(contrib), oc)
    assert np.array_equal(_np(node), on)


@pytest.mark.parametrize("dim", [2, 3])
def test_rope_walk_far_test_edge_betas(gpu, dim):
    # The rope walk decides most votes from a per-call bound (beta diam)^2
    # (1 + 2^-14) and runs the exact ratio only near the threshold; beta <= 0,
    # beta = +inf, NaN and extreme scales take the exact path.  The plain walk
    # (run under stream capture) always evaluates the exact ratio: both must
    # give identical bits for every beta, on a tree with coincident points
    # (zero-diameter nodes) and a tight clump.
    ffi = _ffi()
    p, nv, m = _case(20000, dim, 81, "clump")
    p[:300] = p[5000]
    lo, hi = _corners(p, dim)
    leaf, node, contrib = _build(ffi, dim, p, nv, m, lo, hi)
    rng = np.random.default_rng(9)
    Q = torch.tensor(np.concatenate([p[:1500], p[5000:5064],
                                     rng.uniform(-1.5, 1.5, (500, dim))]),
                     dtype=torch.float32, device="cuda")
    betas = [2.0, 0.5, 0.0, -1.0, float("inf"), float("nan"), 1e-30, 1e30, 3e19]
    eager = [_np(ffi.barnes_hut_force(leaf, node, Q, beta=b)) for b in betas]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ffi.barnes_hut_force(leaf, node, Q, beta=2.0)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = [ffi.barnes_hut_force(leaf, node, Q, beta=b) for b in betas]
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        for b, o, e in zip(betas, outs, eager):
            np.testing.assert_array_equal(_np(o), e, err_msg=f"beta={b}")
    # beta = 1e30 opens every node (its bound overflows, so every vote takes
    # the exact ratio): the sum over all leaves, finite
    assert np.isfinite(eager[betas.index(1e30)]).all()
    assert node.numel() // ffi.NODE_SIZES[(dim, 3)] >= 2048  # the rope path ran
    # Node bytes that no builder of this library writes: internal nodes with
    # negated diameters.  The exact ratio is then negative (the node opens);
    # the rope has no usable sure-far bound for such a node and must agree.
    size = ffi.NODE_SIZES[(dim, 3)]
    off_d, off_leaf = (48, 612) if dim == 2 else (56, 4268)
    nb = node.clone().view(-1, size)
    internal = torch.nonzero(nb[:, off_leaf] == 0).flatten()[1::3]
    assert internal.numel() > 10
    d = nb[internal, off_d:off_d + 4].contiguous().view(torch.float32)
    nb[internal, off_d:off_d + 4] = (-d).contiguous().view(torch.uint8).view(-1, 4)
    neg = nb.view(-1)
    e_neg = _np(ffi.barnes_hut_force(leaf, neg, Q, beta=2.0))
    graph2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph2):
        o_neg = ffi.barnes_hut_force(leaf, neg, Q, beta=2.0)
    graph2.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(o_neg), e_neg)
