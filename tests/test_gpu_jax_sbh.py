"""GPU parity: the reference's tree fields module (python/jax_sbh.py over the
pybind11 GPU module src/sbh_gpu.cu) against the oracle.

Trees Tree2D<1, L> / Tree3D<1, L> for L = 1, 2, 3 built on the device must
equal the oracle's bytes word for word; the vote walk and the sampler (32-query
blocks, the module's seed 0 for every query) take the oracle's paths, so they
differ only by fp32 against fp64 sums; the compensated brute force is within
fp32 rounding of the fp64 sum.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mod():
    from sbh_amd import jax_sbh

    return jax_sbh


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _case(n, dim, seed, clump=True):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (n, dim))
    if clump:
        p[: n // 3] = 0.25 + rng.normal(0, 5e-3, (n // 3, dim))
    p = p.astype(np.float32)
    nv = rng.normal(size=(n, dim))
    nv = (nv / np.linalg.norm(nv, axis=1, keepdims=True)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32) / n
    lo = (p.min(0) - np.float32(1e-3)).astype(np.float32)
    hi = (p.max(0) + np.float32(1e-3)).astype(np.float32)
    return p, nv, m, lo, hi


def _same_words(a, b):
    a = np.ascontiguousarray(a, np.uint8).view(np.uint32)
    b = np.ascontiguousarray(b, np.uint8).view(np.uint32)
    assert a.shape == b.shape
    diff = np.nonzero(a != b)[0]
    if diff.size:
        fa, fb = a[diff].view(np.float32), b[diff].view(np.float32)
        assert (np.isnan(fa) & np.isnan(fb)).all(), (diff[:8], a[diff[:8]], b[diff[:8]])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("level", [1, 2, 3])
def test_module_tree_bytes_match_oracle(gpu, dim, level):
    J = _mod()
    p, nv, m, lo, hi = _case(6000, dim, 100 + level)
    t = J.build_gpu_tree(p, nv, m, lo, hi, alpha=0.7, level=level)
    ol, on, oc = O.nd_build(p, nv, m, lo, hi, alpha=0.7, level=level)
    assert t.node.numel() == on.size
    _same_words(_np(t.leaf), ol)
    _same_words(_np(t.node), on)
    _same_words(_np(t.contrib), oc)


KINDS = [(0, O.L3_POTENTIAL, "gravity"), (1, O.ND_WN, "wn"), (2, O.ND_SMOOTH, "smoothdist")]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("level", [1, 2, 3])
def test_module_walks_match_oracle(gpu, dim, level):
    J = _mod()
    p, nv, m, lo, hi = _case(8000, dim, 200 + level)
    t = J.build_gpu_tree(p, nv, m, lo, hi, alpha=2.0, level=level)
    lb, nb, cb = _np(t.leaf), _np(t.node), _np(t.contrib)
    rng = np.random.default_rng(7)
    q = np.concatenate([p[:1500], rng.uniform(-1.3, 1.3, (500, dim))]).astype(np.float32)
    for _, okind, name in KINDS:
        bh = _np(getattr(J, f"eval_barnes_hut_{name}")(t, q, beta=2.0))
        ref = O.nd_eval(lb, nb, cb, q, okind, beta=2.0, threads=32, level=level)
        assert O.rel_l2(bh, ref) < 2e-5, (name, "walk")
        fn = J.eval_mlpcv_gravity if name == "gravity" else \
            getattr(J, f"multi_level_prefix_control_variate_{name}")
        st = _np(fn(t, q, samples_per_subdomain=2))
        ref = O.nd_eval(lb, nb, cb, q, okind, samples=2, seed=0, threads=32, level=level,
                        seed_per_query=False)
        assert O.rel_l2(st, ref) < 2e-4, (name, "sampler")
        bf = _np(getattr(J, f"eval_brute_force_{name}")(t, q))
        ref = O.nd_brute(lb, q, okind)
        assert O.rel_l2(bf, ref) < 1e-6, (name, "brute force")


def test_module_reference_usage(gpu):
    # python/jax_sbh.py defaults: level 1 (a true quadtree / octree), beta 2,
    # one sample per subdomain; the walk is within a BH error of the exact sum
    J = _mod()
    rng = np.random.default_rng(0)
    for dim in (2, 3):
        p = rng.uniform(-1, 1, (3000, dim)).astype(np.float32)
        nv = np.zeros_like(p)
        m = np.full(3000, 1.0 / 3000, np.float32)
        t = J.build_gpu_tree(p, nv, m)
        assert t.level == 1 and t.leaf.numel() == 3000 * (24 if dim == 2 else 32)
        exact = _np(J.eval_brute_force_gravity(t, p))
        bh = _np(J.eval_barnes_hut_gravity(t, p))
        assert O.rel_l2(bh, exact) < 1e-3
        mc = _np(J.eval_mlpcv_gravity(t, p))
        # one sample per subdomain: a noisy estimate (its paths are checked
        # against the oracle above); a sanity bound only
        assert np.isfinite(mc).all() and O.rel_l2(mc, exact) < 0.5
        w = J.build_gpu_worker(p, nv, m, level=2)
        assert w.level == 2 and O.rel_l2(_np(J.eval_barnes_hut_gravity(w, p)), exact) < 1e-3


def test_module_rejects(gpu):
    J = _mod()
    p, nv, m, lo, hi = _case(200, 2, 300)
    with pytest.raises(ValueError):
        J.build_gpu_tree(p, nv, m, level=4)
    t = J.build_gpu_tree(p, nv, m)
    with pytest.raises(ValueError):
        J.eval_barnes_hut_gravity(t, np.zeros((5, 3), np.float32))
    with pytest.raises(ValueError):
        J.eval_mlpcv_gravity(t, p, samples_per_subdomain=-1)


@pytest.mark.parametrize("level", [1, 3])
def test_module_rope_walks_equal_plain_walks(gpu, level):
    # Eager calls on trees of >= 2048 nodes walk the compact rope records (with
    # normal and alpha for the winding number and the smooth distance); under
    # stream capture the plain walk over the node records runs.  Identical bits.
    J = _mod()
    p, nv, m, lo, hi = _case(8000, 3, 400 + level)
    t = J.build_gpu_tree(p, nv, m, lo, hi, alpha=1.5, level=level)
    Q = torch.tensor(p[:3000], device="cuda")
    names = ["gravity", "wn", "smoothdist"]
    eager = {k: _np(getattr(J, f"eval_barnes_hut_{k}")(t, Q)) for k in names}
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        J.eval_barnes_hut_wn(t, Q)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = {k: getattr(J, f"eval_barnes_hut_{k}")(t, Q) for k in names}
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        for k in names:
            np.testing.assert_array_equal(_np(outs[k]), eager[k])


def test_tree_builder_random_sweep(gpu):
    # 24 random configurations: dim, level, size (0..6000), clustering,
    # coincident points, points outside the given box, masses with zeros
    from sbh_amd import ffi

    rng = np.random.default_rng(4242)
    for case in range(24):
        dim = int(rng.integers(2, 4))
        level = int(rng.integers(1, 4))
        n = int(rng.choice([0, 1, 2, 3, 17, 64, 500, 3000, 6000]))
        p = rng.uniform(-1, 1, (n, dim))
        if n > 4 and rng.random() < 0.5:
            k = n // 2
            p[:k] = rng.uniform(-1, 1, dim) + rng.normal(0, 10.0 ** rng.uniform(-6, -2), (k, dim))
        if n > 4 and rng.random() < 0.3:
            p[: n // 5] = p[-1]
        p = p.astype(np.float32)
        nv = rng.normal(size=(n, dim)).astype(np.float32)
        m = rng.uniform(0, 1, n).astype(np.float32)
        if n > 4:
            m[rng.random(n) < 0.1] = 0.0
        if rng.random() < 0.3 or n == 0:
            lo, hi = np.full(dim, -0.8, np.float32), np.full(dim, 0.9, np.float32)
        else:
            lo = (p.min(0) - np.float32(1e-3)).astype(np.float32)
            hi = (p.max(0) + np.float32(1e-3)).astype(np.float32)
        alpha = float(rng.uniform(0.1, 2.0))
        leaf, node, contrib = ffi._tree_buffers(p, nv, m, lo, hi, alpha, dim, level)
        ol, on, oc = O.nd_build(p, nv, m, lo, hi, alpha=alpha, level=level)
        info = (case, dim, level, n)
        assert node.numel() == on.size, info
        _same_words(_np(leaf), ol)
        _same_words(_np(node), on)
        _same_words(_np(contrib), oc)
