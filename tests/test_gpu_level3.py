"""GPU parity: level-3 (Tree3DNode<1,3>) builder and walks against the oracle.

Builder: leaf records in stable (cell, index) order and every node field
except the mass moments must equal the oracle bit for bit; moments (fp32 cell
sums vs fp64) within a stated tolerance.  Walks: the same warp votes, hence
the same paths, so results differ from the oracle only by fp32 accumulation.
The reference's own chain test (python/test_jax_ffi_tree_builder.py:125-205)
is replicated at the end.
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


def _case(n, seed, clump=False):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (n, 3))
    if clump:
        p[: n // 2] = 0.4 + rng.normal(0, 1e-3, (n // 2, 3))
    p = p.astype(np.float32)
    nv = rng.normal(size=(n, 3)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32)
    return p, nv, m


EXACT_FIELDS = ["start", "end", "buffer_idx", "parent_idx", "bmin", "bmax", "diameter",
                "bitmask", "child", "num_children", "depth", "max_depth", "is_leaf"]


@pytest.mark.parametrize("n,seed,clump", [(5000, 1, False), (3000, 2, True), (1, 3, False),
                                          (0, 4, False), (77, 5, False)])
def test_level3_builder_matches_oracle(gpu, n, seed, clump):
    p, nv, m = _case(n, seed, clump)
    if n > 0:
        lo, hi = O.default_bounds(p)
    else:
        lo, hi = -np.ones(3, np.float32), np.ones(3, np.float32)
    leaf, node, contrib = _ffi().build_octree_buffers_with_contrib(p, nv, m, lo, hi)
    leaf, node, contrib = _np(leaf), _np(node), _np(contrib)
    rl, rn, rc = O.build_level3(p, nv, m, lo, hi)
    np.testing.assert_array_equal(leaf, rl)
    g, r = O.decode_nodes3(node), O.decode_nodes3(rn)
    for f in EXACT_FIELDS:
        np.testing.assert_array_equal(g[f], r[f], err_msg=f)
    gm, rm = g["contrib"], r["contrib"]
    np.testing.assert_array_equal(gm["alpha"][rm["mass"] > 0], 1.0)
    np.testing.assert_allclose(gm["mass"], rm["mass"], rtol=2e-5, atol=0)
    big = rm["mass"] > 0
    for f in ("pos", "normal"):
        if big.any():
            np.testing.assert_allclose(gm[f][big], rm[f][big], rtol=0,
                                       atol=2e-5 * max(np.abs(rm[f][big]).max(), 1.0))
    gc, rcd = O.decode_leaf3(contrib), O.decode_leaf3(rc)
    np.testing.assert_array_equal(gc["mass"], gm["mass"])


@pytest.mark.parametrize("threads", [64, 100, 256])
def test_level3_walks_match_oracle(gpu, threads):
    ffi = _ffi()
    p, nv, m = _case(4000, 11)
    lo, hi = O.default_bounds(p)
    leaf, node, contrib = ffi.build_octree_buffers_with_contrib(p, nv, m, lo, hi)
    lb, nb = _np(leaf), _np(node)
    rng = np.random.default_rng(12)
    # interior queries, far queries, and a mix inside some warps
    far = (rng.normal(size=(300, 3)) * 0.2 + np.array([25.0, -3.0, 1.0])).astype(np.float32)
    q = np.concatenate([p[:400], far, p[400:440], far[:50]]).astype(np.float32)
    out = _np(ffi.barnes_hut_gravity(leaf, node, q, beta=2.0, threads=threads))
    ref = O.l3_eval(lb, nb, q, O.L3_POTENTIAL, beta=2.0, threads=threads)
    assert O.rel_l2(out, ref) < 1e-5
    np.testing.assert_allclose(out, ref, rtol=1e-4)
    out = _np(ffi.barnes_hut_force(leaf, node, q, beta=2.0, threads=threads))
    ref = O.l3_eval(lb, nb, q, O.L3_FORCE, beta=2.0, threads=threads)
    assert O.rel_l2(out, ref) < 1e-5
    out = _np(ffi.softened_barnes_hut_force(leaf, node, q, beta=2.0, softening_scale=0.05,
                                            cutoff_scale=1.5, threads=threads))
    ref = O.l3_eval(lb, nb, q, O.L3_SOFT, beta=2.0, threads=threads, eps=0.05, rc=1.5)
    assert O.rel_l2(out, ref) < 1e-5


def test_level3_far_warps_use_root_monopole(gpu):
    ffi = _ffi()
    p, nv, m = _case(2000, 13)
    leaf, node = ffi.build_octree_buffers(p, nv, m)
    root = O.decode_nodes3(_np(node))[0]
    com = root["contrib"]["pos"].astype(np.float64)
    q = (com + np.array([40.0, 0, 0]) +
         np.random.default_rng(1).normal(0, 0.1, (128, 3))).astype(np.float32)
    q[70] = p[0]  # the warp holding query 70 sums exactly
    out = _np(ffi.barnes_hut_gravity(leaf, node, q, beta=2.0, threads=64))
    mono = -float(root["contrib"]["mass"]) / (np.linalg.norm(q - com, axis=1) + 0.1)
    d = p[None].astype(np.float64) - q[:, None].astype(np.float64)
    exact = (-m / (np.linalg.norm(d, axis=2) + 0.1)).sum(1)
    warp = np.arange(128) // 32
    np.testing.assert_allclose(out[warp != 2], mono[warp != 2], rtol=1e-5)
    np.testing.assert_allclose(out[warp == 2], exact[warp == 2], rtol=1e-4)


def test_level3_stochastic_is_exact_sum(gpu):
    ffi = _ffi()
    p, nv, m = _case(3000, 14)
    leaf, node, contrib = ffi.build_octree_buffers_with_contrib(p, nv, m)
    q = p[:500]
    out = _np(ffi.softened_stochastic_barnes_hut_force(leaf, node, contrib, q,
                                                       samples_per_subdomain=2, seed=7,
                                                       softening_scale=0.05, cutoff_scale=1.5))
    assert O.rel_l2(out, O.direct_force_soft(p, m, q, 0.05, 1.5)) < 1e-5
    g = _np(ffi.stochastic_barnes_hut_gravity(leaf, node, contrib, q, seed=3))
    ref = O.l3_eval(_np(leaf), _np(node), q, O.L3_POTENTIAL, samples=1)
    assert O.rel_l2(g, ref) < 1e-5
    f = _np(ffi.stochastic_barnes_hut_force(leaf, node, contrib, q, samples_per_subdomain=0))
    assert np.isnan(f).all()


def test_level3_rejects_bad_arguments(gpu):
    ffi = _ffi()
    p, nv, m = _case(100, 15)
    leaf, node, contrib = ffi.build_octree_buffers_with_contrib(p, nv, m)
    with pytest.raises(ValueError):
        ffi.barnes_hut_force(leaf, node[:-5], p)
    with pytest.raises(ValueError):
        ffi.barnes_hut_force(leaf, node, p[:, :2])
    with pytest.raises(ValueError):
        ffi.barnes_hut_force(leaf, node, p, threads=2000)
    with pytest.raises(ValueError):
        ffi.build_octree_buffers(p, nv, m, level=4)


def test_reference_octree_builder_ffi_chain_3d(gpu):
    # python/test_jax_ffi_tree_builder.py:125-205
    ffi = _ffi()
    rng = np.random.default_rng(0)
    n = 256
    points = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    normals = np.zeros_like(points)
    masses = np.ones(n, np.float32)
    leaf, node, contrib = ffi.build_octree_buffers_with_contrib(points, normals, masses)
    assert leaf.dtype == node.dtype == contrib.dtype == torch.uint8
    f_bh = ffi.barnes_hut_force(leaf, node, points, beta=2.0)
    f_sbh = ffi.stochastic_barnes_hut_force(leaf, node, contrib, points,
                                            samples_per_subdomain=2, seed=7)
    assert tuple(f_bh.shape) == (n, 3) and tuple(f_sbh.shape) == (n, 3)
    assert np.isfinite(_np(f_bh)).all() and np.isfinite(_np(f_sbh)).all()
    f_soft = ffi.softened_barnes_hut_force(leaf, node, points, beta=2.0, softening_scale=5e-2,
                                           cutoff_scale=1.5)
    f_ssoft = ffi.softened_stochastic_barnes_hut_force(leaf, node, contrib, points,
                                                       samples_per_subdomain=2, seed=11,
                                                       softening_scale=5e-2, cutoff_scale=1.5)
    assert np.isfinite(_np(f_soft)).all() and np.isfinite(_np(f_ssoft)).all()
    # rebuild-and-evaluate step, as the reference's jitted step
    pos = torch.tensor(points, device="cuda")
    lb, nb, cb = ffi.build_octree_buffers_with_contrib(pos, torch.zeros_like(pos),
                                                       torch.ones(n, device="cuda"))
    f = ffi.stochastic_barnes_hut_force(lb, nb, cb, pos, samples_per_subdomain=2, seed=5)
    assert np.isfinite(_np(f)).all()
