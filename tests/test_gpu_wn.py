"""GPU parity: winding numbers (brute force, octree8 classic and stochastic
walks, adaptive-tree group walk) against the oracle.

The octree8 walks make the oracle's MAC decisions and the stochastic walk
draws the oracle's Philox numbers, so both agree with it to fp32
accumulation; the adaptive walk builds the same tree and query groups, so
its interaction counts are identical to the oracle's.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _wn():
    from sbh_amd import wn

    return wn


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _mesh(freq, seed):
    from sbh_amd import synthetic as S

    V, F = S.perturbed_sphere(freq, seed)
    c, n, a = S.face_dipoles(V, F)
    return V, F, c, n, a


def _queries(m, seed, lo=-1.5, hi=1.5):
    return np.random.default_rng(seed).uniform(lo, hi, (m, 3)).astype(np.float32)


def test_bruteforce_wn_matches_oracle(gpu):
    _, _, c, n, a = _mesh(12, 1)
    q = _queries(3000, 2)
    out = _np(_wn().eval_brute_force_wn(c, n, a, q))
    ref = O.bruteforce_wn(c, n, a, q)
    assert O.rel_l2(out, ref) < 1e-5
    np.testing.assert_allclose(out, ref, rtol=1e-3, atol=1e-5)


def test_octree8_wn_walks_match_oracle(gpu):
    from sbh_amd import ffi

    _, _, c, n, a = _mesh(16, 3)
    q = _queries(2000, 4)
    leaf, node = ffi.build_octree8_buffers(c, n, a, max_depth=6)
    lb, nb = _np(leaf), _np(node)
    for beta in (1.0, 2.0):
        out = _np(_wn().eval_barnes_hut_wn(leaf, node, q, beta=beta))
        ref, _ = O.bh_octree8(lb, nb, q, O.KIND_WN, beta)
        assert O.rel_l2(out, ref) < 1e-5, beta
    for samples, seed in ((1, 0), (8, 77)):
        out = _np(_wn().multi_level_prefix_control_variate_wn(
            leaf, node, q, samples_per_subdomain=samples, seed=seed))
        ref = O.sbh_octree8(lb, nb, q, O.KIND_WN, samples, seed, 1.0)
        assert O.rel_l2(out, ref) < 1e-5, (samples, seed)


def test_octree8_wn_huge_beta_is_brute_force(gpu):
    from sbh_amd import ffi

    _, _, c, n, a = _mesh(8, 5)
    q = _queries(500, 6)
    leaf, node = ffi.build_octree8_buffers(c, n, a, max_depth=5)
    out = _np(_wn().eval_barnes_hut_wn(leaf, node, q, beta=1e30))
    assert O.rel_l2(out, O.bruteforce_wn(c, n, a, q)) < 1e-5


@pytest.mark.parametrize("freq,m,leaf_size", [(12, 5000, 16), (24, 20000, 8), (6, 100, 64)])
def test_adaptive_wn_matches_oracle(gpu, freq, m, leaf_size):
    _, _, c, n, a = _mesh(freq, 7)
    q = _queries(m, 8)
    st = torch.zeros(8, dtype=torch.int64, device="cuda")
    out = _np(_wn().winding_number(c, n, a, q, beta=2.0, max_depth=10, leaf_size=leaf_size,
                                   stats=st))
    T = O.Tree(c, a, max_depth=10, leaf_size=leaf_size)
    ref = T.wn_eval(n, q, beta=2.0)
    s = _np(st)
    # same tree, same query groups, same MAC: identical interaction sets
    assert s[3] == T.num_nodes and s[5] == 0  # no capacity overflow, no truncated walk
    assert (s[0], s[1], s[2]) == (ref["far"], ref["near"], ref["groups"])
    assert O.rel_l2(out, ref["wn"]) < 1e-5
    # and within the BH error of the exact dipole sum
    assert O.rel_l2(out, O.bruteforce_wn(c, n, a, q)) < 3e-3


def test_adaptive_wn_is_indicator_of_closed_mesh(gpu):
    V, F, c, n, a = _mesh(32, 9)
    q = _queries(2000, 10)
    out = _np(_wn().winding_number(c, n, a, q))
    exact = O.triangle_wn(V, F, q)
    # away from the surface the dipole sum converges to the indicator
    from scipy.spatial import cKDTree

    d, _ = cKDTree(c).query(q)
    far = d > 0.1
    assert far.sum() > 500
    np.testing.assert_allclose(out[far], exact[far], atol=1e-2)
    assert np.all(np.abs(np.round(exact[far]) - exact[far]) < 1e-9)


def test_adaptive_wn_huge_beta_and_determinism(gpu):
    _, _, c, n, a = _mesh(10, 11)
    q = _queries(3000, 12)
    wn = _wn()
    out = _np(wn.winding_number(c, n, a, q, beta=1e30))
    assert O.rel_l2(out, O.bruteforce_wn(c, n, a, q)) < 1e-5
    ws = wn.new_workspace(c.shape[0], q.shape[0])
    x = _np(wn.winding_number(c, n, a, q, workspace=ws))
    y = _np(wn.winding_number(c, n, a, q, workspace=ws))
    np.testing.assert_array_equal(x, y)


def test_adaptive_wn_edge_cases(gpu):
    wn = _wn()
    _, _, c, n, a = _mesh(4, 13)
    out = _np(wn.winding_number(c, n, a, np.zeros((0, 3), np.float32)))
    assert out.shape == (0,)
    q = _queries(70, 14)
    one = _np(wn.winding_number(c[:1], n[:1], a[:1], q))
    np.testing.assert_allclose(one, O.bruteforce_wn(c[:1], n[:1], a[:1], q), rtol=1e-4,
                               atol=1e-7)
    with pytest.raises(ValueError):
        wn.winding_number(c, n[:-1], a, q)
    with pytest.raises(ValueError):
        wn.winding_number(c, n, a, q, max_depth=11)
