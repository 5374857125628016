"""GPU parity: potential and smooth-distance fields (sbh_amd.fields, the
reference's src/sbh_gpu.cu module functions) against the oracle: brute
force vs the FP64 sums, classic walks vs the oracle's walk (same
interactions), stochastic walks vs the oracle's sample paths."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _F():
    from sbh_amd import fields

    return fields


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _case(n, depth, seed, plummer=False):
    rng = np.random.default_rng(seed)
    if plummer:
        xmax = 20.0 ** 3 / (20.0 ** 2 + 1.0) ** 1.5
        r = 1.0 / np.sqrt(rng.uniform(0, xmax, n) ** (-2.0 / 3.0) - 1.0)
        v = rng.normal(size=(n, 3))
        p = (r[:, None] * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    else:
        p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32)
    lo, hi = O.default_bounds(p)
    leaf, node, _ = O.build_octree8(p, np.zeros_like(p), m, lo, hi, depth)
    return p, m, leaf, node


def test_brute_force_fields_match_fp64(gpu):
    F = _F()
    rng = np.random.default_rng(1)
    p = rng.uniform(-1, 1, (5000, 3)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, 5000).astype(np.float32)
    a = rng.uniform(0.5, 4.0, 5000).astype(np.float32)
    q = rng.uniform(-1.5, 1.5, (700, 3)).astype(np.float32)
    d = np.linalg.norm(p[None].astype(np.float64) - q[:, None].astype(np.float64), axis=2)
    pot = _np(F.eval_brute_force_gravity(p, m, q))
    np.testing.assert_allclose(pot, (-m.astype(np.float64)[None] / (d + 0.1)).sum(1), rtol=1e-5)
    sd = _np(F.eval_brute_force_smoothdist(p, q, alpha=a))
    np.testing.assert_allclose(sd, O.bruteforce_smoothdist(p, a, q), rtol=1e-6)
    sd1 = _np(F.eval_brute_force_smoothdist(p, q))
    np.testing.assert_allclose(sd1, np.exp(-d).sum(1), rtol=1e-6)
    empty = _np(F.eval_brute_force_smoothdist(p, q[:0]))
    assert empty.shape == (0,)


@pytest.mark.parametrize("beta", [0.8, 2.0])
@pytest.mark.parametrize("nq", [600, 5000])
def test_classic_field_walks_match_oracle(gpu, beta, nq):
    # nq = 5000 takes the Morton-ordered schedule; results per query are equal
    F = _F()
    p, m, leaf, node = _case(6000, 6, 2, plummer=True)
    q = np.concatenate([p[::3][: nq - 100],
                        np.random.default_rng(3).normal(0, 3, (100, 3)).astype(np.float32)])
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    out = _np(F.eval_barnes_hut_gravity(lt, nt, q, beta=beta))
    ref, _ = O.bh_octree8(leaf, node, q, O.KIND_POTENTIAL, beta)
    assert O.rel_l2(out, ref) < 1e-5
    out = _np(F.eval_barnes_hut_smoothdist(lt, nt, q, beta=beta))
    ref, _ = O.bh_octree8(leaf, node, q, O.KIND_SMOOTH, beta)
    assert O.rel_l2(out, ref) < 1e-5


def test_classic_field_walk_huge_beta_is_brute_force(gpu):
    F = _F()
    p, m, leaf, node = _case(3000, 5, 4)
    q = np.random.default_rng(5).uniform(-1.5, 1.5, (400, 3)).astype(np.float32)
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    pot = _np(F.eval_barnes_hut_gravity(lt, nt, q, beta=1e30))
    np.testing.assert_allclose(pot, _np(F.eval_brute_force_gravity(p, m, q)), rtol=2e-5)
    sd = _np(F.eval_barnes_hut_smoothdist(lt, nt, q, beta=1e30))
    np.testing.assert_allclose(sd, _np(F.eval_brute_force_smoothdist(p, q)), rtol=2e-5)


@pytest.mark.parametrize("samples,seed", [(1, 0), (4, 99)])
def test_stochastic_field_walks_match_oracle(gpu, samples, seed):
    F = _F()
    p, m, leaf, node = _case(5000, 6, 6)
    q = p[::9].copy()
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    kw = dict(samples_per_subdomain=samples, seed=seed, beta=1.0)
    out = _np(F.multi_level_prefix_control_variate_gravity(lt, nt, q, **kw))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_POTENTIAL, samples, seed, 1.0)
    assert O.rel_l2(out, ref) < 1e-5
    other = O.sbh_octree8(leaf, node, q, O.KIND_POTENTIAL, samples, seed + 1, 1.0)
    assert O.rel_l2(out, other) > 100 * O.rel_l2(out, ref)
    out = _np(F.multi_level_prefix_control_variate_smoothdist(lt, nt, q, **kw))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_SMOOTH, samples, seed, 1.0)
    assert O.rel_l2(out, ref) < 1e-5


def test_field_arguments_are_checked(gpu):
    F = _F()
    p, m, leaf, node = _case(500, 4, 7)
    lt, nt = torch.tensor(leaf, device="cuda"), torch.tensor(node, device="cuda")
    with pytest.raises(ValueError):
        F.eval_barnes_hut_gravity(lt, nt, p[:, :2])
    with pytest.raises(ValueError):
        F.eval_brute_force_smoothdist(p, p, alpha=np.ones(3, np.float32))
    with pytest.raises(ValueError):
        F.multi_level_prefix_control_variate_smoothdist(lt, nt, p, threads=0)
