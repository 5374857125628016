"""GPU parity: stochastic Barnes-Hut walks over the dense octree8 records.

The GPU kernels and the oracle draw the same Philox4x32-10 numbers for each
(query, subdomain, sample, step), so for one seed they follow identical
sample paths and must agree to fp32 accumulation.  Statistical checks
(unbiasedness against the exact sum) and the reference's own tests
(python/test_jax_ffi_tree_builder.py:415-523,
python/test_jax_custom_vjp_recompute.py:27-101) follow.
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


def _case(kind="uniform", n=4000, depth=6, seed=3, box=None):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        lo_v, hi_v = (0.0, box) if box else (-1.0, 1.0)
        p = rng.uniform(lo_v, hi_v, (n, 3)).astype(np.float32)
    else:  # plummer sphere, truncated at 20 scale radii
        xmax = 20.0 ** 3 / (20.0 ** 2 + 1.0) ** 1.5
        r = 1.0 / np.sqrt(rng.uniform(0, xmax, n) ** (-2.0 / 3.0) - 1.0)
        v = rng.normal(size=(n, 3))
        p = (r[:, None] * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32)
    if box:
        lo, hi = np.zeros(3, np.float32), np.full(3, box, np.float32)
    else:
        lo, hi = O.default_bounds(p)
    leaf, node, _ = O.build_octree8(p, np.zeros_like(p), m, lo, hi, depth)
    return p, m, leaf, node


def _dev(*arrs):
    return [torch.tensor(a, device="cuda") for a in arrs]


SOFT = dict(softening_scale=0.03, cutoff_scale=0.6, prune_enabled=True, prune_r_cut_mult=2.5)


@pytest.mark.parametrize("kind,n,depth", [("uniform", 4000, 6), ("plummer", 6000, 7)])
@pytest.mark.parametrize("samples,seed", [(1, 0), (4, 12345)])
def test_stochastic_matches_oracle(gpu, kind, n, depth, samples, seed):
    ffi = _ffi()
    p, m, leaf, node = _case(kind, n, depth)
    q = p[::7].copy()
    c = np.random.default_rng(2).normal(size=q.shape).astype(np.float32)
    lt, nt, qt, ct = _dev(leaf, node, q, c)
    beta = 1.0
    common = dict(beta=beta, samples_per_subdomain=samples, seed=seed)

    out = _np(ffi.stochastic_barnes_hut_force_octree8(lt, nt, qt, **common))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_GRAV, samples, seed, beta)
    assert O.rel_l2(out, ref) < 1e-5
    # the check sees the sample paths: another seed is far outside tolerance
    other = O.sbh_octree8(leaf, node, q, O.KIND_GRAV, samples, seed + 1, beta)
    assert O.rel_l2(out, other) > 100 * O.rel_l2(out, ref)

    out = _np(ffi.stochastic_barnes_hut_force_octree8_vjp(lt, nt, qt, ct, **common))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_GRAV_VJP, samples, seed, beta, cot=c)
    assert O.rel_l2(out, ref) < 1e-5

    out = _np(ffi.softened_stochastic_barnes_hut_force_octree8(lt, nt, qt, **common, **SOFT))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_SOFT, samples, seed, beta, eps=0.03, rc=0.6,
                        prune_enabled=True, prune_r_cut_mult=2.5)
    assert O.rel_l2(out, ref) < 1e-5

    out = _np(ffi.softened_stochastic_barnes_hut_force_octree8_vjp(lt, nt, qt, ct, use_ffi=True,
                                                                  **common, **SOFT))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_SOFT_VJP, samples, seed, beta, eps=0.03, rc=0.6,
                        cot=c, prune_enabled=True, prune_r_cut_mult=2.5)
    assert O.rel_l2(out, ref) < 1e-5


def test_stochastic_periodic_matches_oracle(gpu):
    ffi = _ffi()
    L = 2.0
    p, m, leaf, node = _case("uniform", 3000, 6, seed=17, box=L)
    q = p[:400]
    kw = dict(beta=1.2, samples_per_subdomain=3, seed=7, softening_scale=4e-2, cutoff_scale=0.7,
              prune_enabled=True, prune_r_cut_mult=3.0, periodic=True, box_length=L)
    out = _np(ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, q, **kw))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_SOFT, 3, 7, 1.2, eps=4e-2, rc=0.7,
                        prune_enabled=True, prune_r_cut_mult=3.0, periodic=True, box=L)
    assert O.rel_l2(out, ref) < 1e-5
    out = _np(ffi.stochastic_barnes_hut_force_octree8(leaf, node, q, beta=1.2,
                                                      samples_per_subdomain=3, seed=7,
                                                      periodic=True, box_length=L))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_GRAV, 3, 7, 1.2, periodic=True, box=L)
    assert O.rel_l2(out, ref) < 1e-5


def test_stochastic_seed_determinism(gpu):
    ffi = _ffi()
    p, m, leaf, node = _case("plummer", 5000, 7, seed=8)
    lt, nt, qt = _dev(leaf, node, p)
    a = _np(ffi.stochastic_barnes_hut_force_octree8(lt, nt, qt, samples_per_subdomain=2, seed=5))
    b = _np(ffi.stochastic_barnes_hut_force_octree8(lt, nt, qt, samples_per_subdomain=2, seed=5))
    c = _np(ffi.stochastic_barnes_hut_force_octree8(lt, nt, qt, samples_per_subdomain=2, seed=6))
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    # a query's estimate does not depend on which other queries share the launch
    d = _np(ffi.stochastic_barnes_hut_force_octree8(lt, nt, qt[:129].contiguous(),
                                                    samples_per_subdomain=2, seed=5))
    np.testing.assert_array_equal(a[:129], d)


def test_stochastic_unbiased_over_seeds(gpu):
    """Mean over 256 seeds converges to the exact direct sum."""
    ffi = _ffi()
    p, m, leaf, node = _case("uniform", 2000, 6, seed=21)
    q = p[:96]
    lt, nt, qt = _dev(leaf, node, q)
    runs = np.stack([_np(ffi.softened_stochastic_barnes_hut_force_octree8(
        lt, nt, qt, beta=1.0, samples_per_subdomain=1, seed=s, softening_scale=0.03,
        cutoff_scale=float("inf"), prune_enabled=False)) for s in range(256)])
    exact = O.direct_force_soft(p, m, q, 0.03, np.inf)
    mean = runs.mean(0)
    se = runs.std(0, ddof=1) / np.sqrt(runs.shape[0])
    z = (mean - exact) / se
    assert (z ** 2).mean() < 2.0
    assert np.abs(z).max() < 5.5
    assert O.rel_l2(mean, exact) < O.rel_l2(runs[0], exact)
    # against classic BH (theta = 0.5) in the l2 sense: classic's own error is
    # far below the standard error of the 256-seed mean
    classic = _np(ffi.softened_barnes_hut_force_octree8(
        lt, nt, qt, beta=2.0, softening_scale=0.03, cutoff_scale=float("inf"),
        prune_enabled=False))
    assert np.linalg.norm(mean - classic) < 3.0 * np.linalg.norm(se)


def test_stochastic_shallow_tree_is_exact(gpu):
    # depth 3: every subdomain is a leaf, no sampling happens
    ffi = _ffi()
    p, m, leaf, node = _case("uniform", 1500, 3, seed=4)
    q = p[:200]
    out = _np(ffi.softened_stochastic_barnes_hut_force_octree8(
        leaf, node, q, samples_per_subdomain=3, seed=1, softening_scale=0.05,
        cutoff_scale=float("inf"), prune_enabled=False))
    assert O.rel_l2(out, O.direct_force_soft(p, m, q, 0.05, np.inf)) < 1e-5


def test_stochastic_empty_and_tiny(gpu):
    ffi = _ffi()
    p, m, leaf, node = _case("uniform", 1, 4, seed=4)
    out = _np(ffi.stochastic_barnes_hut_force_octree8(leaf, node, np.zeros((0, 3), np.float32)))
    assert out.shape == (0, 3)
    out = _np(ffi.stochastic_barnes_hut_force_octree8(leaf, node, p + 0.5, seed=3))
    ref = O.sbh_octree8(leaf, node, p + 0.5, O.KIND_GRAV, 1, 3, 1.0)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-7)


def test_stochastic_rejects_bad_arguments(gpu):
    ffi = _ffi()
    p, m, leaf, node = _case("uniform", 100, 4, seed=4)
    with pytest.raises(ValueError):
        ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, p, prune_enabled=True,
                                                         prune_r_cut_mult=0.0)
    with pytest.raises(ValueError):
        ffi.stochastic_barnes_hut_force_octree8_vjp(leaf, node, p, p[:10])
    with pytest.raises(ValueError):
        ffi.stochastic_barnes_hut_force_octree8(leaf, node, p, periodic=True, box_length=0.0)


# ---------------------------------------------------------- reference test replicas
def test_reference_stochastic_prune_large_cutoff_matches_disabled(gpu):
    # python/test_jax_ffi_tree_builder.py:415-465
    ffi = _ffi()
    L = 2.0
    rng = np.random.default_rng(17)
    p = rng.uniform(0, L, (320, 3)).astype(np.float32)
    m = np.ones(320, np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), m, min_corner=[0, 0, 0],
                                           max_corner=[L, L, L], max_depth=6)
    q = p[:96]
    kw = dict(samples_per_subdomain=3, seed=11, softening_scale=5e-2, cutoff_scale=0.8,
              periodic=True, box_length=L)
    a = ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, q, prune_enabled=False,
                                                         **kw)
    b = ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, q, prune_enabled=True,
                                                         prune_r_cut_mult=1e6, **kw)
    np.testing.assert_allclose(_np(a), _np(b), rtol=1e-6, atol=1e-6)


def test_reference_stochastic_prune_periodic_shift_invariance(gpu):
    # python/test_jax_ffi_tree_builder.py:468-523
    ffi = _ffi()
    L = 3.0
    rng = np.random.default_rng(19)
    p = rng.uniform(0, L, (256, 3)).astype(np.float32)
    m = np.ones(256, np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), m, min_corner=[0, 0, 0],
                                           max_corner=[L, L, L], max_depth=6)
    q = p[:48]
    kw = dict(samples_per_subdomain=3, seed=7, softening_scale=4e-2, cutoff_scale=0.7,
              prune_enabled=True, prune_r_cut_mult=3.0, periodic=True, box_length=L)
    f0 = ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, q, **kw)
    f1 = ffi.softened_stochastic_barnes_hut_force_octree8(
        leaf, node, q + np.array([L, 0, 0], np.float32), **kw)
    np.testing.assert_allclose(_np(f0), _np(f1), rtol=3e-2, atol=1e-2)


def test_reference_stochastic_vjp_against_finite_difference(gpu):
    # python/test_jax_custom_vjp_recompute.py:27-101 computes both and leaves
    # the comparison disabled; here beta is huge so the sample weights do not
    # depend on the query and the VJP is the exact derivative of the forward.
    ffi = _ffi()
    rng = np.random.default_rng(0)
    n = 200
    p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    m = rng.uniform(0.3, 1.2, n).astype(np.float32)
    c = rng.normal(size=(6, 3)).astype(np.float32)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), m, max_depth=6)
    q = (p[:6] + 0.01).astype(np.float32)
    kw = dict(beta=1e30, samples_per_subdomain=2, seed=4, softening_scale=0.1,
              cutoff_scale=1.3, prune_enabled=False)
    g_ffi = _np(ffi.softened_stochastic_barnes_hut_force_octree8_vjp(leaf, node, q, c,
                                                                     use_ffi=True, **kw))
    ref = O.sbh_octree8(_np(leaf), _np(node), q, O.KIND_SOFT_VJP, 2, 4, 1e30, eps=0.1, rc=1.3,
                        cot=c)
    assert O.rel_l2(g_ffi, ref) < 1e-5
    g_fd = _np(ffi.softened_stochastic_barnes_hut_force_octree8_vjp(leaf, node, q, c,
                                                                    use_ffi=False, **kw))
    np.testing.assert_allclose(g_fd, g_ffi, rtol=5e-2, atol=5e-2 * np.abs(g_ffi).max())


@pytest.mark.parametrize("softened", [False, True])
def test_stochastic_recompute_vjp_matches_explicit_vjp(gpu, softened):
    ffi = _ffi()
    rng = np.random.default_rng(123)
    n = 300
    p = rng.uniform(-0.7, 0.7, (n, 3)).astype(np.float32)
    m = rng.uniform(0.3, 1.2, n).astype(np.float32)
    c = torch.tensor(rng.normal(size=(n, 3)).astype(np.float32), device="cuda")
    x = torch.tensor(p, device="cuda", requires_grad=True)
    kw = dict(beta=1.3, samples_per_subdomain=2, seed=9)
    leaf, node = ffi.build_octree8_buffers(p, np.zeros_like(p), m, max_depth=6)
    if softened:
        kw.update(softening_scale=0.05, cutoff_scale=1.4)
        f = ffi.softened_stochastic_barnes_hut_force_octree8_recompute_vjp(x, m, max_depth=6, **kw)
        ref_f = ffi.softened_stochastic_barnes_hut_force_octree8(leaf, node, p, **kw)
        ref_g = ffi.softened_stochastic_barnes_hut_force_octree8_vjp(leaf, node, p, c,
                                                                     use_ffi=True, **kw)
    else:
        f = ffi.stochastic_barnes_hut_force_octree8_recompute_vjp(x, m, max_depth=6, **kw)
        ref_f = ffi.stochastic_barnes_hut_force_octree8(leaf, node, p, **kw)
        ref_g = ffi.stochastic_barnes_hut_force_octree8_vjp(leaf, node, p, c, **kw)
    np.testing.assert_array_equal(_np(f), _np(ref_f))
    (f * c).sum().backward()
    np.testing.assert_array_equal(_np(x.grad), _np(ref_g))


def test_stochastic_morton_scheduled_matches_oracle(gpu):
    # 6000 queries: the walk runs in Morton order of the queries; the random
    # numbers are keyed by the query index, so the paths are the oracle's
    ffi = _ffi()
    p, m, leaf, node = _case("plummer", 12000, 6, seed=31)
    rng = np.random.default_rng(32)
    q = p[rng.permutation(p.shape[0])[:6000]].copy()
    lt, nt, qt = _dev(leaf, node, q)
    out = _np(ffi.softened_stochastic_barnes_hut_force_octree8(
        lt, nt, qt, beta=1.0, samples_per_subdomain=2, seed=77, **SOFT))
    ref = O.sbh_octree8(leaf, node, q, O.KIND_SOFT, 2, 77, 1.0, eps=0.03, rc=0.6,
                        prune_enabled=True, prune_r_cut_mult=2.5)
    assert O.rel_l2(out, ref) < 1e-5
