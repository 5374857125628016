"""GPU parity: direct sums (sbh_direct_force_soft, _vjp, _full_vjp,
sbh_bruteforce_gravity) against the FP64 oracle, plus the reference's own
closed-form tests (python/test_jax_ffi_tree_builder.py:20-123)."""
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


def _case(seed, n, m, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    p = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    ms = rng.uniform(0.1, 1.3, n).astype(np.float32)
    q = rng.uniform(0.7 * lo, 0.7 * hi, (m, 3)).astype(np.float32)
    c = rng.normal(size=(m, 3)).astype(np.float32)
    return p, ms, q, c


def test_reference_closed_form_force(gpu):
    # python/test_jax_ffi_tree_builder.py:20-46 (N=96, M=31, rtol 5e-5, atol 5e-6)
    p, m, q, _ = _case(23, 96, 31)
    out = _ffi().softened_direct_force(p, m, q, softening_scale=5e-2, cutoff_scale=0.9)
    ref = O.direct_force_soft(p, m, q, 5e-2, 0.9)
    np.testing.assert_allclose(_np(out), ref, rtol=5e-5, atol=5e-6)


def test_reference_closed_form_vjp(gpu):
    # python/test_jax_ffi_tree_builder.py:49-86 (rtol 1e-4, atol 1e-5)
    p, m, q, c = _case(31, 64, 17)
    out = _ffi().softened_direct_force_vjp(p, m, q, c, softening_scale=7e-2, cutoff_scale=1.4)
    ref = O.direct_force_soft_vjp(p, m, q, c, 7e-2, 1.4)
    np.testing.assert_allclose(_np(out), ref, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("n,m", [(1, 1), (255, 257), (3000, 1000), (4097, 513)])
@pytest.mark.parametrize("rc,periodic", [(np.inf, False), (0.8, False), (0.8, True)])
def test_direct_force_matches_oracle(gpu, n, m, rc, periodic):
    p, ms, q, c = _case(n + m, n, m, 0.0, 2.0)
    kw = dict(softening_scale=0.02, cutoff_scale=float(rc), periodic=periodic, box_length=2.0)
    f = _np(_ffi().softened_direct_force(p, ms, q, **kw))
    ref = O.direct_force_soft(p, ms, q, 0.02, rc, periodic, 2.0)
    # fp32 accumulation over n terms vs fp64: tolerance stated per element
    scale = np.abs(ref).max() + 1e-30
    np.testing.assert_allclose(f, ref, rtol=2e-4, atol=2e-6 * scale)
    assert O.rel_l2(f, ref) < 5e-6
    g = _np(_ffi().softened_direct_force_vjp(p, ms, q, c, **kw))
    gref = O.direct_force_soft_vjp(p, ms, q, c, 0.02, rc, periodic, 2.0)
    np.testing.assert_allclose(g, gref, rtol=2e-4, atol=2e-6 * (np.abs(gref).max() + 1e-30))
    assert O.rel_l2(g, gref) < 5e-6


@pytest.mark.parametrize("n", [1, 100, 2500])
@pytest.mark.parametrize("rc", [np.inf, 0.6])
def test_direct_full_vjp_matches_oracle(gpu, n, rc):
    p, ms, _, _ = _case(n, n, 1)
    c = np.random.default_rng(n + 1).normal(size=(n, 3)).astype(np.float32)
    gx, gm = _ffi().softened_direct_force_full_vjp(p, ms, c, softening_scale=0.03,
                                                   cutoff_scale=float(rc))
    rgx, rgm = O.direct_full_vjp_cut(p, ms, c, 0.03, rc)
    gx, gm = _np(gx), _np(gm)
    if n == 1:
        assert np.all(gx == 0) and np.all(gm == 0)
        return
    assert O.rel_l2(gx, rgx) < 1e-5
    assert O.rel_l2(gm, rgm) < 1e-5


def test_recompute_vjp_modes(gpu):
    # python/test_jax_ffi_tree_builder.py:89-123 pattern: the custom gradient
    # equals the explicit VJP kernel.
    ffi = _ffi()
    p, ms, _, _ = _case(41, 72, 1)
    c = torch.tensor(np.random.default_rng(43).normal(size=(72, 3)).astype(np.float32),
                     device="cuda")
    x = torch.tensor(p, device="cuda", requires_grad=True)
    mt = torch.tensor(ms, device="cuda", requires_grad=True)
    kw = dict(softening_scale=6e-2, cutoff_scale=1.1)
    (ffi.softened_direct_force_recompute_vjp(x, mt, **kw) * c).sum().backward()
    explicit = ffi.softened_direct_force_vjp(p, ms, p, c, **kw)
    np.testing.assert_allclose(_np(x.grad), _np(explicit), rtol=1e-6, atol=1e-6)
    assert np.all(_np(mt.grad) == 0)
    x.grad = None
    mt.grad = None
    (ffi.softened_direct_force_recompute_vjp(x, mt, vjp_mode="full", **kw) * c).sum().backward()
    gx, gm = ffi.softened_direct_force_full_vjp(p, ms, c, **kw)
    np.testing.assert_allclose(_np(x.grad), _np(gx), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(_np(mt.grad), _np(gm), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("periodic", [False, True])
def test_bruteforce_gravity_matches_oracle(gpu, dim, periodic):
    rng = np.random.default_rng(dim)
    p = rng.uniform(0, 2, (1500, dim)).astype(np.float32)
    m = rng.uniform(0.1, 1.0, 1500).astype(np.float32)
    q = rng.uniform(0, 2, (300, dim)).astype(np.float32)
    out = _np(_ffi().brute_force_gravity(p, m, q, periodic=periodic, box_length=2.0))
    ref = O.bruteforce_gravity(p, m, q, periodic, 2.0)
    np.testing.assert_allclose(out, ref, rtol=2e-5)


def test_empty_inputs(gpu):
    ffi = _ffi()
    p = np.zeros((0, 3), np.float32)
    q = np.ones((5, 3), np.float32)
    out = _np(ffi.softened_direct_force(p, np.zeros(0, np.float32), q))
    assert out.shape == (5, 3) and np.all(out == 0)
    out = _np(ffi.softened_direct_force(q, np.ones(5, np.float32), p))
    assert out.shape == (0, 3)
