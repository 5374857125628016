"""GPU: the benchmark workloads at their own parameters.

- Plain classic BH at N = 10,000 uniform in [0, 1)^3, masses 1/N, softening
  0.01, theta = 0.5 (beta 2 on the cell diagonal), leaf size 16: against the
  oracle's walk of the same tree and against the FP64 direct sum (rel_l2 and
  the reference benchmark's outlier fraction at 0.05,
  python/unified_force_benchmark.py:316-328).
- Differentiated gravity at N = 2^18 in 64 Gaussian blobs (sigma 0.02),
  log-normal masses normalised to 1, softening 0.002, loss sum F . c with
  c ~ N(0, 1): gradients through the recompute custom VJP (autograd) against
  the FP64 direct adjoint on sampled rows, and that adjoint against central
  differences of the FP64 direct loss.
- Batched sets: 64 sets of log-uniform sizes in [1,024, 524,288], each a
  mixture of U{8..64} blobs, softening 0.005, cotangents N(0, 1), through the
  graph-captured forest pipeline: every set's forces and both gradients
  against the FP64 direct sum and adjoint on sampled rows, forces bit-equal
  to the per-set pipeline; a reduced batch (64 sets of 64-4,096 points) set
  by set against the oracle.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS_C5 = 0.005


def _G():
    from sbh_amd import gravity

    return gravity


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _fp64_rows(P, M, C, rows, eps):
    """FP64 direct-sum forces and gradients of L = sum_i c_i . F_i at `rows`
    of one point set, written out independently of the product's helpers:
    F_j = sum_k -m_k s^3 d, d = x_j - x_k, s = (|d|^2 + eps^2 + 1e-12)^-1/2;
    dL/dx_j = -sum_k m_k (s^3 c_j - 3 s^5 (d.c_j) d)
              + m_j sum_k (s^3 c_k - 3 s^5 (c_k.d) d);
    dL/dm_j = sum_k s^3 (c_k . d)."""
    e2 = eps * eps + 1e-12
    P, M, C = P.double(), M.double(), C.double()
    fs, gx, gm = [], [], []
    for j in rows.tolist():
        d = P[j][None, :] - P
        r2 = (d * d).sum(1) + e2
        s3 = r2 ** -1.5
        s5 = s3 / r2
        s3[j] = 0.0
        s5[j] = 0.0
        cj = C[j]
        fs.append(-((M * s3)[:, None] * d).sum(0))
        q = -(M[:, None] * (s3[:, None] * cj[None, :] - 3.0 * (s5 * (d @ cj))[:, None] * d)).sum(0)
        src = M[j] * (s3[:, None] * C - 3.0 * (s5 * (C * d).sum(1))[:, None] * d).sum(0)
        gx.append(q + src)
        gm.append((s3 * (C * d).sum(1)).sum())
    return torch.stack(fs), torch.stack(gx), torch.stack(gm)


def _outlier_frac(pred, ref, thr=0.05):
    pred, ref = np.asarray(pred, np.float64), np.asarray(ref, np.float64)
    rel = np.linalg.norm(pred - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-12)
    return float(np.mean(rel > thr))


# ------------------------------------------------------------------ N = 10,000 uniform
@pytest.mark.parametrize("order", [1, 2])
def test_uniform_cube_n10000_leaf16(gpu, order):
    G = _G()
    n, eps, leaf = 10000, 0.01, 16
    p = np.random.default_rng(1000).uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    c = np.random.default_rng(1002).normal(size=(n, 3)).astype(np.float32)
    st = G.new_stats()
    F = _np(G.gravity_force(p, m, beta=2.0, softening_scale=eps, max_depth=10, leaf_size=leaf,
                            multipole_order=order, stats=st))
    gx, gm = (_np(t) for t in G.gravity_force_vjp(p, m, c, beta=2.0, softening_scale=eps,
                                                  max_depth=10, leaf_size=leaf,
                                                  multipole_order=order))
    s = G.stats_dict(st)
    T = O.Tree(p, m, max_depth=10, leaf_size=leaf)
    r = T.eval(beta=2.0, eps=eps, use_quad=order == 2, cot_sorted=c[T.perm])
    # same interaction set as the oracle, fp32 rounding apart
    assert (s["far"], s["near"], s["groups"], s["overflow"]) == (r["far"], r["near"],
                                                                  r["groups"], 0)
    assert O.rel_l2(F[T.perm], r["force"]) < 2e-5
    assert O.rel_l2(gx[T.perm], r["gq"] + r["gsx"]) < 1e-4
    assert O.rel_l2(gm[T.perm], r["gsm"]) < 1e-4
    # the metric's bar against the FP64 direct sum, all N queries
    ref = O.direct_force_soft(p, m, p, eps, np.inf)
    err = O.rel_l2(F, ref)
    out = _outlier_frac(F, ref)
    print(f"order {order}: force rel_l2 {err:.2e}, outlier_frac {out:.4f}")
    # oracle on the same inputs: rel_l2 2.4e-4 (monopole), 2.8e-5 (quadrupole);
    # no row above 0.05 relative error with either
    assert err < 1e-3
    assert out == 0.0


# ------------------------------------------------------------------ N = 2^18 blobs, custom VJP
def _blob_case(n, seed):
    from sbh_amd import synthetic

    p, m = synthetic.blobs(n, 64, seed)  # log-normal(0, 0.5) masses, normalised to 1
    c = np.random.default_rng(seed + 2).normal(size=(n, 3)).astype(np.float32)
    return p, m, c


def _partial_loss(P, M, C, j, xj, eps):
    """The part of L = sum_i c_i . F_i (FP64 direct sum) that depends on x_j:
    sum_{k != j} K(x_j - x_k) . (m_j c_k - m_k c_j), K(d) = d s^3; every other
    term of L is independent of x_j, so central differences of this equal
    those of L."""
    d = xj[None, :] - P
    r2 = (d * d).sum(1) + (eps * eps + 1e-12)
    s3 = r2 ** -1.5
    s3[j] = 0.0
    w = M[j] * C - M[:, None] * C[j][None, :]
    return float(((s3[:, None] * d) * w).sum())


def test_blobs_n262144_custom_vjp_against_direct_adjoint_and_fd(gpu):
    G = _G()
    n, eps = 1 << 18, 0.002
    p, m, c = _blob_case(n, 3000)
    x = torch.tensor(p, device="cuda", requires_grad=True)
    mt = torch.tensor(m, device="cuda", requires_grad=True)
    C = torch.tensor(c, device="cuda")
    F = G.gravity_force_recompute_vjp(x, mt, beta=2.0, softening_scale=eps, max_depth=10,
                                      leaf_size=32, multipole_order=2)
    (F * C).sum().backward()
    rng = np.random.default_rng(3003)
    rows = torch.tensor(rng.choice(n, 64, replace=False), device="cuda")
    Pd, Md, Cd = (t.detach().double() for t in (x, mt, C))
    rf, rx, rm = _fp64_rows(Pd, Md, Cd, rows, eps)
    ef = O.rel_l2(_np(F[rows]), _np(rf))
    ex = O.rel_l2(_np(x.grad[rows]), _np(rx))
    em = O.rel_l2(_np(mt.grad[rows]), _np(rm))
    print(f"BH vs FP64 direct: force {ef:.2e} grad_x {ex:.2e} grad_m {em:.2e}")
    assert ef < 1e-3 and ex < 1e-3 and em < 1e-3
    # the FP64 direct adjoint against central differences of the FP64 direct
    # loss at 64 random (point, axis) coordinates.  h = 1e-3 is half the
    # softening length, so its truncation error is O(h^2) large; the
    # differences must converge to the adjoint at second order.
    js = rng.choice(n, 64, replace=False)
    axes = rng.integers(0, 3, 64)
    _, ax, am = _fp64_rows(Pd, Md, Cd, torch.tensor(js, device="cuda"), eps)
    ax = _np(ax)[np.arange(64), axes]
    errs = {}
    for h in (1e-3, 1e-4, 1e-5):
        fd = []
        for j, a in zip(js.tolist(), axes.tolist()):
            xp, xm = Pd[j].clone(), Pd[j].clone()
            xp[a] += h
            xm[a] -= h
            fd.append((_partial_loss(Pd, Md, Cd, j, xp, eps) -
                       _partial_loss(Pd, Md, Cd, j, xm, eps)) / (2 * h))
        errs[h] = O.rel_l2(np.array(fd), ax)
    print("central differences vs direct adjoint:", errs)
    assert errs[1e-3] / errs[1e-4] > 30 and errs[1e-4] / errs[1e-5] > 30
    assert errs[1e-5] < 1e-4
    # masses: L is linear in m_j, so the central difference is exact
    fdm = []
    for j in js[:16].tolist():
        h = 1e-3 * float(Md[j])
        Mp, Mm = Md.clone(), Md.clone()
        Mp[j] += h
        Mm[j] -= h
        fdm.append((_partial_loss(Pd, Mp, Cd, j, Pd[j], eps) -
                    _partial_loss(Pd, Mm, Cd, j, Pd[j], eps)) / (2 * h))
    assert O.rel_l2(np.array(fdm), _np(am)[:16]) < 1e-9


# ------------------------------------------------------------------ 64 batched sets
def _blob_sets(num_sets, seed, lo, hi):
    from sbh_amd import synthetic

    return synthetic.blob_sets(num_sets, seed, lo, hi)


def test_batched_64_sets_against_fp64_and_per_set(gpu):
    from sbh_amd.graph import Graph

    G = _G()
    sizes, ps, ms, cs = _blob_sets(64, 5000, 1024, 524288)
    assert sizes.min() >= 1024 and sizes.max() <= 524288
    P = torch.tensor(np.concatenate(ps), device="cuda")
    M = torch.tensor(np.concatenate(ms), device="cuda")
    C = torch.tensor(np.concatenate(cs), device="cuda")
    offs = G.set_offsets(sizes)
    kw = dict(beta=2.0, softening_scale=EPS_C5, max_depth=10, leaf_size=32, multipole_order=2)
    ws = G.new_batch_workspace(P.shape[0], 64, 10, 32)
    F, gx, gm = torch.empty_like(P), torch.empty_like(P), torch.empty_like(M)
    st = G.new_stats()

    def step():
        G.gravity_force_batch(P, M, offs, workspace=ws, out=F, check_offsets=False, **kw)
        G.gravity_force_vjp_batch(P, M, C, offs, workspace=ws, grad_points_out=gx,
                                  grad_masses_out=gm, stats=st, check_offsets=False, **kw)

    step()
    g = Graph()
    g.capture(step)
    F.zero_()
    gx.zero_()
    gm.zero_()
    g.replay()
    torch.cuda.synchronize()
    g.close()
    assert G.stats_dict(st)["overflow"] == 0
    o = np.concatenate([[0], np.cumsum(sizes)])
    rng = np.random.default_rng(5999)
    wsd = G.new_workspace(int(sizes.max()), 10, 32)
    keys = ("force", "grad_points", "grad_masses")
    per_set = {key: [] for key in keys}
    pooled = {key: ([], []) for key in keys}
    for k in range(64):
        sl = slice(int(o[k]), int(o[k + 1]))
        Pk, Mk, Ck = P[sl], M[sl], C[sl]
        rows = torch.tensor(rng.choice(int(sizes[k]), 8, replace=False), device="cuda")
        ref = dict(zip(keys, (_np(t) for t in _fp64_rows(Pk, Mk, Ck, rows, EPS_C5))))
        got = {"force": _np(F[sl][rows]), "grad_points": _np(gx[sl][rows]),
               "grad_masses": _np(gm[sl][rows])}
        for key in keys:
            per_set[key].append(O.rel_l2(got[key], ref[key]))
            pooled[key][0].append(got[key])
            pooled[key][1].append(ref[key])
        fk = G.gravity_force(Pk, Mk, workspace=wsd, **kw)
        assert torch.equal(F[sl], fk), f"set {k}: batch forces differ from the per-set call"
    pool = {key: O.rel_l2(np.concatenate(a), np.concatenate(b)) for key, (a, b) in pooled.items()}
    worst = {key: (int(np.argmax(v)), float(np.max(v))) for key, v in per_set.items()}
    print(f"64 sets, {int(sizes.sum())} points: rel_l2 vs FP64 over all 512 sampled rows "
          f"{pool}; worst set (index, rel_l2) {worst}")
    # Forces and position gradients: the metric's bar (1e-3) in every set.
    for key in ("force", "grad_points"):
        k, v = worst[key]
        assert v < 1e-3, f"set {k} ({int(sizes[k])} points): {key} rel_l2 {v:.2e}"
    # Mass gradients dL/dm_j = sum_k s^3 (c_k . d_jk) sum random-sign terms
    # (c ~ N(0, 1)), so the total is small against its terms and the same
    # truncation error of the theta = 0.5 walk is a larger fraction of it
    # (the oracle's walk of the same tree gives the same error): 1e-3 over
    # all sampled rows, 3e-3 in any one set of 8 rows.
    assert pool["grad_masses"] < 1e-3, pool
    k, v = worst["grad_masses"]
    assert v < 3e-3, f"set {k} ({int(sizes[k])} points): grad_masses rel_l2 {v:.2e}"


def test_batched_reduced_sets_against_oracle(gpu):
    G = _G()
    sizes, ps, ms, cs = _blob_sets(64, 5100, 64, 4096)
    P = torch.tensor(np.concatenate(ps), device="cuda")
    M = torch.tensor(np.concatenate(ms), device="cuda")
    C = torch.tensor(np.concatenate(cs), device="cuda")
    offs = G.set_offsets(sizes)
    kw = dict(beta=2.0, softening_scale=EPS_C5, max_depth=10, leaf_size=16, multipole_order=2)
    st = G.new_stats()
    F = torch.empty_like(P)
    gx, gm = G.gravity_force_vjp_batch(P, M, C, offs, forces_out=F, stats=st, **kw)
    F, gx, gm = _np(F), _np(gx), _np(gm)
    o = np.concatenate([[0], np.cumsum(sizes)])
    far = near = groups = 0
    for k in range(64):
        sl = slice(int(o[k]), int(o[k + 1]))
        T = O.Tree(ps[k], ms[k], max_depth=10, leaf_size=16)
        r = T.eval(beta=2.0, eps=EPS_C5, use_quad=True, cot_sorted=cs[k][T.perm])
        far += r["far"]
        near += r["near"]
        groups += r["groups"]
        assert O.rel_l2(F[sl][T.perm], r["force"]) < 2e-5, k
        assert O.rel_l2(gx[sl][T.perm], r["gq"] + r["gsx"]) < 1e-4, k
        assert O.rel_l2(gm[sl][T.perm], r["gsm"]) < 1e-4, k
    s = G.stats_dict(st)
    assert (s["far"], s["near"], s["groups"], s["overflow"]) == (far, near, groups, 0)
    assert math.isclose(float(sizes.sum()), float(P.shape[0]))
