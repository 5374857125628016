"""GPU parity: adaptive Morton octree + group walk + full adjoint
(sbh_gravity_force / sbh_gravity_force_vjp) against the oracle.

- Topology (perm, codes, node ranges, children, parents) and node geometry
  must equal the oracle's adaptive tree bit for bit.
- Interaction counts (far lane-node pairs, near pairs) must be identical,
  i.e. the same MAC decisions; forces and gradients then differ from the
  oracle's FP64 evaluation of the same interaction set only by fp32 rounding
  (tolerance rel_l2 <= 2e-5 forces, <= 1e-4 gradients, stated per test).
- Accuracy against the FP64 direct sum (the metric's bar, rel_l2 <= 1e-3).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _G():
    from sbh_amd import gravity

    return gravity


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def plummer(n, seed):
    rng = np.random.default_rng(seed)
    xmax = 20.0 ** 3 / (20.0 ** 2 + 1.0) ** 1.5
    X = rng.uniform(0, xmax, n)
    r = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (r[:, None] * v).astype(np.float32)


def blobs(n, k, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (k, 3))
    lab = rng.integers(0, k, n)
    return (c[lab] + rng.normal(0, 0.02, (n, 3))).astype(np.float32)


def _case(kind, n, seed):
    rng = np.random.default_rng(seed + 1)
    if kind == "plummer":
        p = plummer(n, seed)
    elif kind == "blobs":
        p = blobs(n, 64, seed)
    elif kind == "uniform":
        p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    elif kind == "same":
        p = np.tile(np.array([[0.25, -0.5, 1.0]], np.float32), (n, 1))
    else:
        raise ValueError(kind)
    m = (rng.lognormal(0, 0.5, n) / n).astype(np.float32)
    c = rng.normal(size=(n, 3)).astype(np.float32)
    return p, m, c


@pytest.mark.parametrize("kind,n,depth,leaf", [
    ("plummer", 20000, 10, 32), ("blobs", 12000, 10, 16), ("uniform", 5000, 6, 8),
    ("same", 300, 10, 32), ("uniform", 1, 10, 32), ("uniform", 40, 10, 32),
])
def test_tree_matches_oracle(gpu, kind, n, depth, leaf):
    G = _G()
    p, m, _ = _case(kind, n, 7)
    ws = G.new_workspace(n, depth, leaf)
    G.gravity_force(p, m, max_depth=depth, leaf_size=leaf, workspace=ws)
    tv = G.read_tree(ws, n, depth, leaf)
    T = O.Tree(p, m, max_depth=depth, leaf_size=leaf)
    assert tv.overflow == 0
    np.testing.assert_array_equal(tv.perm, T.perm)
    np.testing.assert_array_equal(tv.codes, T.codes)
    assert tv.num_nodes == T.num_nodes
    np.testing.assert_array_equal(tv.topo[:, 0], T.start)
    np.testing.assert_array_equal(tv.topo[:, 1], T.count)
    np.testing.assert_array_equal(tv.topo[:, 2], T.first_child)
    np.testing.assert_array_equal(tv.topo[:, 3], T.nchild)
    np.testing.assert_array_equal(tv.parent, T.parent)
    np.testing.assert_array_equal(tv.geo, T.geo)
    np.testing.assert_array_equal(tv.sorted_points[:, :3], T.points_sorted)
    mass, com, quad = T.moments()
    np.testing.assert_allclose(tv.mp[:, 3], mass, rtol=2e-5)
    scale = np.abs(T.points_sorted).max()
    np.testing.assert_allclose(tv.mp[:, :3], com, rtol=0, atol=2e-6 * scale)
    qs = np.abs(quad).max() + 1e-30
    gq = tv.qd[:, [0, 1, 2, 3, 4, 5]]
    np.testing.assert_allclose(gq, quad, rtol=0, atol=2e-4 * qs)


@pytest.mark.parametrize("kind,n", [("plummer", 20000), ("blobs", 16384), ("uniform", 3001)])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_forces_and_adjoint_match_oracle(gpu, kind, n, order):
    # order 3 (second-moment adjoint) is the oracle's mode 2; its forces are
    # the monopole's
    G = _G()
    p, m, c = _case(kind, n, 11)
    eps = 0.005
    stats = G.new_stats()
    gx, gm, f = G.gravity_force_vjp(p, m, c, beta=2.0, softening_scale=eps, max_depth=10,
                                    leaf_size=32, multipole_order=order, stats=stats,
                                    return_forces=True)
    gx, gm, f = _np(gx), _np(gm), _np(f)
    st = G.stats_dict(stats)
    T = O.Tree(p, m, max_depth=10, leaf_size=32)
    r = T.eval(beta=2.0, eps=eps, use_quad={1: 0, 2: 1, 3: 2}[order], cot_sorted=c[T.perm])
    # identical interaction sets
    assert st["far"] == r["far"] and st["near"] == r["near"], (st, r["far"], r["near"])
    assert st["groups"] == r["groups"] and st["overflow"] == 0
    perm = T.perm
    assert O.rel_l2(f[perm], r["force"]) < 2e-5
    assert O.rel_l2(gx[perm], r["gq"] + r["gsx"]) < 1e-4
    assert O.rel_l2(gm[perm], r["gsm"]) < 1e-4
    # forward-only entry point gives the same forces
    f2 = _np(G.gravity_force(p, m, beta=2.0, softening_scale=eps,
                             multipole_order=2 if order == 2 else 1))
    np.testing.assert_allclose(f2, f, rtol=1e-6, atol=1e-6 * np.abs(f).max())


def test_accuracy_vs_fp64_direct(gpu):
    G = _G()
    n = 65536
    p, m, c = _case("plummer", n, 3)
    m = np.full(n, 1.0 / n, np.float32)
    eps = 0.005
    idx = np.random.default_rng(0).choice(n, 2048, replace=False)
    ref = O.direct_force_soft(p, m, p[idx], eps, np.inf)
    for order in (1, 2):
        f = _np(G.gravity_force(p, m, beta=2.0, softening_scale=eps, multipole_order=order))
        err = O.rel_l2(f[idx], ref)
        print(f"order={order} rel_l2={err:.3e} outlier={O.outlier_frac(f[idx], ref):.4f}")
        assert err < 1e-3


def test_gradients_vs_fp64_direct_adjoint(gpu):
    G = _G()
    n = 8192
    p, m, c = _case("blobs", n, 5)
    eps = 0.002
    rgx, rgm = O.direct_full_vjp(p, m, c, eps)
    for order in (1, 2, 3):
        gx, gm = G.gravity_force_vjp(p, m, c, beta=2.0, softening_scale=eps,
                                     multipole_order=order)
        ex, em = O.rel_l2(_np(gx), rgx), O.rel_l2(_np(gm), rgm)
        print(f"order={order} grad_x rel_l2={ex:.3e} grad_m rel_l2={em:.3e}")
        assert ex < 1e-3 and em < (3e-3 if order == 1 else 1e-3)


def test_huge_beta_is_direct_sum(gpu):
    G = _G()
    n = 3000
    p, m, c = _case("plummer", n, 2)
    f = _np(G.gravity_force(p, m, beta=1e30, softening_scale=0.01))
    ref = O.direct_force_soft(p, m, p, 0.01, np.inf)
    assert O.rel_l2(f, ref) < 1e-5
    gx, gm = G.gravity_force_vjp(p, m, c, beta=1e30, softening_scale=0.01)
    rgx, rgm = O.direct_full_vjp(p, m, c, 0.01)
    assert O.rel_l2(_np(gx), rgx) < 1e-5 and O.rel_l2(_np(gm), rgm) < 1e-5


@pytest.mark.parametrize("kind", ["uniform", "plummer"])
def test_full_traversal_of_a_deep_tree(gpu, kind):
    # beta = 1e30 accepts no node, so every group opens every node of a
    # depth-10 tree with leaves of 1 point: the traversal stack's worst case
    # (every pop of 64 nodes opens all of them); it must neither overflow
    # nor change the sum
    G = _G()
    n = 1 << 15
    p, m, c = _case(kind, n, 21)
    stats = G.new_stats()
    gx, gm, f = G.gravity_force_vjp(p, m, c, beta=1e30, softening_scale=0.01, max_depth=10,
                                    leaf_size=1, multipole_order=3, stats=stats,
                                    return_forces=True)
    st = G.stats_dict(stats)
    assert st["overflow"] == 0
    rows = np.random.default_rng(5).choice(n, 512, replace=False)
    ref = O.direct_force_soft(p, m, p[rows], 0.01, np.inf)
    assert O.rel_l2(_np(f)[rows], ref) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129])
def test_small_and_ragged(gpu, n):
    G = _G()
    p, m, c = _case("uniform", n, n)
    f = _np(G.gravity_force(p, m, softening_scale=0.05, leaf_size=4))
    ref = O.direct_force_soft(p, m, p, 0.05, np.inf)
    assert np.isfinite(f).all()
    if n > 1:
        assert O.rel_l2(f, ref) < 1e-3
    rgx, rgm = O.direct_full_vjp(p, m, c, 0.05)
    for order in (2, 3):
        gx, gm = G.gravity_force_vjp(p, m, c, softening_scale=0.05, leaf_size=4,
                                     multipole_order=order)
        if n > 1:
            assert O.rel_l2(_np(gx), rgx) < 2e-3 and O.rel_l2(_np(gm), rgm) < 5e-3
        else:
            assert np.all(_np(gx) == 0) and np.all(_np(gm) == 0)


def test_identical_points_deep_leaf(gpu):
    G = _G()
    n = 500
    p, m, c = _case("same", n, 1)
    p[-1] += np.float32(0.5)
    f = _np(G.gravity_force(p, m, softening_scale=0.01))
    ref = O.direct_force_soft(p, m, p, 0.01, np.inf)
    assert O.rel_l2(f, ref) < 1e-5


def test_recompute_custom_vjp(gpu):
    G = _G()
    n = 4096
    p, m, c = _case("blobs", n, 9)
    x = torch.tensor(p, device="cuda", requires_grad=True)
    mt = torch.tensor(m, device="cuda", requires_grad=True)
    ct = torch.tensor(c, device="cuda")
    kw = dict(softening_scale=0.002, multipole_order=2)
    (G.gravity_force_recompute_vjp(x, mt, **kw) * ct).sum().backward()
    gx, gm = G.gravity_force_vjp(p, m, c, **kw)
    np.testing.assert_allclose(_np(x.grad), _np(gx), rtol=1e-4, atol=1e-4 * np.abs(_np(gx)).max())
    np.testing.assert_allclose(_np(mt.grad), _np(gm), rtol=1e-4, atol=1e-4 * np.abs(_np(gm)).max())


def test_recompute_custom_vjp_with_its_own_adjoint_order(gpu):
    # monopole forward, quadrupole adjoint: the forward equals the order-1
    # forces bit for bit, the gradients the order-2 VJP
    G = _G()
    n = 6000
    p, m, c = _case("plummer", n, 19)
    x = torch.tensor(p, device="cuda", requires_grad=True)
    mt = torch.tensor(m, device="cuda", requires_grad=True)
    ct = torch.tensor(c, device="cuda")
    F = G.gravity_force_recompute_vjp(x, mt, softening_scale=0.005, multipole_order=1,
                                      vjp_multipole_order=2)
    (F * ct).sum().backward()
    np.testing.assert_array_equal(_np(F), _np(G.gravity_force(p, m, softening_scale=0.005,
                                                               multipole_order=1)))
    gx, gm = G.gravity_force_vjp(p, m, c, softening_scale=0.005, multipole_order=2)
    np.testing.assert_allclose(_np(x.grad), _np(gx), rtol=1e-4, atol=1e-4 * np.abs(_np(gx)).max())
    np.testing.assert_allclose(_np(mt.grad), _np(gm), rtol=1e-4, atol=1e-4 * np.abs(_np(gm)).max())
    g1x, _ = G.gravity_force_vjp(p, m, c, softening_scale=0.005, multipole_order=1)
    assert O.rel_l2(_np(x.grad), _np(g1x)) > 1e-6  # the adjoint order did change
    with pytest.raises(ValueError):
        G.gravity_force_recompute_vjp(x, mt, multipole_order=1, vjp_multipole_order=4)
    with pytest.raises(ValueError):
        G.gravity_force_recompute_vjp(x, mt, multipole_order=3)  # not a forward order


def test_determinism_of_forces(gpu):
    G = _G()
    p, m, _ = _case("plummer", 30000, 4)
    a = _np(G.gravity_force(p, m, softening_scale=0.005))
    b = _np(G.gravity_force(p, m, softening_scale=0.005))
    np.testing.assert_array_equal(a, b)


def test_walk_schedule_changes_no_result(gpu):
    # A fresh (zeroed) workspace walks the groups in their own order; the next
    # walks on it are ordered by the costs the previous ones measured.  Every
    # order must walk each group once: same interaction counts, bitwise equal
    # forces, and gradients equal to atomics' rounding.
    G = _G()
    n = 60000
    p, m, c = _case("plummer", n, 21)
    P, M, C = (torch.tensor(a, device="cuda") for a in (p, m, c))
    kw = dict(softening_scale=0.005, max_depth=10, leaf_size=32, multipole_order=1)
    ws = torch.zeros(G.workspace_bytes(n, 10, 32), dtype=torch.uint8, device="cuda")
    runs = []
    for _ in range(3):
        sf, sv = G.new_stats(), G.new_stats()
        f = G.gravity_force(P, M, workspace=ws, stats=sf, **kw).clone()
        gx, gm = (t.clone() for t in G.gravity_force_vjp(P, M, C, workspace=ws, stats=sv,
                                                          **dict(kw, multipole_order=3)))
        runs.append((_np(f), _np(gx), _np(gm), G.stats_dict(sf), G.stats_dict(sv)))
    f0, gx0, gm0, sf0, sv0 = runs[0]
    assert sf0["groups"] > 8 and sf0["far"] > 0 and sf0["near"] > 0
    for f, gx, gm, sf, sv in runs[1:]:
        np.testing.assert_array_equal(f, f0)
        assert sf == sf0 and sv == sv0
        assert O.rel_l2(gx, gx0) < 1e-5 and O.rel_l2(gm, gm0) < 1e-5
    # a workspace full of random bytes (tags, costs) and one whose last walk
    # had another group count: the order is still a permutation of the groups
    junk = torch.randint(0, 256, ws.shape, dtype=torch.uint8, device="cuda")
    sf = G.new_stats()
    np.testing.assert_array_equal(_np(G.gravity_force(P, M, workspace=junk, stats=sf, **kw)), f0)
    assert G.stats_dict(sf) == sf0
    G.gravity_force(P[: n // 3], M[: n // 3], workspace=ws, **kw)
    sf = G.new_stats()
    np.testing.assert_array_equal(_np(G.gravity_force(P, M, workspace=ws, stats=sf, **kw)), f0)
    assert G.stats_dict(sf) == sf0


def test_graph_replay_single_set_matches_eager(gpu):
    # one set through the single-set entry points, captured once and replayed
    # with unrelated eager work in between (the bench's timed step)
    from sbh_amd.graph import Graph

    G = _G()
    n = 50000
    p, m, c = _case("plummer", n, 12)
    P, M, C = (torch.tensor(a, device="cuda") for a in (p, m, c))
    kw = dict(softening_scale=0.005, max_depth=10, leaf_size=32, multipole_order=2)
    ws = G.new_workspace(n, 10, 32)
    Fe = G.gravity_force(P, M, workspace=ws, **kw).clone()
    gxe, gme = (t.clone() for t in G.gravity_force_vjp(P, M, C, workspace=ws, **kw))
    F, gx, gm = torch.empty_like(P), torch.empty_like(P), torch.empty_like(M)

    def step():
        G.gravity_force(P, M, workspace=ws, out=F, **kw)
        G.gravity_force_vjp(P, M, C, workspace=ws, grad_points_out=gx, grad_masses_out=gm,
                            **kw)

    g = Graph()
    g.capture(step)
    for r in range(3):
        F.zero_()
        gx.zero_()
        gm.zero_()
        tmp = torch.randn(1 << 16, device="cuda").sum()
        g.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(tmp)
        assert torch.equal(F, Fe), r
        torch.testing.assert_close(gx, gxe, rtol=1e-4, atol=1e-5 * float(gxe.abs().max()))
        torch.testing.assert_close(gm, gme, rtol=1e-4, atol=1e-5 * float(gme.abs().max()))
    g.close()


def test_capture_without_workspace_is_refused(gpu):
    # A graph must not keep the library's cached per-stream workspace: a later
    # eager call on the same stream handle that needs a larger one frees it.
    # Entry points given no workspace therefore refuse to run under capture
    # (before any launch), and the captured call with a workspace replays the
    # eager result.
    from sbh_amd import ffi, wn
    from sbh_amd._lib import buf, call, stream_handle
    from sbh_amd.graph import Graph

    G = _G()
    p, m, _ = _case("plummer", 5000, 13)
    P, M = (torch.tensor(a, device="cuda") for a in (p, m))
    F = torch.empty_like(P)
    Fe = G.gravity_force(P, M, softening_scale=0.005).clone()  # eager: the cache is fine
    Nrm = torch.nn.functional.normalize(torch.randn_like(P), dim=1)
    out = torch.empty(P.shape[0], device="cuda")
    lo = (P.min(0).values - 1e-3).contiguous()
    hi = (P.max(0).values + 1e-3).contiguous()
    leaf_size, node_size, max_nodes = ffi._get_octree_level3_layout()
    leaf = torch.empty(P.shape[0] * leaf_size, dtype=torch.uint8, device="cuda")
    node = torch.empty(max_nodes * node_size, dtype=torch.uint8, device="cuda")
    contrib = torch.empty(max_nodes * leaf_size, dtype=torch.uint8, device="cuda")
    calls = {
        "gravity": lambda: G.gravity_force(P, M, softening_scale=0.005, out=F),
        "winding_number": lambda: wn.winding_number(P, Nrm, M, P, out=out),
        "level3_builder": lambda: call("sbh_build_octree_level3_buffers", stream_handle(),
                                       buf(P), buf(Nrm), buf(M), buf(lo), buf(hi), buf(leaf),
                                       buf(node), buf(contrib)),
    }
    for name, fn in calls.items():
        g = Graph()
        with pytest.raises(ValueError, match="workspace"):
            g.capture(fn)
        assert not g.captured, name
    ws = G.new_workspace(P.shape[0], 10, 32)
    g = Graph()
    g.capture(lambda: G.gravity_force(P, M, softening_scale=0.005, out=F, workspace=ws))
    F.zero_()
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(F, Fe)
    g.close()


def _direct_adjoint_rows(P, M, C, rows, eps):
    """FP64 direct-sum gradients of L = sum_i c_i . F_i with respect to the
    positions and masses of the points `rows` (closed form of the softened
    kernel, self pair excluded)."""
    e2 = eps * eps + 1e-12
    gx, gm = [], []
    for j in rows.tolist():
        d = P[j][None, :] - P  # x_j - x_k
        r2 = (d * d).sum(1) + e2
        s3 = r2 ** -1.5
        s5 = s3 / r2
        s3[j] = 0.0
        s5[j] = 0.0
        cj = C[j]
        # query side: d/dx_j of c_j . sum_k -m_k s^3 (x_j - x_k)
        q = -(M[:, None] * (s3[:, None] * cj[None, :] -
                            3.0 * (s5 * (d @ cj))[:, None] * d)).sum(0)
        # source side: d/dx_j of sum_i c_i . (-m_j s^3 (x_i - x_j))
        src = M[j] * (s3[:, None] * C - 3.0 * (s5 * (C * d).sum(1))[:, None] * d).sum(0)
        gx.append(q + src)
        gm.append((s3 * (C * d).sum(1)).sum())
    return torch.stack(gx), torch.stack(gm)


def test_maximum_size_against_sampled_direct_sums(gpu):
    # 2^23 points (the path's working size is bounded by HBM only); forces and
    # both gradients checked on sampled points against FP64 direct sums
    G = _G()
    n = 1 << 23
    rng = np.random.default_rng(23)
    p = plummer(n, 23)
    m = np.full(n, 1.0 / n, np.float32)
    c = rng.normal(size=(n, 3)).astype(np.float32)
    P, M, C = (torch.tensor(a, device="cuda") for a in (p, m, c))
    kw = dict(softening_scale=0.005, max_depth=10, leaf_size=32, multipole_order=2)
    F = G.gravity_force(P, M, **kw)
    gx, gm = G.gravity_force_vjp(P, M, C, **kw)
    rows = torch.tensor(rng.choice(n, 96, replace=False), device="cuda")
    Pd, Md, Cd = P.double(), M.double(), C.double()
    e2 = 0.005 ** 2 + 1e-12
    ref = []
    for j in rows.tolist():
        d = Pd[j][None, :] - Pd
        s3 = ((d * d).sum(1) + e2) ** -1.5
        ref.append(-((Md * s3)[:, None] * d).sum(0))
    ref = torch.stack(ref)
    rx, rm = _direct_adjoint_rows(Pd, Md, Cd, rows, 0.005)
    torch.cuda.synchronize()
    err_f = O.rel_l2(F[rows].double().cpu().numpy(), ref.cpu().numpy())
    err_x = O.rel_l2(gx[rows].double().cpu().numpy(), rx.cpu().numpy())
    err_m = O.rel_l2(gm[rows].double().cpu().numpy(), rm.cpu().numpy())
    print(f"n=2^23 force {err_f:.2e} grad_x {err_x:.2e} grad_m {err_m:.2e}")
    assert err_f < 1e-3
    assert err_x < 1e-3 and err_m < 1e-3
    assert torch.isfinite(gx).all() and torch.isfinite(gm).all()
