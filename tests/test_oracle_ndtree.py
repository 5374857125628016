"""Pins of the oracle's CPU-layout trees (Tree2D<1,3>, Tree3D<1,3>) and of its
walks over them (oracle/sbh_ndtree.c).

Builder: an independent numpy breadth-first partition (fp32 arithmetic in the
reference's operation order, include/tree2d.h:237-321, include/aabb.h:36-42)
must give the same nodes, boxes and point order bit for bit; moments are
checked against fp64 mass-weighted means.  Vote walk
(include/barnes_hut_gpu.h:19-72): beta -> inf is the exact direct sum (fp64
closed forms of include/kernel.h:55-60, 89-94 and src/sbh_ffi.cu:142-160),
beta = 0 the root monopole.  Sampler (include/tree_sampler_gpu.h:25-140): a
tree whose subdomains are leaves gives the exact sum, and the mean over
independent blocks (each block draws its own sample points and paths)
matches the direct sum within its standard error.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

from oracle import oracle as O


def _np_build(p, lo, hi, level=3):
    """Breadth-first reference build in numpy fp32 (stable partition)."""
    n, d = p.shape
    S = 1 << level
    order = np.arange(n)
    nodes = [dict(start=0, end=n, parent=0, depth=0, lo=lo.astype(np.float32),
                  hi=hi.astype(np.float32))]
    i = 0
    while i < len(nodes):
        nd = nodes[i]
        cnt = nd["end"] - nd["start"]
        nd["leaf"] = nd["depth"] >= 100 or cnt <= 1
        nd["children"] = []
        if not nd["leaf"]:
            idx = order[nd["start"]:nd["end"]]
            with np.errstate(invalid="ignore", divide="ignore"):
                # zero-width boxes at depth ~50+ give NaN: the x86 conversion
                # yields INT_MIN, clipped to cell 0 (as the reference)
                t = (np.float32(S) * (p[idx] - nd["lo"])) / (nd["hi"] - nd["lo"])
                q = np.clip(np.trunc(t).astype(np.int64), 0, S - 1)
            cell = np.zeros(len(idx), np.int64)
            for a in reversed(range(d)):
                cell = cell * S + q[:, a]
            s = np.argsort(cell, kind="stable")
            order[nd["start"]:nd["end"]] = idx[s]
            cs = cell[s]
            inc = (nd["hi"] - nd["lo"]) / np.float32(S)
            for c in np.unique(cs):
                k = np.array([(c // S ** a) % S for a in range(d)], np.float32)
                clo = (nd["lo"] + k * inc).astype(np.float32)
                first = np.searchsorted(cs, c, "left")
                last = np.searchsorted(cs, c, "right")
                nd["children"].append((int(c), len(nodes)))
                nodes.append(dict(start=nd["start"] + int(first), end=nd["start"] + int(last),
                                  parent=i, depth=nd["depth"] + 1, lo=clo,
                                  hi=(clo + inc).astype(np.float32)))
        i += 1
    return nodes, order


def _decode(node_bytes, dim, level=3):
    return O.decode_nodes(node_bytes, dim, level)


def _leaves(leaf_bytes, dim):
    dt = O.LEAF2_DTYPE if dim == 2 else O.LEAF3_DTYPE
    return np.frombuffer(np.ascontiguousarray(leaf_bytes).tobytes(), dt)


def _data(n, dim, seed, kind="uniform"):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (n, dim))
    if kind == "clump":
        p[: n // 2] = 0.3 + rng.normal(0, 2e-3, (n // 2, dim))
    if kind == "dups":
        p[: n // 4] = p[0]
    p = p.astype(np.float32)
    nv = rng.normal(size=(n, dim)).astype(np.float32)
    m = rng.uniform(0.2, 1.0, n).astype(np.float32)
    lo = p.min(0) - np.float32(1e-3) if n else np.zeros(dim, np.float32)
    hi = p.max(0) + np.float32(1e-3) if n else np.ones(dim, np.float32)
    return p, nv, m, lo.astype(np.float32), hi.astype(np.float32)


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n,seed,kind", [(400, 1, "uniform"), (300, 2, "clump"),
                                         (120, 3, "dups"), (2, 4, "uniform")])
def test_ndtree_build_matches_numpy_partition(dim, n, seed, kind, level):
    p, nv, m, lo, hi = _data(n, dim, seed, kind)
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi, alpha=0.75, level=level)
    N = _decode(node, dim, level)
    ref, order = _np_build(p, lo, hi, level)
    # layout sizes of Tree2DNode<1, L> / Tree3DNode<1, L>
    assert N.dtype.itemsize == {(2, 1): 136, (2, 2): 232, (2, 3): 616, (3, 1): 184,
                                (3, 2): 632, (3, 3): 4272}[(dim, level)]
    assert len(N) == len(ref)
    for i, r in enumerate(ref):
        assert (N[i]["start"], N[i]["end"], N[i]["parent_idx"], N[i]["depth"]) == \
            (r["start"], r["end"], r["parent"], r["depth"])
        assert N[i]["buffer_idx"] == i and N[i]["max_depth"] == 100
        assert bool(N[i]["is_leaf"]) == r["leaf"]
        np.testing.assert_array_equal(N[i]["bmin"], r["lo"])
        np.testing.assert_array_equal(N[i]["bmax"], r["hi"])
        assert N[i]["num_children"] == len(r["children"])
        cells = 1 << (dim * level)
        expect = np.zeros(cells, np.uint64)
        for c, ch in r["children"]:
            expect[c] = ch
        np.testing.assert_array_equal(N[i]["child"], expect)
        mask = np.zeros(max(1, cells // 64), np.uint64)
        for c, _ in r["children"]:
            mask[c >> 6] |= np.uint64(1) << np.uint64(c & 63)
        np.testing.assert_array_equal(N[i]["bitmask"], mask)
    L = _leaves(leaf, dim)
    np.testing.assert_array_equal(L["pos"], p[order])
    np.testing.assert_array_equal(L["normal"], nv[order])
    np.testing.assert_array_equal(L["mass"], m[order])
    assert (L["alpha"] == np.float32(0.75)).all()
    # contrib bytes are the node records' contrib fields
    C = np.frombuffer(contrib.tobytes(), O.LEAF2_DTYPE if dim == 2 else O.LEAF3_DTYPE)
    assert C.tobytes() == N["contrib"].tobytes()


@pytest.mark.parametrize("dim", [2, 3])
def test_ndtree_moments_are_mass_weighted_means(dim):
    p, nv, m, lo, hi = _data(500, dim, 7, "clump")
    leaf, node, _ = O.nd_build(p, nv, m, lo, hi)
    N = _decode(node, dim)
    L = _leaves(leaf, dim)
    for nd in N:
        s, e = int(nd["start"]), int(nd["end"])
        mm = L["mass"][s:e].astype(np.float64)
        M = mm.sum()
        com = (mm[:, None] * L["pos"][s:e]).sum(0) / M
        nrm = (mm[:, None] * L["normal"][s:e]).sum(0) / M
        assert abs(nd["contrib"]["mass"] - M) <= 2e-6 * M * max(1, e - s)
        np.testing.assert_allclose(nd["contrib"]["pos"], com, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(nd["contrib"]["normal"], nrm, rtol=1e-4, atol=1e-5)
        assert nd["contrib"]["alpha"] == np.float32(1.0)
        # every point lies in the node's cell
        if nd["depth"] > 0:
            assert (L["pos"][s:e] >= nd["bmin"] - 1e-6).all()
            assert (L["pos"][s:e] <= nd["bmax"] + 1e-6).all()


def test_ndtree_edge_cases():
    # no points: one root leaf whose moments are 0 / 0
    p, nv, m, lo, hi = _data(0, 2, 0)
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi)
    N = O.decode_nodes2(node)
    assert len(N) == 1 and N[0]["is_leaf"] == 1 and leaf.size == 0
    assert np.isnan(N[0]["contrib"]["pos"]).all() and N[0]["contrib"]["alpha"] == 0
    # coincident points: a chain down to depth 100, one leaf holding them
    p = np.repeat(np.array([[0.25, 0.5]], np.float32), 5, 0)
    p = np.concatenate([p, np.array([[-0.5, -0.5]], np.float32)])
    m = np.ones(6, np.float32)
    lo, hi = np.array([-1, -1], np.float32), np.array([1, 1], np.float32)
    leaf, node, _ = O.nd_build(p, np.zeros_like(p), m, lo, hi)
    N = O.decode_nodes2(node)
    deep = N[N["depth"] == 100]
    assert len(deep) == 1 and deep[0]["is_leaf"] and deep[0]["end"] - deep[0]["start"] == 5
    assert N["depth"].max() == 100


def _direct2(p, m, q, kind, eps=0.0, rc=0.0):
    p = p.astype(np.float64)
    q = q.astype(np.float64)
    m = m.astype(np.float64)
    if kind == O.L3_POTENTIAL:
        r = np.linalg.norm(p[None] - q[:, None], axis=2)
        return (-m[None] / (r + 0.1)).sum(1)
    d = q[:, None] - p[None]
    r2 = (d * d).sum(2)
    if kind == O.L3_FORCE:
        c = -m[None] / (np.sqrt(r2) + 0.1) ** 3
    else:
        s = 1.0 / np.sqrt(r2 + eps * eps + 1e-12)
        att = 1.0
        if rc > 0 and math.isfinite(rc):
            u = np.sqrt(r2 + 1e-12) / (2 * rc)
            att = erfc(u) + 2 / math.sqrt(math.pi) * u * np.exp(-u * u)
        c = -m[None] * att * s ** 3
    return (c[:, :, None] * d).sum(1)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [O.L3_POTENTIAL, O.L3_FORCE, O.L3_SOFT])
def test_vote_walk_limits(dim, kind):
    p, nv, m, lo, hi = _data(400, dim, 11, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi)
    rng = np.random.default_rng(5)
    q = rng.uniform(-1.2, 1.2, (70, dim)).astype(np.float32)
    eps, rc = 0.03, 0.7
    exact = O.nd_eval(leaf, node, contrib, q, kind, beta=1e30, threads=64, eps=eps, rc=rc)
    if dim == 2:
        ref = _direct2(p, m, q, kind, eps, rc)
    elif kind == O.L3_POTENTIAL:
        ref = -(m[None] / (np.linalg.norm(p[None].astype(np.float64) - q[:, None], axis=2)
                           + 0.1)).sum(1)
    elif kind == O.L3_SOFT:
        ref = O.direct_force_soft(p, m, q, eps, rc)
    else:
        d = q[:, None].astype(np.float64) - p[None]
        ref = (-m[None, :, None] * d / (np.linalg.norm(d, axis=2)[..., None] + 0.1) ** 3).sum(1)
    np.testing.assert_allclose(exact, ref, rtol=1e-9, atol=1e-12)
    # beta = 0: the root is a far field for every query
    mono = O.nd_eval(leaf, node, contrib, q, kind, beta=0.0, threads=64, eps=eps, rc=rc)
    root = _decode(node, dim)[0]["contrib"]
    rp = np.asarray(root["pos"], np.float32)[None]
    rm = np.asarray([root["mass"]], np.float32)
    if dim == 2:
        ref0 = _direct2(rp, rm, q, kind, eps, rc)
    elif kind == O.L3_SOFT:
        ref0 = O.direct_force_soft(rp, rm, q, eps, rc)
    elif kind == O.L3_POTENTIAL:
        ref0 = -rm[0] / (np.linalg.norm(q.astype(np.float64) - rp, axis=1) + 0.1)
    else:
        d = q.astype(np.float64) - rp
        ref0 = -rm[0] * d / (np.linalg.norm(d, axis=1)[:, None] + 0.1) ** 3
    np.testing.assert_allclose(mono, ref0, rtol=1e-9, atol=1e-12)


def test_vote_walk_error_order_2d():
    p, nv, m, lo, hi = _data(2000, 2, 13, "uniform")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi)
    q = p[:200]
    ref = _direct2(p, m, q, O.L3_FORCE)
    errs = [O.rel_l2(O.nd_eval(leaf, node, contrib, q, O.L3_FORCE, beta=b, threads=32), ref)
            for b in (1.0, 2.0, 4.0)]
    assert errs[0] > errs[1] > errs[2] and errs[1] < 1e-3


def test_sampler_on_one_level_tree_is_exact():
    # one point per cell of the 8 x 8 split: every subdomain is a leaf
    g = (np.arange(8, dtype=np.float32) + 0.5) / 8
    p = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    m = np.linspace(0.5, 1.5, 64).astype(np.float32)
    leaf, node, contrib = O.nd_build(p, np.zeros_like(p), m, np.zeros(2, np.float32),
                                     np.ones(2, np.float32))
    N = O.decode_nodes2(node)
    assert len(N) == 65 and N[1:]["is_leaf"].all()
    q = np.random.default_rng(0).uniform(-0.5, 1.5, (40, 2)).astype(np.float32)
    for seed in (0, 9):
        out = O.nd_eval(leaf, node, contrib, q, O.L3_FORCE, samples=3, seed=seed, threads=16)
        np.testing.assert_allclose(out, _direct2(p, m, q, O.L3_FORCE), rtol=1e-9)
    nan = O.nd_eval(leaf, node, contrib, q, O.L3_POTENTIAL, samples=0, threads=16)
    assert np.isnan(nan).all()


@pytest.mark.parametrize("dim,kind", [(2, O.L3_POTENTIAL), (2, O.L3_FORCE), (3, O.L3_FORCE)])
def test_sampler_is_unbiased_over_blocks(dim, kind):
    p, nv, m, lo, hi = _data(600, dim, 17, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi)
    q0 = np.full((1, dim), 0.1, np.float32)
    K = 4000
    q = np.repeat(q0, K, 0)
    # threads = 1: every query is its own block, hence its own sample points
    est = O.nd_eval(leaf, node, contrib, q, kind, samples=1, seed=5, threads=1)
    exact = (_direct2(p, m, q0, kind) if dim == 2 else
             (-m[None, :, None] * (q0[:, None] - p[None]).astype(np.float64) /
              (np.linalg.norm((q0[:, None] - p[None]).astype(np.float64), axis=2)[..., None]
               + 0.1) ** 3).sum(1))
    est = est.reshape(K, -1)
    mean = est.mean(0)
    se = est.std(0, ddof=1) / math.sqrt(K)
    assert (se > 0).all()  # the paths are random
    z = np.abs(mean - np.asarray(exact).reshape(-1)) / se
    assert (z < 4.5).all(), (mean, exact, se)


def test_sampler_determinism_and_seed():
    p, nv, m, lo, hi = _data(500, 2, 19, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi)
    q = p[:64]
    a = O.nd_eval(leaf, node, contrib, q, O.L3_FORCE, samples=2, seed=3, threads=32)
    b = O.nd_eval(leaf, node, contrib, q, O.L3_FORCE, samples=2, seed=3, threads=32)
    c = O.nd_eval(leaf, node, contrib, q, O.L3_FORCE, samples=2, seed=4, threads=32)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


def _direct_scalar(p, nv, m, q, kind, alpha=1.0):
    p, nv, q, m = (x.astype(np.float64) for x in (p, nv, q, m))
    d = p[None] - q[:, None]
    r2 = (d * d).sum(2)
    if kind == O.L3_POTENTIAL:
        return (-m[None] / (np.sqrt(r2) + 0.1)).sum(1)
    if kind == O.ND_WN:
        dn = (nv[None] * d).sum(2)
        if p.shape[1] == 2:
            return (m[None] * dn / (2 * np.pi * (r2 + 1e-5))).sum(1)
        return (m[None] * dn / (4 * np.pi * (r2 * np.sqrt(r2) + 1e-5))).sum(1)
    return np.exp(-alpha * np.sqrt(r2)).sum(1)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("level", [1, 3])
@pytest.mark.parametrize("kind", [O.L3_POTENTIAL, O.ND_WN, O.ND_SMOOTH])
def test_module_kinds_limits_and_brute_force(dim, level, kind):
    p, nv, m, lo, hi = _data(300, dim, 23, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi, alpha=0.6, level=level)
    q = np.random.default_rng(9).uniform(-1.2, 1.2, (50, dim)).astype(np.float32)
    ref = _direct_scalar(p, nv, m, q, kind, alpha=float(np.float32(0.6)))  # the records' fp32 alpha
    exact = O.nd_eval(leaf, node, contrib, q, kind, beta=1e30, threads=32, level=level)
    np.testing.assert_allclose(exact, ref, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(O.nd_brute(leaf, q, kind), ref, rtol=1e-9, atol=1e-12)
    # beta = 0: the root record alone
    N = _decode(node, dim, level)
    r = N[0]["contrib"]
    mono = _direct_scalar(np.asarray(r["pos"], np.float32)[None],
                          np.asarray(r["normal"], np.float32)[None],
                          np.asarray([r["mass"]], np.float32), q, kind, alpha=float(r["alpha"]))
    np.testing.assert_allclose(O.nd_eval(leaf, node, contrib, q, kind, beta=0.0, threads=32,
                                         level=level), mono, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("level", [1, 2])
def test_sampler_unbiased_low_levels(level):
    p, nv, m, lo, hi = _data(500, 2, 29, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi, level=level)
    K = 4000
    q0 = np.full((1, 2), -0.2, np.float32)
    est = O.nd_eval(leaf, node, contrib, np.repeat(q0, K, 0), O.L3_POTENTIAL, samples=1,
                    seed=2, threads=1, level=level)
    exact = _direct_scalar(p, nv, m, q0, O.L3_POTENTIAL)[0]
    se = est.std(ddof=1) / math.sqrt(K)
    assert se > 0 and abs(est.mean() - exact) < 4.5 * se, (est.mean(), exact, se)


def test_module_seeding_is_ffi_seeding_at_minus_query():
    # the pybind11 module passes seed 0 to every query (src/sbh_gpu.cu:143-164);
    # the FFI targets seed + query: the two agree at seed = -query
    p, nv, m, lo, hi = _data(400, 3, 31, "clump")
    leaf, node, contrib = O.nd_build(p, nv, m, lo, hi, level=1)
    q = p[:96]
    mod = O.nd_eval(leaf, node, contrib, q, O.L3_POTENTIAL, samples=2, threads=32, level=1,
                    seed=0, seed_per_query=False)
    for i in (0, 5, 40, 95):
        ffi = O.nd_eval(leaf, node, contrib, q, O.L3_POTENTIAL, samples=2, threads=32, level=1,
                        seed=-i, seed_per_query=True)
        assert mod[i] == ffi[i]
    assert not np.all(mod == O.nd_eval(leaf, node, contrib, q, O.L3_POTENTIAL, samples=2,
                                       threads=32, level=1, seed=0, seed_per_query=True))
