"""GPU: batched self-gravity (one forest pipeline for many independent sets)
and HIP graph replay.

The single-set pipeline is pinned against the oracle elsewhere; here every
set of a batch must reproduce it: forces bit for bit (same cube, order,
topology and interaction sequence), gradients to fp32 atomic-order noise,
and the interaction counts must add up.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _G():
    from sbh_amd import gravity

    return gravity


def _sets(sizes, seed):
    from sbh_amd import synthetic

    rng = np.random.default_rng(seed)
    ps, ms, cs = [], [], []
    for k, n in enumerate(sizes):
        if n == 0:
            ps.append(np.zeros((0, 3), np.float32))
            ms.append(np.zeros(0, np.float32))
            cs.append(np.zeros((0, 3), np.float32))
            continue
        p, m = synthetic.blobs(n, int(rng.integers(8, 65)), seed * 1000 + k)
        ps.append(p)
        ms.append(m)
        cs.append(rng.normal(size=(n, 3)).astype(np.float32))
    return ps, ms, cs


def _dev(*a):
    return [torch.tensor(np.concatenate(x), device="cuda") for x in a]


KW = dict(beta=2.0, softening_scale=0.005, max_depth=10, leaf_size=16, multipole_order=2)


@pytest.mark.parametrize("sizes", [[3000, 1, 0, 700, 12000, 64, 65],
                                   [5000], [0, 0, 2, 0]])
def test_batch_matches_per_set_calls(gpu, sizes):
    G = _G()
    ps, ms, cs = _sets(sizes, 5)
    P, M, C = _dev(ps, ms, cs)
    offs = G.set_offsets(sizes)
    sb = G.new_stats()
    F = G.gravity_force_batch(P, M, offs, stats=sb, **KW)
    sv = G.new_stats()
    Fv = torch.empty_like(P)
    gx, gm = G.gravity_force_vjp_batch(P, M, C, offs, stats=sv, forces_out=Fv, **KW)
    torch.cuda.synchronize()
    F, Fv, gx, gm = (t.cpu().numpy() for t in (F, Fv, gx, gm))
    np.testing.assert_array_equal(F, Fv)
    tot = np.zeros(4, np.int64)
    o = np.concatenate([[0], np.cumsum(sizes)])
    for k, n in enumerate(sizes):
        if n == 0:
            continue
        st = G.new_stats()
        f = G.gravity_force(ps[k], ms[k], stats=st, **KW)
        g1, g2 = G.gravity_force_vjp(ps[k], ms[k], cs[k], **KW)
        torch.cuda.synchronize()
        sl = slice(o[k], o[k + 1])
        np.testing.assert_array_equal(F[sl], f.cpu().numpy(), err_msg=f"set {k}")
        ref_gx, ref_gm = g1.cpu().numpy(), g2.cpu().numpy()
        scale = max(np.abs(ref_gx).max(), 1e-30)
        np.testing.assert_allclose(gx[sl], ref_gx, rtol=1e-4, atol=1e-5 * scale)
        scale = max(np.abs(ref_gm).max(), 1e-30)
        np.testing.assert_allclose(gm[sl], ref_gm, rtol=1e-4, atol=1e-5 * scale)
        s = st.cpu().numpy()
        assert s[5] == 0, f"set {k}: overflow or truncated walk"
        tot += s[[0, 1, 2, 3]]
    s = sb.cpu().numpy()
    nonempty = sum(1 for n in sizes if n > 0)
    assert s[5] == 0
    np.testing.assert_array_equal(s[:3], tot[:3])
    # nodes: the per-set trees plus one empty root per empty set
    assert s[3] == tot[3] + (len(sizes) - nonempty)
    np.testing.assert_array_equal(sv.cpu().numpy()[:4], s[:4])
    assert sv.cpu().numpy()[5] == 0


def test_batch_rejects_bad_offsets(gpu):
    G = _G()
    P = torch.zeros((10, 3), device="cuda")
    M = torch.ones(10, device="cuda")
    bad = torch.tensor([0, 6, 4, 10], dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        G.gravity_force_batch(P, M, bad, **KW)
    with pytest.raises(ValueError):
        G.gravity_force_batch(P, M, torch.tensor([0, 9], dtype=torch.int32, device="cuda"), **KW)
    # unchecked bad offsets are caught on the device: every set left empty
    st = G.new_stats()
    G.gravity_force_batch(P, M, bad, stats=st, check_offsets=False, **KW)
    assert int(st.cpu()[5]) == 2
    assert int(st.cpu()[1]) == 0


def test_graph_replay_matches_eager(gpu):
    from sbh_amd.graph import Graph

    G = _G()
    sizes = [4000, 900, 17000, 1, 2500]
    ps, ms, cs = _sets(sizes, 9)
    P, M, C = _dev(ps, ms, cs)
    offs = G.set_offsets(sizes)
    ws = G.new_batch_workspace(P.shape[0], len(sizes), KW["max_depth"], KW["leaf_size"])
    F = torch.empty_like(P)
    gx = torch.empty_like(P)
    gm = torch.empty_like(M)
    Fe = G.gravity_force_batch(P, M, offs, workspace=ws, **KW).clone()
    gxe, gme = (t.clone() for t in G.gravity_force_vjp_batch(P, M, C, offs, workspace=ws, **KW))

    def step():
        G.gravity_force_batch(P, M, offs, workspace=ws, out=F, check_offsets=False, **KW)
        G.gravity_force_vjp_batch(P, M, C, offs, workspace=ws, grad_points_out=gx,
                                  grad_masses_out=gm, check_offsets=False, **KW)

    g = Graph()
    g.capture(step)
    for r in range(3):
        F.zero_()
        gx.zero_()
        gm.zero_()
        # unrelated eager work between replays
        tmp = torch.randn(1 << 16, device="cuda").sum()
        g.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(tmp)
        assert torch.equal(F, Fe), r
        torch.testing.assert_close(gx, gxe, rtol=1e-4, atol=1e-5 * float(gxe.abs().max()))
        torch.testing.assert_close(gm, gme, rtol=1e-4, atol=1e-5 * float(gme.abs().max()))
    g.close()


@pytest.mark.parametrize("threads", [64, 128, 512])
def test_walk_block_sizes_give_the_same_results(gpu, threads):
    # every query group is walked by one wave whatever the block size, so
    # forces and statistics are bit-equal to the default 256-thread walk
    G = _G()
    sizes = [9000, 0, 2500, 17, 30000]
    ps, ms, cs = _sets(sizes, 9)
    P, M, C = _dev(ps, ms, cs)
    offs = G.set_offsets(sizes)
    sa, sb = G.new_stats(), G.new_stats()
    Fa = G.gravity_force_batch(P, M, offs, stats=sa, **KW)
    Fb = G.gravity_force_batch(P, M, offs, stats=sb, walk_threads=threads, walk_blocks=97, **KW)
    ga, gma = G.gravity_force_vjp_batch(P, M, C, offs, **KW)
    gb, gmb = G.gravity_force_vjp_batch(P, M, C, offs, walk_threads=threads, **KW)
    torch.cuda.synchronize()
    assert torch.equal(Fa, Fb)
    assert torch.equal(sa, sb) and int(sa[5]) == 0
    assert float(torch.linalg.norm(ga - gb) / torch.linalg.norm(ga)) < 1e-5
    assert float(torch.linalg.norm(gma - gmb) / torch.linalg.norm(gma)) < 1e-5
    with pytest.raises(ValueError, match="walk_threads"):
        G.gravity_force_batch(P, M, offs, walk_threads=96, **KW)


def test_autotuner_gate_and_search(gpu):
    # a small batch: every candidate of the coordinate search is timed once,
    # the winner passes the FP64 gate and is the fastest that does, the
    # largest sets are checked alone, and the result is cached
    from sbh_amd import autotune

    G = _G()
    sizes = [20000, 1500, 0, 6000, 300, 12000, 2]
    ps, ms, cs = _sets(sizes, 13)
    P, M, C = _dev(ps, ms, cs)
    offs = G.set_offsets(sizes)
    kw = dict(beta=2.0, softening_scale=0.005, max_depth=10)
    res = autotune.tune_batch(P, M, C, offs, leaf_sizes=(16, 32), walk_blocks=(0, 512),
                              walk_threads=(256, 128), reps=1, **kw)
    table = res["table"]
    # adjoint orders x leaves, then the monopole forward, one more grid, one
    # more block size
    assert len(table) == 4 + 1 + 1 + 1
    assert {r["vjp_order"] for r in table} == {2, G.SECOND_MOMENT_ADJOINT}
    assert {r["fwd_order"] for r in table} == {1, 2}
    passing = [r for r in table if r["passes_gate"]]
    assert passing and res["ms"] == min(r["ms"] for r in passing)
    best = [r for r in passing if r["ms"] == res["ms"]][0]
    assert (res["fwd_order"], res["vjp_order"], res["leaf_size"], res["walk_blocks"],
            res["walk_threads"]) == (best["fwd_order"], best["vjp_order"], best["leaf_size"],
                                     best["walk_blocks"], best["walk_threads"])
    assert res["gate"]["rows"] > 0
    assert res["gate"]["sets_checked_alone"] >= 1
    for r in table:
        assert r["passes_gate"] == (max(r["rel_l2"].values()) <= 1e-3)
    assert autotune.tune_batch(P, M, C, offs, leaf_sizes=(16, 32), walk_blocks=(0, 512),
                               walk_threads=(256, 128), reps=1, **kw) is res
    with pytest.raises(RuntimeError, match="no candidate"):
        autotune.tune_batch(P, M, C, offs, leaf_sizes=(16,), walk_blocks=(0,),
                            walk_threads=(256,), reps=1, gate=1e-12, **kw)
    # relaxed: the gate becomes 1.25 x the best accuracy any candidate reached
    with pytest.warns(UserWarning, match="relaxed"):
        rel = autotune.tune_batch(P, M, C, offs, leaf_sizes=(16, 32), walk_blocks=(0,),
                                  walk_threads=(256,), fwd_orders=(2,), reps=1, gate=1e-12,
                                  on_fail="relax", **kw)
    errs = [max(r["rel_l2"].values()) for r in rel["table"]]
    assert rel["gate"]["relaxed"] and rel["gate"]["requested_rel_l2_max"] == 1e-12
    assert rel["gate"]["rel_l2_max"] == pytest.approx(autotune.RELAX_FACTOR * min(errs))
    chosen = [r for r in rel["table"] if r["ms"] == rel["ms"]][0]
    assert chosen["passes_gate"] and max(chosen["rel_l2"].values()) <= rel["gate"]["rel_l2_max"]
