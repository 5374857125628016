#!/usr/bin/env python3
"""Headline benchmark: Barnes-Hut forces/s at N = 2^20, theta = 0.5, fwd + VJP.

Workload (the metric's configuration): N = 1,048,576 fp32 points from a
Plummer sphere (scale radius 1, truncated at r = 20), masses 2^-20, Plummer
softening 0.005, theta = 0.5 (beta = 1/theta = 2 on the cell diagonal),
octree max depth 10, leaf size 32, monopole + quadrupole moments
(--order 1: monopole only).

One step = what a `jax.grad`-style training step does through the recompute
custom VJP: forward (device bounds, Morton codes, radix sort, tree build,
moments, walk) and backward (rebuild + one adjoint walk returning position and
mass gradients).  Inputs are resident in HBM before the timed region.

Prints one JSON line (see the keys below).  The CPU baseline is the oracle's
restatement of the same walk on the host cores (OpenMP), timed on a bounded
sample of query groups and scaled to the full workload; beside it, the
reference's CPU tree path (restated) and its numpy direct sum, forward only.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "barnes-hut_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "BH forces/sec at N=1M θ=0.5 (rel_l2≤1e-3), fwd+VJP; batched vs per-set"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FP32_PEAK_NON_PACKED_TFLOPS = 78.6  # the same VALUs without packed-FP32 instructions
# Algorithmic flops per interaction (one query x one near source or one far
# node), the direct-sum pair counts: 20 forward, 48 full adjoint.  A far node
# counts as one source whatever its expansion order.
FLOPS_ALG_FWD, FLOPS_ALG_VJP = 20, 48
# The implementation's own count: near pairs as above; a far node costs
# 20 / 48 with the monopole, 54 / 142 with the quadrupole terms and 68 in the
# adjoint with monopole terms plus the second-moment adjoint (order 3): the
# FP32 operations, FMA = 2, that the walks' SQ_INSTS_VALU_FLOPS_FP32 counter
# reports per far node (profiles/gravity_walk_plummer_n1048576_valu_counters.md).
FLOPS_NEAR_FWD, FLOPS_NEAR_VJP = 20, 48
FLOPS_FAR_FWD = {1: 20, 2: 54}
FLOPS_FAR_VJP = {1: 48, 2: 142, 3: 68}
OUTLIER_REL_THRESHOLD = 0.05  # per-row relative error counted as an outlier


def walk_flops(stats, vjp, order):
    """(algorithmic, implementation) flops of one walk with interaction
    counts `stats`."""
    near, far = stats["near"], stats["far"]
    if vjp:
        return (FLOPS_ALG_VJP * (near + far),
                FLOPS_NEAR_VJP * near + FLOPS_FAR_VJP[order] * far)
    return (FLOPS_ALG_FWD * (near + far), FLOPS_NEAR_FWD * near + FLOPS_FAR_FWD[order] * far)


def valu_roofline(kernel, flops_alg, flops_impl, ms, traffic=None, **extra):
    """FP32-VALU roofline record of one walk launch: achieved and frac by the
    algorithmic count, the implementation's count beside it, against the
    157.3 TFLOP/s peak (packed FP32) and the 78.6 TFLOP/s of unpacked issue."""
    a = flops_alg / (ms * 1e-3) / 1e12
    b = flops_impl / (ms * 1e-3) / 1e12
    rec = {"bound": "valu", "compute": "fp32 VALU (no matrix op in the walk)",
           "kernel": kernel, "achieved": a, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": a / FP32_PEAK_TFLOPS, "traffic": traffic,
           "flop_model": "20 flop per forward and 48 per adjoint interaction (near pair or "
                         "far node)",
           "flops_per_launch": flops_alg, "ms_per_launch": ms,
           "frac_of_non_packed_peak": a / FP32_PEAK_NON_PACKED_TFLOPS,
           "implementation_count": {
               "flop_model": "near pair 20 / 48, far node 20 / 48 (monopole), 54 / 142 "
                             "(quadrupole), - / 68 (second-moment adjoint)",
               "achieved": b, "frac": b / FP32_PEAK_TFLOPS,
               "frac_of_non_packed_peak": b / FP32_PEAK_NON_PACKED_TFLOPS,
               "flops_per_launch": flops_impl}}
    rec.update(extra)
    return rec


# HBM bytes per launch of the 2^20 Plummer walks from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes: the headline orders, then the quadrupole walks
TRAFFIC_PROFILES = [os.path.join(ROOT, "profiles", f) for f in (
    "gravity_walk_plummer_n1048576_hbm_traffic.md",
    "gravity_walk_plummer_n1048576_quad_hbm_traffic.md")]


def profiled_traffic(vjp, order, forces):
    """HBM bytes per launch of k_walk<vjp, order, forces> at N = 2^20 measured
    by the rocprofv3 --pmc passes recorded in profiles/, and that file (None,
    None if not profiled)."""
    key = f"k_walk<{str(vjp).lower()},{order},{str(forces).lower()}>"
    for path in TRAFFIC_PROFILES:
        try:
            with open(path) as fh:
                for line in fh:
                    if line.startswith(f"traffic_bytes_per_launch {key}:"):
                        return float(line.split(":", 1)[1]), "profiles/" + os.path.basename(path)
        except OSError:
            pass
    return None, None


def direct_fp64(P, M, Q, eps, chunk=16):
    import torch

    out = []
    e2 = eps * eps + 1e-12
    for q in Q.split(chunk):
        d = q[:, None, :] - P[None]
        s3 = ((d * d).sum(-1) + e2) ** -1.5
        out.append(-((M[None] * s3)[..., None] * d).sum(1))
    return torch.cat(out)


def direct_vjp_fp64(P, M, Q, C, eps, chunk=16):
    """FP64 query-side VJP of the softened direct force at queries Q with
    cotangents C: sum_k m_k (3 s^5 (c . d) d - s^3 c), d = q - x_k."""
    import torch

    out = []
    e2 = eps * eps + 1e-12
    for q, c in zip(Q.split(chunk), C.split(chunk)):
        d = q[:, None, :] - P[None]
        r2 = (d * d).sum(-1) + e2
        s3 = r2 ** -1.5
        s5 = s3 / r2
        cd = (d * c[:, None, :]).sum(-1)
        out.append((M[None, :, None] * (3.0 * (s5 * cd)[..., None] * d -
                                        s3[..., None] * c[:, None, :])).sum(1))
    return torch.cat(out)


def rel_l2(a, b):
    import torch

    a = a.double()
    b = b.double()
    return float(torch.linalg.norm(a - b) / torch.clamp(torch.linalg.norm(b), min=1e-30))


def outlier_frac(pred, ref, thr=OUTLIER_REL_THRESHOLD):
    """Fraction of rows whose relative error exceeds thr (the reference
    benchmark's outlier metric, python/unified_force_benchmark.py:316-328)."""
    import torch

    d = torch.linalg.norm((pred.double() - ref.double()).reshape(ref.shape[0], -1), dim=1)
    r = torch.linalg.norm(ref.double().reshape(ref.shape[0], -1), dim=1)
    return float((d / torch.clamp(r, min=1e-12) > thr).double().mean())


def direct_adjoint_rows_fp64(P, M, C, rows, eps, chunk=8):
    """FP64 direct-sum gradients of L = sum_i c_i . F_i (softened force, self
    pair excluded) with respect to the positions and masses of the points
    `rows`: (grad_points[rows], grad_masses[rows])."""
    import torch

    e2 = eps * eps + 1e-12
    gx, gm = [], []
    for r in rows.split(chunk):
        d = P[r][:, None, :] - P[None]  # x_j - x_k
        r2 = (d * d).sum(-1) + e2
        s3 = r2 ** -1.5
        s5 = s3 / r2
        self_mask = torch.arange(P.shape[0], device=P.device)[None, :] == r[:, None]
        s3 = s3.masked_fill(self_mask, 0.0)
        s5 = s5.masked_fill(self_mask, 0.0)
        cj = C[r]
        dc = (d * cj[:, None, :]).sum(-1)  # d_jk . c_j
        q = -(M[None, :, None] * (s3[..., None] * cj[:, None, :] -
                                  3.0 * (s5 * dc)[..., None] * d)).sum(1)
        cd = (d * C[None]).sum(-1)  # c_k . d_jk
        src = M[r][:, None] * (s3[..., None] * C[None] - 3.0 * (s5 * cd)[..., None] * d).sum(1)
        gx.append(q + src)
        gm.append((s3 * cd).sum(1))
    return torch.cat(gx), torch.cat(gm)


def cpu_baseline(p, m, c, args, sample_groups):
    """Oracle restatement of the same fwd + VJP step on the host cores
    (OpenMP over query groups), on `sample_groups` query groups, scaled to all
    groups."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    T = O.Tree(p, m, max_depth=args.depth, leaf_size=args.leaf)
    t_build = time.perf_counter() - t0
    ng = len(T.groups(64)[0])
    gsel = min(sample_groups, ng)
    g0 = (ng - gsel) // 2
    vjp_mode = {1: 0, 2: 1, 3: 2}[args.order]  # the oracle's expansion modes
    cs = c[T.perm]
    t0 = time.perf_counter()
    T.eval_groups(g0, g0 + gsel, beta=args.beta, eps=args.eps, use_quad=args.fwd_order == 2)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    T.eval_groups(0, 0, beta=args.beta, eps=args.eps, use_quad=vjp_mode, cot_sorted=cs)
    t_push = time.perf_counter() - t0
    t0 = time.perf_counter()
    T.eval_groups(g0, g0 + gsel, beta=args.beta, eps=args.eps, use_quad=vjp_mode,
                  cot_sorted=cs)
    t_vjp = time.perf_counter() - t0
    scale = ng / gsel
    t_step = 2 * t_build + t_fwd * scale + t_push + max(t_vjp - t_push, 0.0) * scale
    return {
        "value": T.n / t_step,
        "unit": "forces/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"oracle adaptive-tree walk with the same MAC and expansions (fp64, OpenMP "
                   f"over query groups): tree build x2 timed in full; forward and adjoint "
                   f"walks timed on {gsel} of {ng} query groups (<= 64 queries each) and "
                   f"scaled by {scale:.2f}"),
        "seconds_measured": round(2 * t_build + t_fwd + t_vjp, 2),
        "host_cpu": platform.processor() or platform.machine(),
        "host_threads_available": os.cpu_count(),
    }


def cpu_reference_path(p, m, args, ref_idx, ref, nsample=1 << 15):
    """The reference's CPU path (src/sbh.cpp evalBarnesHut over Tree3D),
    restated in oracle/sbh_cpu.c, like-for-like: a true octree (Tree3D<L,1>)
    with the benchmark's leaf size, the softened force, forward only; the
    walk is timed on `nsample` queries (plus the accuracy sample) and scaled."""
    import numpy as np

    from oracle import oracle as O

    n = p.shape[0]
    t0 = time.perf_counter()
    T = O.CpuTree(p, m, leaf_size=args.leaf, dim_split_log=1, max_depth=args.depth)
    t_build = time.perf_counter() - t0
    sel = np.arange(0, n, max(1, n // nsample))[:nsample]
    t0 = time.perf_counter()
    _, n_int = T.eval(p[sel], O.CPU_SOFT_FORCE, beta=args.beta, eps=args.eps)
    t_eval = (time.perf_counter() - t0) * n / sel.size
    f, _ = T.eval(p[ref_idx], O.CPU_SOFT_FORCE, beta=args.beta, eps=args.eps)
    err = float(np.linalg.norm(f - ref) / np.linalg.norm(ref))
    return {"value": n / (t_build + t_eval), "unit": "forces/s (forward only)",
            "threads": O.num_threads(), "build_s": round(t_build, 3),
            "eval_s_scaled": round(t_eval, 3), "sample_queries": int(sel.size),
            "interactions_per_query": n_int / sel.size,
            "force_rel_l2_vs_fp64_direct": err,
            "variant": f"Tree3D<{args.leaf},1> restatement, MAC |q-com|/diag >= {args.beta}, "
                       "softened force, single-threaded build, strided OpenMP queries"}


def cpu_reference_literal(p, m, args, nsample=1 << 14):
    """The reference's CPU path as it ships (src/sbh.cpp:49-88): 512-ary
    Octree3 (Tree3D<1,3>), leaf size 1, max depth 100, potential
    -m/(r + 0.1), MAC |q - COM| / diagonal >= beta; walk timed on `nsample`
    strided queries and scaled."""
    import numpy as np

    from oracle import oracle as O

    n = p.shape[0]
    t0 = time.perf_counter()
    T = O.CpuTree(p, m, leaf_size=1, dim_split_log=3, max_depth=100)
    t_build = time.perf_counter() - t0
    sel = np.arange(0, n, max(1, n // nsample))[:nsample]
    t0 = time.perf_counter()
    _, n_int = T.eval(p[sel], O.CPU_POTENTIAL, beta=args.beta)
    t_eval = (time.perf_counter() - t0) * n / sel.size
    return {"value": n / (t_build + t_eval), "unit": "potentials/s (forward only)",
            "threads": O.num_threads(), "build_s": round(t_build, 3),
            "eval_s_scaled": round(t_eval, 3), "sample_queries": int(sel.size),
            "interactions_per_query": n_int / sel.size,
            "variant": "Tree3D<1,3> (512-ary Octree3), leaf 1, depth 100, potential "
                       "-m/(r+0.1), single-threaded build, strided OpenMP queries"}


def cpu_numpy_direct(p, m, args, nsample=64):
    """The reference benchmark's numpy direct sum (unified_force_benchmark.py:
    208-230, fp32, one query at a time), timed on `nsample` queries and scaled."""
    import numpy as np

    eps2 = np.float32(args.eps * args.eps)
    t0 = time.perf_counter()
    for i in range(nsample):
        disp = p[i][None, :] - p
        r2 = np.sum(disp * disp, axis=1)
        inv_r = np.float32(1.0) / np.sqrt(r2 + eps2 + np.float32(1e-12))
        coeff = -m * inv_r * inv_r * inv_r
        np.sum(coeff[:, None] * disp, axis=0, dtype=np.float32)
    t = (time.perf_counter() - t0) / nsample
    return {"value": 1.0 / t, "unit": "forces/s (forward only)", "cores": 1,
            "sample": f"{nsample} of {p.shape[0]} queries, fp32 numpy, scaled"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--order", type=int, default=3, choices=(1, 2, 3),
                    help="far-field expansion of the adjoint walk: 1 = monopole, 2 = "
                         "monopole + quadrupole, 3 = monopole terms with the quadrupole "
                         "term's second-moment adjoint (default; the monopole adjoint "
                         "misses rel_l2 <= 1e-3 on the gradients at N = 2^20)")
    ap.add_argument("--fwd-order", type=int, default=1, choices=(1, 2),
                    help="far-field expansion of the forward walk (default 1, the monopole "
                         "of the reference's octree8 node records: forces at rel_l2 ~1e-4 "
                         "at N = 2^20)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB before the timed "
                         "region (reference kernel for HBM byte counters)")
    ap.add_argument("--depth", type=int, default=10)
    ap.add_argument("--leaf", type=int, default=32)
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--beta", type=float, default=2.0)
    ap.add_argument("--accuracy-queries", type=int, default=1024)
    ap.add_argument("--grad-rows", type=int, default=256,
                    help="rows of the FP64 direct adjoint the gradients are checked on")
    ap.add_argument("--no-batched", action="store_true",
                    help="gravity: skip the 64-set batched record of the headline line")
    ap.add_argument("--no-config1", action="store_true",
                    help="gravity: skip the N = 10,000 uniform-cube record")
    ap.add_argument("--cpu-sample-groups", type=int, default=10000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager launches only (default: the step is also captured as a "
                         "HIP graph and the replay is the reported value)")
    ap.add_argument("--workload", default="gravity",
                    choices=("gravity", "wn", "batched", "stochastic", "treepm", "quadtree",
                             "fields", "config1", "vjp_blobs"),
                    help="gravity: the headline fwd+VJP step; wn: winding numbers of a "
                         "204,020-face perturbed geodesic sphere at 2^20 queries; batched: "
                         "fwd+VJP of 64 blob sets (log-uniform sizes 1,024..524,288), one "
                         "graph-captured forest pipeline vs 64 per-set calls; stochastic: "
                         "stochastic BH (forces and query-side VJP) on the adaptive tree at the "
                         "headline Plummer sphere, with the seed-mean unbiasedness check, and "
                         "on the dense octree8 beside it; vjp_blobs: the custom-VJP path on "
                         "2^18 points in 64 blobs (eps 0.002) through torch.autograd; "
                         "config1: 10,000 uniform points with the CPU legs; treepm: the "
                         "short-range TreePM force (periodic, erfc cutoff, prune) on 128^3 "
                         "clustered particles; quadtree: the 2-D targets over Tree2D<1,3> "
                         "buffers (device build + vote walk, and the stochastic walk) on 2^20 "
                         "points in a square; fields: the reference's tree fields module "
                         "(python/jax_sbh.py defaults, level-1 octree) on 2^20 points")
    ap.add_argument("--o8-depth", type=int, default=8)
    ap.add_argument("--treepm-n1d", type=int, default=128,
                    help="treepm: particles per side (n1d^3 particles, box of n1d cells)")
    ap.add_argument("--treepm-depth", type=int, default=8)
    ap.add_argument("--qt-n", type=int, default=1 << 20,
                    help="quadtree / fields: number of points (queries = the points)")
    ap.add_argument("--fields-level", type=int, default=1, choices=(1, 2, 3))
    ap.add_argument("--seeds", type=int, default=256)
    ap.add_argument("--stoch-seeds", type=int, default=4096,
                    help="stochastic: seeds whose spread gives the standard error of the "
                         "--seeds mean")
    ap.add_argument("--no-octree8", action="store_true",
                    help="stochastic, wn: skip the dense octree8 legs")
    ap.add_argument("--no-stochastic", action="store_true",
                    help="wn: skip the stochastic estimate on the adaptive tree")
    ap.add_argument("--stoch-queries", type=int, default=4096)
    ap.add_argument("--l3-queries", type=int, default=65536,
                    help="queries of the timed level-3 walk (stochastic workload)")
    ap.add_argument("--sets", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--wn-freq", type=int, default=101,
                    help="geodesic frequency of the wn mesh (20 f^2 faces)")
    ap.add_argument("--wn-samples", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of "
                                   "it (the query groups would shard by Morton range, one "
                                   "range per GPU, over a replicated tree)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    bench = {"gravity": bench_gravity, "wn": bench_wn, "batched": bench_batched,
             "stochastic": bench_stochastic, "treepm": bench_treepm,
             "quadtree": bench_quadtree, "fields": bench_fields,
             "config1": bench_config1, "vjp_blobs": bench_vjp_blobs}[args.workload]
    rec = bench(args)
    print(json.dumps(rec))


def bench_gravity(args):
    import numpy as np
    import torch

    import sbh_amd
    from sbh_amd import gravity as G
    from sbh_amd import synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = args.n
    p = synthetic.plummer(n, 2000)
    m = np.full(n, 1.0 / n, np.float32)
    c = np.random.default_rng(2002).normal(size=(n, 3)).astype(np.float32)
    P = torch.tensor(p, device=dev)
    M = torch.tensor(m, device=dev)
    C = torch.tensor(c, device=dev)
    kw = dict(beta=args.beta, softening_scale=args.eps, max_depth=args.depth,
              leaf_size=args.leaf, multipole_order=args.order)
    kw_f = dict(kw, multipole_order=args.fwd_order)
    ws = G.new_workspace(n, args.depth, args.leaf, dev)
    F = torch.empty_like(P)
    GX, GM = torch.empty_like(P), torch.empty_like(M)

    def step(w=ws):
        G.gravity_force(P, M, workspace=w, out=F, **kw_f)
        return G.gravity_force_vjp(P, M, C, workspace=w, grad_points_out=GX,
                                   grad_masses_out=GM, **kw)

    if args.copy_calibration_mb > 0:
        src = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib = sbh_amd.load_library()
    lib.sbh_profile_enable(ctypes.c_int32(1))
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        gx, gm = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    lib.sbh_profile_enable(ctypes.c_int32(0))
    prof = (ctypes.c_double * 6)()
    lib.sbh_profile_collect(prof, ctypes.c_int32(6))
    eager_ms = ev0.elapsed_time(ev1) / args.steps

    # the same step captured once as a HIP graph and replayed (no per-launch
    # host work); every replay recomputes everything
    graph_ms = None
    if not args.no_graph:
        from sbh_amd.graph import Graph

        g = Graph()
        g.capture(step)
        for _ in range(max(args.warmup, 1)):
            g.replay()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(args.steps):
            g.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        graph_ms = ev0.elapsed_time(ev1) / args.steps
        g.close()
        gx, gm = GX, GM
    ms_step = graph_ms if graph_ms is not None else eager_ms
    value = n / (ms_step * 1e-3)

    # Each walk deals its query groups heaviest first by the costs that the
    # previous walk of its kind over the same workspace measured (only the
    # schedule depends on them).  A fresh workspace has none: its forward walk
    # takes the groups in their own order and its adjoint walk orders by that
    # forward walk's costs.  The first step on each of 3 zeroed workspaces:
    cold = []
    for _ in range(3):
        wsc = torch.zeros_like(ws)
        torch.cuda.synchronize()
        ev0.record(stream)
        step(wsc)
        ev1.record(stream)
        torch.cuda.synchronize()
        cold.append(ev0.elapsed_time(ev1))
        del wsc
    schedule = {"first_step_on_fresh_workspace_ms": float(np.mean(cold)),
                "note": "timed steps reuse one workspace: each walk orders its groups by "
                        "the costs the previous walk of its kind measured (heaviest "
                        "first within each eighth of the groups); results do not depend "
                        "on the order"}

    # interaction counts (deterministic: identical every step)
    sf, sv = G.new_stats(dev), G.new_stats(dev)
    G.gravity_force(P, M, workspace=ws, stats=sf, **kw_f)
    G.gravity_force_vjp(P, M, C, workspace=ws, stats=sv, **kw)
    sf, sv = G.stats_dict(sf), G.stats_dict(sv)
    walk_vjp_ms = prof[5] / max(prof[3], 1)
    walk_fwd_ms = prof[2] / max(prof[0], 1)
    flops_vjp, flops_vjp_impl = walk_flops(sv, True, args.order)
    flops_fwd, flops_fwd_impl = walk_flops(sf, False, args.fwd_order)

    # accuracy against the FP64 direct sum on a query sample: forces, and both
    # gradients against the FP64 direct adjoint of the same rows
    idx = torch.tensor(np.random.default_rng(7).choice(n, args.accuracy_queries, replace=False),
                       device=dev)
    ref = direct_fp64(P.double(), M.double(), P[idx].double(), args.eps)
    rows = idx[:args.grad_rows]
    rgx, rgm = direct_adjoint_rows_fp64(P.double(), M.double(), C.double(), rows, args.eps)
    acc = {"force_rel_l2_vs_fp64_direct": rel_l2(F[idx], ref),
           "force_outlier_frac_vs_fp64_direct": outlier_frac(F[idx], ref),
           "outlier_rel_threshold": OUTLIER_REL_THRESHOLD,
           "force_sample_queries": args.accuracy_queries,
           "grad_points_rel_l2_vs_fp64_direct_adjoint": rel_l2(gx[rows], rgx),
           "grad_masses_rel_l2_vs_fp64_direct_adjoint": rel_l2(gm[rows], rgm),
           "grad_points_outlier_frac_vs_fp64_direct_adjoint": outlier_frac(gx[rows], rgx),
           "grad_sample_rows": int(rows.numel())}
    # the same step with other expansion orders (graph replay), beside: all
    # quadrupole, and the headline's forward with the full quadrupole adjoint
    alts = []
    if not args.no_graph:
        from sbh_amd.graph import Graph

        F_a, GX_a, GM_a = torch.empty_like(P), torch.empty_like(P), torch.empty_like(M)
        for fo, vo in ((2, 2), (args.fwd_order, 2)):
            if (fo, vo) == (args.fwd_order, args.order) or (fo, vo) in [
                    (x["forward_order"], x["vjp_order"]) for x in alts]:
                continue
            kw_fa, kw_va = dict(kw, multipole_order=fo), dict(kw, multipole_order=vo)

            def step_alt():
                G.gravity_force(P, M, workspace=ws, out=F_a, **kw_fa)
                G.gravity_force_vjp(P, M, C, workspace=ws, grad_points_out=GX_a,
                                    grad_masses_out=GM_a, **kw_va)

            g = Graph()
            g.capture(step_alt)
            g.replay()
            torch.cuda.synchronize()
            ev0.record(stream)
            for _ in range(args.steps):
                g.replay()
            ev1.record(stream)
            torch.cuda.synchronize()
            alt_ms = ev0.elapsed_time(ev1) / args.steps
            g.close()
            alts.append({"forward_order": fo, "vjp_order": vo, "graph_ms_per_step": alt_ms,
                         "forces_per_s": n / (alt_ms * 1e-3),
                         "force_rel_l2_vs_fp64_direct": rel_l2(F_a[idx], ref),
                         "grad_points_rel_l2_vs_fp64_direct_adjoint": rel_l2(GX_a[rows], rgx),
                         "grad_masses_rel_l2_vs_fp64_direct_adjoint": rel_l2(GM_a[rows], rgm)})
        del F_a, GX_a, GM_a
    direct = None
    if not args.no_direct:
        from sbh_amd import ffi

        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t2 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        Fd = ffi.softened_direct_force(P, M, P, softening_scale=args.eps,
                                       cutoff_scale=float("inf"))
        t1.record(stream)
        gxd, gmd = ffi.softened_direct_force_full_vjp(P, M, C, softening_scale=args.eps,
                                                      cutoff_scale=float("inf"))
        t2.record(stream)
        torch.cuda.synchronize()
        d_ms = t0.elapsed_time(t2)
        direct = {"fwd_ms": t0.elapsed_time(t1), "vjp_ms": t1.elapsed_time(t2),
                  "fwd_plus_vjp_ms": d_ms, "speedup_bh_vs_direct": d_ms / ms_step,
                  "direct_fwd_tflops": 20.0 * n * n / (t0.elapsed_time(t1) * 1e-3) / 1e12,
                  "direct_fwd_rel_l2_vs_fp64": rel_l2(Fd[idx], ref)}
        acc["grad_points_rel_l2_vs_fp32_direct"] = rel_l2(gx, gxd)
        acc["grad_masses_rel_l2_vs_fp32_direct"] = rel_l2(gm, gmd)
        acc["force_full_rel_l2_vs_fp32_direct"] = rel_l2(F, Fd)

    batched = None if args.no_batched else measure_batched(args, dev, tune=not args.no_autotune)
    config1 = None if args.no_config1 else measure_config1(args, dev)

    cpu = cpu_ref = cpu_np = cpu_lit = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, m, c, args, args.cpu_sample_groups)
        cpu_ref = cpu_reference_path(p, m, args, idx.cpu().numpy(), ref.cpu().numpy())
        cpu_np = cpu_numpy_direct(p, m, args)
        cpu_lit = cpu_reference_literal(p, m, args)

    vjp_traffic = profiled_traffic(True, args.order, False) if n == 1 << 20 else (None, None)
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "forces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plummer_sphere_n{n}_theta0.5_depth{args.depth}_leaf{args.leaf}"
                        f"_eps{args.eps}_fwd+vjp",
            "n": n, "theta": 1.0 / args.beta, "beta_diag": args.beta, "max_depth": args.depth,
            "leaf_size": args.leaf, "softening": args.eps,
            "multipole_order": {"forward": args.fwd_order, "vjp": args.order},
            "step": "forward (build+walk) + recompute VJP (rebuild+adjoint walk), full "
                    "position and mass gradients",
            "batched": "see the batched record: 64 blob sets, one graph-captured forest "
                       "pipeline vs the 64 per-set pipelines captured as one graph",
            "launch": "HIP graph replay" if graph_ms is not None else "eager",
        },
        "eager_ms_per_step": eager_ms,
        "graph_ms_per_step": graph_ms,
        "schedule": schedule,
        "phases_ms": {"fwd_build": prof[1] / max(prof[0], 1), "fwd_walk_kernel": walk_fwd_ms,
                      "vjp_build": prof[4] / max(prof[3], 1), "vjp_walk_kernel": walk_vjp_ms},
        "interactions": {"fwd": sf, "vjp": sv},
        "accuracy": acc,
        "other_orders": alts,
        "direct_sum": direct,
        "roofline": valu_roofline(
            f"k_walk<vjp=1,order={args.order},forces=0>", flops_vjp, flops_vjp_impl,
            walk_vjp_ms,
            traffic=vjp_traffic[0],
            traffic_source=vjp_traffic[1],
            timing="library HIP events around the walk launch on its stream, mean of the "
                   "timed eager steps"),
        "other_rooflines": other_rooflines(args, n, sf, walk_fwd_ms, flops_fwd, flops_fwd_impl,
                                           prof, direct),
        "batched": batched,
        "config1": config1,
        "cpu_baseline": cpu,
        "cpu_reference_path": cpu_ref,
        "cpu_numpy_direct": cpu_np,
        "cpu_reference_literal": cpu_lit,
    }
    return rec


def build_bytes(n, nodes, vjp):
    """Algorithmic (minimal) HBM bytes of one tree build of n points with
    `nodes` nodes: read position + mass (16 B), write the sorted record
    (16 B), permutation (4 B) and key (4 B) per point, and 48 B of node
    record (topology + monopole) per node; the adjoint build also reads and
    writes the cotangent (12 + 16 B per point)."""
    return (40 + (28 if vjp else 0)) * n + 48 * nodes


def build_bytes_design(n, nodes, vjp):
    """HBM bytes the build as designed must move: bounds (read 12 B/pt),
    Morton codes (read 12, write 8), 4-pass radix sort (per pass: histogram
    reads 4, scatter reads 8 and writes 8), gather (read perm 4 + position 12
    + mass 4 + code 4, write sorted record 16 + code 4; adjoint +12 read +16
    written), leaf moments (read 16), and 84 B written per node (topology 16,
    geometry 16, parent 4, monopole 16, second moment 32)."""
    per_point = 12 + 20 + 4 * 20 + 44 + 16 + (28 if vjp else 0)
    return per_point * n + 84 * nodes


def other_rooflines(args, n, sf, walk_fwd_ms, flops_fwd, flops_fwd_impl, prof, direct):
    """Build (HBM-bound), forward walk (FP32) and direct sum (FP32) against the
    chip's peaks; the VJP walk's HBM rate from the profiled traffic."""
    fwd_build_ms = prof[1] / max(prof[0], 1)
    vjp_build_ms = prof[4] / max(prof[3], 1)
    vjp_walk_ms = prof[5] / max(prof[3], 1)
    out = {}
    for tag, ms, vjp in (("tree_build_fwd", fwd_build_ms, False),
                         ("tree_build_vjp", vjp_build_ms, True)):
        b = build_bytes(n, sf["nodes"], vjp)
        gbs = b / (ms * 1e-3) / 1e9
        out[tag] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "bytes": b, "ms": ms,
                    "design_bytes": build_bytes_design(n, sf["nodes"], vjp),
                    "note": "24 dependent launches per build (10 of them the internal "
                            "moment levels); launch latency and the radix passes' "
                            "look-back set the time, not bandwidth"}
    out["walk_fwd"] = valu_roofline(f"k_walk<vjp=0,order={args.fwd_order},forces=1>",
                                    flops_fwd, flops_fwd_impl, walk_fwd_ms,
                                    traffic=(profiled_traffic(False, args.fwd_order, True)[0]
                                             if n == 1 << 20 else None))
    # walks against HBM: profiled traffic over the kernel time, beside the
    # minimal bytes (40 B/pt + 48 B/node; the adjoint +36 B/pt)
    for tag, vjp, ms, extra, order in (
            ("walk_fwd_hbm", False, walk_fwd_ms, 0, args.fwd_order),
            ("walk_vjp_hbm", True, vjp_walk_ms, 36, args.order)):
        traffic, source = profiled_traffic(vjp, order, not vjp)
        if traffic is None or n != 1 << 20:
            continue
        gbs = traffic / (ms * 1e-3) / 1e9
        out[tag] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": source,
                    "minimal_bytes": (40 + extra) * n + 48 * sf["nodes"]}
    if direct is not None:
        t = direct["direct_fwd_tflops"]
        out["direct_sum_fwd"] = {"bound": "valu", "compute": "fp32 VALU, LDS-tiled",
                                 "achieved": t, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": t / FP32_PEAK_TFLOPS, "flops_per_pair": 20}
    return out

FLOPS_WN = 20  # per lane interaction (element or node dipole), fp32


def wn_fp64(Pc, Nn, A, Q, chunk=8):
    """FP64 dipole sum (signedSolidAngle<3>) on the device for a query sample."""
    import math

    import torch

    out = []
    for q in Q.split(chunk):
        d = Pc[None] - q[:, None, :]
        r = torch.linalg.norm(d, dim=-1)
        out.append((A[None] * (Nn[None] * d).sum(-1) / (4 * math.pi * (r ** 3 + 1e-5))).sum(1))
    return torch.cat(out)


def triangle_wn_fp64(V, F, Q, chunk=8):
    """FP64 exact solid-angle sum of the triangle mesh (Van Oosterom-Strackee)
    on the device for a query sample."""
    import math

    import torch

    A0, B0, C0 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    out = []
    for q in Q.split(chunk):
        a, b, c = A0[None] - q[:, None], B0[None] - q[:, None], C0[None] - q[:, None]
        la, lb, lc = a.norm(dim=-1), b.norm(dim=-1), c.norm(dim=-1)
        det = (a * torch.cross(b, c, dim=-1)).sum(-1)
        den = (la * lb * lc + (a * b).sum(-1) * lc + (a * c).sum(-1) * lb +
               (b * c).sum(-1) * la)
        out.append(2.0 * torch.atan2(det, den).sum(1) / (4 * math.pi))
    return torch.cat(out)


def cpu_baseline_wn(c, nrm, a, q, leaf, sample):
    """Oracle restatement of the adaptive winding-number walk on the host
    (OpenMP threads), on `sample` queries, scaled to the full query count."""
    import numpy as np

    from oracle import oracle as O

    t0 = time.perf_counter()
    T = O.Tree(c, a, max_depth=10, leaf_size=leaf)
    t_build = time.perf_counter() - t0
    sel = np.random.default_rng(11).choice(q.shape[0], sample, replace=False)
    qs = q[np.sort(sel)]
    t0 = time.perf_counter()
    T.wn_eval(nrm, qs, beta=2.0)
    t_walk = time.perf_counter() - t0
    t_all = t_build + t_walk * q.shape[0] / sample
    return {
        "value": q.shape[0] / t_all, "unit": "winding numbers/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
        "kind": "port",
        "sample": (f"oracle adaptive-tree winding-number walk (fp64, OpenMP): element tree "
                   f"build, query tree and walk timed on {sample} of {q.shape[0]} queries"
                   + (" and scaled" if sample < q.shape[0] else "")),
        "seconds_measured": round(t_build + t_walk, 2),
        "host_cpu": platform.processor() or platform.machine(),
        "host_threads_available": os.cpu_count(),
    }


def bench_wn(args):
    """Winding numbers: 20 f^2-face geodesic sphere (f = 101: 204,020 faces)
    with its radius perturbed by real spherical harmonics l = 1..6, one dipole
    per face, 2^20 queries uniform in [-1.5, 1.5]^3; classic theta = 0.5."""
    import numpy as np
    import torch

    from sbh_amd import ffi, synthetic, wn

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    V, F = synthetic.perturbed_sphere(args.wn_freq, 4000)
    c, nrm, a = synthetic.face_dipoles(V, F)
    m = args.n
    q = np.random.default_rng(4003).uniform(-1.5, 1.5, (m, 3)).astype(np.float32)
    Pc, Nn, A, Q = (torch.tensor(x, device=dev) for x in (c, nrm, a, q))
    leaf = args.leaf if args.leaf != 32 else 16
    ws = wn.new_workspace(c.shape[0], m, 10, leaf, dev)
    out = torch.empty(m, dtype=torch.float32, device=dev)

    def step():
        return wn.winding_number(Pc, Nn, A, Q, beta=args.beta, max_depth=10, leaf_size=leaf,
                                 workspace=ws, out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms_step = ev[0].elapsed_time(ev[1]) / args.steps
    # phase times from the library's HIP events (build = element tree,
    # dipoles and query tree; walk = the k_wn_walk launch)
    import sbh_amd

    lib = sbh_amd.load_library()
    lib.sbh_profile_enable(ctypes.c_int32(1))
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    lib.sbh_profile_enable(ctypes.c_int32(0))
    prof = (ctypes.c_double * 9)()
    lib.sbh_profile_collect(prof, ctypes.c_int32(9))
    wn_build_ms = prof[7] / max(prof[6], 1)
    wn_walk_ms = prof[8] / max(prof[6], 1)
    st = torch.zeros(8, dtype=torch.int64, device=dev)
    wn.winding_number(Pc, Nn, A, Q, beta=args.beta, max_depth=10, leaf_size=leaf,
                      workspace=ws, stats=st)
    s = [int(x) for x in st.cpu()]
    flops = FLOPS_WN * (s[0] + s[1])

    # accuracy on a query sample: FP64 dipole sum and FP64 exact triangles
    idx = torch.tensor(np.random.default_rng(7).choice(m, args.accuracy_queries, replace=False),
                       device=dev)
    ref = wn_fp64(Pc.double(), Nn.double(), A.double(), Q[idx].double())
    exact = triangle_wn_fp64(torch.tensor(V, device=dev), torch.tensor(F, device=dev),
                             Q[idx].double())
    acc = {"wn_rel_l2_vs_fp64_dipole_sum": rel_l2(out[idx], ref),
           "dipole_sum_rel_l2_vs_exact_triangles": rel_l2(ref, exact),
           "sample_queries": args.accuracy_queries}

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return r, e0.elapsed_time(e1) / reps

    others = {}
    if not args.no_direct:
        bf, bf_ms = timed(lambda: wn.eval_brute_force_wn(Pc, Nn, A, Q), reps=1)
        others["brute_force"] = {"ms": bf_ms, "speedup_tree_vs_brute": bf_ms / ms_step,
                                 "rel_l2_vs_fp64": rel_l2(bf[idx], ref),
                                 "tflops": FLOPS_WN * m * c.shape[0] / (bf_ms * 1e-3) / 1e12}
    sb_ms = None
    if not args.no_octree8:
        lb, nb = ffi.build_octree8_buffers(Pc, Nn, A, max_depth=7)
        cl, cl_ms = timed(lambda: wn.eval_barnes_hut_wn(lb, nb, Q, beta=args.beta))
        others["octree8_classic_depth7"] = {"ms": cl_ms, "rel_l2_vs_fp64": rel_l2(cl[idx], ref)}
        sb, sb_ms = timed(lambda: wn.multi_level_prefix_control_variate_wn(
            lb, nb, Q, samples_per_subdomain=args.wn_samples, seed=0, beta=1.0), reps=1)
        others["octree8_stochastic_depth7"] = {"ms": sb_ms, "samples_per_subdomain":
                                               args.wn_samples,
                                               "rel_l2_vs_fp64": rel_l2(sb[idx], ref)}
    if not args.no_stochastic:
        # the same estimator on the adaptive element tree (depth 10), one wave per
        # query group sharing each sampled path
        from sbh_amd import stochastic as S

        wss = S.new_workspace(c.shape[0], m, 10, leaf, dev)
        sto = torch.zeros(8, dtype=torch.int64, device=dev)
        kws = dict(beta=args.beta, samples_per_subdomain=args.wn_samples, max_depth=10,
                   leaf_size=leaf, workspace=wss)
        sa, sa_ms = timed(lambda: S.stochastic_winding_number(Pc, Nn, A, Q, seed=0, **kws))
        S.stochastic_winding_number(Pc, Nn, A, Q, seed=0, stats=sto, **kws)
        ss = S.stats_dict(sto)
        lib.sbh_profile_enable(ctypes.c_int32(1))
        S.stochastic_winding_number(Pc, Nn, A, Q, seed=0, **kws)
        torch.cuda.synchronize()
        lib.sbh_profile_enable(ctypes.c_int32(0))
        prof_s = (ctypes.c_double * 12)()
        lib.sbh_profile_collect(prof_s, ctypes.c_int32(12))
        s_walk_ms = prof_s[11] / max(prof_s[9], 1)
        s_flops = FLOPS_WN * (ss["node_terms"] + ss["source_pairs"])
        others["adaptive_stochastic"] = {
            "ms": sa_ms, "samples_per_subdomain": args.wn_samples, "beta_diag": args.beta,
            "max_depth": 10, "leaf_size": leaf, "rel_l2_vs_fp64": rel_l2(sa[idx], ref),
            "terms": ss, "build_ms": prof_s[10] / max(prof_s[9], 1), "walk_kernel_ms": s_walk_ms,
            "speedup_vs_octree8_stochastic": sb_ms / sa_ms if sb_ms else None,
            "roofline": {"bound": "valu", "kernel": "k_stoch_walk (winding numbers)",
                         "achieved": s_flops / (s_walk_ms * 1e-3) / 1e12,
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": s_flops / (s_walk_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                         "flops_per_launch": s_flops,
                         "flop_model": f"{FLOPS_WN} flop per lane node term or element pair"}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_wn(c, nrm, a, q, leaf, m)
    achieved = flops / (wn_walk_ms * 1e-3) / 1e12
    return {
        "metric": "winding numbers/sec (classic BH, theta=0.5)",
        "value": m / (ms_step * 1e-3),
        "unit": "winding numbers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"perturbed_geodesic_sphere_faces{F.shape[0]}_queries{m}"
                               f"_theta0.5_leaf{leaf}",
                   "faces": int(F.shape[0]), "queries": m, "beta_diag": args.beta,
                   "max_depth": 10, "leaf_size": leaf,
                   "step": "element tree build + dipole moments + query tree + group walk"},
        "interactions": {"far": s[0], "near": s[1], "groups": s[2], "nodes": s[3]},
        "phases_ms": {"build": wn_build_ms, "walk_kernel": wn_walk_ms},
        "accuracy": acc,
        "others": others,
        "roofline": {"bound": "valu", "compute": "fp32 VALU (no matrix op in the walk)",
                     "kernel": "k_wn_walk", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }


FLOPS_STOCH_TERM = {0: 20, 1: 34}  # per lane node term or source pair: force, query-side VJP


def stochastic_zstats(sum1, sum2, first, nseeds, batch, ref_scale):
    """z statistics of seed means against a target from running sums of the
    per-seed deviations (sum1, sum2) and the `batch`-seed sum `first`: the
    standard error of a mean comes from all `nseeds` seeds (heavy tails make
    `batch` draws underestimate their own spread)."""
    import numpy as np

    sd = np.sqrt(np.maximum(sum2 - sum1 * sum1 / nseeds, 0.0) / (nseeds - 1))
    ok = sd > 1e-12 * ref_scale
    z_all = (sum1 / nseeds)[ok] / (sd[ok] / np.sqrt(nseeds))
    z_b = (first / batch)[ok] / (sd[ok] / np.sqrt(batch))
    return {f"frac_abs_z_lt_3_{batch}_seed_mean": float(np.mean(np.abs(z_b) < 3)),
            f"rms_z_{batch}_seed_mean": float(np.sqrt(np.mean(z_b ** 2))),
            f"frac_abs_z_lt_3_{nseeds}_seed_mean": float(np.mean(np.abs(z_all) < 3)),
            f"rms_z_{nseeds}_seed_mean": float(np.sqrt(np.mean(z_all ** 2))),
            "standard_error_from_seeds": nseeds}


def bench_stochastic(args):
    """Stochastic BH gravity on the adaptive octree at the Plummer workload of
    the headline (N = 2^20, scale radius 1 truncated at 20, masses 2^-20,
    softening 0.005, theta 0.5, max depth 10, leaf 32, monopole node terms,
    one sample per subdomain): the reference's multi-level prefix control
    variate (src/sbh_ffi.cu:2146-2442) with every point as a query, forward
    and query-side VJP timed; unbiasedness on a query sample against the FP64
    direct sum (the 256-seed mean, its standard error from --stoch-seeds
    seeds).  The reference's dense octree8 targets at depth --o8-depth on the
    same points are recorded beside it (the depth-10 octree8 does not fit)."""
    import numpy as np
    import torch

    import sbh_amd
    from sbh_amd import gravity as G
    from sbh_amd import stochastic as S
    from sbh_amd import synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n, eps, depth, leaf = args.n, args.eps, args.depth, args.leaf
    p = synthetic.plummer(n, 2000)
    m = np.full(n, 1.0 / n, np.float32)
    c = np.random.default_rng(2002).normal(size=(n, 3)).astype(np.float32)
    P, M, C = (torch.tensor(a, device=dev) for a in (p, m, c))
    kw = dict(beta=args.beta, samples_per_subdomain=1, softening_scale=eps, max_depth=depth,
              leaf_size=leaf, multipole_order=1)
    ws = S.new_workspace(n, n, depth, leaf, dev)
    out = torch.empty_like(P)
    lib = sbh_amd.load_library()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    steps = max(1, args.steps)
    fwd = lambda: S.stochastic_gravity_force(P, M, P, seed=1, workspace=ws, out=out, **kw)
    vjp = lambda: S.stochastic_gravity_force_vjp(P, M, P, C, seed=1, workspace=ws, out=out,
                                                 **kw)
    fwd_ms = timed(fwd, steps)
    vjp_ms = timed(vjp, steps)
    phases = {}
    for name, fn in (("fwd", fwd), ("vjp", vjp)):
        lib.sbh_profile_enable(ctypes.c_int32(1))
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        lib.sbh_profile_enable(ctypes.c_int32(0))
        prof = (ctypes.c_double * 12)()
        lib.sbh_profile_collect(prof, ctypes.c_int32(12))
        phases[f"{name}_build"] = prof[10] / max(prof[9], 1)
        phases[f"{name}_walk_kernel"] = prof[11] / max(prof[9], 1)
    stf, stv = torch.zeros(8, dtype=torch.int64, device=dev), torch.zeros(8, dtype=torch.int64,
                                                                          device=dev)
    F1 = S.stochastic_gravity_force(P, M, P, seed=1, workspace=ws, stats=stf, **kw).clone()
    S.stochastic_gravity_force_vjp(P, M, P, C, seed=1, workspace=ws, stats=stv, out=out, **kw)
    sf, sv = S.stats_dict(stf), S.stats_dict(stv)
    flops_f = FLOPS_STOCH_TERM[0] * (sf["node_terms"] + sf["source_pairs"])
    flops_v = FLOPS_STOCH_TERM[1] * (sv["node_terms"] + sv["source_pairs"])

    # accuracy of one seed and of the classic walk on a query sample
    rng = np.random.default_rng(7)
    idx = torch.tensor(rng.choice(n, args.accuracy_queries, replace=False), device=dev)
    ref = direct_fp64(P.double(), M.double(), P[idx].double(), eps)
    Fc = G.gravity_force(P, M, beta=args.beta, softening_scale=eps, max_depth=depth,
                         leaf_size=leaf, multipole_order=1)
    acc = {"stochastic_one_seed_rel_l2_vs_fp64_direct": rel_l2(F1[idx], ref),
           "classic_rel_l2_vs_fp64_direct": rel_l2(Fc[idx], ref),
           "sample_queries": args.accuracy_queries}

    # unbiasedness: 512 of the points (force) and 512 independent draws from
    # the same sphere (VJP; at a query on a source the self pair's m c / eps^3
    # is sampled too rarely to estimate its standard error), seeds
    # 0..nseeds-1 of the same tree
    nseeds, batch = max(args.stoch_seeds, args.seeds), args.seeds
    q = p[rng.choice(n, 512, replace=False)]
    qv = synthetic.plummer(512, 2003)
    cq = rng.normal(size=(512, 3)).astype(np.float32)
    Q, QV, CQ = (torch.tensor(a, device=dev) for a in (q, qv, cq))
    wsq = S.new_workspace(n, 512, depth, leaf, dev)
    refs = {"force": direct_fp64(P.double(), M.double(), Q.double(), eps)}
    refs["vjp"] = direct_vjp_fp64(P.double(), M.double(), QV.double(), CQ.double(), eps)
    s1 = {k: torch.zeros_like(v) for k, v in refs.items()}
    s2 = {k: torch.zeros_like(v) for k, v in refs.items()}
    first = {}
    t0 = time.perf_counter()
    for seed in range(nseeds):
        got = {"force": S.stochastic_gravity_force(P, M, Q, seed=seed, workspace=wsq, **kw),
               "vjp": S.stochastic_gravity_force_vjp(P, M, QV, CQ, seed=seed, workspace=wsq,
                                                     **kw)}
        for k, g in got.items():
            d = g.double() - refs[k]
            s1[k] += d
            s2[k] += d * d
        if seed == batch - 1:
            first = {k: v.cpu().numpy().copy() for k, v in s1.items()}
    torch.cuda.synchronize()
    unbiased = {}
    for k in ("force", "vjp"):
        r = refs[k].cpu().numpy()
        unbiased[k] = stochastic_zstats(s1[k].cpu().numpy(), s2[k].cpu().numpy(), first[k],
                                        nseeds, batch, float(np.abs(r).max()))
        unbiased[k][f"{batch}_seed_mean_rel_l2_vs_fp64_direct"] = float(
            np.linalg.norm(first[k] / batch) / np.linalg.norm(r))
    unbiased["queries"] = {"force": "512 of the points", "vjp": "512 independent draws from "
                           "the same Plummer sphere"}
    unbiased["seconds"] = round(time.perf_counter() - t0, 2)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        T = O.Tree(p, m, max_depth=depth, leaf_size=leaf)
        ns = 1 << 14
        qs = p[rng.choice(n, ns, replace=False)]
        t0 = time.perf_counter()
        T.stoch_eval(qs, kind=0, beta=args.beta, eps=eps, samples=1, seed=1)
        t = (time.perf_counter() - t0) / ns
        cpu = {"value": 1.0 / t, "unit": "forces/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"oracle stochastic walk over the same adaptive tree (same paths, "
                         f"fp64, OpenMP over query groups) on {ns} of {n} queries, query "
                         f"tree included, source tree build excluded, scaled"}
    octree8 = None if args.no_octree8 else measure_octree8_stochastic(args)
    achieved = flops_f / (phases["fwd_walk_kernel"] * 1e-3) / 1e12
    return {
        "metric": "stochastic BH forces/sec at N=1M (adaptive octree, depth 10, leaf 32, 1 "
                  "sample per subdomain)",
        "value": n / (fwd_ms * 1e-3),
        "unit": "forces/s",
        "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": fwd_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plummer_sphere_n{n}_theta{1.0 / args.beta:g}_depth{depth}"
                               f"_leaf{leaf}_eps{eps}_stochastic_s1",
                   "n": n, "queries": n, "max_depth": depth, "leaf_size": leaf,
                   "beta_diag": args.beta, "softening": eps, "samples_per_subdomain": 1,
                   "multipole_order": 1,
                   "step": "source tree build + query tree build + stochastic walk"},
        "vjp_ms_per_step": vjp_ms,
        "phases_ms": phases,
        "terms": {"fwd": sf, "vjp": sv},
        "accuracy": acc,
        "unbiasedness": unbiased,
        "roofline": {"bound": "valu", "compute": "fp32 VALU (no matrix op in the walk)",
                     "kernel": "k_stoch_walk (forces)", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "frac_of_non_packed_peak": achieved / FP32_PEAK_NON_PACKED_TFLOPS,
                     "traffic": None, "flops_per_launch": flops_f,
                     "flop_model": "20 flop per lane node term or source pair (forward), 34 "
                                   "(query-side VJP)",
                     "vjp": {"kernel": "k_stoch_walk (query-side VJP)",
                             "achieved": flops_v / (phases["vjp_walk_kernel"] * 1e-3) / 1e12,
                             "flops_per_launch": flops_v}},
        "cpu_baseline": cpu,
        "octree8_drop_in": octree8,
    }


def measure_octree8_stochastic(args):
    """Classic and stochastic BH gravity on the reference's dense octree8 at
    the headline's Plummer sphere (the octree8 handlers of src/sbh_ffi.cu:
    3183-3248 and 3375-3443): one classic launch, one stochastic launch with
    one sample per subdomain, and the unbiasedness check over `--seeds` seeds
    on a query sample, against classic BH and the FP64 direct sum."""
    import numpy as np
    import torch

    from sbh_amd import ffi, synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n, depth = args.n, args.o8_depth
    p = synthetic.plummer(n, 2000)
    m = np.full(n, 1.0 / n, np.float32)
    P = torch.tensor(p, device=dev)
    M = torch.tensor(m, device=dev)
    Z = torch.zeros_like(P)
    soft = dict(softening_scale=args.eps, cutoff_scale=float("inf"), prune_enabled=False)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return r, e0.elapsed_time(e1) / reps

    (leaf, node), build_ms = timed(lambda: ffi.build_octree8_buffers(P, Z, M, max_depth=depth),
                                   args.steps)
    cl, cl_ms = timed(lambda: ffi.softened_barnes_hut_force_octree8(
        leaf, node, P, beta=args.beta, **soft), max(1, args.steps // 2))
    st, st_ms = timed(lambda: ffi.softened_stochastic_barnes_hut_force_octree8(
        leaf, node, P, beta=args.beta, samples_per_subdomain=1, seed=1, **soft), args.steps)

    # unbiasedness: `seeds` independent estimates on a query sample
    rng = np.random.default_rng(7)
    idx = torch.tensor(np.sort(rng.choice(n, args.stoch_queries, replace=False)), device=dev)
    Qs = P[idx].contiguous()
    runs = torch.stack([ffi.softened_stochastic_barnes_hut_force_octree8(
        leaf, node, Qs, beta=args.beta, samples_per_subdomain=1, seed=s, **soft).double()
        for s in range(args.seeds)])
    mean = runs.mean(0)
    se = runs.std(0) / args.seeds ** 0.5
    ref = direct_fp64(P.double(), M.double(), Qs.double(), args.eps)
    clq = cl[idx].double()

    def ztest(target):
        z = ((mean - target) / se.clamp(min=1e-30)).flatten()
        # per component z, and the l2 norm of (mean - target) over that of the
        # standard error (about 1 for an unbiased estimate of `target`)
        return {"frac_abs_z_lt_3": float((z.abs() < 3).double().mean()),
                "mean_z2": float((z * z).mean()), "max_abs_z": float(z.abs().max()),
                "l2_dev_over_l2_se": float(torch.linalg.norm(mean - target) /
                                           torch.linalg.norm(se))}

    # the reference's level-3 (Tree3DNode<1,3>) targets on the same points:
    # build, and the warp-voted softened walk (interior queries sum every
    # point exactly, so it is a direct sum in cost), timed on enough queries
    # to fill the chip, accuracy on the query sample
    (l3leaf, l3node), l3_build_ms = timed(
        lambda: ffi.build_octree_buffers(P, Z, M), args.steps)
    soft = dict(beta=args.beta, softening_scale=args.eps, cutoff_scale=float("inf"))
    Ql3 = P[: min(n, args.l3_queries)].contiguous()
    _, l3_walk_ms = timed(lambda: ffi.softened_barnes_hut_force(l3leaf, l3node, Ql3, **soft), 3)
    l3f = ffi.softened_barnes_hut_force(l3leaf, l3node, Qs, **soft)
    level3 = {"build_ms": l3_build_ms, "walk_ms": l3_walk_ms, "walk_queries": int(Ql3.shape[0]),
              "walk_pairs_per_s": Ql3.shape[0] * n / (l3_walk_ms * 1e-3),
              "rel_l2_vs_fp64_direct": rel_l2(l3f, ref),
              "accuracy_queries": int(Qs.shape[0])}

    acc = {"classic_rel_l2_vs_fp64_direct": rel_l2(clq, ref),
           "stochastic_one_seed_rel_l2_vs_fp64_direct": rel_l2(st[idx], ref),
           "stochastic_mean_rel_l2_vs_fp64_direct": rel_l2(mean, ref),
           "unbiasedness_vs_classic_bh": ztest(clq),
           "unbiasedness_vs_fp64_direct": ztest(ref),
           "seeds": args.seeds, "sample_queries": args.stoch_queries}

    cpu = cpu_ref = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        lb, nb = leaf.cpu().numpy(), node.cpu().numpy()
        ns = 4096
        qs = p[rng.choice(n, ns, replace=False)]
        t0 = time.perf_counter()
        O.sbh_octree8(lb, nb, qs, O.KIND_SOFT, 1, 1, args.beta, eps=args.eps, rc=float("inf"))
        t = (time.perf_counter() - t0) / ns
        cpu = {"value": 1.0 / t, "unit": "forces/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"oracle octree8 stochastic walk (same sample paths, fp64, OpenMP "
                         f"over queries) on {ns} of {n} queries, scaled"}
        T = O.CpuTree(p, m, leaf_size=args.leaf, dim_split_log=1, max_depth=10)
        ns = 1 << 16
        t0 = time.perf_counter()
        T.mlpcv(p[:ns], O.CPU_SOFT_FORCE, samples=1, seed=1, eps=args.eps)
        t = (time.perf_counter() - t0) / ns
        cpu_ref = {"value": 1.0 / t, "unit": "forces/s (stochastic, 1 sample)",
                   "threads": O.num_threads(),
                   "variant": f"reference CPU MLPCV (src/sbh.cpp) restated over Tree3D<"
                              f"{args.leaf},1>, {ns} queries, tree build excluded"}
    return {
        "metric": "stochastic BH forces/sec at N=1M (octree8, 1 sample per subdomain)",
        "value": n / (st_ms * 1e-3),
        "unit": "forces/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": st_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plummer_sphere_n{n}_octree8_depth{depth}_eps{args.eps}"
                               f"_beta{args.beta}_stochastic_s1",
                   "n": n, "max_depth": depth, "beta": args.beta, "softening": args.eps,
                   "samples_per_subdomain": 1},
        "phases_ms": {"octree8_build": build_ms, "classic_walk": cl_ms,
                      "stochastic_walk": st_ms},
        "level3": level3,
        "accuracy": acc,
        "cpu_baseline": cpu,
        "cpu_reference_path": cpu_ref,
    }


def treepm_particles(n1d, seed=6000):
    """n1d^3 particles in a periodic box of n1d grid cells: a displaced grid
    (sigma 0.3 cells) with 30% of the particles moved into 2000 Gaussian
    halos (sigma 0.5 cells); masses 1."""
    import numpy as np

    rng = np.random.default_rng(seed)
    g = np.arange(n1d, dtype=np.float64)
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    p += rng.normal(0, 0.3, p.shape)
    k = int(0.3 * p.shape[0])
    idx = rng.choice(p.shape[0], k, replace=False)
    centres = rng.uniform(0, n1d, (2000, 3))
    p[idx] = centres[rng.integers(0, 2000, k)] + rng.normal(0, 0.5, (k, 3))
    p = np.mod(p, n1d).astype(np.float32)
    p[p >= n1d] = 0.0
    return p, np.ones(p.shape[0], np.float32)


def treepm_direct_fp64(P, M, Q, eps, rc, box, chunk=8):
    """FP64 min-image direct sum of the erfc-attenuated softened force
    (src/sbh_ffi.cu:142-160, 431-460)."""
    import math

    import torch

    out = []
    for q in Q.split(chunk):
        d = q[:, None, :] - P[None]
        d = d - box * torch.round(d / box)
        r2 = (d * d).sum(-1)
        s3 = (r2 + eps * eps + 1e-12) ** -1.5
        u = torch.sqrt(r2 + 1e-12) / (2.0 * rc)
        att = torch.special.erfc(u) + (2.0 / math.sqrt(math.pi)) * u * torch.exp(-u * u)
        out.append(-((M[None] * att * s3)[..., None] * d).sum(1))
    return torch.cat(out)


def bench_treepm(args):
    """Short-range TreePM force (the cosmology use of the octree8 path,
    python/cosmology/treepm_jaxpm_demo.py:51-79): periodic box, erfc cutoff
    at the PM split scale, AABB prune, Plummer softening."""
    import numpy as np
    import torch

    from sbh_amd import ffi

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n1d = args.treepm_n1d
    box = float(n1d)
    p, m = treepm_particles(n1d)
    n = p.shape[0]
    P, M = torch.tensor(p, device=dev), torch.tensor(m, device=dev)
    Z = torch.zeros_like(P)
    lo = torch.zeros(3, device=dev)
    hi = torch.full((3,), box, device=dev)
    kw = dict(beta=1.7, softening_scale=0.1, cutoff_scale=1.0, prune_enabled=True,
              prune_r_cut_mult=4.0, periodic=True, box_length=box)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return r, e0.elapsed_time(e1) / reps

    depth = args.treepm_depth
    (leaf, node), build_ms = timed(
        lambda: ffi.build_octree8_buffers(P, Z, M, lo, hi, max_depth=depth), args.steps)
    F, walk_ms = timed(lambda: ffi.softened_barnes_hut_force_octree8(leaf, node, P, **kw),
                       args.steps)
    C = torch.tensor(np.random.default_rng(6001).normal(size=(n, 3)).astype(np.float32),
                     device=dev)
    _, vjp_ms = timed(lambda: ffi.softened_barnes_hut_force_octree8_vjp(leaf, node, P, C, **kw),
                      max(1, args.steps // 2))
    ms_step = build_ms + walk_ms
    idx = torch.tensor(np.random.default_rng(7).choice(n, args.accuracy_queries, replace=False),
                       device=dev)
    ref = treepm_direct_fp64(P.double(), M.double(), P[idx].double(), 0.1, 1.0, box)
    acc = {"force_rel_l2_vs_fp64_direct": rel_l2(F[idx], ref),
           "sample_queries": args.accuracy_queries}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        lb, nb = leaf.cpu().numpy(), node.cpu().numpy()
        ns = 1 << 19
        qs = p[np.random.default_rng(8).choice(n, ns, replace=False)]
        t0 = time.perf_counter()
        O.bh_octree8(lb, nb, qs, O.KIND_SOFT, 1.7, eps=0.1, rc=1.0, prune_enabled=True,
                     prune_r_cut_mult=4.0, periodic=True, box=box)
        t = (time.perf_counter() - t0) / ns
        cpu = {"value": 1.0 / t, "unit": "forces/s (walk only)", "cores": O.num_threads(),
               "kind": "port",
               "sample": f"oracle octree8 walk (same MAC, prune and min-image, fp64, OpenMP "
                         f"over queries) on {ns} of {n} queries, scaled"}
    return {
        "metric": "short-range TreePM forces/sec (octree8, periodic, erfc cutoff, prune)",
        "value": n / (ms_step * 1e-3),
        "unit": "forces/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"treepm_short_range_n{n1d}^3_box{n1d}_depth{depth}_rsplit1",
                   "n": n, "box_length": box, "max_depth": depth, "beta": 1.7,
                   "softening": 0.1, "cutoff_scale": 1.0, "prune_r_cut_mult": 4.0,
                   "step": "octree8 build + softened periodic walk (forward)"},
        "phases_ms": {"octree8_build": build_ms, "walk": walk_ms, "vjp_walk": vjp_ms},
        "accuracy": acc,
        "roofline": None,
        "cpu_baseline": cpu,
    }


def direct2_fp64(P, M, Q, chunk=64):
    """FP64 2-D direct sum of gravityGrad<2> (include/kernel.h:89-94)."""
    import torch

    out = []
    for q in Q.split(chunk):
        d = q[:, None, :] - P[None]
        r = torch.sqrt((d * d).sum(-1))
        out.append((-(M[None] / (r + 0.1) ** 3)[..., None] * d).sum(1))
    return torch.cat(out)


def bench_quadtree(args):
    """2-D targets over Tree2D<1,3> buffers (src/sbh_ffi.cu:1422-1680,
    2444-3064): the tree the reference builds on the host (sbh.Quadtree3,
    include/tree2d.h:237-407), here built on the device, then the warp-voted
    force walk; the reference's 2-D example (python/example_jax_bh_compare.py)
    at 2^20 points: uniform in [-1, 1]^2, masses 1/N, queries = the points in
    input order, beta 2, 256 threads."""
    import numpy as np
    import torch

    from sbh_amd import ffi

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n = args.qt_n
    rng = np.random.default_rng(7000)
    p = rng.uniform(-1.0, 1.0, (n, 2)).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    P, M = torch.tensor(p, device=dev), torch.tensor(m, device=dev)
    Z = torch.zeros_like(P)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        # the build synchronizes once per level: host time is the honest clock
        return r, max(e0.elapsed_time(e1), (time.perf_counter() - t0) * 1e3) / reps

    (leaf, node, contrib), build_ms = timed(
        lambda: ffi.build_quadtree_buffers_with_contrib(P, Z, M), args.steps)
    nodes = node.numel() // 616
    F, walk_ms = timed(lambda: ffi.barnes_hut_force(leaf, node, P, beta=2.0), args.steps)
    _, pot_ms = timed(lambda: ffi.barnes_hut_gravity(leaf, node, P, beta=2.0, threads=64),
                      args.steps)
    ns = min(n, 1 << 16)
    _, stoch_ms = timed(lambda: ffi.stochastic_barnes_hut_force(
        leaf, node, contrib, P[:ns], samples_per_subdomain=1, seed=0), max(1, args.steps // 2))
    ms_step = build_ms + walk_ms
    idx = torch.tensor(np.random.default_rng(7001).choice(n, args.accuracy_queries,
                                                          replace=False), device=dev)
    ref = direct2_fp64(P.double(), M.double(), P[idx].double())
    acc = {"force_rel_l2_vs_fp64_direct": rel_l2(F[idx], ref),
           "sample_queries": args.accuracy_queries}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        t0 = time.perf_counter()
        ol, on, oc = O.nd_build(p, np.zeros_like(p), m, *[x.cpu().numpy() for x in
                                                           ffi._default_corners(P, None, None)])
        cpu_build = time.perf_counter() - t0
        nq = 1 << 12
        t0 = time.perf_counter()
        O.nd_eval(ol, on, oc, p[:nq], O.L3_FORCE, beta=2.0, threads=256)
        t_walk = (time.perf_counter() - t0) / nq
        cpu = {"value": n / (cpu_build + t_walk * n), "unit": "forces/s (build + walk)",
               "cores": 1, "kind": "port",
               "sample": f"oracle Tree2D<1,3> build (the reference's host builder restated, "
                         f"{cpu_build:.2f} s for all {n} points) + oracle vote walk (fp64, "
                         f"same warps) on the first {nq} queries, scaled; one thread"}
    return {
        "metric": "2-D BH forces/sec (Tree2D<1,3> quadtree: build + warp-voted walk)",
        "value": n / (ms_step * 1e-3),
        "unit": "forces/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"quadtree_uniform_square_n{n}_beta2_threads256",
                   "n": n, "nodes": nodes, "beta": 2.0, "threads": 256,
                   "step": "Tree2D<1,3> device build + gravityGrad<2> vote walk"},
        "phases_ms": {"tree_build": build_ms, "force_walk": walk_ms,
                      "potential_walk_threads64": pot_ms,
                      f"stochastic_force_walk_{ns}_queries_1_sample": stoch_ms},
        "accuracy": acc,
        "roofline": None,
        "cpu_baseline": cpu,
    }


def bench_fields(args):
    """The reference's tree fields module (python/jax_sbh.py over the pybind11
    GPU module, src/sbh_gpu.cu) with its defaults: build_gpu_tree(level=1), a
    Tree3D<1,1> (a true octree, leaf size 1), then eval_barnes_hut_gravity
    (beta 2, 32-query blocks) for every point, plus the winding-number and
    smooth-distance walks, the sampler and the compensated brute force on a
    sample.  2^20 points uniform in [-1, 1]^3, masses 1/N, unit normals."""
    import numpy as np
    import torch

    from sbh_amd import jax_sbh as J

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n = args.qt_n
    level = args.fields_level
    rng = np.random.default_rng(8000)
    p = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    nv = rng.normal(size=(n, 3))
    nv = (nv / np.linalg.norm(nv, axis=1, keepdims=True)).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    P, NV, M = (torch.tensor(x, device=dev) for x in (p, nv, m))

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return r, max(e0.elapsed_time(e1), (time.perf_counter() - t0) * 1e3) / reps

    tree, build_ms = timed(lambda: J.build_gpu_tree(P, NV, M, level=level), args.steps)
    nodes = tree.node.numel() // {1: 184, 2: 632, 3: 4272}[level]
    phi, walk_ms = timed(lambda: J.eval_barnes_hut_gravity(tree, P), args.steps)
    _, wn_ms = timed(lambda: J.eval_barnes_hut_wn(tree, P), max(1, args.steps // 2))
    _, sd_ms = timed(lambda: J.eval_barnes_hut_smoothdist(tree, P), max(1, args.steps // 2))
    ns = min(n, 1 << 16)
    _, mc_ms = timed(lambda: J.eval_mlpcv_gravity(tree, P[:ns]), max(1, args.steps // 2))
    # one thread per query over all points, as the reference's module: time it
    # on enough queries to fill the chip, check accuracy on the first 4096
    nbt = min(n, 1 << 16)
    bf, bf_ms = timed(lambda: J.eval_brute_force_gravity(tree, P[:nbt]), 1)
    nb = min(n, 4096)
    ms_step = build_ms + walk_ms
    acc = {"potential_rel_l2_vs_compensated_brute_force": rel_l2(phi[:nb], bf[:nb]),
           "sample_queries": nb}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        lo = (p.min(0) - np.float32(1e-3)).astype(np.float32)
        hi = (p.max(0) + np.float32(1e-3)).astype(np.float32)
        t0 = time.perf_counter()
        ol, on, oc = O.nd_build(p, nv, m, lo, hi, level=level)
        cpu_build = time.perf_counter() - t0
        nq = 1 << 12
        t0 = time.perf_counter()
        O.nd_eval(ol, on, oc, p[:nq], O.L3_POTENTIAL, beta=2.0, threads=32, level=level)
        t_walk = (time.perf_counter() - t0) / nq
        cpu = {"value": n / (cpu_build + t_walk * n), "unit": "potentials/s (build + walk)",
               "cores": 1, "kind": "port",
               "sample": f"oracle Tree3D<1,{level}> build (the reference's host builder "
                         f"restated, {cpu_build:.2f} s for all {n} points) + oracle vote walk "
                         f"(fp64, same 32-query warps) on the first {nq} queries, scaled; one "
                         f"thread"}
    return {
        "metric": "tree-module BH potentials/sec (Tree3D<1,L>: build + warp-voted walk)",
        "value": n / (ms_step * 1e-3),
        "unit": "potentials/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"jax_sbh_uniform_cube_n{n}_level{level}_beta2_threads32",
                   "n": n, "level": level, "nodes": nodes, "beta": 2.0, "threads": 32,
                   "step": "Tree3D<1,L> device build + gravityPotential vote walk"},
        "phases_ms": {"tree_build": build_ms, "potential_walk": walk_ms,
                      "winding_number_walk": wn_ms, "smooth_distance_walk": sd_ms,
                      f"sampler_{ns}_queries_1_sample": mc_ms,
                      f"brute_force_{nbt}_queries": bf_ms},
        "accuracy": acc,
        "roofline": None,
        "cpu_baseline": cpu,
    }


def batched_sets(num_sets, seed=5000):
    """The 64-set workload: sizes log-uniform in [1,024, 524,288], blob
    mixtures (sbh_amd.synthetic.blob_sets)."""
    from sbh_amd import synthetic

    return synthetic.blob_sets(num_sets, seed)


def measure_batched(args, dev, tune=True, cpu=False):
    """The 64-set batched fwd + VJP workload: one forest pipeline captured as
    a HIP graph against the 64 per-set pipelines, both eager and captured as
    one graph.  Returns a record (times, speed-ups, accuracy, walk roofline)."""
    import numpy as np
    import torch

    import sbh_amd
    from sbh_amd import autotune
    from sbh_amd import gravity as G
    from sbh_amd.graph import Graph

    sizes, ps, ms, cs = batched_sets(args.sets)
    P = torch.tensor(np.concatenate(ps), device=dev)
    M = torch.tensor(np.concatenate(ms), device=dev)
    C = torch.tensor(np.concatenate(cs), device=dev)
    offs = G.set_offsets(sizes, dev)
    n_total = int(P.shape[0])
    # the autotuner picks both expansion orders with the leaf size, grid and
    # block size
    base = dict(beta=args.beta, softening_scale=args.eps, max_depth=args.depth)
    if tune:
        tuned = autotune.tune_batch(P, M, C, offs, reps=2, on_fail="relax", **base)
    else:
        tuned = {"fwd_order": args.fwd_order, "vjp_order": args.order, "leaf_size": args.leaf,
                 "walk_blocks": 0, "walk_threads": 0, "table": None}
    fwd_order, vjp_order = tuned["fwd_order"], tuned["vjp_order"]
    kw = dict(base, multipole_order=vjp_order)
    kw_f = dict(base, multipole_order=fwd_order)
    kw_t = dict(kw, fwd_multipole_order=fwd_order)  # the autotuner's step
    leaf, blocks, threads = tuned["leaf_size"], tuned["walk_blocks"], tuned["walk_threads"]
    stream = torch.cuda.current_stream()
    steps = max(1, args.steps)

    def timed(fn, nsteps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(nsteps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / nsteps

    # batched: one forest pipeline, captured as a HIP graph
    step, (F, gx, gm), ws = autotune.build_step(P, M, C, offs, leaf, blocks, kw_t, threads)
    step()
    g = Graph()
    g.capture(step)
    ms_graph = timed(g.replay, steps, max(1, args.warmup))
    ms_eager = timed(step, steps, max(1, args.warmup))
    # walk-kernel times of the batched step (library HIP events, eager calls)
    lib = sbh_amd.load_library()
    lib.sbh_profile_enable(ctypes.c_int32(1))
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    lib.sbh_profile_enable(ctypes.c_int32(0))
    prof = (ctypes.c_double * 6)()
    lib.sbh_profile_collect(prof, ctypes.c_int32(6))
    # per set: the single-set pipeline called once per set on one stream, with
    # the same leaf size, eagerly and captured as one graph of all 64 calls
    bounds = list(zip(offs[:-1].tolist(), offs[1:].tolist()))
    Ps = [P[o0:o1] for o0, o1 in bounds]
    Ms = [M[o0:o1] for o0, o1 in bounds]
    Cs = [C[o0:o1] for o0, o1 in bounds]
    wss = G.new_workspace(int(sizes.max()), args.depth, leaf, dev)
    Fs = [torch.empty_like(x) for x in Ps]
    GXs = [torch.empty_like(x) for x in Ps]
    GMs = [torch.empty_like(x) for x in Ms]

    def per_set():
        for k in range(len(Ps)):
            G.gravity_force(Ps[k], Ms[k], leaf_size=leaf, workspace=wss, out=Fs[k], **kw_f)
            G.gravity_force_vjp(Ps[k], Ms[k], Cs[k], leaf_size=leaf, workspace=wss,
                                grad_points_out=GXs[k], grad_masses_out=GMs[k], **kw)

    ms_per_set = timed(per_set, max(1, steps // 2), 1)
    gp = Graph()
    gp.capture(per_set)
    ms_per_set_graph = timed(gp.replay, max(1, steps // 2), 1)
    # parity of the batch with the per-set calls: forces bit-exact, gradients
    # to float-atomic order
    g.replay()
    gp.replay()
    torch.cuda.synchronize()
    same = all(torch.equal(F[o0:o1], Fs[k]) for k, (o0, o1) in enumerate(bounds))
    gx_err = max(rel_l2(gx[o0:o1], GXs[k]) for k, (o0, o1) in enumerate(bounds))
    # accuracy of the largest set against the FP64 direct sum and adjoint
    kbig = int(np.argmax(sizes))
    o0 = int(offs[kbig])
    nb = int(sizes[kbig])
    idx = torch.tensor(np.random.default_rng(7).choice(nb, args.accuracy_queries,
                                                       replace=False), device=dev)
    Pb, Mb, Cb = Ps[kbig].double(), Ms[kbig].double(), Cs[kbig].double()
    ref = direct_fp64(Pb, Mb, Pb[idx], args.eps)
    rows = idx[:args.grad_rows]
    rgx, rgm = direct_adjoint_rows_fp64(Pb, Mb, Cb, rows, args.eps)
    st = G.new_stats(dev)
    G.gravity_force_vjp_batch(P, M, C, offs, stats=st, leaf_size=leaf, walk_blocks=blocks,
                              walk_threads=threads,
                              workspace=ws, grad_points_out=gx, grad_masses_out=gm, **kw)
    sv = G.stats_dict(st)
    stf = G.new_stats(dev)
    G.gravity_force_batch(P, M, offs, stats=stf, leaf_size=leaf, walk_blocks=blocks,
                          walk_threads=threads,
                          workspace=ws, out=F, **kw_f)
    sf = G.stats_dict(stf)
    torch.cuda.synchronize()
    acc = {"largest_set": nb,
           "largest_set_force_rel_l2_vs_fp64_direct": rel_l2(F[o0:o0 + nb][idx], ref),
           "largest_set_force_outlier_frac": outlier_frac(F[o0:o0 + nb][idx], ref),
           "largest_set_grad_points_rel_l2_vs_fp64_direct_adjoint":
               rel_l2(gx[o0:o0 + nb][rows], rgx),
           "largest_set_grad_masses_rel_l2_vs_fp64_direct_adjoint":
               rel_l2(gm[o0:o0 + nb][rows], rgm)}
    g.close()
    gp.close()
    walk_vjp_ms = prof[5] / max(prof[3], 1)
    walk_fwd_ms = prof[2] / max(prof[0], 1)
    flops_vjp, flops_vjp_impl = walk_flops(sv, True, vjp_order)
    rec = {"workload": f"blob_sets{args.sets}_total{n_total}_theta0.5_fwd+vjp",
           "sets": args.sets, "total_points": n_total, "min_set": int(sizes.min()),
           "max_set": int(sizes.max()), "softening": args.eps, "beta_diag": args.beta,
           "max_depth": args.depth,
           "multipole_order": {"forward": fwd_order, "vjp": vjp_order},
           "leaf_size": leaf, "walk_blocks": blocks, "walk_threads": threads,
           "batched_graph_ms": ms_graph, "batched_eager_ms": ms_eager,
           "per_set_eager_ms": ms_per_set, "per_set_graph_ms": ms_per_set_graph,
           "speedup_vs_per_set_graph": ms_per_set_graph / ms_graph,
           "speedup_vs_per_set_eager": ms_per_set / ms_graph,
           "forces_per_s": n_total / (ms_graph * 1e-3),
           "batch_forces_equal_per_set": bool(same),
           "batch_grad_points_max_rel_l2_vs_per_set": gx_err,
           "autotune": tuned.get("table"),
           "autotune_gate": tuned.get("gate"),
           "interactions_vjp": sv, "interactions_fwd": sf,
           "accuracy": acc,
           "phases_ms": {"fwd_build": prof[1] / max(prof[0], 1), "fwd_walk_kernel": walk_fwd_ms,
                         "vjp_build": prof[4] / max(prof[3], 1),
                         "vjp_walk_kernel": walk_vjp_ms},
           "roofline": valu_roofline(
               f"k_walk<vjp=1,order={vjp_order},forces=0> (all sets)", flops_vjp,
               flops_vjp_impl, walk_vjp_ms)}
    if cpu:
        big_args = argparse.Namespace(**vars(args))
        big_args.leaf = leaf
        big_args.order, big_args.fwd_order = vjp_order, fwd_order
        c = cpu_baseline(ps[kbig], ms[kbig], cs[kbig], big_args, args.cpu_sample_groups)
        c["sample"] = (f"largest set ({nb} points) only; its forces/s stands for the batch. "
                       + c["sample"])
        rec["cpu_baseline"] = c
    return rec


def bench_batched(args):
    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    b = measure_batched(args, dev, tune=not args.no_autotune, cpu=not args.no_cpu_baseline)
    return {
        "metric": "batched fwd+VJP forces/sec, 64 sets (graph-captured forest vs per-set "
                  "calls)",
        "value": b["forces_per_s"],
        "unit": "forces/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": b["batched_graph_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {k: b[k] for k in ("workload", "sets", "total_points", "min_set", "max_set",
                                     "softening", "beta_diag", "max_depth", "multipole_order",
                                     "leaf_size", "walk_blocks", "walk_threads")},
        "batched": {k: v for k, v in b.items() if k not in ("roofline", "cpu_baseline")},
        "roofline": b["roofline"],
        "cpu_baseline": b.get("cpu_baseline"),
    }


def uniform_cube(n, seed=1000):
    """N points uniform in [0, 1)^3 (fp32) and equal masses 1/N."""
    import numpy as np

    p = np.random.default_rng(seed).uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    return p, np.full(n, 1.0 / n, np.float32)


def measure_config1(args, dev, n=10000, eps=0.01, leaf=16, steps=20):
    """Plain classic BH at N = 10,000 uniform in [0, 1)^3, masses 1/N,
    softening 0.01, theta = 0.5, leaf size 16: the GPU pipeline (forces, and
    forces + full VJP) against the FP64 direct sum (rel_l2, outlier fraction)
    and the reference benchmark's FP32 numpy direct sum, with the CPU legs
    beside it: that numpy direct sum over all N queries and the reference's
    CPU tree path (src/sbh.cpp) restated, as it ships and like-for-like."""
    import numpy as np
    import torch

    from sbh_amd import gravity as G
    from sbh_amd.graph import Graph

    p, m = uniform_cube(n)
    c = np.random.default_rng(1002).normal(size=(n, 3)).astype(np.float32)
    P, M, C = (torch.tensor(a, device=dev) for a in (p, m, c))
    kw = dict(beta=args.beta, softening_scale=eps, max_depth=10, leaf_size=leaf,
              multipole_order=args.order)
    kw_f = dict(kw, multipole_order=args.fwd_order)
    ws = G.new_workspace(n, 10, leaf, dev)
    F, GX, GM = torch.empty_like(P), torch.empty_like(P), torch.empty_like(M)
    stream = torch.cuda.current_stream()

    def fwd():
        G.gravity_force(P, M, workspace=ws, out=F, **kw_f)

    def fwd_vjp():
        fwd()
        G.gravity_force_vjp(P, M, C, workspace=ws, grad_points_out=GX, grad_masses_out=GM, **kw)

    out = {}
    for tag, fn in (("fwd", fwd), ("fwd_vjp", fwd_vjp)):
        fn()
        g = Graph()
        g.capture(fn)
        for _ in range(3):
            g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            g.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        out[tag + "_ms"] = e0.elapsed_time(e1) / steps
        g.close()
    fwd_vjp()
    torch.cuda.synchronize()
    ref = direct_fp64(P.double(), M.double(), P.double(), eps)
    rgx, rgm = direct_adjoint_rows_fp64(P.double(), M.double(), C.double(),
                                        torch.arange(n, device=dev), eps, chunk=64)
    # the reference benchmark's oracle: its FP32 numpy direct sum, timed in full
    t0 = time.perf_counter()
    ref_np = numpy_direct_softcut(p, m, p, eps)
    t_np = time.perf_counter() - t0
    ref_np_t = torch.tensor(ref_np, device=dev)
    from oracle import oracle as O

    cpu_args = argparse.Namespace(leaf=leaf, beta=args.beta, eps=eps, depth=100)
    lit = cpu_reference_literal(p, m, cpu_args, nsample=n)
    lfl = cpu_reference_path(p, m, cpu_args, np.arange(n), ref.cpu().numpy(), nsample=n)
    out.update({
        "workload": f"uniform_cube_n{n}_eps{eps}_theta0.5_leaf{leaf}",
        "n": n, "softening": eps, "leaf_size": leaf, "beta_diag": args.beta,
        "multipole_order": {"forward": args.fwd_order, "vjp": args.order},
        "fwd_forces_per_s": n / (out["fwd_ms"] * 1e-3),
        "fwd_vjp_forces_per_s": n / (out["fwd_vjp_ms"] * 1e-3),
        "launch": "HIP graph replay",
        "force_rel_l2_vs_fp64_direct": rel_l2(F, ref),
        "force_outlier_frac_vs_fp64_direct": outlier_frac(F, ref),
        "force_rel_l2_vs_fp32_numpy_direct": rel_l2(F, ref_np_t),
        "force_outlier_frac_vs_fp32_numpy_direct": outlier_frac(F, ref_np_t),
        "fp32_numpy_direct_rel_l2_vs_fp64": rel_l2(ref_np_t, ref),
        "grad_points_rel_l2_vs_fp64_direct_adjoint": rel_l2(GX, rgx),
        "grad_masses_rel_l2_vs_fp64_direct_adjoint": rel_l2(GM, rgm),
        "cpu_numpy_direct": {"value": n / t_np, "unit": "forces/s (forward only)", "cores": 1,
                             "seconds": round(t_np, 3),
                             "sample": f"all {n} queries, fp32 numpy with the attenuation "
                                       "(python/unified_force_benchmark.py:208-230)"},
        "cpu_reference_literal": lit,
        "cpu_reference_like_for_like": lfl,
        "host_threads_used": O.num_threads(),
    })
    return out


def numpy_direct_softcut(points, masses, queries, eps, cutoff=float("inf")):
    """The reference benchmark's numpy direct sum (python/unified_force_benchmark
    .py:208-230): fp32, one query at a time, attenuation erfc(u) + 2u/sqrt(pi)
    exp(-u^2) with u = r / (2 r_c) (1 at r_c = inf), non-periodic."""
    import numpy as np
    from scipy import special

    out = np.empty((queries.shape[0], 3), dtype=np.float32)
    eps2 = float(eps) * float(eps)
    for i in range(queries.shape[0]):
        disp = queries[i][None, :] - points
        r2 = np.sum(disp * disp, axis=1)
        inv_r = 1.0 / np.sqrt(r2 + eps2 + 1e-12)
        inv_r3 = inv_r * inv_r * inv_r
        dist = np.sqrt(r2 + 1e-12)
        u = dist / (2.0 * cutoff)
        att = (special.erfc(u) + (2.0 / np.sqrt(np.pi)) * u * np.exp(-(u * u))).astype(
            np.float32)
        coeff = -masses * att * inv_r3
        out[i] = np.sum(coeff[:, None] * disp, axis=0, dtype=np.float32)
    return out


def bench_vjp_blobs(args):
    """Custom-VJP path at 2^18 points in 64 Gaussian blobs (sigma 0.02,
    log-normal(0, 0.5) masses normalised to 1, eps = 0.002, cotangent N(0, 1)):
    forward through `gravity_force_recompute_vjp` and backward through
    torch.autograd (rebuild + adjoint walk), as the reference's recompute test
    python/test_jax_custom_vjp_recompute.py:105-151 drives its custom VJP.
    Forces and both gradients against the FP64 direct sum and adjoint on
    sampled rows; the GPU direct sum fwd + full VJP timed beside."""
    import copy

    import numpy as np
    import torch

    from sbh_amd import ffi
    from sbh_amd import gravity as G
    from sbh_amd import synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    a = copy.copy(args)
    a.eps = 0.002
    n = 1 << 18
    p, m = synthetic.blobs(n, 64, 3000)
    c = np.random.default_rng(3002).normal(size=(n, 3)).astype(np.float32)
    P = torch.tensor(p, device=dev, requires_grad=True)
    M = torch.tensor(m, device=dev, requires_grad=True)
    C = torch.tensor(c, device=dev)
    ws = G.new_workspace(n, a.depth, a.leaf, dev)
    kw = dict(beta=a.beta, softening_scale=a.eps, max_depth=a.depth, leaf_size=a.leaf,
              multipole_order=a.fwd_order, vjp_multipole_order=a.order, workspace=ws)

    def step():
        F = G.gravity_force_recompute_vjp(P, M, **kw)
        gx, gm = torch.autograd.grad(F, (P, M), grad_outputs=C)
        return F, gx, gm

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        F, gx, gm = step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    rows = torch.tensor(np.random.default_rng(3003).choice(n, args.grad_rows, replace=False),
                        device=dev)
    Pd, Md, Cd = (t.detach().double() for t in (P, M, C))
    ref = direct_fp64(Pd, Md, Pd[rows], a.eps)
    rgx, rgm = direct_adjoint_rows_fp64(Pd, Md, Cd, rows, a.eps)
    acc = {"force_rel_l2_vs_fp64_direct": rel_l2(F[rows], ref),
           "force_outlier_frac_vs_fp64_direct": outlier_frac(F[rows], ref),
           "grad_points_rel_l2_vs_fp64_direct_adjoint": rel_l2(gx[rows], rgx),
           "grad_masses_rel_l2_vs_fp64_direct_adjoint": rel_l2(gm[rows], rgm),
           "grad_points_outlier_frac_vs_fp64_direct_adjoint": outlier_frac(gx[rows], rgx),
           "sample_rows": int(rows.numel())}
    Pn, Mn = P.detach(), M.detach()
    t0, t1, t2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    t0.record(stream)
    ffi.softened_direct_force(Pn, Mn, Pn, softening_scale=a.eps, cutoff_scale=float("inf"))
    t1.record(stream)
    ffi.softened_direct_force_full_vjp(Pn, Mn, C, softening_scale=a.eps,
                                       cutoff_scale=float("inf"))
    t2.record(stream)
    torch.cuda.synchronize()
    d_ms = t0.elapsed_time(t2)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, m, c, a, min(args.cpu_sample_groups, 4000))
    return {
        "metric": "BH forces/sec, custom-VJP path (torch.autograd), N = 2^18 blobs, fwd+VJP",
        "value": n / (ms * 1e-3), "unit": "forces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"blobs64_n{n}_sigma0.02_eps0.002_theta0.5_depth{a.depth}"
                               f"_leaf{a.leaf}_custom_vjp",
                   "n": n, "blobs": 64, "softening": a.eps, "beta_diag": a.beta,
                   "max_depth": a.depth, "leaf_size": a.leaf,
                   "multipole_order": {"forward": a.fwd_order, "vjp": a.order},
                   "launch": "eager, torch.autograd"},
        "accuracy": acc,
        "direct_sum": {"fwd_plus_vjp_ms": d_ms, "fwd_ms": t0.elapsed_time(t1),
                       "speedup_bh_vs_direct": d_ms / ms},
        "cpu_baseline": cpu,
    }


def bench_config1(args):
    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    c = measure_config1(args, dev)
    return {
        "metric": "classic BH forces/sec, N = 10,000 uniform cube (forward; fwd+VJP beside)",
        "value": c["fwd_forces_per_s"], "unit": "forces/s", "n_gpus": 1, "steps": 20,
        "warmup": 3, "ms_per_step": c["fwd_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {k: c[k] for k in ("workload", "n", "softening", "leaf_size", "beta_diag",
                                     "multipole_order")},
        "config1": c,
        "cpu_baseline": {**c["cpu_numpy_direct"], "kind": "port"},
    }


if __name__ == "__main__":
    main()
