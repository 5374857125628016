"""Autotuner for the batched gravity pipeline.

Times the captured fwd + VJP graph of the actual batch for candidate
(forward and adjoint expansion orders, leaf size, persistent walk grid, walk
block size) and keeps the fastest candidate whose forces and gradients stay
within rel_l2 <= 1e-3 of an FP64 direct sum on sampled rows: pooled over all
sets and, separately, on every set sampled with at least 16 rows (rows are
dealt to the sets in proportion to their size, 4 to 64 each, so the largest
sets, where the error is largest, are checked on their own).  The search is
coordinate-wise by default (adjoint orders x leaf sizes, then forward orders,
grids and block sizes at the best so far: 10 + 1 + 2 + 3 captures) or the
full product.  Results are cached per workload key (device, total points,
set sizes, a digest of the per-set bounding boxes and the physics
parameters) for the life of the process, and optionally in a JSON file.
"""
from __future__ import annotations

import hashlib
import json
import os
import warnings
from typing import Iterable, Optional

import torch

from . import gravity as G
from .graph import Graph

_CACHE: dict = {}

DEFAULT_LEAF = (8, 16, 24, 32, 48)
DEFAULT_BLOCKS = (0, 1024, 4096)
DEFAULT_THREADS = (256, 64, 128, 512)  # walk block size; 256 is the library default
# adjoint expansions: the second-moment adjoint (cheaper), then the full quadrupole
DEFAULT_VJP_ORDERS = (G.SECOND_MOMENT_ADJOINT, 2)
# forward expansions: the quadrupole first (the more accurate), then the monopole
DEFAULT_FWD_ORDERS = (2, 1)
REL_L2_GATE = 1e-3
RELAX_FACTOR = 1.25  # on_fail="relax": gate = this x the best accuracy reached
GATE_ROWS = 512        # FP64 rows over the whole batch
GATE_ROWS_SET = (4, 64)  # per set, at least / at most
GATE_SET_MIN_ROWS = 16   # sets sampled with this many rows are also checked alone


def _set_boxes(points: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
    """(num_sets, 6) per-set min xyz, max xyz (empty sets: +inf / -inf)."""
    nsets = offsets.shape[0] - 1
    sizes = (offsets[1:] - offsets[:-1]).long()
    sid = torch.repeat_interleave(torch.arange(nsets, device=points.device), sizes)
    lo = torch.full((nsets, 3), float("inf"), device=points.device)
    hi = torch.full((nsets, 3), float("-inf"), device=points.device)
    idx = sid[:, None].expand(-1, 3)
    lo = lo.scatter_reduce(0, idx, points, reduce="amin")
    hi = hi.scatter_reduce(0, idx, points, reduce="amax")
    return torch.cat([lo, hi], 1)


def _key(points: torch.Tensor, offsets: torch.Tensor, kw: dict) -> str:
    h = hashlib.sha1(offsets.cpu().numpy().tobytes())
    h.update(_set_boxes(points, offsets).cpu().numpy().tobytes())
    dev = torch.cuda.get_device_name(points.device).replace(" ", "_")
    phys = ",".join(f"{k}={kw[k]}" for k in sorted(kw))
    return f"{dev}_n{int(offsets[-1])}_s{offsets.shape[0] - 1}_{h.hexdigest()[:16]}_{phys}"


def direct_rows_fp64(P, M, C, rows, eps):
    """FP64 direct-sum forces and gradients of L = sum_i c_i . F_i at `rows`
    of one point set (softened force, self pair excluded):
    (forces[rows], grad_points[rows], grad_masses[rows])."""
    P, M, C = P.double(), M.double(), C.double()
    e2 = eps * eps + 1e-12
    d = P[rows][:, None, :] - P[None]  # x_j - x_k
    r2 = (d * d).sum(-1) + e2
    s3 = r2 ** -1.5
    s5 = s3 / r2
    self_mask = torch.arange(P.shape[0], device=P.device)[None, :] == rows[:, None]
    s3 = s3.masked_fill(self_mask, 0.0)
    s5 = s5.masked_fill(self_mask, 0.0)
    f = -((M[None, :] * s3)[..., None] * d).sum(1)
    cj = C[rows]
    dc = (d * cj[:, None, :]).sum(-1)
    q = -(M[None, :, None] * (s3[..., None] * cj[:, None, :] - 3.0 * (s5 * dc)[..., None] * d)).sum(1)
    cd = (d * C[None]).sum(-1)
    src = M[rows][:, None] * (s3[..., None] * C[None] - 3.0 * (s5 * cd)[..., None] * d).sum(1)
    return f, q + src, (s3 * cd).sum(1)


def _rel_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    a, b = a.double(), b.double()
    return float(torch.linalg.norm(a - b) / torch.clamp(torch.linalg.norm(b), min=1e-300))


class _Gate:
    """FP64 direct-sum rows of every set with at least 2 points (rows dealt in
    proportion to set size, GATE_ROWS_SET[0] to GATE_ROWS_SET[1] per set) and
    the rel_l2 of a candidate's forces and gradients against them: pooled
    over all rows, and alone for each set with >= GATE_SET_MIN_ROWS rows."""

    def __init__(self, P, M, C, offsets, eps, rows=GATE_ROWS, seed=0):
        g = torch.Generator(device="cpu").manual_seed(seed)
        o = offsets.cpu().tolist()
        total = max(1, o[-1])
        self.sets = []  # (slice of the gathered rows) for the sets checked alone
        rows_all, ref_f, ref_x, ref_m = [], [], [], []
        at = 0
        for o0, o1 in zip(o[:-1], o[1:]):
            n = o1 - o0
            if n < 2:
                continue
            k = int(round(rows * n / total))
            k = min(n, max(GATE_ROWS_SET[0], min(GATE_ROWS_SET[1], k)))
            r = torch.randperm(n, generator=g)[:k].to(P.device)
            f, gx, gm = direct_rows_fp64(P[o0:o1], M[o0:o1], C[o0:o1], r, eps)
            rows_all.append(r + o0)
            ref_f.append(f)
            ref_x.append(gx)
            ref_m.append(gm)
            if k >= GATE_SET_MIN_ROWS:
                self.sets.append(slice(at, at + k))
            at += k
        self.rows = torch.cat(rows_all) if rows_all else None
        if self.rows is not None:
            self.ref = (torch.cat(ref_f), torch.cat(ref_x), torch.cat(ref_m))

    def errors(self, F, gx, gm) -> dict:
        """Per quantity the larger of the pooled rel_l2 and the worst
        per-set rel_l2 (sets with enough rows)."""
        if self.rows is None:
            return {"force": 0.0, "grad_points": 0.0, "grad_masses": 0.0}
        got = (F[self.rows], gx[self.rows], gm[self.rows])
        out = {}
        for key, a, b in zip(("force", "grad_points", "grad_masses"), got, self.ref):
            e = _rel_l2(a, b)
            for sl in self.sets:
                e = max(e, _rel_l2(a[sl], b[sl]))
            out[key] = e
        return out


def build_step(P, M, C, offsets, leaf_size: int, walk_blocks: int, kw: dict,
               walk_threads: int = 0):
    """Preallocated buffers and a closure running one batched fwd + VJP.
    kw["multipole_order"] is the adjoint's expansion; kw["fwd_multipole_order"]
    (optional, default the same) the forward walk's."""
    nsets = offsets.shape[0] - 1
    ws = G.new_batch_workspace(P.shape[0], nsets, kw.get("max_depth", 10), leaf_size)
    F = torch.empty_like(P)
    gx = torch.empty_like(P)
    gm = torch.empty_like(M)
    args = dict(kw, leaf_size=leaf_size, walk_blocks=walk_blocks, walk_threads=walk_threads,
                workspace=ws, check_offsets=False)
    fwd_order = args.pop("fwd_multipole_order", None)
    fwd_args = dict(args) if fwd_order is None else dict(args, multipole_order=fwd_order)

    def step():
        G.gravity_force_batch(P, M, offsets, out=F, **fwd_args)
        G.gravity_force_vjp_batch(P, M, C, offsets, grad_points_out=gx, grad_masses_out=gm,
                                  **args)
    return step, (F, gx, gm), ws


def time_graph(step, reps: int = 3) -> float:
    """Milliseconds per replay of the captured `step`."""
    step()  # eager warm-up (first-touch of the workspace)
    g = Graph()
    g.capture(step)
    g.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        g.replay()
    e1.record(s)
    torch.cuda.synchronize()
    g.close()
    return e0.elapsed_time(e1) / reps


def tune_batch(P, M, C, offsets, *, leaf_sizes: Iterable[int] = DEFAULT_LEAF,
               walk_blocks: Iterable[int] = DEFAULT_BLOCKS,
               walk_threads: Iterable[int] = DEFAULT_THREADS,
               vjp_orders: Optional[Iterable[int]] = None,
               fwd_orders: Optional[Iterable[int]] = None, search: str = "coordinate",
               reps: int = 3, cache_file: Optional[str] = None, gate: float = REL_L2_GATE,
               gate_rows: int = GATE_ROWS, on_fail: str = "raise", **kw) -> dict:
    """Best {fwd_order, vjp_order, leaf_size, walk_blocks, walk_threads, ms}
    for this batch, plus the table of every candidate timed.

    Candidates whose forces or gradients miss rel_l2 <= `gate` against the
    FP64 direct sum on `gate_rows` sampled rows (see _Gate) are dropped.  If
    none of the first stage passes: on_fail="raise" raises RuntimeError,
    on_fail="relax" sets the gate to RELAX_FACTOR x the smallest error any
    candidate reached (the timings are kept) and says so in the result's
    "gate".  vjp_orders / fwd_orders: adjoint and forward expansions to try
    (defaults DEFAULT_VJP_ORDERS / DEFAULT_FWD_ORDERS, or kw["multipole_order"]
    / kw["fwd_multipole_order"] alone when given).  search: "coordinate"
    (adjoint order x leaf size at the first forward order, then the other
    forward orders, then grid, then block size, the others at their best so
    far) or "grid" (every combination).  kw: beta, softening_scale,
    max_depth."""
    kw = dict(kw)
    given = kw.pop("multipole_order", None)
    given_fwd = kw.pop("fwd_multipole_order", None)
    if vjp_orders is None:
        vjp_orders = DEFAULT_VJP_ORDERS if given is None else (given,)
    if fwd_orders is None:
        fwd_orders = DEFAULT_FWD_ORDERS if given_fwd is None else (given_fwd,)
    leaf_sizes, walk_blocks, walk_threads = list(leaf_sizes), list(walk_blocks), list(walk_threads)
    vjp_orders, fwd_orders = list(vjp_orders), list(fwd_orders)
    if search not in ("coordinate", "grid"):
        raise ValueError("search must be 'coordinate' or 'grid'")
    if on_fail not in ("raise", "relax"):
        raise ValueError("on_fail must be 'raise' or 'relax'")
    key = (_key(P, offsets, kw) + f"_gate{gate}x{gate_rows}{on_fail}_{search}_o{vjp_orders}"
           f"_f{fwd_orders}_l{leaf_sizes}_b{walk_blocks}_t{walk_threads}").replace(" ", "")
    if key in _CACHE:
        return _CACHE[key]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as fh:
            disk = json.load(fh)
        if key in disk:
            _CACHE[key] = disk[key]
            return disk[key]
    checker = _Gate(P, M, C, offsets, kw.get("softening_scale", 1e-2), gate_rows)
    table = []
    done = {}
    requested = gate

    def run(fwd, order, leaf, blocks, threads):
        cand = (fwd, order, leaf, blocks, threads)
        if cand in done:
            return done[cand]
        step, (F, gx, gm), ws = build_step(
            P, M, C, offsets, leaf, blocks,
            dict(kw, multipole_order=order, fwd_multipole_order=fwd), threads)
        ms = time_graph(step, reps)
        step()
        err = checker.errors(F, gx, gm)
        ok = max(err.values()) <= gate  # the gate in force when timed (see relax below)
        row = {"fwd_order": fwd, "vjp_order": order, "leaf_size": leaf, "walk_blocks": blocks,
               "walk_threads": threads, "ms": ms, "rel_l2": err, "passes_gate": ok}
        table.append(row)
        done[cand] = row
        del ws
        return row

    def best_of(rows):
        passing = [r for r in rows if r["passes_gate"]]
        return min(passing, key=lambda r: r["ms"]) if passing else None

    def relax(rows):
        # no candidate passed: the gate becomes 1.25 x the best accuracy
        # reached (the fastest of the nearly most accurate candidates wins),
        # re-judging the rows already timed
        nonlocal gate
        if best_of(rows) is None and on_fail == "relax" and rows:
            gate = RELAX_FACTOR * min(max(r["rel_l2"].values()) for r in rows)
            for r in table:
                r["passes_gate"] = max(r["rel_l2"].values()) <= gate

    if search == "grid":
        for fwd in fwd_orders:
            for order in vjp_orders:
                for leaf in leaf_sizes:
                    for blocks in walk_blocks:
                        for threads in walk_threads:
                            run(fwd, order, leaf, blocks, threads)
        relax(table)
    else:
        f0, b0, t0 = fwd_orders[0], walk_blocks[0], walk_threads[0]
        first = [run(f0, o, leaf, b0, t0) for o in vjp_orders for leaf in leaf_sizes]
        relax(first)
        cur = best_of(first)
        if cur is not None:
            cur = best_of([run(f, cur["vjp_order"], cur["leaf_size"], b0, t0)
                           for f in fwd_orders])
        if cur is not None:
            cur = best_of([run(cur["fwd_order"], cur["vjp_order"], cur["leaf_size"], b, t0)
                           for b in walk_blocks])
        if cur is not None:
            cur = best_of([run(cur["fwd_order"], cur["vjp_order"], cur["leaf_size"],
                               cur["walk_blocks"], t) for t in walk_threads])
    best = best_of(table)
    if best is None:
        raise RuntimeError(f"no candidate meets rel_l2 <= {gate}: {table}")
    if gate != requested:
        warnings.warn(f"autotune: no candidate met rel_l2 <= {requested}; gate relaxed to "
                      f"{gate}")
    res = {"vjp_order": best["vjp_order"], "fwd_order": best["fwd_order"],
           "leaf_size": best["leaf_size"], "walk_blocks": best["walk_blocks"],
           "walk_threads": best["walk_threads"], "ms": best["ms"], "table": table,
           "search": search,
           "gate": {"rel_l2_max": gate, "requested_rel_l2_max": requested,
                    "relaxed": gate != requested, "rows": 0 if checker.rows is None
                    else int(checker.rows.numel()), "sets_checked_alone": len(checker.sets),
                    "reference": "FP64 direct sum and adjoint over each row's own set"}}
    _CACHE[key] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as fh:
                disk = json.load(fh)
        disk[key] = res
        with open(cache_file, "w") as fh:
            json.dump(disk, fh, indent=1)
    return res
