"""Autotuner for the batched gravity pipeline.

Times the captured fwd + VJP graph of the actual batch for every candidate
(leaf size, persistent walk grid) and keeps the fastest.  Results are cached
per workload key (total points, set count, a digest of the set sizes and the
physics parameters) for the life of the process, and optionally in a JSON
file.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Iterable, Optional

import torch

from . import gravity as G
from .graph import Graph

_CACHE: dict = {}

DEFAULT_LEAF = (8, 16, 24, 32, 48)
DEFAULT_BLOCKS = (0, 1024, 4096)


def _key(offsets: torch.Tensor, kw: dict) -> str:
    h = hashlib.sha1(offsets.cpu().numpy().tobytes()).hexdigest()[:16]
    phys = ",".join(f"{k}={kw[k]}" for k in sorted(kw))
    return f"n{int(offsets[-1])}_s{offsets.shape[0] - 1}_{h}_{phys}"


def build_step(P, M, C, offsets, leaf_size: int, walk_blocks: int, kw: dict):
    """Preallocated buffers and a closure running one batched fwd + VJP."""
    nsets = offsets.shape[0] - 1
    ws = G.new_batch_workspace(P.shape[0], nsets, kw.get("max_depth", 10), leaf_size)
    F = torch.empty_like(P)
    gx = torch.empty_like(P)
    gm = torch.empty_like(M)
    args = dict(kw, leaf_size=leaf_size, walk_blocks=walk_blocks, workspace=ws,
                check_offsets=False)

    def step():
        G.gravity_force_batch(P, M, offsets, out=F, **args)
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
               walk_blocks: Iterable[int] = DEFAULT_BLOCKS, reps: int = 3,
               cache_file: Optional[str] = None, **kw) -> dict:
    """Best {leaf_size, walk_blocks, ms} for this batch, plus the full table.

    kw: beta, softening_scale, max_depth, multipole_order (as gravity_force)."""
    key = _key(offsets, kw)
    if key in _CACHE:
        return _CACHE[key]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as fh:
            disk = json.load(fh)
        if key in disk:
            _CACHE[key] = disk[key]
            return disk[key]
    table = []
    for leaf in leaf_sizes:
        for blocks in walk_blocks:
            step, _, ws = build_step(P, M, C, offsets, leaf, blocks, kw)
            ms = time_graph(step, reps)
            table.append({"leaf_size": leaf, "walk_blocks": blocks, "ms": ms})
            del ws
    best = min(table, key=lambda r: r["ms"])
    res = {"leaf_size": best["leaf_size"], "walk_blocks": best["walk_blocks"],
           "ms": best["ms"], "table": table}
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
