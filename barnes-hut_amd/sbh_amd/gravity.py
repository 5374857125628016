"""Adaptive-octree Barnes-Hut self-gravity with a full adjoint.

Host side of the `sbh_gravity_*` C-ABI targets (include/sbh_amd.h).  The
recompute wrapper mirrors the reference's custom_vjp structure
(python/jax_sbh_ffi.py:1731-1897: residual = inputs, backward rebuilds the
tree and runs the VJP walk) but returns the complete position and mass
gradients instead of the query side only.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._lib import buf, call, f32, i32, stream_handle
from .ffi import _f32

LAYOUT_FIELDS = ("header", "codes", "perm", "sp", "sc", "topo", "prefix", "parent", "geo", "mp",
                 "qd", "adj", "adjq", "qgrad", "sgrad", "total", "cap")

STATS_FIELDS = ("far", "near", "groups", "nodes", "levels", "overflow")

# multipole_order of the VJP functions beyond 1 (monopole) and 2 (monopole +
# quadrupole): monopole terms plus the quadrupole term's adjoint with respect
# to each node's second moment.  Pushed down to the sources it gives their
# position gradients to first and mass gradients to second order in (cell
# size / distance), as 2 does, at about 3/4 of its adjoint work.
SECOND_MOMENT_ADJOINT = 3


def _check(beta, softening_scale, max_depth, leaf_size, multipole_order, vjp=False):
    if beta < 0:
        raise ValueError("beta must be >= 0")
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if not 1 <= max_depth <= 10:
        raise ValueError("max_depth must be in [1, 10]")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    if vjp and multipole_order not in (1, 2, SECOND_MOMENT_ADJOINT):
        raise ValueError("multipole_order must be 1, 2 or 3")
    if not vjp and multipole_order not in (1, 2):
        raise ValueError("multipole_order must be 1 or 2")


def workspace_bytes(n: int, max_depth: int = 10, leaf_size: int = 32) -> int:
    lib = _lib.load_library()
    lib.sbh_gravity_workspace_bytes.restype = ctypes.c_int64
    v = int(lib.sbh_gravity_workspace_bytes(ctypes.c_int64(n), i32(max_depth), i32(leaf_size)))
    if v < 0:
        raise ValueError("invalid workspace query")
    return v


def workspace_layout(n: int, max_depth: int = 10, leaf_size: int = 32) -> dict:
    arr = (ctypes.c_int64 * len(LAYOUT_FIELDS))()
    call("sbh_gravity_workspace_layout", ctypes.c_int64(n), i32(max_depth), i32(leaf_size), arr,
         i32(len(LAYOUT_FIELDS)))
    return dict(zip(LAYOUT_FIELDS, [int(x) for x in arr]))


def new_workspace(n: int, max_depth: int = 10, leaf_size: int = 32,
                  device: Optional[torch.device] = None) -> torch.Tensor:
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.empty(workspace_bytes(n, max_depth, leaf_size), dtype=torch.uint8, device=dev)


def gravity_force(points, masses, *, beta: float = 2.0, softening_scale: float = 1e-2,
                  max_depth: int = 10, leaf_size: int = 32, multipole_order: int = 2,
                  workspace: Optional[torch.Tensor] = None, stats: Optional[torch.Tensor] = None,
                  out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Forces F_i = sum_j -m_j (r^2 + eps^2 + 1e-12)^(-3/2) (x_i - x_j), j != i."""
    _check(beta, softening_scale, max_depth, leaf_size, multipole_order)
    points, masses = _f32(points), _f32(masses)
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    forces = out if out is not None else torch.empty_like(points)
    call("sbh_gravity_force", stream_handle(), buf(points), buf(masses), buf(forces), f32(beta),
         f32(softening_scale), i32(max_depth), i32(leaf_size), i32(multipole_order),
         buf(workspace), buf(stats))
    return forces


def gravity_force_vjp(points, masses, cotangent, *, beta: float = 2.0,
                      softening_scale: float = 1e-2, max_depth: int = 10, leaf_size: int = 32,
                      multipole_order: int = 2, workspace: Optional[torch.Tensor] = None,
                      stats: Optional[torch.Tensor] = None, return_forces: bool = False,
                      grad_points_out: Optional[torch.Tensor] = None,
                      grad_masses_out: Optional[torch.Tensor] = None):
    """(grad_points, grad_masses) of L = sum_i c_i . F_i [, forces].

    A graph capture passes a preallocated `workspace`, `grad_points_out` and
    `grad_masses_out` (return_forces allocates)."""
    _check(beta, softening_scale, max_depth, leaf_size, multipole_order, vjp=True)
    points, masses, cotangent = _f32(points), _f32(masses), _f32(cotangent)
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if cotangent.shape != points.shape:
        raise ValueError("cotangent must have same shape as points")
    gx = grad_points_out if grad_points_out is not None else torch.empty_like(points)
    gm = grad_masses_out if grad_masses_out is not None else torch.empty_like(masses)
    forces = torch.empty_like(points) if return_forces else None
    call("sbh_gravity_force_vjp", stream_handle(), buf(points), buf(masses), buf(cotangent),
         buf(forces), buf(gx), buf(gm), f32(beta), f32(softening_scale), i32(max_depth),
         i32(leaf_size), i32(multipole_order), buf(workspace), buf(stats))
    if return_forces:
        return gx, gm, forces
    return gx, gm


class _GravityRecompute(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, masses, kw, vjp_order):
        ctx.save_for_backward(points, masses)
        ctx.kw = kw
        ctx.vjp_order = vjp_order
        return gravity_force(points, masses, **kw)

    @staticmethod
    def backward(ctx, cot):
        points, masses = ctx.saved_tensors
        kw = dict(ctx.kw, multipole_order=ctx.vjp_order)
        gx, gm = gravity_force_vjp(points, masses, cot.contiguous(), **kw)
        return gx, gm, None, None


def gravity_force_recompute_vjp(points, masses, *, beta: float = 2.0,
                                softening_scale: float = 1e-2, max_depth: int = 10,
                                leaf_size: int = 32, multipole_order: int = 2,
                                vjp_multipole_order: Optional[int] = None,
                                workspace: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable forces: backward rebuilds the tree and runs the adjoint
    walk (no interaction lists are stored).

    multipole_order is the far-field expansion of the forward walk;
    vjp_multipole_order (default: the same) that of the adjoint walk (1, 2
    or SECOND_MOMENT_ADJOINT).  Both approximate the same softened sum, so a
    monopole forward (the expansion of the reference's octree8 node records)
    may be paired with an adjoint of higher order when the gradients need
    it."""
    vjp_order = multipole_order if vjp_multipole_order is None else vjp_multipole_order
    _check(beta, softening_scale, max_depth, leaf_size, multipole_order)
    _check(beta, softening_scale, max_depth, leaf_size, vjp_order, vjp=True)
    p = points if isinstance(points, torch.Tensor) and points.is_cuda else _f32(points)
    m = masses if isinstance(masses, torch.Tensor) and masses.is_cuda else _f32(masses)
    kw = dict(beta=float(beta), softening_scale=float(softening_scale), max_depth=int(max_depth),
              leaf_size=int(leaf_size), multipole_order=int(multipole_order),
              workspace=workspace)
    return _GravityRecompute.apply(p, m, kw, int(vjp_order))


# ------------------------------------------------------------------ batches
def set_offsets(sizes, device=None) -> torch.Tensor:
    """int32 device vector [0, n_0, n_0 + n_1, ...] for sets of the given sizes."""
    sizes = np.asarray(sizes, np.int64)
    if sizes.ndim != 1 or sizes.size < 1 or np.any(sizes < 0):
        raise ValueError("sizes must be a non-empty vector of non-negative counts")
    offs = np.concatenate([[0], np.cumsum(sizes)])
    if offs[-1] > np.iinfo(np.int32).max:
        raise ValueError("too many points")
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.tensor(offs.astype(np.int32), device=dev)


def batch_workspace_bytes(total_points: int, num_sets: int, max_depth: int = 10,
                          leaf_size: int = 32) -> int:
    lib = _lib.load_library()
    lib.sbh_gravity_batch_workspace_bytes.restype = ctypes.c_int64
    v = int(lib.sbh_gravity_batch_workspace_bytes(ctypes.c_int64(total_points), i32(num_sets),
                                                  i32(max_depth), i32(leaf_size)))
    if v < 0:
        raise ValueError("invalid workspace query")
    return v


def new_batch_workspace(total_points: int, num_sets: int, max_depth: int = 10,
                        leaf_size: int = 32, device: Optional[torch.device] = None):
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.empty(batch_workspace_bytes(total_points, num_sets, max_depth, leaf_size),
                       dtype=torch.uint8, device=dev)


def _check_batch(points, masses, offsets, check_offsets):
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if not isinstance(offsets, torch.Tensor) or offsets.dtype != torch.int32 or \
            offsets.dim() != 1 or offsets.shape[0] < 2:
        raise ValueError("offsets must be an int32 tensor of num_sets + 1 entries")
    if check_offsets:
        o = offsets.cpu().numpy()
        if o[0] != 0 or o[-1] != points.shape[0] or np.any(np.diff(o) < 0):
            raise ValueError("offsets must rise from 0 to N")


def gravity_force_batch(points, masses, offsets, *, beta: float = 2.0,
                        softening_scale: float = 1e-2, max_depth: int = 10,
                        leaf_size: int = 32, multipole_order: int = 2, walk_blocks: int = 0,
                        walk_threads: int = 0, workspace: Optional[torch.Tensor] = None,
                        stats: Optional[torch.Tensor] = None,
                        out: Optional[torch.Tensor] = None,
                        check_offsets: bool = True) -> torch.Tensor:
    """Self-gravity of independent sets in one pipeline: set k is rows
    offsets[k]:offsets[k+1] of points/masses; no set sees another.  Inside
    a graph capture pass preallocated `out` and `workspace` and
    check_offsets=False (the check reads the offsets back to the host)."""
    _check(beta, softening_scale, max_depth, leaf_size, multipole_order)
    points, masses = _f32(points), _f32(masses)
    _check_batch(points, masses, offsets, check_offsets)
    forces = out if out is not None else torch.empty_like(points)
    call("sbh_gravity_batch_force", stream_handle(), buf(points), buf(masses), buf(offsets),
         buf(forces), f32(beta), f32(softening_scale), i32(max_depth), i32(leaf_size),
         i32(multipole_order), i32(walk_blocks), i32(walk_threads), buf(workspace), buf(stats))
    return forces


def gravity_force_vjp_batch(points, masses, cotangent, offsets, *, beta: float = 2.0,
                            softening_scale: float = 1e-2, max_depth: int = 10,
                            leaf_size: int = 32, multipole_order: int = 2,
                            walk_blocks: int = 0, walk_threads: int = 0,
                            workspace: Optional[torch.Tensor] = None,
                            stats: Optional[torch.Tensor] = None,
                            forces_out: Optional[torch.Tensor] = None,
                            grad_points_out: Optional[torch.Tensor] = None,
                            grad_masses_out: Optional[torch.Tensor] = None,
                            check_offsets: bool = True):
    """(grad_points, grad_masses) of sum_i c_i . F_i over all sets; also fills
    forces_out when given."""
    _check(beta, softening_scale, max_depth, leaf_size, multipole_order, vjp=True)
    points, masses, cotangent = _f32(points), _f32(masses), _f32(cotangent)
    _check_batch(points, masses, offsets, check_offsets)
    if cotangent.shape != points.shape:
        raise ValueError("cotangent must have same shape as points")
    gx = grad_points_out if grad_points_out is not None else torch.empty_like(points)
    gm = grad_masses_out if grad_masses_out is not None else torch.empty_like(masses)
    call("sbh_gravity_batch_force_vjp", stream_handle(), buf(points), buf(masses),
         buf(cotangent), buf(offsets), buf(forces_out), buf(gx), buf(gm), f32(beta),
         f32(softening_scale), i32(max_depth), i32(leaf_size), i32(multipole_order),
         i32(walk_blocks), i32(walk_threads), buf(workspace), buf(stats))
    return gx, gm


def new_stats(device=None) -> torch.Tensor:
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.zeros(8, dtype=torch.int64, device=dev)


def stats_dict(stats: torch.Tensor) -> dict:
    v = stats.detach().cpu().numpy().tolist()
    return dict(zip(STATS_FIELDS, v))


@dataclass
class TreeView:
    """Host copy of the tree arrays left in a workspace by the last build."""

    num_nodes: int
    geom: np.ndarray
    level_start: np.ndarray
    level_count: np.ndarray
    overflow: int
    codes: np.ndarray
    perm: np.ndarray
    sorted_points: np.ndarray
    topo: np.ndarray
    prefix: np.ndarray
    parent: np.ndarray
    geo: np.ndarray
    mp: np.ndarray
    qd: np.ndarray


def read_tree(workspace: torch.Tensor, n: int, max_depth: int = 10,
              leaf_size: int = 32) -> TreeView:
    torch.cuda.synchronize()
    L = workspace_layout(n, max_depth, leaf_size)
    ws = workspace.cpu().numpy()
    hdr = ws[L["header"]: L["header"] + 256]
    geom = hdr[:32].view(np.float32).copy()
    ints = hdr[32: 32 + 4 * (24 + 24 + 2)].view(np.int32)
    level_start, level_count = ints[:24].copy(), ints[24:48].copy()
    nn, overflow = int(ints[48]), int(ints[49])

    def arr(name, dtype, count, width=1):
        a = ws[L[name]: L[name] + count * width * np.dtype(dtype).itemsize].view(dtype)
        return a.reshape(count, width).copy() if width > 1 else a.copy()

    return TreeView(
        num_nodes=nn, geom=geom, level_start=level_start, level_count=level_count,
        overflow=overflow, codes=arr("codes", np.uint32, n), perm=arr("perm", np.uint32, n),
        sorted_points=arr("sp", np.float32, n, 4), topo=arr("topo", np.int32, nn, 4),
        prefix=arr("prefix", np.uint32, nn), parent=arr("parent", np.int32, nn),
        geo=arr("geo", np.float32, nn, 4), mp=arr("mp", np.float32, nn, 4),
        qd=arr("qd", np.float32, nn, 8))
