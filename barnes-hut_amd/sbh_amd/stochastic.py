"""Stochastic Barnes-Hut on the adaptive octree (sbh_stochastic_* targets).

The reference's multi-level prefix control-variate estimator
(stochastic_barnes_hut_force_octree8 and its _vjp, python/jax_sbh_ffi.py:
1450-1535 over src/sbh_ffi.cu:2146-2442; multi_level_prefix_control_variate_wn,
src/sbh_gpu.cu:166-187) on the adaptive tree, whose depth is not limited by a
dense 8^D node array.  Same keyword names and defaults as the reference's
stochastic wrappers (beta 1.0, samples_per_subdomain 1, seed 0); the tree
parameters are those of the adaptive gravity targets.  The expectation of
every estimate is the exact sum; queries of one group of <= 64 (their own
Morton tree) share the random draws.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from ._lib import buf, call, f32, i32, stream_handle
from .ffi import _f32

STATS_FIELDS = ("node_terms", "source_pairs", "groups", "nodes", "levels", "overflow",
                "subdomains")


def _check(points, queries, beta, max_depth, leaf_size, samples_per_subdomain):
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if queries.dim() != 2 or queries.shape[1] != 3:
        raise ValueError("queries must have shape (M, 3)")
    if not beta >= 0:
        raise ValueError("beta must be >= 0")
    if not 1 <= max_depth <= 10:
        raise ValueError("max_depth must be in [1, 10]")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    if samples_per_subdomain < 0:
        raise ValueError("samples_per_subdomain must be >= 0")


def workspace_bytes(num_sources: int, num_queries: int, max_depth: int = 10,
                    leaf_size: int = 32) -> int:
    lib = _lib.load_library()
    lib.sbh_stochastic_workspace_bytes.restype = ctypes.c_int64
    v = int(lib.sbh_stochastic_workspace_bytes(ctypes.c_int64(num_sources),
                                               ctypes.c_int64(num_queries), i32(max_depth),
                                               i32(leaf_size)))
    if v < 0:
        raise ValueError("invalid workspace query")
    return v


def new_workspace(num_sources: int, num_queries: int, max_depth: int = 10,
                  leaf_size: int = 32, device=None) -> torch.Tensor:
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.empty(workspace_bytes(num_sources, num_queries, max_depth, leaf_size),
                       dtype=torch.uint8, device=dev)


def stochastic_gravity_force(points, masses, queries, *, beta: float = 1.0,
                             samples_per_subdomain: int = 1, seed: int = 0,
                             softening_scale: float = 1e-2, max_depth: int = 10,
                             leaf_size: int = 32, multipole_order: int = 1,
                             workspace: Optional[torch.Tensor] = None,
                             stats: Optional[torch.Tensor] = None,
                             out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Estimate of F(q) = sum_j -m_j (|q - x_j|^2 + eps^2 + 1e-12)^(-3/2) (q - x_j)."""
    points, masses, queries = _f32(points), _f32(masses), _f32(queries)
    _check(points, queries, beta, max_depth, leaf_size, samples_per_subdomain)
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if multipole_order not in (1, 2):
        raise ValueError("multipole_order must be 1 or 2")
    out = out if out is not None else torch.empty_like(queries)
    call("sbh_stochastic_gravity_force", stream_handle(), buf(points), buf(masses),
         buf(queries), buf(out), f32(beta), f32(softening_scale), i32(max_depth),
         i32(leaf_size), i32(multipole_order), i32(samples_per_subdomain), i32(seed),
         buf(workspace), buf(stats))
    return out


def stochastic_gravity_force_vjp(points, masses, queries, cotangent, *, beta: float = 1.0,
                                 samples_per_subdomain: int = 1, seed: int = 0,
                                 softening_scale: float = 1e-2, max_depth: int = 10,
                                 leaf_size: int = 32, multipole_order: int = 1,
                                 workspace: Optional[torch.Tensor] = None,
                                 stats: Optional[torch.Tensor] = None,
                                 out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Estimate of J(q)^T c, the force's Jacobian with respect to the query
    position, along the same sample paths as the forward (the reference's
    stochastic VJP: query side only)."""
    points, masses = _f32(points), _f32(masses)
    queries, cotangent = _f32(queries), _f32(cotangent)
    _check(points, queries, beta, max_depth, leaf_size, samples_per_subdomain)
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have the shape of queries")
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if multipole_order not in (1, 2):
        raise ValueError("multipole_order must be 1 or 2")
    out = out if out is not None else torch.empty_like(queries)
    call("sbh_stochastic_gravity_force_vjp", stream_handle(), buf(points), buf(masses),
         buf(queries), buf(cotangent), buf(out), f32(beta), f32(softening_scale),
         i32(max_depth), i32(leaf_size), i32(multipole_order), i32(samples_per_subdomain),
         i32(seed), buf(workspace), buf(stats))
    return out


def stochastic_winding_number(points, normals, areas, queries, *, beta: float = 1.0,
                              samples_per_subdomain: int = 1, seed: int = 0,
                              max_depth: int = 10, leaf_size: int = 16,
                              workspace: Optional[torch.Tensor] = None,
                              stats: Optional[torch.Tensor] = None,
                              out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Estimate of sum a n . (p - q) / (4 pi (|p - q|^3 + 1e-5)) over the
    oriented area elements (centroid p, unit normal n, area a)."""
    points, normals, areas, queries = _f32(points), _f32(normals), _f32(areas), _f32(queries)
    _check(points, queries, beta, max_depth, leaf_size, samples_per_subdomain)
    if normals.shape != points.shape:
        raise ValueError("normals must have shape (N, 3)")
    if areas.dim() != 1 or areas.shape[0] != points.shape[0]:
        raise ValueError("areas must have shape (N,)")
    out = out if out is not None else torch.empty((queries.shape[0],), dtype=torch.float32,
                                                  device=queries.device)
    call("sbh_stochastic_winding_number", stream_handle(), buf(points), buf(normals),
         buf(areas), buf(queries), buf(out), f32(beta), i32(max_depth), i32(leaf_size),
         i32(samples_per_subdomain), i32(seed), buf(workspace), buf(stats))
    return out


def stats_dict(stats: torch.Tensor) -> dict:
    v = stats.detach().cpu().numpy().tolist()
    return dict(zip(STATS_FIELDS, v))
