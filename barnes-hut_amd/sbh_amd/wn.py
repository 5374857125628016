"""Winding numbers of oriented area elements (dipoles) on the GPU.

The reference computes these through its pybind11 module over CPU-built
trees (src/sbh_gpu.cu: eval_brute_force_wn, eval_barnes_hut_wn,
multi_level_prefix_control_variate_wn), returning (values, kernel ms).  Here
the same arithmetic (signedSolidAngle<3>, include/kernel.h:73-80) runs on
device tensors through libsbh_amd.so:

* eval_brute_force_wn: exact compensated sum over all elements;
* eval_barnes_hut_wn / multi_level_prefix_control_variate_wn: classic and
  stochastic walks over octree8 records (ffi.build_octree8_buffers with
  points = centroids, normals = unit normals, masses = areas);
* winding_number: the adaptive-tree group walk (large meshes, many queries).

All return a float32 tensor of shape (M,).
"""
from __future__ import annotations

from typing import Optional

import torch

from . import ffi
from ._lib import buf, call, f32, i32, stream_handle

_f32 = ffi._f32
_u8 = ffi._u8


def _check_elements(points, normals, areas):
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if normals.shape != points.shape:
        raise ValueError("normals must have shape (N, 3)")
    if areas.dim() != 1 or areas.shape[0] != points.shape[0]:
        raise ValueError("areas must have shape (N,)")


def _check_queries(queries):
    if queries.dim() != 2 or queries.shape[1] != 3:
        raise ValueError("queries must have shape (M, 3)")


def eval_brute_force_wn(points, normals, areas, queries, *, threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:53-64, 250-282 (Kahan sum of include/brute_force_gpu.h:18-66)."""
    points, normals, areas, queries = _f32(points), _f32(normals), _f32(areas), _f32(queries)
    _check_elements(points, normals, areas)
    _check_queries(queries)
    ffi._check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call("sbh_bruteforce_wn", stream_handle(), buf(points), buf(normals), buf(areas),
         buf(queries), buf(out), i32(threads))
    return out


def eval_barnes_hut_wn(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                       threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:100-120, 363-397 on octree8 records."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_queries(queries)
    ffi._check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call("sbh_barnes_hut_wn_octree8", stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), f32(beta), i32(threads))
    return out


def multi_level_prefix_control_variate_wn(leaf_bytes, node_bytes, queries, *,
                                          samples_per_subdomain: int = 1, seed: int = 0,
                                          beta: float = 1.0,
                                          threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:166-187, 474-510 with the octree8 stochastic walk."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_queries(queries)
    ffi._check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call("sbh_stochastic_bh_wn_octree8", stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), i32(samples_per_subdomain), i32(seed), f32(beta),
         i32(threads))
    return out


def workspace_bytes(num_elements: int, num_queries: int, max_depth: int = 10,
                    leaf_size: int = 16) -> int:
    from ._lib import load_library

    v = int(load_library().sbh_winding_number_workspace_bytes(
        int(num_elements), int(num_queries), int(max_depth), int(leaf_size)))
    if v < 0:
        raise ValueError("invalid workspace query")
    return v


def new_workspace(num_elements: int, num_queries: int, max_depth: int = 10,
                  leaf_size: int = 16, device=None) -> torch.Tensor:
    dev = device if device is not None else ffi._device()
    return torch.empty((workspace_bytes(num_elements, num_queries, max_depth, leaf_size),),
                       dtype=torch.uint8, device=dev)


def winding_number(points, normals, areas, queries, *, beta: float = 2.0, max_depth: int = 10,
                   leaf_size: int = 16, workspace: Optional[torch.Tensor] = None,
                   stats: Optional[torch.Tensor] = None,
                   out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Adaptive-tree winding numbers (group walk; new target)."""
    points, normals, areas, queries = _f32(points), _f32(normals), _f32(areas), _f32(queries)
    _check_elements(points, normals, areas)
    _check_queries(queries)
    if not beta >= 0:
        raise ValueError("beta must be >= 0")
    if max_depth < 1 or max_depth > 10:
        raise ValueError("max_depth must be in [1, 10]")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    if out is None:
        out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call("sbh_winding_number", stream_handle(), buf(points), buf(normals), buf(areas),
         buf(queries), buf(out), f32(beta), i32(max_depth), i32(leaf_size),
         buf(workspace), buf(stats))
    return out
