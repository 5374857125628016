"""Scalar fields of the reference's pybind11 GPU module (src/sbh_gpu.cu) on
device tensors: potential, winding number and smooth distance, each by brute
force, classic Barnes-Hut and the multi-level prefix control variate
(stochastic) walk.

The reference builds its trees on the CPU and returns (values, kernel ms).
sbh_amd.jax_sbh evaluates these kernels over the reference's own trees
(Tree2D/3D<1, L>, as python/jax_sbh.py); this module runs them over octree8
records (ffi.build_octree8_buffers, every
record with alpha = 1 as the reference's octree8 builder writes) and each
function returns a float32 tensor of shape (M,).  Kernels
(include/kernel.h): gravityPotential -m / (r + 0.1), signedSolidAngle<3>,
smoothDistExp exp(-alpha r).
"""
from __future__ import annotations

import torch

from . import ffi
from ._lib import buf, call, f32, i32, stream_handle
from .wn import (eval_barnes_hut_wn, eval_brute_force_wn,  # noqa: F401
                 multi_level_prefix_control_variate_wn)

_f32 = ffi._f32
_u8 = ffi._u8

__all__ = [
    "eval_brute_force_gravity", "eval_barnes_hut_gravity",
    "multi_level_prefix_control_variate_gravity",
    "eval_brute_force_wn", "eval_barnes_hut_wn", "multi_level_prefix_control_variate_wn",
    "eval_brute_force_smoothdist", "eval_barnes_hut_smoothdist",
    "multi_level_prefix_control_variate_smoothdist",
]


def _check_queries(queries):
    if queries.dim() != 2 or queries.shape[1] != 3:
        raise ValueError("queries must have shape (M, 3)")


def _walk(name, leaf_bytes, node_bytes, queries, beta, threads):
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_queries(queries)
    ffi._check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call(name, stream_handle(), buf(leaf_bytes), buf(node_bytes), buf(queries), buf(out),
         f32(beta), i32(threads))
    return out


def _stochastic(name, leaf_bytes, node_bytes, queries, samples, seed, beta, threads):
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_queries(queries)
    ffi._check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call(name, stream_handle(), buf(leaf_bytes), buf(node_bytes), buf(queries), buf(out),
         i32(samples), i32(seed), f32(beta), i32(threads))
    return out


def eval_brute_force_gravity(points, masses, queries, *, threads: int = 64) -> torch.Tensor:
    """src/sbh_gpu.cu:39-51, 213-248: sum of -m / (r + 0.1) over all points."""
    return ffi.brute_force_gravity(points, masses, queries, threads=threads)


def eval_barnes_hut_gravity(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                            threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:79-98, 324-361 on octree8 records."""
    return _walk("sbh_barnes_hut_potential_octree8", leaf_bytes, node_bytes, queries, beta,
                 threads)


def multi_level_prefix_control_variate_gravity(leaf_bytes, node_bytes, queries, *,
                                               samples_per_subdomain: int = 1, seed: int = 0,
                                               beta: float = 1.0,
                                               threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:143-164, 436-472 with the octree8 stochastic walk."""
    return _stochastic("sbh_stochastic_bh_potential_octree8", leaf_bytes, node_bytes, queries,
                       samples_per_subdomain, seed, beta, threads)


def eval_brute_force_smoothdist(points, queries, *, alpha=1.0,
                                threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:66-77, 287-322 (Kahan sum): sum of exp(-alpha_j r_j).

    alpha: a scalar for every point (the reference tree's constructor
    argument, include/py_tree.h:60-67) or a per-point (N,) tensor."""
    points, queries = _f32(points), _f32(queries)
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    _check_queries(queries)
    ffi._check_threads(threads)
    if isinstance(alpha, torch.Tensor) or hasattr(alpha, "__len__"):
        alphas = _f32(alpha)
        if alphas.dim() != 1 or alphas.shape[0] != points.shape[0]:
            raise ValueError("alpha must be a scalar or have shape (N,)")
    else:
        alphas = torch.full((points.shape[0],), float(alpha), dtype=torch.float32,
                            device=points.device)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    call("sbh_bruteforce_smoothdist", stream_handle(), buf(points), buf(alphas), buf(queries),
         buf(out), i32(threads))
    return out


def eval_barnes_hut_smoothdist(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                               threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:122-141, 400-434 on octree8 records (alpha = 1)."""
    return _walk("sbh_barnes_hut_smoothdist_octree8", leaf_bytes, node_bytes, queries, beta,
                 threads)


def multi_level_prefix_control_variate_smoothdist(leaf_bytes, node_bytes, queries, *,
                                                  samples_per_subdomain: int = 1, seed: int = 0,
                                                  beta: float = 1.0,
                                                  threads: int = 256) -> torch.Tensor:
    """src/sbh_gpu.cu:189-211, 512-550 with the octree8 stochastic walk."""
    return _stochastic("sbh_stochastic_bh_smoothdist_octree8", leaf_bytes, node_bytes,
                       queries, samples_per_subdomain, seed, beta, threads)
