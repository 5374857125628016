"""The reference's tree fields module (python/jax_sbh.py over the pybind11 GPU
module src/sbh_gpu.cu) on device tensors.

The reference builds a Tree2D<1, L> / Tree3D<1, L> (leaf size 1, 2^L cells per
axis, L = level, default 1) on the host, copies it to the GPU and evaluates
gravityPotential, signedSolidAngle<N> and smoothDistExp over it by brute
force (compensated sums), the warp-voted Barnes-Hut walk and the multi-level
control-variate sampler, all in 32-query blocks.  Here the same tree bytes are
built on the device (sbh_tree_build) and the same walks run in libsbh_amd.so.
Every evaluation returns a float32 tensor of shape (M,) (the module also
returns a kernel time, which python/jax_sbh.py drops).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import ffi
from ._lib import buf, call, f32, i32, stream_handle

_f32 = ffi._f32

__all__ = [
    "SbhGpuTree", "build_gpu_tree", "build_gpu_worker",
    "eval_brute_force_gravity", "eval_barnes_hut_gravity", "eval_mlpcv_gravity",
    "eval_brute_force_wn", "eval_barnes_hut_wn", "multi_level_prefix_control_variate_wn",
    "eval_brute_force_smoothdist", "eval_barnes_hut_smoothdist",
    "multi_level_prefix_control_variate_smoothdist",
    "multi_level_prefix_control_variate_gravity",
]

_GRAVITY, _WN, _SMOOTHDIST = 0, 1, 2


@dataclass
class SbhGpuTree:
    """python/jax_sbh.py:15-22: a built tree and its parameters; the device
    buffers are the reference's leaf_data / node_buffer / contrib_data bytes."""
    dim: int
    level: int
    alpha: float
    leaf: torch.Tensor
    node: torch.Tensor
    contrib: torch.Tensor

    def close(self) -> None:
        """python/jax_sbh.py:32-47 (SbhGpuWorker.close): nothing to stop here."""


def build_gpu_tree(points, normals, masses, min_corner=None, max_corner=None, *,
                   alpha: float = 1.0, level: int = 1) -> SbhGpuTree:
    """python/jax_sbh.py:132-162 (sbh.QuadtreeL / sbh.OctreeL, then the GPU
    copy): corners default to the points' bounds -/+ 1e-3."""
    points = _f32(points)
    if points.dim() != 2:
        raise ValueError("points must be (N, dim)")
    dim = int(points.shape[1])
    if dim not in (2, 3):
        raise ValueError(f"Unsupported dim: {dim}")
    if level not in (1, 2, 3):
        raise ValueError(f"Unsupported level: {level}")
    leaf, node, contrib = ffi._tree_buffers(points, normals, masses, min_corner, max_corner,
                                            float(alpha), dim, int(level))
    return SbhGpuTree(dim=dim, level=int(level), alpha=float(alpha), leaf=leaf, node=node,
                      contrib=contrib)


def build_gpu_worker(points, normals, masses, min_corner=None, max_corner=None, *,
                     alpha: float = 1.0, level: int = 1) -> SbhGpuTree:
    """python/jax_sbh.py:165-200.  The reference runs its CUDA module in a
    separate process to keep it away from JAX's context; this library shares
    the caller's HIP context, so the worker is the tree itself."""
    return build_gpu_tree(points, normals, masses, min_corner, max_corner, alpha=alpha,
                          level=level)


def _queries(tree: SbhGpuTree, queries):
    queries = _f32(queries)
    if queries.dim() != 2 or queries.shape[1] != tree.dim:
        raise ValueError(f"queries must be (M, {tree.dim})")
    return queries


def _eval(tree: SbhGpuTree, queries, kind: int, beta: float, samples: int) -> torch.Tensor:
    q = _queries(tree, queries)
    out = torch.empty((q.shape[0],), dtype=torch.float32, device=q.device)
    call("sbh_tree_eval", stream_handle(), i32(tree.dim), i32(tree.level), i32(kind),
         buf(tree.leaf), buf(tree.node), buf(tree.contrib), buf(q), buf(out), f32(beta),
         i32(samples))
    return out


def _brute(tree: SbhGpuTree, queries, kind: int) -> torch.Tensor:
    q = _queries(tree, queries)
    out = torch.empty((q.shape[0],), dtype=torch.float32, device=q.device)
    call("sbh_tree_bruteforce", stream_handle(), i32(tree.dim), i32(kind), buf(tree.leaf),
         buf(q), buf(out))
    return out


def _samples(samples_per_subdomain: int) -> int:
    s = int(samples_per_subdomain)
    if s < 0:
        raise ValueError("samples_per_subdomain must be >= 0")
    return s


# python/jax_sbh.py:241-257
def eval_brute_force_gravity(tree: SbhGpuTree, queries) -> torch.Tensor:
    """src/sbh_gpu.cu:39-51, 213-249."""
    return _brute(tree, queries, _GRAVITY)


def eval_barnes_hut_gravity(tree: SbhGpuTree, queries, *, beta: float = 2.0) -> torch.Tensor:
    """src/sbh_gpu.cu:79-98, 324-362."""
    return _eval(tree, queries, _GRAVITY, beta, -1)


def eval_mlpcv_gravity(tree: SbhGpuTree, queries, *,
                       samples_per_subdomain: int = 1) -> torch.Tensor:
    """src/sbh_gpu.cu:143-164, 436-473."""
    return _eval(tree, queries, _GRAVITY, 0.0, _samples(samples_per_subdomain))


multi_level_prefix_control_variate_gravity = eval_mlpcv_gravity


# the module's other kernels (bindings in src/sbh_gpu.cu:566-650)
def eval_brute_force_wn(tree: SbhGpuTree, queries) -> torch.Tensor:
    """src/sbh_gpu.cu:53-64, 250-286: signedSolidAngle<dim>."""
    return _brute(tree, queries, _WN)


def eval_barnes_hut_wn(tree: SbhGpuTree, queries, *, beta: float = 2.0) -> torch.Tensor:
    """src/sbh_gpu.cu:100-120, 363-399."""
    return _eval(tree, queries, _WN, beta, -1)


def multi_level_prefix_control_variate_wn(tree: SbhGpuTree, queries, *,
                                          samples_per_subdomain: int = 1) -> torch.Tensor:
    """src/sbh_gpu.cu:166-187, 474-511."""
    return _eval(tree, queries, _WN, 0.0, _samples(samples_per_subdomain))


def eval_brute_force_smoothdist(tree: SbhGpuTree, queries) -> torch.Tensor:
    """src/sbh_gpu.cu:66-77, 287-323: exp(-alpha |p - q|), alpha the tree's."""
    return _brute(tree, queries, _SMOOTHDIST)


def eval_barnes_hut_smoothdist(tree: SbhGpuTree, queries, *, beta: float = 2.0) -> torch.Tensor:
    """src/sbh_gpu.cu:122-141, 400-435."""
    return _eval(tree, queries, _SMOOTHDIST, beta, -1)


def multi_level_prefix_control_variate_smoothdist(tree: SbhGpuTree, queries, *,
                                                  samples_per_subdomain: int = 1
                                                  ) -> torch.Tensor:
    """src/sbh_gpu.cu:189-211, 512-550."""
    return _eval(tree, queries, _SMOOTHDIST, 0.0, _samples(samples_per_subdomain))
