"""Drop-in mirror of the reference's Python FFI surface (python/jax_sbh_ffi.py).

Same function names, positional order, keyword defaults and ValueError checks
as the reference; arrays are PyTorch device tensors (numpy / lists are
accepted and moved to the current HIP device) and every computation runs in
libsbh_amd.so through the C ABI of include/sbh_amd.h.

`register_*` functions exist for call-site compatibility.  There is no XLA
registry to populate without JAX, so they only check that the library loads.

The reference's `jax.custom_vjp` recompute wrappers are mirrored as
`torch.autograd.Function`s with the same residual/recompute structure
(forward builds and walks; backward rebuilds the tree and runs the VJP walk).
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from ._lib import buf, call, f32, i32, stream_handle

_FD_VJP_EPS = 2e-5


# ------------------------------------------------------------------ helpers
def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("sbh_amd needs a HIP device (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def _f32(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x.detach()
        if t.dtype != torch.float32:
            t = t.float()
        if not t.is_cuda:
            t = t.to(_device())
        return t.contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device=_device()).contiguous()


def _u8(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x.detach()
        if t.dtype != torch.uint8:
            t = t.to(torch.uint8)
        if not t.is_cuda:
            t = t.to(_device())
        return t.contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.uint8), device=_device()).contiguous()


def _check_threads(threads: int) -> None:
    if threads < 1:
        raise ValueError("threads must be >= 1")


def _check_periodic(periodic: bool, box_length: float) -> None:
    if periodic and box_length <= 0:
        raise ValueError("box_length must be > 0 when periodic=True")


def _get_octree8_layout(path: Optional[str] = None) -> tuple[int, int]:
    lib = _lib.load_library(path)
    return int(lib.sbh_leaf3_size()), int(lib.sbh_octree8_node_size())


def _get_octree_level3_layout(path: Optional[str] = None) -> tuple[int, int, int]:
    lib = _lib.load_library(path)
    return (int(lib.sbh_leaf3_size()), int(lib.sbh_node3_size()),
            int(lib.sbh_octree_level3_max_nodes()))


def _octree8_total_nodes(max_depth: int) -> int:
    total, level = 0, 1
    for _ in range(max_depth + 1):
        total += level
        level *= 8
    return total


def _register(path: Optional[str] = None) -> None:
    _lib.load_library(path)


register_bruteforce_gravity = _register
register_softened_direct_force = _register
register_softened_direct_force_vjp = _register
register_build_octree8_buffers = _register
register_barnes_hut_force_octree8 = _register
register_softened_barnes_hut_force_octree8 = _register
register_softened_barnes_hut_force_octree8_vjp = _register


# ------------------------------------------------------------------ direct sums
def brute_force_gravity(points, masses, queries, *, periodic: bool = False,
                        box_length: float = 1.0, threads: int = 64) -> torch.Tensor:
    """python/jax_sbh_ffi.py:341-378."""
    points, masses, queries = _f32(points), _f32(masses), _f32(queries)
    if points.dim() != 2 or queries.dim() != 2:
        raise ValueError("points and queries must be rank-2")
    if masses.dim() != 1:
        raise ValueError("masses must be rank-1")
    if points.shape[0] != masses.shape[0]:
        raise ValueError("masses length mismatch")
    if points.shape[1] != queries.shape[1]:
        raise ValueError("dim mismatch")
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=points.device)
    call("sbh_bruteforce_gravity", stream_handle(), buf(points), buf(masses), buf(queries),
         buf(out), i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def _check_direct(points, masses, queries, softening_scale, cutoff_scale, periodic,
                  box_length, threads):
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must be (N, 3)")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if queries.dim() != 2 or queries.shape[1] != 3:
        raise ValueError("queries must be (M, 3)")
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if cutoff_scale <= 0:
        raise ValueError("cutoff_scale must be > 0")
    _check_periodic(periodic, box_length)
    _check_threads(threads)


def softened_direct_force(points, masses, queries, *, softening_scale: float = 1e-2,
                          cutoff_scale: float = 1.0, periodic: bool = False,
                          box_length: float = 1.0, threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:381-426."""
    points, masses, queries = _f32(points), _f32(masses), _f32(queries)
    _check_direct(points, masses, queries, softening_scale, cutoff_scale, periodic,
                  box_length, threads)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=points.device)
    call("sbh_direct_force_soft", stream_handle(), buf(points), buf(masses), buf(queries),
         buf(out), f32(softening_scale), f32(cutoff_scale), i32(1 if periodic else 0),
         f32(box_length), i32(threads))
    return out


def softened_direct_force_vjp(points, masses, queries, cotangent, *,
                              softening_scale: float = 1e-2, cutoff_scale: float = 1.0,
                              periodic: bool = False, box_length: float = 1.0,
                              threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:429-479 (query-side J^T g)."""
    points, masses, queries = _f32(points), _f32(masses), _f32(queries)
    cotangent = _f32(cotangent)
    _check_direct(points, masses, queries, softening_scale, cutoff_scale, periodic,
                  box_length, threads)
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have same shape as queries")
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=points.device)
    call("sbh_direct_force_soft_vjp", stream_handle(), buf(points), buf(masses), buf(queries),
         buf(cotangent), buf(out), f32(softening_scale), f32(cutoff_scale),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def softened_direct_force_full_vjp(points, masses, cotangent, *,
                                   softening_scale: float = 1e-2, cutoff_scale: float = 1.0,
                                   periodic: bool = False, box_length: float = 1.0,
                                   threads: int = 256) -> tuple[torch.Tensor, torch.Tensor]:
    """Full adjoint of sum_i c_i . F_i for self-gravity: (grad_points, grad_masses).

    No reference counterpart (the reference returns the query side only)."""
    points, masses, cotangent = _f32(points), _f32(masses), _f32(cotangent)
    _check_direct(points, masses, points, softening_scale, cutoff_scale, periodic,
                  box_length, threads)
    if cotangent.shape != points.shape:
        raise ValueError("cotangent must have same shape as points")
    gx = torch.empty_like(points)
    gm = torch.empty_like(masses)
    call("sbh_direct_force_soft_full_vjp", stream_handle(), buf(points), buf(masses),
         buf(cotangent), buf(gx), buf(gm), f32(softening_scale), f32(cutoff_scale),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return gx, gm


class _DirectRecompute(torch.autograd.Function):
    """custom_vjp of python/jax_sbh_ffi.py:482-594 (residual = inputs)."""

    @staticmethod
    def forward(ctx, points, masses, softening_scale, cutoff_scale, periodic, box_length,
                threads, vjp_mode):
        ctx.save_for_backward(points, masses)
        ctx.params = (softening_scale, cutoff_scale, periodic, box_length, threads, vjp_mode)
        return softened_direct_force(points, masses, points, softening_scale=softening_scale,
                                     cutoff_scale=cutoff_scale, periodic=periodic,
                                     box_length=box_length, threads=threads)

    @staticmethod
    def backward(ctx, cot):
        points, masses = ctx.saved_tensors
        softening_scale, cutoff_scale, periodic, box_length, threads, vjp_mode = ctx.params
        kw = dict(softening_scale=softening_scale, cutoff_scale=cutoff_scale,
                  periodic=periodic, box_length=box_length, threads=threads)
        if vjp_mode == "full":
            gx, gm = softened_direct_force_full_vjp(points, masses, cot.contiguous(), **kw)
        else:
            gx = softened_direct_force_vjp(points, masses, points, cot.contiguous(), **kw)
            gm = torch.zeros_like(masses)
        return gx, gm, None, None, None, None, None, None


def softened_direct_force_recompute_vjp(points, masses, *, softening_scale: float = 1e-2,
                                        cutoff_scale: float = 1.0, periodic: bool = False,
                                        box_length: float = 1.0, threads: int = 256,
                                        vjp_mode: str = "query") -> torch.Tensor:
    """python/jax_sbh_ffi.py:502-530.  vjp_mode="query" reproduces the
    reference's backward (query side, zero mass gradient); "full" returns the
    complete position and mass gradients."""
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if cutoff_scale <= 0:
        raise ValueError("cutoff_scale must be > 0")
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    if vjp_mode not in ("query", "full"):
        raise ValueError("vjp_mode must be 'query' or 'full'")
    p = points if isinstance(points, torch.Tensor) and points.is_cuda else _f32(points)
    m = masses if isinstance(masses, torch.Tensor) and masses.is_cuda else _f32(masses)
    return _DirectRecompute.apply(p, m, float(softening_scale), float(cutoff_scale),
                                  bool(periodic), float(box_length), int(threads), vjp_mode)


# ------------------------------------------------------------------ finite differences
def _force_vjp_finite_difference(forward_force_fn: Callable, queries, cotangent, *,
                                 eps: float = _FD_VJP_EPS) -> torch.Tensor:
    """python/jax_sbh_ffi.py:597-623."""
    queries, cotangent = _f32(queries), _f32(cotangent)
    if queries.dim() != 2 or queries.shape[1] not in (2, 3):
        raise ValueError("queries must be (M, 2) or (M, 3)")
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have same shape as queries")
    if eps <= 0:
        raise ValueError("eps must be > 0")
    dim = int(queries.shape[1])
    e = torch.tensor(np.float32(eps), device=queries.device)
    cols = []
    for axis in range(dim):
        delta = torch.zeros((1, dim), dtype=torch.float32, device=queries.device)
        delta[0, axis] = e
        lp = (forward_force_fn(queries + delta) * cotangent).sum(dim=1)
        lm = (forward_force_fn(queries - delta) * cotangent).sum(dim=1)
        cols.append(((lp - lm) / (2.0 * e))[:, None])
    return torch.cat(cols, dim=1)


def _force_vjp_finite_difference_elementwise(forward_force_fn: Callable, queries, cotangent,
                                             *, eps: float = _FD_VJP_EPS) -> torch.Tensor:
    """python/jax_sbh_ffi.py:626-655 (6*M forward calls)."""
    q = _f32(queries).cpu().numpy()
    c = _f32(cotangent)
    if q.ndim != 2 or q.shape[1] not in (2, 3):
        raise ValueError("queries must be (M, 2) or (M, 3)")
    if tuple(c.shape) != q.shape:
        raise ValueError("cotangent must have same shape as queries")
    if eps <= 0:
        raise ValueError("eps must be > 0")
    grad = np.zeros_like(q, dtype=np.float32)
    for i in range(q.shape[0]):
        for d in range(q.shape[1]):
            qp, qm = q.copy(), q.copy()
            qp[i, d] += eps
            qm[i, d] -= eps
            lp = float((forward_force_fn(_f32(qp)) * c).sum())
            lm = float((forward_force_fn(_f32(qm)) * c).sum())
            grad[i, d] = (lp - lm) / (2.0 * eps)
    return _f32(grad)


# ------------------------------------------------------------------ octree8 (drop-in mode)
def _default_corners(points, min_corner, max_corner):
    # python/jax_sbh_ffi.py:879-884: fp32 min - 1e-3 / max + 1e-3.
    if (min_corner is None or max_corner is None) and points.shape[0] == 0:
        # numpy's min/max of an empty array raise ValueError there too
        raise ValueError("zero-size array: pass min_corner and max_corner for empty points")
    if min_corner is None:
        min_corner = points.min(dim=0).values - np.float32(1e-3)
    if max_corner is None:
        max_corner = points.max(dim=0).values + np.float32(1e-3)
    return _f32(min_corner), _f32(max_corner)


def build_octree8_buffers_with_contrib(points, normals, masses, min_corner=None,
                                       max_corner=None, *, max_depth: int = 7):
    """python/jax_sbh_ffi.py:913-960."""
    leaf_size, node_size = _get_octree8_layout()
    points, normals, masses = _f32(points), _f32(normals), _f32(masses)
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if normals.shape != points.shape:
        raise ValueError("normals shape mismatch")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    if max_depth < 1 or max_depth > 10:
        raise ValueError("max_depth must be in [1, 10]")
    mn, mx = _default_corners(points, min_corner, max_corner)
    n = points.shape[0]
    total = _octree8_total_nodes(int(max_depth))
    dev = points.device
    leaf = torch.empty((n * leaf_size,), dtype=torch.uint8, device=dev)
    node = torch.empty((total * node_size,), dtype=torch.uint8, device=dev)
    contrib = torch.empty((total * leaf_size,), dtype=torch.uint8, device=dev)
    call("sbh_build_octree8_buffers", stream_handle(), buf(points), buf(normals), buf(masses),
         buf(mn), buf(mx), buf(leaf), buf(node), buf(contrib), i32(max_depth))
    return leaf, node, contrib


def build_octree8_buffers(points, normals, masses, min_corner=None, max_corner=None, *,
                          max_depth: int = 7):
    """python/jax_sbh_ffi.py:864-910."""
    leaf, node, _ = build_octree8_buffers_with_contrib(points, normals, masses, min_corner,
                                                       max_corner, max_depth=max_depth)
    return leaf, node


def _check_walk(queries, periodic, box_length, threads):
    if queries.dim() != 2 or queries.shape[1] != 3:
        raise ValueError("queries must be (M, 3)")
    _check_periodic(periodic, box_length)
    _check_threads(threads)


def barnes_hut_force_octree8(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                             periodic: bool = False, box_length: float = 1.0,
                             threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1336-1369."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_walk(queries, periodic, box_length, threads)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_barnes_hut_force_octree8", stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), f32(beta), i32(1 if periodic else 0), f32(box_length),
         i32(threads))
    return out


def barnes_hut_force_octree8_vjp(leaf_bytes, node_bytes, queries, cotangent, *,
                                 beta: float = 2.0, periodic: bool = False,
                                 box_length: float = 1.0, threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1372-1395 (central differences)."""
    return _force_vjp_finite_difference(
        lambda q: barnes_hut_force_octree8(leaf_bytes, node_bytes, q, beta=beta,
                                           periodic=periodic, box_length=box_length,
                                           threads=threads),
        queries, cotangent)


def _check_soft(softening_scale, cutoff_scale, prune_enabled, prune_r_cut_mult):
    if softening_scale < 0:
        raise ValueError("softening_scale must be >= 0")
    if cutoff_scale <= 0:
        raise ValueError("cutoff_scale must be > 0")
    if prune_enabled and prune_r_cut_mult <= 0:
        raise ValueError("prune_r_cut_mult must be > 0 when prune_enabled=True")


def softened_barnes_hut_force_octree8(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                                      softening_scale: float = 1e-2, cutoff_scale: float = 1.0,
                                      prune_enabled: bool = True,
                                      prune_r_cut_mult: float = 4.0, periodic: bool = False,
                                      box_length: float = 1.0,
                                      threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1398-1447."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_walk(queries, periodic, box_length, threads)
    _check_soft(softening_scale, cutoff_scale, prune_enabled, prune_r_cut_mult)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_barnes_hut_force_octree8_soft", stream_handle(), buf(leaf_bytes),
         buf(node_bytes), buf(queries), buf(out), f32(beta), f32(softening_scale),
         f32(cutoff_scale), i32(1 if prune_enabled else 0), f32(prune_r_cut_mult),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def softened_barnes_hut_force_octree8_vjp(leaf_bytes, node_bytes, queries, cotangent, *,
                                          beta: float = 2.0, softening_scale: float = 1e-2,
                                          cutoff_scale: float = 1.0,
                                          prune_enabled: bool = True,
                                          prune_r_cut_mult: float = 4.0,
                                          periodic: bool = False, box_length: float = 1.0,
                                          threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1674-1728 (query-side J^T g, tree fixed)."""
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    queries, cotangent = _f32(queries), _f32(cotangent)
    _check_walk(queries, periodic, box_length, threads)
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have same shape as queries")
    _check_soft(softening_scale, cutoff_scale, prune_enabled, prune_r_cut_mult)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_barnes_hut_force_octree8_soft_vjp", stream_handle(), buf(leaf_bytes),
         buf(node_bytes), buf(queries), buf(cotangent), buf(out), f32(beta),
         f32(softening_scale), f32(cutoff_scale), i32(1 if prune_enabled else 0),
         f32(prune_r_cut_mult), i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


class _Octree8SoftRecompute(torch.autograd.Function):
    """custom_vjp of python/jax_sbh_ffi.py:1731-1897: forward builds the tree
    and walks; backward rebuilds the tree and runs the VJP walk."""

    @staticmethod
    def forward(ctx, points, masses, kw, max_depth):
        ctx.save_for_backward(points, masses)
        ctx.kw = kw
        ctx.max_depth = max_depth
        normals = torch.zeros_like(points)
        leaf, node = build_octree8_buffers(points, normals, masses, max_depth=max_depth)
        return softened_barnes_hut_force_octree8(leaf, node, points, **kw)

    @staticmethod
    def backward(ctx, cot):
        points, masses = ctx.saved_tensors
        normals = torch.zeros_like(points)
        leaf, node = build_octree8_buffers(points, normals, masses, max_depth=ctx.max_depth)
        gx = softened_barnes_hut_force_octree8_vjp(leaf, node, points, cot.contiguous(),
                                                   **ctx.kw)
        return gx, torch.zeros_like(masses), None, None


def softened_barnes_hut_force_octree8_recompute_vjp(points, masses, *, beta: float = 2.0,
                                                    softening_scale: float = 1e-2,
                                                    cutoff_scale: float = 1.0,
                                                    max_depth: int = 7,
                                                    prune_enabled: bool = True,
                                                    prune_r_cut_mult: float = 4.0,
                                                    periodic: bool = False,
                                                    box_length: float = 1.0,
                                                    threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1769-1801 (reference gradient semantics: query
    side, zero mass gradient).  The full adjoint lives in sbh_amd.gravity."""
    if prune_enabled and prune_r_cut_mult <= 0:
        raise ValueError("prune_r_cut_mult must be > 0 when prune_enabled=True")
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    p = points if isinstance(points, torch.Tensor) and points.is_cuda else _f32(points)
    m = masses if isinstance(masses, torch.Tensor) and masses.is_cuda else _f32(masses)
    kw = dict(beta=float(beta), softening_scale=float(softening_scale),
              cutoff_scale=float(cutoff_scale), prune_enabled=bool(prune_enabled),
              prune_r_cut_mult=float(prune_r_cut_mult), periodic=bool(periodic),
              box_length=float(box_length), threads=int(threads))
    return _Octree8SoftRecompute.apply(p, m, kw, int(max_depth))


# ------------------------------------------------------------------ stochastic octree8
register_stochastic_barnes_hut_force_octree8 = _register
register_stochastic_barnes_hut_force_octree8_vjp = _register
register_softened_stochastic_barnes_hut_force_octree8 = _register
register_softened_stochastic_barnes_hut_force_octree8_vjp = _register


def stochastic_barnes_hut_force_octree8(leaf_bytes, node_bytes, queries, *, beta: float = 1.0,
                                        samples_per_subdomain: int = 1, seed: int = 0,
                                        periodic: bool = False, box_length: float = 1.0,
                                        threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1450-1487."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_walk(queries, periodic, box_length, threads)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_stochastic_bh_force_octree8", stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), i32(samples_per_subdomain), i32(seed), f32(beta),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def stochastic_barnes_hut_force_octree8_vjp(leaf_bytes, node_bytes, queries, cotangent, *,
                                            beta: float = 1.0, samples_per_subdomain: int = 1,
                                            seed: int = 0, periodic: bool = False,
                                            box_length: float = 1.0,
                                            threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1490-1534."""
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    queries, cotangent = _f32(queries), _f32(cotangent)
    _check_walk(queries, periodic, box_length, threads)
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have same shape as queries")
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_stochastic_bh_force_octree8_vjp", stream_handle(), buf(leaf_bytes),
         buf(node_bytes), buf(queries), buf(cotangent), buf(out), i32(samples_per_subdomain),
         i32(seed), f32(beta), i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def softened_stochastic_barnes_hut_force_octree8(leaf_bytes, node_bytes, queries, *,
                                                 beta: float = 1.0,
                                                 samples_per_subdomain: int = 1, seed: int = 0,
                                                 softening_scale: float = 1e-2,
                                                 cutoff_scale: float = 1.0,
                                                 prune_enabled: bool = True,
                                                 prune_r_cut_mult: float = 4.0,
                                                 periodic: bool = False,
                                                 box_length: float = 1.0,
                                                 threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1537-1588."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _f32(queries)
    _check_walk(queries, periodic, box_length, threads)
    _check_soft(softening_scale, cutoff_scale, prune_enabled, prune_r_cut_mult)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_stochastic_bh_force_octree8_soft", stream_handle(), buf(leaf_bytes),
         buf(node_bytes), buf(queries), buf(out), i32(samples_per_subdomain), i32(seed),
         f32(beta), f32(softening_scale), f32(cutoff_scale), i32(1 if prune_enabled else 0),
         f32(prune_r_cut_mult), i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def softened_stochastic_barnes_hut_force_octree8_vjp(leaf_bytes, node_bytes, queries,
                                                     cotangent, *, beta: float = 1.0,
                                                     samples_per_subdomain: int = 1,
                                                     seed: int = 0,
                                                     softening_scale: float = 1e-2,
                                                     cutoff_scale: float = 1.0,
                                                     prune_enabled: bool = True,
                                                     prune_r_cut_mult: float = 4.0,
                                                     periodic: bool = False,
                                                     box_length: float = 1.0,
                                                     threads: int = 256,
                                                     use_ffi: bool = False) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1591-1671: use_ffi=False is the reference's
    elementwise central difference of the forward call (6*M calls)."""
    kw = dict(beta=beta, samples_per_subdomain=samples_per_subdomain, seed=seed,
              softening_scale=softening_scale, cutoff_scale=cutoff_scale,
              prune_enabled=prune_enabled, prune_r_cut_mult=prune_r_cut_mult,
              periodic=periodic, box_length=box_length, threads=threads)
    if not use_ffi:
        return _force_vjp_finite_difference_elementwise(
            lambda q: softened_stochastic_barnes_hut_force_octree8(leaf_bytes, node_bytes, q,
                                                                   **kw),
            queries, cotangent)
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    queries, cotangent = _f32(queries), _f32(cotangent)
    _check_walk(queries, periodic, box_length, threads)
    if cotangent.shape != queries.shape:
        raise ValueError("cotangent must have same shape as queries")
    _check_soft(softening_scale, cutoff_scale, prune_enabled, prune_r_cut_mult)
    out = torch.empty((queries.shape[0], 3), dtype=torch.float32, device=queries.device)
    call("sbh_stochastic_bh_force_octree8_soft_vjp", stream_handle(), buf(leaf_bytes),
         buf(node_bytes), buf(queries), buf(cotangent), buf(out), i32(samples_per_subdomain),
         i32(seed), f32(beta), f32(softening_scale), f32(cutoff_scale),
         i32(1 if prune_enabled else 0), f32(prune_r_cut_mult), i32(1 if periodic else 0),
         f32(box_length), i32(threads))
    return out


class _StochasticRecompute(torch.autograd.Function):
    """custom_vjp of python/jax_sbh_ffi.py:1900-2046 and 2049-2248: forward
    builds and samples; backward rebuilds and runs the VJP walk with the same
    seed (query side; zero mass gradient, as the reference)."""

    @staticmethod
    def forward(ctx, points, masses, kw, max_depth, softened):
        ctx.save_for_backward(points, masses)
        ctx.kw, ctx.max_depth, ctx.softened = kw, max_depth, softened
        leaf, node = build_octree8_buffers(points, torch.zeros_like(points), masses,
                                           max_depth=max_depth)
        if softened:
            return softened_stochastic_barnes_hut_force_octree8(leaf, node, points, **kw)
        return stochastic_barnes_hut_force_octree8(leaf, node, points, **kw)

    @staticmethod
    def backward(ctx, cot):
        points, masses = ctx.saved_tensors
        leaf, node = build_octree8_buffers(points, torch.zeros_like(points), masses,
                                           max_depth=ctx.max_depth)
        if ctx.softened:
            gx = softened_stochastic_barnes_hut_force_octree8_vjp(
                leaf, node, points, cot.contiguous(), use_ffi=True, **ctx.kw)
        else:
            gx = stochastic_barnes_hut_force_octree8_vjp(leaf, node, points, cot.contiguous(),
                                                         **ctx.kw)
        return gx, torch.zeros_like(masses), None, None, None


def stochastic_barnes_hut_force_octree8_recompute_vjp(points, masses, *, beta: float = 1.0,
                                                      samples_per_subdomain: int = 1,
                                                      seed: int = 0, max_depth: int = 7,
                                                      periodic: bool = False,
                                                      box_length: float = 1.0,
                                                      threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1926-1946."""
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    p = points if isinstance(points, torch.Tensor) and points.is_cuda else _f32(points)
    m = masses if isinstance(masses, torch.Tensor) and masses.is_cuda else _f32(masses)
    kw = dict(beta=float(beta), samples_per_subdomain=int(samples_per_subdomain),
              seed=int(seed), periodic=bool(periodic), box_length=float(box_length),
              threads=int(threads))
    return _StochasticRecompute.apply(p, m, kw, int(max_depth), False)


def softened_stochastic_barnes_hut_force_octree8_recompute_vjp(
        points, masses, *, beta: float = 1.0, samples_per_subdomain: int = 1, seed: int = 0,
        softening_scale: float = 1e-2, cutoff_scale: float = 1.0, max_depth: int = 7,
        prune_enabled: bool = True, prune_r_cut_mult: float = 4.0, periodic: bool = False,
        box_length: float = 1.0, threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:2079-2113."""
    if prune_enabled and prune_r_cut_mult <= 0:
        raise ValueError("prune_r_cut_mult must be > 0 when prune_enabled=True")
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    p = points if isinstance(points, torch.Tensor) and points.is_cuda else _f32(points)
    m = masses if isinstance(masses, torch.Tensor) and masses.is_cuda else _f32(masses)
    kw = dict(beta=float(beta), samples_per_subdomain=int(samples_per_subdomain),
              seed=int(seed), softening_scale=float(softening_scale),
              cutoff_scale=float(cutoff_scale), prune_enabled=bool(prune_enabled),
              prune_r_cut_mult=float(prune_r_cut_mult), periodic=bool(periodic),
              box_length=float(box_length), threads=int(threads))
    return _StochasticRecompute.apply(p, m, kw, int(max_depth), True)


# ------------------------------------------------------------------ level-3 tree (drop-in mode)
register_build_octree_level3_buffers = _register
register_barnes_hut_gravity = _register
register_barnes_hut_force = _register
register_softened_barnes_hut_force = _register
register_stochastic_barnes_hut_gravity = _register
register_stochastic_barnes_hut_force = _register
register_softened_stochastic_barnes_hut_force = _register


def _l3_queries(queries):
    queries = _f32(queries)
    if queries.dim() != 2 or queries.shape[1] not in (2, 3):
        raise ValueError("queries must be (M, 2) or (M, 3)")
    return queries


NODE_SIZES = {(2, 1): 136, (2, 2): 232, (2, 3): 616, (3, 1): 184, (3, 2): 632, (3, 3): 4272}


def _tree_buffers(points, normals, masses, min_corner, max_corner, alpha, dim, level=3):
    """Tree2D<1,L> / Tree3D<1,L> bytes (the reference's sbh.QuadtreeL /
    sbh.OctreeL, include/tree2d.h:237-407, include/tree3d.h:238-413), built on
    the device: (leaf, node, contrib) uint8 tensors."""
    points, normals, masses = _f32(points), _f32(normals), _f32(masses)
    if points.dim() != 2 or points.shape[1] != dim:
        raise ValueError(f"points must have shape (N, {dim})")
    if normals.shape != points.shape:
        raise ValueError("normals shape mismatch")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    mn, mx = _default_corners(points, min_corner, max_corner)
    nodes = ctypes.c_int64(0)
    s = stream_handle()
    call("sbh_tree_build", s, i32(dim), i32(level), buf(points), buf(normals), buf(masses),
         buf(mn), buf(mx), f32(alpha), ctypes.byref(nodes))
    lsz, nsz = (24 if dim == 2 else 32), NODE_SIZES[(dim, level)]
    dev = points.device
    leaf = torch.empty((points.shape[0] * lsz,), dtype=torch.uint8, device=dev)
    node = torch.empty((nodes.value * nsz,), dtype=torch.uint8, device=dev)
    contrib = torch.empty((nodes.value * lsz,), dtype=torch.uint8, device=dev)
    call("sbh_tree_emit", s, buf(leaf), buf(node), buf(contrib))
    return leaf, node, contrib


def build_quadtree_buffers_with_contrib(points, normals, masses, min_corner=None,
                                        max_corner=None, *, alpha: float = 1.0, level: int = 3):
    """python/jax_sbh_ffi.py:692-724 (Tree2DNode<1,3> buffers of sbh.Quadtree3,
    built on the device)."""
    if level != 3:
        raise ValueError("level must be 3 (FFI expects Tree2DNode<1, 3>)")
    return _tree_buffers(points, normals, masses, min_corner, max_corner, alpha, 2)


def build_quadtree_buffers(points, normals, masses, min_corner=None, max_corner=None, *,
                           alpha: float = 1.0, level: int = 3):
    """python/jax_sbh_ffi.py:658-689."""
    leaf, node, _ = build_quadtree_buffers_with_contrib(points, normals, masses, min_corner,
                                                        max_corner, alpha=alpha, level=level)
    return leaf, node


def build_octree_buffers_with_contrib(points, normals, masses, min_corner=None,
                                      max_corner=None, *, alpha: float = 1.0, level: int = 3,
                                      backend: str = "ffi"):
    """python/jax_sbh_ffi.py:805-861: backend "ffi" gives the level-3 buffers
    (leaf N*32, node 513*4272, contrib 513*32 bytes); backend "cpu" the
    multi-level Tree3D<1,3> buffers of sbh.Octree3 (alpha 1, as the reference
    passes), here built on the device."""
    del alpha
    if level != 3:
        raise ValueError("level must be 3 (FFI expects Tree3DNode<1, 3>)")
    if backend == "cpu":
        return _tree_buffers(points, normals, masses, min_corner, max_corner, 1.0, 3)
    if backend != "ffi":
        raise ValueError("backend must be 'ffi' or 'cpu'")
    leaf_size, node_size, max_nodes = _get_octree_level3_layout()
    points, normals, masses = _f32(points), _f32(normals), _f32(masses)
    mn, mx = _default_corners(points, min_corner, max_corner)
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    if normals.shape != points.shape:
        raise ValueError("normals shape mismatch")
    if masses.dim() != 1 or masses.shape[0] != points.shape[0]:
        raise ValueError("masses must have shape (N,)")
    dev = points.device
    leaf = torch.empty((points.shape[0] * leaf_size,), dtype=torch.uint8, device=dev)
    node = torch.empty((max_nodes * node_size,), dtype=torch.uint8, device=dev)
    contrib = torch.empty((max_nodes * leaf_size,), dtype=torch.uint8, device=dev)
    call("sbh_build_octree_level3_buffers", stream_handle(), buf(points), buf(normals),
         buf(masses), buf(mn), buf(mx), buf(leaf), buf(node), buf(contrib))
    return leaf, node, contrib


def build_octree_buffers(points, normals, masses, min_corner=None, max_corner=None, *,
                         alpha: float = 1.0, level: int = 3, backend: str = "ffi"):
    """python/jax_sbh_ffi.py:744-802."""
    leaf, node, _ = build_octree_buffers_with_contrib(points, normals, masses, min_corner,
                                                      max_corner, alpha=alpha, level=level,
                                                      backend=backend)
    return leaf, node


def barnes_hut_gravity(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                       periodic: bool = False, box_length: float = 1.0,
                       threads: int = 64) -> torch.Tensor:
    """python/jax_sbh_ffi.py:963-997 (potential, (M,))."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    target = "sbh_barnes_hut_gravity_3d" if queries.shape[1] == 3 else "sbh_barnes_hut_gravity"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), f32(beta), i32(1 if periodic else 0), f32(box_length),
         i32(threads))
    return out


def barnes_hut_force(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                     periodic: bool = False, box_length: float = 1.0,
                     threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1000-1034."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty(tuple(queries.shape), dtype=torch.float32, device=queries.device)
    target = "sbh_barnes_hut_force_3d" if queries.shape[1] == 3 else "sbh_barnes_hut_force"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), f32(beta), i32(1 if periodic else 0), f32(box_length),
         i32(threads))
    return out


def softened_barnes_hut_force(leaf_bytes, node_bytes, queries, *, beta: float = 2.0,
                              softening_scale: float = 1e-2, cutoff_scale: float = 1.0,
                              periodic: bool = False, box_length: float = 1.0,
                              threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1063-1105."""
    leaf_bytes, node_bytes, queries = _u8(leaf_bytes), _u8(node_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty(tuple(queries.shape), dtype=torch.float32, device=queries.device)
    target = "sbh_barnes_hut_force_soft_3d" if queries.shape[1] == 3 else \
        "sbh_barnes_hut_force_soft"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(queries), buf(out), f32(beta), f32(softening_scale), f32(cutoff_scale),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def stochastic_barnes_hut_gravity(leaf_bytes, node_bytes, contrib_bytes, queries, *,
                                  samples_per_subdomain: int = 1, seed: int = 0,
                                  periodic: bool = False, box_length: float = 1.0,
                                  threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1138-1177."""
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    contrib_bytes, queries = _u8(contrib_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty((queries.shape[0],), dtype=torch.float32, device=queries.device)
    target = "sbh_stochastic_bh_gravity_3d" if queries.shape[1] == 3 else \
        "sbh_stochastic_bh_gravity"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(contrib_bytes), buf(queries), buf(out), i32(samples_per_subdomain), i32(seed),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def stochastic_barnes_hut_force(leaf_bytes, node_bytes, contrib_bytes, queries, *,
                                samples_per_subdomain: int = 1, seed: int = 0,
                                periodic: bool = False, box_length: float = 1.0,
                                threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1180-1219."""
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    contrib_bytes, queries = _u8(contrib_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty(tuple(queries.shape), dtype=torch.float32, device=queries.device)
    target = "sbh_stochastic_bh_force_3d" if queries.shape[1] == 3 else \
        "sbh_stochastic_bh_force"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(contrib_bytes), buf(queries), buf(out), i32(samples_per_subdomain), i32(seed),
         i32(1 if periodic else 0), f32(box_length), i32(threads))
    return out


def softened_stochastic_barnes_hut_force(leaf_bytes, node_bytes, contrib_bytes, queries, *,
                                         samples_per_subdomain: int = 1, seed: int = 0,
                                         softening_scale: float = 1e-2,
                                         cutoff_scale: float = 1.0, periodic: bool = False,
                                         box_length: float = 1.0,
                                         threads: int = 256) -> torch.Tensor:
    """python/jax_sbh_ffi.py:1252-1299."""
    leaf_bytes, node_bytes = _u8(leaf_bytes), _u8(node_bytes)
    contrib_bytes, queries = _u8(contrib_bytes), _l3_queries(queries)
    _check_periodic(periodic, box_length)
    _check_threads(threads)
    out = torch.empty(tuple(queries.shape), dtype=torch.float32, device=queries.device)
    target = "sbh_stochastic_bh_force_soft_3d" if queries.shape[1] == 3 else \
        "sbh_stochastic_bh_force_soft"
    call(target, stream_handle(), buf(leaf_bytes), buf(node_bytes),
         buf(contrib_bytes), buf(queries), buf(out), i32(samples_per_subdomain), i32(seed),
         f32(softening_scale), f32(cutoff_scale), i32(1 if periodic else 0), f32(box_length),
         i32(threads))
    return out


def barnes_hut_force_vjp(leaf_bytes, node_bytes, queries, cotangent, *, beta: float = 2.0,
                         periodic: bool = False, box_length: float = 1.0, threads: int = 256):
    """python/jax_sbh_ffi.py:1037-1060 (central differences of the forward)."""
    return _force_vjp_finite_difference(
        lambda q: barnes_hut_force(leaf_bytes, node_bytes, q, beta=beta, periodic=periodic,
                                   box_length=box_length, threads=threads),
        queries, cotangent)


def softened_barnes_hut_force_vjp(leaf_bytes, node_bytes, queries, cotangent, *,
                                  beta: float = 2.0, softening_scale: float = 1e-2,
                                  cutoff_scale: float = 1.0, periodic: bool = False,
                                  box_length: float = 1.0, threads: int = 256):
    """python/jax_sbh_ffi.py:1108-1135."""
    return _force_vjp_finite_difference(
        lambda q: softened_barnes_hut_force(leaf_bytes, node_bytes, q, beta=beta,
                                            softening_scale=softening_scale,
                                            cutoff_scale=cutoff_scale, periodic=periodic,
                                            box_length=box_length, threads=threads),
        queries, cotangent)


def stochastic_barnes_hut_force_vjp(leaf_bytes, node_bytes, contrib_bytes, queries, cotangent,
                                    *, samples_per_subdomain: int = 1, seed: int = 0,
                                    periodic: bool = False, box_length: float = 1.0,
                                    threads: int = 256):
    """python/jax_sbh_ffi.py:1222-1249."""
    return _force_vjp_finite_difference(
        lambda q: stochastic_barnes_hut_force(leaf_bytes, node_bytes, contrib_bytes, q,
                                              samples_per_subdomain=samples_per_subdomain,
                                              seed=seed, periodic=periodic,
                                              box_length=box_length, threads=threads),
        queries, cotangent)


def softened_stochastic_barnes_hut_force_vjp(leaf_bytes, node_bytes, contrib_bytes, queries,
                                             cotangent, *, samples_per_subdomain: int = 1,
                                             seed: int = 0, softening_scale: float = 1e-2,
                                             cutoff_scale: float = 1.0, periodic: bool = False,
                                             box_length: float = 1.0, threads: int = 256):
    """python/jax_sbh_ffi.py:1302-1333."""
    return _force_vjp_finite_difference(
        lambda q: softened_stochastic_barnes_hut_force(
            leaf_bytes, node_bytes, contrib_bytes, q,
            samples_per_subdomain=samples_per_subdomain, seed=seed,
            softening_scale=softening_scale, cutoff_scale=cutoff_scale, periodic=periodic,
            box_length=box_length, threads=threads),
        queries, cotangent)
