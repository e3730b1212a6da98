"""Batched tridiagonal factorization and solve on the GPU, as autograd Functions.

The reference defines the two operators of the acoustic step's column solve as
custom autograd Functions (``mpas_pytorch/autodiff/tridiagonal_factor.py:145``,
``tridiagonal_solve.py:177``): the Thomas factorization

    alpha(k) = 1 / (b(k) - a(k) gamma(k-1)),  gamma(k) = c(k) alpha(k),  k = 1..L-1

of (nCells, L) coefficient arrays (alpha and gamma zero at level 0), and the
two sweeps of the solve on a (nCells, L + 1) right-hand side

    u(0) = r(0),  u(k) = (r(k) - a(k) u(k-1)) alpha(k),  k = 1..L-1
    x(L) = r(L),  x(k) = u(k) - gamma(k) x(k+1),         k = L-1..0

with hand-written reverse (``backward``) and forward (``jvp``) rules.  The
classes below keep their names, argument order, shape checks and error types;
forward, backward and jvp each run one HIP kernel (one lane per column, see
``csrc/kernels_tridiag.hip``).  The inputs are float64 tensors on a ROCm
device; anything else is rejected.
"""

from __future__ import annotations

import ctypes as C
from typing import Any

import torch

from . import _native as N
from .device import stream_handle
from .errors import MpasDeviceError

__all__ = ["TridiagonalFactorize", "TridiagonalSolve", "tridiagonal_factorize", "tridiagonal_solve"]


def _device_of(*tensors: torch.Tensor) -> torch.device:
    dev = tensors[0].device
    for t in tensors:
        if t.device.type != "cuda" or t.device != dev:
            raise MpasDeviceError(f"the tridiagonal operators run on one ROCm GPU; got a tensor on {t.device}",
                                  requested=str(t.device))
        if t.dtype != torch.float64:
            raise MpasDeviceError(f"the tridiagonal operators take float64 tensors, not {t.dtype}",
                                  requested=str(t.dtype))
    return dev


def _p(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def _c(t: torch.Tensor | None) -> torch.Tensor | None:
    return None if t is None else t.detach().contiguous()


def _call(name: str, dev: torch.device, n: int, levels: int, *ptrs) -> None:
    lib = N.load()
    with torch.cuda.device(dev):
        rc = getattr(lib, name)(n, levels, *ptrs, C.c_void_p(stream_handle(dev)))
    N.check(rc, name)


class TridiagonalFactorize(torch.autograd.Function):
    """alpha, gamma of the Thomas factorization of the (a, b, c) diagonals,
    each (nCells, nVertLevels)."""

    @staticmethod
    def forward(ctx: Any, a_tri: torch.Tensor, b_tri: torch.Tensor,
                c_tri: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        if a_tri.shape != b_tri.shape or a_tri.shape != c_tri.shape:
            raise ValueError("a_tri, b_tri, c_tri must have identical shapes")
        if a_tri.ndim != 2:
            raise ValueError("expected shape (nCells, nVertLevels)")
        dev = _device_of(a_tri, b_tri, c_tri)
        a, b, c = _c(a_tri), _c(b_tri), _c(c_tri)
        n, levels = a.shape
        alpha = torch.empty_like(a)
        gamma = torch.empty_like(a)
        if levels > 0:
            _call("mpas_hip_tridiagonal_factorize", dev, n, levels, _p(a), _p(b), _p(c), _p(alpha), _p(gamma))
        ctx.save_for_backward(a, c, alpha, gamma)
        ctx.save_for_forward(a, c, alpha, gamma)
        return alpha, gamma

    @staticmethod
    def backward(ctx: Any, grad_alpha: torch.Tensor | None,
                 grad_gamma: torch.Tensor | None) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        a, c, alpha, gamma = ctx.saved_tensors
        n, levels = a.shape
        ga, gb, gc = torch.empty_like(a), torch.empty_like(a), torch.empty_like(a)
        if levels == 0:
            return ga, gb, gc
        _call("mpas_hip_tridiagonal_factorize_vjp", a.device, n, levels, _p(a), _p(c), _p(alpha), _p(gamma),
              _p(_c(grad_alpha)), _p(_c(grad_gamma)), _p(ga), _p(gb), _p(gc))
        return ga, gb, gc

    @staticmethod
    def jvp(ctx: Any, grad_a: torch.Tensor | None, grad_b: torch.Tensor | None,
            grad_c: torch.Tensor | None) -> tuple[torch.Tensor, torch.Tensor]:
        a, c, alpha, gamma = ctx.saved_tensors
        n, levels = a.shape
        d_alpha, d_gamma = torch.empty_like(a), torch.empty_like(a)
        if levels == 0:
            return d_alpha, d_gamma
        _call("mpas_hip_tridiagonal_factorize_jvp", a.device, n, levels, _p(a), _p(c), _p(alpha), _p(gamma),
              _p(_c(grad_a)), _p(_c(grad_b)), _p(_c(grad_c)), _p(d_alpha), _p(d_gamma))
        return d_alpha, d_gamma


class TridiagonalSolve(torch.autograd.Function):
    """Solution (nCells, nVertLevels + 1) of the factored system for the
    right-hand side ``rhs``, whose last level is the boundary value the back
    substitution starts from."""

    @staticmethod
    def forward(ctx: Any, a_tri: torch.Tensor, alpha_tri: torch.Tensor, gamma_tri: torch.Tensor,
                rhs: torch.Tensor) -> torch.Tensor:
        if a_tri.shape != alpha_tri.shape or a_tri.shape != gamma_tri.shape:
            raise ValueError("a_tri, alpha_tri, gamma_tri must have identical shapes")
        if a_tri.ndim != 2:
            raise ValueError("expected tridiagonal tensors with shape (nCells, nVertLevels)")
        if rhs.ndim != 2 or rhs.shape[0] != a_tri.shape[0] or rhs.shape[1] != a_tri.shape[1] + 1:
            raise ValueError("rhs must have shape (nCells, nVertLevels + 1)")
        dev = _device_of(a_tri, alpha_tri, gamma_tri, rhs)
        a, al, ga, r = _c(a_tri), _c(alpha_tri), _c(gamma_tri), _c(rhs)
        n, levels = a.shape
        out = torch.empty_like(r)
        if levels == 0:
            out.copy_(r)
        else:
            _call("mpas_hip_tridiagonal_solve", dev, n, levels, _p(a), _p(al), _p(ga), _p(r), _p(out))
        ctx.save_for_backward(a, al, ga, r, out)
        ctx.save_for_forward(a, al, ga, r, out)
        return out

    @staticmethod
    def backward(ctx: Any, grad_out: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
        a, al, ga, r, out = ctx.saved_tensors
        n, levels = a.shape
        g_a, g_al, g_ga, g_r = torch.empty_like(a), torch.empty_like(al), torch.empty_like(ga), torch.empty_like(r)
        if levels == 0:
            return g_a, g_al, g_ga, grad_out.clone()
        _call("mpas_hip_tridiagonal_solve_vjp", a.device, n, levels, _p(a), _p(al), _p(ga), _p(r), _p(out),
              _p(_c(grad_out)), _p(g_a), _p(g_al), _p(g_ga), _p(g_r))
        return g_a, g_al, g_ga, g_r

    @staticmethod
    def jvp(ctx: Any, grad_a: torch.Tensor | None, grad_alpha: torch.Tensor | None,
            grad_gamma: torch.Tensor | None, grad_rhs: torch.Tensor | None) -> torch.Tensor:
        a, al, ga, r, out = ctx.saved_tensors
        n, levels = a.shape
        d_out = torch.empty_like(r)
        if levels == 0:
            return d_out.zero_() if grad_rhs is None else grad_rhs.clone()
        _call("mpas_hip_tridiagonal_solve_jvp", a.device, n, levels, _p(a), _p(al), _p(ga), _p(r), _p(out),
              _p(_c(grad_a)), _p(_c(grad_alpha)), _p(_c(grad_gamma)), _p(_c(grad_rhs)), _p(d_out))
        return d_out


def tridiagonal_factorize(a_tri: torch.Tensor, b_tri: torch.Tensor,
                          c_tri: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """``TridiagonalFactorize.apply(a_tri, b_tri, c_tri)``."""
    return TridiagonalFactorize.apply(a_tri, b_tri, c_tri)


def tridiagonal_solve(a_tri: torch.Tensor, alpha_tri: torch.Tensor, gamma_tri: torch.Tensor,
                      rhs: torch.Tensor) -> torch.Tensor:
    """``TridiagonalSolve.apply(a_tri, alpha_tri, gamma_tri, rhs)``."""
    return TridiagonalSolve.apply(a_tri, alpha_tri, gamma_tri, rhs)
