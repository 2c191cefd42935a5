"""Differentiable projectors: the Compton cone operator T(sigma) and the
single-photon system matrix A behind torch.autograd.Function, with
hand-written adjoint kernels.

  cone_project(x, sigma, store)   f_i = T_i(sigma) . x for the kept events of a
                                  LazyListStore; the backward recomputes T
                                  (jscc_cone_proj, coefficient mode) and returns
                                  dL/dx = T^T g and dL/dsigma = sum_i g_i df_i/dsigma
                                  (fixed angular blur model; the kept-event set and
                                  the support filter are held fixed)
  system_project(x, A)            f = A x; backward A^T g (jscc_proj)

Poisson log-likelihoods on top of them are plain torch code on the projections:
  list mode   L = sum_i log f_i - s . x,  s = sum_i T_i  (so s . x = sum_i f_i)
  bin mode    L = sum_m y_m log (A x)_m - sum_m (A x)_m
No counterpart in the reference (its reconstructions are EM only).
"""
from __future__ import annotations

import dataclasses

import torch

from . import ops
from .engine import LazyListStore


class _ConeProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, sigma, store: LazyListStore):
        sig = float(sigma)
        st = dataclasses.replace(store, fixed_sigma=sig)
        xv = x.detach().reshape(-1).float().contiguous()
        f, _ = st.forward(xv)
        ctx.save_for_backward(xv)
        ctx.st = st
        ctx.x_shape = x.shape
        ctx.sigma_like = sigma if isinstance(sigma, torch.Tensor) else None
        return f.clone()

    @staticmethod
    def backward(ctx, grad_f):
        (xv,) = ctx.saved_tensors
        need_sigma = ctx.needs_input_grad[1] and ctx.sigma_like is not None
        w, g = ctx.st.adjoint(xv, grad_f.detach().float().contiguous(), dsig=need_sigma)
        grad_x = w.view(ctx.x_shape) if ctx.needs_input_grad[0] else None
        grad_sigma = None
        if need_sigma:
            grad_sigma = (grad_f.double() * g.double()).sum().to(ctx.sigma_like.dtype).reshape(ctx.sigma_like.shape)
        return grad_x, grad_sigma, None


def cone_project(x: torch.Tensor, sigma, store: LazyListStore) -> torch.Tensor:
    """f = T(sigma) x (normalised rows of the kept events), differentiable in x and sigma."""
    if store.sigma_model != ops.SIG_FIXED:
        raise ValueError("the differentiable cone operator uses the fixed angular blur model")
    return _ConeProjection.apply(x, sigma, store)


class _SystemProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, A):
        xv = x.detach().reshape(-1).float().contiguous()
        mat = ops.proj_ready(A, ops.dtype_code(A.dtype))
        f = torch.empty(max(A.shape[0], 1), dtype=torch.float32, device=x.device)
        ops.proj(ops.MODE_FWD, mat, x=xv, fwd_out=f)
        ctx.mat = mat
        ctx.x_shape = x.shape
        return f[: A.shape[0]].clone()

    @staticmethod
    def backward(ctx, grad_f):
        mat = ctx.mat
        w = torch.zeros(mat.shape[1], dtype=torch.float32, device=grad_f.device)
        ops.proj(ops.MODE_COEF, mat, coef=grad_f.detach().float().contiguous(), w=w)
        return w.view(ctx.x_shape), None


def system_project(x: torch.Tensor, A: torch.Tensor) -> torch.Tensor:
    """f = A x with the adjoint A^T in the backward."""
    return _SystemProjection.apply(x, A)


def list_loglik(f: torch.Tensor) -> torch.Tensor:
    """Poisson list-mode log-likelihood with s = sum_i T_i: sum_i log f_i - sum_i f_i (FP64)."""
    f64 = f.double()
    return torch.log(f64).sum() - f64.sum()


def bin_loglik(f: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Poisson bin-mode log-likelihood sum_m y_m log f_m - f_m (FP64)."""
    f64 = f.double()
    return (y.double() * torch.log(f64)).sum() - f64.sum()
