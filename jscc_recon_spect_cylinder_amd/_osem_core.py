"""Shared plumbing of the OSEM drop-ins: device placement of matrices, images and
rotation tables, and the calls into the fused projection / EM-update kernels."""
from __future__ import annotations

import torch

from . import ops

_perm_cache = ops.SourceCache(64)


def gpu_device(t: torch.Tensor) -> torch.device:
    if not t.is_cuda:
        raise ValueError("the image must be on a GPU: this engine runs the OSEM kernels on the GPU only")
    return t.device


def perm_on(perm: torch.Tensor, dev: torch.device) -> torch.Tensor:
    """1-based [P, R] rotation table as int64 on `dev` (cached per source tensor)."""
    if perm.device == dev and perm.dtype == torch.int64:
        return perm
    hit = _perm_cache.get(perm, (str(dev),))
    if hit is None:
        hit = _perm_cache.put(perm, (str(dev),), perm.to(device=dev, dtype=torch.int64))
    return hit


def matrix_on(mat: torch.Tensor, dev: torch.device) -> torch.Tensor:
    """A system-matrix or T block on the GPU in a layout the projection kernel reads."""
    if mat.device != dev:
        mat = mat.to(dev, non_blocking=True)
    if mat.dtype not in (torch.float32, torch.bfloat16):
        mat = mat.float()
    return ops.proj_ready(mat, ops.dtype_code(mat.dtype))


def image_vec(img: torch.Tensor) -> torch.Tensor:
    x = img.reshape(-1)
    if x.dtype != torch.float32 or not x.is_contiguous():
        x = x.float().contiguous()
    return x


def add_weight(w: torch.Tensor, mat: torch.Tensor, x: torch.Tensor, *, y=None, x_perm=None, inv_perm=None,
               eps=None, scale: float = 1.0, rows=None, row_base: int = 0, n_rows=None, row_scale=None,
               y_by_row: bool = False, sanitize: bool = False, n_cols=None, dtype_c=None) -> None:
    """w (+)= scale * G^T [ rows^T (y / max(rows x_rot, eps)) ]   (one get_weight_* call)."""
    if n_rows is None and rows is None and mat.shape[0] == 0:
        return
    ops.proj(ops.MODE_RATIO, mat, dtype_c=dtype_c, n_cols=n_cols, rows=rows, n_rows=n_rows, row_base=row_base,
             row_scale=row_scale, x=x, x_perm=x_perm, y=y, eps=eps, inv_perm=inv_perm, out_scale=scale,
             accumulate=True, w=w, y_by_row=y_by_row, sanitize=sanitize)


def em(x: torch.Tensor, w: torch.Tensor, s: torch.Tensor, mode: int = ops.EM_PLAIN, c1: float = 1.0, w2=None,
       c2: float = 0.0, zero_after=None) -> None:
    ops.em_update(x, w, image_vec(s), mode=mode, c1=c1, w2=w2, c2=c2, zero_after=zero_after)


def apply_denoiser(img: torch.Tensor, model_denoiser) -> torch.Tensor:
    """The optional image hook of recon_osem_plane.py:56-59 (a user torch module)."""
    n = int(img.shape[0] ** 0.5)
    out = model_denoiser(img.view(1, 1, n, n)).view(n * n, 1)
    return torch.clamp(out, min=0)
