"""Device-resident list-mode engine: T built by the cone kernel straight into
HBM and consumed there by the fused projection kernel.

The reference builds T chunk by chunk, copies it to the host, splits it into
subsets / rotations / blocks and copies blocks back per use
(process_list_plane_strict.py:249-251; recon_osem_plane.py:195-243).  Here the
rows stay where the cone kernel wrote them, unnormalised and uncompacted; the
kept-row index list and the per-row scale (1/S, or the FP8 scale) carry the
reference's compaction and row normalisation into the projection kernel.
Subsets are contiguous ranges of the kept list with the reference's
split_tensor_even bounds (recon_osem_plane.py:8-18, 204).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import ops
from .process_list_plane_strict import ComptonRows


@dataclass
class ListModeStore:
    mat: torch.Tensor          # [n_sel, ld] rows (FP32 / BF16 / FP8 codes)
    dtype_c: int
    n_vox: int
    kept: torch.Tensor         # int32 stored-row ids of the kept events, in order
    n_kept: int
    row_scale: torch.Tensor    # per stored row: value = row_scale * stored

    @classmethod
    def from_rows(cls, rows: ComptonRows, dtype="bf16") -> "ListModeStore":
        code = ops.dtype_code(dtype)
        if code == ops.DT_FP8:
            codes, scale = ops.quantize_fp8(rows.store, rows.n_vox, rows.n_sel, rows.stats)
            return cls(codes, ops.DT_FP8, rows.n_vox, rows.kept, rows.n_kept, scale)
        if ops.dtype_code(rows.store.dtype) != code:
            raise ValueError("the rows were built in another storage dtype")
        return cls(rows.store, code, rows.n_vox, rows.kept, rows.n_kept, rows.inv_sum)

    def nbytes_kept(self) -> int:
        per = {ops.DT_F32: 4, ops.DT_BF16: 2, ops.DT_FP8: 1}[self.dtype_c]
        return self.n_kept * self.n_vox * per

    def subset_bounds(self, k: int):
        return [(self.n_kept * i // k, self.n_kept * (i + 1) // k) for i in range(k)]

    def backproject_ones(self) -> torch.Tensor:
        """sum_i T_i (column sums of the normalised kept rows) as [P] FP32."""
        w = torch.zeros(self.n_vox, dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_ONES, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept,
                 n_rows=self.n_kept, row_scale=self.row_scale, w=w)
        return w

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """T x for the kept rows (normalised) as [n_kept] FP32."""
        f = torch.empty(max(self.n_kept, 1), dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_FWD, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept,
                 n_rows=self.n_kept, row_scale=self.row_scale, x=x, fwd_out=f)
        return f[: self.n_kept]

    def subset_weight(self, x, w, lo, hi, eps=None, x_perm=None, inv_perm=None, scale=1.0):
        """w (+)= scale * T_s^T (1 / T_s x) for kept rows [lo, hi)."""
        if hi <= lo:
            return
        ops.proj(ops.MODE_RATIO, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept, row_base=lo,
                 n_rows=hi - lo, row_scale=self.row_scale, x=x, eps=eps, x_perm=x_perm, inv_perm=inv_perm,
                 out_scale=scale, accumulate=True, w=w)

    def osem(self, x: torch.Tensor, s: torch.Tensor, n_iter: int, n_subsets: int, eps=None,
             em_mode: int = ops.EM_PLAIN) -> torch.Tensor:
        """List-mode OSEM (osem_list_mode semantics, one rotation, one energy); x updated in place."""
        w = torch.zeros_like(x)
        bounds = self.subset_bounds(n_subsets)
        for _ in range(n_iter):
            for lo, hi in bounds:
                if hi > lo:  # back-projection and EM update in one call (jscc_proj_em)
                    ops.proj(ops.MODE_RATIO, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept,
                             row_base=lo, n_rows=hi - lo, row_scale=self.row_scale, x=x, eps=eps, accumulate=True,
                             w=w, em=(x, s, em_mode, 1e-12, True))
                else:
                    ops.em_update(x, w, s, mode=em_mode, zero_after=w)
        return x


@dataclass
class LazyListStore:
    """List-mode data kept as per-event cone parameters (64 B per event) instead of
    T: every projection re-evaluates the rows in registers (jscc_cone_proj).

    The natural home of lists whose T exceeds HBM and of the lazy per-iteration
    recompute of distributed/python/recon_osem_dist_sparse_jsccsd_only_lazy.py
    (46-125); the kept set and the row normalisation are those of the stored form.
    """
    events: torch.Tensor      # [n_in, 16] event records
    ev_kept: torch.Tensor     # int32 record indices of the kept events, in order
    n_kept: int
    vox: ops.VoxelSoA
    sysmat: torch.Tensor | None
    sigma_model: int
    fixed_sigma: float

    @classmethod
    def from_rows(cls, rows: ComptonRows, sysmat=None, sigma_model=ops.SIG_FIXED, fixed_sigma=0.0):
        """Keep the event records of built (or statistics-only) rows; T itself may be dropped."""
        kept_rows = rows.kept[: rows.n_kept].long()
        ev_kept = rows.sel_idx[kept_rows].to(torch.int32).contiguous()
        return cls(rows.events, ev_kept, rows.n_kept, rows.vox, sysmat, sigma_model, fixed_sigma)

    def _call(self, mode, x, lo=0, hi=None, **kw):
        hi = self.n_kept if hi is None else hi
        ops.cone_proj(mode, self.events, self.ev_kept[lo:hi], hi - lo, self.vox, x, sysmat=self.sysmat,
                      sigma_model=self.sigma_model, fixed_sigma=self.fixed_sigma, **kw)

    def subset_bounds(self, k: int):
        return [(self.n_kept * i // k, self.n_kept * (i + 1) // k) for i in range(k)]

    def forward(self, x: torch.Tensor, dsig: bool = False):
        """(T x, d(T x)/d sigma or None) for the kept events."""
        f = torch.empty(max(self.n_kept, 1), dtype=torch.float32, device=x.device)
        g = torch.empty_like(f) if dsig else None
        self._call(ops.CP_FWD, x, fwd_out=f, dsig_out=g)
        return f[: self.n_kept], (g[: self.n_kept] if dsig else None)

    def adjoint(self, x: torch.Tensor, coef: torch.Tensor, dsig: bool = False):
        """(T^T coef, d(T x)/d sigma or None) in one recompute pass."""
        w = torch.zeros(self.vox.n, dtype=torch.float32, device=x.device)
        g = torch.empty(max(self.n_kept, 1), dtype=torch.float32, device=x.device) if dsig else None
        self._call(ops.CP_COEF, x, coef=coef.contiguous(), dsig_out=g, w=w)
        return w, (g[: self.n_kept] if dsig else None)

    def backproject_ones(self, like: torch.Tensor) -> torch.Tensor:
        ones = torch.ones(max(self.n_kept, 1), dtype=torch.float32, device=like.device)
        return self.adjoint(like, ones)[0]

    def subset_weight(self, x, w, lo, hi, eps=None, scale=1.0):
        if hi > lo:
            self._call(ops.CP_RATIO, x, lo, hi, eps=eps, out_scale=scale, accumulate=True, w=w)

    def osem(self, x, s, n_iter, n_subsets, eps=None, em_mode: int = ops.EM_PLAIN):
        w = torch.zeros_like(x)
        bounds = self.subset_bounds(n_subsets)
        for _ in range(n_iter):
            for lo, hi in bounds:
                self.subset_weight(x, w, lo, hi, eps=eps)
                ops.em_update(x, w, s, mode=em_mode, zero_after=w)
        return x
