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
        if code in (ops.DT_FP8, ops.DT_FP16):
            codes, scale = ops.quantize_rows(rows.store, rows.n_vox, rows.n_sel, rows.stats, code)
            return cls(codes, code, rows.n_vox, rows.kept, rows.n_kept, scale)
        if ops.dtype_code(rows.store.dtype) != code:
            raise ValueError("the rows were built in another storage dtype")
        return cls(rows.store, code, rows.n_vox, rows.kept, rows.n_kept, rows.inv_sum)

    def nbytes_kept(self) -> int:
        per = {ops.DT_F32: 4, ops.DT_BF16: 2, ops.DT_FP8: 1, ops.DT_FP16: 2}[self.dtype_c]
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
class SystemStore:
    """A single-photon system matrix A [M, P] (detector bins x pixels) held for
    the fused projection in FP32, BF16, FP16 (per-row scale = the row's largest
    value / 2^14) or FP8-e4m3 (per-row scale = largest value / 448), always
    accumulated in FP32.  The reference keeps A in
    FP32 (main_plane.py:174-196); the low-precision forms are the engine's
    precision modes.  Subsets are contiguous bin ranges with split_tensor_even
    bounds; the data y is indexed by bin, so no copy of A or y is made."""
    mat: torch.Tensor
    dtype_c: int
    n_vox: int
    n_rows: int
    row_scale: torch.Tensor | None

    @classmethod
    def from_dense(cls, A: torch.Tensor, dtype="bf16") -> "SystemStore":
        ops.require_gpu(A, "sysmat")
        code = ops.dtype_code(dtype)
        M, P = A.shape
        if A.dtype != torch.float32:
            A = A.float()
        ld = ops.round_up(P, 64)
        if code in (ops.DT_FP8, ops.DT_FP16):
            A32 = A if (A.stride(1) == 1 and A.stride(0) % 4 == 0) else A.contiguous()
            codes, scale = ops.quantize_rows(A32, P, M, None, code)
            return cls(codes, code, P, M, scale)
        tdt = torch.float32 if code == ops.DT_F32 else torch.bfloat16
        mat = torch.zeros((M, ld), dtype=tdt, device=A.device)
        mat[:, :P] = A
        return cls(mat, code, P, M, None)

    def subset_bounds(self, k: int):
        return [(self.n_rows * i // k, self.n_rows * (i + 1) // k) for i in range(k)]

    def column_sums(self) -> torch.Tensor:
        """s = A^T 1 (the sensitivity of main_plane.py:320-323 for explicitly stacked rotations)."""
        w = torch.zeros(self.n_vox, dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_ONES, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, n_rows=self.n_rows,
                 row_scale=self.row_scale, w=w)
        return w

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        f = torch.empty(self.n_rows, dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_FWD, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, n_rows=self.n_rows,
                 row_scale=self.row_scale, x=x, fwd_out=f)
        return f

    def subset_weight(self, x, w, lo, hi, y, eps=None, scale=1.0):
        """w (+)= scale * A_s^T (y_s / A_s x) for bins [lo, hi) (get_weight_single, recon_osem_plane.py:28-30)."""
        if hi <= lo:
            return
        ops.proj(ops.MODE_RATIO, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, row_base=lo, n_rows=hi - lo,
                 row_scale=self.row_scale, x=x, y=y, y_by_row=True, eps=eps, out_scale=scale, accumulate=True, w=w)

    def osem(self, x, y, s, n_iter, n_subsets, eps=None, em_mode: int = ops.EM_PLAIN) -> torch.Tensor:
        """SC OSEM over bin subsets (osem_bin_mode with one rotation block per subset); x updated in place."""
        w = torch.zeros_like(x)
        for _ in range(n_iter):
            for lo, hi in self.subset_bounds(n_subsets):
                if hi > lo:
                    ops.proj(ops.MODE_RATIO, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, row_base=lo,
                             n_rows=hi - lo, row_scale=self.row_scale, x=x, y=y, y_by_row=True, eps=eps,
                             accumulate=True, w=w, em=(x, s, em_mode, 1e-12, True))
        return x


def joint_osem(a: SystemStore, y: torch.Tensor, t: "ListModeStore", x: torch.Tensor, s_joint: torch.Tensor,
               alpha: float, n_iter: int, n_subsets: int, eps=None, em_mode: int = ops.EM_PLAIN) -> torch.Tensor:
    """JSCC-SD OSEM (osem_joint_mode, recon_osem_plane.py:100-125): per subset
    w = (2 - alpha) * T_s^T(1/T_s x) + alpha * A_s^T(y_s/A_s x), x <- x * w / s_joint.
    Subset k takes the k-th split_tensor_even range of the bins and of the
    kept events.  Two launches per subset: the A pass accumulates the
    single-photon weight, the T pass finishes the Compton weight and applies the
    combined update in its reduction (jscc_proj_em); x is updated in place."""
    ws = torch.zeros_like(x)
    wc = torch.zeros_like(x)
    a_b = a.subset_bounds(n_subsets)
    t_b = t.subset_bounds(n_subsets)
    for _ in range(n_iter):
        for (alo, ahi), (tlo, thi) in zip(a_b, t_b):
            a.subset_weight(x, ws, alo, ahi, y, eps=eps)
            if thi > tlo:
                ops.proj(ops.MODE_RATIO, t.mat, dtype_c=t.dtype_c, n_cols=t.n_vox, rows=t.kept, row_base=tlo,
                         n_rows=thi - tlo, row_scale=t.row_scale, x=x, eps=eps, accumulate=True, w=wc,
                         em=(x, s_joint, em_mode, 1e-12, True, 2.0 - alpha, ws, alpha))
            else:
                ops.em_update(x, wc, s_joint, mode=em_mode, c1=2.0 - alpha, w2=ws, c2=alpha, zero_after=wc)
                ws.zero_()
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
