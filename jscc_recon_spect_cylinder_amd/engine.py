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
    # rows built with sync_counts=False: the kept count as a device int32 (n_kept is -1 until resolve())
    # and its upper bound; backproject_ones and osem then launch with device-side row ranges
    count_dev: torch.Tensor | None = None
    cap: int = 0

    @classmethod
    def from_rows(cls, rows: ComptonRows, dtype="bf16") -> "ListModeStore":
        code = ops.dtype_code(dtype)
        if code == ops.DT_FP16 and rows.row_scale is not None:  # FP16 codes written by the cone pass itself
            if rows.counts_dev is not None:
                return cls(rows.store, code, rows.n_vox, rows.kept, -1, rows.row_scale, rows.counts_dev[1],
                           rows.kept.shape[0])
            return cls(rows.store, code, rows.n_vox, rows.kept, rows.n_kept, rows.row_scale)
        if code in (ops.DT_FP8, ops.DT_FP16):
            rows.resolve_counts()  # quantisation takes the selected rows
            codes, scale = ops.quantize_rows(rows.store, rows.n_vox, rows.n_sel, rows.stats, code)
            return cls(codes, code, rows.n_vox, rows.kept, rows.n_kept, scale)
        if ops.dtype_code(rows.store.dtype) != code:
            raise ValueError("the rows were built in another storage dtype")
        if rows.counts_dev is not None:
            return cls(rows.store, code, rows.n_vox, rows.kept, -1, rows.inv_sum, rows.counts_dev[1],
                       rows.kept.shape[0])
        return cls(rows.store, code, rows.n_vox, rows.kept, rows.n_kept, rows.inv_sum)

    def resolve(self) -> "ListModeStore":
        """Read a deferred kept count (one synchronisation)."""
        if self.count_dev is not None:
            (self.n_kept,) = ops.read_counts_checked(self.count_dev)
            self.count_dev = None
        return self

    def nbytes_kept(self) -> int:
        self.resolve()
        per = {ops.DT_F32: 4, ops.DT_BF16: 2, ops.DT_FP8: 1, ops.DT_FP16: 2}[self.dtype_c]
        return self.n_kept * self.n_vox * per

    def subset_bounds(self, k: int):
        self.resolve()
        return [(self.n_kept * i // k, self.n_kept * (i + 1) // k) for i in range(k)]

    def backproject_ones(self) -> torch.Tensor:
        """sum_i T_i (column sums of the normalised kept rows) as [P] FP32."""
        w = torch.zeros(self.n_vox, dtype=torch.float32, device=self.mat.device)
        if self.count_dev is not None:
            ops.ProjLauncher(ops.MODE_ONES, self.mat, n_rows_max=self.cap, dtype_c=self.dtype_c, n_cols=self.n_vox,
                             rows=self.kept, row_scale=self.row_scale, w=w,
                             count=(self.count_dev, self.cap, 1)).launch_subset(0)
            return w
        ops.proj(ops.MODE_ONES, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept,
                 n_rows=self.n_kept, row_scale=self.row_scale, w=w)
        return w

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """T x for the kept rows (normalised) as [n_kept] FP32."""
        self.resolve()
        f = torch.empty(max(self.n_kept, 1), dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_FWD, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept,
                 n_rows=self.n_kept, row_scale=self.row_scale, x=x, fwd_out=f)
        return f[: self.n_kept]

    def subset_weight(self, x, w, lo, hi, eps=None, x_perm=None, inv_perm=None, scale=1.0):
        """w (+)= scale * T_s^T (1 / T_s x) for kept rows [lo, hi)."""
        self.resolve()
        if hi <= lo:
            return
        ops.proj(ops.MODE_RATIO, self.mat, dtype_c=self.dtype_c, n_cols=self.n_vox, rows=self.kept, row_base=lo,
                 n_rows=hi - lo, row_scale=self.row_scale, x=x, eps=eps, x_perm=x_perm, inv_perm=inv_perm,
                 out_scale=scale, accumulate=True, w=w)

    def _iteration(self, x, s, n_subsets, eps, em_mode):
        """The launches of one list-mode OSEM iteration (one fused projection + EM launch per subset) and
        the buffers they read: the weight and a workspace owned by this iteration.  With a device-side kept
        count the subset bounds are taken on the device (no host read); otherwise an empty subset applies
        the update with a zero weight, as osem_list_mode does for an empty split (recon_osem_plane.py:63-80)."""
        w = torch.zeros_like(x)
        if self.count_dev is not None:
            per = -(-self.cap // n_subsets)
            ws = torch.empty(ops.lib_proj_workspace_bytes(max(per, 1), self.n_vox), dtype=torch.uint8,
                             device=x.device)
            step = ops.ProjLauncher(ops.MODE_RATIO, self.mat, n_rows_max=self.cap, dtype_c=self.dtype_c,
                                    n_cols=self.n_vox, rows=self.kept, row_scale=self.row_scale, x=x, eps=eps,
                                    accumulate=True, w=w, em=(x, s, em_mode, 1e-12, True),
                                    count=(self.count_dev, self.cap, n_subsets), workspace=ws)

            def run():
                for i in range(n_subsets):
                    step.launch_subset(i)
            return run, (w, ws, s, self.mat, self.kept, self.row_scale, self.count_dev, step)
        bounds = self.subset_bounds(n_subsets)
        n_max = max((hi - lo for lo, hi in bounds), default=0)
        ws = torch.empty(ops.lib_proj_workspace_bytes(max(n_max, 1), self.n_vox), dtype=torch.uint8,
                         device=x.device)
        step = ops.ProjLauncher(ops.MODE_RATIO, self.mat, n_rows_max=n_max, dtype_c=self.dtype_c,
                                n_cols=self.n_vox, rows=self.kept, row_scale=self.row_scale, x=x, eps=eps,
                                accumulate=True, w=w, em=(x, s, em_mode, 1e-12, True),
                                workspace=ws) if n_max > 0 else None

        def run():
            for lo, hi in bounds:
                if hi > lo:
                    step.launch(lo, hi - lo)
                else:
                    ops.em_update(x, w, s, mode=em_mode, zero_after=w)
        return run, (w, ws, s, self.mat, self.kept, self.row_scale, step)

    def capture(self, x, s, n_subsets, eps=None, em_mode: int = ops.EM_PLAIN) -> "GraphedIteration":
        """One list-mode OSEM iteration captured as a HIP graph (x updated in place at each replay); the
        returned object owns every buffer the graph reads -- keep it while replaying.  Rows built with
        device-side counts capture without any host read."""
        run, bufs = self._iteration(x, s, n_subsets, eps, em_mode)
        x_keep = x.clone()
        run()  # warm-up on the real buffers, then restore the image
        x.copy_(x_keep)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            run()
        return GraphedIteration(g, bufs, x, s)

    def osem(self, x: torch.Tensor, s: torch.Tensor, n_iter: int, n_subsets: int, eps=None,
             em_mode: int = ops.EM_PLAIN, persistent: bool = False, graph: bool = False) -> torch.Tensor:
        """List-mode OSEM (osem_list_mode semantics, one rotation, one energy); x updated in place:
        one fused projection + one reduction/update launch per subset (jscc_proj_em, or
        jscc_proj_em_counted for rows built with device-side counts).  graph=True captures one
        iteration as a HIP graph and replays it (10 x 8 subsets over 164758 BF16 rows x 16384:
        8.75 vs 8.85 ms eager).  persistent=True runs all subsets and iterations in one cooperative
        launch (jscc_osem_persistent) instead; on MI355X its two grid barriers per subset cost more
        than the launches they replace (153 vs 117 us per subset at 2e4 BF16 rows x 16384), so it
        is not the default."""
        if graph and not persistent:
            gi = self.capture(x, s, n_subsets, eps, em_mode)
            for _ in range(n_iter):
                gi.replay()
            return x
        if persistent:
            self.resolve()
            if self.n_kept > 0:
                off = torch.tensor([b[0] for b in self.subset_bounds(n_subsets)] + [self.n_kept],
                                   dtype=torch.int64, device=self.mat.device)
                ops.osem_persistent(self.mat, self.dtype_c, self.n_vox, self.kept, self.row_scale, off, n_iter, x,
                                    s, eps=eps, em_mode=em_mode)
                ops.raise_on_index_error(self.mat.device)
                return x
        run, _bufs = self._iteration(x, s, n_subsets, eps, em_mode)
        for _ in range(n_iter):
            run()
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

    def _iteration(self, x, y, s, n_subsets, eps, em_mode):
        """The launches of one OSEM iteration (one fused projection + EM launch pair per subset) and the
        buffers they read: the weight and a workspace owned by this iteration.  An empty subset (more
        subsets than bins) applies the update with a zero weight, x <- x * 0 / s, as the reference's
        osem_bin_mode does for an empty block of rows (recon_osem_plane.py:45-54)."""
        bounds = self.subset_bounds(n_subsets)
        w = torch.zeros_like(x)
        n_max = max((hi - lo for lo, hi in bounds), default=0)
        ws = torch.empty(ops.lib_proj_workspace_bytes(max(n_max, 1), self.n_vox), dtype=torch.uint8,
                         device=x.device)

        # validated once; each subset launch only sets its row range
        step = ops.ProjLauncher(ops.MODE_RATIO, self.mat, n_rows_max=n_max, dtype_c=self.dtype_c,
                                n_cols=self.n_vox, row_scale=self.row_scale, x=x, y=y, y_by_row=True, eps=eps,
                                accumulate=True, w=w, em=(x, s, em_mode, 1e-12, True),
                                workspace=ws) if n_max > 0 else None

        def run():
            for lo, hi in bounds:
                if hi > lo:
                    step.launch(lo, hi - lo)
                else:
                    ops.em_update(x, w, s, mode=em_mode, zero_after=w)
        return run, (w, ws, y, s, self.mat, self.row_scale, step)

    def capture(self, x, y, s, n_subsets, eps=None, em_mode: int = ops.EM_PLAIN) -> "GraphedIteration":
        """One SC OSEM iteration captured as a HIP graph (x updated in place at each replay).  The
        returned object owns every buffer the graph reads; keep it while replaying."""
        run, bufs = self._iteration(x, y, s, n_subsets, eps, em_mode)
        x_keep = x.clone()
        run()  # warm-up on the real buffers, then restore the image
        x.copy_(x_keep)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            run()
        return GraphedIteration(g, bufs, x, s)

    def osem(self, x, y, s, n_iter, n_subsets, eps=None, em_mode: int = ops.EM_PLAIN,
             graph: bool = False) -> torch.Tensor:
        """SC OSEM over bin subsets (osem_bin_mode with one rotation block per subset); x updated in place.
        graph=True captures one iteration as a HIP graph and replays it (the launches of a small A are
        short enough for the launch overhead to matter)."""
        if graph:
            gi = self.capture(x, y, s, n_subsets, eps, em_mode)
            for _ in range(n_iter):
                gi.replay()
            return x
        run, _bufs = self._iteration(x, y, s, n_subsets, eps, em_mode)
        for _ in range(n_iter):
            run()
        return x



@dataclass
class FrameStore:
    """K single-photon data sets that share one system matrix A [M, P] (the time frames of a dynamic
    acquisition, bootstrap replicates): their OSEM subset steps as GEMMs on the matrix cores.  The
    reference reconstructs one data set per call (osem_bin_mode, recon_osem_plane.py:45-54, with
    get_weight_single, :28-30); per frame this computes the same update, for all frames in one pass over
    the subset's rows in each direction:
        F = A_s X  (forward, [n_s, K]),  C = Y_s / F,  W = A_s^T C  (back, [P, K]),  X <- X * W / s.
    A is held in BF16 twice, row-major for the forward GEMM and transposed for the back-projection (each
    GEMM reduces over a contiguous index); images and coefficients enter the GEMMs as two BF16 terms
    (hi = bf16(v), lo = bf16(v - hi)), accumulation FP32.  Frames are padded to 16, 32 or 64 with
    zero-data frames.  Subsets are contiguous bin ranges with split_tensor_even bounds, as SystemStore.osem."""
    mat: torch.Tensor    # [M, ldp] BF16, ldp = P rounded up to ops.FRAMES_ALIGN, zero-padded
    matt: torch.Tensor   # [P, ldm] BF16, the transpose, ldm = M rounded up to ops.FRAMES_ALIGN
    n_vox: int
    n_rows: int

    @classmethod
    def from_dense(cls, A: torch.Tensor) -> "FrameStore":
        ops.require_gpu(A, "sysmat")
        M, P = A.shape
        ldp, ldm = ops.round_up(P, ops.FRAMES_ALIGN), ops.round_up(M, ops.FRAMES_ALIGN)
        mat = torch.zeros((M, ldp), dtype=torch.bfloat16, device=A.device)
        mat[:, :P] = A
        return cls(mat, ops.frames_transpose(mat, M, P, ldm), P, M)

    def subset_bounds(self, k: int):
        return [(self.n_rows * i // k, self.n_rows * (i + 1) // k) for i in range(k)]

    def column_sums(self) -> torch.Tensor:
        """s = A^T 1 of the stored (BF16) matrix, FP32 (the fused projection's column-sum mode)."""
        w = torch.zeros(self.n_vox, dtype=torch.float32, device=self.mat.device)
        ops.proj(ops.MODE_ONES, self.mat, dtype_c=ops.DT_BF16, n_cols=self.n_vox, n_rows=self.n_rows, w=w)
        return w

    def osem(self, X: torch.Tensor, Y: torch.Tensor, s: torch.Tensor, n_iter: int, n_subsets: int, eps=None,
             em_mode: int = ops.EM_PLAIN) -> torch.Tensor:
        """OSEM of every frame: X [P, K] (frames contiguous) updated in place from the data Y [M, K], with the
        sensitivity s [P] shared by the frames."""
        P, M = self.n_vox, self.n_rows
        if X.ndim != 2 or X.shape[0] != P or Y.ndim != 2 or Y.shape != (M, X.shape[1]):
            raise ValueError(f"frames: X must be [{P}, K] and Y [{M}, K]")
        for t, nm in ((X, "images"), (Y, "data"), (s, "sensitivity")):
            ops.require_gpu(t, nm)
        K = X.shape[1]
        if K > ops.FRAMES_K[-1]:
            raise ValueError(f"frames: at most {ops.FRAMES_K[-1]} frames per store call")
        Kp = next(k for k in ops.FRAMES_K if k >= K)
        dev = X.device
        ldp, ldm = self.mat.stride(0), self.matt.stride(0)
        xp = torch.ones((P, Kp), dtype=torch.float32, device=dev)
        xp[:, :K] = X
        yp = torch.zeros((M, Kp), dtype=torch.float32, device=dev)
        yp[:, :K] = Y
        s32 = s.float().contiguous()
        bounds = self.subset_bounds(n_subsets)
        al = ops.FRAMES_ALIGN
        pads = [((lo // al) * al, ops.round_up(hi, al)) for lo, hi in bounds]
        n_max = max((hi - lo for lo, hi in bounds), default=0)
        c_pad = max((b - a for a, b in pads), default=al)
        splits_f = [ops.frames_split_count(hi - lo, ldp, dev) if hi > lo else 0 for lo, hi in bounds]
        splits_b = [ops.frames_split_count(P, b - a, dev) if hi > lo else 0 for (lo, hi), (a, b) in zip(bounds, pads)]
        part = torch.empty(max([n * Kp * (hi - lo) for n, (lo, hi) in zip(splits_f, bounds)] +
                               [n * Kp * P for n in splits_b] + [1]), dtype=torch.float32, device=dev)
        xt = torch.empty((2, Kp, ldp), dtype=torch.bfloat16, device=dev)
        ct = torch.empty((2, Kp, c_pad), dtype=torch.bfloat16, device=dev)
        ops.frames_split(xp, P, ldp, xt[0], xt[1])
        for _ in range(n_iter):
            for (lo, hi), (a0, a1), nf, nb in zip(bounds, pads, splits_f, splits_b):
                if hi <= lo:
                    ops.frames_update(None, 0, s32, xp, em_mode, xt[0], xt[1])
                    continue
                ops.frames_gemm(self.mat[lo:], hi - lo, ldp, xt[0], xt[1], nf, part)
                ops.frames_ratio(part, nf, hi - lo, a1 - a0, lo - a0, yp[lo:], eps, ct[0], ct[1])
                ops.frames_gemm(self.matt[:, a0:], P, a1 - a0, ct[0], ct[1], nb, part)
                ops.frames_update(part, nb, s32, xp, em_mode, xt[0], xt[1])
        X.copy_(xp[:, :K])
        return X

def joint_osem(a: SystemStore, y: torch.Tensor, t: "ListModeStore", x: torch.Tensor, s_joint: torch.Tensor,
               alpha: float, n_iter: int, n_subsets: int, eps=None, em_mode: int = ops.EM_PLAIN) -> torch.Tensor:
    """JSCC-SD OSEM (osem_joint_mode, recon_osem_plane.py:100-125): per subset
    w = (2 - alpha) * T_s^T(1/T_s x) + alpha * A_s^T(y_s/A_s x), x <- x * w / s_joint.
    Subset k takes the k-th split_tensor_even range of the bins and of the
    kept events.  When A and T hold the same storage type, one projection
    launch per subset streams the A rows (numerators alpha * y) and then the T
    rows (numerators 2 - alpha) into the same partial sums and its reduction
    applies the update (jscc_proj_joint_em); otherwise the A pass accumulates
    the single-photon weight and the T pass finishes the Compton weight and
    applies the combined update (jscc_proj_em).  x is updated in place."""
    ws = torch.zeros_like(x)
    wc = torch.zeros_like(x)
    a_b = a.subset_bounds(n_subsets)
    t_b = t.subset_bounds(n_subsets)
    if a.mat.dtype == t.mat.dtype and a.dtype_c == t.dtype_c:
        for _ in range(n_iter):
            for (alo, ahi), (tlo, thi) in zip(a_b, t_b):
                if ahi > alo or thi > tlo:
                    ops.proj_joint_em(a.mat, t.mat, dtype_c=a.dtype_c, n_cols=a.n_vox, x=x, w=ws, s=s_joint,
                                      c1=alpha, c2=2.0 - alpha, n_rows=ahi - alo, row_base=alo,
                                      row_scale=a.row_scale, y=y, y_by_row=True, n_rows2=thi - tlo, rows2=t.kept,
                                      row_base2=tlo, row_scale2=t.row_scale, eps=eps, em_mode=em_mode)
                else:
                    ops.em_update(x, ws, s_joint, mode=em_mode, zero_after=ws)
        return x
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
        rows.resolve_counts()
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


@dataclass
class BatchedLazyStore:
    """B independent list-mode reconstructions on one GPU, T never stored: every
    subset step evaluates the cone rows of all B lists in registers in ONE
    launch (jscc_cone_proj_batched) whose work-groups split the concatenated
    rows evenly, so lists of 1e4 and 2e6 events share the chip, followed by one
    reduction launch that applies the B EM updates.  One iteration (2 launches
    per subset) can be captured as a HIP graph and replayed.  Kept sets and row
    normalisation are those of the stored form (process_list_plane_strict.py:
    137-148, 241-248); subsets are split_tensor_even ranges of each list."""
    events: torch.Tensor          # concatenated event records of all reconstructions
    kept: list                    # per reconstruction: int32 record indices of its kept events, in order
    vox: ops.VoxelSoA
    sysmat: torch.Tensor | None
    sigma_model: int
    fixed_sigma: float
    groups_per_cu: int = 0
    tile: int = 0

    @classmethod
    def from_rows(cls, rows_list, sysmat=None, sigma_model=ops.SIG_FIXED, fixed_sigma=0.0):
        """Concatenate the event records of per-reconstruction ComptonRows (built with store_dtype=None:
        statistics and kept sets only)."""
        evs, kept, base = [], [], 0
        for r in rows_list:
            r.resolve_counts()
            k = r.sel_idx[r.kept[: r.n_kept].long()].to(torch.int32) + base
            evs.append(r.events)
            kept.append(k.contiguous())
            base += r.events.shape[0]
        return cls(torch.cat(evs).contiguous(), kept, rows_list[0].vox, sysmat, sigma_model, fixed_sigma)

    @property
    def n_recon(self) -> int:
        return len(self.kept)

    @property
    def n_kept(self) -> list:
        return [int(k.numel()) for k in self.kept]

    def tables(self, n_subsets: int):
        """Per subset: (concatenated int32 record list, int64 row offsets [B+1]) on the device."""
        dev = self.events.device
        out = []
        for i in range(n_subsets):
            parts, off = [], [0]
            for k in self.kept:
                n = k.numel()
                lo, hi = n * i // n_subsets, n * (i + 1) // n_subsets
                parts.append(k[lo:hi])
                off.append(off[-1] + hi - lo)
            out.append((torch.cat(parts).contiguous(), torch.tensor(off, dtype=torch.int64, device=dev)))
        return out

    def _call(self, mode, ev_idx, row_off, **kw):
        ops.cone_proj_batched(mode, self.events, ev_idx, row_off, self.vox, sysmat=self.sysmat,
                              sigma_model=self.sigma_model, fixed_sigma=self.fixed_sigma,
                              groups_per_cu=kw.pop("groups_per_cu", self.groups_per_cu),
                              tile=kw.pop("tile", self.tile), **kw)

    def sensitivity(self) -> torch.Tensor:
        """[B, P] list-mode sensitivities: column sums of each reconstruction's normalised rows."""
        (ev_idx, off), = self.tables(1)
        w = torch.zeros((self.n_recon, ops.round_up(self.vox.n, 4)), dtype=torch.float32,
                        device=self.events.device)
        self._call(ops.BP_SUM, ev_idx, off, w=w)
        return w

    def autotune(self, x: torch.Tensor, s: torch.Tensor, n_subsets: int, candidates=(1, 2, 4, 8),
                 tiles=None, reps: int = 2):
        """Pick the launch shape of the projection: every tile shape of ops.cone_batch_tiles (threads per
        work-group, voxel quads per thread, rows per step) times every work-groups-per-CU candidate, each
        timed on one subset step (the fastest of `reps` timed runs after a warm-up) on a copy of the
        images.  The choice is kept (tile, groups_per_cu) and returned with its time in ms."""
        ev_idx, off = self.tables(n_subsets)[0]
        n_tiles = len(ops.cone_batch_tiles(self.vox.n))
        tiles = range(n_tiles) if tiles is None else tiles
        best, best_t = (self.tile, self.groups_per_cu), float("inf")
        xs = x.clone()
        for t in tiles:
            for c in candidates:
                ws = ops.cone_batch_workspace(x.device, ev_idx.numel(), self.n_recon, self.vox.n, c)
                self._call(ops.BP_OSEM, ev_idx, off, x=xs, sens=s, workspace=ws, groups_per_cu=c, tile=t)  # warm
                ts = []
                for _ in range(reps):
                    xs.copy_(x)
                    t0 = torch.cuda.Event(enable_timing=True)
                    t1 = torch.cuda.Event(enable_timing=True)
                    t0.record()
                    self._call(ops.BP_OSEM, ev_idx, off, x=xs, sens=s, workspace=ws, groups_per_cu=c, tile=t)
                    t1.record()
                    t1.synchronize()
                    ts.append(t0.elapsed_time(t1))
                if min(ts) < best_t:
                    best, best_t = (t, c), min(ts)
        self.tile, self.groups_per_cu = best
        return self.tile, self.groups_per_cu, best_t

    def _iteration(self, x, s, n_subsets, eps, em_mode):
        """The launches of one OSEM iteration and the buffers they read (tables, workspace)."""
        tabs = self.tables(n_subsets)
        n_max = max(t[0].numel() for t in tabs)
        ws = ops.cone_batch_workspace(x.device, n_max, self.n_recon, self.vox.n, self.groups_per_cu)

        def run():
            for ev_idx, off in tabs:
                if ev_idx.numel():
                    self._call(ops.BP_OSEM, ev_idx, off, x=x, sens=s, eps=eps, em_mode=em_mode, workspace=ws)
        return run, (tabs, ws)

    def capture(self, x: torch.Tensor, s: torch.Tensor, n_subsets: int, eps=None,
                em_mode: int = ops.EM_PLAIN) -> "GraphedIteration":
        """One OSEM iteration of all B images captured as a HIP graph (x updated in place at each
        replay).  The returned object owns every buffer the graph reads; keep it while replaying."""
        run, bufs = self._iteration(x, s, n_subsets, eps, em_mode)
        x_keep = x.clone()
        run()  # warm-up on the real buffers (library and allocator state), then restore the images
        x.copy_(x_keep)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            run()
        return GraphedIteration(g, bufs, x, s)

    def osem(self, x: torch.Tensor, s: torch.Tensor, n_iter: int, n_subsets: int, eps=None,
             em_mode: int = ops.EM_PLAIN, graph: bool = False) -> torch.Tensor:
        """OSEM of all B images x [B, >= P rounded up to 4] (updated in place; the columns past P are read,
        never written, and must be finite) with sensitivities s [B, >=P].
        graph=True captures one iteration (2 launches per subset) as a HIP graph and replays it."""
        if graph:
            gi = self.capture(x, s, n_subsets, eps, em_mode)
            for _ in range(n_iter):
                gi.replay()
            return x
        run, _bufs = self._iteration(x, s, n_subsets, eps, em_mode)
        for _ in range(n_iter):
            run()
        return x


@dataclass
class GraphedIteration:
    """A captured OSEM iteration together with the device buffers its kernels read."""
    graph: torch.cuda.CUDAGraph
    buffers: tuple
    x: torch.Tensor
    s: torch.Tensor

    def replay(self) -> None:
        self.graph.replay()
