"""Drop-in for the reference's strict Compton cone back-projection.

Reference: process_list_plane_strict.py (lpz7777777/JSCC_Recon_SPECT_Cylinder).
`get_compton_backproj_list_single` keeps the reference's name, argument order,
defaults and return arity (process_list_plane_strict.py:151), and its
behaviour: energy smearing with two CPU `torch.randn` draws moved to `device`
(:158-161), the caller's list columns 1 and 3 mutated in place (:154-161),
the same event filters (:163-198, :30-36), the strict angular sigma model
(:39-52, :99-134), first-hit weighting (:236-240), NaN / zero-sum / effective
support filters (:241-248, :137-148) and row-normalised FP32 results returned
on the CPU (:249-253).  The work runs in the gfx950 kernels of libjscc_hip.

`build_compton_rows` is the engine-facing form: it leaves the rows
unnormalised and uncompacted in device memory (FP32 or BF16) with per-row
statistics and the kept-row index list, which the OSEM kernels consume
directly.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import ops

FRONT_LAYER_CRYSTAL_MM = (3.0, 3.0, 3.0)   # process_list_plane_strict.py:5
REAR_LAYER_CRYSTAL_MM = (2.0, 6.0, 2.0)    # :6
ELECTRON_REST_MEV = 0.511                  # :7
MIN_EVENT_EFFECTIVE_SUPPORT = 50.0         # :8


def get_coor_plane(pixel_num_x, pixel_num_y, pixel_l_x, pixel_l_y, fov_z):
    """Cartesian plane of voxel centres (process_list_plane_strict.py:11-21)."""
    xs = torch.linspace(-(pixel_num_x / 2 - 0.5) * pixel_l_x, (pixel_num_x / 2 - 0.5) * pixel_l_x, pixel_num_x)
    ys = torch.linspace(-(pixel_num_y / 2 - 0.5) * pixel_l_y, (pixel_num_y / 2 - 0.5) * pixel_l_y, pixel_num_y)
    coor = torch.empty((pixel_num_x, pixel_num_y, 3), dtype=torch.float32)
    coor[:, :, 0] = xs.reshape(1, -1)
    coor[:, :, 1] = ys.reshape(-1, 1)
    coor[:, :, 2] = fov_z
    return coor.reshape(-1, 3)


def _device_of(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError(f"device {dev} is not a GPU: this engine runs the cone kernels on the GPU only")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


@dataclass
class ComptonRows:
    """Cone rows of one list chunk, resident in device memory.

    store[r]    unnormalised row r (r < n_sel) of the events that passed the
                energy / kinematic / layer filters, FP32 or BF16, row stride ld
    stats[r]    (sum t, sum t^2, sum t_c, max t) in FP64
    kept[:n_kept]  rows that pass the NaN / zero / support filters, in order
    inv_sum[r]  1 / sum t (the reference's row normalisation)
    """
    store: torch.Tensor | None
    tc_store: torch.Tensor | None
    ld: int
    n_vox: int
    stats: torch.Tensor
    events: torch.Tensor
    sel_idx: torch.Tensor
    n_sel: int
    kept: torch.Tensor
    n_kept: int
    inv_sum: torch.Tensor
    device: torch.device
    vox: ops.VoxelSoA | None = None
    # sync_counts=False builds: the selected and kept counts as device int32 until resolve_counts() reads them
    # (n_sel, n_kept are -1 until then; store, stats, keep cover every input event)
    counts_dev: tuple | None = None
    # want_sens builds: column sums of the kept rows divided by their sums (the list-mode sensitivity), [P] FP32
    sens: torch.Tensor | None = None
    # FP16 rows (want_sens with store_dtype=torch.float16): per stored row, value / S = row_scale * code
    row_scale: torch.Tensor | None = None

    def resolve_counts(self) -> "ComptonRows":
        """Read deferred counts (one synchronisation, with the index check of the event pass)."""
        if self.counts_dev is not None:
            self.n_sel, self.n_kept = ops.read_counts_checked(*self.counts_dev)
            self.counts_dev = None
        return self

    def kept_arows(self) -> torch.Tensor:
        self.resolve_counts()
        ev_i32 = self.events.view(torch.int32)
        src = self.sel_idx[self.kept[:self.n_kept].long()].long()
        return ev_i32[src, 14].contiguous()

    def kept_event_index(self) -> torch.Tensor:
        """Input positions (int64) of the kept events, in order."""
        self.resolve_counts()
        ev_i32 = self.events.view(torch.int32)
        src = self.sel_idx[self.kept[:self.n_kept].long()].long()
        return ev_i32[src, 15].long()

    def normalized_t(self) -> torch.Tensor:
        self.resolve_counts()
        if self.row_scale is not None:  # FP16 codes: value / S = row_scale * code
            k = self.kept[:self.n_kept].long()
            return self.store[k, :self.n_vox].float() * self.row_scale[k, None]
        return ops.gather_normalize(self.store, self.n_vox, self.kept, self.n_kept, self.stats, 0)

    def normalized_tc(self) -> torch.Tensor:
        self.resolve_counts()
        return ops.gather_normalize(self.tc_store, self.n_vox, self.kept, self.n_kept, self.stats, 2)

    def normalized_single(self, sysmat_dev: torch.Tensor) -> torch.Tensor:
        self.resolve_counts()
        return ops.arows_normalize(sysmat_dev, self.kept_arows(), self.n_kept)


def _noise(noise, n: int, dev: torch.device):
    if noise is None or noise == "none":
        return None, None
    if noise == "cpu":  # process_list_plane_strict.py:160-161
        n1 = torch.randn(n).to(dev)
        n2 = torch.randn(n).to(dev)
        return n1, n2
    if noise == "device":  # process_list_plane_sparse.py:47-48
        n1 = torch.randn(n, device=dev)
        n2 = torch.randn(n, device=dev)
        return n1, n2
    n1, n2 = noise
    return n1.to(dev, torch.float32).contiguous(), n2.to(dev, torch.float32).contiguous()


def prepare_list(list_origin: torch.Tensor, dev: torch.device):
    """Device float32 row-contiguous list; returns (list_dev, copy_back) where copy_back
    propagates the in-place energy smearing to the caller's tensor."""
    if list_origin.ndim != 2 or list_origin.shape[1] < 4:
        raise ValueError("list_origin must have shape [num_events, >=4]")
    if list_origin.device == dev and list_origin.dtype == torch.float32 and list_origin.stride(1) == 1:
        return list_origin, None
    lst = list_origin.to(device=dev, dtype=torch.float32).contiguous()

    def copy_back():
        list_origin[:, 1] = lst[:, 1].to(list_origin.device, list_origin.dtype)
        list_origin[:, 3] = lst[:, 3].to(list_origin.device, list_origin.dtype)
    return lst, copy_back


def build_compton_rows(sysmat, detector, coor, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                       ene_threshold_max, ene_threshold_min, ene_threshold_sum, device, *, noise="cpu",
                       sigma_model="strict", fixed_sigma=None, store_dtype=torch.float32, want_tc=False,
                       taps=None, sysmat_dev=None, coor_dev=None, mutate_list=True,
                       sync_counts=True, want_sens=False) -> ComptonRows:
    """Cone rows of a list chunk on the GPU (see ComptonRows).

    noise: "cpu" (two torch.randn on the CPU generator, as the strict
    reference), "device" (as the sparse reference), "none", or a pair of
    N(0,1) tensors.  sigma_model "strict" or "fixed" (fixed_sigma in rad).
    sysmat may be None (first-hit factor 1).  mutate_list: write the smeared
    energies back into list_origin (the strict reference does, the sparse one
    does not).  sync_counts=False: no host synchronisation -- the cone launch reads the selected count on
    the device and covers every input event (rows past the count are dropped by the keep pass), and the
    selected and kept counts stay on the device (ComptonRows.counts_dev) until resolve_counts(), which also
    raises the index errors of the event pass.  want_sens: the cone pass also forms ComptonRows.sens, the column
    sums of the kept rows divided by their sums (the list-mode sensitivity of main_plane.py:348-355 before its
    scale), from each row as stored, so no second pass over T is needed (FP32 / BF16 store, no taps, no t_c).
    """
    dev = _device_of(device)
    if mutate_list:
        lst, copy_back = prepare_list(list_origin, dev)
    else:
        if list_origin.ndim != 2 or list_origin.shape[1] < 4:
            raise ValueError("list_origin must have shape [num_events, >=4]")
        lst, copy_back = list_origin.to(device=dev, dtype=torch.float32).clone(memory_format=torch.contiguous_format), None
    n = lst.shape[0]
    if sysmat_dev is None and sysmat is not None:
        sysmat_dev = ops.aligned_rows(sysmat, device=dev) if taps is None else ops.aligned_rows(sysmat, 1, dev)
    vox = coor_dev if coor_dev is not None else ops.voxel_soa(coor, dev)
    if sysmat_dev is not None and taps is None and sysmat_dev.shape[1] != vox.n:
        raise ValueError(f"sysmat has {sysmat_dev.shape[1]} columns for {vox.n} voxels")
    n1, n2 = _noise(noise, n, dev)
    det = detector.to(dev)
    events, flags = ops.events_prepare(lst, n1, n2, n1 is not None, det, delta_r1, delta_r2, e0, ene_resolution,
                                       ene_threshold_max, ene_threshold_min, ene_threshold_sum,
                                       n_sysmat_rows=sysmat_dev.shape[0] if sysmat_dev is not None else 0)
    if copy_back is not None and n1 is not None:
        copy_back()
    sel_idx, sel_count = ops.compact(flags)
    sig = ops.SIG_STRICT if sigma_model == "strict" else ops.SIG_FIXED
    if sig == ops.SIG_FIXED and not fixed_sigma:
        raise ValueError("sigma_model='fixed' needs fixed_sigma (rad)")
    if want_sens and (taps is not None or want_tc or store_dtype not in (torch.float32, torch.bfloat16,
                                                                         torch.float16)):
        raise ValueError("want_sens needs an FP32, BF16 or FP16 store, no tap tables and no t_c rows")
    if store_dtype == torch.float16 and not want_sens:
        raise ValueError("FP16 rows come from the cone pass with the sensitivity (want_sens=True); otherwise build "
                         "FP32 rows and quantise them (ListModeStore.from_rows(rows, 'fp16'))")

    def rows_of(n_rows, n_dev):
        if want_sens:
            store, stats, ld, sens, scale = ops.cone_rows_sens(events, sel_idx, n_rows, vox,
                                                               MIN_EVENT_EFFECTIVE_SUPPORT, sysmat=sysmat_dev,
                                                               sigma_model=sig, fixed_sigma=fixed_sigma or 0.0,
                                                               out_dtype=store_dtype, n_dev=n_dev)
            return store, None, stats, ld, sens, scale
        store, tc_store, stats, ld = ops.cone_rows(events, sel_idx, n_rows, vox, sysmat=sysmat_dev, taps=taps,
                                                   sigma_model=sig, fixed_sigma=fixed_sigma or 0.0,
                                                   out_dtype=store_dtype, want_tc=want_tc, n_dev=n_dev)
        return store, tc_store, stats, ld, None, None

    if not sync_counts and n > 0:
        store, tc_store, stats, ld, sens, scale = rows_of(n, sel_count)
        keep, inv_sum = ops.rows_keep(stats, n, MIN_EVENT_EFFECTIVE_SUPPORT, want_inv_sum=True)
        kept, kept_count = ops.compact(keep)
        return ComptonRows(store, tc_store, ld, vox.n, stats, events, sel_idx, -1, kept, -1, inv_sum, dev, vox,
                           (sel_count, kept_count), sens, scale)
    n_sel = ops.read_count_checked(sel_count)  # one synchronisation for the count and the index check
    store, tc_store, stats, ld, sens, scale = rows_of(n_sel, None)
    keep, inv_sum = ops.rows_keep(stats, n_sel, MIN_EVENT_EFFECTIVE_SUPPORT, want_inv_sum=True)
    kept, kept_count = ops.compact(keep)
    n_kept = int(kept_count.item())
    return ComptonRows(store, tc_store, ld, vox.n, stats, events, sel_idx, n_sel, kept, n_kept, inv_sum, dev, vox,
                       None, sens, scale)


def get_compton_backproj_list_single(sysmat, detector, coor, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                                     ene_threshold_max, ene_threshold_min, ene_threshold_sum, device,
                                     model_compton_generator=None, *, keep_on_device=False):
    """Row-normalised (t, t_compton, t_single) of the kept events, FP32 on the CPU
    (process_list_plane_strict.py:151-253); keep_on_device=True (engine option)
    returns them on `device` without the host copy."""
    if model_compton_generator is not None:
        raise NotImplementedError("model_compton_generator hooks are not supported by the GPU cone kernels")
    dev = _device_of(device)
    sysmat_dev = ops.aligned_rows(sysmat, device=dev)
    rows = build_compton_rows(sysmat, detector, coor, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                              ene_threshold_max, ene_threshold_min, ene_threshold_sum, dev, noise="cpu",
                              want_tc=True, sysmat_dev=sysmat_dev)
    t = rows.normalized_t()
    t_compton = rows.normalized_tc()
    t_single = rows.normalized_single(sysmat_dev)
    if keep_on_device:
        return t, t_compton, t_single
    return t.cpu(), t_compton.cpu(), t_single.cpu()
