"""Drop-in for the reference's coarse-grid cone builder (process_list_plane_sparse.py).

`get_compton_backproj_list_single_sparse` keeps the reference's signature and
result (process_list_plane_sparse.py:25-123): the strict cone model evaluated
on the projector's coarse voxels, the first-hit factor being the 2-tap-reduced
system-matrix row (:105), the same event and row filters, and packed rows
[float(cpnum1), row-normalised t_c] of the kept events returned on the CPU.
Its smearing noise is drawn on the device generator and, unlike the strict
builder, the caller's list is left unchanged (:45-48).
"""
from __future__ import annotations

import torch

from . import ops
from .compton_sparse_ops import _tables
from .process_list_plane_strict import MIN_EVENT_EFFECTIVE_SUPPORT, _device_of, build_compton_rows  # noqa: F401


def build_sparse_rows(sysmat, detector, sparse_projector, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                      ene_threshold_max, ene_threshold_min, ene_threshold_sum, device, *, noise="device",
                      sigma_model="strict", fixed_sigma=None):
    """Cone rows on the coarse grid (device resident, see ComptonRows) plus the
    per-row crystal numbers used by the packed format."""
    dev = _device_of(device)
    proj = sparse_projector
    t = _tables(proj, dev)
    if t.taps is None:
        raise ValueError("the projector's fine-to-coarse reduce has more than 2 taps per pixel")
    A = ops.aligned_rows(sysmat, 1, dev)
    if A.shape[1] != proj.fine_pixel_num:
        raise ValueError("sysmat size does not match sparse projector.")
    rows = build_compton_rows(A, detector, proj.coor_coarse, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                              ene_threshold_max, ene_threshold_min, ene_threshold_sum, dev, noise=noise,
                              sigma_model=sigma_model, fixed_sigma=fixed_sigma, store_dtype=torch.float32,
                              want_tc=True, taps=t.taps, sysmat_dev=A, mutate_list=False)
    return rows


def get_compton_backproj_list_single_sparse(sysmat, detector, sparse_projector, list_origin, delta_r1, delta_r2, e0,
                                            ene_resolution, ene_threshold_max, ene_threshold_min, ene_threshold_sum,
                                            device, model_compton_generator=None, *, keep_on_device=False):
    """Packed coarse rows (event_rows, None, None) on the CPU (process_list_plane_sparse.py:25-123);
    keep_on_device=True (engine option) returns them on `device` without the host copy."""
    if model_compton_generator is not None:
        raise NotImplementedError("model_compton_generator is not supported in sparse Compton mode.")
    dev = _device_of(device)
    pc = sparse_projector.coarse_pixel_num
    rows = build_sparse_rows(sysmat, detector, sparse_projector, list_origin, delta_r1, delta_r2, e0,
                             ene_resolution, ene_threshold_max, ene_threshold_min, ene_threshold_sum, dev)
    if rows.n_kept == 0:
        return torch.empty((0, pc + 1), dtype=torch.float32, device=dev if keep_on_device else "cpu"), None, None
    lst = list_origin.to(dev)
    src = rows.events.view(torch.int32)[rows.sel_idx[:rows.n_sel].long(), 15].long()
    lead = lst[src, 0].to(torch.float32).int().to(torch.float32).contiguous()  # float(cpnum1)
    packed = ops.gather_normalize(rows.tc_store, pc, rows.kept, rows.n_kept, rows.stats, 2, lead=lead)
    return (packed if keep_on_device else packed.cpu()), None, None
