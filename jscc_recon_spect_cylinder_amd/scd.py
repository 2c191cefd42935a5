"""Numeric steps on either side of OSEM: the equal-count single-photon data for
the SCD mode, count / list downsampling, and the sensitivity maps.

Reference: main_plane.py:300-336 (proj_d with as many counts as Compton events,
s = (1/R) sum_r colsum(A[:, RotMatInv[:, r]-1]), d = s * N_c / N_s) and
:348-355 (Sensi_d from the column sums of T renormalised to sum(s) * N_c / N_s);
sparse_main_utils.py:130-160 (downsample_projection_and_list).

The random draws are the reference's own host calls (torch.randperm on the
global CPU generator, in the same order), so a caller that seeds torch the same
way gets the same thinned data; the thinning itself (the reference expands every
count into a Python list and compares it against every bin) runs as one
histogram kernel on the implicit expansion (jscc_thin_counts).
"""
from __future__ import annotations

import torch

from . import ops


def _device(t: torch.Tensor, device) -> torch.device:
    if device is not None:
        return torch.device(device)
    return t.device if t.is_cuda else torch.device("cuda", torch.cuda.current_device())


def equal_count_projection(proj: torch.Tensor, compton_event_count_list, device=None) -> torch.Tensor:
    """proj_d [bins, R]: per rotation i, compton_event_count_list[i] counts drawn without replacement from
    the round(proj[:, i]) counts (main_plane.py:300-310).  Returns FP32 on the CPU, like the reference."""
    dev = _device(proj, device)
    P = proj.to(dev, torch.float32)
    out = torch.zeros(P.shape, dtype=torch.float32, device=dev)
    for i in range(P.shape[1]):
        n = int(torch.round(P[:, i].double()).sum().item())
        perm = torch.randperm(n)  # the reference's draw (global CPU generator)
        sel = perm[0:compton_event_count_list[i]]
        ops.thin_counts(P[:, i], sel, out=out[:, i])
    ops.raise_on_index_error(dev)
    return out.cpu()


def downsample_projection_and_list(proj: torch.Tensor, list_origin, ds: float, device=None):
    """sparse_main_utils.py:130-160: keep a fraction ds of the counts of every rotation (drawn without
    replacement) and of the events of every rotation's list (first int(n*ds) of a permutation, at least 1)."""
    if ds >= 0.999999:
        return proj, list_origin
    dev = _device(proj, device)
    proj_out = proj.clone()
    lists = [chunk.clone() for chunk in list_origin]
    P = proj.to(dev, torch.float32)
    col = torch.empty(P.shape[0], dtype=torch.float32, device=dev)
    for r in range(proj.shape[1]):
        n = int(torch.round(P[:, r].double()).sum().item())
        if n > 0:
            selected = int(torch.round(proj[:, r].sum() * ds).item())
            selected = min(selected, n)
            sel = torch.randperm(n)[:selected]
            ops.thin_counts(P[:, r], sel, out=col)
            proj_out[:, r] = col.to(proj_out.device, proj_out.dtype)
        else:
            proj_out[:, r] = 0
        lst = lists[r]
        if lst.size(0) > 0:
            k = int(lst.size(0) * ds)
            k = min(max(k, 1), lst.size(0))
            lists[r] = lst[torch.randperm(lst.size(0))[:k], :]
    ops.raise_on_index_error(dev)
    return proj_out, lists


def sensitivity_single(sysmat: torch.Tensor, rotmat_inv: torch.Tensor, device=None) -> torch.Tensor:
    """s = (1/R) sum_r colsum(sysmat[:, RotMatInv[:, r]-1]) as [P, 1] FP32 on the device (main_plane.py:320-323),
    column sums in FP64."""
    dev = _device(sysmat, device)
    return ops.sensitivity(sysmat.to(dev), rotmat_inv).view(-1, 1)


def sensitivity_compton_scaled(s: torch.Tensor, compton_event_count: int, single_event_count: int) -> torch.Tensor:
    """d = s * N_c / N_s (main_plane.py:324)."""
    return s * compton_event_count / single_event_count


def sensitivity_compton_from_t(t_all, s: torch.Tensor, compton_event_count_total: int,
                               single_event_count_total: int) -> torch.Tensor:
    """Sensi_d (main_plane.py:348-355): for every T (a row-normalised tensor, or an engine ListModeStore),
    its column sums rescaled to sum(s) * N_c / N_s, summed over the T's.  Returns [P, 1] FP32."""
    from .engine import ListModeStore
    dev = s.device if s.is_cuda else torch.device("cuda", torch.cuda.current_device())
    sv = s.to(dev, torch.float32).reshape(-1).contiguous()
    s_sum = ops.sum_f64(sv)
    out = torch.zeros_like(sv)
    for t in t_all:
        if isinstance(t, ListModeStore):
            cs = t.backproject_ones()
        else:
            mat = ops.proj_ready(t.to(dev, torch.float32), ops.DT_F32)
            cs = torch.zeros(mat.shape[1], dtype=torch.float32, device=dev)
            ops.proj(ops.MODE_ONES, mat, w=cs)
        cs_sum = ops.sum_f64(cs)
        scale = (s_sum / cs_sum * compton_event_count_total / single_event_count_total).float()
        out += cs * scale
    return out.view(-1, 1)
