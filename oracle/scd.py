"""SCD-data and sensitivity oracle (numpy + torch's CPU generator).  TEST INFRASTRUCTURE ONLY.

Follows main_plane.py:300-336, 348-355 and sparse_main_utils.py:130-160
(lpz7777777/JSCC_Recon_SPECT_Cylinder).  The random permutations are the
reference's own torch.randperm calls on the global CPU generator, made in the
same order, so a seeded run reproduces the reference's draws.
"""
from __future__ import annotations

import numpy as np
import torch


def thin_literal(counts, perm_prefix):
    """The reference's literal algorithm (main_plane.py:303-310): expand bin j into round(count_j) copies of
    j, pick the copies at the permutation prefix, count each bin with a comparison.  Tiny inputs only."""
    expanded = [j for j in range(len(counts)) for _ in range(round(float(counts[j])))]
    picked = [expanded[p] for p in perm_prefix]
    return np.array([sum(1 for v in picked if v == j) for j in range(len(counts))], dtype=np.float32)


def thin(counts, perm_prefix):
    """The same with the expansion implicit: copy p lies in the bin whose inclusive cumulative count first
    exceeds p."""
    c = np.rint(np.asarray(counts, np.float64)).astype(np.int64)  # rint: round half to even, as Python round
    cum = np.cumsum(c)
    bins = np.searchsorted(cum, np.asarray(perm_prefix, np.int64), side="right")
    return np.bincount(bins, minlength=len(c)).astype(np.float32)


def equal_count_projection(proj, compton_event_count_list):
    """proj_d of main_plane.py:300-310 (proj [bins, R] float32)."""
    proj = np.asarray(proj, np.float32)
    out = np.zeros_like(proj)
    for i in range(proj.shape[1]):
        n = int(np.rint(proj[:, i].astype(np.float64)).sum())
        perm = torch.randperm(n).numpy()
        out[:, i] = thin(proj[:, i], perm[0:compton_event_count_list[i]])
    return out


def downsample_projection_and_list(proj, lists, ds):
    """sparse_main_utils.py:130-160."""
    if ds >= 0.999999:
        return proj, lists
    proj = torch.as_tensor(proj).clone()
    lists = [torch.as_tensor(l).clone() for l in lists]
    for r in range(proj.shape[1]):
        col = proj[:, r].numpy()
        n = int(np.rint(col.astype(np.float64)).sum())
        new = np.zeros_like(col)
        if n > 0:
            k = int(torch.round(proj[:, r].sum() * ds).item())
            k = min(k, n)
            new = thin(col, torch.randperm(n)[:k].numpy())
        proj[:, r] = torch.from_numpy(new)
        lst = lists[r]
        if lst.size(0) > 0:
            k = min(max(int(lst.size(0) * ds), 1), lst.size(0))
            lists[r] = lst[torch.randperm(lst.size(0))[:k], :]
    return proj, lists


def sensitivity_compton_from_t(t_list, s, n_c, n_s):
    """main_plane.py:348-355 in FP64: sum_t colsum(t) * sum(s) / sum(colsum(t)) * n_c / n_s."""
    s = np.asarray(s, np.float64).ravel()
    out = np.zeros_like(s)
    for t in t_list:
        cs = np.asarray(t, np.float64).sum(0)
        out += cs * s.sum() / cs.sum() * n_c / n_s
    return out
