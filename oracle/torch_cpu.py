"""The reference's FP32 torch CPU path for the JSCCD list-mode workload.  TEST
INFRASTRUCTURE ONLY: bench.py's CPU-baseline leg times it on the host cores;
nothing in the product imports it, and the FP64 oracle (cone.py, osem.py) is
the checker it is compared against.

It keeps the reference's torch operation sequence in FP32 on the CPU, with the
two substitutions the workload defines:
* process_list_plane_strict.py:151-253 (energy smearing :157-161, windows
  :163-176, kinematic filter :30-36 / :177-179, gathers and same-layer filter
  :181-198, cone weight :202-228, NaN / zero filter :241-248, normalisation and
  support filter :137-148), with the strict angular sigma (:217-226) replaced
  by the workload's constant sigma and no system-matrix factor (the workload's
  first-hit factor is 1, :236-240);
* osem_list_mode (recon_osem_plane.py:63-80) with get_weight_compton (:33-42)
  for one rotation and one energy, over split_tensor_even subsets (:8-18) of
  the kept rows, and s = the column sums of the normalised rows.
Events are processed in chunks (the reference chunks its list per worker,
process_list_plane_strict.py:271-272); every filter decides per event, so the
chunking does not change the rows.
"""
from __future__ import annotations

import torch

ELECTRON_REST_MEV = 0.511
MIN_EVENT_EFFECTIVE_SUPPORT = 50.0
COS_CLAMP = 1.0 - 1e-7


def _theta(e1, e0):
    c = 1 - ((ELECTRON_REST_MEV * e1) / ((e0 - e1) * e0))
    return torch.acos(torch.clamp(c, -COS_CLAMP, COS_CLAMP))


def cone_rows_fixed_sigma(detector, coor, lst, e0, ene_resolution, thr_max, thr_min, thr_sum, noise1, noise2,
                          sigma):
    """Normalised kept rows [n_kept, P] (FP32) and the kept positions within `lst` (int64)."""
    idx = torch.arange(lst.shape[0])
    cpnum1 = lst[:, 0].int()
    cpnum2 = lst[:, 2].int()
    e1 = lst[:, 1].clone()
    e2 = lst[:, 3].clone()
    if noise1 is not None:
        e1 += e1 * ene_resolution / 2.355 * (e0 / e1) ** 0.5 * noise1
        e2 += e2 * ene_resolution / 2.355 * (e0 / e2) ** 0.5 * noise2
    flag = (e1 < thr_max) & (e1 > thr_min) & (e2 > thr_min) & ((e1 + e2) > thr_sum)
    cpnum1, cpnum2, e1, idx = cpnum1[flag], cpnum2[flag], e1[flag], idx[flag]
    c_raw = 1 - ((ELECTRON_REST_MEV * e1) / ((e0 - e1) * e0))
    ok = (c_raw > -1.0 + 1e-6) & (c_raw < 1.0 - 1e-6)
    cpnum1, cpnum2, e1, idx = cpnum1[ok], cpnum2[ok], e1[ok], idx[ok]
    pos = detector[:, :3]
    pos1 = pos[cpnum1 - 1, :]
    pos2 = pos[cpnum2 - 1, :]
    layer = torch.abs(pos1[:, 1] - pos2[:, 1]) > 0.1
    pos1, pos2, e1, idx = pos1[layer], pos2[layer], e1[layer], idx[layer]

    v01 = pos1.unsqueeze(1) - coor.unsqueeze(0)
    v12 = (pos2 - pos1).unsqueeze(1)
    d01 = torch.norm(v01, dim=2)
    d12 = torch.norm(v12, dim=2)
    theta = _theta(e1, e0)
    kn = e0 / (e0 - e1) + (e0 - e1) / e0
    bcos = (v01 * v12).sum(2) / torch.clamp(d01 * d12, min=1e-7)
    beta = torch.acos(torch.clamp(bcos, -COS_CLAMP, COS_CLAMP))
    t = torch.exp(-(beta - theta.unsqueeze(-1)) ** 2 / (2 * sigma ** 2)) * (kn.unsqueeze(-1) - torch.sin(beta) ** 2)

    good = (torch.isnan(t).sum(dim=1) + (t.sum(dim=1) == 0)) == 0
    t, idx = t[good], idx[good]
    tn = t / t.sum(dim=1, keepdim=True)
    stable = 1.0 / torch.sum(tn ** 2, dim=1) >= MIN_EVENT_EFFECTIVE_SUPPORT
    return tn[stable], idx[stable]


def list_osem(T, n_iter, n_subsets):
    """x after n_iter list-mode OSEM iterations over the kept rows T (FP32) with s = sum of the rows."""
    s = T.sum(dim=0, keepdim=True).t()
    x = torch.ones((T.shape[1], 1), dtype=T.dtype)
    n = T.shape[0]
    bounds = [(n * i // n_subsets, n * (i + 1) // n_subsets) for i in range(n_subsets)]
    for _ in range(n_iter):
        for lo, hi in bounds:
            tb = T[lo:hi]
            w = torch.matmul(tb.transpose(0, 1), 1 / torch.matmul(tb, x))
            x = x * w / s
    return x.view(-1)
