"""OSEM oracle (numpy, FP64).  TEST INFRASTRUCTURE ONLY.

Follows recon_osem_plane.py:8-125 (lpz7777777/JSCC_Recon_SPECT_Cylinder),
recon_osem_local_cntstat.py:6-45, recon_osem_plane_sparse.py:60-66,
recon_osem_local_sparse_jsccsd_only.py:13-28, 58-100,
distributed/python/recon_osem_dist.py:46-51 and main_plane.py:320-323.
Images are [P] vectors; permutation tables are the reference's 1-based
[P, R] RotMat / RotMatInv.
"""
from __future__ import annotations

import numpy as np


def split_bounds_even(total: int, parts: int):
    """recon_osem_plane.py:8-18 (floor bounds)."""
    return [(total * i // parts, total * (i + 1) // parts) for i in range(parts)]


def chunk_bounds(total: int, parts: int):
    """torch.chunk: ceil-sized chunks, possibly fewer than `parts`."""
    size = -(-total // parts) if total > 0 else 0
    out = []
    start = 0
    while start < total:
        out.append((start, min(total, start + size)))
        start += size
    return out


def weight_single(A, y, x, eps=None):
    """A^T (y / A x); eps clamps A x from below (recon_osem_local_cntstat.py:6-8)."""
    A = np.asarray(A, np.float64)
    f = A @ np.asarray(x, np.float64)
    if eps is not None:
        f = np.maximum(f, eps)
    with np.errstate(divide="ignore", invalid="ignore"):
        return A.T @ (np.asarray(y, np.float64) / f)


def weight_compton(T, x, eps=None):
    """T^T (1 / T x) (recon_osem_plane.py:33-42)."""
    T = np.asarray(T, np.float64)
    f = T @ np.asarray(x, np.float64)
    if eps is not None:
        f = np.maximum(f, eps)
    with np.errstate(divide="ignore", invalid="ignore"):
        return T.T @ (1.0 / f)


def rot_gather(x, perm_col):
    """torch.index_select(x, 0, perm[:, r] - 1)."""
    return np.asarray(x)[np.asarray(perm_col, np.int64) - 1]


def em_plain(x, w, s):
    with np.errstate(divide="ignore", invalid="ignore"):
        return x * w / s


def nan_to_num(v, nan=0.0, posinf=0.0, neginf=0.0):
    return np.nan_to_num(v, nan=nan, posinf=posinf, neginf=neginf)


def em_safe(x, w, s, eps=1e-12):
    """recon_osem_plane_sparse.py:60-66."""
    ws = nan_to_num(w)
    ss = nan_to_num(s)
    valid = ss > eps
    out = np.zeros_like(x, dtype=np.float64)
    out[valid] = x[valid] * np.maximum(ws[valid], 0.0) / ss[valid]
    return out


def em_apply(x, w, s, eps=1e-12):
    """distributed/python/recon_osem_dist.py:46-51."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ss = np.where(np.isnan(s), s, np.maximum(s, eps))
        u = nan_to_num(w / ss, nan=1.0, posinf=1e6, neginf=0.0)
        return nan_to_num(x * u, nan=0.0, posinf=1e12, neginf=0.0)


def osem_bin_mode(A_subsets, y_subsets, rot, inv, x, s, eps=None):
    """recon_osem_plane.py:45-61 (one energy): A_subsets[k] [M_k, P], y_subsets[k] [M_k, R]."""
    R = rot.shape[1]
    x = np.asarray(x, np.float64).copy()
    for A, Y in zip(A_subsets, y_subsets):
        w = np.zeros_like(x)
        for r in range(R):
            wt = weight_single(A, Y[:, r], rot_gather(x, rot[:, r]), eps)
            w += rot_gather(wt, inv[:, r])
        x = em_plain(x, w, s)
    return x


def osem_list_mode(T_subsets, rot, inv, x, s, eps=None):
    """recon_osem_plane.py:63-80 (one energy): T_subsets[k][r] = list of blocks of rotation r."""
    x = np.asarray(x, np.float64).copy()
    for t_sub in T_subsets:
        w = np.zeros_like(x)
        for r, blocks in enumerate(t_sub):
            for T in blocks:
                wt = weight_compton(T, rot_gather(x, rot[:, r]), eps)
                w += rot_gather(wt, inv[:, r])
        x = em_plain(x, w, s)
    return x


def osem_joint_mode(A_subsets, y_subsets, T_subsets, rot, inv, x, s, alpha, eps=None):
    """recon_osem_plane.py:100-125 (one energy)."""
    R = rot.shape[1]
    x = np.asarray(x, np.float64).copy()
    for A, Y, t_sub in zip(A_subsets, y_subsets, T_subsets):
        wc = np.zeros_like(x)
        ws = np.zeros_like(x)
        for r in range(R):
            wt = weight_single(A, Y[:, r], rot_gather(x, rot[:, r]), eps)
            ws += rot_gather(wt, inv[:, r])
            for T in t_sub[r]:
                wt = weight_compton(T, rot_gather(x, rot[:, r]), eps)
                wc += rot_gather(wt, inv[:, r])
        x = em_plain(x, (2 - alpha) * wc + alpha * ws, s)
    return x


def sensitivity(A, inv):
    """(1/R) sum_r colsum(A[:, inv[:, r]-1]) (main_plane.py:320-323)."""
    A = np.asarray(A, np.float64)
    cs = A.sum(axis=0)
    R = inv.shape[1]
    return sum(cs[np.asarray(inv[:, r], np.int64) - 1] for r in range(R)) / R
