"""Coarse/fine polar-grid operator oracle (numpy, FP64).  TEST INFRASTRUCTURE ONLY.

Follows compton_sparse_ops.py (lpz7777777/JSCC_Recon_SPECT_Cylinder):
layout inference :47-92, 1-D reduce/expand operators :95-185, ring stride
choice :188-209, projector :212-290, row/image transforms :293-356, packing
:359-377 and materialisation :380-400.
"""
from __future__ import annotations

import math

import numpy as np


def infer_layout(coor, rtol=1e-4, ztol=1e-4):
    coor = np.asarray(coor, np.float64)
    z = coor[:, 2]
    zs, counts = [z[0]], [1]
    for v in z[1:]:
        if abs(v - zs[-1]) <= ztol:
            counts[-1] += 1
        else:
            zs.append(v)
            counts.append(1)
    if len(set(counts)) != 1:
        raise ValueError("inconsistent z layout")
    L = counts[0]
    r = np.sqrt((coor[:L, :2] ** 2).sum(1))
    lengths, radii = [], []
    i = 0
    while i < L:
        j = i + 1
        while j < L and abs(r[j] - r[i]) <= rtol:
            j += 1
        lengths.append(j - i)
        radii.append(r[i:j].mean())
        i = j
    return np.array(zs), L, lengths, np.array(radii)


def periodic_reduce(n, s):
    m = np.zeros((n // s, n))
    m[np.arange(n // s), np.arange(n // s) * s] = 1.0
    return m


def periodic_expand(n, s):
    nc = n // s
    m = np.zeros((n, nc))
    for f in range(n):
        left = (f // s) % nc
        right = (left + 1) % nc
        lpos, rpos, padj = left * s, left * s + s, f
        if right == 0 and f < lpos:
            padj += n
            rpos = n
        wr = (padj - lpos) / max(rpos - lpos, 1e-12)
        m[f, left] += 1 - wr
        m[f, right] += wr
    return m


def linear_reduce(n, s):
    nc = n // s
    m = np.zeros((nc, n))
    for k in range(nc):
        p = (k + 0.5) * s - 0.5
        left = max(0, min(n - 1, math.floor(p)))
        right = min(n - 1, left + 1)
        if right == left:
            m[k, left] = 1.0
        else:
            m[k, left] += 1 - (p - left)
            m[k, right] += p - left
    return m


def linear_expand(n, s):
    nc = n // s
    cpos = (np.arange(nc) + 0.5) * s - 0.5
    m = np.zeros((n, nc))
    for f in range(n):
        ins = int(np.searchsorted(cpos, float(f), side="left"))
        if ins <= 0:
            m[f, 0] = 1.0
        elif ins >= nc:
            m[f, nc - 1] = 1.0
        else:
            wr = (f - cpos[ins - 1]) / max(cpos[ins] - cpos[ins - 1], 1e-12)
            m[f, ins - 1] += 1 - wr
            m[f, ins] += wr
    return m


def ring_stride(length, rotate_num, target):
    if target <= 1:
        return 1
    if rotate_num and rotate_num > 0 and length % rotate_num == 0:
        step = length // rotate_num
        for c in range(target, 0, -1):
            if length % c == 0 and step % c == 0:
                return c
        return 1
    for c in range(target, 0, -1):
        if length % c == 0:
            return c
    return 1


def block_diag(blocks):
    rows = sum(b.shape[0] for b in blocks)
    cols = sum(b.shape[1] for b in blocks)
    out = np.zeros((rows, cols))
    i = j = 0
    for b in blocks:
        out[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return out


def build(coor_full, theta_stride=2, z_stride=2, rotate_num=None):
    zs, L, lengths, radii = infer_layout(coor_full)
    strides = [ring_stride(n, rotate_num, theta_stride) if rotate_num is not None else theta_stride for n in lengths]
    lr = block_diag([periodic_reduce(n, s) for n, s in zip(lengths, strides)])
    le = block_diag([periodic_expand(n, s) for n, s in zip(lengths, strides)])
    zr = linear_reduce(len(zs), z_stride)
    ze = linear_expand(len(zs), z_stride)
    parts = []
    for n, s, rad in zip(lengths, strides, radii):
        th = np.arange(n // s) * s * 2 * math.pi / n
        parts.append(np.stack([rad * np.cos(th), rad * np.sin(th)], axis=1))
    layer = np.concatenate(parts)
    cz = zr @ zs
    coor_coarse = np.concatenate([np.hstack([layer, np.full((layer.shape[0], 1), v)]) for v in cz])
    return {"ring_strides": strides, "layer_reduce": lr, "layer_expand": le, "z_reduce": zr, "z_expand": ze,
            "coor_coarse": coor_coarse, "fine_z": len(zs), "fine_layer": L, "coarse_layer": layer.shape[0]}


def reduce_rows(rows, pj):
    """rows [n, Zf*Lf] -> [n, Zc*Lc] (compton_sparse_ops.py:317-328)."""
    n = rows.shape[0]
    r3 = np.asarray(rows, np.float64).reshape(n, pj["fine_z"], pj["fine_layer"])
    t = r3 @ pj["layer_reduce"].T
    return np.einsum("cz,nzl->ncl", pj["z_reduce"], t).reshape(n, -1)


def upsample_rows(rows, pj):
    """rows [n, Zc*Lc] -> [n, Zf*Lf] (compton_sparse_ops.py:331-342)."""
    n = rows.shape[0]
    zc = pj["z_reduce"].shape[0]
    r3 = np.asarray(rows, np.float64).reshape(n, zc, pj["coarse_layer"])
    t = r3 @ pj["layer_expand"].T
    return np.einsum("fz,nzl->nfl", pj["z_expand"], t).reshape(n, -1)


def adjoint_rows(rows, pj):
    """Exact adjoint of upsample_rows (compton_sparse_ops.py:345-356)."""
    n = rows.shape[0]
    r3 = np.asarray(rows, np.float64).reshape(n, pj["fine_z"], pj["fine_layer"])
    t = r3 @ pj["layer_expand"]
    return np.einsum("fz,nfl->nzl", pj["z_expand"], t).reshape(n, -1)


def materialize(event_rows, A, pj):
    """(t_norm, valid) of compton_sparse_ops.py:380-400."""
    er = np.asarray(event_rows, np.float64)
    cp = np.clip(np.round(er[:, 0]).astype(np.int64) - 1, 0, A.shape[0] - 1)
    tf = np.nan_to_num(upsample_rows(er[:, 1:], pj) * np.asarray(A, np.float64)[cp], nan=0.0, posinf=0.0,
                       neginf=0.0)
    s = tf.sum(1)
    valid = np.isfinite(s) & (s > 0)
    return tf[valid] / s[valid, None], valid
