"""Polar image grid and rotation maps (oracle).

Restates Reproduction/Step1_GenerateFactors/genSysMatPolar_3D.m:84-157:
rings at radii r_k, each ring with theta_num(k) equally spaced angles starting
at 0, the per-ring angle count taken from groups `theta_num_value` by
ceil(k / (n_rings / n_groups)); the full grid repeats the layer for every z
(layer-major).  RotMat[:, r] maps a rotated pixel to its source pixel
(1-based): within each ring, mod((r)*interval + i, theta_num) with
interval = theta_num / rotate_num; RotMatInv is the column-wise argsort
(inverse permutation), also 1-based.
"""
from __future__ import annotations

import math

import numpy as np


def ring_counts(n_rings: int, theta_num_value) -> list[int]:
    groups = len(theta_num_value)
    return [int(theta_num_value[math.ceil(k / (n_rings / groups)) - 1]) for k in range(1, n_rings + 1)]


def polar_layer(r_values, theta_num_value) -> np.ndarray:
    """[L, 2] (x, y) of one z-layer, ring-major then angle-minor (genSysMatPolar_3D.m:84-99)."""
    counts = ring_counts(len(r_values), theta_num_value)
    pts = []
    for r, n in zip(r_values, counts):
        ang = np.arange(n) * 360.0 / n
        # MATLAB cosd/sind are exact at multiples of 90 degrees
        c = np.cos(np.deg2rad(ang))
        s = np.sin(np.deg2rad(ang))
        q = np.round(ang / 90.0)
        exact = np.isclose(ang, q * 90.0, rtol=0, atol=1e-12)
        c = np.where(exact, np.array([1.0, 0.0, -1.0, 0.0])[(q.astype(int)) % 4], c)
        s = np.where(exact, np.array([0.0, 1.0, 0.0, -1.0])[(q.astype(int)) % 4], s)
        pts.append(np.stack([r * c, r * s], axis=1))
    return np.concatenate(pts, axis=0)


def polar_full(r_values, theta_num_value, z_values) -> np.ndarray:
    """[L*Z, 3] grid, z-layer major (genSysMatPolar_3D.m:101-108)."""
    layer = polar_layer(r_values, theta_num_value)
    return np.concatenate([np.hstack([layer, np.full((layer.shape[0], 1), z)]) for z in z_values], axis=0)


def rotmat_layer(r_values, theta_num_value, rotate_num: int) -> np.ndarray:
    """[L, R] 1-based RotMat of one layer (genSysMatPolar_3D.m:111-132)."""
    counts = ring_counts(len(r_values), theta_num_value)
    cols = []
    for rot in range(rotate_num):
        parts = []
        off = 0
        for n in counts:
            interval = n / rotate_num
            idx = np.mod(rot * interval + np.arange(n), n) + 1 + off
            parts.append(idx)
            off += n
        cols.append(np.concatenate(parts))
    return np.stack(cols, axis=1).astype(np.int64)


def inverse_columns(rotmat: np.ndarray) -> np.ndarray:
    """Column-wise 1-based argsort (genSysMatPolar_3D.m:134-139)."""
    return np.stack([np.argsort(rotmat[:, r], kind="stable") + 1 for r in range(rotmat.shape[1])], axis=1)


def full_maps(rot_layer: np.ndarray, inv_layer: np.ndarray, n_z: int):
    """Per-layer maps stacked for n_z layers with offsets (genSysMatPolar_3D.m:146-152)."""
    L = rot_layer.shape[0]
    rot = np.concatenate([rot_layer + z * L for z in range(n_z)], axis=0)
    inv = np.concatenate([inv_layer + z * L for z in range(n_z)], axis=0)
    return rot, inv


# Grids of the reference's factor sets (radii mm, angle groups, z mm, rotations)
GRIDS = {
    "511keV_RotateNum20": (np.arange(6, 151, 6), [20, 40, 60, 80], np.arange(-28.5, 28.6, 3.0), 20),
    "511keV_RotateNum10": (np.arange(6, 151, 6), list(range(10, 81, 10)), np.arange(-28.5, 28.6, 3.0), 10),
    "511keV_RotateNum5": (np.arange(6, 151, 6), list(range(10, 81, 10)), np.arange(-28.5, 28.6, 3.0), 5),
    "511keV_RotateNum40": (np.arange(6, 151, 6), [40, 80], np.arange(-28.5, 28.6, 3.0), 40),
}


def reference_grid(name: str):
    r, th, z, R = GRIDS[name]
    coor = polar_full(r, th, z)
    rl = rotmat_layer(r, th, R)
    il = inverse_columns(rl)
    rot, inv = full_maps(rl, il, len(z))
    return coor, rot, inv
