"""Seeded synthetic workloads for the benchmark and the tests.

There are no measured SysMat / CntStat / List files for the reference's
detector, so the engine is exercised on synthetic inputs of the documented
shapes:

  sc_osem(...)        single-photon OSEM: polar grid, dense A with a sparse
                      log-normal value law, rod phantom, Poisson data
  compton_list(...)   two-hit Compton list: sources drawn from a rod phantom,
                      isotropic emission onto a cylinder, Klein-Nishina
                      scattering, second hit along the scattered ray
  polar_grid_3d(...), packed_coarse_rows(...)
                      the multi-slice factor grid and packed coarse event
                      rows of the sparse (coarse-T) mode
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

E0_511 = 0.511
ME_C2 = 0.511


def polar_grid(n_rings: int, n_angles: int, r_max: float, z: float = 0.0) -> np.ndarray:
    """[n_rings*n_angles, 3] ring-major polar grid, radii (k+1)*r_max/n_rings."""
    r = (np.arange(n_rings) + 1) * r_max / n_rings
    a = np.arange(n_angles) * 2.0 * np.pi / n_angles
    rr, aa = np.meshgrid(r, a, indexing="ij")
    return np.stack([rr * np.cos(aa), rr * np.sin(aa), np.full(rr.shape, z)], axis=-1).reshape(-1, 3)


def polar_grid_3d(radii, angle_groups, z_values) -> np.ndarray:
    """[L*Z, 3] multi-slice polar grid in the reference's factor layout: ring k (1-based)
    takes angle_groups[ceil(k / (n_rings / n_groups)) - 1] equally spaced angles from 0,
    rings in order within a slice, slices in order (genSysMatPolar_3D.m:84-108)."""
    radii = np.asarray(radii, np.float64)
    n, g = len(radii), len(angle_groups)
    layer = []
    for k, r in enumerate(radii, start=1):
        m = int(angle_groups[math.ceil(k / (n / g)) - 1])
        a = np.arange(m) * 2.0 * np.pi / m
        layer.append(np.stack([r * np.cos(a), r * np.sin(a)], axis=1))
    layer = np.concatenate(layer)
    return np.concatenate([np.hstack([layer, np.full((layer.shape[0], 1), z)]) for z in z_values])


def packed_coarse_rows(n_rows: int, n_coarse: int, n_crystals: int, device, seed: int = 0, chunk: int = 8192):
    """[n_rows, 1 + n_coarse] packed sparse event rows [float(cpnum1), t_c] with cpnum1
    uniform over 1..n_crystals and row-normalised t_c ~ U(0,1)^3 (synthetic, on `device`)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.empty((n_rows, n_coarse + 1), dtype=torch.float32, device=device)
    for lo in range(0, n_rows, chunk):
        hi = min(lo + chunk, n_rows)
        t = torch.rand((hi - lo, n_coarse), generator=g, device=device).pow_(3)
        t.div_(t.sum(1, keepdim=True))
        out[lo:hi, 1:] = t
        out[lo:hi, 0] = torch.randint(1, n_crystals + 1, (hi - lo,), generator=g, device=device).float()
    return out


def rod_phantom(coor: np.ndarray, r_max: float, rods=None, background: float = 1.0) -> np.ndarray:
    """Uniform disk (activity `background`) with hot rods [(cx, cy, radius, activity)]."""
    if rods is None:
        radii = [2.0, 2.8, 3.6, 4.4, 5.2, 6.0]
        rods = [(25.0 * math.cos(-i * math.pi / 3 + math.pi / 2), 25.0 * math.sin(-i * math.pi / 3 + math.pi / 2),
                 radii[i], 4.0) for i in range(6)]
    x = np.where(np.hypot(coor[:, 0], coor[:, 1]) <= r_max + 1e-9, background, 0.0)
    for cx, cy, rad, act in rods:
        x = np.where(np.hypot(coor[:, 0] - cx, coor[:, 1] - cy) <= rad, act, x)
    return x


@dataclass
class ScProblem:
    coor: np.ndarray      # [P, 3]
    A: np.ndarray         # [M, P] float32
    y: np.ndarray         # [M] float32 counts
    x_true: np.ndarray    # [P]
    subsets: list         # list of (start, end) row ranges


def sc_osem(n_rings=32, n_angles=128, r_max=50.0, n_rot=5, bins_per_rot=1024, density=0.3, mu=-9.0, sd=1.5,
            total_counts=1e6, seed=20260331) -> ScProblem:
    """Single-photon problem: A [n_rot*bins, P], one subset per rotation block."""
    rng = np.random.default_rng(seed)
    coor = polar_grid(n_rings, n_angles, r_max)
    P = coor.shape[0]
    M = n_rot * bins_per_rot
    mask = rng.random((M, P), dtype=np.float32) < density
    vals = np.exp(rng.normal(mu, sd, size=(M, P)).astype(np.float32))
    A = np.where(mask, vals, 0.0).astype(np.float32)
    x_true = rod_phantom(coor, r_max)
    f = A.astype(np.float64) @ x_true
    c = total_counts / f.sum()
    y = rng.poisson(c * f).astype(np.float32)
    subsets = [(k * bins_per_rot, (k + 1) * bins_per_rot) for k in range(n_rot)]
    return ScProblem(coor, A, y, x_true, subsets)


@dataclass
class ComptonList:
    coor: np.ndarray       # [P, 3] image grid
    detector: np.ndarray   # [2N, 3]: first hits then second hits
    list_origin: np.ndarray  # [N, 4] (cpnum1, e1, cpnum2, e2), 1-based crystal numbers
    x_true: np.ndarray
    e0: float


def _kn_cos_theta(rng, n: int, e0: float) -> np.ndarray:
    """Klein-Nishina scattering-angle cosines by rejection sampling."""
    k = e0 / ME_C2
    out = np.empty(0)
    # dsigma/dOmega ~ P^2 (P + 1/P - sin^2), P = 1/(1 + k(1-c)); max at c = 1 (value 1)
    while out.size < n:
        m = int((n - out.size) * 2.2) + 64
        c = rng.uniform(-1.0, 1.0, m)
        p = 1.0 / (1.0 + k * (1.0 - c))
        f = 0.5 * p * p * (p + 1.0 / p - (1.0 - c * c))
        acc = rng.uniform(0.0, 1.0, m) < f
        out = np.concatenate([out, c[acc]])
    return out[:n]


def compton_list(n_events=200_000, n_rings=64, n_angles=256, r_max=50.0, e0=E0_511, cyl_r=100.0, cyl_half_z=20.0,
                 d_min=10.0, d_max=40.0, rods=None, seed=20260332) -> ComptonList:
    """Two-hit list on a cylinder of radius cyl_r (|z| <= cyl_half_z)."""
    rng = np.random.default_rng(seed)
    coor = polar_grid(n_rings, n_angles, r_max)
    x_true = rod_phantom(coor, r_max, rods)
    prob = x_true / x_true.sum()
    src = np.empty((0, 3))
    dirs = np.empty((0, 3))
    # isotropic emission, keep photons whose first hit lies on the cylinder band
    while src.shape[0] < n_events:
        m = int((n_events - src.shape[0]) * 2.5) + 64
        vox = rng.choice(coor.shape[0], size=m, p=prob)
        rr = np.hypot(coor[vox, 0], coor[vox, 1])
        aa = np.arctan2(coor[vox, 1], coor[vox, 0])
        rr = rr + rng.uniform(-0.5, 0.5, m) * (r_max / n_rings)
        aa = aa + rng.uniform(-0.5, 0.5, m) * (2 * np.pi / n_angles)
        s = np.stack([rr * np.cos(aa), rr * np.sin(aa), np.zeros(m)], axis=1)
        u = rng.normal(size=(m, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        # intersection |s_xy + t u_xy| = cyl_r, t > 0
        a = u[:, 0] ** 2 + u[:, 1] ** 2
        b = 2 * (s[:, 0] * u[:, 0] + s[:, 1] * u[:, 1])
        c = s[:, 0] ** 2 + s[:, 1] ** 2 - cyl_r ** 2
        t = (-b + np.sqrt(np.maximum(b * b - 4 * a * c, 0.0))) / np.maximum(2 * a, 1e-300)
        hit = s + t[:, None] * u
        ok = (a > 1e-12) & (np.abs(hit[:, 2]) <= cyl_half_z)
        src = np.concatenate([src, hit[ok]])
        dirs = np.concatenate([dirs, u[ok]])
    pos1 = src[:n_events]
    u = dirs[:n_events]
    ct = _kn_cos_theta(rng, n_events, e0)
    st = np.sqrt(np.maximum(1 - ct * ct, 0.0))
    phi = rng.uniform(0, 2 * np.pi, n_events)
    # orthonormal frame around u
    helper = np.where(np.abs(u[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    e1v = np.cross(u, helper)
    e1v /= np.linalg.norm(e1v, axis=1, keepdims=True)
    e2v = np.cross(u, e1v)
    v = ct[:, None] * u + st[:, None] * (np.cos(phi)[:, None] * e1v + np.sin(phi)[:, None] * e2v)
    d = rng.uniform(d_min, d_max, n_events)
    pos2 = pos1 + d[:, None] * v
    e_scat = e0 / (1.0 + (e0 / ME_C2) * (1.0 - ct))
    e1 = e0 - e_scat
    e2 = e_scat
    det = np.concatenate([pos1, pos2], axis=0).astype(np.float32)
    lst = np.stack([np.arange(1, n_events + 1), e1, np.arange(n_events + 1, 2 * n_events + 1), e2],
                   axis=1).astype(np.float32)
    return ComptonList(coor.astype(np.float32), det, lst, x_true, e0)


def energy_thresholds(e0: float):
    """(max, min, sum) windows of main_plane.py:157-159 with the 511 keV sum window."""
    return 2 * e0 ** 2 / (0.511 + 2 * e0) - 0.001, 0.05, 0.46


def random_rods(rng, n_min=3, n_max=12, r_max=50.0, rad=(2.0, 8.0), contrast=(2.0, 8.0)):
    """3-12 hot rods of random radius and contrast inside the disk (batched-reconstruction phantoms)."""
    n = int(rng.integers(n_min, n_max + 1))
    rods = []
    for _ in range(n):
        rr = float(rng.uniform(rad[0], rad[1]))
        d = float(rng.uniform(0.0, r_max - rr))
        a = float(rng.uniform(0.0, 2 * np.pi))
        rods.append((d * math.cos(a), d * math.sin(a), rr, float(rng.uniform(contrast[0], contrast[1]))))
    return rods


def compton_list_device(n_events: int, coor: np.ndarray, x_true: np.ndarray, n_rings: int, n_angles: int,
                        device, seed: int, r_max=50.0, e0=E0_511, cyl_r=100.0, cyl_half_z=20.0, d_min=10.0,
                        d_max=40.0):
    """The generator of compton_list with torch on `device` (FP64), for lists of millions of events.
    Returns (detector [2N, 3] float32, list [N, 4] float32) device tensors, crystal numbers 1..2N."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=device)
    cxy = torch.from_numpy(coor[:, :2].astype(np.float64)).to(device)
    prob = torch.from_numpy((x_true / x_true.sum()).astype(np.float64)).to(device)
    srcs, dirs, have = [], [], 0
    while have < n_events:
        m = int((n_events - have) * 2.5) + 64
        vox = torch.multinomial(prob, m, replacement=True, generator=g)
        rr = torch.hypot(cxy[vox, 0], cxy[vox, 1]) + (torch.rand(m, generator=g, **f64) - 0.5) * (r_max / n_rings)
        aa = torch.atan2(cxy[vox, 1], cxy[vox, 0]) + (torch.rand(m, generator=g, **f64) - 0.5) * (
            2 * np.pi / n_angles)
        s = torch.stack([rr * torch.cos(aa), rr * torch.sin(aa), torch.zeros_like(rr)], dim=1)
        u = torch.randn((m, 3), generator=g, **f64)
        u = u / u.norm(dim=1, keepdim=True)
        a = u[:, 0] ** 2 + u[:, 1] ** 2
        b = 2 * (s[:, 0] * u[:, 0] + s[:, 1] * u[:, 1])
        c = s[:, 0] ** 2 + s[:, 1] ** 2 - cyl_r ** 2
        t = (-b + torch.sqrt(torch.clamp(b * b - 4 * a * c, min=0.0))) / torch.clamp(2 * a, min=1e-300)
        hit = s + t[:, None] * u
        ok = (a > 1e-12) & (hit[:, 2].abs() <= cyl_half_z)
        srcs.append(hit[ok])
        dirs.append(u[ok])
        have += int(ok.sum())
    pos1 = torch.cat(srcs)[:n_events]
    u = torch.cat(dirs)[:n_events]
    k = e0 / ME_C2
    cts, have = [], 0
    while have < n_events:
        m = int((n_events - have) * 2.2) + 64
        cc = torch.rand(m, generator=g, **f64) * 2 - 1
        p = 1.0 / (1.0 + k * (1.0 - cc))
        f = 0.5 * p * p * (p + 1.0 / p - (1.0 - cc * cc))
        acc = torch.rand(m, generator=g, **f64) < f
        cts.append(cc[acc])
        have += int(acc.sum())
    ct = torch.cat(cts)[:n_events]
    st = torch.sqrt(torch.clamp(1 - ct * ct, min=0.0))
    phi = torch.rand(n_events, generator=g, **f64) * (2 * np.pi)
    helper = torch.where(u[:, 2:3].abs() < 0.9, torch.tensor([[0.0, 0.0, 1.0]], **f64),
                         torch.tensor([[1.0, 0.0, 0.0]], **f64))
    e1v = torch.linalg.cross(u, helper)
    e1v = e1v / e1v.norm(dim=1, keepdim=True)
    e2v = torch.linalg.cross(u, e1v)
    v = ct[:, None] * u + st[:, None] * (torch.cos(phi)[:, None] * e1v + torch.sin(phi)[:, None] * e2v)
    d = d_min + (d_max - d_min) * torch.rand(n_events, generator=g, **f64)
    pos2 = pos1 + d[:, None] * v
    e_scat = e0 / (1.0 + (e0 / ME_C2) * (1.0 - ct))
    det = torch.cat([pos1, pos2]).float()
    idx = torch.arange(1, n_events + 1, **f64)
    lst = torch.stack([idx, e0 - e_scat, idx + n_events, e_scat], dim=1).float()
    return det, lst
