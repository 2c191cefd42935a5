"""Compton cone back-projection oracle (numpy).  TEST INFRASTRUCTURE ONLY.

Follows process_list_plane_strict.py (lpz7777777/JSCC_Recon_SPECT_Cylinder):
  * event stage in FP32 with the reference's operation order, so that the
    surviving event set and the smeared energies are bit-identical
    (:157-198, :24-36);
  * everything per (event, voxel) in FP64, with the reference's formulas and
    clamps (:39-52, :99-134, :201-248, :137-148).
"""
from __future__ import annotations

import numpy as np

F = np.float32
ELECTRON_REST_MEV = 0.511          # process_list_plane_strict.py:7
MIN_EVENT_EFFECTIVE_SUPPORT = 50.0  # :8
FRONT = (3.0, 3.0, 3.0)             # :5
REAR = (2.0, 6.0, 2.0)              # :6
# torch clamps an FP32 tensor, so the bounds are the FP32 roundings
COS_MAX = float(F(1.0 - 1e-7))


def theta_from_e1(e1, e0, ee=ELECTRON_REST_MEV):
    """process_list_plane_strict.py:24-27 (FP64)."""
    c = 1.0 - (ee * np.asarray(e1, np.float64)) / ((e0 - np.asarray(e1, np.float64)) * e0)
    return np.arccos(np.clip(c, -COS_MAX, COS_MAX))


def klein_nishina(e1, e0):
    """process_list_plane_strict.py:210."""
    e1 = np.asarray(e1, np.float64)
    return e0 / (e0 - e1) + (e0 - e1) / e0


def angle_sigma_ene_pm(e1, e0, ene_resolution):
    """(sigma_plus, sigma_minus) of process_list_plane_strict.py:39-49 (FP64)."""
    e1 = np.asarray(e1, np.float64)
    se = e1 * ene_resolution / 2.355 * np.sqrt(e0 / e1)
    lo = np.clip(e1 - se, 1e-7, e0 - 1e-7)
    hi = np.clip(e1 + se, 1e-7, e0 - 1e-7)
    th = theta_from_e1(e1, e0)
    sm = np.maximum(th - theta_from_e1(lo, e0), 1e-7)
    sp = np.maximum(theta_from_e1(hi, e0) - th, 1e-7)
    return sp, sm


def detector_sigma_sq(detector, extra_sigma):
    """process_list_plane_strict.py:55-96: per-crystal position variance [D, 3]."""
    det = np.asarray(detector, F)
    ay = np.abs(det[:, 1])
    layers = np.unique(ay)
    if layers.size != 4:
        return np.full((det.shape[0], 3), max(extra_sigma, 0.0) ** 2, dtype=F)
    front = np.array([s * s / 12.0 for s in FRONT], dtype=F)
    rear = np.array([s * s / 12.0 for s in REAR], dtype=F)
    out = np.zeros((det.shape[0], 3), dtype=F)
    for i in range(4):
        out[ay == layers[i]] = front if i < 3 else rear
    if extra_sigma > 0:
        out = (out + F(extra_sigma ** 2)).astype(F)
    return out


def prepare_events(list_origin, detector, e0, ene_resolution, thr_max, thr_min, thr_sum, noise1=None, noise2=None):
    """FP32 event stage of process_list_plane_strict.py:152-198.

    Returns dict with the smeared list (as the reference leaves it), the input
    positions of the surviving events and their gathered quantities.
    """
    lst = np.array(list_origin, dtype=F, copy=True)
    n = lst.shape[0]
    cp1 = lst[:, 0].astype(np.int32)  # .int(): truncation toward zero
    cp2 = lst[:, 2].astype(np.int32)
    e1 = lst[:, 1].copy()
    e2 = lst[:, 3].copy()
    if noise1 is not None:
        n1 = np.asarray(noise1, F)
        n2 = np.asarray(noise2, F)
        # sigma = ((e*R)/2.355) * sqrt(reciprocal(e)*E0)  (torch: E0/e == reciprocal(e)*E0)
        s1 = ((e1 * F(ene_resolution)) / F(2.355)) * np.sqrt((F(1.0) / e1) * F(e0))
        s2 = ((e2 * F(ene_resolution)) / F(2.355)) * np.sqrt((F(1.0) / e2) * F(e0))
        e1 = (e1 + (s1 * n1).astype(F)).astype(F)
        e2 = (e2 + (s2 * n2).astype(F)).astype(F)
        lst[:, 1] = e1
        lst[:, 3] = e2
    keep = (e1 < F(thr_max)) & (e1 > F(thr_min)) & (e2 > F(thr_min)) & ((e1 + e2).astype(F) > F(thr_sum))
    with np.errstate(divide="ignore", invalid="ignore"):
        cr = F(1.0) - ((F(ELECTRON_REST_MEV) * e1) / ((F(e0) - e1) * F(e0)))
    keep &= (cr > F(-1.0 + 1e-6)) & (cr < F(1.0 - 1e-6))
    det = np.asarray(detector, F)
    D = det.shape[0]
    i1 = cp1.astype(np.int64) - 1
    i2 = cp2.astype(np.int64) - 1
    bad = keep & ((i1 < -D) | (i1 >= D) | (i2 < -D) | (i2 >= D))
    if bad.any():
        raise IndexError("crystal number out of range of the detector table")
    i1 = np.where(i1 < 0, i1 + D, i1)
    i2 = np.where(i2 < 0, i2 + D, i2)
    i1s = np.where(keep, i1, 0)
    i2s = np.where(keep, i2, 0)
    y1 = det[i1s, 1]
    y2 = det[i2s, 1]
    keep &= np.abs((y1 - y2).astype(F)) > F(0.1)
    idx = np.nonzero(keep)[0]
    return {"list": lst, "idx": idx, "e1": e1[idx], "e2": e2[idx], "arow": i1[idx], "crow": i2[idx],
            "pos1": det[i1[idx], :3], "pos2": det[i2[idx], :3]}


def cone_tc(pos1, pos2, s1sq, s2sq, theta, kn, sig_plus, sig_minus, coor, fixed_sigma=None):
    """t_c [N, P] in FP64 (process_list_plane_strict.py:201-228, 39-52, 99-134).

    fixed_sigma: if given, the angular sigma is that constant (rad) instead of
    the strict energy + position model.
    """
    p1 = np.asarray(pos1, np.float64)[:, None, :]
    p2 = np.asarray(pos2, np.float64)[:, None, :]
    c = np.asarray(coor, np.float64)[None, :, :]
    v01 = p1 - c
    v12 = p2 - p1
    d01 = np.linalg.norm(v01, axis=2)
    d12 = np.linalg.norm(v12, axis=2)
    bc = (v01 * v12).sum(2) / np.maximum(d01 * d12, 1e-7)
    beta = np.arccos(np.clip(bc, -COS_MAX, COS_MAX))
    th = np.asarray(theta, np.float64)[:, None]
    if fixed_sigma is None:
        u01 = v01 / np.maximum(d01, 1e-7)[..., None]
        u12 = v12 / np.maximum(d12, 1e-7)[..., None]
        cb = (u01 * u12).sum(2)
        sb = np.sqrt(np.maximum(1.0 - cb ** 2, 1e-12))
        jac = (u01 - cb[..., None] * u12) / np.maximum(d12, 1e-7)[..., None]
        g = jac / sb[..., None]
        w = (np.asarray(s1sq, np.float64) + np.asarray(s2sq, np.float64))[:, None, :]
        spos2 = (g ** 2 * w).sum(2)
        spos = np.sqrt(np.maximum(spos2, 1e-12))
        delta = beta - th
        sene = np.where(delta >= 0, np.asarray(sig_plus, np.float64)[:, None], np.asarray(sig_minus, np.float64)[:, None])
        sig = np.sqrt(np.maximum(spos ** 2 + sene ** 2, 1e-12))
    else:
        sig = float(fixed_sigma)
    return np.exp(-(beta - th) ** 2 / (2.0 * sig ** 2)) * (np.asarray(kn, np.float64)[:, None] - np.sin(beta) ** 2)


def beta_of(pos1, pos2, coor):
    """beta = acos(u01 . u12) without clamping (used by the finite-difference check)."""
    v01 = pos1 - coor
    v12 = pos2 - pos1
    return np.arccos(np.dot(v01, v12) / (np.linalg.norm(v01) * np.linalg.norm(v12)))


def sigma_pos_sq(pos1, pos2, coor, wsum):
    """Strict position variance of beta for one (event, voxel), FP64 (:99-134)."""
    v01 = pos1 - coor
    v12 = pos2 - pos1
    d01, d12 = np.linalg.norm(v01), np.linalg.norm(v12)
    u01, u12 = v01 / d01, v12 / d12
    cb = float(np.dot(u01, u12))
    sb = np.sqrt(max(1.0 - cb * cb, 1e-12))
    g = (u01 - cb * u12) / d12 / sb
    return float(np.sum(g * g * wsum))


def as_fp32_rows(t):
    """The rows as the reference holds them: FP32 values (tails below the FP32
    subnormal range are exactly zero), carried on in FP64."""
    with np.errstate(over="ignore", under="ignore"):
        return np.asarray(t, np.float64).astype(np.float32).astype(np.float64)


def row_filters(t):
    """NaN / zero-sum / effective-support decisions of :241-248, :137-148.

    The reference evaluates the rows in FP32, so a row whose every weight lies
    below the FP32 range sums to zero and is dropped; the decisions are taken on
    the FP32-rounded rows (sums in FP64).  Returns (keep [N] bool, S [N],
    support [N])."""
    t = as_fp32_rows(t)
    with np.errstate(invalid="ignore", divide="ignore"):
        S = t.sum(axis=1)
        nan = np.isnan(t).any(axis=1)
        # rows entirely in the FP32 subnormal range count as zero rows (the
        # reference's outcome there depends on the exponential's subnormal handling)
        zero = (S == 0) | (np.nan_to_num(t, nan=0.0).max(axis=1, initial=0.0) < np.finfo(np.float32).tiny)
        tn = t / S[:, None]
        support = 1.0 / np.sum(tn ** 2, axis=1)
    keep = (~nan) & (~zero) & (support >= MIN_EVENT_EFFECTIVE_SUPPORT)
    return keep, S, support


def backproj_list_single(sysmat, detector, coor, list_origin, delta_r1, delta_r2, e0, ene_resolution,
                         thr_max, thr_min, thr_sum, noise1=None, noise2=None, fixed_sigma=None):
    """Full strict chain (process_list_plane_strict.py:151-253) with FP64 rows.

    sysmat may be None (first-hit factor 1).  Returns a dict with the smeared
    list, kept input positions, the row-normalised t / t_compton / t_single
    and the effective supports (for borderline accounting)."""
    ev = prepare_events(list_origin, detector, e0, ene_resolution, thr_max, thr_min, thr_sum, noise1, noise2)
    s1 = detector_sigma_sq(detector, delta_r1)[ev["arow"]]
    s2 = detector_sigma_sq(detector, delta_r2)[ev["crow"]]
    th = theta_from_e1(ev["e1"], e0)
    kn = klein_nishina(ev["e1"], e0)
    sp, sm = angle_sigma_ene_pm(ev["e1"], e0, ene_resolution)
    tc = as_fp32_rows(cone_tc(ev["pos1"], ev["pos2"], s1, s2, th, kn, sp, sm, coor, fixed_sigma))
    a = np.ones_like(tc) if sysmat is None else np.asarray(sysmat, np.float64)[ev["arow"]]
    t = as_fp32_rows(tc * a)
    keep, S, support = row_filters(t)
    kk = np.nonzero(keep)[0]
    with np.errstate(invalid="ignore", divide="ignore"):
        tn = t[kk] / S[kk, None]
        tcn = tc[kk] / tc[kk].sum(1, keepdims=True)
        tsn = a[kk] / a[kk].sum(1, keepdims=True)
    return {"list": ev["list"], "sel": ev["idx"], "kept": ev["idx"][kk], "t": tn, "t_compton": tcn,
            "t_single": tsn, "support": support, "S": S, "stats_keep": keep, "theta": th}
