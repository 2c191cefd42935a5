"""On-disk formats of the reference's factor, data, image and T-shard files.

Reference: main_plane.py:161-196 (SysMat_polar: raw float32, pixel-major with
the bins fastest, transposed on load; Detector.csv columns 1..3;
coor_polar*.csv; RotMat*_full.csv (1-based); Sensi_s / Sensi_d raw float32 [P];
CntStat_<name>.csv rows = rotations; List_<name>/<r>.csv columns
(cpnum1, e1, cpnum2, e2)), recon_osem_plane.py:128-147 (Image_* raw float32),
sparse_main_utils.py:123-128 (single-row list files) and
distributed/python/t_shard_dist.py:32-38, 118-175 (T shards: manifest.json +
rotate_NNN/chunk_NNNNNN.bin raw float32 [rows, pixels], and the subset plan
that groups chunks into OSEM subsets).

CSV files are parsed with pandas' round-trip float parser, which gives the
same values as numpy.genfromtxt (the correctly rounded decimal conversion).
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch


def _read_csv(path: str, dtype) -> np.ndarray:
    import pandas as pd
    df = pd.read_csv(path, header=None, float_precision="round_trip")
    return df.to_numpy(dtype=np.float64 if dtype is not int else np.int64).astype(dtype)


# ----------------------------------------------------------------------------- factors
def read_sysmat_polar(path: str, pixel_num: int, intensity: float = 1.0) -> torch.Tensor:
    """[bins, pixels] FP32: the file is [pixels, bins] (bins fastest); returned as the reference's transposed view."""
    a = np.fromfile(path, dtype=np.float32).reshape(pixel_num, -1)
    t = torch.from_numpy(a).transpose(0, 1)
    return t * intensity if intensity != 1.0 else t


def write_sysmat_polar(path: str, sysmat: torch.Tensor) -> None:
    """Inverse of read_sysmat_polar: writes A [bins, pixels] pixel-major, bins fastest."""
    np.ascontiguousarray(sysmat.detach().cpu().numpy().astype(np.float32).T).tofile(path)


def read_detector_csv(path: str) -> torch.Tensor:
    """Crystal centres [D, 3] FP32 (columns 1..3 of Detector.csv)."""
    return torch.from_numpy(np.ascontiguousarray(_read_csv(path, np.float32)[:, 1:4]))


def read_coor_polar_csv(path: str) -> torch.Tensor:
    return torch.from_numpy(_read_csv(path, np.float32))


def read_rotmat_csv(path: str) -> torch.Tensor:
    """RotMat / RotMatInv (1-based pixel indices) as int64."""
    return torch.from_numpy(_read_csv(path, int))


def read_sensi(path: str, pixel_num: int, intensity: float = 1.0) -> torch.Tensor:
    """Sensi_s / Sensi_d: raw float32 [P] -> [P, 1]."""
    t = torch.from_numpy(np.fromfile(path, dtype=np.float32).reshape(pixel_num, 1))
    return t * intensity if intensity != 1.0 else t


def write_sensi(path: str, s: torch.Tensor) -> None:
    s.detach().cpu().numpy().astype("float32").tofile(path)


# ----------------------------------------------------------------------------- data
def read_cntstat(path: str, rotate_num: int) -> torch.Tensor:
    """CntStat_<name>.csv (one row per rotation) -> [bins, rotate_num] FP32."""
    a = _read_csv(path, np.float32)
    return torch.from_numpy(np.reshape(a, [rotate_num, -1])).transpose(0, 1)


def read_list_csv(path: str) -> torch.Tensor:
    """List_<name>/<r>.csv -> [N, 4] FP32 (cpnum1, e1, cpnum2, e2); a single event gives [1, 4]."""
    a = _read_csv(path, np.float32)
    if a.ndim == 1:
        a = a[None, :]
    return torch.from_numpy(np.ascontiguousarray(a[:, 0:4]))


def read_lists(list_dir: str, rotate_num: int) -> list:
    """The per-rotation list files <list_dir>/1.csv .. <rotate_num>.csv."""
    return [read_list_csv(os.path.join(list_dir, f"{i + 1}.csv")) for i in range(rotate_num)]


def write_list_csv(path: str, lst: torch.Tensor) -> None:
    np.savetxt(path, lst.detach().cpu().numpy().astype(np.float32), delimiter=",", fmt="%.9g")


# ----------------------------------------------------------------------------- T shards
def write_t_shards(energy_dir: str, t_chunks, pixel_num: int, rotate_num: int, meta: dict | None = None) -> dict:
    """Write row-normalised T chunks [(rotate_idx, tensor [rows, pixels]), ...] as
    rotate_NNN/chunk_NNNNNN.bin raw float32 plus manifest.json (t_shard_dist.py:118-160)."""
    os.makedirs(energy_dir, exist_ok=True)
    records, rows_per_rotate = [], [0] * rotate_num
    for chunk_index, (r, t) in enumerate(t_chunks):
        rows = int(t.shape[0])
        rel = os.path.join(f"rotate_{r + 1:03d}", f"chunk_{chunk_index:06d}.bin")
        if rows > 0:
            os.makedirs(os.path.join(energy_dir, f"rotate_{r + 1:03d}"), exist_ok=True)
            t.detach().cpu().numpy().astype(np.float32).tofile(os.path.join(energy_dir, rel))
        records.append({"rotate": r, "chunk_index": chunk_index, "rows": rows, "cols": pixel_num,
                        "path": rel.replace("\\", "/")})
        rows_per_rotate[r] += rows
    manifest = {"version": 1, "pixel_num": pixel_num, "rotate_num": rotate_num,
                "rows_per_rotate": rows_per_rotate, "total_rows": int(sum(rows_per_rotate)), "records": records}
    if meta:
        manifest.update(meta)
    with open(os.path.join(energy_dir, "manifest.json"), "w", encoding="utf-8") as f:
        json.dump(manifest, f, indent=2)
    return manifest


def load_t_manifest(manifest_path: str) -> dict:
    with open(manifest_path, "r", encoding="utf-8") as f:
        return json.load(f)


def build_subset_plan_for_manifest(manifest: dict, subset_num: int, manifest_dir: str) -> list:
    """plan[subset][rotate] = chunk records (t_shard_dist.py:166-201): per rotation, chunks in order are
    assigned greedily so each subset gets about 1/subset_num of the rows while every later subset keeps
    at least one chunk."""
    rotate_num = int(manifest["rotate_num"])
    plan = [[[] for _ in range(rotate_num)] for _ in range(subset_num)]
    by_rotate = [[] for _ in range(rotate_num)]
    for rec in manifest["records"]:
        if rec["rows"] > 0:
            by_rotate[rec["rotate"]].append(rec)
    for r in range(rotate_num):
        recs = sorted(by_rotate[r], key=lambda it: it["chunk_index"])
        total = sum(it["rows"] for it in recs)
        if total == 0:
            continue
        target = total / subset_num
        k, acc = 0, 0
        for idx, rec in enumerate(recs):
            c = dict(rec)
            c["path"] = os.path.join(manifest_dir, rec["path"])
            plan[k][r].append(c)
            acc += rec["rows"]
            if k < subset_num - 1 and acc >= target and len(recs) - idx - 1 >= subset_num - k - 1:
                k += 1
                acc = 0
    return plan


def read_t_chunk(record: dict, device=None) -> torch.Tensor:
    """One chunk [rows, cols] FP32 (on `device` when given)."""
    a = np.fromfile(record["path"], dtype=np.float32).reshape(record["rows"], record["cols"])
    t = torch.from_numpy(a)
    return t.to(device) if device is not None else t


def load_t_list_from_plan(plan: list, device=None) -> list:
    """t_list_all[subset][rotation][divide][energy] of recon_osem_plane.osem_list_mode /
    osem_joint_mode for a single-energy manifest: each chunk of the plan becomes one divide."""
    return [[[[read_t_chunk(rec, device)] for rec in rot] for rot in sub] for sub in plan]
