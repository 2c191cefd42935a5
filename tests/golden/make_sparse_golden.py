"""Writes tests/golden/sparse_projector_golden.npz (development container only).

Runs the reference's own build_compton_sparse_projector
(compton_sparse_ops.py:212-290) on the 511keV_RotateNum20 polar grid for a few
(theta_stride, z_stride, rotate_num) settings, including the branchy
ring-stride choice for theta_stride > 1 that only an import pins, and stores
the ring strides, shapes and SHA-256 digests of every FP32 operator and of the
coarse coordinates.  Usage: python tests/golden/make_sparse_golden.py
"""
import hashlib
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
SETTINGS = [(1, 2, None), (2, 2, 20), (3, 2, 20), (4, 4, 20), (2, 5, None), (3, 1, 7)]


def digest(t: torch.Tensor) -> str:
    return hashlib.sha256(np.ascontiguousarray(t.detach().cpu().numpy().astype(np.float32)).tobytes()).hexdigest()


def main():
    sys.path.insert(0, REF)
    import compton_sparse_ops as ref  # the reference's module (CPU run only)
    coor = torch.from_numpy(np.genfromtxt(os.path.join(REF, "Factors", "511keV_RotateNum20", "coor_polar_full.csv"),
                                          delimiter=",").astype(np.float32))
    out = {}
    for k, (ts, zs, rot) in enumerate(SETTINGS):
        p = ref.build_compton_sparse_projector(coor, theta_stride=ts, z_stride=zs, rotate_num=rot)
        out[f"s{k}_ring_strides"] = np.array(p.ring_strides)
        out[f"s{k}_shape"] = np.array([p.fine_layer_num, p.fine_z_num, p.coarse_layer_num, p.coarse_z_num])
        for name in ("layer_reduce", "z_reduce", "layer_expand", "z_expand", "coor_coarse"):
            out[f"s{k}_{name}"] = np.array(digest(getattr(p, name)))
        out[f"s{k}_z_expand_full"] = p.z_expand.numpy()
        out[f"s{k}_z_reduce_full"] = p.z_reduce.numpy()
    out["settings"] = np.array([(a, b, -1 if c is None else c) for a, b, c in SETTINGS])
    np.savez_compressed(os.path.join(HERE, "sparse_projector_golden.npz"), **out)
    print("wrote", len(SETTINGS), "projector settings")


if __name__ == "__main__":
    main()
