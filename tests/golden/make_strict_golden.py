"""Writes tests/golden/strict_cone_small.npz and tests/golden/factors_511keV_RotateNum20.npz.

Run in the development container only (it reads /root/reference, which the GPU
machines do not have):  python tests/golden/make_strict_golden.py

* strict_cone_small.npz: inputs and outputs of the reference's own
  get_compton_backproj_list_single (process_list_plane_strict.py:151-253), run
  on the CPU: the FP32 event filters and the smeared energies can only be pinned
  this way.  The system matrix is not stored: it is the closed-form
  `golden_sysmat` of golden_inputs.py, recomputed by the tests.
* factors_511keV_RotateNum20.npz: the reference's detector table and checksums
  of its polar grid / RotMat / RotMatInv files (data only).
"""
import hashlib
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from golden_inputs import golden_sysmat  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
FACT = os.path.join(REF, "Factors", "511keV_RotateNum20")


def grid_checksum(a: np.ndarray, decimals: int) -> str:
    return hashlib.sha256(np.ascontiguousarray(np.round(a, decimals)).tobytes()).hexdigest()


def main():
    sys.path.insert(0, REF)
    import process_list_plane_strict as ref  # the reference's module (CPU run only)

    det = np.genfromtxt(os.path.join(FACT, "Detector.csv"), delimiter=",", dtype=np.float32)[:, 1:4]
    coor_full = np.genfromtxt(os.path.join(FACT, "coor_polar_full.csv"), delimiter=",")
    rot = np.genfromtxt(os.path.join(FACT, "RotMat_full.csv"), delimiter=",", dtype=np.int64)
    inv = np.genfromtxt(os.path.join(FACT, "RotMatInv_full.csv"), delimiter=",", dtype=np.int64)
    np.savez_compressed(
        os.path.join(HERE, "factors_511keV_RotateNum20.npz"),
        detector=det,
        coor_sha=np.array(grid_checksum(coor_full, 9)),
        rot_sha=np.array(hashlib.sha256(np.ascontiguousarray(rot).tobytes()).hexdigest()),
        inv_sha=np.array(hashlib.sha256(np.ascontiguousarray(inv).tobytes()).hexdigest()),
        coor_layer0=coor_full[:1280],
        rot_layer0=rot[:1280].astype(np.int16),
    )

    # ---- strict cone golden vectors --------------------------------------------------------
    rng = np.random.default_rng(20260331)
    coor = coor_full[10 * 1280:11 * 1280].astype(np.float32)  # the z = +1.5 mm layer
    n = 160
    ay = np.abs(det[:, 1])
    front = np.nonzero(ay < 289.0)[0]
    cp1 = rng.choice(front, n) + 1
    cp2 = rng.choice(det.shape[0], n) + 1
    cp2[:6] = cp1[:6] + 1                                     # same-layer neighbours (filtered)
    e0 = 0.511
    e1 = rng.uniform(0.04, 0.36, n).astype(np.float32)
    e1[6:12] = np.float32(2 * e0 ** 2 / (0.511 + 2 * e0) - 0.001)  # at the energy window edge
    e1[12:16] = np.float32(0.05)                                    # at the lower threshold
    e2 = (e0 - e1 + rng.normal(0, 0.004, n)).astype(np.float32)
    lst = np.stack([cp1, e1, cp2, e2], axis=1).astype(np.float32)
    sysmat = golden_sysmat(det, coor)

    torch.manual_seed(1234)
    noise1 = torch.randn(n)
    noise2 = torch.randn(n)
    torch.manual_seed(1234)
    ene_resolution = 0.1 * (0.662 / e0) ** 0.5
    thr_max = 2 * e0 ** 2 / (0.511 + 2 * e0) - 0.001
    lst_t = torch.from_numpy(lst.copy())
    t, tc, ts = ref.get_compton_backproj_list_single(
        torch.from_numpy(sysmat), torch.from_numpy(det), torch.from_numpy(coor), lst_t, 2.0, 2.0, e0,
        ene_resolution, thr_max, 0.05, 0.46, torch.device("cpu"))
    np.savez_compressed(
        os.path.join(HERE, "strict_cone_small.npz"),
        list_in=lst, noise1=noise1.numpy(), noise2=noise2.numpy(), coor=coor,
        list_out=lst_t.numpy(), t=t.numpy(), t_compton=tc.numpy(), t_single_sum=ts.numpy().sum(1),
        params=np.array([2.0, 2.0, e0, ene_resolution, thr_max, 0.05, 0.46]),
    )
    print("kept", t.shape[0], "of", n)


if __name__ == "__main__":
    main()
