"""On-disk formats (CPU): the readers on files the reference itself holds (Factors/511keV_RotateNum20:
Detector.csv, coor_polar.csv, RotMat.csv, RotMatInv.csv, Sensi_d, copied as fixtures), the reference's file
layouts round-trip, and the T-shard subset plan."""
import json
import os

import numpy as np
import torch

from jscc_recon_spect_cylinder_amd import formats
from tests.conftest import GOLDEN

REF_FILES = os.path.join(GOLDEN, "reference_files", "511keV_RotateNum20")
FACTORS = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")


def test_readers_on_the_reference_factor_files():
    """Detector.csv (main_plane.py:174-196 reads columns 1..3), coor_polar.csv, RotMat.csv / RotMatInv.csv
    (1-based, one column per rotation) and Sensi_d (raw FP32, one value per pixel) as the reference stores them;
    the expected arrays come from the factors fixture (written from the same files with pandas / numpy)."""
    g = np.load(FACTORS)
    det = formats.read_detector_csv(os.path.join(REF_FILES, "Detector.csv"))
    assert det.dtype == torch.float32 and det.shape == (10496, 3)
    np.testing.assert_array_equal(det.numpy(), g["detector"])
    coor = formats.read_coor_polar_csv(os.path.join(REF_FILES, "coor_polar.csv"))
    assert coor.shape == (1280, 2)
    np.testing.assert_array_equal(coor.numpy(), g["coor_layer0"][:, :2].astype(np.float32))
    rot = formats.read_rotmat_csv(os.path.join(REF_FILES, "RotMat.csv"))
    inv = formats.read_rotmat_csv(os.path.join(REF_FILES, "RotMatInv.csv"))
    assert rot.dtype == torch.int64 and rot.shape == (1280, 20)
    np.testing.assert_array_equal(rot.numpy(), g["rot_layer0"].astype(np.int64))
    # RotMatInv undoes RotMat rotation by rotation: rot[inv[p, r] - 1, r] == p + 1
    r = rot.numpy()
    v = inv.numpy()
    for k in range(20):
        np.testing.assert_array_equal(r[v[:, k] - 1, k], np.arange(1, 1281))
    s = formats.read_sensi(os.path.join(REF_FILES, "Sensi_d"), 25600)
    assert s.shape == (25600, 1) and s.dtype == torch.float32
    raw = np.fromfile(os.path.join(REF_FILES, "Sensi_d"), dtype="<f4")
    np.testing.assert_array_equal(s.numpy()[:, 0], raw)
    assert np.isfinite(raw).all() and (raw > 0).all()
    s2 = formats.read_sensi(os.path.join(REF_FILES, "Sensi_d"), 25600, intensity=2.0)
    np.testing.assert_array_equal(s2.numpy(), 2 * s.numpy())


def test_sysmat_polar_is_pixel_major_and_transposed_on_load(tmp_path):
    A = np.arange(6 * 10, dtype=np.float32).reshape(6, 10)  # [bins, pixels]
    p = str(tmp_path / "SysMat_polar")
    formats.write_sysmat_polar(p, torch.from_numpy(A))
    raw = np.fromfile(p, dtype=np.float32)
    np.testing.assert_array_equal(raw[:6], A[:, 0])  # bins fastest
    B = formats.read_sysmat_polar(p, 10, intensity=2.0)
    assert B.shape == (6, 10)
    np.testing.assert_array_equal(B.numpy(), 2 * A)


def test_csv_readers_follow_the_reference_layouts(tmp_path):
    det = tmp_path / "Detector.csv"
    det.write_text("1,10.5,-200,3\n2,0.25,230.125,-4.5\n")
    d = formats.read_detector_csv(str(det))
    np.testing.assert_array_equal(d.numpy(), np.array([[10.5, -200, 3], [0.25, 230.125, -4.5]], np.float32))
    cnt = tmp_path / "CntStat.csv"
    cnt.write_text("1,2,3,4\n5,6,7,8\n")  # rows = rotations
    c = formats.read_cntstat(str(cnt), 2)
    np.testing.assert_array_equal(c.numpy(), np.array([[1, 5], [2, 6], [3, 7], [4, 8]], np.float32))
    one = tmp_path / "1.csv"
    one.write_text("3,0.1234567,9,0.387654321,7\n")  # one event, extra column ignored
    l1 = formats.read_list_csv(str(one))
    assert l1.shape == (1, 4) and l1[0, 1] == np.float32(0.1234567)
    rot = tmp_path / "RotMat_full.csv"
    rot.write_text("1,2\n2,1\n")
    assert formats.read_rotmat_csv(str(rot)).dtype == torch.int64
    lst = torch.from_numpy(np.random.default_rng(0).random((5, 4)).astype(np.float32))
    formats.write_list_csv(str(tmp_path / "2.csv"), lst)
    assert torch.equal(formats.read_list_csv(str(tmp_path / "2.csv")), lst)  # %.9g round-trips FP32


def test_sensi_round_trip(tmp_path):
    s = torch.from_numpy(np.random.default_rng(1).random((25, 1)).astype(np.float32))
    formats.write_sensi(str(tmp_path / "Sensi_d"), s)
    assert os.path.getsize(tmp_path / "Sensi_d") == 100
    assert torch.equal(formats.read_sensi(str(tmp_path / "Sensi_d"), 25), s)


def test_t_shards_round_trip_and_subset_plan(tmp_path):
    rng = np.random.default_rng(2)
    chunks = [(r, torch.from_numpy(rng.random((n, 8)).astype(np.float32)))
              for r, n in [(0, 5), (0, 7), (0, 0), (0, 3), (1, 4), (1, 4), (1, 4), (1, 4)]]
    d = str(tmp_path / "511keV")
    man = formats.write_t_shards(d, chunks, 8, 2)
    assert man["rows_per_rotate"] == [15, 16] and man["records"][2]["rows"] == 0
    m2 = formats.load_t_manifest(os.path.join(d, "manifest.json"))
    assert m2 == json.loads(json.dumps(man))
    plan = formats.build_subset_plan_for_manifest(m2, 2, d)
    # every non-empty chunk exactly once, in chunk order, each subset getting rows of both rotations
    for r in range(2):
        got = [rec["chunk_index"] for sub in plan for rec in sub[r]]
        want = [i for i, (rr, t) in enumerate(chunks) if rr == r and t.shape[0] > 0]
        assert got == want
        assert all(len(sub[r]) > 0 for sub in plan)
    assert [rec["chunk_index"] for rec in plan[0][1]] == [4, 5]  # 8 >= 16/2 after two chunks
    t_all = formats.load_t_list_from_plan(plan)
    assert torch.equal(t_all[0][0][1][0], chunks[1][1])
