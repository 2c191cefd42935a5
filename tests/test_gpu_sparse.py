"""GPU parity of the coarse-grid path: row transforms, the sparse cone builder,
materialisation and the sparse OSEM weights, against the FP64 oracle."""
import os

import numpy as np
import pytest
import torch

from oracle import cone, layout, osem, sparse
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FIX = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def grid():
    coor, rot, inv = layout.reference_grid("511keV_RotateNum20")
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    p = build_compton_sparse_projector(torch.from_numpy(coor.astype(np.float32)), 2, 2, 20)
    pj = sparse.build(coor.astype(np.float32), 2, 2, 20)
    return coor.astype(np.float32), rot, inv, p, pj


def test_row_transforms_vs_oracle(grid):
    from jscc_recon_spect_cylinder_amd import compton_sparse_ops as cso
    coor, rot, inv, p, pj = grid
    rng = np.random.default_rng(0)
    fine = rng.random((5, p.fine_pixel_num)).astype(np.float32)
    coarse = rng.random((5, p.coarse_pixel_num)).astype(np.float32)
    pd = p.to(DEV)
    f = torch.from_numpy(fine).to(DEV)
    c = torch.from_numpy(coarse).to(DEV)
    assert rel(cso.reduce_fine_rows_to_coarse(f, pd).cpu(), sparse.reduce_rows(fine, pj)) < 1e-6
    assert rel(cso.upsample_coarse_rows_to_fine(c, pd).cpu(), sparse.upsample_rows(coarse, pj)) < 1e-6
    assert rel(cso.project_fine_rows_to_coarse_adjoint(f, pd).cpu(), sparse.adjoint_rows(fine, pj)) < 1e-6
    img = cso.reduce_fine_image_to_coarse(f[0].reshape(-1, 1), pd)
    assert rel(img.cpu().ravel(), sparse.reduce_rows(fine[:1], pj)[0]) < 1e-6
    up = cso.upsample_coarse_weight_to_fine(c[0].reshape(-1, 1), pd)
    ref = (pj["z_reduce"].T @ coarse[0].reshape(pj["z_reduce"].shape[0], -1) @ pj["layer_reduce"]).ravel()
    assert rel(up.cpu().ravel(), ref) < 1e-6


def test_materialize_vs_oracle(grid):
    from jscc_recon_spect_cylinder_amd import compton_sparse_ops as cso
    coor, rot, inv, p, pj = grid
    rng = np.random.default_rng(2)
    n = 40
    A = rng.random((30, p.fine_pixel_num)).astype(np.float32)
    ev = np.hstack([rng.integers(1, 31, (n, 1)).astype(np.float32),
                    (rng.random((n, p.coarse_pixel_num)) ** 3).astype(np.float32)])
    ev[3, 1:] = 0.0  # an all-zero row is invalid
    t, valid = cso.materialize_sparse_event_rows_to_fine(torch.from_numpy(ev).to(DEV), torch.from_numpy(A).to(DEV),
                                                         p.to(DEV))
    tr, vr = sparse.materialize(ev, A, pj)
    assert np.array_equal(valid.cpu().numpy(), vr)
    assert rel(t.cpu().numpy(), tr) < 1e-6


def test_sparse_cone_builder_vs_oracle(grid):
    """get_compton_backproj_list_single_sparse on the coarse grid of the reference's
    511keV_RotateNum20 layout with its detector table."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_sparse import build_sparse_rows
    coor, rot, inv, p, pj = grid
    det = np.load(FIX)["detector"]
    rng = np.random.default_rng(9)
    A = rng.random((det.shape[0], p.fine_pixel_num)).astype(np.float32) * 0.01 + 0.001
    n = 200
    ay = np.abs(det[:, 1])
    cp1 = rng.choice(np.nonzero(ay < 289)[0], n) + 1
    cp2 = rng.choice(det.shape[0], n) + 1
    e1 = rng.uniform(0.06, 0.33, n)
    lst = np.stack([cp1, e1, cp2, 0.511 - e1], axis=1).astype(np.float32)
    e0, res = 0.511, 0.1138
    tmax, tmin, tsum = 2 * e0 ** 2 / (0.511 + 2 * e0) - 0.001, 0.05, 0.46
    rows = build_sparse_rows(torch.from_numpy(A), torch.from_numpy(det), p, torch.from_numpy(lst), 2.0, 2.0, e0, res,
                             tmax, tmin, tsum, DEV, noise="none")
    # oracle: strict cone on the coarse voxels with the reduced A rows
    ev = cone.prepare_events(lst, det, e0, res, tmax, tmin, tsum)
    s1 = cone.detector_sigma_sq(det, 2.0)[ev["arow"]]
    s2 = cone.detector_sigma_sq(det, 2.0)[ev["crow"]]
    th = cone.theta_from_e1(ev["e1"], e0)
    sp, sm = cone.angle_sigma_ene_pm(ev["e1"], e0, res)
    # the coarse voxel centres are the projector's (FP32, as the reference computes them)
    tc = cone.cone_tc(ev["pos1"], ev["pos2"], s1, s2, th, cone.klein_nishina(ev["e1"], e0), sp, sm,
                      p.coor_coarse.numpy())
    a = sparse.reduce_rows(A[ev["arow"]], pj)
    keep, S, sup = cone.row_filters(tc * a)
    kept = rows.kept_event_index().cpu().numpy()
    border = np.abs(sup - 50.0) / 50.0 < 1e-4
    assert set(kept.tolist()) - set(ev["idx"][border].tolist()) == set(ev["idx"][keep & ~border].tolist())
    common = np.isin(ev["idx"][keep], kept)
    tcn = tc[keep][common] / tc[keep][common].sum(1, keepdims=True)
    from jscc_recon_spect_cylinder_amd import ops
    got = ops.gather_normalize(rows.tc_store, p.coarse_pixel_num, rows.kept, rows.n_kept, rows.stats, 2)
    got = got.cpu().numpy()[np.isin(kept, ev["idx"][keep])]
    assert rel(got, tcn) < 1e-4


def test_sparse_list_weight_equals_fine_row_weight(grid):
    """Grouped sparse weight == T_f^T (1/max(T_f x, 1e-12)) with unnormalised fine rows."""
    from jscc_recon_spect_cylinder_amd import recon_osem_plane_sparse as rps
    from jscc_recon_spect_cylinder_amd import recon_osem_local_sparse_jsccsd_only as loc
    coor, rot, inv, p, pj = grid
    rng = np.random.default_rng(3)
    A = rng.random((12, p.fine_pixel_num)).astype(np.float32)
    ev = np.hstack([rng.integers(1, 13, (30, 1)).astype(np.float32),
                    (rng.random((30, p.coarse_pixel_num)) ** 4).astype(np.float32)])
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    cp = np.round(ev[:, 0]).astype(int) - 1
    tf = sparse.upsample_rows(ev[:, 1:], pj) * A[cp]
    ref = osem.weight_compton(tf, x, eps=1e-12)
    pd = p.to(DEV)
    w = rps.get_weight_compton_sparse(torch.from_numpy(ev), torch.from_numpy(A).to(DEV),
                                      torch.from_numpy(x[:, None]).to(DEV), pd)
    assert rel(w.cpu().numpy().ravel(), ref) < 1e-5
    w2 = loc.get_weight_compton_sparse(torch.from_numpy(ev).to(DEV), torch.from_numpy(A).to(DEV),
                                       torch.from_numpy(x[:, None]).to(DEV), pd)
    tn, _ = sparse.materialize(ev, A, pj)
    assert rel(w2.cpu().numpy().ravel(), osem.weight_compton(tn, x, eps=1e-12)) < 1e-5


def test_sparse_builder_drop_in_packed_rows(grid):
    """get_compton_backproj_list_single_sparse returns (event_rows, None, None): packed
    [float(cpnum1), row-normalised t_c] rows of the kept events, on the CPU by default and
    on the device with keep_on_device=True (same device draws, same rows); the caller's
    list is left unchanged (process_list_plane_sparse.py:25-123)."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_sparse import get_compton_backproj_list_single_sparse
    coor, rot, inv, p, pj = grid
    det = np.load(FIX)["detector"]
    rng = np.random.default_rng(10)
    A = torch.from_numpy(rng.random((det.shape[0], p.fine_pixel_num)).astype(np.float32) * 0.01 + 0.001).to(DEV)
    n = 150
    ay = np.abs(det[:, 1])
    cp1 = rng.choice(np.nonzero(ay < 289)[0], n) + 1
    cp2 = rng.choice(det.shape[0], n) + 1
    e1 = rng.uniform(0.06, 0.33, n)
    lst = torch.from_numpy(np.stack([cp1, e1, cp2, 0.511 - e1], axis=1).astype(np.float32)).to(DEV)
    lst0 = lst.clone()
    e0, res = 0.511, 0.1138
    args = (A, torch.from_numpy(det).to(DEV), p.to(DEV), lst, 2.0, 2.0, e0, res, 2 * e0 ** 2 / (0.511 + 2 * e0) - 0.001,
            0.05, 0.46, torch.device(DEV))
    torch.cuda.manual_seed(7)
    er, a, b = get_compton_backproj_list_single_sparse(*args)
    torch.cuda.manual_seed(7)
    er2, _, _ = get_compton_backproj_list_single_sparse(*args, keep_on_device=True)
    assert a is None and b is None
    assert er.device.type == "cpu" and er2.device.type == "cuda"
    assert torch.equal(er2.cpu(), er)
    assert torch.equal(lst, lst0)
    assert er.shape[1] == p.coarse_pixel_num + 1 and 0 < er.shape[0] <= n
    sums = er[:, 1:].double().sum(1).numpy()
    np.testing.assert_allclose(sums, 1.0, rtol=1e-5)
    assert set(er[:, 0].numpy().astype(int).tolist()) <= set(cp1.tolist())
