"""The cone pass that also forms the list-mode sensitivity (jscc_cone_rows_sens, build_compton_rows(want_sens=True)):
the rows are the bits jscc_cone_rows stores, the statistics equal up to the order of the per-row sums, the kept
rows are the same, and the sensitivity -- torch.sum(t, dim=0) of the normalised kept T that main_plane.py:348-355
scales into Sensi_d -- matches the FP64 column sums of the stored rows, the projection kernel's column-sum pass and
the reference-written golden T."""
import os

import numpy as np
import pytest
import torch

from oracle import cone
from tests.conftest import GOLDEN
from tests.golden.golden_inputs import golden_sysmat

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FIX = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")
STRICT = os.path.join(GOLDEN, "strict_cone_small.npz")


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_against_plain_pass(ra, rb, P, sens_bar=2e-6):
    """rb (want_sens) against ra (jscc_cone_rows) built from the same list."""
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    ra.resolve_counts()
    rb.resolve_counts()
    assert ra.n_sel == rb.n_sel and ra.n_kept == rb.n_kept
    n = ra.n_sel
    assert torch.equal(ra.store[:n, :P], rb.store[:n, :P])          # the same row bits
    sa, sb = ra.stats[:n].cpu().numpy(), rb.stats[:n].cpu().numpy()
    assert np.array_equal(sa[:, 3], sb[:, 3])                      # max: exact
    # S, Q, Sc: per-thread FP32 partial sums over another thread layout, combined in FP64
    for k, rtol in ((0, 1e-6), (1, 1e-6), (2, 1e-6)):
        np.testing.assert_allclose(sb[:, k], sa[:, k], rtol=rtol, atol=0)
    assert torch.equal(ra.kept[:ra.n_kept], rb.kept[:rb.n_kept])
    kept = rb.kept[:rb.n_kept].long()
    rows64 = rb.store[kept, :P].double()
    inv = rb.inv_sum[kept].double()
    ref = (rows64 * inv[:, None]).sum(0).cpu().numpy()            # FP64 column sums of the stored rows
    got = rb.sens.cpu().numpy()
    assert got.shape == (P,)
    assert _rel(got, ref) < sens_bar
    name = "bf16" if rb.store.dtype == torch.bfloat16 else "fp32"
    s_pass = ListModeStore.from_rows(rb, name).backproject_ones().cpu().numpy()
    assert _rel(got, s_pass) < sens_bar                            # the column-sum pass over T
    return got


@pytest.mark.parametrize("store", [torch.float32, torch.bfloat16])
def test_strict_sensitivity_matches_reference_golden_t(store):
    """Strict sigma model with the first-hit factor on the reference grid layout (1280 voxels): the fused
    sensitivity against the column sums of the reference-written normalised T (98 kept rows)."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    built = []
    for want in (False, True):
        built.append(build_compton_rows(torch.from_numpy(A), torch.from_numpy(det), torch.from_numpy(coor),
                                        torch.from_numpy(g["list_in"].copy()), dr1, dr2, e0, res, tmax, tmin, tsum,
                                        DEV, noise=(torch.from_numpy(g["noise1"]).to(DEV),
                                                    torch.from_numpy(g["noise2"]).to(DEV)),
                                        store_dtype=store, want_sens=want))
    got = _check_against_plain_pass(built[0], built[1], coor.shape[0])
    golden = g["t"].astype(np.float64).sum(0)
    assert _rel(got, golden) < (1e-5 if store == torch.float32 else 2e-3)


@pytest.mark.parametrize("rings,angles,store,sync", [
    (16, 64, torch.float32, True),
    (16, 64, torch.bfloat16, False),
    (64, 256, torch.bfloat16, False),    # the 16384-voxel workload grid, device-side counts
    (64, 256, torch.float32, True),
    (100, 256, torch.bfloat16, False),   # 25600 voxels: 16 voxel quads per thread
    (100, 256, torch.float32, True),
])
def test_fixed_sigma_sensitivity_matches_plain_pass_and_oracle(rings, angles, store, sync):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = synth.compton_list(n_events=1500, n_rings=rings, n_angles=angles, seed=rings + angles)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    sig = float(np.deg2rad(3.0))
    P = cl.coor.shape[0]
    print(f"P {P} store {store} sync {sync}", flush=True)
    built = []
    for want in (False, True):
        built.append(build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                        torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                        tsum, DEV, noise="none", sigma_model="fixed", fixed_sigma=sig,
                                        store_dtype=store, sync_counts=sync, want_sens=want))
    rb = built[1]
    if not sync:  # rows of the grid past the device count: NaN statistics, as jscc_cone_rows_counted
        rb.resolve_counts()
        assert rb.n_sel < cl.list_origin.shape[0]
        assert torch.isnan(rb.stats[rb.n_sel:]).all()
    got = _check_against_plain_pass(built[0], rb, P)
    if store == torch.float32 and P <= 16384:
        ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.02, tmax,
                                        tmin, tsum, None, None, fixed_sigma=sig)
        kept = rb.kept_event_index().cpu().numpy()
        border = set(np.nonzero(np.abs(ref["support"] - 50.0) / 50.0 < 1e-4)[0].tolist())
        assert set(kept.tolist()) - border == set(ref["kept"].tolist()) - border
        if np.array_equal(kept, ref["kept"]):
            assert _rel(got, np.asarray(ref["t"], np.float64).sum(0)) < 1e-5


def test_sensitivity_of_no_rows_is_zero_and_bad_requests_raise():
    from jscc_recon_spect_cylinder_amd import _lib, ops, synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = synth.compton_list(n_events=50, n_rings=8, n_angles=32, seed=3)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    sig = float(np.deg2rad(3.0))
    rows = build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model="fixed", fixed_sigma=sig, store_dtype=torch.float32,
                              want_sens=True)
    _, _, _, sens, _ = ops.cone_rows_sens(rows.events, rows.sel_idx, 0, rows.vox, 50.0, sigma_model=ops.SIG_FIXED,
                                          fixed_sigma=sig, out_dtype=torch.float32)
    assert sens.shape == (256,) and not torch.any(sens != 0)
    for kw in ({"want_tc": True}, {"store_dtype": None}, {"store_dtype": torch.float16, "want_sens": False}):
        with pytest.raises(ValueError):
            build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                               torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                               noise="none", sigma_model="fixed", fixed_sigma=sig, **{"want_sens": True, **kw})
    big = ops.voxel_soa(torch.from_numpy(np.random.default_rng(0).normal(size=(32772, 3)).astype(np.float32) * 40.0),
                        DEV)
    with pytest.raises(_lib.JsccError, match="voxels"):
        ops.cone_rows_sens(rows.events, rows.sel_idx, 4, big, 50.0, sigma_model=ops.SIG_FIXED, fixed_sigma=sig)


@pytest.mark.parametrize("rings,angles,sync", [(16, 64, True), (64, 256, False), (100, 256, False)])
def test_fp16_rows_of_the_cone_pass_equal_quantised_fp32_rows(rings, angles, sync):
    """FP16 rows written by the cone pass (the FP32 values quantised right after their group's statistics, when
    the row maximum and sum are known): codes bit-identical to quantize_rows of the FP32 rows for every kept row, the same row
    scales, the sensitivity equal to the column-sum pass over the FP16 store, and the OSEM image of the FP16 store
    equal to the one of the quantised FP32 rows (1e-4 voxels, device-side counts included)."""
    from jscc_recon_spect_cylinder_amd import ops, synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = synth.compton_list(n_events=1500, n_rings=rings, n_angles=angles, seed=3 * rings + angles)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    sig = float(np.deg2rad(3.0))
    P = cl.coor.shape[0]
    built = {}
    for sd in (torch.float32, torch.float16):
        built[sd] = build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                       torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                       tsum, DEV, noise="none", sigma_model="fixed", fixed_sigma=sig, store_dtype=sd,
                                       sync_counts=sync, want_sens=True)
    r32, r16 = built[torch.float32].resolve_counts(), built[torch.float16].resolve_counts()
    assert r16.store.dtype == torch.float16 and r16.row_scale is not None
    assert r32.n_sel == r16.n_sel and torch.equal(r32.kept[:r32.n_kept], r16.kept[:r16.n_kept])
    n = r16.n_sel
    codes, scale = ops.quantize_rows(r32.store, P, n, r32.stats, "fp16")
    kept = r16.kept[:r16.n_kept].long()
    assert torch.equal(r16.store[kept, :P].view(torch.int16), codes[kept, :P].view(torch.int16))
    # the same scales for every selected row (FP32 statistics of both passes agree to the rounding of their sums)
    np.testing.assert_allclose(r16.row_scale[:n].cpu().numpy(), scale[:n].cpu().numpy(), rtol=1e-6)
    st16 = ListModeStore.from_rows(r16, "fp16")
    s_pass = st16.backproject_ones()
    assert _rel(r16.sens.cpu().numpy(), s_pass.cpu().numpy()) < 2e-6
    st_q = ListModeStore.from_rows(r32, "fp16")
    x16 = st16.osem(torch.ones(P, device=DEV), r16.sens, 3, 4)
    xq = st_q.osem(torch.ones(P, device=DEV), st_q.backproject_ones(), 3, 4)
    assert _rel(x16.cpu().numpy(), xq.cpu().numpy()) < 1e-4


def test_fp16_strict_rows_with_first_hit_factor_equal_quantised_fp32_rows():
    """The strict sigma model with the first-hit factor on the reference grid layout (the reference-written
    golden list, 1280 voxels): FP16 rows written by the cone pass have the codes and scales of quantize_rows on
    the FP32 rows, the sensitivity of the column-sum pass over them, and column sums within FP16 rounding
    (2^-11 per element) of the reference-written normalised T."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    built = {}
    for sd in (torch.float32, torch.float16):
        built[sd] = build_compton_rows(torch.from_numpy(A), torch.from_numpy(det), torch.from_numpy(coor),
                                       torch.from_numpy(g["list_in"].copy()), dr1, dr2, e0, res, tmax, tmin, tsum,
                                       DEV, noise=(torch.from_numpy(g["noise1"]).to(DEV),
                                                   torch.from_numpy(g["noise2"]).to(DEV)),
                                       store_dtype=sd, want_sens=True)
    r32, r16 = built[torch.float32].resolve_counts(), built[torch.float16].resolve_counts()
    P = coor.shape[0]
    assert r32.n_kept == r16.n_kept == g["t"].shape[0]
    assert torch.equal(r32.kept[:r32.n_kept], r16.kept[:r16.n_kept])
    n = r16.n_sel
    codes, scale = ops.quantize_rows(r32.store, P, n, r32.stats, "fp16")
    kept = r16.kept[:r16.n_kept].long()
    assert torch.equal(r16.store[kept, :P].view(torch.int16), codes[kept, :P].view(torch.int16))
    np.testing.assert_allclose(r16.row_scale[:n].cpu().numpy(), scale[:n].cpu().numpy(), rtol=1e-6)
    s_pass = ListModeStore.from_rows(r16, "fp16").backproject_ones().cpu().numpy()
    assert _rel(r16.sens.cpu().numpy(), s_pass) < 2e-6
    assert _rel(r16.sens.cpu().numpy(), g["t"].astype(np.float64).sum(0)) < 5e-4


@pytest.mark.parametrize("store", [torch.float32, torch.bfloat16, torch.float16])
def test_strict_first_hit_pass_at_the_reference_grid_size(store):
    """The strict sigma model with a first-hit factor at 25600 voxels (the size of the reference's
    511keV_RotateNum20 grid; 1024-thread work-groups): FP32 / BF16 rows and statistics as the plain pass,
    the sensitivity as its column sums; FP16 codes and scales as quantize_rows of the FP32 rows."""
    from jscc_recon_spect_cylinder_amd import ops, synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    print(f"strict first-hit 25600 {store}", flush=True)
    rng = np.random.default_rng(100)
    P = 100 * 256
    A = torch.from_numpy((rng.random((800, P)) + 0.05).astype(np.float32))  # 2 crystals per event
    cl = synth.compton_list(n_events=400, n_rings=100, n_angles=256, seed=140)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)

    def build(sd, want):
        return build_compton_rows(A, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                  torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin, tsum,
                                  DEV, noise="none", sigma_model="strict", store_dtype=sd, want_sens=want)
    if store != torch.float16:
        got = _check_against_plain_pass(build(store, False), build(store, True), P)
        assert np.all(np.isfinite(got))
        return
    r32, r16 = build(torch.float32, True).resolve_counts(), build(torch.float16, True).resolve_counts()
    assert r16.n_kept > 20 and torch.equal(r32.kept[:r32.n_kept], r16.kept[:r16.n_kept])
    n = r16.n_sel
    codes, scale = ops.quantize_rows(r32.store, P, n, r32.stats, "fp16")
    kept = r16.kept[:r16.n_kept].long()
    assert torch.equal(r16.store[kept, :P].view(torch.int16), codes[kept, :P].view(torch.int16))
    np.testing.assert_allclose(r16.row_scale[:n].cpu().numpy(), scale[:n].cpu().numpy(), rtol=1e-6)
    s_pass = ListModeStore.from_rows(r16, "fp16").backproject_ones().cpu().numpy()
    assert _rel(r16.sens.cpu().numpy(), s_pass) < 2e-6
