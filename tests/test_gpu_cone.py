"""GPU parity of the cone back-projection (libjscc_hip through the drop-in API)
against the reference's golden vectors and the FP64 oracle."""
import os

import numpy as np
import pytest
import torch

from oracle import cone
from tests.conftest import GOLDEN
from tests.golden.golden_inputs import golden_sysmat

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")
STRICT = os.path.join(GOLDEN, "strict_cone_small.npz")
DEV = "cuda:0"


def _golden_sysmat(det, coor):
    return golden_sysmat(det, coor)


def _rel_rows(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return (np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max()


def test_strict_dropin_draws_cpu_noise_and_mutates_list():
    """The drop-in draws its smearing noise with two CPU torch.randn calls (as the
    reference) and mutates the caller's list; checked against the FP32 oracle with
    the same draws.  (The CPU normal stream itself differs across host ISAs at the
    ulp level, so it is re-drawn here rather than taken from the fixture.)"""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import get_compton_backproj_list_single
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = _golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    n = g["list_in"].shape[0]
    torch.manual_seed(1234)
    n1 = torch.randn(n).numpy()
    n2 = torch.randn(n).numpy()
    ref = cone.backproj_list_single(A, det, coor, g["list_in"], dr1, dr2, e0, res, tmax, tmin, tsum, n1, n2)
    lst = torch.from_numpy(g["list_in"].copy()).to(DEV)
    torch.manual_seed(1234)
    t, tc, ts = get_compton_backproj_list_single(torch.from_numpy(A).to(DEV), torch.from_numpy(det).to(DEV),
                                                 torch.from_numpy(coor).to(DEV), lst, dr1, dr2, e0, res, tmax,
                                                 tmin, tsum, torch.device(DEV))
    assert np.array_equal(lst.cpu().numpy(), ref["list"])  # in-place mutation, bit-identical
    assert t.device.type == "cpu" and t.dtype == torch.float32
    assert t.shape == ref["t"].shape
    assert _rel_rows(t.numpy(), ref["t"]) < 1e-4
    assert _rel_rows(tc.numpy(), ref["t_compton"]) < 1e-4
    np.testing.assert_allclose(ts.numpy(), ref["t_single"], rtol=1e-5, atol=1e-9)
    # keep_on_device: the same rows, left on the GPU (same draws again)
    lst2 = torch.from_numpy(g["list_in"].copy()).to(DEV)
    torch.manual_seed(1234)
    t2, tc2, ts2 = get_compton_backproj_list_single(torch.from_numpy(A).to(DEV), torch.from_numpy(det).to(DEV),
                                                    torch.from_numpy(coor).to(DEV), lst2, dr1, dr2, e0, res, tmax,
                                                    tmin, tsum, torch.device(DEV), keep_on_device=True)
    assert t2.device.type == "cuda" and tc2.device.type == "cuda" and ts2.device.type == "cuda"
    assert torch.equal(t2.cpu(), t) and torch.equal(tc2.cpu(), tc) and torch.equal(ts2.cpu(), ts)


def test_strict_kernels_match_reference_golden_vectors():
    """With the fixture's noise injected, the GPU path reproduces the reference's own
    output: smeared energies bit-exact, the same kept events, rows within 1e-4."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = _golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    lst = torch.from_numpy(g["list_in"].copy()).to(DEV)
    rows = build_compton_rows(torch.from_numpy(A), torch.from_numpy(det), torch.from_numpy(coor), lst, dr1, dr2, e0,
                              res, tmax, tmin, tsum, DEV,
                              noise=(torch.from_numpy(g["noise1"]), torch.from_numpy(g["noise2"])), want_tc=True)
    assert np.array_equal(lst.cpu().numpy(), g["list_out"])
    t = rows.normalized_t().cpu().numpy()
    assert t.shape == g["t"].shape
    assert _rel_rows(t, g["t"]) < 1e-4
    assert _rel_rows(rows.normalized_tc().cpu().numpy(), g["t_compton"]) < 1e-4


def test_non_finite_hit_positions_drop_their_events():
    """A crystal with a non-finite coordinate makes the reference's rows NaN, which its
    row filter drops: the events using it are not kept, the others are unchanged."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"].copy()
    coor = g["coor"]
    A = _golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    lst = g["list_in"]
    base = cone.backproj_list_single(A, det, coor, lst, dr1, dr2, e0, res, tmax, tmin, tsum, g["noise1"],
                                     g["noise2"])["kept"]
    hit1 = int(lst[base[0], 0]) - 1   # the first-hit crystal of the first kept event: NaN x
    hit2 = int(lst[base[1], 2]) - 1   # the second-hit crystal of the second kept event: infinite z
    det[hit1, 0] = np.nan
    det[hit2, 2] = np.inf
    with np.errstate(invalid="ignore", over="ignore"):
        ref = cone.backproj_list_single(A, det, coor, lst, dr1, dr2, e0, res, tmax, tmin, tsum, g["noise1"],
                                        g["noise2"])
    rows = build_compton_rows(torch.from_numpy(A), torch.from_numpy(det), torch.from_numpy(coor),
                              torch.from_numpy(lst.copy()), dr1, dr2, e0, res, tmax, tmin, tsum, DEV,
                              noise=(torch.from_numpy(g["noise1"]), torch.from_numpy(g["noise2"])), want_tc=True)
    kept = rows.kept_event_index().cpu().numpy()
    uses = (np.round(lst[:, 0]).astype(int) - 1 == hit1) | (np.round(lst[:, 2]).astype(int) - 1 == hit2)
    assert np.isin(np.nonzero(uses)[0], base).any() and not np.isin(np.nonzero(uses)[0], kept).any()
    assert np.array_equal(kept, ref["kept"]) and len(kept) < len(base)
    assert _rel_rows(rows.normalized_t().cpu().numpy(), ref["t"]) < 1e-4


def _axis_geometry(u_dir):
    """Hits p1 and p2 = p1 + 30 u and voxels placed at known angles beta from the axis:
    60 angles from 1e-5 to 0.6 rad (dense around sin^2 beta = 1/16) next to beta = 0 and
    beta = pi, at two distances, plus 96 voxels on each of the cone surfaces beta = 0.522
    and beta = 2.537 (the scattering angles of the two test events)."""
    p1 = np.array([90.0, 10.0, 3.0])
    u = np.asarray(u_dir, np.float64)
    u /= np.linalg.norm(u)
    p2 = p1 + 30.0 * u
    h = np.array([0.0, 0.0, 1.0]) if abs(u[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    w1 = np.cross(u, h)
    w1 /= np.linalg.norm(w1)
    w2 = np.cross(u, w1)
    betas = np.unique(np.concatenate([np.logspace(-5, np.log10(0.6), 48), np.linspace(0.24, 0.27, 12)]))
    pts = []
    for s in (20.0, 80.0):
        for b in betas:
            pts.append(p1 - s * (np.cos(b) * u + np.sin(b) * w1))  # angle b between p1 - c and p2 - p1
            pts.append(p1 + s * (np.cos(b) * u - np.sin(b) * w2))  # angle pi - b
    for th in (0.522, 2.537):  # the two events' cone surfaces
        for ph in np.linspace(0.0, 2.0 * np.pi, 96, endpoint=False):
            pts.append(p1 - 50.0 * (np.cos(th) * u + np.sin(th) * (np.cos(ph) * w1 + np.sin(ph) * w2)))
    det = np.stack([p1, p2]).astype(np.float32)
    return det, np.asarray(pts, np.float32), betas


@pytest.mark.parametrize("u_dir", [(-0.6, -0.75, 0.1), (0.2, 0.9, -0.4)])
def test_cone_rows_near_the_cone_axis(u_dir):
    """Voxels within 1e-5..0.6 rad of the cone axis (both ends) against the FP64 oracle,
    element by element.  The kernel forms |v01 x v12|^2 as d01^2 d12^2 - (v01 . v12)^2
    away from the axis and from the cross product near it; the difference form alone
    cancels near the axis (beta off by up to 1.3e-4 rad here, t by 0.7 %)."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    det, coor, betas = _axis_geometry(u_dir)
    assert betas.min() < 1e-4 and (np.abs(np.sin(betas) ** 2 - 0.0625) < 2e-3).sum() >= 2
    e0, tmax, tmin, tsum = 0.511, 0.3397, 0.05, 0.46
    # theta = 0.522 rad (voxels near beta = 0 carry weight) and 2.537 rad (near beta = pi)
    lst = np.array([[1.0, 0.06, 2.0, 0.451], [1.0, 0.33, 2.0, 0.181]], np.float32)
    sig = 0.1  # steep enough that an angle error near the axis (delta ~ 0.52) shows: d t / t = delta / sig^2 * d beta
    ref = cone.backproj_list_single(None, det, coor, lst, 0.0, 0.0, e0, 0.02, tmax, tmin, tsum, None, None,
                                    fixed_sigma=sig)
    assert ref["kept"].tolist() == [0, 1]
    rows = build_compton_rows(None, torch.from_numpy(det), torch.from_numpy(coor), torch.from_numpy(lst.copy()),
                              0.0, 0.0, e0, 0.02, tmax, tmin, tsum, DEV, noise="none", sigma_model="fixed",
                              fixed_sigma=sig, want_tc=True)
    assert rows.kept_event_index().cpu().tolist() == [0, 1]
    t = rows.normalized_tc().cpu().numpy().astype(np.float64)
    r = np.asarray(ref["t_compton"], np.float64)
    for i in range(2):
        m = r[i] > 1e-30
        assert m.sum() > 200
        np.testing.assert_allclose(t[i][m], r[i][m], rtol=1e-4, atol=0)
        assert np.all(t[i][~m] < 1e-30)


def test_rows_of_tiny_weights_keep_their_effective_support():
    """A row whose weights all lie near 1e-21 and below (every voxel 0.49-0.60 rad off the cone, sigma 0.05
    rad): the FP32 squares of its values fall under the FP32 normal range, so the kernels take the sum of
    squares of a row whose largest value is below 2^-40 again in FP64.  Its effective support S^2/Q must
    match the oracle's scale-free 1/sum((t/S)^2) (process_list_plane_strict.py:145-146), in the plain cone
    pass and in the pass that also forms the sensitivity."""
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    p1 = np.array([90.0, 10.0, 3.0])
    u = np.array([-0.6, -0.75, 0.1])
    u /= np.linalg.norm(u)
    w1 = np.cross(u, [0.0, 0.0, 1.0])
    w1 /= np.linalg.norm(w1)
    w2 = np.cross(u, w1)
    th = 0.522  # the scattering angle of the event below
    pts = [p1 - 50.0 * (np.cos(b) * u + np.sin(b) * (np.cos(ph) * w1 + np.sin(ph) * w2))
           for b in np.linspace(th + 0.49, th + 0.60, 64) for ph in np.linspace(0.0, 2 * np.pi, 16, endpoint=False)]
    coor = np.asarray(pts, np.float32)
    det = np.stack([p1, p1 + 30.0 * u]).astype(np.float32)
    lst = np.array([[1.0, 0.06, 2.0, 0.451]], np.float32)
    e0, tmax, tmin, tsum, sig = 0.511, 0.3397, 0.05, 0.46, 0.05
    ref = cone.backproj_list_single(None, det, coor, lst, 0.0, 0.0, e0, 0.02, tmax, tmin, tsum, None, None,
                                    fixed_sigma=sig)
    assert abs(ref["theta"][0] - th) < 1e-3
    for want_sens in (False, True):
        rows = build_compton_rows(None, torch.from_numpy(det), torch.from_numpy(coor), torch.from_numpy(lst.copy()),
                                  0.0, 0.0, e0, 0.02, tmax, tmin, tsum, DEV, noise="none", sigma_model="fixed",
                                  fixed_sigma=sig, store_dtype=torch.float32, want_sens=want_sens)
        S, Q, _, M = rows.stats[0].cpu().numpy()
        assert 1e-30 < M < 2.0 ** -40  # the FP64 sum-of-squares path
        assert abs(S / ref["S"][0] - 1.0) < 1e-4
        assert abs((S * S / Q) / ref["support"][0] - 1.0) < 1e-4


def test_non_finite_voxel_coordinates_raise():
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"].copy()
    coor[3, 1] = np.nan
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    with pytest.raises(ValueError):
        build_compton_rows(None, torch.from_numpy(det), torch.from_numpy(coor), torch.from_numpy(g["list_in"].copy()),
                           dr1, dr2, e0, res, tmax, tmin, tsum, DEV, noise="none")


def test_strict_rows_match_fp64_oracle_and_kept_indices_bit_exact():
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = _golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    ref = cone.backproj_list_single(A, det, coor, g["list_in"], dr1, dr2, e0, res, tmax, tmin, tsum,
                                    g["noise1"], g["noise2"])
    n1 = torch.from_numpy(g["noise1"]).to(DEV)
    n2 = torch.from_numpy(g["noise2"]).to(DEV)
    rows = build_compton_rows(torch.from_numpy(A), torch.from_numpy(det), torch.from_numpy(coor),
                              torch.from_numpy(g["list_in"].copy()), dr1, dr2, e0, res, tmax, tmin, tsum, DEV,
                              noise=(n1, n2), want_tc=True)
    kept = rows.kept_event_index().cpu().numpy()
    borderline = np.abs(ref["support"] - 50.0) / 50.0 < 1e-5
    assert not borderline.any()
    assert np.array_equal(kept, ref["kept"])
    t = rows.normalized_t().cpu().numpy()
    assert _rel_rows(t, ref["t"]) < 1e-4


def _small_list(n=600, rings=16, angles=64, seed=5):
    from jscc_recon_spect_cylinder_amd import synth
    return synth.compton_list(n_events=n, n_rings=rings, n_angles=angles, seed=seed)


@pytest.mark.parametrize("store", ["fp32", "bf16"])
def test_fixed_sigma_rows_match_oracle(store):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = _small_list()
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    sig = np.deg2rad(3.0)
    ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                    tsum, None, None, fixed_sigma=sig)
    rows = build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model="fixed", fixed_sigma=sig,
                              store_dtype=torch.float32 if store == "fp32" else torch.bfloat16)
    kept = rows.kept_event_index().cpu().numpy()
    border = np.abs(ref["support"] - 50.0) / 50.0 < 1e-4
    ok_ref = set(ref["kept"].tolist()) - set(np.nonzero(border)[0].tolist())
    ok_gpu = set(kept.tolist()) - set(np.nonzero(border)[0].tolist())
    assert ok_ref == ok_gpu
    if store == "fp32":
        common = np.isin(ref["kept"], kept)
        t = rows.normalized_t().cpu().numpy()
        tg = t[np.isin(kept, ref["kept"])]
        assert _rel_rows(tg, ref["t"][common]) < 1e-4
    else:
        t = rows.normalized_t().cpu().numpy()
        tg = t[np.isin(kept, ref["kept"])]
        common = np.isin(ref["kept"], kept)
        assert _rel_rows(tg, ref["t"][common]) < 5e-3  # BF16 storage: 2^-9 relative per element


def test_cpu_list_is_mutated_and_device_list_too():
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = _small_list(n=64)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    lst_cpu = torch.from_numpy(cl.list_origin.copy())
    torch.manual_seed(7)
    build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor), lst_cpu, 0.0, 0.0, cl.e0,
                       0.1, tmax, tmin, tsum, DEV, noise="cpu", sigma_model="fixed", fixed_sigma=0.05)
    torch.manual_seed(7)
    n1 = torch.randn(64).numpy()
    n2 = torch.randn(64).numpy()
    ev = cone.prepare_events(cl.list_origin, cl.detector, cl.e0, 0.1, tmax, tmin, tsum, n1, n2)
    assert np.array_equal(lst_cpu.numpy(), ev["list"])


def test_index_error_raises():
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = _small_list(n=32)
    lst = cl.list_origin.copy()
    lst[3, 0] = cl.detector.shape[0] + 5  # crystal number beyond the table
    with pytest.raises(IndexError):
        build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor), torch.from_numpy(lst),
                           0.0, 0.0, cl.e0, 0.02, 0.34, 0.05, 0.46, DEV, noise="none", sigma_model="fixed",
                           fixed_sigma=0.05)


def test_cpu_device_rejected():
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import get_compton_backproj_list_single
    with pytest.raises(ValueError):
        get_compton_backproj_list_single(torch.ones(2, 4), torch.ones(2, 3), torch.ones(4, 3), torch.ones(1, 4), 0, 0,
                                         0.511, 0.1, 0.34, 0.05, 0.46, torch.device("cpu"))


def test_fp8_hardware_decode_matches_ocp_software_decode():
    from jscc_recon_spect_cylinder_amd import ops
    out = ops.fp8_selftest(torch.device(DEV)).cpu().numpy()
    hw, sw = out[:256], out[256:]
    nan = np.isnan(sw)
    assert nan.sum() == 2                       # 0x7f and 0xff
    assert np.array_equal(np.isnan(hw), nan)
    assert np.array_equal(hw[~nan], sw[~nan])
    assert sw[0x7E] == 448.0 and sw[0x01] == 2.0 ** -9 and sw[0x08] == 2.0 ** -6
