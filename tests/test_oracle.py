"""CPU tests that pin the oracle (oracle/) against the reference's fixtures,
golden vectors and closed forms.  No GPU needed."""
import hashlib
import math
import os

import numpy as np
import pytest

from oracle import cone, layout, osem
from tests.conftest import GOLDEN
from tests.golden.golden_inputs import golden_sysmat

FIX = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")
STRICT = os.path.join(GOLDEN, "strict_cone_small.npz")


def _golden_sysmat(det, coor):
    return golden_sysmat(det, coor)


# ------------------------------------------------------------------ layout (genSysMatPolar_3D.m)
def test_layout_reproduces_reference_factor_files():
    f = np.load(FIX)
    coor, rot, inv = layout.reference_grid("511keV_RotateNum20")
    assert coor.shape == (25600, 3)
    assert hashlib.sha256(np.ascontiguousarray(np.round(coor, 9)).tobytes()).hexdigest() == str(f["coor_sha"])
    assert hashlib.sha256(np.ascontiguousarray(rot).tobytes()).hexdigest() == str(f["rot_sha"])
    assert hashlib.sha256(np.ascontiguousarray(inv).tobytes()).hexdigest() == str(f["inv_sha"])
    np.testing.assert_allclose(coor[:1280], f["coor_layer0"], atol=1e-12)
    assert np.array_equal(rot[:1280], f["rot_layer0"].astype(np.int64))


def test_rotmat_inverse_is_exact_inverse():
    _, rot, inv = layout.reference_grid("511keV_RotateNum20")
    for r in range(rot.shape[1]):
        assert np.array_equal(rot[inv[:, r] - 1, r], np.arange(1, rot.shape[0] + 1))
    assert np.array_equal(rot[:, 0], np.arange(1, rot.shape[0] + 1))  # column 1 is the identity


# ------------------------------------------------------------------ cone closed forms
def test_theta_closed_forms():
    e0 = 0.511
    assert math.isclose(cone.theta_from_e1(e0 / 3, e0), math.pi / 3, rel_tol=1e-12)
    assert math.isclose(cone.theta_from_e1(e0 / 2, e0), math.pi / 2, rel_tol=1e-12)
    edge = 2 * e0 ** 2 / (0.511 + 2 * e0)
    # at the Compton edge cos(theta) = -1, clamped to the FP32 bound
    assert math.isclose(cone.theta_from_e1(edge, e0), math.acos(-cone.COS_MAX), rel_tol=1e-12)
    c = 1 - 0.511 * (edge - 0.001) / ((e0 - edge + 0.001) * e0)
    assert abs(c + 0.98249) < 5e-5


def test_klein_nishina_closed_form():
    e0 = 0.511
    assert math.isclose(cone.klein_nishina(e0 / 3, e0), 1.5 + 2 / 3, rel_tol=1e-12)
    assert math.isclose(cone.klein_nishina(e0 / 2, e0), 2.5, rel_tol=1e-12)


def test_sigma_ene_is_theta_of_perturbed_energy():
    e0, res = 0.511, 0.1138
    e1 = np.array([0.1, 0.2, 0.3])
    sp, sm = cone.angle_sigma_ene_pm(e1, e0, res)
    se = e1 * res / 2.355 * np.sqrt(e0 / e1)
    th = np.arccos(1 - 0.511 * e1 / ((e0 - e1) * e0))
    thp = np.arccos(1 - 0.511 * (e1 + se) / ((e0 - e1 - se) * e0))
    thm = np.arccos(1 - 0.511 * (e1 - se) / ((e0 - e1 + se) * e0))
    np.testing.assert_allclose(sp, thp - th, rtol=1e-12)
    np.testing.assert_allclose(sm, th - thm, rtol=1e-12)


def test_sigma_pos_matches_finite_difference_of_beta():
    """g = d beta / d pos2 (and -d beta / d pos1 through the inter-crystal leg): the
    strict sigma_pos^2 equals sum_k (d beta/d pos2_k)^2 (s1_k + s2_k)."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        p1 = rng.normal(0, 100, 3) + np.array([0, 220, 0])
        p2 = p1 + rng.normal(0, 30, 3)
        c = rng.normal(0, 40, 3)
        w = rng.uniform(0.2, 3.0, 3)
        h = 1e-6
        grad = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            grad[k] = (cone.beta_of(p1, p2 + e, c) - cone.beta_of(p1, p2 - e, c)) / (2 * h)
        fd = float(np.sum(grad ** 2 * w))
        assert math.isclose(cone.sigma_pos_sq(p1, p2, c, w), fd, rel_tol=2e-5)


def test_detector_variances_from_reference_table():
    det = np.load(FIX)["detector"]
    s = cone.detector_sigma_sq(det, 0.0)
    ay = np.abs(det[:, 1])
    assert sorted(np.unique(ay).tolist()) == [200.0, 230.0, 260.0, 290.0]
    np.testing.assert_allclose(s[ay < 289], 0.75, rtol=1e-7)
    np.testing.assert_allclose(s[ay > 289], np.tile([1 / 3, 3.0, 1 / 3], (int((ay > 289).sum()), 1)), rtol=1e-6)
    s2 = cone.detector_sigma_sq(det, 2.0)
    np.testing.assert_allclose(s2 - s, 4.0, rtol=1e-6)
    # any other layer count: isotropic extra sigma only
    odd = det[ay < 250]
    np.testing.assert_allclose(cone.detector_sigma_sq(odd, 1.5), 2.25)
    np.testing.assert_allclose(cone.detector_sigma_sq(odd, -1.0), 0.0)


def test_oracle_matches_reference_golden_vectors():
    """Golden vectors written by the reference's own get_compton_backproj_list_single."""
    g = np.load(STRICT)
    det = np.load(FIX)["detector"]
    coor = g["coor"]
    A = _golden_sysmat(det, coor)
    dr1, dr2, e0, res, tmax, tmin, tsum = g["params"]
    out = cone.backproj_list_single(A, det, coor, g["list_in"], dr1, dr2, e0, res, tmax, tmin, tsum,
                                    g["noise1"], g["noise2"])
    assert np.array_equal(out["list"], g["list_out"])           # in-place smeared energies, bit-exact
    assert out["t"].shape == g["t"].shape                       # same kept events
    rel = np.linalg.norm(out["t"] - g["t"], axis=1) / np.linalg.norm(g["t"], axis=1)
    assert rel.max() < 2e-5
    rel = np.linalg.norm(out["t_compton"] - g["t_compton"], axis=1) / np.linalg.norm(g["t_compton"], axis=1)
    assert rel.max() < 2e-5
    np.testing.assert_allclose(out["t_single"].sum(1), g["t_single_sum"], rtol=1e-5)


def test_oracle_filters_catch_planted_mistake():
    """A planted op-order change in the FP32 event stage (E0/e instead of
    reciprocal(e)*E0) must change the smeared energies of the golden list."""
    g = np.load(STRICT)
    e1 = g["list_in"][:, 1]
    F = np.float32
    res = F(g["params"][3])
    good = ((e1 * res) / F(2.355)) * np.sqrt((F(1.0) / e1) * F(0.511))
    planted = ((e1 * res) / F(2.355)) * np.sqrt(F(0.511) / e1)
    assert not np.array_equal(good, planted)
    assert np.array_equal((e1 + (good * g["noise1"]).astype(F)).astype(F), g["list_out"][:, 1])


def test_cone_rotation_equivariance_under_rotmat():
    """Rotating the event by -2*pi*r/R about z permutes t_c columns exactly as RotMat[:, r]."""
    f = np.load(FIX)
    coor = f["coor_layer0"]  # [1280, 3], z = -28.5 mm
    rot = f["rot_layer0"].astype(np.int64)
    rng = np.random.default_rng(3)
    p1 = np.array([[12.0, 210.0, 3.0], [-150.0, -201.0, -5.0]])
    p2 = p1 + rng.normal(0, 25, (2, 3))
    th = np.array([0.9, 1.7])
    kn = np.array([2.3, 2.1])
    w = np.ones((2, 3))
    t0 = cone.cone_tc(p1, p2, w, w, th, kn, th * 0 + 0.03, th * 0 + 0.02, coor)
    for r in (1, 7):
        a = -2 * math.pi * r / 20
        R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        t1 = cone.cone_tc(p1 @ R.T, p2 @ R.T, w, w, th, kn, th * 0 + 0.03, th * 0 + 0.02, coor)
        np.testing.assert_allclose(t1, t0[:, rot[:, r] - 1], rtol=1e-7, atol=1e-12)


def test_cone_peak_on_cone_surface():
    """Fixed sigma: a voxel exactly on the cone (beta = theta) gets KN - sin^2(theta)."""
    p1 = np.array([[0.0, 200.0, 0.0]])
    p2 = np.array([[0.0, 230.0, 0.0]])
    th = 0.8
    # voxel on the cone: direction from p1 making angle th with v12 = +y
    vox = p1[0] - 100.0 * np.array([math.sin(th), math.cos(th), 0.0])
    t = cone.cone_tc(p1, p2, None, None, np.array([th]), np.array([2.2]), None, None, vox[None, :], fixed_sigma=0.05)
    assert math.isclose(t[0, 0], 2.2 - math.sin(th) ** 2, rel_tol=1e-10)


def test_row_filters_support_definition():
    t = np.zeros((3, 100))
    t[0, :60] = 1.0      # support 60 -> kept
    t[1, :40] = 1.0      # support 40 -> dropped
    t[2, 0] = np.nan     # NaN -> dropped
    keep, S, sup = cone.row_filters(t)
    assert keep.tolist() == [True, False, False]
    assert math.isclose(sup[0], 60.0) and math.isclose(sup[1], 40.0)


# ------------------------------------------------------------------ OSEM oracle
def test_weight_single_is_the_adjoint_of_the_forward_projection():
    """With y = r * (A x) the oracle's weight_single is A^T r, so <x, w> = <A x, r>; with a
    transpose-free mistake (A instead of A^T on the way back) the shapes or the identity fail."""
    rng = np.random.default_rng(0)
    A = rng.random((50, 40))
    x, r = rng.random(40) + 0.1, rng.random(50)
    w = osem.weight_single(A, r * (A @ x), x)
    assert w.shape == (40,)
    assert math.isclose(float(x @ w), float((A @ x) @ r), rel_tol=1e-12)
    np.testing.assert_allclose(w, np.einsum("ij,i->j", A, r), rtol=1e-12)


def test_em_fixed_point_and_count_preservation():
    rng = np.random.default_rng(1)
    A = rng.random((60, 30)) * (rng.random((60, 30)) < 0.5)
    xs = rng.random(30) + 0.5
    y = A @ xs
    s = A.sum(0)
    x1 = osem.osem_bin_mode([A], [y[:, None]], np.arange(1, 31)[:, None], np.arange(1, 31)[:, None], xs, s)
    np.testing.assert_allclose(x1, xs, rtol=1e-12)
    x0 = rng.random(30) + 0.1
    x2 = osem.osem_bin_mode([A], [y[:, None]], np.arange(1, 31)[:, None], np.arange(1, 31)[:, None], x0, s)
    assert math.isclose(float(s @ x2), float(y.sum()), rel_tol=1e-12)


def test_list_mode_weight_row_scale_invariance():
    rng = np.random.default_rng(2)
    T = rng.random((20, 16))
    x = rng.random(16)
    d = rng.uniform(0.1, 10, 20)
    np.testing.assert_allclose(osem.weight_compton(T, x), osem.weight_compton(T * d[:, None], x), rtol=1e-12)


def test_safe_and_apply_updates_special_cases():
    x = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    w = np.array([2.0, np.nan, -1.0, np.inf, 1.0])
    s = np.array([1.0, 1.0, 1.0, 1.0, 0.0])
    np.testing.assert_array_equal(osem.em_safe(x, w, s), [2.0, 0.0, 0.0, 0.0, 0.0])
    out = osem.em_apply(x, w, s)
    # w/s: 2 ; nan -> 1 ; -1 ; inf -> 1e6 ; 1/1e-12 = 1e12
    np.testing.assert_allclose(out, [2.0, 2.0, -3.0, 4e6, 5e12])


def test_subset_bounds_match_torch_semantics():
    torch = pytest.importorskip("torch")
    for total, parts in [(10, 3), (7, 8), (25000, 8), (1023, 5)]:
        t = torch.arange(total)
        ch = torch.chunk(t, parts)
        assert [(int(c[0]), int(c[-1]) + 1) for c in ch] == osem.chunk_bounds(total, parts)
    assert osem.split_bounds_even(10, 3) == [(0, 3), (3, 6), (6, 10)]


def test_sensitivity_closed_form():
    A = np.arange(12, dtype=float).reshape(3, 4)
    inv = np.array([[1, 2], [2, 1], [3, 4], [4, 3]])
    s = osem.sensitivity(A, inv)
    cs = A.sum(0)
    np.testing.assert_allclose(s, (cs + cs[[1, 0, 3, 2]]) / 2)


def test_scd_thinning_oracle_equals_the_literal_expansion():
    """The implicit-expansion thinning equals the reference's literal expand/compare algorithm."""
    import torch
    from oracle import scd
    rng = np.random.default_rng(12)
    for _ in range(20):
        counts = rng.integers(0, 6, 17).astype(np.float32)
        counts[rng.random(17) < 0.2] += 0.5  # half counts round to even, like Python's round
        n = int(sum(round(float(c)) for c in counts))
        if n == 0:
            continue
        k = int(rng.integers(0, n + 1))
        perm = torch.randperm(n).numpy()[:k]
        np.testing.assert_array_equal(scd.thin(counts, perm), scd.thin_literal(counts, perm))


def test_scd_equal_count_projection_properties():
    import torch
    from oracle import scd
    rng = np.random.default_rng(13)
    proj = rng.poisson(3.0, (40, 3)).astype(np.float32)
    torch.manual_seed(5)
    k = [10, int(proj[:, 1].sum()), 0]
    out = scd.equal_count_projection(proj, k)
    assert out[:, 0].sum() == 10 and np.all(out <= proj)
    np.testing.assert_array_equal(out[:, 1], proj[:, 1])  # all counts kept
    assert not out[:, 2].any()
