"""CPU pins of the oracle functions that the golden vectors do not reach:
the rotation scatter and alpha weighting of the OSEM modes, the sparse
materialisation, the SCD down-sampling and the list-mode sensitivity.

Each check is built from something other than the oracle's own formula: a
brute-force loop, a permuted-matrix restatement derived from the algebra, a
closed-form invariant, or the reference's own outputs (golden vectors).  Where
a plausible mistake could slip past, the test also shows that the mistaken call
fails the same check."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import osem, scd, sparse
from tests.conftest import GOLDEN


def _perm_pair(rng, P):
    """A random permutation as 1-based (RotMat, RotMatInv) columns, not an involution."""
    while True:
        p = rng.permutation(P)
        if not np.array_equal(p[p], np.arange(P)):
            break
    rot = p + 1
    inv = np.empty(P, np.int64)
    inv[p] = np.arange(P)
    return rot, inv + 1


def _tables(rng, P, R):
    rots, invs = zip(*(_perm_pair(rng, P) for _ in range(R)))
    return np.stack(rots, 1), np.stack(invs, 1)


# ------------------------------------------------------------------ GEMV weights (brute force)
def _loops_weight(M, num, x, eps):
    n, P = M.shape
    w = [0.0] * P
    for i in range(n):
        f = sum(float(M[i, j]) * float(x[j]) for j in range(P))
        if eps is not None:
            f = max(f, eps)
        r = float(num[i]) / f
        for j in range(P):
            w[j] += float(M[i, j]) * r
    return np.array(w)


def test_weight_functions_equal_brute_force_loops():
    """recon_osem_plane.py:28-42 (+ the clamp of recon_osem_local_cntstat.py:6-8)."""
    rng = np.random.default_rng(40)
    A = rng.random((9, 7)) * (rng.random((9, 7)) < 0.6)
    A[3] = 0.0                                   # a zero row: A x = 0, clamped or inf/nan
    x = rng.random(7) + 0.1
    y = rng.poisson(3.0, 9).astype(float) + 1.0
    np.testing.assert_allclose(osem.weight_single(A, y, x, 1e-12), _loops_weight(A, y, x, 1e-12), rtol=1e-12)
    np.testing.assert_allclose(osem.weight_compton(np.delete(A, 3, 0), x), _loops_weight(np.delete(A, 3, 0),
                                                                                         np.ones(8), x, None),
                               rtol=1e-12)
    # unclamped: the zero row's 0 * inf turns every weight into NaN, as in torch
    assert np.isnan(osem.weight_single(A, y, x)).all()


# ------------------------------------------------------------------ rotations (recon_osem_plane.py:45-125)
def _bin_brute(A, Y, rot, inv, x, s):
    """One bin-mode subset: index_select(x, rot-1) through A equals A[:, inv-1] applied to x, and the
    RotMatInv gather of the weight equals the transpose of that same permuted matrix."""
    w = np.zeros_like(x)
    for r in range(rot.shape[1]):
        B = A[:, inv[:, r] - 1]
        w += B.T @ (Y[:, r] / (B @ x))
    return x * w / s


def _list_brute(blocks_per_rot, inv, x, s):
    w = np.zeros_like(x)
    for r, blocks in enumerate(blocks_per_rot):
        for T in blocks:
            B = T[:, inv[:, r] - 1]
            w += B.T @ (1.0 / (B @ x))
    return x * w / s


@pytest.fixture(scope="module")
def rot_problem():
    rng = np.random.default_rng(41)
    P, R, K = 30, 3, 2
    rot, inv = _tables(rng, P, R)
    A = [rng.random((12, P)) * (rng.random((12, P)) < 0.5) + 1e-3 for _ in range(K)]
    Y = [rng.poisson(4.0, (12, R)).astype(float) + 1.0 for _ in range(K)]
    T = [[[rng.random((5, P)) + 1e-3, rng.random((3, P)) + 1e-3] for _ in range(R)] for _ in range(K)]
    x0 = rng.random(P) + 0.5
    s = rng.random(P) + 1.0
    return dict(rot=rot, inv=inv, A=A, Y=Y, T=T, x0=x0, s=s, R=R, K=K)


def test_bin_mode_rotations_equal_permuted_matrices(rot_problem):
    p = rot_problem
    x = p["x0"]
    for A, Y in zip(p["A"], p["Y"]):
        x = _bin_brute(A, Y, p["rot"], p["inv"], x, p["s"])
    got = osem.osem_bin_mode(p["A"], p["Y"], p["rot"], p["inv"], p["x0"], p["s"])
    np.testing.assert_allclose(got, x, rtol=1e-12)
    # a rot/inv swap (gathering x with RotMatInv and w with RotMat) fails the same check
    swapped = osem.osem_bin_mode(p["A"], p["Y"], p["inv"], p["rot"], p["x0"], p["s"])
    assert not np.allclose(swapped, x, rtol=1e-6)


def test_list_mode_rotations_and_blocks_equal_permuted_matrices(rot_problem):
    p = rot_problem
    x = p["x0"]
    for t_sub in p["T"]:
        x = _list_brute(t_sub, p["inv"], x, p["s"])
    got = osem.osem_list_mode(p["T"], p["rot"], p["inv"], p["x0"], p["s"])
    np.testing.assert_allclose(got, x, rtol=1e-12)
    swapped = osem.osem_list_mode(p["T"], p["inv"], p["rot"], p["x0"], p["s"])
    assert not np.allclose(swapped, x, rtol=1e-6)
    # a dropped block fails it too
    short = [[blocks[:1] for blocks in t_sub] for t_sub in p["T"]]
    assert not np.allclose(osem.osem_list_mode(short, p["rot"], p["inv"], p["x0"], p["s"]), x, rtol=1e-6)


def test_joint_mode_alpha_limits(rot_problem):
    """alpha = 2 leaves only the single-photon weight (times 2), alpha = 0 only the Compton weight (times 2):
    recon_osem_plane.py:117 (weight = (2 - alpha) * weight_compton + alpha * weight_single)."""
    p = rot_problem
    args = (p["rot"], p["inv"], p["x0"])
    j2 = osem.osem_joint_mode(p["A"], p["Y"], p["T"], *args, p["s"], 2.0)
    np.testing.assert_allclose(j2, osem.osem_bin_mode(p["A"], p["Y"], *args, p["s"] / 2), rtol=1e-12)
    j0 = osem.osem_joint_mode(p["A"], p["Y"], p["T"], *args, p["s"], 0.0)
    np.testing.assert_allclose(j0, osem.osem_list_mode(p["T"], *args, p["s"] / 2), rtol=1e-12)


def test_joint_mode_count_identity(rot_problem):
    """One subset: sum_j s_j x+_j = sum_j x_j w_j = (2 - alpha) * (number of list rows over all rotations
    and blocks) + alpha * (sum of the subset's bin counts over all rotations), whatever the rotations."""
    p = rot_problem
    for alpha in (0.3, 0.5, 1.7):
        A, Y, T = p["A"][:1], p["Y"][:1], p["T"][:1]
        x1 = osem.osem_joint_mode(A, Y, T, p["rot"], p["inv"], p["x0"], p["s"], alpha)
        n_rows = sum(b.shape[0] for blocks in T[0] for b in blocks)
        expect = (2 - alpha) * n_rows + alpha * Y[0].sum()
        assert math.isclose(float(p["s"] @ x1), expect, rel_tol=1e-12)
        # the weights the other way round ((2 - alpha) on the bins) break it
        assert not math.isclose(alpha * n_rows + (2 - alpha) * Y[0].sum(), expect, rel_tol=1e-6)


def test_bin_mode_rotated_fixed_point():
    """With y_r = B_r x* for every rotation and s = sum_r B_r^T 1, x* is a fixed point of the rotated bin mode."""
    rng = np.random.default_rng(42)
    P, R = 20, 4
    rot, inv = _tables(rng, P, R)
    A = rng.random((15, P)) * (rng.random((15, P)) < 0.5) + 1e-3
    xs = rng.random(P) + 0.5
    Y = np.stack([A @ xs[rot[:, r] - 1] for r in range(R)], 1)
    s = sum(A[:, inv[:, r] - 1].T @ np.ones(15) for r in range(R))
    np.testing.assert_allclose(osem.osem_bin_mode([A], [Y], rot, inv, xs, s), xs, rtol=1e-12)


# ------------------------------------------------------------------ materialise (compton_sparse_ops.py:380-400)
def _tiny_grid():
    """4 z slices x (a 4-pixel ring at r = 10 and an 8-pixel ring at r = 20): 48 fine pixels; theta stride 2
    and z stride 2 give 2 x (2 + 4) = 12 coarse pixels."""
    pts = []
    for z in (-3.0, -1.0, 1.0, 3.0):
        for r, n in ((10.0, 4), (20.0, 8)):
            for k in range(n):
                a = 2 * math.pi * k / n
                pts.append((r * math.cos(a), r * math.sin(a), z))
    return np.array(pts, np.float32)


def _materialize_loops(er, A, pj):
    """Per element: fine(zf, lf) = sum_zc sum_lc z_expand[zf, zc] layer_expand[lf, lc] t_c[zc, lc] as Python
    floats (0 * nan and 0 * inf give nan, as in the reference's dense matmuls), times A[clamp(round(c) - 1)],
    non-finite products set to 0, rows kept where the sum is finite and positive, then divided by it."""
    ze, le = pj["z_expand"], pj["layer_expand"]
    Zf, Zc = ze.shape
    Lf, Lc = le.shape
    rows, valid = [], []
    for row in er:
        c = int(round(float(row[0]))) - 1           # Python round: half to even, like torch.round
        c = min(max(c, 0), A.shape[0] - 1)
        t = []
        for zf in range(Zf):
            for lf in range(Lf):
                v = 0.0
                for zc in range(Zc):
                    for lc in range(Lc):
                        v += float(ze[zf, zc]) * float(le[lf, lc]) * float(row[1 + zc * Lc + lc])
                pr = v * float(A[c, zf * Lf + lf])
                t.append(pr if math.isfinite(pr) else 0.0)
        S = sum(t)
        ok = math.isfinite(S) and S > 0
        valid.append(ok)
        if ok:
            rows.append([v / S for v in t])
    return np.array(rows).reshape(-1, Zf * Lf), np.array(valid)


def test_materialize_equals_per_element_loops():
    coor = _tiny_grid()
    pj = sparse.build(coor, theta_stride=2, z_stride=2)
    assert pj["coor_coarse"].shape[0] == 12 and pj["fine_layer"] == 12 and pj["fine_z"] == 4
    rng = np.random.default_rng(43)
    M = 6
    A = rng.uniform(0.5, 2.0, (M, 48))
    A[4, 7] = 0.0
    A[2, 9] = np.nan           # non-finite A entries: only those products are zeroed, the row stays valid
    A[2, 11] = np.inf
    tc = rng.random((11, 12))
    cp = np.array([1, 2, 3, 0, M + 5, 2.5, 3.5, 1, 2, 3, 4], np.float64)  # 0 and M+5 clamp; .5 rounds to even
    tc[7, 5] = np.nan          # poisons the whole fine row -> dropped
    tc[8, 2] = np.inf          # 0 * inf = nan -> whole row dropped
    tc[9] = 0.0                # zero sum -> dropped
    tc[10] = -tc[10]           # negative sum -> dropped
    er = np.concatenate([cp[:, None], tc], 1)
    got, valid = sparse.materialize(er, A, pj)
    ref, ref_valid = _materialize_loops(er, A, pj)
    assert valid.tolist() == ref_valid.tolist() == [True] * 7 + [False] * 4
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)
    assert got[2, 9] == 0.0 and got[2, 11] == 0.0 and got[2].sum() == pytest.approx(1.0)
    # the clamps and the rounding matter: rows 3, 4, 5 and 6 select A rows 0, M-1, 1 and 3
    for i, a_row in ((3, 0), (4, M - 1), (5, 1), (6, 3)):
        one, _ = sparse.materialize(np.concatenate([[a_row + 1.0], tc[i]])[None], A, pj)
        np.testing.assert_allclose(got[i], one[0], rtol=1e-12)
    floor_row, _ = sparse.materialize(np.concatenate([[3.0], tc[6]])[None], A, pj)  # floor(3.5) -> A row 2
    assert not np.allclose(floor_row[0], got[6], rtol=1e-6)


# ------------------------------------------------------------------ SCD down-sampling (sparse_main_utils.py:130-160)
SCD_GOLD = os.path.join(GOLDEN, "scd_downsample_golden.npz")


def test_downsample_matches_reference_golden_draws():
    """The reference's own outputs for three seeded cases: same bin counts and the same list rows in the
    same order, which fixes the order of the permutations drawn from the global generator."""
    g = np.load(SCD_GOLD)
    for k in range(int(g["n_cases"])):
        proj = g[f"c{k}_proj_in"]
        lists = [g[f"c{k}_list{r}_in"] for r in range(proj.shape[1])]
        torch.manual_seed(int(g[f"c{k}_seed"]))
        p_out, l_out = scd.downsample_projection_and_list(proj, lists, float(g[f"c{k}_ds"]))
        np.testing.assert_array_equal(p_out.numpy(), g[f"c{k}_proj_out"])
        for r in range(proj.shape[1]):
            np.testing.assert_array_equal(l_out[r].numpy(), g[f"c{k}_list{r}_out"])


def test_downsample_count_rules():
    rng = np.random.default_rng(44)
    proj = rng.poisson(3.0, (31, 5)).astype(np.float32)
    proj[0, 0] = 2.5                                        # rounds to 2 (half to even)
    lists = [rng.normal(size=(n, 4)).astype(np.float32) for n in (40, 1, 7, 0, 13)]
    same_p, same_l = scd.downsample_projection_and_list(proj, lists, 0.9999995)
    assert same_p is proj and same_l is lists               # identity at ds >= 0.999999
    for ds in (0.37, 0.05):
        torch.manual_seed(3)
        p_out, l_out = scd.downsample_projection_and_list(proj, lists, ds)
        p_out = p_out.numpy()
        for r in range(proj.shape[1]):
            n = int(sum(round(float(v)) for v in proj[:, r]))
            want = min(int(torch.round(torch.from_numpy(proj[:, r]).sum() * ds).item()), n)
            assert p_out[:, r].sum() == want
            assert np.all(p_out[:, r] <= np.array([round(float(v)) for v in proj[:, r]]))  # a thinning
            m = lists[r].shape[0]
            assert l_out[r].shape[0] == (min(max(int(m * ds), 1), m) if m else 0)
            # rows drawn without replacement from the input list
            src = {tuple(v) for v in lists[r].tolist()}
            got = [tuple(v) for v in l_out[r].numpy().tolist()]
            assert len(set(got)) == len(got) and set(got) <= src


# ------------------------------------------------------------------ list-mode sensitivity (main_plane.py:348-355)
def test_sensitivity_from_t_invariants():
    rng = np.random.default_rng(45)
    s = rng.random(16) + 0.2
    blocks = [rng.random((9, 16)), rng.random((4, 16)) * 3.0]
    n_c, n_s = 1234, 98765
    d = scd.sensitivity_compton_from_t(blocks, s, n_c, n_s)
    # each block contributes sum(s) * n_c / n_s in total
    assert math.isclose(d.sum(), len(blocks) * s.sum() * n_c / n_s, rel_tol=1e-12)
    # a block's contribution is its column sum up to a scale: scaling a block changes nothing
    d2 = scd.sensitivity_compton_from_t([blocks[0] * 17.0, blocks[1] * 1e-3], s, n_c, n_s)
    np.testing.assert_allclose(d2, d, rtol=1e-12)
    # one block whose column sums are proportional to s gives s * n_c / n_s
    one = scd.sensitivity_compton_from_t([np.outer(np.arange(1.0, 6.0), s)], s, n_c, n_s)
    np.testing.assert_allclose(one, s * n_c / n_s, rtol=1e-12)
    # the count ratio the other way round (n_s / n_c) fails the total
    assert not math.isclose(d.sum(), len(blocks) * s.sum() * n_s / n_c, rel_tol=1e-6)


# ------------------------------------------------------------------ the CPU baseline's torch restatement
def test_torch_cpu_baseline_agrees_with_the_fp64_oracle():
    """The FP32 torch restatement that bench.py times as the reference's CPU path keeps the oracle's events
    and reproduces its rows and its 10 x 8 list-mode image (FP32 against FP64)."""
    from jscc_recon_spect_cylinder_amd import synth
    from oracle import cone, torch_cpu
    cl = synth.compton_list(n_events=700, n_rings=16, n_angles=64, seed=9)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    sig = float(np.deg2rad(3.0))
    g = torch.Generator().manual_seed(4)
    n1, n2 = torch.randn(700, generator=g), torch.randn(700, generator=g)
    ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                    tsum, n1.numpy(), n2.numpy(), fixed_sigma=sig)
    T, kept = torch_cpu.cone_rows_fixed_sigma(torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                              torch.from_numpy(cl.list_origin.copy()), cl.e0, 0.02, tmax, tmin, tsum,
                                              n1, n2, sig)
    assert np.array_equal(kept.numpy(), ref["kept"])
    assert np.abs(T.double().numpy() - ref["t"]).max() / ref["t"].max() < 1e-4
    x = torch_cpu.list_osem(T, 10, 8).double().numpy()
    xr = np.ones(T.shape[1])
    s = ref["t"].sum(0)
    for _ in range(10):
        for lo, hi in osem.split_bounds_even(ref["t"].shape[0], 8):
            xr = osem.em_plain(xr, osem.weight_compton(ref["t"][lo:hi], xr), s)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-4
