"""Batched reconstructions (one launch for B lists of uneven length, T recomputed):
parity with the single-list no-store path, HIP-graph replay against eager runs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SIG = float(np.deg2rad(3.0))


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _lists(counts=(300, 1700, 40, 900), rings=16, angles=64, seed=3):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    rng = np.random.default_rng(seed)
    out = []
    for b, n in enumerate(counts):
        cl = synth.compton_list(n_events=n, n_rings=rings, n_angles=angles, seed=seed + b,
                                rods=synth.random_rods(rng))
        tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
        out.append(build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                      torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin,
                                      tsum, DEV, noise="none", sigma_model="fixed", fixed_sigma=SIG, store_dtype=None))
    return out


def test_batched_sensitivity_and_osem_match_single_lists():
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore, LazyListStore
    rows = _lists()
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    assert bs.n_kept == [r.n_kept for r in rows]
    P = rows[0].n_vox
    s = bs.sensitivity()
    x = torch.ones((len(rows), P), device=DEV)
    bs.osem(x, s, 3, 4)
    for b, r in enumerate(rows):
        lz = LazyListStore.from_rows(r, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
        sb = lz.backproject_ones(torch.ones(P, device=DEV))
        assert rel(s[b, :P].cpu().numpy(), sb.cpu().numpy()) < 1e-5
        xb = lz.osem(torch.ones(P, device=DEV), sb, 3, 4)
        assert rel(x[b, :P].cpu().numpy(), xb.cpu().numpy()) < 1e-5


def test_batched_graph_replays_equal_eager_iterations():
    """One captured iteration replayed 3 times, with eager work and synchronisation in
    between, reproduces each eager iteration exactly."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore
    rows = _lists(counts=(500, 1200, 200), seed=9)
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    P = rows[0].n_vox
    s = bs.sensitivity()
    xe = torch.ones((len(rows), P), device=DEV)
    eager = []
    for _ in range(3):
        bs.osem(xe, s, 1, 4)
        eager.append(xe.clone())
    xg = torch.ones((len(rows), P), device=DEV)
    gi = bs.capture(xg, s, 4)
    for it in range(3):
        gi.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device=DEV).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(xg, eager[it]), f"replay {it}"
        del junk


def test_batched_autotune_choice_is_consistent():
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore
    rows = _lists(counts=(800, 100, 1500), seed=21)
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    P = rows[0].n_vox
    s = bs.sensitivity()
    x0 = torch.ones((len(rows), P), device=DEV)
    ref = bs.osem(x0.clone(), s, 2, 4)
    c = bs.autotune(x0, s, 4)
    assert c in (1, 2, 4, 8) and bs.groups_per_cu == c
    out = bs.osem(x0.clone(), s, 2, 4)
    assert rel(out.cpu().numpy(), ref.cpu().numpy()) < 1e-6  # only the partial-sum grouping differs
