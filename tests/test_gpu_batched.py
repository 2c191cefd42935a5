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
    tile, c, ms = bs.autotune(x0, s, 4)
    assert c in (1, 2, 4, 8) and bs.groups_per_cu == c
    assert 0 <= tile < len(ops.cone_batch_tiles(P)) and bs.tile == tile and ms > 0
    out = bs.osem(x0.clone(), s, 2, 4)
    assert rel(out.cpu().numpy(), ref.cpu().numpy()) < 1e-6  # only the partial-sum grouping differs


@pytest.mark.parametrize("rings,angles", [(16, 64), (16, 320), (32, 256), (64, 256)])
def test_batched_every_tile_shape_matches_the_single_list_path(rings, angles):
    """Each tile shape (threads x voxel quads x rows per step) at 1024, 5120, 8192 and 16384 voxels,
    with 1 and 8 work-groups per CU, against the single-list no-store path of every list."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore, LazyListStore
    rows = _lists(counts=(700, 60, 1300), rings=rings, angles=angles, seed=rings + angles)
    P = rows[0].n_vox
    shapes = ops.cone_batch_tiles(P)
    assert len(shapes) == 3 and all(t * q * 4 >= P for t, q, _ in shapes)
    refs = []
    for r in rows:
        lz = LazyListStore.from_rows(r, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
        sb = lz.backproject_ones(torch.ones(P, device=DEV))
        refs.append((sb, lz.osem(torch.ones(P, device=DEV), sb, 2, 3)))
    for tile in range(len(shapes)):
        for gpc in (1, 8):
            print(f"P {P} tile {tile} {shapes[tile]} groups/CU {gpc}", flush=True)
            bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
            bs.tile, bs.groups_per_cu = tile, gpc
            s = bs.sensitivity()
            x = torch.ones((len(rows), P), device=DEV)
            bs.osem(x, s, 2, 3)
            for b, (sb, xb) in enumerate(refs):
                assert rel(s[b, :P].cpu().numpy(), sb.cpu().numpy()) < 1e-5
                assert rel(x[b, :P].cpu().numpy(), xb.cpu().numpy()) < 1e-5


def test_batched_tile_queries_and_rejection():
    from jscc_recon_spect_cylinder_amd import _lib, ops
    assert [len(ops.cone_batch_tiles(p)) for p in (4, 4096, 4100, 8192, 16384, 16388, 32768, 32772)] == \
        [3, 3, 3, 3, 3, 1, 1, 0]
    assert ops.cone_batch_tiles(1024)[0] == (256, 4, 2)
    rows = _lists(counts=(50, 80), seed=4)
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    bs.tile = 3
    with pytest.raises(RuntimeError, match="tile"):
        bs.sensitivity()
