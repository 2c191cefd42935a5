"""GPU parity of the synchronisation-free list-mode path (build_compton_rows(sync_counts=False),
jscc_cone_rows_counted, jscc_proj_counted / jscc_proj_em_counted) against the path that reads the
counts on the host: the same rows bit for bit, the same kept set, and the same sensitivity and OSEM
image (the split_tensor_even subset bounds of recon_osem_plane.py:8-18 taken on the device)."""
import numpy as np
import pytest
import torch

from oracle import cone, osem

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SIG = np.deg2rad(3.0)


def _build(cl, store, sync, lst=None):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    src = cl.list_origin if lst is None else lst
    return build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(src.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model="fixed", fixed_sigma=SIG, store_dtype=store,
                              sync_counts=sync)


@pytest.mark.parametrize("store", [torch.float32, torch.bfloat16], ids=["fp32", "bf16"])
@pytest.mark.parametrize("shape", [(16, 64, 600), (13, 40, 500), (64, 256, 3000)],
                         ids=["16x64", "13x40", "64x256"])
def test_deferred_counts_build_the_same_rows(store, shape):
    from jscc_recon_spect_cylinder_amd import synth
    rings, angles, n_ev = shape
    cl = synth.compton_list(n_events=n_ev, n_rings=rings, n_angles=angles, seed=31 + rings)
    a = _build(cl, store, True)
    b = _build(cl, store, False)
    assert b.n_sel == -1 and b.n_kept == -1 and b.counts_dev is not None
    assert b.store.shape[0] == n_ev and b.stats.shape[0] == n_ev
    stats_b = b.stats.cpu().numpy()
    b.resolve_counts()
    assert (b.n_sel, b.n_kept) == (a.n_sel, a.n_kept)
    assert a.n_sel < n_ev  # the energy windows drop events: the cone launch read its count on the device
    n = a.n_sel
    view = torch.int16 if store == torch.bfloat16 else torch.int32
    assert torch.equal(a.store[:n, :a.n_vox].contiguous().view(view), b.store[:n, :b.n_vox].contiguous().view(view))
    assert np.array_equal(a.stats[:n].cpu().numpy(), stats_b[:n])  # same kernel, same reduction order
    assert np.isnan(stats_b[n:]).all()  # rows past the count: dropped by the keep pass
    assert torch.equal(a.kept[:a.n_kept], b.kept[:b.n_kept])
    assert torch.equal(a.inv_sum[:n], b.inv_sum[:n])
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                    tsum, None, None, fixed_sigma=SIG)
    border = np.abs(ref["support"] - 50.0) / 50.0 < 1e-4
    if not border.any():
        assert np.array_equal(b.kept_event_index().cpu().numpy(), ref["kept"])


@pytest.mark.parametrize("n_ev,n_sub", [(4000, 4), (12000, 8), (8, 8)], ids=["4000x4", "12000x8", "8x8"])
def test_counted_sensitivity_and_osem_equal_the_host_counted_path(n_ev, n_sub):
    """backproject_ones and osem with the kept count on the device against the same calls with the count
    read first; subsets below the row count of a full grid use another grid, so those agree to rounding.
    8 events over 8 subsets leave some subsets empty (osem_list_mode then applies w = 0)."""
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    cl = synth.compton_list(n_events=n_ev, n_rings=16, n_angles=64, seed=77)
    out = {}
    for sync in (True, False):
        rows = _build(cl, torch.bfloat16, sync)
        st = ListModeStore.from_rows(rows, "bf16")
        assert (st.count_dev is None) == sync
        s = st.backproject_ones()
        x = torch.ones(st.n_vox, dtype=torch.float32, device=DEV)
        st.osem(x, s, 3, n_sub)
        out[sync] = (s.cpu().numpy().astype(np.float64), x.cpu().numpy().astype(np.float64), st.resolve().n_kept)
    assert out[True][2] == out[False][2]
    if n_ev == 8:
        # empty subsets: the update with w = 0 zeroes the image, the next subset divides by T x = 0, and the
        # image turns NaN / inf, as the reference's torch operations do; both paths must agree element-wise
        assert 0 < out[True][2] < n_sub
        np.testing.assert_array_equal(out[False][0], out[True][0])
        np.testing.assert_allclose(out[False][1], out[True][1], rtol=1e-5, atol=0, equal_nan=True)
        return
    rs = np.linalg.norm(out[False][0] - out[True][0]) / np.linalg.norm(out[True][0])
    rx = np.linalg.norm(out[False][1] - out[True][1]) / max(np.linalg.norm(out[True][1]), 1e-30)
    print(f"counted vs host-counted: sensitivity {rs:.3e}, image {rx:.3e} ({out[True][2]} kept rows)")
    assert rs <= 1e-6 and rx <= 1e-5


def test_counted_osem_matches_the_fp64_oracle():
    """The synchronisation-free FP32 path against the FP64 oracle of osem_list_mode (recon_osem_plane.py:63-80)."""
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    cl = synth.compton_list(n_events=3000, n_rings=16, n_angles=64, seed=91)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                    tsum, None, None, fixed_sigma=SIG)
    T = ref["t"].astype(np.float64)
    P = T.shape[1]
    ident = np.arange(1, P + 1, dtype=np.int64)[:, None]  # one rotation, identity tables
    subsets = [[[T[lo:hi]]] for lo, hi in osem.split_bounds_even(T.shape[0], 4)]
    xr = np.ones(P)
    for _ in range(5):
        xr = osem.osem_list_mode(subsets, ident, ident, xr, T.sum(0))
    rows = _build(cl, torch.float32, False)
    st = ListModeStore.from_rows(rows, "fp32")
    s = st.backproject_ones()
    x = torch.ones(st.n_vox, dtype=torch.float32, device=DEV)
    st.osem(x, s, 5, 4)
    xg = x.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(xg - xr) / np.linalg.norm(xr)
    print(f"counted FP32 OSEM vs FP64 oracle: {rel:.3e}")
    assert rel <= 1e-4


def test_deferred_index_error_raises_on_resolve():
    from jscc_recon_spect_cylinder_amd import synth
    cl = synth.compton_list(n_events=32, n_rings=16, n_angles=64, seed=5)
    lst = cl.list_origin.copy()
    lst[3, 0] = cl.detector.shape[0] + 5  # crystal number beyond the table
    rows = _build(cl, torch.float32, False, lst=lst)
    with pytest.raises(IndexError):
        rows.resolve_counts()


def test_quantised_store_from_deferred_rows_reads_the_counts():
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    cl = synth.compton_list(n_events=800, n_rings=16, n_angles=64, seed=12)
    a = ListModeStore.from_rows(_build(cl, torch.float32, True), "fp16")
    b = ListModeStore.from_rows(_build(cl, torch.float32, False), "fp16")
    assert a.n_kept == b.n_kept and b.count_dev is None
    sa, sb = a.backproject_ones(), b.backproject_ones()
    assert torch.equal(sa, sb)


@pytest.mark.parametrize("sync", [True, False], ids=["host_counts", "device_counts"])
def test_list_mode_graph_replays_equal_eager_iterations(sync):
    """One captured list-mode iteration replayed 3 times, with eager work and synchronisation in between,
    reproduces each eager iteration exactly -- with the kept count read on the host, and left on the device
    (the captured launches take their subset bounds from it)."""
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    cl = synth.compton_list(n_events=6000, n_rings=32, n_angles=128, seed=23)
    st = ListModeStore.from_rows(_build(cl, torch.bfloat16, sync), "bf16")
    assert (st.count_dev is None) == sync
    s = st.backproject_ones()
    xe = torch.ones(st.n_vox, dtype=torch.float32, device=DEV)
    eager = []
    for _ in range(3):
        st.osem(xe, s, 1, 8)
        eager.append(xe.clone())
    xg = torch.ones(st.n_vox, dtype=torch.float32, device=DEV)
    gi = st.capture(xg, s, 8)
    for it in range(3):
        gi.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device=DEV).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(xg, eager[it]), f"replay {it}"
        del junk
    x3 = torch.ones(st.n_vox, dtype=torch.float32, device=DEV)
    st.osem(x3, s, 3, 8, graph=True)
    assert torch.equal(x3, eager[2])
