"""GPU parity of the no-store and batched cone projections on rows wider than 16384 voxels
(25600 = 100 x 256 and 32768 = 128 x 256: the shapes whose row values are staged in LDS instead
of registers) against the stored-row path and the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import cone, osem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SIG = float(np.deg2rad(3.0))


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _rows(n, rings, angles, seed, store=torch.float32, sysmat=None, model="fixed"):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = synth.compton_list(n_events=n, n_rings=rings, n_angles=angles, seed=seed)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    rows = build_compton_rows(sysmat, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model=model, fixed_sigma=SIG, store_dtype=store)
    return cl, rows


@pytest.mark.parametrize("rings,model", [(100, "fixed"), (100, "strict"), (128, "fixed")],
                         ids=["25600-fixed", "25600-strict-A", "32768-fixed"])
def test_lazy_store_wide_rows_equal_stored_rows(rings, model):
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore, ListModeStore
    rng = np.random.default_rng(rings)
    P = rings * 256
    A = None
    if model == "strict":
        A = torch.from_numpy((rng.random((800, P)) + 0.05).astype(np.float32))
    cl, rows = _rows(300, rings, 256, 40 + rings, sysmat=A, model=model)
    assert rows.n_vox == P and rows.n_kept > 20
    st = ListModeStore.from_rows(rows, "fp32")
    lz = LazyListStore.from_rows(rows, sysmat=A.to(DEV) if A is not None else None,
                                 sigma_model=ops.SIG_FIXED if model == "fixed" else ops.SIG_STRICT, fixed_sigma=SIG)
    x = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    f_st = st.forward(x).cpu().numpy()
    f_lz, _ = lz.forward(x)
    assert rel(f_lz.cpu().numpy(), f_st) < 1e-5
    s_st = st.backproject_ones()
    s_lz = lz.backproject_ones(x)
    assert rel(s_lz.cpu().numpy(), s_st.cpu().numpy()) < 1e-5
    x1 = st.osem(torch.ones_like(s_st), s_st, 2, 4)
    x2 = lz.osem(torch.ones_like(s_st), s_st, 2, 4)
    err = rel(x2.cpu().numpy(), x1.cpu().numpy())
    print(f"lazy {P} {model}: osem vs stored rows {err:.3e}")
    assert err < 1e-5
    if model == "fixed":
        # and the no-store OSEM against the FP64 oracle of osem_list_mode (recon_osem_plane.py:63-80)
        from jscc_recon_spect_cylinder_amd import synth
        tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
        ref = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin, 0.0, 0.0, cl.e0, 0.05, tmax, tmin,
                                        tsum, None, None, fixed_sigma=SIG)
        T = ref["t"]
        ident = np.arange(1, P + 1, dtype=np.int64)[:, None]
        subsets = [[[T[lo:hi]]] for lo, hi in osem.split_bounds_even(T.shape[0], 4)]
        xr = np.ones(P)
        for _ in range(2):
            xr = osem.osem_list_mode(subsets, ident, ident, xr, T.sum(0))
        e_or = rel(x2.cpu().numpy(), xr)
        print(f"lazy {P}: osem vs FP64 oracle {e_or:.3e}")
        assert e_or < 1e-4


def test_no_store_adjoint_wide_rows_matches_stored_rows():
    """The coefficient-mode adjoint of the no-store rows (the differentiable operator's backward) at 25600
    voxels against T^T c of the stored rows in FP64."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore
    rng = np.random.default_rng(9)
    cl, rows = _rows(300, 100, 256, 77)
    P = rows.n_vox
    lz = LazyListStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    x = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    coef = torch.from_numpy(rng.random(rows.n_kept).astype(np.float32)).to(DEV)
    w_lz, _ = lz.adjoint(x, coef)
    T = rows.normalized_t()
    w_ref = (T.double().t() @ coef.double()).cpu().numpy()
    assert rel(w_lz.cpu().numpy(), w_ref) < 1e-5


def test_batched_wide_rows_match_single_lists():
    from jscc_recon_spect_cylinder_amd import ops, synth
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore, LazyListStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    rng = np.random.default_rng(4)
    rows = []
    for b, n in enumerate((250, 90, 400)):
        cl = synth.compton_list(n_events=n, n_rings=100, n_angles=256, seed=60 + b, rods=synth.random_rods(rng))
        tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
        rows.append(build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                                       torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin,
                                       tsum, DEV, noise="none", sigma_model="fixed", fixed_sigma=SIG,
                                       store_dtype=None))
    P = rows[0].n_vox
    assert P == 25600
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    s = bs.sensitivity()
    x = torch.ones((len(rows), P), device=DEV)
    bs.osem(x, s, 2, 4)
    for b, r in enumerate(rows):
        lz = LazyListStore.from_rows(r, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
        sb = lz.backproject_ones(torch.ones(P, device=DEV))
        assert rel(s[b, :P].cpu().numpy(), sb.cpu().numpy()) < 1e-5
        xb = lz.osem(torch.ones(P, device=DEV), sb, 2, 4)
        err = rel(x[b, :P].cpu().numpy(), xb.cpu().numpy())
        print(f"batched 25600 list {b} ({r.n_kept} rows): vs single list {err:.3e}")
        assert err < 1e-5
