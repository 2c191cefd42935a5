"""No-store cone projection (rows recomputed in registers) and the differentiable
operators: parity with the stored-T path, and gradients against FP64 central
finite differences of the oracle's log-likelihood."""
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


def _build(n=400, rings=16, angles=64, seed=21, store=torch.float32, sysmat=None, sigma_model="fixed"):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = synth.compton_list(n_events=n, n_rings=rings, n_angles=angles, seed=seed)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    rows = build_compton_rows(sysmat, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model=sigma_model, fixed_sigma=SIG, store_dtype=store)
    return cl, rows


@pytest.mark.parametrize("model", ["fixed", "strict"])
def test_lazy_store_equals_stored_rows(model):
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore, ListModeStore
    rng = np.random.default_rng(0)
    A = None
    if model == "strict":
        A = torch.from_numpy((rng.random((800, 1024)) + 0.05).astype(np.float32))
    cl, rows = _build(sysmat=A, sigma_model=model)
    st = ListModeStore.from_rows(rows, "fp32")
    lz = LazyListStore.from_rows(rows, sysmat=A.to(DEV) if A is not None else None,
                                 sigma_model=ops.SIG_FIXED if model == "fixed" else ops.SIG_STRICT, fixed_sigma=SIG)
    x = torch.from_numpy((rng.random(1024) + 0.5).astype(np.float32)).to(DEV)
    f_st = st.forward(x).cpu().numpy()
    f_lz, _ = lz.forward(x)
    assert rel(f_lz.cpu().numpy(), f_st) < 1e-5
    s_st = st.backproject_ones()
    s_lz = lz.backproject_ones(x)
    assert rel(s_lz.cpu().numpy(), s_st.cpu().numpy()) < 1e-5
    x1 = st.osem(torch.ones_like(s_st), s_st, 2, 4)
    x2 = lz.osem(torch.ones_like(s_st), s_st, 2, 4)
    assert rel(x2.cpu().numpy(), x1.cpu().numpy()) < 1e-5


def _oracle_L(ev, keep_idx, coor, e0, x, sigma):
    sel = np.isin(ev["idx"], keep_idx)
    th = cone.theta_from_e1(ev["e1"][sel], e0)
    kn = cone.klein_nishina(ev["e1"][sel], e0)
    tc = cone.cone_tc(ev["pos1"][sel], ev["pos2"][sel], None, None, th, kn, None, None, coor, fixed_sigma=sigma)
    T = tc / tc.sum(1, keepdims=True)
    f = T @ x
    return np.sum(np.log(f)) - np.sum(f), T


def test_cone_autograd_matches_fp64_finite_differences():
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.differentiable import cone_project, list_loglik
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore
    cl, rows = _build(n=300, store=None)
    lz = LazyListStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    rng = np.random.default_rng(5)
    x0 = (rng.random(cl.coor.shape[0]) + 0.5)
    x = torch.tensor(x0, dtype=torch.float32, device=DEV, requires_grad=True)
    sigma = torch.tensor(SIG, dtype=torch.float64, device=DEV, requires_grad=True)
    L = list_loglik(cone_project(x, sigma, lz))
    L.backward()
    gx = x.grad.cpu().numpy().astype(np.float64)
    gs = float(sigma.grad)

    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    ev = cone.prepare_events(cl.list_origin, cl.detector, cl.e0, 0.05, tmax, tmin, tsum)
    keep = rows.kept_event_index().cpu().numpy()
    xs = x0.astype(np.float32).astype(np.float64)
    L0, T = _oracle_L(ev, keep, cl.coor, cl.e0, xs, SIG)
    assert abs(float(L.detach()) - L0) / abs(L0) < 1e-5
    # analytic x-gradient of the oracle (T^T(1/Tx) - T^T 1) and 64 directional FD checks
    ref_gx = osem.weight_compton(T, xs) - T.sum(0)
    assert rel(gx, ref_gx) < 1e-4
    h = 1e-6
    gnorm = np.linalg.norm(gx)
    for _ in range(64):
        v = rng.normal(size=xs.shape)
        v /= np.linalg.norm(v)
        fd = (np.sum(np.log(T @ (xs + h * v))) - np.sum(T @ (xs + h * v))
              - np.sum(np.log(T @ (xs - h * v))) + np.sum(T @ (xs - h * v))) / (2 * h)
        # FP32 gradient: error scale is |grad| (random directions nearly cancel)
        assert abs(fd - gx @ v) <= 1e-5 * gnorm + 1e-4 * abs(fd)
    # sigma gradient: FP64 central difference through the normalised rows
    fd_s = (_oracle_L(ev, keep, cl.coor, cl.e0, xs, SIG + h)[0] - _oracle_L(ev, keep, cl.coor, cl.e0, xs, SIG - h)[0]) \
        / (2 * h)
    assert abs(gs - fd_s) <= 1e-3 * abs(fd_s) + 1e-6


@pytest.mark.parametrize("deg", [1.0, 6.0])
def test_cone_sigma_gradient_range(deg):
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.differentiable import cone_project, list_loglik
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore
    sig0 = float(np.deg2rad(deg))
    cl, rows = _build(n=250, store=None, seed=33)
    lz = LazyListStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=sig0)
    x0 = np.ones(cl.coor.shape[0])
    x = torch.ones(cl.coor.shape[0], device=DEV)
    sigma = torch.tensor(sig0, dtype=torch.float64, device=DEV, requires_grad=True)
    list_loglik(cone_project(x, sigma, lz)).backward()
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    ev = cone.prepare_events(cl.list_origin, cl.detector, cl.e0, 0.05, tmax, tmin, tsum)
    keep = rows.kept_event_index().cpu().numpy()
    h = 1e-6
    fd = (_oracle_L(ev, keep, cl.coor, cl.e0, x0, sig0 + h)[0] - _oracle_L(ev, keep, cl.coor, cl.e0, x0, sig0 - h)[0]) \
        / (2 * h)
    assert abs(float(sigma.grad) - fd) <= 2e-3 * abs(fd) + 1e-6


def test_system_autograd_matches_textbook_gradient():
    from jscc_recon_spect_cylinder_amd.differentiable import bin_loglik, system_project
    rng = np.random.default_rng(8)
    A = (rng.random((512, 1024)) * (rng.random((512, 1024)) < 0.3)).astype(np.float32) + 1e-4
    xs = rng.random(1024) + 0.5
    y = rng.poisson(A @ xs * 5).astype(np.float32)
    x = torch.tensor(xs, dtype=torch.float32, device=DEV, requires_grad=True)
    L = bin_loglik(system_project(x, torch.from_numpy(A).to(DEV)), torch.from_numpy(y).to(DEV))
    L.backward()
    x32 = xs.astype(np.float32).astype(np.float64)
    ref = osem.weight_single(A, y, x32) - A.astype(np.float64).sum(0)
    assert rel(x.grad.cpu().numpy(), ref) < 1e-5
