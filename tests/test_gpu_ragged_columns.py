"""Pixel counts that are not a multiple of the projection kernels' vector width (4 FP32, 8 BF16 / FP16 and 16
FP8 values): the rows are read in whole vectors, the ragged last one is cleared past the pixel count in the
thread that holds it, the image's last quad is read by element, and only the first P columns of the partial
sums are reduced.  Every check fills the rows' padding with NaN, so any leak of it into a sum shows, and
compares with FP64 products of the stored (dequantised) rows.  Column layouts cover one, two and four chunks
per thread, a tail in the last thread of a chunk and a tail in a chunk before the last.

Reference: recon_osem_plane.py:28-42 (weights), :45-80 (bin and list modes), :100-125 (joint mode) -- which
take any pixel count."""
import numpy as np
import pytest
import torch

from oracle import osem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SIG = float(np.deg2rad(3.0))


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _poison_padding(mat: torch.Tensor, P: int) -> None:
    """NaN in every stored column past P (FP8 e4m3fn: code 0x7f)."""
    if mat.shape[1] > P:
        mat[:, P:] = 0x7f if mat.dtype == torch.uint8 else float("nan")


def _dequant(mat: torch.Tensor, P: int, scale) -> np.ndarray:
    """FP64 values of stored rows [:, :P] times their per-row scale (None: 1)."""
    m = mat[:, :P].cpu()
    v = (m.view(torch.float8_e4m3fn).float() if m.dtype == torch.uint8 else m.float()).numpy().astype(np.float64)
    if scale is not None:
        v = v * scale.cpu().numpy().astype(np.float64)[: v.shape[0], None]
    return v


def _store(A: np.ndarray, dtype: str):
    """A [M, P] stored as `dtype` in rows of round_up(P, 64) with NaN padding: (mat, row_scale, FP64 rows)."""
    from jscc_recon_spect_cylinder_amd import ops
    M, P = A.shape
    At = torch.from_numpy(A).to(DEV)
    if dtype in ("fp32", "bf16"):
        tdt = torch.float32 if dtype == "fp32" else torch.bfloat16
        mat = torch.empty((M, ops.round_up(P, 64)), dtype=tdt, device=DEV)
        mat[:, :P] = At.to(tdt)
        scale = None
    else:
        mat, scale = ops.quantize_rows(At, P, M, None, dtype)
    _poison_padding(mat, P)
    return mat, scale, _dequant(mat, P, scale)


# 13: below every vector width; 1001: one chunk; 4093 (FP32) / 8190 (BF16, FP16): the tail in the last thread
# of the only chunk; 9001: two chunks (FP32 four, the tail in the third); 20001: BF16 / FP16 four chunks with
# the tail in the third, FP32 eight, FP8 two
@pytest.mark.parametrize("P", [13, 1001, 4093, 8190, 9001, 20001])
@pytest.mark.parametrize("dtype", ["fp32", "bf16", "fp16", "fp8"])
def test_projection_modes_on_ragged_columns(dtype, P):
    from jscc_recon_spect_cylinder_amd import ops
    print(f"ragged projection {dtype} P={P}", flush=True)
    rng = np.random.default_rng(P)
    M = 300
    A = (rng.random((M, P)) ** 3 + 1e-3).astype(np.float32)
    mat, sc, dec = _store(A, dtype)
    code = ops.dtype_code(dtype)
    x = (rng.random(P) + 0.5).astype(np.float32)
    y = (rng.poisson(20 * rng.random(M)) + 1).astype(np.float32)
    xd, yd = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    f = torch.empty(M, device=DEV)
    ops.proj(ops.MODE_FWD, mat, dtype_c=code, n_cols=P, n_rows=M, row_scale=sc, x=xd, fwd_out=f)
    fr = dec @ x
    assert rel(f.cpu().numpy(), fr) < 1e-5
    w = torch.zeros(P, device=DEV)
    ops.proj(ops.MODE_RATIO, mat, dtype_c=code, n_cols=P, n_rows=M, row_scale=sc, x=xd, y=yd, w=w)
    assert rel(w.cpu().numpy(), dec.T @ (y / fr)) < 1e-5
    # a row range past a row offset, through a row list
    rows = torch.from_numpy(rng.permutation(M).astype(np.int32)).to(DEV)
    w.zero_()
    ops.proj(ops.MODE_RATIO, mat, dtype_c=code, n_cols=P, rows=rows, row_base=37, n_rows=200, row_scale=sc, x=xd,
             w=w)
    sel = rows.cpu().numpy()[37:237]
    assert rel(w.cpu().numpy(), dec[sel].T @ (1.0 / (dec[sel] @ x))) < 1e-5
    w.zero_()
    ops.proj(ops.MODE_ONES, mat, dtype_c=code, n_cols=P, n_rows=M, row_scale=sc, w=w)
    assert rel(w.cpu().numpy(), dec.sum(0)) < 1e-5


@pytest.mark.parametrize("P", [1001, 9001])
@pytest.mark.parametrize("dtype", ["fp32", "bf16", "fp8"])
def test_rotation_gathers_on_ragged_columns(dtype, P):
    """x_rot = x[RotMat - 1] into the dot products and w = wt[RotMatInv - 1] out of the reduction
    (recon_osem_plane.py:45-61) with a ragged last vector."""
    from jscc_recon_spect_cylinder_amd import ops
    print(f"ragged rotation {dtype} P={P}", flush=True)
    rng = np.random.default_rng(7 + P)
    M = 256
    A = (rng.random((M, P)) ** 3 + 1e-3).astype(np.float32)
    mat, sc, dec = _store(A, dtype)
    rot = rng.permutation(P).astype(np.int64) + 1
    inv = np.argsort(rot - 1, kind="stable").astype(np.int64) + 1
    x = (rng.random(P) + 0.5).astype(np.float32)
    y = (rng.poisson(20 * rng.random(M)) + 1).astype(np.float32)
    w = torch.zeros(P, device=DEV)
    ops.proj(ops.MODE_RATIO, mat, dtype_c=ops.dtype_code(dtype), n_cols=P, n_rows=M, row_scale=sc,
             x=torch.from_numpy(x).to(DEV), x_perm=torch.from_numpy(rot).to(DEV), y=torch.from_numpy(y).to(DEV),
             inv_perm=torch.from_numpy(inv).to(DEV), w=w)
    xr = x.astype(np.float64)[rot - 1]
    wt = dec.T @ (y / (dec @ xr))
    assert rel(w.cpu().numpy(), wt[inv - 1]) < 1e-5
    ops.raise_on_index_error(torch.device(DEV))


def _ragged_list(n=900, seed=31):
    """A two-hit list on a 13 x 77 polar grid: 1001 voxels."""
    from jscc_recon_spect_cylinder_amd import synth
    return synth.compton_list(n_events=n, n_rings=13, n_angles=77, seed=seed)


@pytest.mark.parametrize("dtype,sync", [("fp32", True), ("fp32", False), ("bf16", False), ("fp16", False),
                                        ("fp8", True)])
def test_list_mode_osem_on_a_ragged_grid(dtype, sync):
    """T built on the GPU for a 1001-voxel grid, the rows' padding poisoned, then list-mode OSEM (3 x 4,
    device-side counts where sync is False) against osem_list_mode on the FP64 values of the rows the store
    projects; the column-sum pass against their column sums."""
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    print(f"ragged list-mode OSEM {dtype} sync={sync}", flush=True)
    cl = _ragged_list()
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    P = cl.coor.shape[0]
    assert P == 1001
    sd = {"fp32": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16, "fp8": torch.float32}[dtype]
    rows = build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model="fixed", fixed_sigma=SIG, store_dtype=sd, sync_counts=sync,
                              want_sens=dtype != "fp8")
    _poison_padding(rows.store, P)
    st = ListModeStore.from_rows(rows, dtype)
    if dtype == "fp8":
        _poison_padding(st.mat, P)
    x = torch.ones(P, device=DEV)
    s = st.backproject_ones()
    st.osem(x, s, 3, 4)
    st.resolve()
    kept = st.kept[:st.n_kept].long()
    T = _dequant(st.mat[kept], P, st.row_scale[kept])
    assert rel(s.cpu().numpy(), T.sum(0)) < 1e-5
    if dtype != "fp8":
        assert rel(rows.sens.cpu().numpy(), T.sum(0)) < 2e-6  # the cone pass's own sums
    xr = osem.osem_list_mode([[[T[lo:hi]]] for lo, hi in osem.split_bounds_even(T.shape[0], 4)],
                             np.arange(1, P + 1)[:, None], np.arange(1, P + 1)[:, None], np.ones(P), T.sum(0))
    for _ in range(2):
        xr = osem.osem_list_mode([[[T[lo:hi]]] for lo, hi in osem.split_bounds_even(T.shape[0], 4)],
                                 np.arange(1, P + 1)[:, None], np.arange(1, P + 1)[:, None], xr, T.sum(0))
    assert rel(x.cpu().numpy(), xr) < 1e-4


@pytest.mark.parametrize("dtype", ["fp32", "bf16", "fp16", "fp8"])
def test_system_store_osem_on_ragged_columns(dtype):
    """SC OSEM of a SystemStore with 1001 pixels (2 subsets x 3 iterations) against FP64 on its stored rows."""
    from jscc_recon_spect_cylinder_amd.engine import SystemStore
    print(f"ragged system store {dtype}", flush=True)
    rng = np.random.default_rng(5)
    M, P = 512, 1001
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.3) + 1e-4).astype(np.float32)
    y = rng.poisson(A @ (rng.random(P) * 30)).astype(np.float32)
    a = SystemStore.from_dense(torch.from_numpy(A).to(DEV), dtype)
    _poison_padding(a.mat, P)
    dec = _dequant(a.mat[:M], P, a.row_scale)
    s = a.column_sums()
    assert rel(s.cpu().numpy(), dec.sum(0)) < 1e-5
    x = a.osem(torch.ones(P, device=DEV), torch.from_numpy(y).to(DEV), s, 3, 2)
    xr = np.ones(P)
    for _ in range(3):
        for lo, hi in ((0, 256), (256, 512)):
            xr = xr * osem.weight_single(dec[lo:hi], y[lo:hi].astype(np.float64), xr) / dec.sum(0)
    assert rel(x.cpu().numpy(), xr) < 1e-5


def test_drop_in_modes_with_rotations_on_ragged_columns():
    """osem_bin_mode / osem_list_mode / osem_joint_mode (recon_osem_plane.py) on 1001 pixels with random
    rotation tables: the reference's contiguous [M, 1001] FP32 matrices go through zero-padded copies."""
    from jscc_recon_spect_cylinder_amd import recon_osem_plane as rp
    rng = np.random.default_rng(11)
    P, R, K = 1001, 2, 3
    rot = np.stack([rng.permutation(P) for _ in range(R)], axis=1).astype(np.int64) + 1
    inv = np.stack([np.argsort(rot[:, r] - 1, kind="stable") for r in range(R)], axis=1).astype(np.int64) + 1
    A = [(rng.random((64, P)) * (rng.random((64, P)) < 0.2) + 1e-4).astype(np.float32) for _ in range(K)]
    Y = [rng.poisson(20 * rng.random((64, R))).astype(np.float32) for _ in range(K)]
    T = [[[(rng.random((40, P)) ** 4).astype(np.float32)] for _ in range(R)] for _ in range(K)]
    s = (rng.random(P) + 1.0).astype(np.float32)
    x0 = np.ones(P, np.float32)
    rot_t, inv_t = torch.from_numpy(rot), torch.from_numpy(inv)
    img = torch.ones((P, 1), device=DEV)
    s_t = torch.from_numpy(s[:, None]).to(DEV)
    sys_l = [[torch.from_numpy(a).to(DEV)] for a in A]
    proj_l = [[torch.from_numpy(y).to(DEV)] for y in Y]
    t_l = [[[[torch.from_numpy(b).to(DEV)] for b in rb] for rb in ts] for ts in T]
    out = rp.osem_bin_mode(sys_l, proj_l, [rot_t], [inv_t], img, s_t, R)
    assert rel(out.cpu().numpy(), osem.osem_bin_mode(A, Y, rot, inv, x0, s)) < 1e-5
    out = rp.osem_list_mode(t_l, [rot_t], [inv_t], img, s_t, R)
    assert rel(out.cpu().numpy(), osem.osem_list_mode(T, rot, inv, x0, s)) < 1e-5
    out = rp.osem_joint_mode(sys_l, proj_l, t_l, [rot_t], [inv_t], img, s_t, 0.5, R)
    assert rel(out.cpu().numpy(), osem.osem_joint_mode(A, Y, T, rot, inv, x0, s, 0.5)) < 1e-5


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_joint_osem_on_ragged_columns(dtype):
    """engine.joint_osem (one projection launch per subset over the A and T rows) on 1001 pixels against
    osem_joint_mode on the FP64 values of both stores' rows."""
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore, SystemStore, joint_osem
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    print(f"ragged joint OSEM {dtype}", flush=True)
    cl = _ragged_list(seed=41)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    P = cl.coor.shape[0]
    rows = build_compton_rows(None, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model="fixed", fixed_sigma=SIG,
                              store_dtype=torch.bfloat16 if dtype == "bf16" else torch.float16, want_sens=True)
    _poison_padding(rows.store, P)
    t = ListModeStore.from_rows(rows, dtype)
    rng = np.random.default_rng(2)
    M = 400
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.3) + 1e-4).astype(np.float32)
    y = rng.poisson(A @ (rng.random(P) * 30)).astype(np.float32)
    a = SystemStore.from_dense(torch.from_numpy(A).to(DEV), dtype)
    _poison_padding(a.mat, P)
    decA = _dequant(a.mat[:M], P, a.row_scale)
    kept = t.kept[:t.n_kept].long()
    decT = _dequant(t.mat[kept], P, t.row_scale[kept])
    s = decA.sum(0) + decT.sum(0)
    x = torch.ones(P, device=DEV)
    joint_osem(a, torch.from_numpy(y).to(DEV), t, x, torch.from_numpy(s.astype(np.float32)).to(DEV), 0.5, 2, 3)
    ident = np.arange(1, P + 1)[:, None]
    ab, tb = osem.split_bounds_even(M, 3), osem.split_bounds_even(decT.shape[0], 3)
    xr = np.ones(P)
    for _ in range(2):
        xr = osem.osem_joint_mode([decA[lo:hi] for lo, hi in ab], [y[lo:hi, None].astype(np.float64) for lo, hi in ab],
                                  [[[decT[lo:hi]]] for lo, hi in tb], ident, ident, xr, s, 0.5)
    assert rel(x.cpu().numpy(), xr) < 1e-4


# ------------------------------------------------------------------ no-store projections (T recomputed)
def _ragged_rows(n, seed, store, sysmat=None, model="fixed"):
    from jscc_recon_spect_cylinder_amd import synth
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    cl = _ragged_list(n=n, seed=seed)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    return build_compton_rows(sysmat, torch.from_numpy(cl.detector), torch.from_numpy(cl.coor),
                              torch.from_numpy(cl.list_origin.copy()), 0.0, 0.0, cl.e0, 0.05, tmax, tmin, tsum, DEV,
                              noise="none", sigma_model=model, fixed_sigma=SIG, store_dtype=store)


@pytest.mark.parametrize("model", ["fixed", "strict"])
def test_lazy_store_on_a_ragged_grid_equals_stored_rows(model):
    """The no-store projection on 1001 voxels (whole quads: the padded voxels weighted zero by a factor row --
    the ragged factor row without a system matrix, the zero-padded system-matrix rows with one) against the
    stored FP32 rows: forward, column sums and 2 x 4 OSEM."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore, ListModeStore
    print(f"ragged lazy store {model}", flush=True)
    rng = np.random.default_rng(3)
    A = torch.from_numpy((rng.random((800, 1001)) + 0.05).astype(np.float32)) if model == "strict" else None
    rows = _ragged_rows(400, 21, torch.float32, sysmat=A, model=model)  # 800 crystals: one A row each
    P = rows.n_vox
    assert P % 4 == 1
    st = ListModeStore.from_rows(rows, "fp32")
    lz = LazyListStore.from_rows(rows, sysmat=A.to(DEV) if A is not None else None,
                                 sigma_model=ops.SIG_FIXED if model == "fixed" else ops.SIG_STRICT, fixed_sigma=SIG)
    x = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    f_lz, _ = lz.forward(x)
    assert rel(f_lz.cpu().numpy(), st.forward(x).cpu().numpy()) < 1e-5
    s_st = st.backproject_ones()
    s_lz = lz.backproject_ones(x)
    assert s_lz.shape == (P,)
    assert rel(s_lz.cpu().numpy(), s_st.cpu().numpy()) < 1e-5
    x1 = st.osem(torch.ones_like(s_st), s_st, 2, 4)
    x2 = lz.osem(torch.ones_like(s_st), s_st, 2, 4)
    assert rel(x2.cpu().numpy(), x1.cpu().numpy()) < 1e-5


def test_batched_store_on_a_ragged_grid_matches_single_lists():
    """Three lists of uneven length on 1001 voxels in one launch per subset step (images [B, 1004]) against
    each list's stored-row OSEM."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore, ListModeStore
    print("ragged batched store", flush=True)
    built = [_ragged_rows(n, 50 + b, torch.float32) for b, n in enumerate((500, 1400, 80))]
    bs = BatchedLazyStore.from_rows(built, sigma_model=ops.SIG_FIXED, fixed_sigma=SIG)
    P = built[0].n_vox
    s = bs.sensitivity()
    x = torch.ones((len(built), ops.round_up(P, 4)), device=DEV)
    bs.osem(x, s, 3, 4)
    for b, r in enumerate(built):
        st = ListModeStore.from_rows(r, "fp32")
        sb = st.backproject_ones()
        assert rel(s[b, :P].cpu().numpy(), sb.cpu().numpy()) < 1e-5
        xb = st.osem(torch.ones(P, device=DEV), sb, 3, 4)
        assert rel(x[b, :P].cpu().numpy(), xb.cpu().numpy()) < 1e-5
        assert torch.all(x[b, P:] == 1.0)  # the padding is never updated
