"""GPU parity of the fused coarse-row weight (jscc_coarse_weight): packed
[cpnum1, t_c] rows expanded on the fly, A-row factor, ratio back-projection,
against the FP64 oracle (dense expansion of compton_sparse_ops.py:331-342,
materialisation of :380-400, weights of recon_osem_plane_sparse.py:22-57 and
recon_osem_local_sparse_jsccsd_only.py:13-19), on the reference's
511keV_RotateNum20 polar grid."""
import numpy as np
import pytest
import torch

from oracle import layout, osem, sparse

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _odd_grid(n_z):
    """A layer of 300 pixels (10 rings x 30 angles, not a multiple of 64) over n_z slices, with
    identity rotation maps (every pixel its own source)."""
    from jscc_recon_spect_cylinder_amd import synth
    coor = synth.polar_grid_3d(np.arange(1, 11) * 5.0, [30], np.arange(n_z) * 3.0).astype(np.float32)
    ident = np.arange(1, coor.shape[0] + 1, dtype=np.int64)[:, None].repeat(8, axis=1)
    return coor, ident, ident


def _deep_grid():
    """The 1280-pixel layer of the 511keV factors over 24 slices: 30720 voxels (the kernel's
    32-slot register layout)."""
    from jscc_recon_spect_cylinder_amd import synth
    coor = synth.polar_grid_3d(np.arange(6, 151, 6), [20, 40, 60, 80], np.arange(24) * 3.0).astype(np.float32)
    ident = np.arange(1, coor.shape[0] + 1, dtype=np.int64)[:, None].repeat(8, axis=1)
    return coor, ident, ident


def _layer_grid(n_z):
    """The 1280-pixel layer of the 511keV factors over n_z slices (identity rotation maps): 12, 6 and 2
    slices take the 8-, 4- and 2-quad instantiations of the prefetching per-row kernel."""
    from jscc_recon_spect_cylinder_amd import synth
    coor = synth.polar_grid_3d(np.arange(6, 151, 6), [20, 40, 60, 80], np.arange(n_z) * 3.0).astype(np.float32)
    ident = np.arange(1, coor.shape[0] + 1, dtype=np.int64)[:, None].repeat(8, axis=1)
    return coor, ident, ident


@pytest.fixture(scope="module",
                params=[("ref", 1, 2), ("ref", 2, 2), ("ref", 2, 4), ("odd", 1, 2), ("odd", 2, 2), ("deep", 1, 2),
                        ("layer12", 1, 2), ("layer6", 1, 2), ("layer2", 1, 2)],
                ids=["theta1_z2", "theta2_z2", "theta2_z4", "odd_theta1_z2", "odd_theta2_z2", "deep_theta1_z2",
                     "slices12_theta1_z2", "slices6_theta1_z2", "slices2_theta1_z2"])
def grid(request):
    which, ts, zs = request.param
    if which == "ref":
        coor, rot, inv = layout.reference_grid("511keV_RotateNum20")
        rnum = 20
    elif which == "odd":
        coor, rot, inv = _odd_grid(6)
        rnum = None
    elif which.startswith("layer"):
        coor, rot, inv = _layer_grid(int(which[5:]))
        rnum = None
    else:
        coor, rot, inv = _deep_grid()
        rnum = None
    coor = coor.astype(np.float32)
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    p = build_compton_sparse_projector(torch.from_numpy(coor), ts, zs, rnum)
    pj = sparse.build(coor, ts, zs, rnum)
    return rot, inv, p, pj


def oracle_weight(ev, A, x_rot, pj, normalize):
    """nan_to_num(T^T (1 / max(T x_rot, 1e-12))) of the packed rows, FP64."""
    if normalize:
        with np.errstate(invalid="ignore"):
            rows, _ = sparse.materialize(ev, A, pj)
    else:
        cp = np.clip(np.round(ev[:, 0]).astype(np.int64) - 1, 0, A.shape[0] - 1)
        with np.errstate(invalid="ignore"):
            rows = sparse.upsample_rows(ev[:, 1:], pj) * np.asarray(A, np.float64)[cp]
    if rows.shape[0] == 0:
        return np.zeros(x_rot.shape[0])
    with np.errstate(invalid="ignore", over="ignore"):
        return osem.nan_to_num(osem.weight_compton(rows, x_rot, eps=1e-12))


def make_rows(p, n, n_arows, seed, power=3):
    rng = np.random.default_rng(seed)
    cp = rng.integers(1, n_arows + 1, n).astype(np.float32)
    tc = (rng.random((n, p.coarse_pixel_num)) ** power).astype(np.float32)
    return np.hstack([cp[:, None], tc]), rng


FORMS = pytest.mark.parametrize("form", [True, False], ids=["grouped_form", "per_row_form"])


@FORMS
@pytest.mark.parametrize("normalize", [True, False], ids=["materialised", "grouped"])
def test_fused_weight_vs_oracle(grid, normalize, form):
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    ax = p.tables.axis
    assert ops.coarse_weight_fits(ax)
    n, M = 700, 40
    ev, rng = make_rows(p, n, M, 11)
    ev[5, 1:] = 0.0          # all-zero row: dropped / contributes nothing
    ev[6, 0] = 0.0           # cpnum1 below range: clamped to the first A row
    ev[7, 0] = M + 9.0       # above range: clamped to the last
    ev[8, 0] = 3.5           # rounds half to even
    A = (rng.random((M, p.fine_pixel_num)) + 0.05).astype(np.float32)
    A[2, ::7] = 0.0
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    r = 7
    w0 = rng.random(p.fine_pixel_num).astype(np.float32)
    w = torch.from_numpy(w0.copy()).to(DEV)
    ops.coarse_weight(torch.from_numpy(ev).to(DEV), ax.to(DEV), torch.from_numpy(A).to(DEV),
                      torch.from_numpy(x).to(DEV), w, normalize=normalize, eps=1e-12,
                      x_perm=torch.from_numpy(rot[:, r].copy()).to(DEV),
                      inv_perm=torch.from_numpy(inv[:, r].copy()).to(DEV), out_scale=0.7, accumulate=True,
                      grouped=form)
    ops.raise_on_index_error(torch.device(DEV))
    ref = oracle_weight(ev, A, osem.rot_gather(x, rot[:, r]), pj, normalize)
    expect = w0 + 0.7 * osem.rot_gather(ref, inv[:, r])
    assert rel(w.cpu().numpy() - w0, expect - w0) < 2e-6


@FORMS
@pytest.mark.parametrize("normalize", [True, False], ids=["materialised", "grouped"])
def test_fused_weight_non_finite_rows(grid, normalize, form):
    """A non-finite coarse value spreads over its whole fine row in the reference's dense
    expansion: the row is dropped (materialised variant) or the block weight becomes
    NaN and is zeroed by the final nan_to_num (grouped variant)."""
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    ax = p.tables.axis.to(DEV)
    n, M = 300, 16
    ev, rng = make_rows(p, n, M, 12)
    ev[10, 1 + p.coarse_pixel_num // 3] = np.inf
    ev[200, 1 + 5] = np.nan
    A = (rng.random((M, p.fine_pixel_num)) + 0.05).astype(np.float32)
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    w = torch.zeros(p.fine_pixel_num, device=DEV)
    ops.coarse_weight(torch.from_numpy(ev).to(DEV), ax, torch.from_numpy(A).to(DEV), torch.from_numpy(x).to(DEV), w,
                      normalize=normalize, eps=1e-12, grouped=form)
    ref = oracle_weight(ev, A, x, pj, normalize)
    got = w.cpu().numpy()
    if normalize:
        keep = np.ones(n, bool)
        keep[[10, 200]] = False
        assert rel(got, oracle_weight(ev[keep], A, x, pj, True)) < 2e-6
        assert rel(got, ref) < 2e-6
    else:
        assert not np.any(ref) and not np.any(got)


@FORMS
@pytest.mark.parametrize("normalize", [True, False], ids=["materialised", "grouped"])
def test_fused_weight_non_finite_system_matrix(grid, normalize, form):
    """Infinite and NaN entries in the first-hit A rows with finite coarse values: the
    materialised variant's element-wise nan_to_num zeroes those products and keeps the
    rows (the kernel filters only when a row's sum comes out non-finite), the grouped
    variant has no element-wise filter."""
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    ax = p.tables.axis.to(DEV)
    n, M = 300, 16
    ev, rng = make_rows(p, n, M, 13)
    A = (rng.random((M, p.fine_pixel_num)) + 0.05).astype(np.float32)
    A[3, 11] = np.inf
    A[3, p.fine_pixel_num - 2] = -np.inf
    A[9, p.fine_pixel_num // 2] = np.nan
    used = np.isin(np.round(ev[:, 0]).astype(int) - 1, [3, 9])
    assert used.sum() >= 10
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    w = torch.zeros(p.fine_pixel_num, device=DEV)
    ops.coarse_weight(torch.from_numpy(ev).to(DEV), ax, torch.from_numpy(A).to(DEV), torch.from_numpy(x).to(DEV), w,
                      normalize=normalize, eps=1e-12, grouped=form)
    ref = oracle_weight(ev, A, x, pj, normalize)
    got = w.cpu().numpy()
    if normalize:
        assert np.all(np.isfinite(got)) and rel(got, ref) < 2e-6
        # the rows through the poisoned A rows contribute (nan_to_num per element, not a dropped row)
        assert rel(got, oracle_weight(ev[~used], A, x, pj, True)) > 1e-3
    else:
        # a NaN (or inf - inf) row sum makes every pixel of the block weight NaN: zero after nan_to_num
        assert not np.any(ref) and not np.any(got)


def test_fused_weight_equals_materialise_then_project(grid):
    """The fused pass and the two-kernel path (materialise fine rows, fused projection)
    give the same weight."""
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    pd = p.to(DEV)
    n, M = 257, 9
    ev, rng = make_rows(p, n, M, 13, power=5)
    A = torch.from_numpy((rng.random((M, p.fine_pixel_num)) + 0.05).astype(np.float32)).to(DEV)
    x = torch.from_numpy((rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)).to(DEV)
    evd = torch.from_numpy(ev).to(DEV)
    for normalize in (True, False):
        w1 = torch.zeros(p.fine_pixel_num, device=DEV)
        ops.coarse_weight(evd, pd.tables.axis, A, x, w1, normalize=normalize)
        rows, flags = ops.materialize(evd, pd.tables.expand, A, normalize=normalize)
        w2 = torch.zeros(p.fine_pixel_num, device=DEV)
        ops.proj(ops.MODE_RATIO, ops.proj_ready(rows, ops.DT_F32), x=x, eps=1e-12, w=w2, accumulate=True,
                 row_scale=flags.float() if normalize else None, sanitize=True)
        assert rel(w1.cpu(), w2.cpu()) < 1e-6


@FORMS
def test_fused_weight_row_counts_and_empty(grid, form):
    """Fewer rows than work-groups, one row, and no rows (weight untouched / zeroed)."""
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    ax = p.tables.axis.to(DEV)
    M = 5
    ev, rng = make_rows(p, 3, M, 14)
    A = (rng.random((M, p.fine_pixel_num)) + 0.05).astype(np.float32)
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    Ad, xd = torch.from_numpy(A).to(DEV), torch.from_numpy(x).to(DEV)
    for k in (1, 3):
        w = torch.zeros(p.fine_pixel_num, device=DEV)
        ops.coarse_weight(torch.from_numpy(ev[:k]).to(DEV), ax, Ad, xd, w, normalize=True, grouped=form)
        assert rel(w.cpu(), oracle_weight(ev[:k], A, x, pj, True)) < 2e-6
    w = torch.full((p.fine_pixel_num,), 2.0, device=DEV)
    ops.coarse_weight(torch.zeros((0, p.coarse_pixel_num + 1), device=DEV), ax, Ad, xd, w, normalize=True)
    assert torch.all(w == 2.0)
    ops.coarse_weight(torch.zeros((0, p.coarse_pixel_num + 1), device=DEV), ax, Ad, xd, w, normalize=True,
                      accumulate=False)
    assert torch.all(w == 0.0)


def test_sparse_drop_ins_on_a_grid_beyond_the_fused_kernel():
    """A 64 x 256 single-plane grid with theta stride 2 needs a 16384-entry layer table,
    more LDS than the fused kernel has: the drop-ins materialise and project instead,
    with the same result as the oracle."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd import recon_osem_local_sparse_jsccsd_only as loc
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    r = (np.arange(64) + 1) * 50.0 / 64
    coor = layout.polar_full(r, [256], np.array([0.0])).astype(np.float32)
    p = build_compton_sparse_projector(torch.from_numpy(coor), 2, 1, None)
    pj = sparse.build(coor, 2, 1, None)
    assert not ops.coarse_weight_fits(p.tables.axis)
    ev, rng = make_rows(p, 50, 6, 15)
    A = (rng.random((6, p.fine_pixel_num)) + 0.05).astype(np.float32)
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    w = loc.get_weight_compton_sparse(torch.from_numpy(ev).to(DEV), torch.from_numpy(A).to(DEV),
                                      torch.from_numpy(x[:, None]).to(DEV), p.to(DEV))
    assert rel(w.cpu().numpy().ravel(), oracle_weight(ev, A, x, pj, True)) < 2e-6


@pytest.mark.parametrize("normalize", [True, False], ids=["materialised", "grouped"])
@pytest.mark.parametrize("n_cryst", [3, 40])
def test_grouped_form_equals_per_row_form(grid, normalize, n_cryst):
    """Rows in first-hit crystal order (as a row list, and physically sorted): the grouped form (adjoints
    once per run of one crystal, runs spanning work-group boundaries) gives the per-row form's weight and the
    oracle's; with a rotation gather, both scatters and accumulation."""
    from jscc_recon_spect_cylinder_amd import ops
    rot, inv, p, pj = grid
    ax = p.tables.axis.to(DEV)
    n = 2000
    ev, rng = make_rows(p, n, n_cryst, 16)
    A = (rng.random((n_cryst, p.fine_pixel_num)) + 0.05).astype(np.float32)
    x = (rng.random(p.fine_pixel_num) + 0.2).astype(np.float32)
    evd, Ad, xd = torch.from_numpy(ev).to(DEV), torch.from_numpy(A).to(DEV), torch.from_numpy(x).to(DEV)
    order = torch.from_numpy(np.argsort(ev[:, 0], kind="stable").astype(np.int32)).to(DEV)
    r = 5
    kw = dict(normalize=normalize, eps=1e-12, x_perm=torch.from_numpy(rot[:, r].copy()).to(DEV),
              inv_perm=torch.from_numpy(inv[:, r].copy()).to(DEV), out_scale=1.3, accumulate=True)
    w0 = rng.random(p.fine_pixel_num).astype(np.float32)
    outs = []
    for rows, grouped, src in ((order, True, evd), (None, True, evd[order.long()].contiguous()),
                               (order, False, evd)):
        w = torch.from_numpy(w0.copy()).to(DEV)
        ops.coarse_weight(src, ax, Ad, xd, w, rows=rows, grouped=grouped, **kw)
        outs.append(w.cpu().numpy().astype(np.float64) - w0)
    ops.raise_on_index_error(torch.device(DEV))
    ref = 1.3 * osem.rot_gather(oracle_weight(ev, A, osem.rot_gather(x, rot[:, r]), pj, normalize), inv[:, r])
    for o in outs:
        assert rel(o, ref) < 2e-6
    assert rel(outs[0], outs[2]) < 1e-6 and rel(outs[1], outs[2]) < 1e-6
