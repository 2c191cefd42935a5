"""Coarse/fine projector: the engine's builder equals the reference's bit for bit
(golden digests), the oracle's builder equals hand-computed operators, and the
transforms satisfy the adjoint identity.  CPU only."""
import hashlib
import os

import numpy as np
import pytest
import torch

from oracle import layout, sparse
from tests.conftest import GOLDEN

G = os.path.join(GOLDEN, "sparse_projector_golden.npz")


def digest(t):
    return hashlib.sha256(np.ascontiguousarray(np.asarray(t, np.float32)).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def coor_full():
    coor, _, _ = layout.reference_grid("511keV_RotateNum20")
    return coor.astype(np.float32)


def test_hand_computed_operators_fine4_stride2():
    np.testing.assert_array_equal(sparse.periodic_reduce(4, 2), [[1, 0, 0, 0], [0, 0, 1, 0]])
    np.testing.assert_array_equal(sparse.periodic_expand(4, 2), [[1, 0], [0.5, 0.5], [0, 1], [0.5, 0.5]])
    np.testing.assert_array_equal(sparse.linear_reduce(4, 2), [[0.5, 0.5, 0, 0], [0, 0, 0.5, 0.5]])
    np.testing.assert_array_equal(sparse.linear_expand(4, 2), [[1, 0], [0.75, 0.25], [0.25, 0.75], [0, 1]])
    assert sparse.ring_stride(20, 20, 2) == 1      # one angle per rotation step: no sub-sampling
    assert sparse.ring_stride(80, 20, 3) == 2      # largest stride dividing the ring and its step of 4
    assert sparse.ring_stride(60, 7, 4) == 4       # rings not split by the rotations: divisor of the ring
    assert sparse.ring_stride(21, None, 4) == 3


def test_engine_builder_matches_reference_digests(coor_full):
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    g = np.load(G)
    for k, (ts, zs, rot) in enumerate(g["settings"]):
        p = build_compton_sparse_projector(torch.from_numpy(coor_full), int(ts), int(zs),
                                           None if rot < 0 else int(rot))
        assert list(p.ring_strides) == g[f"s{k}_ring_strides"].tolist()
        assert [p.fine_layer_num, p.fine_z_num, p.coarse_layer_num, p.coarse_z_num] == g[f"s{k}_shape"].tolist()
        for name in ("layer_reduce", "z_reduce", "layer_expand", "z_expand", "coor_coarse"):
            assert digest(getattr(p, name).numpy()) == str(g[f"s{k}_{name}"]), (k, name)


def test_oracle_builder_matches_reference(coor_full):
    g = np.load(G)
    for k, (ts, zs, rot) in enumerate(g["settings"]):
        pj = sparse.build(coor_full, int(ts), int(zs), None if rot < 0 else int(rot))
        assert pj["ring_strides"] == g[f"s{k}_ring_strides"].tolist()
        np.testing.assert_allclose(pj["z_expand"], g[f"s{k}_z_expand_full"], atol=1e-7)
        np.testing.assert_allclose(pj["z_reduce"], g[f"s{k}_z_reduce_full"], atol=1e-7)
        assert digest(pj["layer_reduce"]) == str(g[f"s{k}_layer_reduce"])
        assert digest(pj["layer_expand"]) == str(g[f"s{k}_layer_expand"])


def test_expand_adjoint_identity(coor_full):
    pj = sparse.build(coor_full, 2, 2, 20)
    rng = np.random.default_rng(0)
    a = rng.random((3, pj["coor_coarse"].shape[0]))
    b = rng.random((3, coor_full.shape[0]))
    lhs = np.sum(sparse.upsample_rows(a, pj) * b)
    rhs = np.sum(a * sparse.adjoint_rows(b, pj))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_engine_csr_tables_equal_dense_operators(coor_full):
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    p = build_compton_sparse_projector(torch.from_numpy(coor_full), 2, 2, 20)
    pj = sparse.build(coor_full, 2, 2, 20)
    rng = np.random.default_rng(1)
    x = rng.random(coor_full.shape[0])
    red = p.tables.reduce
    dense = np.zeros((red.n_out, red.n_in))
    ptr, col, val = red.ptr.numpy(), red.col.numpy(), red.val.numpy()
    for j in range(red.n_out):
        dense[j, col[ptr[j]:ptr[j + 1]]] = val[ptr[j]:ptr[j + 1]]
    np.testing.assert_allclose(dense @ x, sparse.reduce_rows(x[None, :], pj)[0], rtol=1e-6)
    assert p.tables.taps is not None


def _dense_of_taps(idx, w, n_coarse):
    m = np.zeros((idx.shape[0], n_coarse), np.float64)
    for f in range(idx.shape[0]):
        m[f, idx[f, 0]] += w[f, 0]
        m[f, idx[f, 1]] += w[f, 1]
    return m


def test_axis_taps_rebuild_the_expand_operators(coor_full):
    """The per-axis taps of the fused coarse-row kernel are the projector's z_expand and
    layer_expand entry for entry (identity layer maps carry no table), for every golden
    projector setting."""
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    g = np.load(G)
    for k, (ts, zs, rot) in enumerate(g["settings"]):
        p = build_compton_sparse_projector(torch.from_numpy(coor_full), int(ts), int(zs),
                                           None if rot < 0 else int(rot))
        ax = p.tables.axis
        ze = p.z_expand.double().numpy()
        le = p.layer_expand.double().numpy()
        assert (ax.n_zf, ax.n_zc, ax.n_lf, ax.n_lc) == (ze.shape[0], ze.shape[1], le.shape[0], le.shape[1])
        np.testing.assert_array_equal(_dense_of_taps(ax.z_idx.numpy(), ax.z_w.numpy(), ax.n_zc), ze)
        if ax.l_idx is None:
            np.testing.assert_array_equal(le, np.eye(le.shape[0]))
        else:
            np.testing.assert_array_equal(_dense_of_taps(ax.l_idx.numpy(), ax.l_w.numpy(), ax.n_lc), le)
        assert (ax.l_idx is None) == all(s == 1 for s in p.ring_strides), k


def test_axis_taps_reject_three_tap_rows():
    from jscc_recon_spect_cylinder_amd import ops
    with pytest.raises(ValueError):
        ops.axis_taps(np.eye(3), np.array([[0.5, 0.25, 0.25], [0, 1, 0], [0, 0, 1]]))


def test_axis_taps_adjoint_table_is_the_transpose_of_z_expand(coor_full):
    """The grouped weight's z adjoint table lists, per coarse slice, exactly the nonzero column entries of
    z_expand (padded with weight 0), so E^T computed from it is the transpose of the expand; it is absent
    for non-identity layer maps and for strides whose columns have more than four entries."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    p = build_compton_sparse_projector(torch.from_numpy(coor_full), 1, 2, 20)
    ax = p.tables.axis
    assert ax.l_idx is None and ax.za_idx is not None
    ze = p.z_expand.numpy().astype(np.float64)
    rebuilt = np.zeros_like(ze)
    for zc in range(ax.n_zc):
        for j in range(ops.ZADJ):
            rebuilt[ax.za_idx[zc, j], zc] += ax.za_w[zc, j]
    np.testing.assert_array_equal(rebuilt, ze)
    assert build_compton_sparse_projector(torch.from_numpy(coor_full), 2, 2, 20).tables.axis.za_idx is None
    assert build_compton_sparse_projector(torch.from_numpy(coor_full), 1, 4, 20).tables.axis.za_idx is None
