"""End-to-end parity of the reconstruction drivers against the FP64 oracle:
run_recon_osem (recon_osem_plane.py:161-403), run_recon_osem_sparse
(recon_osem_plane_sparse.py:144-348) and run_recon_osem_local_sparse_jsccsd_only
(recon_osem_local_sparse_jsccsd_only.py:111-202), with 20 rotations, bin
subsets drawn as the reference draws them, T split per rotation, subset and
divide, and the image files they write."""
import os
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import layout, osem, sparse
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FIX = os.path.join(GOLDEN, "factors_511keV_RotateNum20.npz")


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def read_img(path, name, P):
    return np.fromfile(os.path.join(path, name), dtype=np.float32).reshape(-1, P)


def blocks_of(rows, K, D, bounds):
    """[subset][divide] row blocks of one rotation's T: `bounds` (split_tensor_even or
    torch.chunk) over the subsets, then over the divides."""
    out = []
    for lo, hi in bounds(rows.shape[0], K):
        sub = rows[lo:hi]
        out.append([sub[a:b] for a, b in bounds(sub.shape[0], D)])
    while len(out) < K:
        out.append([])
    return out


def test_run_recon_osem_end_to_end(tmp_path):
    from jscc_recon_spect_cylinder_amd import recon_osem_plane as rp
    rot = np.load(FIX)["rot_layer0"].astype(np.int64)
    inv = np.stack([np.argsort(rot[:, r], kind="stable") + 1 for r in range(rot.shape[1])], axis=1)
    P, R = rot.shape
    rng = np.random.default_rng(21)
    M, K, D, alpha = 90, 3, 2, 0.5
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.2) + 1e-4).astype(np.float32)
    Y = rng.poisson(20 * rng.random((M, R))).astype(np.float32)
    Yd = rng.poisson(10 * rng.random((M, R))).astype(np.float32)
    T = [(rng.random((7 + r % 4, P)) ** 4 + 1e-6).astype(np.float32) for r in range(R)]
    s = (rng.random(P) + 1.0).astype(np.float32)
    d = (rng.random(P) + 0.5).astype(np.float32)
    it = SimpleNamespace(sc=2, jsccd=2, jsccsd=2, save_iter_step=1, osem_subset_num=K, ene_num=1, t_divide_num=D)
    sm = SimpleNamespace(s=torch.from_numpy(s[:, None]), d=torch.from_numpy(d[:, None]))
    out = str(tmp_path) + "/"
    torch.manual_seed(5)
    rp.run_recon_osem([torch.from_numpy(A)], [torch.from_numpy(rot)], [torch.from_numpy(inv)],
                      [torch.from_numpy(Y)], [torch.from_numpy(Yd)], [[torch.from_numpy(t) for t in T]], it, sm,
                      alpha, out, 1)
    # the same bin subsets (global generator, split_tensor_even) and T blocks
    torch.manual_seed(5)
    perm = torch.arange(M)[torch.randperm(M)].numpy()
    bins = [perm[lo:hi] for lo, hi in osem.split_bounds_even(M, K)]
    A_s, Y_s, Yd_s = [A[b] for b in bins], [Y[b] for b in bins], [Yd[b] for b in bins]
    per_rot = [blocks_of(T[r], K, D, osem.split_bounds_even) for r in range(R)]
    T_s = [[per_rot[r][k] for r in range(R)] for k in range(K)]
    x = np.ones(P)
    stack = []
    for _ in range(2):
        x = osem.osem_bin_mode(A_s, Y_s, rot, inv, x, s)
        stack.append(x)
    assert rel(read_img(out, "Image_SC", P), x) < 1e-5
    assert rel(read_img(out, "Image_SC_Iter_2_2", P), np.stack(stack)) < 1e-5
    x = np.ones(P)
    for _ in range(2):
        x = osem.osem_bin_mode(A_s, Yd_s, rot, inv, x, s)
    assert rel(read_img(out, "Image_SCD", P), x) < 1e-5
    x = np.ones(P)
    for _ in range(2):
        x = osem.osem_list_mode(T_s, rot, inv, x, d)
    assert rel(read_img(out, "Image_JSCCD", P), x) < 1e-5
    x = np.ones(P)
    for _ in range(2):
        x = osem.osem_joint_mode(A_s, Y_s, T_s, rot, inv, x, alpha * s + (2 - alpha) * d, alpha)
    assert rel(read_img(out, "Image_JSCCSD", P), x) < 1e-5


@pytest.fixture(scope="module")
def ref_grid():
    coor, rot, inv = layout.reference_grid("511keV_RotateNum20")
    coor = coor.astype(np.float32)
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    p = build_compton_sparse_projector(torch.from_numpy(coor), 1, 2, 20)
    return coor, rot, inv, p, sparse.build(coor, 1, 2, 20)


def packed_rows(rng, n, Pc, M):
    cp = rng.integers(1, M + 1, n).astype(np.float32)
    tc = (rng.random((n, Pc)) ** 3).astype(np.float32)
    return np.hstack([cp[:, None], tc / tc.sum(1, keepdims=True)])


def grouped_weight(ev, A, x_rot, pj):
    """nan_to_num(T^T (1 / max(T x, 1e-12))) with the unnormalised fine rows (recon_osem_plane_sparse.py:22-57)."""
    if ev.shape[0] == 0:
        return np.zeros(x_rot.shape[0])
    cp = np.clip(np.round(ev[:, 0]).astype(np.int64) - 1, 0, A.shape[0] - 1)
    rows = sparse.upsample_rows(ev[:, 1:], pj) * np.asarray(A, np.float64)[cp]
    return osem.nan_to_num(osem.weight_compton(rows, x_rot, eps=1e-12))


def materialised_weight(ev, A, x_rot, pj):
    """recon_osem_local_sparse_jsccsd_only.py:13-19 on materialised, normalised rows."""
    rows, _ = sparse.materialize(ev, A, pj)
    if rows.shape[0] == 0:
        return np.zeros(x_rot.shape[0])
    return osem.nan_to_num(osem.weight_compton(rows, x_rot, eps=1e-12))


def test_run_recon_osem_sparse_end_to_end(tmp_path, ref_grid):
    from jscc_recon_spect_cylinder_amd import recon_osem_plane_sparse as rps
    coor, rot, inv, p, pj = ref_grid
    P, R = rot.shape
    Pc = p.coarse_pixel_num
    rng = np.random.default_rng(22)
    M, K, D, alpha = 40, 2, 2, 0.5
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.3) + 1e-3).astype(np.float32)
    Y = rng.poisson(20 * rng.random((M, R))).astype(np.float32)
    Yd = rng.poisson(10 * rng.random((M, R))).astype(np.float32)
    T = [packed_rows(rng, 5 + r % 3, Pc, M) for r in range(R)]
    s = (rng.random(P) + 1.0).astype(np.float32)
    d = (rng.random(P) + 0.5).astype(np.float32)
    it = SimpleNamespace(sc=1, jsccd=1, jsccsd=1, save_iter_step=1, osem_subset_num=K, ene_num=1, t_divide_num=D)
    sm = SimpleNamespace(s=torch.from_numpy(s[:, None]), d=torch.from_numpy(d[:, None]))
    out = str(tmp_path) + "/"
    torch.manual_seed(6)
    rps.run_recon_osem_sparse([torch.from_numpy(A)], [torch.from_numpy(rot)], [torch.from_numpy(inv)],
                              [torch.from_numpy(Y)], [torch.from_numpy(Yd)], [[torch.from_numpy(t) for t in T]], [p],
                              it, sm, alpha, out, 1)
    torch.manual_seed(6)
    perm = torch.arange(M)[torch.randperm(M)].numpy()
    bins = [perm[lo:hi] for lo, hi in osem.split_bounds_even(M, K)]
    per_rot = [blocks_of(T[r], K, D, osem.split_bounds_even) for r in range(R)]

    def compton(x, k):
        w = np.zeros(P)
        for r in range(R):
            xr = osem.rot_gather(x, rot[:, r])
            for blk in per_rot[r][k]:
                w += osem.rot_gather(grouped_weight(blk, A, xr, pj), inv[:, r])
        return w

    def single(x, k):
        w = np.zeros(P)
        for r in range(R):
            wt = osem.weight_single(A[bins[k]], Y[bins[k], r], osem.rot_gather(x, rot[:, r]))
            w += osem.rot_gather(wt, inv[:, r])
        return w

    x = np.ones(P)
    for k in range(K):
        x = osem.em_safe(x, compton(x, k), d)
    assert rel(read_img(out, "Image_JSCCD", P), x) < 1e-5
    x = np.ones(P)
    sj = alpha * s + (2 - alpha) * d
    for k in range(K):
        x = osem.em_safe(x, (2 - alpha) * compton(x, k) + alpha * single(x, k), sj)
    assert rel(read_img(out, "Image_JSCCSD", P), x) < 1e-5


def test_run_recon_osem_local_sparse_jsccsd_only_end_to_end(tmp_path, ref_grid):
    from jscc_recon_spect_cylinder_amd import recon_osem_local_sparse_jsccsd_only as loc
    coor, rot, inv, p, pj = ref_grid
    P, R = rot.shape
    Pc = p.coarse_pixel_num
    rng = np.random.default_rng(23)
    M, K, D, alpha = 40, 2, 2, 0.6
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.3) + 1e-3).astype(np.float32)
    Y = rng.poisson(20 * rng.random((M, R))).astype(np.float32)
    T = [packed_rows(rng, 5 + r % 3, Pc, M) for r in range(R)]
    T[3][1, 1:] = 0.0  # an all-zero row: dropped by the materialisation
    sj = (rng.random(P) + 1.0).astype(np.float32)
    it = SimpleNamespace(jsccsd=2, save_iter_step=1, osem_subset_num=K, ene_num=1, t_divide_num=D, seed=77)
    sm = SimpleNamespace(j=torch.from_numpy(sj[:, None]))
    out = str(tmp_path) + "/"
    loc.run_recon_osem_local_sparse_jsccsd_only([torch.from_numpy(A)], [torch.from_numpy(A)], [torch.from_numpy(rot)],
                                                [torch.from_numpy(inv)], [torch.from_numpy(Y)],
                                                [[torch.from_numpy(t) for t in T]], [p], it, sm, alpha, out,
                                                torch.device(DEV))
    gen = torch.Generator(device=DEV)
    gen.manual_seed(77)
    per_rot = [blocks_of(T[r], K, D, osem.chunk_bounds) for r in range(R)]
    x = np.ones(P)
    for _ in range(2):
        perm = torch.randperm(M, generator=gen, device=DEV).cpu().numpy()
        for k, (lo, hi) in enumerate(osem.chunk_bounds(M, K)):
            ids = perm[lo:hi]
            w = np.zeros(P)
            for r in range(R):
                xr = osem.rot_gather(x, rot[:, r])
                w += osem.rot_gather(alpha * osem.weight_single(A[ids], Y[ids, r], xr, eps=1e-12), inv[:, r])
            for r in range(R):
                xr = osem.rot_gather(x, rot[:, r])
                for blk in per_rot[r][k]:
                    w += osem.rot_gather((2 - alpha) * materialised_weight(blk, A, xr, pj), inv[:, r])
            x = osem.em_safe(x, w, sj)
    assert rel(read_img(out, "Image_JSCCSD", P), x) < 1e-5
