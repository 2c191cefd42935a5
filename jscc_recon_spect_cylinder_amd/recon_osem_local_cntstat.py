"""Drop-in for the reference's local SC OSEM (recon_osem_local_cntstat.py).

Same names and signatures as recon_osem_local_cntstat.py:6-100.  The random
bin subsets are drawn with the same torch generator calls (seeded randperm +
torch.chunk, :11-25); inside `run_recon_osem_local_cntstat` the subsets are
handed to the fused projection kernel as row-index lists, so the system
matrix is never copied per iteration.
"""
from __future__ import annotations

import time

import torch

from . import _osem_core as core
from . import ops


def get_weight_single(sysmat, proj, img_rotate):
    """A^T (y / max(A x, 1e-12)) (recon_osem_local_cntstat.py:6-8)."""
    dev = core.gpu_device(img_rotate)
    A = core.matrix_on(sysmat, dev)
    y = proj.to(dev).reshape(-1).float()
    w = torch.zeros(A.shape[1], dtype=torch.float32, device=dev)
    core.add_weight(w, A, core.image_vec(img_rotate), y=y, eps=1e-12)
    return w.reshape(-1, 1)


def build_random_bin_subsets(sysmat_all_device, proj_all_device, subset_num, generator):
    """Seeded random bin subsets (recon_osem_local_cntstat.py:11-25)."""
    local_bin_num = proj_all_device[0].size(0)
    cpnum_list = torch.randperm(local_bin_num, generator=generator, device=proj_all_device[0].device)
    chunks = list(torch.chunk(cpnum_list, subset_num, dim=0))
    sysmat_list_all = [[None for _ in range(len(sysmat_all_device))] for _ in range(subset_num)]
    proj_list_all = [[None for _ in range(len(proj_all_device))] for _ in range(subset_num)]
    for e in range(len(sysmat_all_device)):
        for k in range(subset_num):
            ids = chunks[k] if k < len(chunks) else cpnum_list[:0]
            sysmat_list_all[k][e] = sysmat_all_device[e][ids, :]
            proj_list_all[k][e] = proj_all_device[e][ids, :]
    return sysmat_list_all, proj_list_all


def osem_bin_mode_local(sysmat_list_all, proj_list_all, rotmat_all, rotmat_inv_all, img, s_map):
    """recon_osem_local_cntstat.py:28-45 (clamped weights, empty subsets skipped)."""
    dev = core.gpu_device(img)
    rotate_num = rotmat_all[0].size(1)
    x = core.image_vec(img).clone()
    w = torch.zeros_like(x)
    for sysmat_list, proj_list in zip(sysmat_list_all, proj_list_all):
        for r in range(rotate_num):
            for sysmat, proj, rotmat, rotmat_inv in zip(sysmat_list, proj_list, rotmat_all, rotmat_inv_all):
                if sysmat.size(0) == 0:
                    continue
                A = core.matrix_on(sysmat, dev)
                Y = proj.to(dev).float()
                core.add_weight(w, A, x, y=Y[:, r], eps=1e-12, x_perm=core.perm_on(rotmat, dev)[:, r],
                                inv_perm=core.perm_on(rotmat_inv, dev)[:, r])
        core.em(x, w, s_map, zero_after=w)
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def osem_bin_mode_local_rows(sysmat_all, proj_all, subset_rows, rotmat_all, rotmat_inv_all, x, w, s_map):
    """osem_bin_mode_local over subsets given as int32 row lists of the full matrices
    (no per-iteration copy of the system matrix); x and w are updated in place."""
    dev = x.device
    rotate_num = rotmat_all[0].size(1)
    for rows in subset_rows:
        n = rows.numel()
        for r in range(rotate_num):
            for A, Y, rotmat, rotmat_inv in zip(sysmat_all, proj_all, rotmat_all, rotmat_inv_all):
                if n == 0:
                    continue
                core.add_weight(w, A, x, y=Y[:, r], y_by_row=True, rows=rows, n_rows=n, eps=1e-12,
                                x_perm=core.perm_on(rotmat, dev)[:, r], inv_perm=core.perm_on(rotmat_inv, dev)[:, r])
        core.em(x, w, s_map, zero_after=w)


def save_img_local_cntstat(img_sc, img_sc_iter, iter_arg, save_path):
    """recon_osem_local_cntstat.py:48-53."""
    img_sc.detach().cpu().numpy().astype("float32").tofile(save_path + "Image_SC")
    img_sc_iter.detach().cpu().numpy().astype("float32").tofile(
        save_path + "Image_SC_Iter_%d_%d" % (iter_arg.sc, iter_arg.sc / iter_arg.save_iter_step))
    print("Images saved to local disk.")


def run_recon_osem_local_cntstat(sysmat_all, rotmat_all, rotmat_inv_all, proj_all, iter_arg, s_map_arg, save_path,
                                 device, *, save=True):
    """recon_osem_local_cntstat.py:56-100; returns the final image (the reference
    only writes it to disk)."""
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("run_recon_osem_local_cntstat: this engine runs on a GPU device only")
    pixel_num = s_map_arg.s.size(0)
    x = torch.ones(pixel_num, dtype=torch.float32, device=device)
    w = torch.zeros_like(x)
    img_sc_iter = torch.ones([round(iter_arg.sc / iter_arg.save_iter_step), pixel_num], dtype=torch.float32)

    generator = torch.Generator(device=device if device.type == "cuda" else "cpu")
    generator.manual_seed(int(getattr(iter_arg, "seed", 20260331)))

    sysmat_dev = [core.matrix_on(s, device) for s in sysmat_all]
    proj_dev = [p.to(device).float() for p in proj_all]
    rot_dev = [core.perm_on(r, device) for r in rotmat_all]
    inv_dev = [core.perm_on(r, device) for r in rotmat_inv_all]
    s_map = s_map_arg.s.to(device)

    print("\n" + "=" * 50)
    print("Starting local single-photon reconstruction")
    print("=" * 50)
    time_start = time.time()
    id_save = 0
    for it in range(iter_arg.sc):
        perm = torch.randperm(proj_dev[0].size(0), generator=generator, device=proj_dev[0].device)
        chunks = list(torch.chunk(perm, iter_arg.osem_subset_num, dim=0))
        rows = [c.to(torch.int32) for c in chunks]
        osem_bin_mode_local_rows(sysmat_dev, proj_dev, rows, rot_dev, inv_dev, x, w, s_map)
        if (it + 1) % iter_arg.save_iter_step == 0:
            img_sc_iter[id_save, :] = x.detach().cpu()
            id_save += 1
            print(f"[SC] Iter: {it + 1}/{iter_arg.sc} | Time: {time.time() - time_start:.2f}s")
    ops.raise_on_index_error(device)
    img_sc = x.view(-1, 1)
    if save:
        save_img_local_cntstat(img_sc, img_sc_iter, iter_arg, save_path)
    return img_sc
