"""Drop-in for the reference's sparse (coarse-T) OSEM core (recon_osem_plane_sparse.py).

Same names and signatures as recon_osem_plane_sparse.py:22-348.  The grouped
list-mode weight (:22-57) equals T_f^T (1 / max(T_f x, 1e-12)) with the
unnormalised fine rows T_f = expand(t_c) * A[cpnum1] followed by nan_to_num;
here one gfx950 kernel streams the packed coarse rows of a block, expands each
row with index arithmetic, applies its A-row factor and accumulates the ratio
back-projection (jscc_coarse_weight), instead of a Python loop over first-hit
crystals.  Grids beyond that kernel's limits materialise the fine rows first
and run the fused projection over them.  safe_em_update is the kernel's mode 1.
"""
from __future__ import annotations

import os
import time

import torch

from . import _osem_core as core
from . import ops
from .compton_sparse_ops import _tables, materialize_sparse_event_rows_to_fine
from .recon_osem_plane import (get_gpu_memory_usage, get_t_budget_bytes_per_gpu, get_weight_single,  # noqa: F401
                               osem_bin_mode, save_img, split_tensor_even)


def _sparse_block_weight(w, event_block, sysmat_full, x, sparse_projector, *, x_perm=None, inv_perm=None,
                         scale=1.0):
    dev = x.device
    if event_block.numel() == 0 or event_block.size(0) == 0:
        return
    rows_blk = event_block.to(dev, non_blocking=True)
    A = sysmat_full if sysmat_full.device == dev else sysmat_full.to(dev, non_blocking=True)
    proj = sparse_projector if sparse_projector.coor_coarse.device == dev else sparse_projector.to(dev)
    axis = _tables(proj, dev).axis
    if axis is not None and ops.coarse_weight_fits(axis):
        # one pass over the packed rows: expand, A-row factor and the ratio back-projection fused; taken in
        # first-hit crystal order, the rows of one crystal share its A row (the grouping of :39-51)
        ops.coarse_weight(rows_blk, axis, A, x, w, normalize=False, eps=1e-12, x_perm=x_perm, inv_perm=inv_perm,
                          out_scale=scale, accumulate=True, rows=ops.crystal_order(rows_blk))
        return
    t_fine, _ = materialize_sparse_event_rows_to_fine(rows_blk, A.float(), proj, normalize=False)
    core.add_weight(w, ops.proj_ready(t_fine, ops.DT_F32), x, eps=1e-12, x_perm=x_perm, inv_perm=inv_perm,
                    scale=scale, sanitize=True)


def get_weight_compton_sparse(event_block, sysmat_full, img_fine, sparse_projector):
    """recon_osem_plane_sparse.py:22-57."""
    dev = core.gpu_device(img_fine)
    x = core.image_vec(img_fine)
    w = torch.zeros_like(x)
    _sparse_block_weight(w, event_block, sysmat_full, x, sparse_projector)
    return w.view(-1, 1)


def safe_em_update(img, weight, s_map, eps=1e-12):
    """recon_osem_plane_sparse.py:60-66."""
    core.gpu_device(img)
    x = core.image_vec(img).clone()
    ops.em_update(x, core.image_vec(weight), core.image_vec(s_map), mode=ops.EM_SAFE, eps=eps)
    return x.view_as(img) if img.is_contiguous() else x.view(-1, 1)


def osem_list_mode_sparse(t_list_all, sysmat_all, sparse_projector_all, rotmat_all, rotmat_inv_all, img, s_map,
                          rotate_num, model_denoiser=None):
    """recon_osem_plane_sparse.py:69-88."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    w = torch.zeros_like(x)
    for t_list in t_list_all:
        for r in range(rotate_num):
            for t in t_list[r]:
                for t_tmp, sysmat_full, rotmat, rotmat_inv, sp in zip(t, sysmat_all, rotmat_all, rotmat_inv_all,
                                                                      sparse_projector_all):
                    _sparse_block_weight(w, t_tmp, sysmat_full, x, sp, x_perm=core.perm_on(rotmat, dev)[:, r],
                                         inv_perm=core.perm_on(rotmat_inv, dev)[:, r])
        core.em(x, w, s_map, mode=ops.EM_SAFE, zero_after=w)
        if model_denoiser is not None:
            x = core.image_vec(core.apply_denoiser(x.view(-1, 1), model_denoiser)).clone()
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def osem_joint_mode_sparse(sysmat_list_all, proj_list_all, t_list_all, sysmat_all, sparse_projector_all, rotmat_all,
                           rotmat_inv_all, img, s_map, alpha, rotate_num, model_denoiser=None):
    """recon_osem_plane_sparse.py:91-118."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    wc = torch.zeros_like(x)
    ws = torch.zeros_like(x)
    for sysmat_list, proj_list, t_list in zip(sysmat_list_all, proj_list_all, t_list_all):
        for r in range(rotate_num):
            for sysmat, proj, rotmat, rotmat_inv in zip(sysmat_list, proj_list, rotmat_all, rotmat_inv_all):
                Y = proj.to(dev).float()
                core.add_weight(ws, core.matrix_on(sysmat, dev), x, y=Y[:, r], x_perm=core.perm_on(rotmat, dev)[:, r],
                                inv_perm=core.perm_on(rotmat_inv, dev)[:, r])
            for t in t_list[r]:
                for t_tmp, sysmat_full, rotmat, rotmat_inv, sp in zip(t, sysmat_all, rotmat_all, rotmat_inv_all,
                                                                      sparse_projector_all):
                    _sparse_block_weight(wc, t_tmp, sysmat_full, x, sp, x_perm=core.perm_on(rotmat, dev)[:, r],
                                         inv_perm=core.perm_on(rotmat_inv, dev)[:, r])
        core.em(x, wc, s_map, mode=ops.EM_SAFE, c1=2.0 - alpha, w2=ws, c2=alpha, zero_after=wc)
        ws.zero_()
        if model_denoiser is not None:
            x = core.image_vec(core.apply_denoiser(x.view(-1, 1), model_denoiser)).clone()
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def run_recon_osem_sparse(sysmat_all, rotmat_all, rotmat_inv_all, proj_all, proj_d_all, t_all, sparse_projector_all,
                          iter_arg, s_map_arg, alpha, save_path, num_gpus, model_denoiser=None):
    """recon_osem_plane_sparse.py:144-348 on the single-GPU engine."""
    if not os.path.exists(save_path):
        os.makedirs(save_path)
    if num_gpus != 1:
        print(f"run_recon_osem_sparse: num_gpus={num_gpus} -> single-GPU engine on cuda:0")
    dev = torch.device("cuda:0")
    pixel_num = sysmat_all[0].size(1)
    rotate_num = rotmat_all[0].size(1)
    step = iter_arg.save_iter_step
    imgs = {k: torch.ones([pixel_num, 1], dtype=torch.float32, device=dev) for k in ("sc", "scd", "jsccd", "jsccsd")}
    stacks = {"sc": torch.ones([round(iter_arg.sc / step), pixel_num]),
              "scd": torch.ones([round(iter_arg.jsccd / step), pixel_num]),
              "jsccd": torch.ones([round(iter_arg.jsccd / step), pixel_num]),
              "jsccsd": torch.ones([round(iter_arg.jsccsd / step), pixel_num])}
    K, E, D = iter_arg.osem_subset_num, iter_arg.ene_num, iter_arg.t_divide_num
    sysmat_list_all = [[[] for _ in range(E)] for _ in range(K)]
    proj_list_all = [[[] for _ in range(E)] for _ in range(K)]
    proj_d_list_all = [[[] for _ in range(E)] for _ in range(K)]
    t_list_all = [[[[[] for _ in range(E)] for _ in range(D)] for _ in range(rotate_num)] for _ in range(K)]
    cpnum_list = torch.arange(0, proj_all[0].size(dim=0))
    cpnum_list = split_tensor_even(cpnum_list[torch.randperm(proj_all[0].size(dim=0))], K, dim=0)
    budget = get_t_budget_bytes_per_gpu(1)[0]
    used = 0
    for e, (sysmat, proj, proj_d, t) in enumerate(zip(sysmat_all, proj_all, proj_d_all, t_all)):
        for r in range(rotate_num):
            for j, chunk_j in enumerate(split_tensor_even(t[r], K, dim=0)):
                for k, chunk in enumerate(split_tensor_even(chunk_j, D, dim=0)):
                    size = chunk.nelement() * chunk.element_size()
                    if used + size < budget:
                        t_list_all[j][r][k][e] = chunk.to(dev, non_blocking=True)
                        used += size
                    else:
                        t_list_all[j][r][k][e] = chunk.pin_memory()
        rotmat_all[e] = rotmat_all[e].to(dev, non_blocking=True)
        rotmat_inv_all[e] = rotmat_inv_all[e].to(dev, non_blocking=True)
        for i in range(K):
            sysmat_list_all[i][e] = sysmat[cpnum_list[i], :].to(dev, non_blocking=True)
            proj_list_all[i][e] = proj[cpnum_list[i], :].to(dev, non_blocking=True)
            proj_d_list_all[i][e] = proj_d[cpnum_list[i], :].to(dev, non_blocking=True)
    sysmat_all = [s.to(dev, non_blocking=True) for s in sysmat_all]
    del t_all
    s_map_arg.s = s_map_arg.s.to(dev, non_blocking=True)
    s_map_arg.d = s_map_arg.d.to(dev, non_blocking=True)
    s_map_arg.j = alpha * s_map_arg.s + (2 - alpha) * s_map_arg.d
    if model_denoiser is not None:
        model_denoiser.to(dev)
    proj_gpu = [p.to(dev) for p in sparse_projector_all]
    get_gpu_memory_usage(1)
    t0 = time.time()

    def loop(label, key, n_iter, fn):
        print(f"{label} OSEM starts")
        sid = 0
        for it in range(n_iter):
            imgs[key] = fn(imgs[key])
            if (it + 1) % step == 0:
                stacks[key][sid, :] = torch.squeeze(imgs[key]).cpu()
                sid += 1
                print(f"{label} Iteration {it + 1} ends, time used: {time.time() - t0} s")

    loop("Self-Collimation", "sc", iter_arg.sc, lambda im: osem_bin_mode(
        sysmat_list_all, proj_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.s, rotate_num, model_denoiser))
    loop("SC-D", "scd", iter_arg.jsccd, lambda im: osem_bin_mode(
        sysmat_list_all, proj_d_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.s, rotate_num, model_denoiser))
    loop("JSCC-D Sparse", "jsccd", iter_arg.jsccd, lambda im: osem_list_mode_sparse(
        t_list_all, sysmat_all, proj_gpu, rotmat_all, rotmat_inv_all, im, s_map_arg.d, rotate_num, model_denoiser))
    loop("JSCC-SD Sparse", "jsccsd", iter_arg.jsccsd, lambda im: osem_joint_mode_sparse(
        sysmat_list_all, proj_list_all, t_list_all, sysmat_all, proj_gpu, rotmat_all, rotmat_inv_all, im,
        s_map_arg.j, alpha, rotate_num, model_denoiser))
    save_img(imgs["sc"], imgs["scd"], imgs["jsccd"], imgs["jsccsd"], stacks["sc"], stacks["scd"], stacks["jsccd"],
             stacks["jsccsd"], iter_arg, save_path)
