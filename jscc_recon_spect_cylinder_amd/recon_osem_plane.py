"""Drop-in for the reference's dense OSEM core (recon_osem_plane.py).

Same names, argument order, defaults and return values as
recon_osem_plane.py:8-403 (lpz7777777/JSCC_Recon_SPECT_Cylinder).  Every
A.x / A^T(y/Ax) / T^T(1/Tx) pair, the RotMat / RotMatInv gathers, the alpha
weighting and the EM update run in the fused gfx950 kernels of libjscc_hip:
one pass over each matrix block per call.  The reference's multi-process
(several-GPU) branches are replaced by the single-GPU engine.
"""
from __future__ import annotations

import os
import time

import torch

from . import _osem_core as core
from . import ops


def split_tensor_even(tensor, parts, dim=0):
    """recon_osem_plane.py:8-18 (floor bounds, exactly `parts` pieces)."""
    if parts <= 0:
        raise ValueError("parts must be positive")
    total = tensor.size(dim)
    return [tensor.narrow(dim, total * i // parts, total * (i + 1) // parts - total * i // parts)
            for i in range(parts)]


def get_t_budget_bytes_per_gpu(num_gpus, usage_ratio=0.8):
    """recon_osem_plane.py:21-26."""
    return [max(1, int(torch.cuda.get_device_properties(g).total_memory * usage_ratio)) for g in range(num_gpus)]


def get_weight_single(sysmat_list_tmp, proj_list_tmp, img_tmp):
    """A^T (y / (A x)) without clamping (recon_osem_plane.py:28-30)."""
    dev = core.gpu_device(img_tmp)
    A = core.matrix_on(sysmat_list_tmp, dev)
    y = proj_list_tmp.to(dev).reshape(-1).float()
    w = torch.zeros(A.shape[1], dtype=torch.float32, device=dev)
    core.add_weight(w, A, core.image_vec(img_tmp), y=y)
    ops.raise_on_index_error(dev)
    return w.reshape(-1, 1)


def get_weight_compton(t_block, img):
    """T^T (1 / (T x)) without clamping; host blocks are copied in per call (recon_osem_plane.py:33-42)."""
    dev = core.gpu_device(img)
    T = core.matrix_on(t_block, dev)
    w = torch.zeros(T.shape[1], dtype=torch.float32, device=dev)
    core.add_weight(w, T, core.image_vec(img))
    return w.reshape(-1, 1)


def _bin_subset_weight(w, sysmat_list, proj_list, rotmat_all, rotmat_inv_all, x, rotate_num, dev):
    for i in range(rotate_num):
        for sysmat, proj, rotmat, rotmat_inv in zip(sysmat_list, proj_list, rotmat_all, rotmat_inv_all):
            A = core.matrix_on(sysmat, dev)
            Y = proj.to(dev)
            if Y.dtype != torch.float32:
                Y = Y.float()
            core.add_weight(w, A, x, y=Y[:, i], x_perm=core.perm_on(rotmat, dev)[:, i],
                            inv_perm=core.perm_on(rotmat_inv, dev)[:, i])


def _list_rotation_weight(w, t_list_i, rotmat_all, rotmat_inv_all, x, i, dev, scale=1.0):
    for t in t_list_i:
        for t_tmp, rotmat, rotmat_inv in zip(t, rotmat_all, rotmat_inv_all):
            if t_tmp.numel() == 0:
                continue
            T = core.matrix_on(t_tmp, dev)
            core.add_weight(w, T, x, x_perm=core.perm_on(rotmat, dev)[:, i],
                            inv_perm=core.perm_on(rotmat_inv, dev)[:, i], scale=scale)


def osem_bin_mode(sysmat_list_all, proj_list_all, rotmat_all, rotmat_inv_all, img, s_map, rotate_num,
                  model_denoiser=None):
    """SC / SCD subset updates x <- x * sum_r P_r^-1 A^T(y/A P_r x) / s (recon_osem_plane.py:45-61)."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    w = torch.zeros_like(x)
    for sysmat_list, proj_list in zip(sysmat_list_all, proj_list_all):
        _bin_subset_weight(w, sysmat_list, proj_list, rotmat_all, rotmat_inv_all, x, rotate_num, dev)
        core.em(x, w, s_map, zero_after=w)
        if model_denoiser is not None:
            x = core.image_vec(core.apply_denoiser(x.view(-1, 1), model_denoiser)).clone()
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def osem_list_mode(t_list_all, rotmat_all, rotmat_inv_all, img, s_map, rotate_num, model_denoiser=None):
    """JSCC-D subset updates with T^T(1/Tx) over rotations and T blocks (recon_osem_plane.py:63-80)."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    w = torch.zeros_like(x)
    for t_list in t_list_all:
        for i in range(rotate_num):
            _list_rotation_weight(w, t_list[i], rotmat_all, rotmat_inv_all, x, i, dev)
        core.em(x, w, s_map, zero_after=w)
        if model_denoiser is not None:
            x = core.image_vec(core.apply_denoiser(x.view(-1, 1), model_denoiser)).clone()
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def osem_joint_mode(sysmat_list_all, proj_list_all, t_list_all, rotmat_all, rotmat_inv_all, img, s_map, alpha,
                    rotate_num, model_denoiser=None):
    """JSCC-SD: x <- x * ((2-alpha) w_c + alpha w_s) / s (recon_osem_plane.py:100-125)."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    wc = torch.zeros_like(x)
    ws = torch.zeros_like(x)
    for sysmat_list, proj_list, t_list in zip(sysmat_list_all, proj_list_all, t_list_all):
        for i in range(rotate_num):
            for sysmat, proj, rotmat, rotmat_inv in zip(sysmat_list, proj_list, rotmat_all, rotmat_inv_all):
                A = core.matrix_on(sysmat, dev)
                Y = proj.to(dev)
                if Y.dtype != torch.float32:
                    Y = Y.float()
                core.add_weight(ws, A, x, y=Y[:, i], x_perm=core.perm_on(rotmat, dev)[:, i],
                                inv_perm=core.perm_on(rotmat_inv, dev)[:, i])
            _list_rotation_weight(wc, t_list[i], rotmat_all, rotmat_inv_all, x, i, dev)
        core.em(x, wc, s_map, c1=2.0 - alpha, w2=ws, c2=alpha, zero_after=wc)
        ws.zero_()
        if model_denoiser is not None:
            x = core.image_vec(core.apply_denoiser(x.view(-1, 1), model_denoiser)).clone()
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def save_img(img_sc, img_scd, img_jsccd, img_jsccsd, img_sc_iter, img_scd_iter, img_jsccd_iter, img_jsccsd_iter,
             iter_arg, save_path):
    """Raw float32 image files with the reference's names (recon_osem_plane.py:128-147)."""
    for name, img in (("Image_SC", img_sc), ("Image_SCD", img_scd), ("Image_JSCCD", img_jsccd),
                      ("Image_JSCCSD", img_jsccsd)):
        img.detach().cpu().numpy().astype("float32").tofile(save_path + name)
    step = iter_arg.save_iter_step
    for name, n_iter, stack in (("SC", iter_arg.sc, img_sc_iter), ("SCD", iter_arg.jsccd, img_scd_iter),
                                ("JSCCD", iter_arg.jsccd, img_jsccd_iter),
                                ("JSCCSD", iter_arg.jsccsd, img_jsccsd_iter)):
        stack.detach().cpu().numpy().astype("float32").tofile(save_path + "Image_%s_Iter_%d_%d" % (name, n_iter,
                                                                                                n_iter / step))


def get_gpu_memory_usage(num_gpus):
    """recon_osem_plane.py:150-158."""
    for i in range(num_gpus):
        print(f"cuda:{i}")
        print(f"allocated_GB:{round(torch.cuda.memory_allocated(i) / 1024 ** 3, 2)}")
        print(f"reserved_GB:{round(torch.cuda.memory_reserved(i) / 1024 ** 3, 2)}")


def run_recon_osem(sysmat_all, rotmat_all, rotmat_inv_all, proj_all, proj_d_all, t_all, iter_arg, s_map_arg, alpha,
                   save_path, num_gpus, model_denoiser=None):
    """SC -> SCD -> JSCC-D -> JSCC-SD reconstruction and image files (recon_osem_plane.py:161-403).

    T blocks are placed on the GPU up to 80% of its memory and kept in pinned host
    memory beyond that (:195-243).  num_gpus > 1 runs on the single GPU engine
    (the per-subset weight sums are the same; only their summation order differs)."""
    if not os.path.exists(save_path):
        os.makedirs(save_path)
    if num_gpus != 1:
        print(f"run_recon_osem: num_gpus={num_gpus} -> single-GPU engine on cuda:0")
    dev = torch.device("cuda:0")
    pixel_num = sysmat_all[0].size(1)
    rotate_num = rotmat_all[0].size(1)

    img_sc = torch.ones([pixel_num, 1], dtype=torch.float32, device=dev)
    img_scd = torch.ones([pixel_num, 1], dtype=torch.float32, device=dev)
    img_jsccd = torch.ones([pixel_num, 1], dtype=torch.float32, device=dev)
    img_jsccsd = torch.ones([pixel_num, 1], dtype=torch.float32, device=dev)
    step = iter_arg.save_iter_step
    img_sc_iter = torch.ones([round(iter_arg.sc / step), pixel_num], dtype=torch.float32)
    img_scd_iter = torch.ones([round(iter_arg.jsccd / step), pixel_num], dtype=torch.float32)
    img_jsccd_iter = torch.ones([round(iter_arg.jsccd / step), pixel_num], dtype=torch.float32)
    img_jsccsd_iter = torch.ones([round(iter_arg.jsccsd / step), pixel_num], dtype=torch.float32)

    K, E, D = iter_arg.osem_subset_num, iter_arg.ene_num, iter_arg.t_divide_num
    sysmat_list_all = [[[] for _ in range(E)] for _ in range(K)]
    proj_list_all = [[[] for _ in range(E)] for _ in range(K)]
    proj_d_list_all = [[[] for _ in range(E)] for _ in range(K)]
    t_list_all = [[[[[] for _ in range(E)] for _ in range(D)] for _ in range(rotate_num)] for _ in range(K)]

    cpnum_list = torch.arange(0, proj_all[0].size(dim=0))
    cpnum_list = cpnum_list[torch.randperm(proj_all[0].size(dim=0))]
    cpnum_list = split_tensor_even(cpnum_list, K, dim=0)

    budget = get_t_budget_bytes_per_gpu(1)[0]
    used = 0
    for e, (sysmat, proj, proj_d, t) in enumerate(zip(sysmat_all, proj_all, proj_d_all, t_all)):
        for i in range(rotate_num):
            t_chunks = split_tensor_even(t[i], K, dim=0)
            for j in range(K):
                for k, chunk in enumerate(split_tensor_even(t_chunks[j], D, dim=0)):
                    size = chunk.nelement() * chunk.element_size()
                    if used + size < budget:
                        t_list_all[j][i][k][e] = chunk.to(dev, non_blocking=True)
                        used += size
                    else:
                        t_list_all[j][i][k][e] = chunk.pin_memory()
        rotmat_all[e] = rotmat_all[e].to(dev, non_blocking=True)
        rotmat_inv_all[e] = rotmat_inv_all[e].to(dev, non_blocking=True)
        print(f"T placed on the GPU: {round(100 * used / budget, 2)} % of the T budget")
        for i in range(K):
            sysmat_list_all[i][e] = sysmat[cpnum_list[i], :].to(dev, non_blocking=True)
            proj_list_all[i][e] = proj[cpnum_list[i], :].to(dev, non_blocking=True)
            proj_d_list_all[i][e] = proj_d[cpnum_list[i], :].to(dev, non_blocking=True)
    del sysmat_all, t_all

    s_map_arg.s = s_map_arg.s.to(dev, non_blocking=True)
    s_map_arg.d = s_map_arg.d.to(dev, non_blocking=True)
    s_map_arg.j = alpha * s_map_arg.s + (2 - alpha) * s_map_arg.d
    if model_denoiser is not None:
        model_denoiser.to(dev)

    get_gpu_memory_usage(1)
    time_start = time.time()

    def loop(label, n_iter, fn, img, stack):
        print(f"{label} OSEM starts")
        id_save = 0
        for it in range(n_iter):
            img = fn(img)
            if (it + 1) % step == 0:
                stack[id_save, :] = torch.squeeze(img).cpu()
                id_save += 1
                print(f"{label} Iteration {it + 1} ends, time used: {time.time() - time_start} s")
        print(f"{label} OSEM ends, time used: {time.time() - time_start}")
        return img

    img_sc = loop("Self-Collimation", iter_arg.sc, lambda im: osem_bin_mode(
        sysmat_list_all, proj_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.s, rotate_num, model_denoiser),
        img_sc, img_sc_iter)
    img_scd = loop("SC-D", iter_arg.jsccd, lambda im: osem_bin_mode(
        sysmat_list_all, proj_d_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.s, rotate_num, model_denoiser),
        img_scd, img_scd_iter)
    img_jsccd = loop("JSCC-D", iter_arg.jsccd, lambda im: osem_list_mode(
        t_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.d, rotate_num, model_denoiser), img_jsccd,
        img_jsccd_iter)
    img_jsccsd = loop("JSCC-SD", iter_arg.jsccsd, lambda im: osem_joint_mode(
        sysmat_list_all, proj_list_all, t_list_all, rotmat_all, rotmat_inv_all, im, s_map_arg.j, alpha, rotate_num,
        model_denoiser), img_jsccsd, img_jsccsd_iter)

    save_img(img_sc, img_scd, img_jsccd, img_jsccsd, img_sc_iter, img_scd_iter, img_jsccd_iter, img_jsccsd_iter,
             iter_arg, save_path)
