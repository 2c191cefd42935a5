"""Drop-in for the reference's local sparse JSCCSD OSEM
(recon_osem_local_sparse_jsccsd_only.py).

Same names and signatures as recon_osem_local_sparse_jsccsd_only.py:8-202.
Per T block and iteration the coarse rows are expanded on the fine grid
(normalised, invalid rows dropped) and T^T(1/max(Tx, 1e-12)) is formed with
nan_to_num (:13-19) in one gfx950 pass over the packed rows
(jscc_coarse_weight; grids beyond its limits materialise the fine rows and run
the fused projection); the system-matrix term uses the clamped weight, each
term is scaled by alpha or 2-alpha before accumulation (:58-100) and the
update is safe_em_update.  Invalid rows are kept in place with a zero row
scale instead of being compacted: they contribute nothing, as in the
reference.
"""
from __future__ import annotations

import time

import torch

from . import _osem_core as core
from . import ops
from .compton_sparse_ops import _tables
from .recon_osem_local_cntstat import build_random_bin_subsets  # noqa: F401  (same helper, :40-55)
from .recon_osem_local_cntstat import get_weight_single  # noqa: F401  (clamped, :8-10)
from .recon_osem_plane_sparse import safe_em_update  # noqa: F401  (:22-28)


def summarize_image_tensor(img):
    """recon_osem_local_sparse_jsccsd_only.py:31-37."""
    c = img.detach().float().cpu()
    return (f"min={c.min().item():.6e} max={c.max().item():.6e} mean={c.mean().item():.6e} "
            f"sum={c.sum().item():.6e} zero={int((c == 0).sum().item())}/{c.numel()}")


def _materialized_weight(w, event_block, sysmat_full, x, projector, *, x_perm=None, inv_perm=None, scale=1.0):
    if event_block.numel() == 0 or event_block.size(0) == 0:
        return
    dev = x.device
    t = _tables(projector, dev)
    if t.axis is not None and ops.coarse_weight_fits(t.axis):
        # materialise + normalise + ratio back-projection in one pass over the packed rows, taken in first-hit
        # crystal order so that the rows of one crystal share its A row
        blk = event_block.to(dev)
        ops.coarse_weight(blk, t.axis, sysmat_full, x, w, normalize=True, eps=1e-12, x_perm=x_perm,
                          inv_perm=inv_perm, out_scale=scale, accumulate=True, rows=ops.crystal_order(blk))
        return
    rows, flags = ops.materialize(event_block.to(dev), t.expand, sysmat_full, normalize=True)
    row_scale = flags.to(torch.float32)
    core.add_weight(w, ops.proj_ready(rows, ops.DT_F32), x, eps=1e-12, row_scale=row_scale, x_perm=x_perm,
                    inv_perm=inv_perm, scale=scale, sanitize=True)


def get_weight_compton_sparse(event_block, sysmat_full, img_rotate, sparse_projector):
    """recon_osem_local_sparse_jsccsd_only.py:13-19."""
    core.gpu_device(img_rotate)
    x = core.image_vec(img_rotate)
    w = torch.zeros_like(x)
    _materialized_weight(w, event_block, sysmat_full.float(), x, sparse_projector)
    return w.view(-1, 1)


def osem_joint_mode_local_sparse(sysmat_l_all, proj_l_all, t_l_all, sysmat_full_all, sparse_projector_all,
                                 rotmat_all, rotmat_inv_all, img, s_map, alpha, rotate_num, device):
    """recon_osem_local_sparse_jsccsd_only.py:58-100."""
    dev = core.gpu_device(img)
    x = core.image_vec(img).clone()
    w = torch.zeros_like(x)
    for sysmat_l, proj_l, t_l in zip(sysmat_l_all, proj_l_all, t_l_all):
        for r in range(rotate_num):
            for sysmat, proj, rotmat, rotmat_inv in zip(sysmat_l, proj_l, rotmat_all, rotmat_inv_all):
                if sysmat.size(0) == 0:
                    continue
                core.add_weight(w, core.matrix_on(sysmat, dev), x, y=proj.to(dev).float()[:, r], eps=1e-12,
                                x_perm=core.perm_on(rotmat, dev)[:, r], inv_perm=core.perm_on(rotmat_inv, dev)[:, r],
                                scale=alpha)
        for r in range(rotate_num):
            for t_rotate in t_l[r]:
                for t_block, sysmat_full, rotmat, rotmat_inv, sp in zip(t_rotate, sysmat_full_all, rotmat_all,
                                                                        rotmat_inv_all, sparse_projector_all):
                    _materialized_weight(w, t_block, sysmat_full, x, sp, x_perm=core.perm_on(rotmat, dev)[:, r],
                                         inv_perm=core.perm_on(rotmat_inv, dev)[:, r], scale=2.0 - alpha)
        core.em(x, w, s_map, mode=ops.EM_SAFE, zero_after=w)
    ops.raise_on_index_error(dev)
    return x.view(-1, 1)


def save_img_local_sparse_jsccsd_only(img_jsccsd, img_jsccsd_iter, iter_arg, save_path):
    """recon_osem_local_sparse_jsccsd_only.py:103-108."""
    img_jsccsd.detach().cpu().numpy().astype("float32").tofile(save_path + "Image_JSCCSD")
    img_jsccsd_iter.detach().cpu().numpy().astype("float32").tofile(
        save_path + "Image_JSCCSD_Iter_%d_%d" % (iter_arg.jsccsd, iter_arg.jsccsd / iter_arg.save_iter_step))
    print("Images saved to local disk.")


def run_recon_osem_local_sparse_jsccsd_only(sysmat_all, sysmat_full_all, rotmat_all, rotmat_inv_all, proj_all,
                                            t_local_all, sparse_projector_all, iter_arg, s_map_arg, alpha, save_path,
                                            device, *, save=True):
    """recon_osem_local_sparse_jsccsd_only.py:111-202; returns the final image."""
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("run_recon_osem_local_sparse_jsccsd_only: this engine runs on a GPU device only")
    pixel_num = s_map_arg.j.size(0)
    rotate_num = rotmat_all[0].size(1)
    img = torch.ones([pixel_num, 1], dtype=torch.float32, device=device)
    stack = torch.ones([round(iter_arg.jsccsd / iter_arg.save_iter_step), pixel_num], dtype=torch.float32)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(getattr(iter_arg, "seed", 20260331)))
    sysmat_dev = [core.matrix_on(s, device) for s in sysmat_all]
    sysmat_full_dev = [ops.aligned_rows(s, 1, device) for s in sysmat_full_all]
    proj_dev = [p.to(device).float() for p in proj_all]
    rot_dev = [core.perm_on(r, device) for r in rotmat_all]
    inv_dev = [core.perm_on(r, device) for r in rotmat_inv_all]
    proj_gpu = [p.to(device) for p in sparse_projector_all]
    s_map = s_map_arg.j.to(device)
    empty = torch.zeros((0, proj_gpu[0].coarse_pixel_num + 1), dtype=torch.float32, device=device)
    K, D, E = iter_arg.osem_subset_num, iter_arg.t_divide_num, iter_arg.ene_num
    t_list_all = [[[[empty for _ in range(E)] for _ in range(D)] for _ in range(rotate_num)] for _ in range(K)]
    for e in range(E):
        for r in range(rotate_num):
            for k, chunk in enumerate(torch.chunk(t_local_all[e][r], K, dim=0)):
                for d, blk in enumerate(torch.chunk(chunk, D, dim=0)):
                    t_list_all[k][r][d][e] = blk.to(device)
    print("\n" + "=" * 50)
    print("Starting local sparse JSCCSD-only reconstruction")
    print("=" * 50)
    t0 = time.time()
    sid = 0
    for it in range(iter_arg.jsccsd):
        sl, pl = build_random_bin_subsets(sysmat_dev, proj_dev, K, gen)
        img = osem_joint_mode_local_sparse(sl, pl, t_list_all, sysmat_full_dev, proj_gpu, rot_dev, inv_dev, img,
                                           s_map, alpha, rotate_num, device)
        if (it + 1) % iter_arg.save_iter_step == 0:
            stack[sid, :] = img.squeeze().detach().cpu()
            sid += 1
            print(f"[JSCCSD-Sparse-Only] Iter: {it + 1}/{iter_arg.jsccsd} | Time: {time.time() - t0:.2f}s | "
                  f"{summarize_image_tensor(img)}")
    if save:
        save_img_local_sparse_jsccsd_only(img, stack, iter_arg, save_path)
    return img
