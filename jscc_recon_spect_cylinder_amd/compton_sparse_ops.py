"""Drop-in for the reference's coarse-grid Compton operators (compton_sparse_ops.py).

`build_compton_sparse_projector` returns a `ComptonSparseProjector` with the
reference's fields and values (compton_sparse_ops.py:11-44, 212-290): the ring
layout inferred from the fine coordinates, per-ring periodic theta
reduce/expand (sample every s-th angle / periodic linear interpolation), the
linear z reduce (2-tap average) / expand (clamped linear interpolation), the
ring stride chosen compatible with the rotation count, and the coarse voxel
coordinates.  The dense operators are built once on the host, as in the
reference; the projector additionally carries their CSR forms (at most 4
taps per output pixel), which the gfx950 kernels apply: the row transforms
(:293-356), packing (:359-377) and materialisation (:380-400) below run in
libjscc_hip.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import ops

RADIUS_TOL_MM = 1e-4   # compton_sparse_ops.py:7
Z_TOL_MM = 1e-4        # :8


@dataclass
class SparseTables:
    """CSR operators of a projector (engine addition; no counterpart in the reference)."""
    reduce: ops.CSR        # fine -> coarse rows  (z_reduce (x) layer_reduce)
    expand: ops.CSR        # coarse -> fine rows  (z_expand (x) layer_expand)
    adjoint: ops.CSR       # fine -> coarse, exact adjoint of expand
    reduce_t: ops.CSR      # coarse -> fine, transpose of reduce (upsample_coarse_weight_to_fine)
    taps: ops.ATaps | None  # the <=2-tap reduce as cone-kernel tables
    axis: ops.AxisTaps | None = None  # the separable expand as per-axis taps (fused coarse-row weight)

    def to(self, device):
        taps = None
        if self.taps is not None:
            taps = ops.ATaps(*(t.to(device) for t in (self.taps.i0, self.taps.i1, self.taps.w0, self.taps.w1)))
        return SparseTables(self.reduce.to(device), self.expand.to(device), self.adjoint.to(device),
                            self.reduce_t.to(device), taps, self.axis.to(device) if self.axis is not None else None)


@dataclass
class ComptonSparseProjector:
    theta_stride: int
    z_stride: int
    ring_strides: tuple
    fine_layer_num: int
    fine_z_num: int
    coarse_layer_num: int
    coarse_z_num: int
    fine_pixel_num: int
    coarse_pixel_num: int
    layer_reduce: torch.Tensor
    z_reduce: torch.Tensor
    layer_expand: torch.Tensor
    z_expand: torch.Tensor
    coor_coarse: torch.Tensor
    tables: SparseTables | None = field(default=None, repr=False)

    def to(self, device):
        return ComptonSparseProjector(
            theta_stride=self.theta_stride, z_stride=self.z_stride, ring_strides=self.ring_strides,
            fine_layer_num=self.fine_layer_num, fine_z_num=self.fine_z_num, coarse_layer_num=self.coarse_layer_num,
            coarse_z_num=self.coarse_z_num, fine_pixel_num=self.fine_pixel_num,
            coarse_pixel_num=self.coarse_pixel_num, layer_reduce=self.layer_reduce.to(device),
            z_reduce=self.z_reduce.to(device), layer_expand=self.layer_expand.to(device),
            z_expand=self.z_expand.to(device), coor_coarse=self.coor_coarse.to(device),
            tables=self.tables.to(device) if self.tables is not None else None)


# ----------------------------------------------------------------------------- layout inference
def _runs(values: torch.Tensor, tol: float):
    """Consecutive runs of values within tol of the run's first value: (firsts, counts)."""
    firsts, counts = [values[0]], [1]
    for i in range(1, values.numel()):
        if torch.abs(values[i] - firsts[-1]) <= tol:
            counts[-1] += 1
        else:
            firsts.append(values[i])
            counts.append(1)
    return torch.stack(firsts), counts


def _infer_polar_layout(coor_full: torch.Tensor):
    """compton_sparse_ops.py:59-92."""
    if coor_full.ndim != 2 or coor_full.size(1) != 3:
        raise ValueError("coor_full must have shape [pixel_num, 3].")
    if coor_full.size(0) == 0:
        raise ValueError("coor_full is empty.")
    z_values, z_counts = _runs(coor_full[:, 2], Z_TOL_MM)
    if len(set(z_counts)) != 1:
        raise ValueError("coor_full does not contain a consistent z-plane layout.")
    fine_layer_num = z_counts[0]
    radii = torch.sqrt(torch.sum(coor_full[:fine_layer_num, :2] ** 2, dim=1))
    ring_lengths, ring_radii = [], []
    i = 0
    while i < fine_layer_num:
        j = i + 1
        while j < fine_layer_num and torch.abs(radii[j] - radii[i]) <= RADIUS_TOL_MM:
            j += 1
        ring_lengths.append(j - i)
        ring_radii.append(torch.mean(radii[i:j]))
        i = j
    return {"z_values": z_values, "fine_z_num": len(z_counts), "fine_layer_num": fine_layer_num,
            "ring_lengths": ring_lengths, "ring_radii": torch.stack(ring_radii)}


# ----------------------------------------------------------------------------- 1-D operators
def _check_div(n, s):
    if n % s != 0:
        raise ValueError(f"fine_num={n} is not divisible by stride={s}.")


def _periodic_reduce(n, s, dtype):
    """Sample every s-th fine angle (compton_sparse_ops.py:95-106)."""
    _check_div(n, s)
    m = torch.zeros((n // s, n), dtype=dtype)
    m[torch.arange(n // s), torch.arange(n // s) * s] = 1.0
    return m


def _linear_reduce(n, s, dtype):
    """2-tap average at fine position (k + 0.5) s - 0.5 (compton_sparse_ops.py:109-129)."""
    _check_div(n, s)
    nc = n // s
    m = torch.zeros((nc, n), dtype=dtype)
    pos = (torch.arange(nc, dtype=dtype) + 0.5) * s - 0.5
    for k in range(nc):
        p = pos[k].item()
        left = max(0, min(n - 1, math.floor(p)))
        right = min(n - 1, left + 1)
        if right == left:
            m[k, left] = 1.0
            continue
        wr = p - left
        m[k, left] += 1.0 - wr
        m[k, right] += wr
    return m


def _periodic_expand(n, s, dtype):
    """Periodic linear interpolation between coarse samples (compton_sparse_ops.py:132-154)."""
    _check_div(n, s)
    nc = n // s
    m = torch.zeros((n, nc), dtype=dtype)
    for f in range(n):
        left = (f // s) % nc
        right = (left + 1) % nc
        lpos = left * s
        rpos = lpos + s
        padj = f
        if right == 0 and f < lpos:
            padj += n
            rpos = n
        wr = (padj - lpos) / max(rpos - lpos, 1e-12)
        m[f, left] += 1.0 - wr
        m[f, right] += wr
    return m


def _linear_expand(n, s, dtype):
    """Clamped linear interpolation from coarse centres (compton_sparse_ops.py:157-185)."""
    _check_div(n, s)
    nc = n // s
    m = torch.zeros((n, nc), dtype=dtype)
    cpos = (torch.arange(nc, dtype=dtype) + 0.5) * s - 0.5
    for f in range(n):
        p = float(f)
        ins = torch.searchsorted(cpos, torch.tensor(p, dtype=dtype)).item()
        if ins <= 0:
            m[f, 0] = 1.0
            continue
        if ins >= nc:
            m[f, nc - 1] = 1.0
            continue
        lp, rp = cpos[ins - 1].item(), cpos[ins].item()
        wr = (p - lp) / max(rp - lp, 1e-12)
        m[f, ins - 1] += 1.0 - wr
        m[f, ins] += wr
    return m


def _choose_compatible_ring_stride(ring_length, rotate_num, target_stride):
    """Largest stride <= target dividing the ring (and the rotation step when the ring
    splits evenly into rotations) (compton_sparse_ops.py:188-209)."""
    if target_stride <= 1:
        return 1
    if rotate_num > 0 and ring_length % rotate_num == 0:
        step = ring_length // rotate_num
        for c in range(target_stride, 0, -1):
            if ring_length % c == 0 and step % c == 0:
                return c
        return 1
    for c in range(target_stride, 0, -1):
        if ring_length % c == 0:
            return c
    return 1


# ----------------------------------------------------------------------------- builder
def _csr_kron(Z: torch.Tensor, L: torch.Tensor) -> ops.CSR:
    import scipy.sparse as sp
    K = sp.kron(sp.csr_matrix(Z.double().numpy()), sp.csr_matrix(L.double().numpy()), format="csr")
    K.eliminate_zeros()
    K.sort_indices()
    return ops.CSR(torch.from_numpy(K.indptr.astype(np.int32)), torch.from_numpy(K.indices.astype(np.int32)),
                   torch.from_numpy(K.data.astype(np.float32)), K.shape[0], K.shape[1])


def _csr_transpose(c: ops.CSR) -> ops.CSR:
    import scipy.sparse as sp
    K = sp.csr_matrix((c.val.double().numpy(), c.col.numpy(), c.ptr.numpy()), shape=(c.n_out, c.n_in)).T.tocsr()
    K.sort_indices()
    return ops.CSR(torch.from_numpy(K.indptr.astype(np.int32)), torch.from_numpy(K.indices.astype(np.int32)),
                   torch.from_numpy(K.data.astype(np.float32)), K.shape[0], K.shape[1])


def _taps_of(reduce: ops.CSR) -> ops.ATaps | None:
    ptr = reduce.ptr.numpy()
    if (np.diff(ptr) > 2).any():
        return None
    n = reduce.n_out
    npad = ops.round_up(max(n, 1), 4)
    i0 = np.zeros(npad, np.int32)
    i1 = np.zeros(npad, np.int32)
    w0 = np.zeros(npad, np.float32)
    w1 = np.zeros(npad, np.float32)
    col, val = reduce.col.numpy(), reduce.val.numpy()
    for j in range(n):
        a, b = ptr[j], ptr[j + 1]
        if b > a:
            i0[j], w0[j] = col[a], val[a]
            i1[j] = col[a]
        if b > a + 1:
            i1[j], w1[j] = col[a + 1], val[a + 1]
    return ops.ATaps(torch.from_numpy(i0), torch.from_numpy(i1), torch.from_numpy(w0), torch.from_numpy(w1))


def build_compton_sparse_projector(coor_full, theta_stride=2, z_stride=2, rotate_num=None, dtype=torch.float32):
    """compton_sparse_ops.py:212-290 (plus the CSR tables used by the kernels)."""
    if theta_stride <= 0 or z_stride <= 0:
        raise ValueError("theta_stride and z_stride must be positive.")
    coor_full = torch.as_tensor(coor_full).detach().cpu()
    layout = _infer_polar_layout(coor_full)
    z_values = layout["z_values"].to(dtype=dtype)
    fine_z_num = layout["fine_z_num"]
    fine_layer_num = layout["fine_layer_num"]
    ring_radii = layout["ring_radii"].to(dtype=dtype)
    if fine_z_num % z_stride != 0:
        raise ValueError(f"fine_z_num={fine_z_num} is not divisible by z_stride={z_stride}.")
    red_blocks, exp_blocks, coarse_parts, strides = [], [], [], []
    for ring_length, ring_radius in zip(layout["ring_lengths"], ring_radii):
        s = _choose_compatible_ring_stride(ring_length, rotate_num, theta_stride) if rotate_num is not None \
            else theta_stride
        red_blocks.append(_periodic_reduce(ring_length, s, dtype))
        exp_blocks.append(_periodic_expand(ring_length, s, dtype))
        strides.append(s)
        theta = torch.arange(ring_length // s, dtype=dtype) * s * (2.0 * math.pi / ring_length)
        coarse_parts.append(torch.stack((ring_radius * torch.cos(theta), ring_radius * torch.sin(theta)), dim=1))
    layer_reduce = torch.block_diag(*red_blocks)
    layer_expand = torch.block_diag(*exp_blocks)
    coarse_layer = torch.cat(coarse_parts, dim=0)
    z_reduce = _linear_reduce(fine_z_num, z_stride, dtype)
    z_expand = _linear_expand(fine_z_num, z_stride, dtype)
    coarse_z = torch.matmul(z_reduce, z_values.unsqueeze(1)).squeeze(1)
    coor_coarse = torch.cat([torch.cat((coarse_layer, torch.full((coarse_layer.size(0), 1), zv.item(), dtype=dtype)),
                                       dim=1) for zv in coarse_z], dim=0)
    reduce = _csr_kron(z_reduce, layer_reduce)
    expand = _csr_kron(z_expand, layer_expand)
    tables = SparseTables(reduce, expand, _csr_transpose(expand), _csr_transpose(reduce), _taps_of(reduce),
                          ops.axis_taps(z_expand, layer_expand))
    return ComptonSparseProjector(
        theta_stride=theta_stride, z_stride=z_stride, ring_strides=tuple(strides), fine_layer_num=fine_layer_num,
        fine_z_num=fine_z_num, coarse_layer_num=coarse_layer.size(0), coarse_z_num=coarse_z.numel(),
        fine_pixel_num=fine_layer_num * fine_z_num, coarse_pixel_num=coarse_layer.size(0) * coarse_z.numel(),
        layer_reduce=layer_reduce, z_reduce=z_reduce, layer_expand=layer_expand, z_expand=z_expand,
        coor_coarse=coor_coarse, tables=tables)


# ----------------------------------------------------------------------------- transforms (GPU)
def _tables(projector, device) -> SparseTables:
    t = projector.tables
    if t is None:
        raise ValueError("projector was not built by this engine's build_compton_sparse_projector")
    if t.reduce.ptr.device != device:
        t = t.to(device)
        projector.tables = t
    return t


def _rows_in(rows, n_cols, what):
    if rows.ndim != 2:
        raise ValueError(f"{what} must have shape [num_rows, pixel_num].")
    if rows.size(1) != n_cols:
        raise ValueError(f"{what} size does not match the projector.")
    ops.require_gpu(rows, what)


def reduce_fine_image_to_coarse(img_fine, projector):
    """compton_sparse_ops.py:293-302."""
    if img_fine.ndim != 2 or img_fine.size(1) != 1:
        raise ValueError("img_fine must have shape [fine_pixel_num, 1].")
    if img_fine.size(0) != projector.fine_pixel_num:
        raise ValueError("img_fine size does not match projector.fine_pixel_num.")
    ops.require_gpu(img_fine, "img_fine")
    return ops.csr_map_rows(img_fine.reshape(1, -1), _tables(projector, img_fine.device).reduce).reshape(-1, 1)


def upsample_coarse_weight_to_fine(weight_coarse, projector):
    """compton_sparse_ops.py:305-314 (the transpose of the reduce)."""
    if weight_coarse.ndim != 2 or weight_coarse.size(1) != 1:
        raise ValueError("weight_coarse must have shape [coarse_pixel_num, 1].")
    if weight_coarse.size(0) != projector.coarse_pixel_num:
        raise ValueError("weight_coarse size does not match projector.coarse_pixel_num.")
    ops.require_gpu(weight_coarse, "weight_coarse")
    t = _tables(projector, weight_coarse.device)
    return ops.csr_map_rows(weight_coarse.reshape(1, -1), t.reduce_t).reshape(-1, 1)


def reduce_fine_rows_to_coarse(rows_fine, projector):
    """compton_sparse_ops.py:317-328."""
    _rows_in(rows_fine, projector.fine_pixel_num, "rows_fine")
    return ops.csr_map_rows(rows_fine, _tables(projector, rows_fine.device).reduce)


def upsample_coarse_rows_to_fine(rows_coarse, projector):
    """compton_sparse_ops.py:331-342."""
    _rows_in(rows_coarse, projector.coarse_pixel_num, "rows_coarse")
    return ops.csr_map_rows(rows_coarse, _tables(projector, rows_coarse.device).expand)


def project_fine_rows_to_coarse_adjoint(rows_fine, projector):
    """compton_sparse_ops.py:345-356."""
    _rows_in(rows_fine, projector.fine_pixel_num, "rows_fine")
    return ops.csr_map_rows(rows_fine, _tables(projector, rows_fine.device).adjoint)


def pack_sparse_event_rows(cpnum1, t_compton):
    """[float(cpnum1), t_c...] rows (compton_sparse_ops.py:359-368)."""
    if cpnum1.ndim != 1:
        raise ValueError("cpnum1 must have shape [num_events].")
    if t_compton.ndim != 2:
        raise ValueError("t_compton must have shape [num_events, coarse_pixel_num].")
    if cpnum1.size(0) != t_compton.size(0):
        raise ValueError("cpnum1 and t_compton must have the same number of rows.")
    return torch.cat((cpnum1.to(dtype=t_compton.dtype).unsqueeze(1), t_compton), dim=1)


def unpack_sparse_event_rows(event_rows):
    """compton_sparse_ops.py:371-377."""
    if event_rows.ndim != 2 or event_rows.size(1) < 2:
        raise ValueError("event_rows must have shape [num_events, 1 + coarse_pixel_num].")
    return torch.round(event_rows[:, 0]).to(torch.long), event_rows[:, 1:]


def materialize_sparse_event_rows_to_fine(event_rows, sysmat, projector, normalize=True):
    """(t_norm, valid): fine rows nan_to_num(expand(t_c) * A[cp]) normalised, and the
    row mask (compton_sparse_ops.py:380-400).  normalize=False returns every row
    unnormalised (used by the grouped list-mode weight)."""
    if event_rows.ndim != 2 or event_rows.size(1) != projector.coarse_pixel_num + 1:
        raise ValueError("event_rows size does not match sparse projector.")
    if sysmat.ndim != 2 or sysmat.size(1) != projector.fine_pixel_num:
        raise ValueError("sysmat size does not match sparse projector.")
    if event_rows.device != sysmat.device:
        raise ValueError("event_rows and sysmat must be on the same device.")
    if projector.coor_coarse.device != event_rows.device:
        raise ValueError("projector must be on the same device as event_rows.")
    ops.require_gpu(event_rows, "event_rows")
    t = _tables(projector, event_rows.device)
    rows, flags = ops.materialize(event_rows, t.expand, sysmat, normalize=normalize)
    valid = flags.bool()
    if not normalize:
        return rows, valid
    idx, cnt = ops.compact(flags)
    n = int(cnt.item())
    return ops.gather_normalize(rows, projector.fine_pixel_num, idx, n, None, 0), valid
