"""Tensor-level wrappers of the libjscc_hip entry points.

PyTorch provides device memory, streams and host plumbing only; every
numerical step runs in the gfx950 kernels.  All tensors handed to these
functions must live on a GPU; workspaces are allocated here (through the
torch caching allocator) so the library itself never allocates.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

DT_F32, DT_BF16, DT_FP8, DT_FP16 = 0, 1, 2, 3
SIG_STRICT, SIG_FIXED = 0, 1
EVENT_RECORD_FLOATS = 16  # 64-byte per-event record

_DTYPE_CODE = {torch.float32: DT_F32, torch.bfloat16: DT_BF16, torch.float8_e4m3fn: DT_FP8, torch.float16: DT_FP16}
_CODE_TORCH = {DT_F32: torch.float32, DT_BF16: torch.bfloat16, DT_FP8: torch.uint8, DT_FP16: torch.float16}


def dtype_code(dtype) -> int:
    if isinstance(dtype, int):
        return dtype
    if isinstance(dtype, str):
        return {"fp32": DT_F32, "f32": DT_F32, "float32": DT_F32, "bf16": DT_BF16, "bfloat16": DT_BF16,
                "fp8": DT_FP8, "e4m3": DT_FP8, "fp16": DT_FP16, "float16": DT_FP16, "half": DT_FP16}[dtype.lower()]
    return _DTYPE_CODE[dtype]


def require_gpu(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a tensor on a GPU (this engine has no CPU path)")


def _p(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def round_up(n: int, m: int) -> int:
    return (n + m - 1) // m * m


# ----------------------------------------------------------------------------- error word
_err_words: dict[int, torch.Tensor] = {}

_ERR_TEXT = {1: "crystal number out of range of the detector table",
             2: "image gather index (RotMat) out of range",
             4: "weight gather index (RotMatInv) out of range",
             8: "sensitivity gather index (RotMatInv) out of range",
             16: "thinning position out of range of the expanded counts",
             32: "grid barrier of the persistent OSEM kernel timed out (results invalid)",
             64: "expand tap table index out of range of the coarse grid"}


def err_word(device: torch.device) -> torch.Tensor:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    w = _err_words.get(idx)
    if w is None:
        w = torch.zeros(1, dtype=torch.int32, device=torch.device("cuda", idx))
        _err_words[idx] = w
    return w


def raise_on_index_error(device: torch.device) -> None:
    """Synchronising check of the device error word (raises IndexError)."""
    w = err_word(device)
    code = int(w.item())
    if code:
        w.zero_()
        msgs = [txt for bit, txt in _ERR_TEXT.items() if code & bit]
        raise IndexError("; ".join(msgs))


def read_count_checked(count: torch.Tensor) -> int:
    """The device int32 count of a compaction read back together with the
    device error word in one synchronisation (raises IndexError as
    raise_on_index_error)."""
    w = err_word(count.device)
    count_v, code = torch.cat((count.view(1), w)).tolist()
    if code:
        w.zero_()
        msgs = [txt for bit, txt in _ERR_TEXT.items() if code & bit]
        raise IndexError("; ".join(msgs))
    return int(count_v)


# ----------------------------------------------------------------------------- geometry caches
@dataclass
class VoxelSoA:
    x: torch.Tensor
    y: torch.Tensor
    z: torch.Tensor
    n: int


class SourceCache:
    """Results derived from a source tensor, reused only for that same tensor object at
    the same version.  Each entry holds a reference to its source, so the source's memory
    cannot be freed and handed to another tensor while the entry lives (a key made of the
    data pointer alone would then return a stale result)."""

    def __init__(self, capacity: int):
        self.capacity = capacity
        self.entries: dict = {}

    def get(self, src: torch.Tensor, extra: tuple = ()):
        e = self.entries.get((id(src),) + extra)
        if e is not None and e[0] is src and e[1] == src._version:
            return e[2]
        return None

    def put(self, src: torch.Tensor, extra: tuple, value):
        if len(self.entries) >= self.capacity:
            self.entries.clear()
        self.entries[(id(src),) + extra] = (src, src._version, value)
        return value


_voxel_cache = SourceCache(16)


def voxel_soa(coor: torch.Tensor, device: torch.device) -> VoxelSoA:
    """[P,3] voxel centres -> three FP32 arrays padded to a multiple of 4 on `device`."""
    hit = _voxel_cache.get(coor, (str(device),))
    if hit is not None:
        return hit
    if coor.ndim != 2 or coor.shape[1] < 3:
        raise ValueError("coor must have shape [pixel_num, 3]")
    if not bool(torch.isfinite(coor[:, :3]).all()):
        # (the reference's rows would all hold NaN and be dropped)
        raise ValueError("voxel coordinates must be finite")
    P = coor.shape[0]
    Pp = round_up(max(P, 1), 4)
    c = coor[:, :3].to(device=device, dtype=torch.float32)
    soa = torch.zeros((3, Pp), dtype=torch.float32, device=device)
    soa[:, :P] = c.t()
    return _voxel_cache.put(coor, (str(device),), VoxelSoA(soa[0], soa[1], soa[2], P))


_det_cache = SourceCache(16)


def detector_layers(detector: torch.Tensor) -> tuple[int, float]:
    """(four |y| layers?, rear-layer |y|) of a crystal table (process_list_plane_strict.py:59-71)."""
    hit = _det_cache.get(detector)
    if hit is not None:
        return hit
    y = np.abs(detector[:, 1].detach().to("cpu", torch.float32).numpy())
    layers = np.unique(y)
    return _det_cache.put(detector, (), (1, float(layers[-1])) if layers.size == 4 else (0, 0.0))


_aligned_cache = SourceCache(2)


def aligned_rows(A: torch.Tensor, vec: int = 4, device: torch.device | None = None) -> torch.Tensor:
    """Row-major FP32 copy/view of A on `device` whose row stride is a multiple of `vec`
    and whose base is 16-byte aligned.

    The reference loads the system matrix as a transposed view (main_plane.py:174)
    and passes it to every chunk; such inputs (or host tensors) are copied once into
    a padded row-major device buffer that later calls on the same tensor reuse.
    """
    dev = device if device is not None else A.device
    if A.device == dev and A.dtype == torch.float32 and A.stride(1) == 1 and A.stride(0) % vec == 0 \
            and A.stride(0) >= A.shape[1] and A.data_ptr() % 16 == 0:
        return A
    hit = _aligned_cache.get(A, (str(dev), vec))
    if hit is not None:
        return hit
    M, P = A.shape
    buf = torch.zeros((M, round_up(P, vec)), dtype=torch.float32, device=dev)
    buf[:, :P] = A.to(dev)
    return _aligned_cache.put(A, (str(dev), vec), buf[:, :P])


# ----------------------------------------------------------------------------- events
def events_prepare(list_dev: torch.Tensor, noise1, noise2, writeback: bool, detector: torch.Tensor,
                   delta_r1: float, delta_r2: float, e0: float, ene_resolution: float,
                   thr_max: float, thr_min: float, thr_sum: float, n_sysmat_rows: int = 0):
    """Per-event records [n,16] (FP32 view of 64-byte structs) and keep flags [n] int32."""
    require_gpu(list_dev, "list_origin")
    dev = list_dev.device
    n = list_dev.shape[0]
    if list_dev.dtype != torch.float32 or list_dev.stride(1) != 1:
        raise ValueError("list rows must be contiguous float32")
    for nz in (noise1, noise2):
        if nz is not None and (nz.numel() < n or nz.dtype != torch.float32 or not nz.is_contiguous()):
            raise ValueError("noise tensors must be contiguous float32 with one value per event")
    det = detector
    if det.dtype != torch.float32 or det.stride(1) != 1:
        det = det.float().contiguous()
    det = det.to(dev)
    layers4, rear = detector_layers(det)
    events = torch.empty((max(n, 1), EVENT_RECORD_FLOATS), dtype=torch.float32, device=dev)
    flags = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    err = err_word(dev)
    f32 = lambda v: float(np.float32(v))
    rc = _lib.lib().jscc_events_prepare(
        list_dev.data_ptr(), n, list_dev.stride(0), _p(noise1), _p(noise2),
        list_dev.data_ptr() if writeback else None, det.data_ptr(), det.shape[0], det.stride(0),
        int(n_sysmat_rows), layers4, f32(rear), f32(delta_r1), f32(delta_r2),
        f32(e0), f32(ene_resolution), f32(thr_max), f32(thr_min), f32(thr_sum), float(e0), float(ene_resolution),
        events.data_ptr(), flags.data_ptr(), err.data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_events_prepare")
    return events, flags[:n]


def compact(flags: torch.Tensor):
    """Order-preserving indices of the set flags: (idx int32 [n] (valid prefix), count int32 [1] device)."""
    dev = flags.device
    n = flags.shape[0]
    idx = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    count = torch.empty(1, dtype=torch.int32, device=dev)
    ws = torch.empty(max(_lib.lib().jscc_compact_workspace_bytes(n), 4), dtype=torch.uint8, device=dev)
    rc = _lib.lib().jscc_compact_flags(flags.data_ptr(), n, idx.data_ptr(), count.data_ptr(), ws.data_ptr(),
                                       ws.numel(), stream_of(dev))
    _lib.check(rc, "jscc_compact_flags")
    return idx, count


# ----------------------------------------------------------------------------- cone rows
@dataclass
class ATaps:
    """First-hit factor on coarse voxels: w0*A[row, i0] + w1*A[row, i1]."""
    i0: torch.Tensor
    i1: torch.Tensor
    w0: torch.Tensor
    w1: torch.Tensor


def cone_rows(events: torch.Tensor, ev_idx: torch.Tensor, n_rows: int, vox: VoxelSoA, sysmat=None,
              taps: ATaps | None = None, sigma_model: int = SIG_STRICT, fixed_sigma: float = 0.0,
              out_dtype=None, want_tc: bool = False, ld: int | None = None, n_dev: torch.Tensor | None = None):
    """Evaluate cone rows; returns (t_store | None, t_c | None, stats [n,4] f64, ld).  n_dev: the row count
    as a device int32 (jscc_cone_rows_counted), n_rows then its upper bound: rows past the count get NaN
    statistics (dropped by rows_keep) and no values."""
    if n_dev is not None and (n_dev.dtype != torch.int32 or n_dev.device != events.device or n_rows < 1
                              or ev_idx.numel() < n_rows):
        raise ValueError("the row count must be a device int32 and the event list hold n_rows entries")
    dev = events.device
    P = vox.n
    ld = ld or round_up(P, 64)
    out_t = None
    code = -1
    if out_dtype is not None:
        code = dtype_code(out_dtype)
        if code not in (DT_F32, DT_BF16):
            raise ValueError("cone rows are stored as FP32 or BF16 (FP8 is a re-quantisation)")
        out_t = torch.empty((max(n_rows, 1), ld), dtype=_CODE_TORCH[code], device=dev)
    out_tc = torch.empty((max(n_rows, 1), ld), dtype=torch.float32, device=dev) if want_tc else None
    stats = torch.empty((max(n_rows, 1), 4), dtype=torch.float64, device=dev)
    A = None
    lda = 0
    if sysmat is not None:
        A = aligned_rows(sysmat) if taps is None else sysmat
        lda = A.stride(0)
    tail = (vox.x.data_ptr(), vox.y.data_ptr(), vox.z.data_ptr(), P, _p(A), lda,
            _p(taps.i0) if taps else None, _p(taps.i1) if taps else None,
            _p(taps.w0) if taps else None, _p(taps.w1) if taps else None,
            int(sigma_model), float(fixed_sigma), _p(out_t), code, ld, _p(out_tc), ld, stats.data_ptr(), stream_of(dev))
    if n_dev is None:
        rc = _lib.lib().jscc_cone_rows(events.data_ptr(), ev_idx.data_ptr(), n_rows, *tail)
        _lib.check(rc, "jscc_cone_rows")
    else:
        rc = _lib.lib().jscc_cone_rows_counted(events.data_ptr(), ev_idx.data_ptr(), n_dev.data_ptr(), n_rows, *tail)
        _lib.check(rc, "jscc_cone_rows_counted")
    return out_t, out_tc, stats, ld


def cone_rows_sens(events: torch.Tensor, ev_idx: torch.Tensor, n_rows: int, vox: VoxelSoA, min_support: float,
                   sysmat=None, sigma_model: int = SIG_STRICT, fixed_sigma: float = 0.0, out_dtype=torch.bfloat16,
                   ld: int | None = None, n_dev: torch.Tensor | None = None):
    """cone_rows (FP32, BF16 or FP16-with-row-scale rows, no t_c, no tap tables) that also returns the list-mode
    sensitivity: the column sums of the kept rows divided by their sums, formed from each row as stored while it
    is still in cache (jscc_cone_rows_sens).  FP16 rows are the codes of quantize_rows(..., "fp16") of the FP32
    rows (dropped rows: zeros) with their row scales.  Returns (t_store, stats [n,4] f64, ld, sens [P] f32,
    row_scale [n] f32 or None)."""
    if n_dev is not None and (n_dev.dtype != torch.int32 or n_dev.device != events.device or n_rows < 1
                              or ev_idx.numel() < n_rows):
        raise ValueError("the row count must be a device int32 and the event list hold n_rows entries")
    dev = events.device
    P = vox.n
    ld = ld or round_up(P, 64)
    code = dtype_code(out_dtype)
    if code not in (DT_F32, DT_BF16, DT_FP16):
        raise ValueError("cone rows with the sensitivity are stored as FP32, BF16 or FP16")
    out_t = torch.empty((max(n_rows, 1), ld), dtype=_CODE_TORCH[code], device=dev)
    stats = torch.empty((max(n_rows, 1), 4), dtype=torch.float64, device=dev)
    sens = torch.empty(P, dtype=torch.float32, device=dev)
    row_scale = torch.empty(max(n_rows, 1), dtype=torch.float32, device=dev) if code == DT_FP16 else None
    A = aligned_rows(sysmat) if sysmat is not None else None
    ws_bytes = int(_lib.lib().jscc_cone_sens_workspace_bytes(P, code))
    ws = _workspace(dev, ws_bytes)
    rc = _lib.lib().jscc_cone_rows_sens(events.data_ptr(), ev_idx.data_ptr(), _p(n_dev), n_rows, vox.x.data_ptr(),
                                        vox.y.data_ptr(), vox.z.data_ptr(), P, _p(A), A.stride(0) if A is not None else 0,
                                        int(sigma_model), float(fixed_sigma), out_t.data_ptr(), code, ld,
                                        _p(row_scale), stats.data_ptr(), float(min_support), sens.data_ptr(),
                                        ws.data_ptr(), ws_bytes, stream_of(dev))
    _lib.check(rc, "jscc_cone_rows_sens")
    return out_t, stats, ld, sens, row_scale


def read_counts_checked(*counts: torch.Tensor) -> list:
    """Several device int32 counts read back together with the device error word in one synchronisation
    (raises IndexError as raise_on_index_error)."""
    w = err_word(counts[0].device)
    vals = torch.cat([c.view(1) for c in counts] + [w]).tolist()
    if vals[-1]:
        w.zero_()
        msgs = [txt for bit, txt in _ERR_TEXT.items() if vals[-1] & bit]
        raise IndexError("; ".join(msgs))
    return [int(v) for v in vals[:-1]]


def rows_keep(stats: torch.Tensor, n: int, min_support: float, want_inv_sum: bool = False):
    """Keep flags [n] int32 (and 1/S per row when want_inv_sum)."""
    flags = torch.empty(max(n, 1), dtype=torch.int32, device=stats.device)
    inv = torch.empty(max(n, 1), dtype=torch.float32, device=stats.device) if want_inv_sum else None
    rc = _lib.lib().jscc_rows_keep_flags(stats.data_ptr(), n, float(min_support), flags.data_ptr(), _p(inv),
                                         stream_of(stats.device))
    _lib.check(rc, "jscc_rows_keep_flags")
    return (flags[:n], inv) if want_inv_sum else flags[:n]


def gather_normalize(src: torch.Tensor, P: int, idx: torch.Tensor | None, n_out: int, stats, field: int,
                     lead: torch.Tensor | None = None) -> torch.Tensor:
    """Compacted, row-normalised FP32 rows [n_out, P] (or [n_out, 1+P] with a leading column)."""
    dev = src.device
    off = 1 if lead is not None else 0
    dst = torch.empty((n_out, P + off), dtype=torch.float32, device=dev)
    if n_out == 0:
        return dst
    code = DT_F32 if src.dtype == torch.float32 else DT_BF16
    rc = _lib.lib().jscc_rows_gather_normalize(src.data_ptr(), code, src.stride(0), P, _p(idx), n_out, _p(stats),
                                               field, _p(lead), dst.data_ptr(), dst.stride(0), off, stream_of(dev))
    _lib.check(rc, "jscc_rows_gather_normalize")
    return dst


def arows_normalize(A: torch.Tensor, arow: torch.Tensor, n_out: int) -> torch.Tensor:
    dev = A.device
    P = A.shape[1]
    dst = torch.empty((n_out, P), dtype=torch.float32, device=dev)
    if n_out == 0:
        return dst
    rc = _lib.lib().jscc_arows_normalize(A.data_ptr(), A.stride(0), P, arow.data_ptr(), n_out, dst.data_ptr(),
                                         dst.stride(0), stream_of(dev))
    _lib.check(rc, "jscc_arows_normalize")
    return dst


def quantize_rows(src: torch.Tensor, P: int, n_rows: int, stats: torch.Tensor | None, dtype="fp8"):
    """FP8 e4m3fn (uint8 codes) or FP16 copy of the rows plus per-row scales
    (scale * code ~ t / S from the row statistics; stats None: scale * code ~ row)."""
    dev = src.device
    code_out = dtype_code(dtype)
    if code_out not in (DT_FP8, DT_FP16):
        raise ValueError("quantised storage is FP8 or FP16")
    ld = round_up(P, 64)
    dst = torch.empty((max(n_rows, 1), ld), dtype=_CODE_TORCH[code_out], device=dev)
    scale = torch.empty(max(n_rows, 1), dtype=torch.float32, device=dev)
    code = DT_F32 if src.dtype == torch.float32 else DT_BF16
    rc = _lib.lib().jscc_quantize_rows(src.data_ptr(), code, src.stride(0), P, n_rows, _p(stats), dst.data_ptr(),
                                       code_out, ld, scale.data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_quantize_rows")
    return dst, scale


def quantize_fp8(src: torch.Tensor, P: int, n_rows: int, stats: torch.Tensor | None):
    """FP8 e4m3fn copy of the rows plus per-row scales (scale * code ~ t / S; stats None: scale * code ~ row)."""
    return quantize_rows(src, P, n_rows, stats, "fp8")


# ----------------------------------------------------------------------------- CSR row maps
@dataclass
class CSR:
    ptr: torch.Tensor
    col: torch.Tensor
    val: torch.Tensor
    n_out: int
    n_in: int

    def to(self, device):
        return CSR(self.ptr.to(device), self.col.to(device), self.val.to(device), self.n_out, self.n_in)


def csr_from_dense(M: np.ndarray, tol: float = 0.0) -> CSR:
    """CSR (on CPU) of a dense operator M [n_out, n_in]: out = M @ in."""
    M = np.asarray(M, dtype=np.float64)
    nz = np.abs(M) > tol
    counts = nz.sum(axis=1)
    ptr = np.zeros(M.shape[0] + 1, dtype=np.int32)
    ptr[1:] = np.cumsum(counts)
    rows, cols = np.nonzero(nz)
    return CSR(torch.from_numpy(ptr), torch.from_numpy(cols.astype(np.int32)),
               torch.from_numpy(M[rows, cols].astype(np.float32)), M.shape[0], M.shape[1])


def csr_map_rows(inp: torch.Tensor, csr: CSR) -> torch.Tensor:
    """out[r, :] = csr @ inp[r, :] for every row."""
    require_gpu(inp, "rows")
    if inp.dtype != torch.float32 or inp.stride(-1) != 1:
        inp = inp.float().contiguous()
    n = inp.shape[0]
    out = torch.empty((n, csr.n_out), dtype=torch.float32, device=inp.device)
    rc = _lib.lib().jscc_rows_csr_map(inp.data_ptr(), inp.stride(0), n, csr.ptr.data_ptr(), csr.col.data_ptr(),
                                      csr.val.data_ptr(), csr.n_out, out.data_ptr(), out.stride(0),
                                      stream_of(inp.device))
    _lib.check(rc, "jscc_rows_csr_map")
    return out


def materialize(packed: torch.Tensor, expand: CSR, sysmat: torch.Tensor, normalize: bool = True):
    """Fine rows of packed coarse rows: (rows [n, P_f] (normalised where valid), flags int32 [n])."""
    dev = packed.device
    n = packed.shape[0]
    Pf = expand.n_out
    if packed.dtype != torch.float32 or packed.stride(1) != 1:
        packed = packed.float().contiguous()
    A = sysmat if (sysmat.dtype == torch.float32 and sysmat.stride(1) == 1) else sysmat.float().contiguous()
    out = torch.empty((max(n, 1), Pf), dtype=torch.float32, device=dev)
    flags = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    rc = _lib.lib().jscc_materialize_rows(packed.data_ptr(), packed.stride(0), n, expand.n_in, expand.ptr.data_ptr(),
                                          expand.col.data_ptr(), expand.val.data_ptr(), Pf, A.data_ptr(),
                                          A.stride(0), A.shape[0], out.data_ptr(), out.stride(0),
                                          1 if normalize else 0, flags.data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_materialize_rows")
    return out[:n], flags[:n]


@dataclass
class AxisTaps:
    """z_expand (x) layer_expand as at most two (index, weight) taps per fine z
    slice and per fine layer pixel: the form jscc_coarse_weight expands coarse
    rows with.  l_idx / l_w are None for an identity layer map."""
    z_idx: torch.Tensor          # int32 [n_zf, 2] coarse z indices
    z_w: torch.Tensor            # float32 [n_zf, 2]
    n_zc: int
    l_idx: torch.Tensor | None   # int32 [n_lf, 2] coarse layer indices
    l_w: torch.Tensor | None     # float32 [n_lf, 2]
    n_lf: int
    n_lc: int
    za_idx: torch.Tensor | None = None  # int32 [n_zc, ZADJ]: fine z slices with a tap on each coarse slice
    za_w: torch.Tensor | None = None    # float32 [n_zc, ZADJ]: their weights (padding: weight 0)

    @property
    def n_zf(self) -> int:
        return self.z_idx.shape[0]

    @property
    def n_fine(self) -> int:
        return self.n_zf * self.n_lf

    @property
    def n_coarse(self) -> int:
        return self.n_zc * self.n_lc

    def to(self, device):
        mv = lambda t: None if t is None else t.to(device)  # noqa: E731
        return AxisTaps(mv(self.z_idx), mv(self.z_w), self.n_zc, mv(self.l_idx), mv(self.l_w), self.n_lf, self.n_lc,
                        mv(self.za_idx), mv(self.za_w))


def _two_taps(m: np.ndarray, what: str):
    """(idx [n, 2] int32, w [n, 2] float32) of a dense interpolation matrix whose rows
    hold one or two nonzeros (a single tap repeats its index with weight 0)."""
    n = m.shape[0]
    idx = np.zeros((n, 2), np.int32)
    w = np.zeros((n, 2), np.float32)
    for f in range(n):
        nz = np.flatnonzero(m[f])
        if nz.size == 0 or nz.size > 2:
            raise ValueError(f"{what} row {f} has {nz.size} taps (1 or 2 expected)")
        idx[f, 0], w[f, 0] = nz[0], m[f, nz[0]]
        idx[f, 1] = nz[-1]
        if nz.size == 2:
            w[f, 1] = m[f, nz[1]]
    return idx, w


ZADJ = 4  # fine z slices per coarse slice in the grouped weight's adjoint table (linear expand, z stride 2)


def _adjoint_taps(m: np.ndarray):
    """(idx [n_coarse, ZADJ] int32, w [n_coarse, ZADJ] float32): the nonzeros of each column of a dense
    interpolation matrix m [fine, coarse] (the fine indices with a tap on that coarse index), padded with
    weight 0; None when a column has more than ZADJ nonzeros."""
    n_c = m.shape[1]
    idx = np.zeros((n_c, ZADJ), np.int32)
    w = np.zeros((n_c, ZADJ), np.float32)
    for c in range(n_c):
        nz = np.flatnonzero(m[:, c])
        if nz.size > ZADJ:
            return None
        idx[c, :nz.size] = nz
        w[c, :nz.size] = m[nz, c]
    return idx, w


def axis_taps(z_expand, layer_expand) -> AxisTaps:
    """Tap tables of the separable expand kron(z_expand, layer_expand) (dense [fine, coarse]), and for an
    identity layer map the z adjoint table of the grouped weight."""
    ze = np.asarray(torch.as_tensor(z_expand).detach().cpu().float().numpy())
    le = np.asarray(torch.as_tensor(layer_expand).detach().cpu().float().numpy())
    zi, zw = _two_taps(ze, "z_expand")
    identity = le.shape[0] == le.shape[1] and np.array_equal(le, np.eye(le.shape[0], dtype=le.dtype))
    li = lw = None
    za = None
    if not identity:
        a, b = _two_taps(le, "layer_expand")
        li, lw = torch.from_numpy(a), torch.from_numpy(b)
    else:
        za = _adjoint_taps(ze)
    return AxisTaps(torch.from_numpy(zi), torch.from_numpy(zw), ze.shape[1], li, lw, le.shape[0], le.shape[1],
                    None if za is None else torch.from_numpy(za[0]), None if za is None else torch.from_numpy(za[1]))


_order_cache = SourceCache(64)


def crystal_order(packed: torch.Tensor) -> torch.Tensor:
    """int32 order of packed rows [cpnum1, t_c] by first-hit crystal (stable), cached per rows tensor and
    version: the row order that gives jscc_coarse_weight's grouped form its longest runs."""
    hit = _order_cache.get(packed)
    if hit is not None:
        return hit
    order = torch.argsort(torch.round(packed[:, 0]), stable=True).to(torch.int32).contiguous()
    return _order_cache.put(packed, (), order)


# Rows per run of one first-hit crystal above which jscc_coarse_weight's grouped form beats its per-row form.
# Grids the prefetching per-row kernel takes (identity layer map, 256 | layer size, <= 26624 fine pixels: the
# reference grids) read each A row once per run in both forms; measured on the 25600-voxel grid (100000 rows):
# 50 rows per run 1.453 (grouped) vs 1.177 ms, 100: 1.228 vs 1.140, 250: 1.126 vs 1.137, 500: 1.106 vs 1.142;
# 250000 rows at 23.8 per run 3.99 vs 2.85 ms.  Other grids: the first per-row kernel reloads the A row of
# every row (25000 rows: 2.6 rows per run 1661 vs 618 us per launch, 25: 452 vs 546, 250: 327 vs 537).
GROUPED_MIN_RUN = 16.0
GROUPED_MIN_RUN_PF = 250.0
_PF_MAX_K4 = 13


def coarse_pf_fits(taps: "AxisTaps") -> bool:
    """Whether jscc_coarse_weight's per-row form is the prefetching kernel on this grid (the grid conditions;
    the row count must also leave the LDS room for the work-group's crystal table, about 14000 rows per
    work-group on the reference grid)."""
    P, Pc = taps.n_fine, taps.n_coarse
    k4 = -(-P // 2048)
    kv = 2 if k4 <= 2 else 4 if k4 <= 4 else 8 if k4 <= 8 else _PF_MAX_K4
    return (taps.l_idx is None and taps.n_lf % 256 == 0 and k4 <= _PF_MAX_K4 and Pc <= (kv + 1) // 2 * 2048 and
            (2 * round_up(Pc, 256) + 72 + 4 * 8 * _PF_MAX_K4) * 4 <= 160 * 1024)


def grouped_min_run(taps: "AxisTaps") -> float:
    """The crystal run length from which coarse_weight(grouped="auto") takes the grouped form on this grid."""
    return GROUPED_MIN_RUN_PF if coarse_pf_fits(taps) else GROUPED_MIN_RUN

_runs_cache = SourceCache(64)


def crystal_run_length(packed: torch.Tensor, rows: torch.Tensor | None = None) -> float:
    """Mean number of consecutive rows (in the launch order: `rows`, or storage order) that share a first-hit
    crystal, cached per rows tensor and version (one synchronisation when first computed)."""
    key = (id(rows) if rows is not None else None,)
    hit = _runs_cache.get(packed, key)
    if hit is not None and (rows is None or hit[1] is rows):
        return hit[0]
    if packed.shape[0] == 0:
        return 0.0
    cp = torch.round(packed[:, 0] if rows is None else packed[rows.long(), 0])
    runs = 1 + int((cp[1:] != cp[:-1]).sum().item())
    return _runs_cache.put(packed, key, (packed.shape[0] / runs, rows))[0]


_COARSE_MAX_FINE = 32 * 1024
_COARSE_MAX_COARSE = 16 * 1024
_COARSE_LDS = 160 * 1024


def coarse_weight_fits(taps: AxisTaps) -> bool:
    """Whether jscc_coarse_weight takes this grid: fine pixels <= 32768 (a multiple
    of 4), coarse pixels <= 16384, and both coarse row buffers plus the tap tables
    within the 160 KB of LDS of a work-group."""
    P, Pc = taps.n_fine, taps.n_coarse
    lds = 2 * round_up(Pc, 4) * 4 + 16 * taps.n_zf + (0 if taps.l_idx is None else 16 * taps.n_lf) + 3 * 3 * 16 * 4
    return P <= _COARSE_MAX_FINE and P % 4 == 0 and Pc <= _COARSE_MAX_COARSE and lds <= _COARSE_LDS


def coarse_weight(packed: torch.Tensor, taps: AxisTaps, sysmat: torch.Tensor, x: torch.Tensor, w: torch.Tensor, *,
                  normalize: bool, eps: float | None = 1e-12, x_perm: torch.Tensor | None = None,
                  inv_perm: torch.Tensor | None = None, out_scale: float = 1.0, accumulate: bool = True,
                  rows: torch.Tensor | None = None, grouped: bool | str = "auto") -> None:
    """w (+)= out_scale * nan_to_num(T^T (1 / max(T x_rot, eps))) for packed coarse rows
    [cpnum1, t_c] expanded on the fly (jscc_coarse_weight, include/jscc_hip.h).  rows: optional int32
    order of the packed rows (rows ordered by first-hit crystal let the grouped form share each A row
    over a run of rows; the weight is a sum over rows, so the order changes only rounding).
    grouped: True for the grouped form where the grid allows it, False for the per-row form, "auto" for the
    grouped form when runs of one crystal average grouped_min_run(taps) rows or more."""
    dev = packed.device
    if packed.ndim != 2 or packed.shape[1] != taps.n_coarse + 1:
        raise ValueError("packed rows must have shape [n, 1 + coarse_pixel_num]")
    if packed.dtype != torch.float32 or packed.stride(1) != 1:
        packed = packed.float().contiguous()
    A = sysmat if (sysmat.dtype == torch.float32 and sysmat.stride(1) == 1) else sysmat.float().contiguous()
    if A.ndim != 2 or A.shape[1] != taps.n_fine:
        raise ValueError("sysmat columns must equal the fine pixel count")
    if x_perm is None and x.shape[0] != taps.n_fine:
        raise ValueError(f"image has {x.shape[0]} pixels, the fine grid {taps.n_fine}")
    if x_perm is not None and x_perm.shape[0] < taps.n_fine:
        raise ValueError(f"RotMat column has {x_perm.shape[0]} entries for {taps.n_fine} pixels")
    if inv_perm is not None and inv_perm.shape[0] < w.shape[0]:
        raise ValueError("RotMatInv column shorter than the weight vector")
    for t in (A, x, w, x_perm, inv_perm, taps.z_idx, taps.za_idx, rows):
        if t is not None and t.device != dev:
            raise ValueError("all operands must be on the rows' device")
    if rows is not None and (rows.dtype != torch.int32 or rows.numel() != packed.shape[0] or
                             not rows.is_contiguous()):
        raise ValueError("rows must be a contiguous int32 order of the packed rows")
    if x.dtype != torch.float32 or not x.is_contiguous() or w.dtype != torch.float32 or not w.is_contiguous():
        raise ValueError("image and weight must be contiguous float32")
    n = packed.shape[0]
    ws = _workspace(dev, _lib.lib().jscc_coarse_weight_workspace_bytes(max(n, 1), taps.n_fine))
    if grouped == "auto":
        grouped = taps.za_idx is not None and n > 0 and crystal_run_length(packed, rows) >= grouped_min_run(taps)
    za_i, za_w = (taps.za_idx, taps.za_w) if grouped else (None, None)
    rc = _lib.lib().jscc_coarse_weight(
        packed.data_ptr() if n else None, packed.stride(0), _p(rows), n, taps.z_idx.data_ptr(), taps.z_w.data_ptr(),
        taps.n_zf, taps.n_zc, _p(za_i), _p(za_w), ZADJ if za_i is not None else 0, _p(taps.l_idx), _p(taps.l_w),
        taps.n_lf, taps.n_lc, A.data_ptr(), A.stride(0),
        A.shape[0], x.data_ptr(), _p(x_perm), x_perm.stride(0) if x_perm is not None else 0, x.shape[0],
        float(eps) if eps is not None else 0.0, 1 if eps is not None else 0, 1 if normalize else 0, _p(inv_perm),
        inv_perm.stride(0) if inv_perm is not None else 0, float(out_scale), 1 if accumulate else 0, w.data_ptr(),
        w.shape[0], ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_coarse_weight")


# ----------------------------------------------------------------------------- projections
MODE_RATIO, MODE_FWD, MODE_COEF, MODE_ONES = 0, 1, 2, 3

_ws_cache: dict = {}


def lib_proj_workspace_bytes(n_rows: int, n_cols: int) -> int:
    """Workspace bytes of one projection launch over n_rows rows of n_cols pixels."""
    return int(_lib.lib().jscc_proj_workspace_bytes(max(n_rows, 1), n_cols))


def _workspace(dev: torch.device, nbytes: int) -> torch.Tensor:
    """Scratch of the launches issued on the device's current stream (one buffer per
    stream: launches on different streams may overlap, launches on one stream are ordered)."""
    key = (dev.index if dev.index is not None else torch.cuda.current_device(), stream_of(dev))
    ws = _ws_cache.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        _ws_cache[key] = ws
    return ws


def proj_ready(mat: torch.Tensor, dtype_c: int) -> torch.Tensor:
    """Matrix rows in the layout the projection kernel streams: unit column stride,
    16-byte aligned rows whose stride is a multiple of the vector width.  The
    kernel reads each row in whole vectors, so the rows must also extend (in
    memory) to the column count rounded up to the vector width; a matrix that
    does not -- e.g. a contiguous [M, P] with P not a multiple of it -- is copied
    into zero-padded rows and returned as the [M, P] view of them."""
    vw = _lib.lib().jscc_proj_vec_width(dtype_c)
    M, P = mat.shape
    Pv = round_up(P, vw)
    esz = mat.element_size()
    if (mat.stride(1) == 1 and mat.stride(0) % vw == 0 and mat.data_ptr() % 16 == 0 and
            (M == 0 or (mat.storage_offset() + (M - 1) * mat.stride(0) + Pv) * esz <= mat.untyped_storage().nbytes())):
        return mat
    if P % vw == 0:
        # a fresh allocation is 16-byte aligned; a contiguous but offset view must be copied
        return mat.clone(memory_format=torch.contiguous_format) if mat.is_contiguous() else mat.contiguous()
    padded = torch.zeros((M, Pv), dtype=mat.dtype, device=mat.device)
    padded[:, :P] = mat
    return padded[:, :P]


def proj(mode: int, mat: torch.Tensor, **kw) -> None:
    """Launch jscc_proj on the current stream (see include/jscc_hip.h).

    Keywords: dtype_c, n_cols, rows, n_rows, row_base, row_scale, x, x_perm, y,
    coef, eps, fwd_out, inv_perm, out_scale, accumulate, w, y_by_row, sanitize,
    em, workspace.  em = (x_em, s, em_mode, em_eps, zero_w[, c1, w2, c2])
    (mode MODE_RATIO only) runs jscc_proj_em instead: the EM update of x_em with
    the weight c1*w (+ c2*w2) and sensitivity s is applied in the same reduction
    pass that finishes w.  workspace: a caller-owned buffer of at least
    jscc_proj_workspace_bytes (graph capture); else the per-device cache."""
    name, args, _ = _proj_call(mode, mat, **kw)
    _lib.check(getattr(_lib.lib(), name)(*args), name)


class ProjLauncher:
    """One jscc_proj / jscc_proj_em call validated once and launched many times
    over row ranges of the same operands (OSEM subset loops): the operand
    checks, pointer and workspace lookups of `proj` happen at construction,
    `launch(row_base, n_rows)` only replaces the two integers and takes the
    device's current stream (so that a graph capture records it).  The
    workspace is the caller's, or a buffer the launcher allocates for
    n_rows_max rows and owns for its lifetime; with its own buffer the
    launcher only launches on the stream it was built on (launches on other
    streams could overlap and share that buffer)."""

    def __init__(self, mode: int, mat: torch.Tensor, *, n_rows_max: int, **kw):
        if n_rows_max < 1:
            raise ValueError("n_rows_max must be positive")
        kw = dict(kw)
        kw["n_rows"] = n_rows_max
        kw.setdefault("row_base", 0)
        self._count = kw.get("count")
        self._own_ws = kw.get("workspace") is None
        if self._own_ws:
            P = kw.get("n_cols") if kw.get("n_cols") is not None else mat.shape[1]
            per = n_rows_max if self._count is None else -(-self._count[1] // self._count[2])
            kw["workspace"] = torch.empty(max(lib_proj_workspace_bytes(per, P), 256), dtype=torch.uint8,
                                          device=mat.device)
        self.name, args, (self._i_rows, self._i_base) = _proj_call(mode, mat, **kw)
        self._fn = getattr(_lib.lib(), self.name)
        self._args = list(args)
        self._dev = mat.device
        self._stream = stream_of(mat.device)
        self._n_max = n_rows_max
        self._n_total = kw["rows"].shape[0] if kw.get("rows") is not None else mat.shape[0]
        self._keep = (mat, kw)  # the operands (and the workspace) stay alive as long as the launcher

    def launch(self, row_base: int, n_rows: int) -> None:
        if self._count is not None:
            raise ValueError("a counted launcher takes launch_subset(i)")
        if n_rows < 0 or n_rows > self._n_max or row_base < 0 or row_base + n_rows > self._n_total:
            raise ValueError("row range outside the launcher's matrix or workspace")
        self._go(n_rows, row_base)

    def launch_subset(self, sub_i: int) -> None:
        """Subset sub_i of a counted launcher (count=(n_dev, cap, sub_k)): its bounds are taken on the device."""
        if self._count is None or not 0 <= sub_i < self._count[2]:
            raise ValueError("launch_subset needs a counted launcher and 0 <= sub_i < sub_k")
        self._go(sub_i, self._args[self._i_base])

    def _go(self, first: int, row_base: int) -> None:
        """Launch with the row count (or, counted, the subset index) `first` and row_base."""
        s = stream_of(self._dev)
        if self._own_ws and s != self._stream:
            raise ValueError("this launcher owns its workspace and launches only on the stream it was built on "
                             "(pass workspace= to launch on others)")
        a = self._args
        a[self._i_rows] = first
        a[self._i_base] = row_base
        a[-1] = s
        rc = self._fn(*a)
        if rc:
            _lib.check(rc, self.name)


def _segment_check(mat, rows, row_base, n_rows, row_scale, y, y_by_row, P, dev, what):
    """Operand extents of one row segment of a joint projection."""
    if n_rows < 0 or row_base < 0:
        raise ValueError(f"{what}: negative row range")
    if mat.stride(1) != 1 or mat.stride(0) < P:
        raise ValueError(f"{what}: rows must be contiguous and cover the {P} columns")
    if rows is not None:
        if rows.dtype != torch.int32 or row_base + n_rows > rows.shape[0]:
            raise ValueError(f"{what}: row list shorter than row_base + n_rows (int32 expected)")
    elif row_base + n_rows > mat.shape[0]:
        raise ValueError(f"{what}: row range exceeds the matrix")
    if row_scale is not None and row_scale.numel() < mat.shape[0]:
        raise ValueError(f"{what}: row_scale needs one entry per stored row")
    if y is not None and y.shape[0] < (mat.shape[0] if y_by_row else n_rows):
        raise ValueError(f"{what}: numerators shorter than the rows")
    for t in (rows, row_scale, y):
        if t is not None and t.device != dev:
            raise ValueError(f"{what}: all operands must be on the image's device")


def proj_joint_em(mat: torch.Tensor, mat2: torch.Tensor, *, dtype_c: int, n_cols: int, x: torch.Tensor,
                  w: torch.Tensor, s: torch.Tensor, c1: float, c2: float, n_rows: int, n_rows2: int,
                  rows=None, row_base: int = 0, row_scale=None, y=None, y_by_row: bool = False,
                  rows2=None, row_base2: int = 0, row_scale2=None, y2=None, y2_by_row: bool = False,
                  eps: float | None = None, em_mode: int = 0, em_eps: float = 1e-12, zero_w: bool = True,
                  accumulate: bool = True, workspace: torch.Tensor | None = None) -> None:
    """One JSCC-SD subset step in one projection launch (jscc_proj_joint_em):
    w (+)= c1 * A^T(y / A x) + c2 * T^T(y2 / T x) over the two row segments,
    then x <- update(x, w, s) and w cleared when zero_w.  Both matrices hold
    the storage type dtype_c; y2 None means numerators of one."""
    dev = x.device
    P = int(n_cols)
    if n_rows + n_rows2 <= 0:
        raise ValueError("proj_joint_em: empty subset (use em_update)")
    for m in (mat, mat2):
        if m.device != dev:
            raise ValueError("proj_joint_em: matrices must be on the image's device")
    if mat.dtype != mat2.dtype:
        raise ValueError("proj_joint_em: both matrices must hold the same storage type")
    _segment_check(mat, rows, row_base, n_rows, row_scale, y, y_by_row, P, dev, "proj_joint_em (first segment)")
    _segment_check(mat2, rows2, row_base2, n_rows2, row_scale2, y2, y2_by_row, P, dev,
                   "proj_joint_em (second segment)")
    for t in (x, w, s):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.device != dev or t.numel() != P:
            raise ValueError(f"proj_joint_em: image, weight and sensitivity must be contiguous float32 of {P} pixels")
    need = _lib.lib().jscc_proj_workspace_bytes(n_rows + n_rows2, P)
    if workspace is not None:
        if workspace.numel() * workspace.element_size() < need or workspace.device != dev:
            raise ValueError(f"projection workspace needs {need} bytes on the image's device")
        ws = workspace
    else:
        ws = _workspace(dev, need)
    rc = _lib.lib().jscc_proj_joint_em(
        mat.data_ptr(), int(dtype_c), mat.stride(0), P, _p(rows), int(n_rows), int(row_base), _p(row_scale), _p(y),
        y.stride(0) if y is not None else 0, 1 if y_by_row else 0, float(c1),
        mat2.data_ptr(), mat2.stride(0), _p(rows2), int(n_rows2), int(row_base2), _p(row_scale2), _p(y2),
        y2.stride(0) if y2 is not None else 0, 1 if y2_by_row else 0, float(c2),
        x.data_ptr(), float(eps) if eps is not None else 0.0, 1 if eps is not None else 0, 1 if accumulate else 0,
        w.data_ptr(), P, x.data_ptr(), s.data_ptr(), int(em_mode), float(em_eps), 1 if zero_w else 0,
        ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_proj_joint_em")


def _proj_call(mode: int, mat: torch.Tensor, *, dtype_c: int | None = None, n_cols: int | None = None,
               rows: torch.Tensor | None = None, n_rows: int | None = None, row_base: int = 0,
               row_scale: torch.Tensor | None = None, x: torch.Tensor | None = None,
               x_perm: torch.Tensor | None = None, y: torch.Tensor | None = None, coef: torch.Tensor | None = None,
               eps: float | None = None, fwd_out: torch.Tensor | None = None, inv_perm: torch.Tensor | None = None,
               out_scale: float = 1.0, accumulate: bool = False, w: torch.Tensor | None = None,
               y_by_row: bool = False, sanitize: bool = False, em: tuple | None = None,
               workspace: torch.Tensor | None = None, count: tuple | None = None):
    """Checks the operands of a projection launch and returns (entry point name,
    argument list, (index of n_rows, index of row_base) in that list).

    count = (n_dev, cap, sub_k): a counted launch (jscc_proj_counted /
    jscc_proj_em_counted) over subset sub_i of sub_k of the first n_dev rows
    past row_base (n_dev a device int32 clamped to cap); the first index
    returned is then that of sub_i (0 in the list built here)."""
    dev = mat.device
    if count is not None:
        n_dev, cap, sub_k = count
        if n_dev.dtype != torch.int32 or n_dev.device != dev or cap < 1 or sub_k < 1:
            raise ValueError("counted launch: a device int32 count, a positive cap and subset count")
        if mode not in (MODE_RATIO, MODE_ONES) or coef is not None or fwd_out is not None:
            raise ValueError("counted launches take the ratio and column-sum modes")
        if y is not None and not y_by_row:
            raise ValueError("counted launches index their numerators by the stored row")
        n_rows = cap
    if dtype_c is None:
        dtype_c = dtype_code(mat.dtype) if mat.dtype != torch.uint8 else DT_FP8
    P = n_cols if n_cols is not None else mat.shape[1]
    if n_rows is None:
        n_rows = rows.shape[0] - row_base if rows is not None else mat.shape[0] - row_base
    ld = mat.stride(0)
    wl = 0
    if w is not None:
        wl = w.shape[0]
    # operand extents the kernels rely on
    if P > mat.shape[1] and mat.stride(0) < P:
        raise ValueError("n_cols exceeds the matrix row")
    if rows is not None:
        if row_base + n_rows > rows.shape[0] or rows.dtype != torch.int32:
            raise ValueError("row list shorter than row_base + n_rows (int32 expected)")
    elif row_base + n_rows > mat.shape[0]:
        raise ValueError("row range exceeds the matrix")
    if row_scale is not None and row_scale.numel() < mat.shape[0]:
        raise ValueError("row_scale needs one entry per stored row")
    if x_perm is not None and x_perm.shape[0] < P:
        raise ValueError(f"RotMat column has {x_perm.shape[0]} entries for {P} pixels")
    if x is not None and x_perm is None and x.shape[0] != P:
        raise ValueError(f"image has {x.shape[0]} pixels, matrix {P} columns")
    if inv_perm is not None and inv_perm.shape[0] < wl:
        raise ValueError("RotMatInv column shorter than the weight vector")
    if y is not None:
        need = mat.shape[0] if y_by_row else n_rows
        if y.shape[0] < need:
            raise ValueError(f"projection data has {y.shape[0]} bins for {need} rows")
    if coef is not None and coef.shape[0] < n_rows:
        raise ValueError("coefficients shorter than the rows")
    if fwd_out is not None and fwd_out.shape[0] < n_rows:
        raise ValueError("forward output shorter than the rows")
    for t in (x, x_perm, y, coef, fwd_out, inv_perm, w, row_scale, rows):
        if t is not None and t.device != dev:
            raise ValueError("all operands must be on the matrix's device")
    n_launch = n_rows if count is None else -(-count[1] // count[2])  # rows one launch is sized for
    need = _lib.lib().jscc_proj_workspace_bytes(max(n_launch, 1), P)
    if workspace is not None:
        if workspace.numel() * workspace.element_size() < need or workspace.device != dev:
            raise ValueError(f"projection workspace needs {need} bytes on the matrix's device")
        ws = workspace
    else:
        ws = _workspace(dev, need)
    xl = x.shape[0] if x is not None else 0
    flags = (1 if y_by_row else 0) | (2 if sanitize else 0)
    if em is not None:
        x_em, s_em, em_mode, em_eps, zero_w = em[:5]
        c1, w2, c2 = em[5:] if len(em) > 5 else (1.0, None, 0.0)
        if mode != MODE_RATIO:
            raise ValueError("the EM epilogue follows the ratio back-projection only")
        for t in (x_em, s_em) + ((w2,) if w2 is not None else ()):
            if (t.dtype != torch.float32 or not t.is_contiguous() or t.device != dev or w is None
                    or t.numel() < w.shape[0]):
                raise ValueError("EM operands must be contiguous float32 covering the weight, on the matrix's device")
        rng = [n_rows, row_base] if count is None else [count[0].data_ptr(), count[1], 0, count[2], row_base]
        args = [
            mat.data_ptr(), dtype_c, ld, P, _p(rows), *rng, _p(row_scale), _p(x), _p(x_perm),
            x_perm.stride(0) if x_perm is not None else 0, xl, _p(y), y.stride(0) if y is not None else 0,
            float(eps) if eps is not None else 0.0, 1 if eps is not None else 0, _p(inv_perm),
            inv_perm.stride(0) if inv_perm is not None else 0, float(out_scale), 1 if accumulate else 0, flags,
            _p(w), wl, x_em.data_ptr(), s_em.data_ptr(), int(em_mode), float(em_eps), float(c1), _p(w2), float(c2),
            1 if zero_w else 0,
            ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev)]
        if count is not None:
            return "jscc_proj_em_counted", args, (7, 9)
        return "jscc_proj_em", args, (5, 6)
    if count is not None:
        args = [
            mode, mat.data_ptr(), dtype_c, ld, P, _p(rows), count[0].data_ptr(), count[1], 0, count[2], row_base,
            _p(row_scale), _p(x), _p(x_perm), x_perm.stride(0) if x_perm is not None else 0, xl, _p(y),
            y.stride(0) if y is not None else 0, float(eps) if eps is not None else 0.0, 1 if eps is not None else 0,
            _p(inv_perm), inv_perm.stride(0) if inv_perm is not None else 0, float(out_scale),
            1 if accumulate else 0, flags, _p(w), wl,
            ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev)]
        return "jscc_proj_counted", args, (8, 10)
    args = [
        mode, mat.data_ptr(), dtype_c, ld, P, _p(rows), n_rows, row_base, _p(row_scale), _p(x), _p(x_perm),
        x_perm.stride(0) if x_perm is not None else 0, xl, _p(y), y.stride(0) if y is not None else 0, _p(coef),
        float(eps) if eps is not None else 0.0, 1 if eps is not None else 0, _p(fwd_out), _p(inv_perm),
        inv_perm.stride(0) if inv_perm is not None else 0, float(out_scale), 1 if accumulate else 0,
        flags, _p(w), wl,
        ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev)]
    return "jscc_proj", args, (6, 7)


EM_PLAIN, EM_SAFE, EM_APPLY = 0, 1, 2


def em_update(x: torch.Tensor, w1: torch.Tensor, s: torch.Tensor, mode: int = EM_PLAIN, c1: float = 1.0,
              w2: torch.Tensor | None = None, c2: float = 0.0, eps: float = 1e-12,
              zero_after: torch.Tensor | None = None) -> None:
    """In-place x <- update(x, c1*w1 + c2*w2, s)."""
    n = x.numel()
    for t in (w1, s, w2, zero_after):
        if t is not None and (t.numel() < n or t.device != x.device or t.dtype != torch.float32
                              or not t.is_contiguous()):
            raise ValueError("EM update operands must be contiguous float32 of the image's size and device")
    rc = _lib.lib().jscc_em_update(x.data_ptr(), w1.data_ptr(), float(c1), _p(w2), float(c2), s.data_ptr(),
                                   x.numel(), int(mode), float(eps), _p(zero_after), stream_of(x.device))
    _lib.check(rc, "jscc_em_update")


def sensitivity(A: torch.Tensor, inv: torch.Tensor | None) -> torch.Tensor:
    """(1/R) sum_r colsum(A[:, inv[:, r]-1]) as [P, 1] FP32 (reductions in FP64)."""
    require_gpu(A, "sysmat")
    dev = A.device
    if A.dtype != torch.float32 or A.stride(1) != 1:
        A = A.float().contiguous()
    M, P = A.shape
    R = inv.shape[1] if inv is not None else 1
    if inv is not None:
        if inv.ndim != 2 or inv.shape[0] != P:
            raise ValueError(f"RotMatInv must be [{P}, R]")
        inv = inv.to(device=dev, dtype=torch.int64)
    out = torch.empty((P, 1), dtype=torch.float32, device=dev)
    ws = torch.empty(_lib.lib().jscc_sensitivity_workspace_bytes(M, P), dtype=torch.uint8, device=dev)
    rc = _lib.lib().jscc_sensitivity(A.data_ptr(), M, A.stride(0), P, _p(inv),
                                     inv.stride(0) if inv is not None else 0, inv.stride(1) if inv is not None else 0,
                                     R, out.data_ptr(), ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(),
                                     stream_of(dev))
    _lib.check(rc, "jscc_sensitivity")
    return out


CP_RATIO, CP_FWD, CP_COEF = 0, 1, 2


_ragged_rows: dict = {}


def ragged_factor_row(P: int, device) -> torch.Tensor:
    """[1, P rounded up to 4] FP32: ones on the P voxels, zeros past them.  The no-store projections run on
    whole voxel quads; for a voxel count that is not a multiple of 4 (and no system-matrix factor, whose rows
    are zero-padded already) this row, read with stride 0 as every event's first-hit factor, gives the padded
    voxels of the last quad zero weight."""
    key = (int(P), str(device))
    row = _ragged_rows.get(key)
    if row is None:
        row = torch.zeros((1, round_up(P, 4)), dtype=torch.float32, device=device)
        row[0, :P] = 1.0
        _ragged_rows[key] = row
    return row


def cone_proj(mode: int, events: torch.Tensor, ev_idx: torch.Tensor, n_rows: int, vox: VoxelSoA, x: torch.Tensor,
              *, sysmat=None, sigma_model: int = SIG_FIXED, fixed_sigma: float = 0.0, coef=None, eps=None,
              normalize: bool = True, fwd_out=None, dsig_out=None, out_scale: float = 1.0, accumulate=False,
              w=None) -> None:
    """No-store cone projection (rows recomputed in registers), see include/jscc_hip.h."""
    dev = events.device
    P = vox.n
    if x.numel() != P or x.dtype != torch.float32 or not x.is_contiguous() or x.device != dev:
        raise ValueError("image must be a contiguous float32 vector with one value per voxel")
    if ev_idx.dtype != torch.int32 or ev_idx.numel() < n_rows or not ev_idx.is_contiguous():
        raise ValueError("event index list must be contiguous int32 with n_rows entries")
    for t, need in ((coef, n_rows), (fwd_out, n_rows), (dsig_out, n_rows), (w, P)):
        if t is not None and (t.numel() < need or t.dtype != torch.float32 or t.device != dev):
            raise ValueError("cone_proj operand has the wrong size, dtype or device")
    A, lda = None, 0
    if sysmat is not None:
        A = aligned_rows(sysmat, 4, dev)
        if A.shape[1] != P:
            raise ValueError("sysmat columns do not match the voxels")
        lda = A.stride(0)
    elif P % 4:
        A = ragged_factor_row(P, dev)
    if P % 4 or x.data_ptr() % 16:
        # the kernels read the image in whole 16-byte quads: a zero-padded copy
        xq = torch.zeros(round_up(P, 4), dtype=torch.float32, device=dev)
        xq[:P] = x
        x = xq
    ws = None
    nbytes = 0
    if mode != CP_FWD:
        nbytes = _lib.lib().jscc_cone_proj_workspace_bytes(max(n_rows, 1), P)
        ws = _workspace(dev, nbytes)
        nbytes = ws.numel()
    rc = _lib.lib().jscc_cone_proj(
        mode, events.data_ptr(), ev_idx.data_ptr(), n_rows, vox.x.data_ptr(), vox.y.data_ptr(), vox.z.data_ptr(),
        P, _p(A), lda, int(sigma_model), float(fixed_sigma), x.data_ptr(), _p(coef),
        float(eps) if eps is not None else 0.0, 1 if eps is not None else 0, 1 if normalize else 0, _p(fwd_out),
        _p(dsig_out), float(out_scale), 1 if accumulate else 0, _p(w), w.numel() if w is not None else 0,
        _p(ws), nbytes, err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_cone_proj")


BP_OSEM, BP_SUM = 0, 1


def cone_proj_batched(mode: int, events: torch.Tensor, ev_idx: torch.Tensor, row_off: torch.Tensor, vox: VoxelSoA,
                      *, x: torch.Tensor | None = None, sens: torch.Tensor | None = None, w: torch.Tensor | None = None,
                      sysmat=None, sigma_model: int = SIG_FIXED, fixed_sigma: float = 0.0, eps=None,
                      normalize: bool = True, em_mode: int = 0, em_eps: float = 1e-12, groups_per_cu: int = 0,
                      tile: int = 0, workspace: torch.Tensor | None = None) -> None:
    """Batched no-store cone projection over B reconstructions (see include/jscc_hip.h):
    BP_OSEM updates the images x [B, ldx] in place (one OSEM subset step each),
    BP_SUM writes the column sums of the normalised rows to w [B, ldw].  groups_per_cu and tile (one of
    cone_batch_tiles(P) shapes, 0 the default) choose the launch shape."""
    dev = events.device
    P = vox.n
    B = row_off.numel() - 1
    n_rows = ev_idx.numel()
    if row_off.dtype != torch.int64 or row_off.device != dev or ev_idx.dtype != torch.int32 or B < 1:
        raise ValueError("row offsets must be int64 [B+1] and the row list int32, on the events' device")
    for t in (x, sens, w):
        if t is not None and (t.ndim != 2 or t.shape[0] != B or t.shape[1] < P or t.dtype != torch.float32
                              or t.stride(1) != 1 or t.device != dev):
            raise ValueError("images / sensitivities / weights must be float32 [B, >= P] with unit column stride")
    if x is not None and (x.shape[1] < round_up(P, 4) or x.stride(0) % 4 != 0):
        raise ValueError("images must be [B, >= P rounded up to 4] with a row stride that is a multiple of 4 "
                         "(the kernels read whole voxel quads; values past P must be finite)")
    A, lda = None, 0
    if sysmat is not None:
        A = aligned_rows(sysmat, 4, dev)
        lda = A.stride(0)
    elif P % 4:
        A = ragged_factor_row(P, dev)
    nbytes = _lib.lib().jscc_cone_batch_workspace_bytes(max(n_rows, 1), B, P, int(groups_per_cu))
    ws = workspace if workspace is not None else _workspace(dev, nbytes)
    if ws.numel() < nbytes:
        raise ValueError("workspace too small")
    rc = _lib.lib().jscc_cone_proj_batched(
        mode, events.data_ptr(), ev_idx.data_ptr(), n_rows, row_off.data_ptr(), B, vox.x.data_ptr(),
        vox.y.data_ptr(), vox.z.data_ptr(), P, _p(A), lda, int(sigma_model),
        float(fixed_sigma), _p(x), x.stride(0) if x is not None else 0, float(eps) if eps is not None else 0.0,
        1 if eps is not None else 0, 1 if normalize else 0, _p(sens), sens.stride(0) if sens is not None else 0,
        int(em_mode), float(em_eps), _p(w), w.stride(0) if w is not None else 0, int(groups_per_cu), int(tile),
        ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_cone_proj_batched")


def cone_batch_tiles(n_vox: int) -> list:
    """The tile shapes (threads per work-group, voxel quads per thread, rows per step) of the batched
    no-store projection at n_vox voxels; index 0 is the default."""
    import ctypes
    out = []
    for t in range(int(_lib.lib().jscc_cone_batch_tiles(n_vox))):
        v = [ctypes.c_int(0) for _ in range(3)]
        _lib.check(_lib.lib().jscc_cone_batch_tile_shape(n_vox, t, *[ctypes.byref(c) for c in v]),
                   "jscc_cone_batch_tile_shape")
        out.append(tuple(c.value for c in v))
    return out


def cone_batch_workspace(dev, n_rows: int, n_recon: int, P: int, groups_per_cu: int = 0) -> torch.Tensor:
    """A dedicated workspace for the batched projection (stable address, e.g. for graph capture)."""
    nbytes = _lib.lib().jscc_cone_batch_workspace_bytes(max(n_rows, 1), n_recon, P, int(groups_per_cu))
    return torch.empty(nbytes, dtype=torch.uint8, device=dev)


_persist_ws: dict = {}


def osem_persistent(mat: torch.Tensor, dtype_c: int, n_cols: int, rows: torch.Tensor, row_scale, sub_off: torch.Tensor,
                    n_iter: int, x: torch.Tensor, sens: torch.Tensor, eps=None, em_mode: int = 0,
                    em_eps: float = 1e-12) -> None:
    """All n_iter x K list-mode OSEM subset steps in one cooperative launch (jscc_osem_persistent);
    sub_off: int64 [K+1] device offsets into the kept-row list `rows`.  x is updated in place."""
    dev = mat.device
    for t in (x, sens):
        if t.dtype != torch.float32 or t.numel() < n_cols or not t.is_contiguous() or t.device != dev:
            raise ValueError("image and sensitivity must be contiguous float32 with one value per pixel")
    if rows.dtype != torch.int32 or sub_off.dtype != torch.int64 or sub_off.device != dev:
        raise ValueError("row list int32 and subset offsets int64 on the matrix's device")
    K = sub_off.numel() - 1
    nbytes = _lib.lib().jscc_osem_persistent_workspace_bytes(n_cols)
    key = (str(dev), stream_of(dev), nbytes)
    ws = _persist_ws.get(key)
    if ws is None:
        _persist_ws.clear()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _persist_ws[key] = ws
    rc = _lib.lib().jscc_osem_persistent(
        mat.data_ptr(), dtype_c, mat.stride(0), n_cols, rows.data_ptr(), _p(row_scale), sub_off.data_ptr(), K,
        n_iter, x.data_ptr(), sens.data_ptr(), float(eps) if eps is not None else 0.0, 1 if eps is not None else 0,
        int(em_mode), float(em_eps), ws.data_ptr(), ws.numel(), err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_osem_persistent")


def thin_counts(counts: torch.Tensor, positions: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Histogram over the bins of the copy positions `positions` (int64, device) in the implicit expansion
    of round(counts) (bin j repeated round(counts[j]) times); FP32 [n_bins] (see jscc_thin_counts)."""
    dev = counts.device
    require_gpu(counts, "counts")
    cnt = torch.round(counts.double()).to(torch.int64)
    cum = torch.cumsum(cnt, 0)
    n = counts.numel()
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=dev)
    pos = positions.to(device=dev, dtype=torch.int64).contiguous()
    ws = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    rc = _lib.lib().jscc_thin_counts(cum.data_ptr(), n, pos.data_ptr(), pos.numel(), out.data_ptr(), out.stride(0),
                                     ws.data_ptr(), err_word(dev).data_ptr(), stream_of(dev))
    _lib.check(rc, "jscc_thin_counts")
    return out


def sum_f64(v: torch.Tensor) -> torch.Tensor:
    """FP64 sum of a float32 vector as a device scalar (deterministic order)."""
    if v.dtype != torch.float32 or v.ndim != 1:
        raise ValueError("sum_f64 takes a 1-D float32 tensor")
    out = torch.empty((), dtype=torch.float64, device=v.device)
    rc = _lib.lib().jscc_sum_f64(v.data_ptr(), v.numel(), v.stride(0), out.data_ptr(), stream_of(v.device))
    _lib.check(rc, "jscc_sum_f64")
    return out


def fp8_selftest(device) -> torch.Tensor:
    out = torch.empty(512, dtype=torch.float32, device=device)
    rc = _lib.lib().jscc_fp8_selftest(out.data_ptr(), stream_of(out.device))
    _lib.check(rc, "jscc_fp8_selftest")
    return out


# ----------------------------------------------------------------------------- K frames on the matrix cores
FRAMES_K = (16, 32, 64)   # frame counts of one multi-frame GEMM (pad with zero-data frames)
FRAMES_ALIGN = 256        # reduced-index granularity of a multi-frame GEMM


def _check_bf16_operand(t: torch.Tensor, name: str) -> None:
    if t.dtype != torch.bfloat16 or t.ndim != 2 or t.stride(1) != 1 or t.stride(0) % 8 or t.data_ptr() % 16:
        raise ValueError(f"{name} must be a 2-D BF16 tensor, rows contiguous, row stride a multiple of 8, "
                         "16-byte aligned")


def frames_transpose(mat: torch.Tensor, n_rows: int, n_cols: int, ld_out: int) -> torch.Tensor:
    """The transpose [n_cols, ld_out] of the BF16 matrix mat[:n_rows, :n_cols] (columns past n_rows zero)."""
    require_gpu(mat, "matrix")
    _check_bf16_operand(mat, "matrix")
    out = torch.empty((n_cols, ld_out), dtype=torch.bfloat16, device=mat.device)
    _lib.check(_lib.lib().jscc_frames_transpose_bf16(mat.data_ptr(), n_rows, n_cols, mat.stride(0), out.data_ptr(),
                                                     ld_out, stream_of(mat.device)), "jscc_frames_transpose_bf16")
    return out


def frames_split(x: torch.Tensor, n: int, n_pad: int, hi: torch.Tensor, lo: torch.Tensor) -> None:
    """hi/lo[f, i] = the BF16 terms of x[i, f] (i < n; zero up to n_pad) for a frames-minor FP32 array x."""
    _check_bf16_operand(hi, "hi")
    _check_bf16_operand(lo, "lo")
    if x.dtype != torch.float32 or x.stride(1) != 1 or hi.shape[0] != x.shape[1] or hi.shape != lo.shape:
        raise ValueError("frames_split: x must be FP32 [n, K] (frames contiguous) and hi/lo [K, >= n_pad]")
    _lib.check(_lib.lib().jscc_frames_split(x.data_ptr(), n, n_pad, x.shape[1], x.stride(0), hi.data_ptr(),
                                            lo.data_ptr(), hi.stride(0), stream_of(x.device)), "jscc_frames_split")


_cu_counts: dict[int, int] = {}


def device_cu_count(device: torch.device) -> int:
    """Compute units of a GPU (jscc_device_cu_count on it, cached)."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _cu_counts:
        import ctypes
        v = ctypes.c_int(0)
        with torch.cuda.device(idx):
            _lib.check(_lib.lib().jscc_device_cu_count(ctypes.byref(v)), "jscc_device_cu_count")
        _cu_counts[idx] = v.value
    return _cu_counts[idx]


def frames_split_count(m: int, kdim: int, device: torch.device) -> int:
    """Chunks of the reduced index of a multi-frame GEMM with m output rows (about 16 waves per CU)."""
    return int(_lib.lib().jscc_frames_split_count(m, kdim, device_cu_count(device)))


def frames_gemm(a: torch.Tensor, m: int, kdim: int, b_hi: torch.Tensor, b_lo: torch.Tensor, n_split: int,
                partial: torch.Tensor) -> None:
    """partial[sp, i, f] = sum over chunk sp of the reduced index of a[i, k] * (b_hi[f, k] + b_lo[f, k])
    (jscc_frames_gemm): a [>= m, >= kdim] and b [K, >= kdim] BF16, kdim a multiple of FRAMES_ALIGN."""
    for t, nm in ((a, "a"), (b_hi, "b_hi"), (b_lo, "b_lo")):
        require_gpu(t, nm)
        _check_bf16_operand(t, nm)
    K = b_hi.shape[0]
    if K not in FRAMES_K or b_lo.shape != b_hi.shape or b_hi.stride(0) != b_lo.stride(0):
        raise ValueError(f"frames_gemm: the frame operand must hold {FRAMES_K} frames in matching hi/lo tensors")
    if a.shape[0] < m or a.shape[1] < kdim or b_hi.shape[1] < kdim or kdim % FRAMES_ALIGN:
        raise ValueError("frames_gemm: operand shapes do not cover (m, kdim)")
    if partial.dtype != torch.float32 or partial.numel() * 4 < n_split * m * K * 4:
        raise ValueError("frames_gemm: partial too small")
    _lib.check(_lib.lib().jscc_frames_gemm(a.data_ptr(), a.stride(0), m, kdim, b_hi.data_ptr(), b_lo.data_ptr(),
                                           b_hi.stride(0), K, n_split, partial.data_ptr(), partial.numel() * 4,
                                           stream_of(a.device)), "jscc_frames_gemm")


def frames_ratio(partial: torch.Tensor, n_split: int, n: int, n_pad: int, c_off: int, y: torch.Tensor, eps,
                 c_hi: torch.Tensor, c_lo: torch.Tensor, f_out: torch.Tensor | None = None) -> None:
    """C^T terms of c = y / F (/ max(F, eps)) for a subset of n rows, F summed from the forward GEMM's partials
    (jscc_frames_ratio); y [n, K] FP32 frames-minor, c_hi/c_lo [K, >= n_pad]."""
    K = c_hi.shape[0]
    if y.dtype != torch.float32 or y.stride(1) != 1 or y.shape[0] < n or y.shape[1] != K:
        raise ValueError("frames_ratio: y must be FP32 [>= n, K] with frames contiguous")
    if c_hi.shape != c_lo.shape or c_hi.shape[1] < n_pad or not 0 <= c_off <= n_pad - n:
        raise ValueError("frames_ratio: coefficient buffers do not cover the subset")
    _lib.check(_lib.lib().jscc_frames_ratio(partial.data_ptr(), n_split, n, n_pad, c_off, K, y.data_ptr(), y.stride(0),
                                            float(eps) if eps is not None else 0.0, 1 if eps is not None else 0,
                                            c_hi.data_ptr(), c_lo.data_ptr(), c_hi.stride(0), _p(f_out),
                                            stream_of(y.device)), "jscc_frames_ratio")


def frames_update(partial: torch.Tensor | None, n_split: int, s: torch.Tensor, x: torch.Tensor, em_mode: int = EM_PLAIN,
                  t_hi: torch.Tensor | None = None, t_lo: torch.Tensor | None = None, em_eps: float = 1e-12) -> None:
    """x[j, f] <- update(x[j, f], W[j, f], s[j]) for every frame, W summed from the back-projection GEMM's partials
    (zero when partial is None), and the BF16 terms of the new image into t_hi/t_lo [K, >= P]
    (jscc_frames_update)."""
    P, K = x.shape
    if x.dtype != torch.float32 or x.stride(1) != 1 or s.shape[0] < P:
        raise ValueError("frames_update: x must be FP32 [P, K] with frames contiguous, s [P]")
    _lib.check(_lib.lib().jscc_frames_update(_p(partial), n_split, P, K, s.data_ptr(), x.data_ptr(), x.stride(0),
                                             em_mode, em_eps, _p(t_hi), _p(t_lo),
                                             t_hi.stride(0) if t_hi is not None else 0, stream_of(x.device)),
               "jscc_frames_update")
