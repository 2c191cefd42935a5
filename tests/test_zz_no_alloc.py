"""The allocation-free OSEM subset step (jscc_proj_em_noalloc, called through the
C ABI) against jscc_proj_em with a caller workspace: the same kernels and
reduction order, so the images and weights agree bit for bit, over two calls
in a row (the second on a smaller subset, so the static buffer holds the
first call's partials when the second starts) and for every storage type."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _noalloc(lib, mat, dtype_c, P, rows, n_rows, row_base, scale, x, y, w, s, err, *, em_mode=0, c1=1.0, w2=None,
             c2=0.0, x_perm=None, inv_perm=None, eps=1e-12, out_scale=1.0):
    p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    stream = torch.cuda.current_stream(torch.device(DEV)).cuda_stream
    return lib.jscc_proj_em_noalloc(
        mat.data_ptr(), dtype_c, mat.stride(0), P, p(rows), n_rows, row_base, p(scale), x.data_ptr(), p(x_perm),
        x_perm.stride(0) if x_perm is not None else 0, x.numel(), p(y), 1 if y is not None else 0, eps, 1, p(inv_perm),
        inv_perm.stride(0) if inv_perm is not None else 0, out_scale, 1, 1 if y is not None else 0, w.data_ptr(),
        w.numel(), x.data_ptr(), s.data_ptr(), em_mode, 1e-12, c1, p(w2), c2, 1, err.data_ptr(), stream)


def _matrix(dtype, M, P, rng):
    from jscc_recon_spect_cylinder_amd import ops
    base = torch.from_numpy((rng.random((M, P)) ** 3 + 1e-3).astype(np.float32)).to(DEV)
    if dtype in ("fp8", "fp16"):
        codes, scale = ops.quantize_rows(base, P, M, None, dtype)
        return codes, scale
    scale = torch.from_numpy((rng.random(M) + 0.5).astype(np.float32)).to(DEV)
    m = base.to(torch.bfloat16) if dtype == "bf16" else base
    if P % 64:  # rows padded to whole vectors (their padding poisoned: it never reaches the sums)
        padded = torch.full((M, ops.round_up(P, 64)), float("nan"), dtype=m.dtype, device=DEV)
        padded[:, :P] = m
        m = padded
    return m, scale


@pytest.mark.parametrize("dtype,P,M", [("fp32", 16384, 3000), ("bf16", 16384, 3000), ("fp16", 25600, 3000),
                                       ("fp8", 4096, 3000), ("bf16", 32768, 600), ("fp32", 1001, 1500),
                                       ("bf16", 9001, 1500)])
def test_noalloc_subset_step_equals_workspace_form(dtype, P, M):
    from jscc_recon_spect_cylinder_amd import _lib, ops
    lib = _lib.lib()
    rng = np.random.default_rng(P + len(dtype))
    mat, scale = _matrix(dtype, M, P, rng)
    dt = ops.dtype_code(dtype)
    rows = torch.from_numpy(rng.permutation(M).astype(np.int32)).to(DEV)
    y = torch.from_numpy(rng.poisson(5.0, M).astype(np.float32)).to(DEV)
    s = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    x0 = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    err = ops.err_word(torch.device(DEV))
    xa, wa = x0.clone(), torch.zeros(P, device=DEV)
    xb, wb = x0.clone(), torch.zeros(P, device=DEV)
    # two calls in a row, the second on a smaller subset
    first = M * 13 // 15
    for lo, n in ((0, first), (first, M - first)):
        ops.proj(ops.MODE_RATIO, mat, dtype_c=dt, n_cols=P, rows=rows, row_base=lo, n_rows=n, row_scale=scale, x=xa,
                 y=y, y_by_row=True, eps=1e-12, accumulate=True, w=wa, em=(xa, s, ops.EM_PLAIN, 1e-12, True))
        rc = _noalloc(lib, mat, dt, P, rows, n, lo, scale, xb, y, wb, s, err)
        assert rc == 0, lib.jscc_last_error()
        torch.cuda.synchronize()
        assert torch.equal(xa, xb), (lo, n)
        assert torch.equal(wa, wb) and not wb.any()
    ops.raise_on_index_error(torch.device(DEV))
    assert not torch.equal(xb, x0)


def test_noalloc_rotated_joint_weight_form():
    """With the rotation gathers and a second weight (the JSCCSD combination (2 - a) w_c + a w_s)."""
    from jscc_recon_spect_cylinder_amd import _lib, ops
    lib = _lib.lib()
    rng = np.random.default_rng(5)
    P, M = 16384, 2000
    mat, scale = _matrix("bf16", M, P, rng)
    perm = rng.permutation(P)
    inv = np.empty(P, np.int64)
    inv[perm] = np.arange(P)
    x_perm = torch.from_numpy(perm.astype(np.int64) + 1).to(DEV)
    inv_perm = torch.from_numpy(inv + 1).to(DEV)
    s = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    x0 = torch.from_numpy((rng.random(P) + 0.5).astype(np.float32)).to(DEV)
    ws0 = torch.from_numpy(rng.random(P).astype(np.float32)).to(DEV)
    err = ops.err_word(torch.device(DEV))
    xa, wa, w2a = x0.clone(), torch.zeros(P, device=DEV), ws0.clone()
    xb, wb, w2b = x0.clone(), torch.zeros(P, device=DEV), ws0.clone()
    ops.proj(ops.MODE_RATIO, mat, dtype_c=ops.DT_BF16, n_cols=P, n_rows=M, row_scale=scale, x=xa, x_perm=x_perm,
             inv_perm=inv_perm, eps=1e-12, accumulate=True, w=wa, em=(xa, s, ops.EM_SAFE, 1e-12, True, 1.5, w2a, 0.5))
    rc = _noalloc(lib, mat, ops.DT_BF16, P, None, M, 0, scale, xb, None, wb, s, err, em_mode=ops.EM_SAFE, c1=1.5,
                  w2=w2b, c2=0.5, x_perm=x_perm, inv_perm=inv_perm)
    assert rc == 0, lib.jscc_last_error()
    torch.cuda.synchronize()
    assert torch.equal(xa, xb)
    assert not wb.any() and not w2b.any()
    ops.raise_on_index_error(torch.device(DEV))


def test_noalloc_rejects_subsets_beyond_its_buffer():
    from jscc_recon_spect_cylinder_amd import _lib, ops
    lib = _lib.lib()
    err = ops.err_word(torch.device(DEV))
    need = lib.jscc_proj_workspace_bytes(100000, 32768)  # the widest BF16 projection, at a full grid
    assert need > (24 << 20)
    stream = torch.cuda.current_stream(torch.device(DEV)).cuda_stream
    dummy = torch.zeros(16, device=DEV)
    rc = lib.jscc_proj_em_noalloc(dummy.data_ptr(), 1, 32768, 32768, None, 100000, 0, None, dummy.data_ptr(), None, 0,
                                  32768, None, 0, 1e-12, 1, None, 0, 1.0, 1, 0, dummy.data_ptr(), 32768,
                                  dummy.data_ptr(), dummy.data_ptr(), 0, 1e-12, 1.0, None, 0.0, 1, err.data_ptr(),
                                  stream)
    assert rc != 0 and b"static buffer" in lib.jscc_last_error()
    assert lib.jscc_proj_em_noalloc(dummy.data_ptr(), 1, 64, 64, None, 0, 0, None, dummy.data_ptr(), None, 0, 64, None,
                                    0, 1e-12, 1, None, 0, 1.0, 1, 0, dummy.data_ptr(), 64, dummy.data_ptr(),
                                    dummy.data_ptr(), 0, 1e-12, 1.0, None, 0.0, 1, err.data_ptr(), stream) != 0
    assert int(err.item()) == 0
