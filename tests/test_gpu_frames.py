"""K single-photon data sets sharing one system matrix, projected on the matrix cores
(engine.FrameStore, jscc_frames_*): the GEMMs against FP64 products, and the OSEM of every
frame against the FP64 oracle of osem_bin_mode / get_weight_single
(recon_osem_plane.py:28-54) run with the stored (BF16-rounded) matrix."""
import numpy as np
import pytest
import torch

from oracle import osem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def bf16_round(a):
    """FP32 -> BF16 -> FP64, round to nearest even (torch's conversion)."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


@pytest.mark.parametrize("K", [16, 32, 64])
@pytest.mark.parametrize("m,kdim", [(1000, 512), (33, 256), (4096, 1536)])
def test_frames_gemm_vs_fp64_product(m, kdim, K):
    """partial sums of a [m, kdim] (BF16) times b [K, kdim] (FP32 as BF16 hi + lo): the product to the
    16 significant bits of b; with b exact in BF16 (lo = 0) to FP32 accumulation."""
    from jscc_recon_spect_cylinder_amd import ops
    rng = np.random.default_rng(m + kdim + K)
    dev = torch.device(DEV)
    a = torch.from_numpy(rng.normal(size=(m, kdim)).astype(np.float32)).to(dev).to(torch.bfloat16)
    a64 = a.double().cpu().numpy()
    for exact in (False, True):
        b = rng.random((kdim, K)).astype(np.float32) + 0.1
        if exact:
            b = bf16_round(b).astype(np.float32)
        hi = torch.empty((K, kdim), dtype=torch.bfloat16, device=dev)
        lo = torch.empty_like(hi)
        ops.frames_split(torch.from_numpy(b).to(dev), kdim, kdim, hi, lo)
        ns = ops.frames_split_count(m, kdim, dev)
        part = torch.full((ns * m * K,), np.nan, dtype=torch.float32, device=dev)
        ops.frames_gemm(a, m, kdim, hi, lo, ns, part)
        got = part.view(ns, m, K).double().sum(0).cpu().numpy()
        ref = a64 @ b.astype(np.float64)
        err = rel(got, ref)
        print(f"m {m} kdim {kdim} K {K} split {ns} exact {exact}: rel {err:.2e}")
        assert err < (2e-6 if exact else 2e-5)


def test_frames_transpose_is_exact_and_pads_with_zeros():
    from jscc_recon_spect_cylinder_amd import ops
    rng = np.random.default_rng(1)
    a = torch.from_numpy(rng.normal(size=(300, 200)).astype(np.float32)).to(DEV).to(torch.bfloat16)
    buf = torch.zeros((300, 256), dtype=torch.bfloat16, device=DEV)
    buf[:, :200] = a
    t = ops.frames_transpose(buf, 300, 200, 384)
    assert t.shape == (200, 384)
    assert torch.equal(t[:, :300].cpu(), a.t().cpu())
    assert not torch.any(t[:, 300:].float() != 0)


def _frames_problem(K, seed=21):
    from jscc_recon_spect_cylinder_amd import synth
    sc = synth.sc_osem(n_rings=32, n_angles=128, n_rot=2, bins_per_rot=1024, total_counts=2e5, seed=seed)
    rng = np.random.default_rng(seed)
    f = sc.A.astype(np.float64) @ sc.x_true
    Y = np.stack([rng.poisson(f * (10.0 ** rng.uniform(4, 6) / f.sum())) for _ in range(K)], axis=1)
    return sc.A, Y.astype(np.float32)


def _oracle_frames(Ab, Y, n_iter, bounds):
    s = Ab.sum(0)
    out = []
    for k in range(Y.shape[1]):
        x = np.ones(Ab.shape[1])
        y = Y[:, k].astype(np.float64)
        for _ in range(n_iter):
            for lo, hi in bounds:
                x = osem.em_plain(x, osem.weight_single(Ab[lo:hi], y[lo:hi], x), s)
        out.append(x)
    return np.stack(out, axis=1)


@pytest.mark.parametrize("K", [16, 20, 64])
def test_frame_store_osem_vs_fp64_oracle(K):
    """Every frame against the FP64 oracle with the matrix as stored (BF16-rounded): the only
    difference is the 16-bit frame operand of the GEMMs (bar 1e-4; observed 1.1e-6), and against
    the single-frame fused projection with BF16 rows (SystemStore.osem)."""
    from jscc_recon_spect_cylinder_amd.engine import FrameStore, SystemStore
    A, Y = _frames_problem(K)
    Ad = torch.from_numpy(A).to(DEV)
    fs = FrameStore.from_dense(Ad)
    s = fs.column_sums()
    X = torch.ones((A.shape[1], K), device=DEV)
    fs.osem(X, torch.from_numpy(Y).to(DEV), s, 3, 2)
    got = X.cpu().numpy()
    Ab = bf16_round(A)
    assert rel(s.cpu().numpy(), Ab.sum(0)) < 1e-6
    ref = _oracle_frames(Ab, Y, 3, fs.subset_bounds(2))
    errs = [rel(got[:, k], ref[:, k]) for k in range(K)]
    print(f"K {K}: worst frame rel err vs FP64 oracle {max(errs):.2e}")
    assert max(errs) < 1e-4
    ss = SystemStore.from_dense(Ad, "bf16")
    for k in (0, K - 1):
        x1 = torch.ones(A.shape[1], device=DEV)
        ss.osem(x1, torch.from_numpy(np.ascontiguousarray(Y[:, k])).to(DEV), s, 3, 2)
        assert rel(got[:, k], x1.cpu().numpy()) < 1e-4


def test_frame_store_empty_subsets_and_safe_update():
    """More subsets than bins (an empty subset applies x * 0 / s, as SystemStore.osem and the
    reference's osem_bin_mode do), and the safe EM update (recon_osem_plane_sparse.py:60-66) with a
    zero-sensitivity pixel, against the single-frame path frame by frame."""
    from jscc_recon_spect_cylinder_amd import ops
    from jscc_recon_spect_cylinder_amd.engine import FrameStore, SystemStore
    rng = np.random.default_rng(5)
    M, P, K = 40, 1000, 3
    A = (rng.random((M, P)) * (rng.random((M, P)) < 0.4) + 1e-3).astype(np.float32)
    Y = rng.poisson(50.0, size=(M, K)).astype(np.float32)
    A0 = A.copy()
    A0[:, 7] = 0.0   # a pixel no bin sees: s = 0 there
    for A_, n_sub, mode in ((A, 60, ops.EM_PLAIN), (A0, 4, ops.EM_SAFE)):
        Ad = torch.from_numpy(A_).to(DEV)
        fs = FrameStore.from_dense(Ad)
        s = fs.column_sums()
        X = torch.ones((P, K), device=DEV)
        fs.osem(X, torch.from_numpy(Y).to(DEV), s, 2, n_sub, eps=1e-12, em_mode=mode)
        ss = SystemStore.from_dense(Ad, "bf16")
        for k in range(K):
            x1 = torch.ones(P, device=DEV)
            ss.osem(x1, torch.from_numpy(np.ascontiguousarray(Y[:, k])).to(DEV), s, 2, n_sub, eps=1e-12,
                    em_mode=mode)
            a, b = X[:, k].cpu().numpy(), x1.cpu().numpy()
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
            if mode == ops.EM_PLAIN:
                assert not np.any(a) and not np.any(b)   # 40 bins, 60 subsets: the first subset is empty
            else:
                assert a[7] == 0.0 and b[7] == 0.0
                assert rel(a, b) < 1e-4


def test_frames_reject_bad_operands():
    from jscc_recon_spect_cylinder_amd import _lib, ops
    from jscc_recon_spect_cylinder_amd.engine import FrameStore
    dev = torch.device(DEV)
    a = torch.zeros((64, 128), dtype=torch.bfloat16, device=dev)
    hi = torch.zeros((16, 128), dtype=torch.bfloat16, device=dev)
    part = torch.zeros(64 * 16, device=dev)
    with pytest.raises(ValueError):
        ops.frames_gemm(a, 64, 100, hi, hi, 1, part)           # reduced length not a multiple of 128
    with pytest.raises(ValueError):
        ops.frames_gemm(a, 64, 128, hi[:15], hi[:15], 1, part)  # 15 frames
    rc = _lib.lib().jscc_frames_gemm(a.data_ptr(), 128, 64, 128, hi.data_ptr(), hi.data_ptr(), 128, 17, 1,
                                     part.data_ptr(), part.numel() * 4, ops.stream_of(dev))
    assert rc != 0 and b"frames" in _lib.lib().jscc_last_error()
    fs = FrameStore.from_dense(torch.ones((8, 256), device=dev))
    with pytest.raises(ValueError):
        fs.osem(torch.ones((256, 65), device=dev), torch.ones((8, 65), device=dev), torch.ones(256, device=dev), 1, 1)
