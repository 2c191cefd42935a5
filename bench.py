"""Benchmark of the reconstruction hot path on one MI355X.

Default workload (the JSCCD list-mode case): 2e5 synthetic two-hit events from
a rod phantom on a 64 x 256 polar grid (16384 voxels), cone blur sigma = 3 deg,
T (2e5 x 16384) built on the GPU, then 10 OSEM iterations x 8 subsets.

One step = build T from the list (event filters + cone back-projection) +
list-mode sensitivity (column sums of T) + 10 OSEM iterations, all on inputs
resident in HBM.  The step is timed for BF16-stored T (the reported value) and
for FP32-stored T.

Prints ONE JSON line: value = events reconstructed per second (Mevents/s) of
the BF16 path, the roofline of the dominant kernel (one OSEM subset pass
T_s^T(1/T_s x), timed by the library's HIP-event kernel timer on its launch
stream over the timed region), its HBM traffic from the committed rocprofv3
counter summary, and a bounded CPU baseline (the numpy oracle on the host).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 matrix-core peak of MI355X (no sparsity)
# MI355X VALU issue rate without packed FP32 (kept out of every kernel): one 64-lane instruction per 4 cycles
# per SIMD, 1024 SIMDs x 16 lanes x 2.4 GHz = 3.93e13 lane instructions/s (the 157.3 TFLOP/s FP32 figure
# counts packed FMAs: 2 lanes x 2 flops per lane instruction)
FP32_LANE_PEAK = 157.3e12 / 4
HERE = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_FILE = os.path.join(HERE, "profiles", "k_proj_fused_jsccd_2e5x16384_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--events", type=int, default=200_000)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--subsets", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=12000, help="events in the CPU-baseline sample (0: skip)")
    ap.add_argument("--skip-fp32", action="store_true", help="time the BF16 path only")
    ap.add_argument("--workload", default="jsccd",
                    choices=["jsccd", "sc", "jsccsd", "grad", "batched", "sparse", "frames"],
                    help="jsccd (default, the reported metric): JSCCD list mode, 2e5 events; sc: single-photon "
                         "OSEM, A 5120x4096; jsccsd: joint mode, A 81920x16384 + 1e6 events, FP32/BF16/FP8; "
                         "grad: Poisson log-likelihood gradient through the no-store cone operator; batched: 48 "
                         "reconstructions of 1e4-2e6 events in one launch per subset step; sparse: coarse-T "
                         "list-mode OSEM with on-the-fly expansion, 2e5 packed rows on the 25600-voxel grid; "
                         "frames: 16 single-photon data sets sharing the 81920x16384 matrix, projected as GEMMs on "
                         "the matrix cores")
    ap.add_argument("--headline-only", action="store_true",
                    help="the default (jsccd) line without the summaries of the other workloads")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch
    calibrate the counters against the bytes actually moved (mib MiB read + mib MiB written)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))


def _host_threads():
    """Host threads the CPU leg uses: this process's CPU affinity, capped by OMP_NUM_THREADS when it is set (the
    share of the host a run is given)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def fp64_oracle_sample(cl, noise1, noise2, n_sample, n_iter, n_sub, sigma, chunk=1000):
    """The checker of the error term: the FP64 oracle's kept rows of the first n_sample events (the same noise
    draws as the GPU run) and its n_iter x n_sub list-mode OSEM image."""
    sys.path.insert(0, HERE)
    from oracle import cone, osem
    from jscc_recon_spect_cylinder_amd import synth
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    blocks, kept = [], []
    for lo in range(0, n_sample, chunk):
        hi = min(lo + chunk, n_sample)
        out = cone.backproj_list_single(None, cl.detector, cl.coor, cl.list_origin[lo:hi], 0.0, 0.0, cl.e0, 0.02,
                                        tmax, tmin, tsum, noise1[lo:hi], noise2[lo:hi], fixed_sigma=sigma)
        blocks.append(out["t"])
        kept.append(out["kept"] + lo)
    T = np.concatenate(blocks)
    del blocks
    s = T.sum(0)
    x = np.ones(T.shape[1])
    bounds = osem.split_bounds_even(T.shape[0], n_sub)
    for _ in range(n_iter):
        for lo, hi in bounds:
            x = osem.em_plain(x, osem.weight_compton(T[lo:hi], x), s)
    return np.concatenate(kept), x


def cpu_baseline(cl, noise1, noise2, n_sample, n_sample_1t, n_iter, n_sub, sigma, x_ref, chunk=1000):
    """The reference's CPU path: the FP32 torch restatement of process_list_plane_strict.py:151-253 (fixed
    sigma, first-hit factor 1) and osem_list_mode (recon_osem_plane.py:63-80), timed on the host cores on the
    first n_sample events (T build in chunks + n_iter x n_sub OSEM over their rows) at the host share of
    threads, and on the first n_sample_1t events at one thread; extrapolated linearly per event (both parts
    scale with the event count).  The image of the threaded run is compared with the FP64 oracle's."""
    import torch
    sys.path.insert(0, HERE)
    from oracle import torch_cpu
    from jscc_recon_spect_cylinder_amd import synth
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    det = torch.from_numpy(cl.detector)
    coor = torch.from_numpy(cl.coor)
    lst = torch.from_numpy(cl.list_origin)
    threads_before = torch.get_num_threads()
    runs = {}
    try:
        for threads, n in ((_host_threads(), n_sample), (1, n_sample_1t)):
            torch.set_num_threads(threads)
            t0 = time.perf_counter()
            blocks = []
            for lo in range(0, n, chunk):
                hi = min(lo + chunk, n)
                blocks.append(torch_cpu.cone_rows_fixed_sigma(det, coor, lst[lo:hi].clone(), cl.e0, 0.02, tmax,
                                                              tmin, tsum, noise1[lo:hi], noise2[lo:hi], sigma)[0])
            T = torch.cat(blocks)
            del blocks
            t_build = time.perf_counter() - t0
            t0 = time.perf_counter()
            x = torch_cpu.list_osem(T, n_iter, n_sub)
            t_osem = time.perf_counter() - t0
            runs[threads] = {"events": n, "kept_rows": int(T.shape[0]), "t_build_s": round(t_build, 3),
                             "t_osem_s": round(t_osem, 3), "Mevents_s": n / (t_build + t_osem) / 1e6}
            if threads != 1 or n == n_sample:
                runs[threads]["rel_err_vs_fp64_oracle"] = _rel(x.numpy(), x_ref)
            del T
    finally:
        torch.set_num_threads(threads_before)
    nt = _host_threads()
    main_run = runs[nt]
    one = runs[1]
    return {"value": main_run["Mevents_s"], "unit": "Mevents/s", "cores": nt, "kind": "port",
            "sample": f"FP32 torch restatement of the reference's CPU path (process_list_plane_strict.py:151-253 with "
                      f"the workload's fixed sigma, osem_list_mode) on the first {n_sample} events at {nt} threads: "
                      f"T build {main_run['t_build_s']:.2f} s + {n_iter} x {n_sub} OSEM {main_run['t_osem_s']:.2f} s, "
                      f"extrapolated linearly per event; image vs the FP64 oracle "
                      f"{main_run.get('rel_err_vs_fp64_oracle', float('nan')):.2e}",
            "single_thread": {"value": one["Mevents_s"], "unit": "Mevents/s", "cores": 1,
                              "sample": f"first {one['events']} events: T build {one['t_build_s']:.2f} s + OSEM "
                                        f"{one['t_osem_s']:.2f} s"},
            "host_cpu": _cpu_model(), "host_threads_visible": os.cpu_count()}


# --------------------------------------------------------------------------- other workloads
def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_baseline_sc(pb, n_iter):
    """The numpy oracle (FP64, one thread): the full SC OSEM run; also the reference image."""
    sys.path.insert(0, HERE)
    from threadpoolctl import threadpool_limits
    from oracle import osem
    A = pb.A.astype(np.float64)
    s = A.sum(0)
    x = np.ones(A.shape[1])
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for _ in range(n_iter):
            for a, b in pb.subsets:
                x = osem.em_plain(x, osem.weight_single(A[a:b], pb.y[a:b].astype(np.float64), x), s)
        dt = time.perf_counter() - t0
    return {"value": n_iter / dt, "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle (1 thread), the full {n_iter}-iteration x {len(pb.subsets)}-subset run",
            "host_cpu": _cpu_model()}, x


def run_sc(args):
    """Single-photon OSEM (osem_bin_mode, one subset per rotation block): A 5120 x 4096 (32 x 128 polar grid,
    5 rotation blocks x 1024 bins, density 0.3, values exp(N(-9, 1.5))), 10 iterations x 5 subsets."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, synth
    from jscc_recon_spect_cylinder_amd.engine import SystemStore
    dev = torch.device("cuda:0")
    lib = _lib.load()
    pb = synth.sc_osem()
    M, P = pb.A.shape
    n_iter, K = 10, len(pb.subsets)
    A = torch.from_numpy(pb.A).to(dev)
    y = torch.from_numpy(pb.y).to(dev)
    res = {}
    imgs = {}
    for name in ("fp32", "bf16", "fp16", "fp8"):
        a = SystemStore.from_dense(A, name)
        s = a.column_sums()
        _prof_read(lib, _lib.PROF_PROJ)
        lib.jscc_prof_enable(1)
        dt, x = _timed(lambda: a.osem(torch.ones(P, device=dev), y, s, n_iter, K), args.steps, args.warmup)
        lib.jscc_prof_enable(0)
        c, t = _prof_read(lib, _lib.PROF_PROJ)
        b = {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}[name]
        sub_bytes = M / K * P * b
        # the same run with each iteration replayed from a HIP graph captured once
        xg = torch.ones(P, device=dev)
        gi = a.capture(xg, y, s, K)

        def replayed():
            xg.fill_(1.0)
            for _ in range(n_iter):
                gi.replay()
            return xg
        dtg, xg = _timed(replayed, args.steps, args.warmup)
        res[name] = {"iter_s": round(n_iter / dt, 2), "ms_per_run": round(dt * 1e3, 3),
                     "graph_iter_s": round(n_iter / dtg, 2), "graph_ms_per_run": round(dtg * 1e3, 3),
                     "graph_equals_eager": bool(torch.equal(xg, x)),
                     "subset_launch_us": round(t / max(c, 1) * 1e3, 2),
                     "subset_GBs": round(sub_bytes / (t / max(c, 1) * 1e-3) / 1e9, 1)}
        imgs[name] = x.cpu().numpy()
    cpu, xref = cpu_baseline_sc(pb, n_iter)
    for name in res:
        res[name]["rel_err_vs_fp64_oracle"] = _rel(imgs[name], xref)
    return {"metric": "SC OSEM iter/s (osem_bin_mode, A 5120x4096 FP32, 10 iterations x 5 subsets)",
            "value": res["fp32"]["graph_iter_s"], "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["fp32"]["graph_ms_per_run"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded)",
            "config": {"workload": "SC OSEM: 32x128 polar grid, A 5120x4096 (5 rotation blocks x 1024 bins), "
                                   "10 iterations x 5 subsets, each iteration replayed from a captured HIP graph "
                                   "(eager launches in components); A (84 MB FP32) is Infinity-Cache resident"},
            "components": res, "cpu_baseline": cpu}


def _big_sysmat(M, P, seed, dev, density=0.3, mu=-9.0, sd=1.5, chunk=4096):
    """Dense A [M, P] with the single-photon value law, drawn on the device in row chunks (synthetic data)."""
    import torch
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A = torch.empty((M, P), dtype=torch.float32, device=dev)
    for lo in range(0, M, chunk):
        hi = min(lo + chunk, M)
        v = torch.randn((hi - lo, P), generator=g, device=dev).mul_(sd).add_(mu).exp_()
        v.mul_(torch.rand((hi - lo, P), generator=g, device=dev) < density)
        A[lo:hi] = v
    return A


def _noise(n):
    """The smearing draws (host generator, like the reference's CPU torch.randn calls)."""
    import torch
    g = torch.Generator().manual_seed(7)
    return torch.randn(n, generator=g), torch.randn(n, generator=g)


def cpu_baseline_jsccsd(A, y, cl, keep_idx, n_iter, K, alpha, sigma, noise):
    """The numpy oracle (FP64, one thread) of osem_joint_mode on the reduced instance: T of the kept events
    built by the oracle in chunks, then n_iter joint iterations.  Returns the timing and the FP64 image."""
    sys.path.insert(0, HERE)
    from threadpoolctl import threadpool_limits
    from oracle import cone, osem
    from jscc_recon_spect_cylinder_amd import synth
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        ev = cone.prepare_events(cl.list_origin, cl.detector, cl.e0, 0.02, tmax, tmin, tsum, noise[0], noise[1])
        sel = np.flatnonzero(np.isin(ev["idx"], keep_idx))
        blocks = []
        for lo in range(0, sel.size, 1000):
            ii = sel[lo:lo + 1000]
            th = cone.theta_from_e1(ev["e1"][ii], cl.e0)
            kn = cone.klein_nishina(ev["e1"][ii], cl.e0)
            tc = cone.cone_tc(ev["pos1"][ii], ev["pos2"][ii], None, None, th, kn, None, None, cl.coor,
                              fixed_sigma=sigma)
            blocks.append(tc / tc.sum(1, keepdims=True))
        T = np.concatenate(blocks)
        t_build = time.perf_counter() - t0
        s_j = alpha * A.sum(0) + (2 - alpha) * T.sum(0)
        x = np.ones(A.shape[1])
        ab = [(A.shape[0] * i // K, A.shape[0] * (i + 1) // K) for i in range(K)]
        tb = [(T.shape[0] * i // K, T.shape[0] * (i + 1) // K) for i in range(K)]
        t0 = time.perf_counter()
        for _ in range(n_iter):
            for (alo, ahi), (tlo, thi) in zip(ab, tb):
                ws = osem.weight_single(A[alo:ahi], y[alo:ahi], x)
                wc = osem.weight_compton(T[tlo:thi], x)
                x = x * ((2 - alpha) * wc + alpha * ws) / s_j
        t_iter = (time.perf_counter() - t0) / n_iter
    return {"value": 1.0 / t_iter, "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle (1 thread) on the reduced instance (A {A.shape[0]}x{A.shape[1]}, "
                      f"{T.shape[0]} kept events): {n_iter} joint iterations x {K} subsets, {t_iter:.2f} s per "
                      f"iteration (T build {t_build:.1f} s not counted)", "host_cpu": _cpu_model()}, x


def run_jsccsd(args):
    """Joint JSCCSD OSEM with A 81920 x 16384 (20 rotations x 4096 bins, single-photon value law) and a 1e6-event
    list (64 x 256 grid, sigma 3 deg), alpha = 0.5, A and T stored in FP32 / BF16 / FP8 (per-row scales), FP32
    accumulation; one iteration = 10 subsets (8192 bins + 1/10 of the kept events each).  The image error is
    measured against the FP64 oracle on a reduced instance (A 8192 x 16384, 1e4 events, 10 iterations x 4
    subsets) computed in the CPU-baseline leg."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore, SystemStore, joint_osem
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    dev = torch.device("cuda:0")
    lib = _lib.load()
    alpha, sigma = 0.5, float(np.deg2rad(3.0))

    def problem(M, n_ev, seed):
        cl = synth.compton_list(n_events=n_ev, n_rings=64, n_angles=256, seed=seed)
        P = cl.coor.shape[0]
        A = _big_sysmat(M, P, seed, dev)
        x_true = torch.from_numpy(cl.x_true.astype(np.float32)).to(dev)
        a32 = SystemStore.from_dense(A, "fp32")
        f = a32.forward(x_true)
        y = torch.poisson(f * (1e6 * M / 81920 / float(f.sum())))
        return cl, A, y

    def rows_of(cl, dt):
        tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
        n1, n2 = _noise(cl.list_origin.shape[0])
        return build_compton_rows(None, torch.from_numpy(cl.detector).to(dev), torch.from_numpy(cl.coor).to(dev),
                                  torch.from_numpy(cl.list_origin).to(dev), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum,
                                  dev, noise=(n1.to(dev), n2.to(dev)), sigma_model="fixed", fixed_sigma=sigma,
                                  store_dtype=dt)

    def stores(A, rows32, rows16, name):
        a = SystemStore.from_dense(A, name)
        t = ListModeStore.from_rows(rows16 if name == "bf16" else rows32, name)
        return a, t

    # full size: timing
    n_ev = args.events if args.events != 200_000 else 1_000_000
    cl, A, y = problem(81920, n_ev, 20260333)
    P = A.shape[1]
    res = {}
    for name in ("bf16", "fp16", "fp8", "fp32"):
        rows = rows_of(cl, torch.bfloat16 if name == "bf16" else torch.float32)
        a, t = stores(A, rows, rows, name)
        del rows
        s_j = (alpha * a.column_sums() + (2 - alpha) * t.backproject_ones()).contiguous()
        K = 10
        dt, _ = _timed(lambda: joint_osem(a, y, t, torch.ones(P, device=dev), s_j, alpha, 1, K), args.steps,
                       args.warmup)
        b = {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}[name]
        nbytes = (a.n_rows + t.n_kept) * P * b
        res[name] = {"iter_s": round(1.0 / dt, 3), "ms_per_iter": round(dt * 1e3, 3), "kept_events": t.n_kept,
                     "bytes_per_iter": nbytes, "GBs": round(nbytes / dt / 1e9, 1),
                     "hbm_frac": round(nbytes / dt / 1e9 / HBM_PEAK_GBS, 4)}
        del a, t
        torch.cuda.empty_cache()
    del A, y
    torch.cuda.empty_cache()
    # reduced instance: error against the FP64 oracle
    cl_r, A_r, y_r = problem(8192, 10_000, 20260334)
    rows32, rows16 = rows_of(cl_r, torch.float32), rows_of(cl_r, torch.bfloat16)
    imgs = {}
    for name in ("fp32", "bf16", "fp16", "fp8"):
        a, t = stores(A_r, rows32, rows16, name)
        s_j = (alpha * a.column_sums() + (2 - alpha) * t.backproject_ones()).contiguous()
        imgs[name] = joint_osem(a, y_r, t, torch.ones(P, device=dev), s_j, alpha, 10, 4).cpu().numpy()
    cpu, xref = cpu_baseline_jsccsd(A_r.cpu().numpy().astype(np.float64), y_r.cpu().numpy().astype(np.float64), cl_r,
                                    rows32.kept_event_index().cpu().numpy(), 10, 4, alpha, sigma,
                                    [v.numpy() for v in _noise(cl_r.list_origin.shape[0])])
    for name in imgs:
        e = _rel(imgs[name], xref)
        bar = 1e-4 if name == "fp32" else 1e-3
        res[name]["rel_err_vs_fp64_oracle_reduced"] = e
        res[name]["error_bar"] = bar
        res[name]["meets_error_bar"] = e <= bar
    return {"metric": "JSCCSD joint OSEM iter/s (A 81920x16384 + 1e6-event T, BF16 / FP16 / FP8 / FP32 storage)",
            "value": res["bf16"]["iter_s"], "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["bf16"]["ms_per_iter"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16 storage, fp32 accumulate",
            "data": "synthetic (seeded)",
            "config": {"workload": f"JSCCSD joint: A 81920x16384, {n_ev} events on 64x256, alpha 0.5, 1 iteration "
                                   "= 10 subsets"},
            "components": res, "cpu_baseline": cpu}


def cpu_baseline_sparse(coor, packed, A, n_iter, K, eps=1e-12):
    """The numpy oracle (FP64, one thread) on a bounded sample of the packed rows: the materialised-variant
    weight (dense expansion, A-row factor, nan_to_num, row normalisation, clamped ratio back-projection) and
    the safe EM update for n_iter iterations x K subsets; returns the baseline record and its image."""
    sys.path.insert(0, HERE)
    from threadpoolctl import threadpool_limits
    from oracle import osem, sparse
    pj = sparse.build(coor, 1, 2, 20)
    P = coor.shape[0]
    s = np.ones(P)
    x = np.ones(P)
    bounds = osem.chunk_bounds(packed.shape[0], K)
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for _ in range(n_iter):
            for lo, hi in bounds:
                rows, _ = sparse.materialize(packed[lo:hi], A, pj)
                w = osem.nan_to_num(osem.weight_compton(rows, x, eps=eps))
                x = osem.em_safe(x, w, s)
        dt = time.perf_counter() - t0
    per_row_iter = dt / (n_iter * packed.shape[0])
    return {"value": 1e-6 / per_row_iter, "unit": "Mevent-iterations/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle (1 thread): {packed.shape[0]} packed rows, {n_iter} iterations x {K} "
                      f"subsets in {dt:.2f} s (expand, A-row factor, normalisation, weight, safe EM update)",
            "host_cpu": _cpu_model()}, x


def run_sparse(args):
    """Sparse (coarse-T) list-mode OSEM, the reference's production mode (recon_osem_local_sparse_jsccsd_only):
    the 25600-voxel multi-slice grid of the 511keV_RotateNum20 factors (25 rings, 20 slices), theta stride 1,
    z stride 2 (12800 coarse pixels), A 10496 x 25600 FP32, 2e5 packed coarse rows [cpnum1, t_c] expanded on the
    fly by the fused weight kernel (materialised variant: normalised rows, clamped ratio, safe EM update),
    10 iterations x 8 subsets; rows ordered by first-hit crystal within each subset."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, ops, synth
    from jscc_recon_spect_cylinder_amd.compton_sparse_ops import build_compton_sparse_projector
    dev = torch.device("cuda:0")
    lib = _lib.load()
    coor = synth.polar_grid_3d(np.arange(6, 151, 6), [20, 40, 60, 80], np.arange(-28.5, 28.6, 3.0))
    proj = build_compton_sparse_projector(torch.from_numpy(coor.astype(np.float32)), 1, 2, 20).to(dev)
    ax = proj.tables.axis
    P, Pc, M = proj.fine_pixel_num, proj.coarse_pixel_num, 10496
    n = args.events
    n_iter, K = args.iters, args.subsets
    A = _big_sysmat(M, P, 20260335, dev)
    packed = synth.packed_coarse_rows(n, Pc, M, dev, seed=20260335)
    bounds = [(lo, min(n, lo + -(-n // K))) for lo in range(0, n, -(-n // K))]  # torch.chunk sizes
    for lo, hi in bounds:  # first-hit crystal order within each subset (the weight is a sum over rows)
        order = torch.argsort(packed[lo:hi, 0], stable=True)
        packed[lo:hi] = packed[lo:hi][order]
    s = torch.ones(P, device=dev)
    w = torch.zeros(P, device=dev)
    # the weight's form per subset (grouped when runs of one crystal are long), decided once outside the loop
    run_len = [ops.crystal_run_length(packed[lo:hi]) for lo, hi in bounds]
    forms = [bool(ax.za_idx is not None and r >= ops.grouped_min_run(ax)) for r in run_len]

    def run(x):
        for _ in range(n_iter):
            for (lo, hi), form in zip(bounds, forms):
                ops.coarse_weight(packed[lo:hi], ax, A, x, w, normalize=True, eps=1e-12, accumulate=True,
                                  grouped=form)
                ops.em_update(x, w, s, mode=ops.EM_SAFE, zero_after=w)
        return x

    _prof_read(lib, _lib.PROF_COARSE)
    lib.jscc_prof_enable(1)
    dt, x = _timed(lambda: run(torch.ones(P, device=dev)), args.steps, args.warmup)
    lib.jscc_prof_enable(0)
    ops.raise_on_index_error(dev)
    cnt, tot = _prof_read(lib, _lib.PROF_COARSE)
    launches = n_iter * len(bounds)
    cnt_step = cnt / (args.steps + args.warmup)
    assert abs(cnt_step - launches) < 1e-6, (cnt, launches)
    launch_ms = tot / cnt
    # algorithmic bytes of one launch: its packed rows once, and each distinct A row it touches once
    distinct = [int(torch.unique(packed[lo:hi, 0]).numel()) for lo, hi in bounds]
    nbytes = sum((hi - lo) * (Pc + 1) * 4 + d * P * 4 for (lo, hi), d in zip(bounds, distinct)) / len(bounds)
    achieved = nbytes / (launch_ms * 1e-3) / 1e9
    # parity and CPU baseline on a bounded sample (the first 1000 rows)
    ns = 1000
    sample = packed[:ns].contiguous()
    bs = [(lo, min(ns, lo + -(-ns // 2))) for lo in range(0, ns, -(-ns // 2))]
    xg = torch.ones(P, device=dev)
    for _ in range(3):
        for lo, hi in bs:
            ops.coarse_weight(sample[lo:hi], ax, A, xg, w, normalize=True, eps=1e-12, accumulate=True,
                              grouped=forms[0])
            ops.em_update(xg, w, s, mode=ops.EM_SAFE, zero_after=w)
    # the oracle gets only the A rows the sample touches (crystal numbers remapped to them)
    cps, inv_idx = torch.unique(sample[:, 0], return_inverse=True)
    sample_c = sample.clone()
    sample_c[:, 0] = inv_idx.float() + 1.0
    cpu, xref = cpu_baseline_sparse(coor.astype(np.float32), sample_c.cpu().numpy(),
                                    A[cps.long() - 1].cpu().numpy(), 3, 2)
    return {"metric": "sparse list-mode OSEM Mevent-iterations/s (packed coarse rows expanded in the weight pass)",
            "value": round(n * n_iter / dt / 1e6, 3), "unit": "Mevent-iterations/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded): random row-normalised coarse rows, cpnum1 uniform over 10496 crystals, "
                    "A with the single-photon value law",
            "config": {"workload": f"sparse OSEM: {n} packed rows, 25600-voxel 511keV_RotateNum20 grid, theta 1 / "
                                   f"z 2 coarse grid ({Pc} pixels), A {M}x{P} FP32, {n_iter} iterations x {K} "
                                   "subsets"},
            "iter_s": round(n_iter / dt, 3),
            "roofline": {"kernel": ("k_coarse_grouped<25, normalised>" if all(forms) else
                                    "k_coarse_pf<13, normalised> (prefetching per-row form)") + ": one subset",
                         "form": "grouped" if all(forms) else ("per-row" if not any(forms) else "mixed"),
                         "rows_per_crystal_run": round(float(np.mean(run_len)), 2),
                         "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         # the committed counter summary covers the per-row form on 2e5 rows only
                         "traffic": (_traffic("f32", os.path.join(HERE, "profiles",
                                                                  "k_coarse_pf_sparse_2e5x25600_traffic.json"))
                                     if n == 200_000 and not any(forms) else None),
                         "launch_us": round(launch_ms * 1e3, 2), "bytes_per_launch": int(nbytes),
                         "distinct_arows_per_subset": int(np.mean(distinct))},
            "rel_err_vs_fp64_oracle_sample": _rel(xg.cpu().numpy(), xref),
            "cpu_baseline": cpu}


def run_frames(args):
    """K = 16 single-photon data sets sharing one system matrix (the time frames of a dynamic acquisition):
    A 81920 x 16384 with the single-photon value law (the configuration-2 matrix) stored in BF16, Poisson data
    of the rod phantom on the 64 x 256 grid at count levels 1e5-1e7, OSEM with 10 subsets per iteration.  The
    frames' subset products run as GEMMs on the matrix cores (engine.FrameStore: forward over the row-major
    matrix, back-projection over its transpose), beside the same frames one at a time through the fused
    single-frame projection with BF16 rows (SystemStore.osem).  The error is measured per frame against the
    FP64 oracle with the stored matrix on a reduced instance (A 4096 x 16384, 2 frames, 2 iterations x 4
    subsets), which is also the CPU baseline's sample."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, synth
    from jscc_recon_spect_cylinder_amd.engine import FrameStore, SystemStore
    dev = torch.device("cuda:0")
    lib = _lib.load()
    K, M, n_sub = 16, 81920, 10
    coor = synth.polar_grid(64, 256, 50.0)
    P = coor.shape[0]
    x_true = torch.from_numpy(synth.rod_phantom(coor, 50.0).astype(np.float32)).to(dev)

    def data(M_, seed):
        A = _big_sysmat(M_, P, seed, dev)
        f = SystemStore.from_dense(A, "fp32").forward(x_true)
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        levels = torch.logspace(5, 7, K, device=dev)
        Y = torch.poisson(f[:, None] * (levels / f.sum())[None, :], generator=g).contiguous()
        return A, Y

    A, Y = data(M, 20260336)
    fs = FrameStore.from_dense(A)
    ss = SystemStore.from_dense(A, "bf16")
    del A
    torch.cuda.empty_cache()
    s = fs.column_sums()
    X = torch.ones((P, K), device=dev)
    _prof_read(lib, _lib.PROF_FRAMES)
    lib.jscc_prof_enable(1)
    dt, _ = _timed(lambda: fs.osem(X.fill_(1.0), Y, s, 1, n_sub), args.steps, args.warmup)
    lib.jscc_prof_enable(0)
    cnt, tot = _prof_read(lib, _lib.PROF_FRAMES)
    assert cnt == 2 * n_sub * (args.steps + args.warmup), cnt
    gemm_ms = tot / cnt
    x1 = torch.ones(P, device=dev)
    y0 = Y[:, 0].contiguous()
    dt1, _ = _timed(lambda: ss.osem(x1.fill_(1.0), y0, s, 1, n_sub), args.steps, args.warmup)
    # one GEMM launch (forward or back-projection of one subset, their mean): its matrix rows in BF16 once,
    # the frame operand (two BF16 terms, 4 B per element and frame: K x P forward, K x n_s back) and the
    # FP32 result (n_s x K forward, P x K back); 2 MFMA products (hi, lo) of 2 flops per element and frame
    n_s = M // n_sub
    nbytes = n_s * P * 2 + K * (n_s + P) * 4
    flops = 2 * 2 * n_s * P * K
    gbs = nbytes / (gemm_ms * 1e-3) / 1e9
    tfl = flops / (gemm_ms * 1e-3) / 1e12
    del fs, ss, Y, X
    torch.cuda.empty_cache()
    # reduced instance: error per frame against the FP64 oracle with the stored (BF16) matrix; CPU baseline
    sys.path.insert(0, HERE)
    from threadpoolctl import threadpool_limits
    from oracle import osem
    Mr, n_it, ks = 4096, 2, 4
    Ar, Yr = data(Mr, 20260337)
    fr = FrameStore.from_dense(Ar)
    sr = fr.column_sums()
    Xr = torch.ones((P, K), device=dev)
    fr.osem(Xr, Yr, sr, n_it, ks)
    Ab = Ar.to(torch.bfloat16).double().cpu().numpy()
    Yn = Yr.double().cpu().numpy()
    s64 = Ab.sum(0)
    errs = []
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for k in (0, K - 1):
            x = np.ones(P)
            for _ in range(n_it):
                for lo, hi in fr.subset_bounds(ks):
                    x = osem.em_plain(x, osem.weight_single(Ab[lo:hi], Yn[lo:hi, k], x), s64)
            errs.append(_rel(Xr[:, k].cpu().numpy(), x))
        t_cpu = (time.perf_counter() - t0) / (2 * n_it)  # seconds per frame-iteration of the reduced matrix
    cpu_value = 1.0 / (t_cpu * M / Mr)
    return {"metric": "multi-frame SC OSEM frame-iterations/s (16 frames sharing A 81920x16384, BF16, MFMA)",
            "value": round(K / dt, 2), "unit": "frame-iterations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16 matrix and frame terms, fp32 accumulate", "data": "synthetic (seeded)",
            "config": {"workload": f"{K} frames, A {M}x{P} BF16 (single-photon value law), 64x256 rod phantom, "
                                   f"counts 1e5-1e7 per frame, 1 iteration = {n_sub} subsets"},
            "single_frame_fused": {"frame_iterations_s": round(1.0 / dt1, 2), "ms_per_frame_iteration":
                                   round(dt1 * 1e3, 3)},
            "speedup_vs_single_frame": round(dt1 * K / dt, 2),
            "roofline": {"kernel": "k_frames_gemm<1> (v_mfma_f32_16x16x32_bf16): one subset's forward or "
                                   "back-projection GEMM, 16 frames", "bound": "hbm", "achieved": round(gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": None, "launch_us": round(gemm_ms * 1e3, 2), "bytes_per_launch": int(nbytes),
                         "mfma_tflops": round(tfl, 1), "mfma_peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                         "mfma_frac": round(tfl / MFMA_BF16_PEAK_TFLOPS, 4)},
            "rel_err_vs_fp64_oracle_reduced": {"worst_frame": max(errs), "bar": 1e-4,
                                               "sample": f"A {Mr}x{P} (stored BF16 matrix), frames 0 and {K - 1}, "
                                                         f"{n_it} iterations x {ks} subsets"},
            "cpu_baseline": {"value": round(cpu_value, 4), "unit": "frame-iterations/s", "cores": 1, "kind": "port",
                             "sample": f"numpy FP64 oracle (1 thread) on A {Mr}x{P}: {t_cpu:.3f} s per frame-"
                                       f"iteration, scaled by rows to {M}", "host_cpu": _cpu_model()}}


def run_grad(args):
    """Gradient of the list-mode Poisson log-likelihood w.r.t. the image and the angular blur sigma through the
    no-store cone operator (T recomputed in the backward), on the 2e5-event list at sigma = 3 deg."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, ops, synth
    from jscc_recon_spect_cylinder_amd.differentiable import cone_project, list_loglik
    from jscc_recon_spect_cylinder_amd.engine import LazyListStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    dev = torch.device("cuda:0")
    lib = _lib.load()
    sigma0 = float(np.deg2rad(3.0))
    cl = synth.compton_list(n_events=args.events, n_rings=64, n_angles=256)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    rows = build_compton_rows(None, torch.from_numpy(cl.detector).to(dev), torch.from_numpy(cl.coor).to(dev),
                              torch.from_numpy(cl.list_origin).to(dev), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, dev,
                              noise="none", sigma_model="fixed", fixed_sigma=sigma0, store_dtype=None)
    lz = LazyListStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=sigma0)
    x0 = torch.from_numpy((cl.x_true + 0.5).astype(np.float32)).to(dev)

    def grad():
        x = x0.clone().requires_grad_(True)
        sig = torch.tensor(sigma0, dtype=torch.float64, device=dev, requires_grad=True)
        L = list_loglik(cone_project(x, sig, lz))
        L.backward()
        return x.grad, sig.grad

    for _ in range(args.warmup):
        grad()
    torch.cuda.synchronize()
    _prof_read(lib, _lib.PROF_CONE_PROJ)
    lib.jscc_prof_enable(1)  # the timer is per thread: it sees the forward passes (autograd runs the backward
    dt, (gx, gs) = _timed(grad, args.steps, 0)  # on its own device thread)
    lib.jscc_prof_enable(0)
    c, t = _prof_read(lib, _lib.PROF_CONE_PROJ)
    return {"metric": "list-mode log-likelihood gradient (d/dx, d/dsigma) per second, T recomputed",
            "value": round(1.0 / dt, 3), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded)",
            "config": {"workload": f"{lz.n_kept} kept events x {cl.coor.shape[0]} voxels, sigma 3 deg, forward + "
                                   "backward (x and sigma) with T recomputed"},
            "components": {"forward_cone_proj_ms_avg": round(t / max(c, 1), 3),
                           "pairs_per_s_per_pass": round(lz.n_kept * cl.coor.shape[0] / (t / max(c, 1) * 1e-3), 1),
                           "grad_sigma": float(gs), "grad_x_norm": float(gx.norm())}}


def run_batched(args):
    """48 independent reconstructions (64 x 256 polar phantoms with 3-12 random rods of radius 2-8 mm and
    contrast 2-8x; event counts log-uniform in [1e4, 2e6]) reconstructed together: 10 iterations x 8 subsets,
    T recomputed in every pass (it would be ~0.6 TB in BF16), each subset step of all 48 images in one launch,
    iterations replayed from a HIP graph, work-groups per CU autotuned."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, ops, synth
    from jscc_recon_spect_cylinder_amd.engine import BatchedLazyStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows
    dev = torch.device("cuda:0")
    lib = _lib.load()
    rng = np.random.default_rng(20260335)
    B, n_iter, K = 48, 10, 8
    counts = np.exp(rng.uniform(np.log(1e4), np.log(2e6), B)).astype(np.int64)
    coor = synth.polar_grid(64, 256, 50.0).astype(np.float32)
    coor_t = torch.from_numpy(coor).to(dev)
    sigma = float(np.deg2rad(3.0))
    tmax, tmin, tsum = synth.energy_thresholds(synth.E0_511)
    rows = []
    t0 = time.perf_counter()
    for b in range(B):
        x_true = synth.rod_phantom(coor, 50.0, synth.random_rods(rng))
        det, lst = synth.compton_list_device(int(counts[b]), coor, x_true, 64, 256, dev, seed=1000 + b)
        g = torch.Generator(device=dev)
        g.manual_seed(2000 + b)
        n1 = torch.randn(lst.shape[0], generator=g, device=dev)
        n2 = torch.randn(lst.shape[0], generator=g, device=dev)
        rows.append(build_compton_rows(None, det, coor_t, lst, 0.0, 0.0, synth.E0_511, 0.02, tmax, tmin, tsum, dev,
                                       noise=(n1, n2), sigma_model="fixed", fixed_sigma=sigma, store_dtype=None))
        del det, lst, n1, n2
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    bs = BatchedLazyStore.from_rows(rows, sigma_model=ops.SIG_FIXED, fixed_sigma=sigma)
    n_kept = bs.n_kept
    del rows
    P = coor.shape[0]
    s = bs.sensitivity()
    x = torch.ones((B, P), device=dev)
    tile, groups, _ = bs.autotune(x, s, K)
    _prof_read(lib, _lib.PROF_CONE_PROJ)
    gi = bs.capture(x, s, K)
    torch.cuda.synchronize()
    lib.jscc_prof_enable(0)
    t0 = time.perf_counter()
    for _ in range(n_iter):
        gi.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    total = int(sum(n_kept))
    ev_it = total * n_iter / dt
    return {"metric": "batched OSEM events x iterations per second (48 reconstructions, T recomputed)",
            "value": round(ev_it / 1e6, 3), "unit": "Mevent-iterations/s", "n_gpus": 1, "steps": n_iter,
            "warmup": 1, "ms_per_step": round(dt / n_iter * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (rows recomputed in registers)", "data": "synthetic (seeded)",
            "config": {"workload": f"{B} reconstructions, 64x256 polar grid, event counts log-uniform [1e4, 2e6] "
                                   f"(total {int(counts.sum())}, kept {total}), {n_iter} iterations x {K} subsets, "
                                   "one launch per subset step for all images, HIP-graph replay"},
            "components": {"groups_per_cu": groups, "tile": dict(zip(("threads", "voxel_quads", "rows_per_step"),
                                                                     ops.cone_batch_tiles(P)[tile])),
                           "kept_min": int(min(n_kept)), "kept_max": int(max(n_kept)),
                           "pairs_per_s": round(total * P * n_iter / dt, 1),
                           "list_generation_and_filters_s": round(t_prep, 2),
                           "image_min": float(x.min()), "image_max": float(x.max())}}


class _null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _copy_bandwidth(dev, mib=2048, reps=5):
    """Measured HBM bandwidth of a plain device copy of `mib` MiB (bytes read + written per second, GB/s),
    reported beside the nominal 8 TB/s."""
    import torch
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    gbs = 2 * src.numel() * 4 * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return round(gbs, 1)


def _prof_read(lib, tag):
    c = ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    if lib.jscc_prof_read(tag, ctypes.byref(c), ctypes.byref(t)) != 0:
        raise RuntimeError("jscc_prof_read failed")
    return c.value, t.value


COS_CLAMP = float(np.float32(1.0 - 1e-7))


def plain_fp32_cone(rows, coor, sigma, chunk=2048):
    """The cone rows of the selected events as plain FP32 PyTorch operations on the GPU, broadcast over
    chunks of events (the form a torch implementation of the cone backprojection takes on a GPU):
    t = exp(-(beta - theta)^2 / (2 sigma^2)) * (kn - sin^2 beta) with beta = acos(clamp(v01 . v12 /
    (|v01| |v12|))), written to an [n_sel, P] FP32 matrix together with the row sums.  Returns (seconds
    for one pass over all selected events, the rows)."""
    import torch
    ev = rows.events[rows.sel_idx[: rows.n_sel].long()]
    p1, a, th, kn = ev[:, 0:3], ev[:, 3:6], ev[:, 7:8], ev[:, 8:9]
    C = coor.reshape(-1, 3).float()
    out = torch.empty((rows.n_sel, C.shape[0]), dtype=torch.float32, device=C.device)
    sums = torch.empty(rows.n_sel, dtype=torch.float32, device=C.device)

    def run():
        for i in range(0, rows.n_sel, chunk):
            j = min(i + chunk, rows.n_sel)
            v = p1[i:j, None, :] - C[None, :, :]
            d12 = torch.linalg.vector_norm(a[i:j], dim=1, keepdim=True)
            cosb = (v * a[i:j, None, :]).sum(2) / torch.clamp(torch.linalg.vector_norm(v, dim=2) * d12, min=1e-7)
            beta = torch.acos(torch.clamp(cosb, -COS_CLAMP, COS_CLAMP))
            t = torch.exp(-(beta - th[i:j]) ** 2 / (2.0 * sigma * sigma)) * (kn[i:j] - torch.sin(beta) ** 2)
            out[i:j] = t
            sums[i:j] = t.sum(1)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def _traffic(dtype_name, path=TRAFFIC_FILE):
    """HBM bytes per subset launch from the committed rocprofv3 counter summary (None when absent)."""
    try:
        with open(path) as f:
            rec = json.load(f)[dtype_name]
    except (OSError, KeyError, ValueError):
        return None
    kib = rec["fetch_kib"] * rec["fetch_factor"] + rec["write_kib"] * rec["write_factor"]
    return kib * 1024.0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the list shards by events with one all-reduce of the image weight per subset).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.workload != "jsccd":
        print(json.dumps(OTHER_WORKLOADS[args.workload](args)))
        return
    out = run_jsccd(args)
    if not args.headline_only:
        out["workloads"] = other_workloads(args)
    print(json.dumps(out))


def _summary(res):
    """The figures of a workload's line kept in the default line: metric, value, timing, roofline, error terms and
    the CPU baseline's value (the full line comes from bench.py --workload NAME)."""
    keep = {k: res[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "warmup", "dtype", "roofline",
                                "speedup_vs_single_frame", "single_frame_fused") if k in res}
    keep["config"] = res.get("config", {}).get("workload")
    keep.update({k: v for k, v in res.items() if k.startswith("rel_err")})
    if "components" in res and isinstance(res["components"], dict):
        per = ("iter_s", "ms_per_iter", "hbm_frac", "rel_err_vs_fp64_oracle_reduced", "meets_error_bar")
        comp = {}
        for n, c in res["components"].items():
            if isinstance(c, dict):  # per storage type: the figures above; a launch shape: as it is
                comp[n] = {k: v for k, v in c.items() if k in per} or c
            elif n in ("groups_per_cu", "kept_min", "kept_max"):
                comp[n] = c
        keep["components"] = comp
    cb = res.get("cpu_baseline")
    if isinstance(cb, dict):
        keep["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind") if k in cb}
    return keep


def other_workloads(args):
    """Every other workload of this file, run after the headline in the same process, one summary each (a
    failure is recorded, not raised: the headline line is printed whatever happens here)."""
    import torch
    out = {}
    for name, fn in OTHER_WORKLOADS.items():
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        try:
            out[name] = _summary(fn(args))
        except Exception as e:  # noqa: BLE001 -- reported in the line
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["wall_s"] = round(time.perf_counter() - t0, 1)
    return out


def run_jsccd(args):
    """The headline workload (the metric's configuration): 2e5-event JSCCD list, 64 x 256 grid, sigma 3 deg, T
    built on the GPU, 10 iterations x 8 subsets; BF16, FP32 and FP16 storage."""
    import torch
    from jscc_recon_spect_cylinder_amd import _lib, synth
    from jscc_recon_spect_cylinder_amd.engine import ListModeStore
    from jscc_recon_spect_cylinder_amd.process_list_plane_strict import build_compton_rows

    dev = torch.device("cuda:0")
    lib = _lib.load()
    sigma = float(np.deg2rad(3.0))
    cl = synth.compton_list(n_events=args.events, n_rings=64, n_angles=256)
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    det = torch.from_numpy(cl.detector).to(dev)
    coor = torch.from_numpy(cl.coor).to(dev)
    lst0 = torch.from_numpy(cl.list_origin).to(dev)
    g = torch.Generator().manual_seed(20260332)
    n1_host = torch.randn(args.events, generator=g)
    n2_host = torch.randn(args.events, generator=g)
    n1, n2 = n1_host.to(dev), n2_host.to(dev)
    P = cl.coor.shape[0]

    last_rows = [None]

    def step(store_dtype, name):
        # the selected and kept counts stay on the device (sync_counts=False), so the step has no host
        # synchronisation; the sensitivity (column sums of the kept, normalised rows) is formed by the cone pass
        # itself (want_sens), from the rows as stored, instead of a second pass over T; FP16 rows are quantised
        # by the same pass (row maximum at 2^14, right after each group's statistics) and their sensitivity is
        # the column sums of the codes the OSEM projects
        rows = build_compton_rows(None, det, coor, lst0.clone(), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, dev,
                                  noise=(n1, n2), sigma_model="fixed", fixed_sigma=sigma, store_dtype=store_dtype,
                                  sync_counts=False, want_sens=True)
        last_rows[0] = rows
        st = ListModeStore.from_rows(rows, name)
        s = rows.sens
        x = torch.ones_like(s)
        st.osem(x, s, args.iters, args.subsets)
        return st, x

    results = {}
    for name, sd in (("bf16", torch.bfloat16), ("fp32", torch.float32), ("fp16", torch.float16)):
        if name != "bf16" and args.skip_fp32:
            continue
        for _ in range(args.warmup):
            step(sd, name)
        torch.cuda.synchronize()
        for tag in (_lib.PROF_CONE, _lib.PROF_PROJ, _lib.PROF_PROJ_SUM):
            _prof_read(lib, tag)  # drop pairs of the warm-up
        lib.jscc_prof_enable(1)
        st = x = None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            # the previous step's T is released before the next is built (one T in HBM, as in a
            # reconstruction), so its memory is reused instead of a second allocation of the same size
            st = x = None
            st, x = step(sd, name)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lib.jscc_prof_enable(0)
        ccnt, ctot = _prof_read(lib, _lib.PROF_CONE)
        pcnt, ptot = _prof_read(lib, _lib.PROF_PROJ)
        scnt, stot = _prof_read(lib, _lib.PROF_PROJ_SUM)
        ms_step = (t1 - t0) * 1e3 / args.steps
        n_sel = [last_rows[0].resolve_counts().n_sel]  # counts (and the index check) of the last step
        st.resolve()
        # the OSEM part alone (T already built): iterations per second
        s = st.backproject_ones()
        xi = torch.ones_like(s)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        st.osem(xi, s, args.iters, args.subsets)
        torch.cuda.synchronize()
        osem_s = (time.perf_counter() - t2) / args.iters
        # the same iterations replayed from one captured iteration (ListModeStore.capture)
        xg = torch.ones_like(s)
        gi = st.capture(xg, s, args.subsets)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(args.iters):
            gi.replay()
        torch.cuda.synchronize()
        osem_graph_s = (time.perf_counter() - t3) / args.iters
        graph_equal = bool(torch.equal(xg, xi))
        del gi, xg
        b = {"bf16": 2, "fp32": 4, "fp16": 2}[name]
        # algorithmic bytes of the OSEM launches of one step: every kept row read once per iteration
        # (+ the image read and the weight written per subset), divided over the subset launches
        osem_bytes = args.iters * (st.n_kept * P * b + args.subsets * 2 * P * 4)
        launches = pcnt / args.steps
        subset_bytes = osem_bytes / launches
        proj_ms = ptot / max(pcnt, 1)
        results[name] = {
            "ms_per_step": ms_step,
            "img": x.cpu().numpy(),
            "n_kept": st.n_kept,
            "cone_ms": ctot / max(ccnt, 1),
            "cone_Mevents_s": args.events / (ctot / max(ccnt, 1) * 1e-3) / 1e6,
            "cone_rows": n_sel[0],
            "cone_pairs_s": n_sel[0] * P / (ctot / max(ccnt, 1) * 1e-3),
            "osem_launch_ms": proj_ms,
            "osem_launches_per_step": launches,
            "osem_bytes_per_launch": subset_bytes,
            "osem_GBs": subset_bytes / (proj_ms * 1e-3) / 1e9,
            "sensitivity": "formed in the cone pass (want_sens)",
            "osem_iter_s": 1.0 / osem_s,
            "osem_graph_iter_s": 1.0 / osem_graph_s,
            "osem_graph_equals_eager": graph_equal,
        }
        del st, x

    bf = results["bf16"]
    rel_err = None
    if "fp32" in results:
        f = results["fp32"]["img"]
        rel_err = float(np.linalg.norm(bf["img"] - f) / np.linalg.norm(f))
        results["fp16"]["rel_err_vs_fp32_image"] = float(np.linalg.norm(results["fp16"]["img"] - f) / np.linalg.norm(f))
    traffic = _traffic("bf16")
    roof = {"kernel": "k_proj_fused<bf16,2,2,ratio>: one OSEM subset pass T_s^T(1/T_s x)", "bound": "hbm",
            "achieved": round(bf["osem_GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bf["osem_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else int(traffic),
            "launch_ms": round(bf["osem_launch_ms"], 4), "bytes_per_launch": int(bf["osem_bytes_per_launch"])}
    value = args.events / (bf["ms_per_step"] * 1e-3) / 1e6
    # the cone pass (rows, statistics and the sensitivity's read-back and column sums) is VALU-bound: its
    # instruction count per (event, voxel) pair on this workload, from SQ_INSTS_VALU x 64 / pairs of
    # k_cone_rows_sens<fixed sigma, no A factor, BF16> (47.6, 5 of them sqrt / rsq / rcp / exp), times the
    # measured pair rate, against the non-packed issue rate at the 2.4 GHz engine clock (one wave64 instruction
    # per 4 cycles per SIMD).  Under load the pass is power-limited to ~2.15-2.3 GHz (GRBM_GUI_ACTIVE per
    # dispatch), which caps the fraction below 1.
    cone_roof = {"kernel": "k_cone_rows_sens<fixed sigma, BF16>: rows + statistics + sensitivity", "bound": "valu",
                 "valu_instr_per_pair": 47.6, "transcendental_per_pair": 5,
                 "pairs_s": round(bf["cone_pairs_s"], -8),
                 "valu_lane_instr_s": round(bf["cone_pairs_s"] * 47.6, -10),
                 "peak_lane_instr_s": FP32_LANE_PEAK,
                 "frac": round(bf["cone_pairs_s"] * 47.6 / FP32_LANE_PEAK, 4),
                 "clock_GHz_profiled": 2.15}
    comp = {}
    for k, r in results.items():
        comp[k] = {kk: (float(f"{v:.5g}") if isinstance(v, float) else v) for kk, v in r.items() if kk != "img"}
    out = {
        "metric": "Compton backproj Mevents/s + OSEM iter/s, BF16 vs FP32 path, rel err<=1e-3",
        "value": round(value, 4),
        "unit": "Mevents/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bf["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 storage, fp32 accumulate",
        "data": "synthetic (seeded rod phantom, Klein-Nishina list; no measured factors exist for this geometry)",
        "config": {"workload": f"JSCCD list-mode: {args.events} events, 64x256 polar grid ({P} voxels), sigma 3 deg, "
                               f"T built on the GPU, {args.iters} iterations x {args.subsets} subsets",
                   "events": args.events, "voxels": P, "iterations": args.iters, "subsets": args.subsets},
        "components": comp,
        "rel_err_bf16_vs_fp32_image": rel_err,
        "roofline": roof,
        "cone_roofline": cone_roof,
        "hbm_copy_GBs": _copy_bandwidth(dev),
    }
    if not args.skip_fp32:
        # the same rows as plain FP32 torch operations on the GPU, against the FP32 rows of the cone kernel
        rows32 = build_compton_rows(None, det, coor, lst0.clone(), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, dev,
                                    noise=(n1, n2), sigma_model="fixed", fixed_sigma=sigma,
                                    store_dtype=torch.float32)
        sec, plain = plain_fp32_cone(rows32, coor, sigma)
        k = min(512, rows32.n_sel)
        ref = rows32.store[:k, :P]
        rel_rows = float(torch.linalg.vector_norm(plain[:k] - ref) / torch.linalg.vector_norm(ref))
        plain_mev = rows32.n_sel / sec / 1e6
        out["plain_fp32_gpu"] = {
            "form": "cone rows + row sums as plain FP32 torch operations on the GPU (chunks of 2048 events)",
            "cone_ms": round(sec * 1e3, 2), "cone_Mevents_s": round(plain_mev, 3),
            "bf16_kernel_speedup": round(bf["cone_Mevents_s"] / plain_mev, 2),
            "rel_l2_rows_vs_fp32_kernel": rel_rows}
        del plain, rows32
    if args.cpu_sample > 0:
        # the metric's error term: the GPU images of the first cpu_sample events (same noise) against the FP64
        # oracle's, then the reference's CPU path timed on the host (checked against the same oracle image)
        n_s = min(args.cpu_sample, args.events)
        kept_ref, x_ref = fp64_oracle_sample(cl, n1_host.numpy(), n2_host.numpy(), n_s, args.iters, args.subsets,
                                             sigma)
        errs = {}
        for name, sd in (("bf16", torch.bfloat16), ("fp16", torch.float16), ("fp32", torch.float32)):
            rows_s = build_compton_rows(None, det, coor, lst0[:n_s].clone(), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum,
                                        dev, noise=(n1[:n_s], n2[:n_s]), sigma_model="fixed", fixed_sigma=sigma,
                                        store_dtype=sd, sync_counts=False, want_sens=True)
            kept_ok = bool(np.array_equal(rows_s.kept_event_index().cpu().numpy(), kept_ref))
            st = ListModeStore.from_rows(rows_s, name)
            s_s = rows_s.sens
            x_s = st.osem(torch.ones_like(s_s), s_s, args.iters, args.subsets)
            e = _rel(x_s.cpu().numpy(), x_ref)
            errs[name] = {"rel_l2": e, "bar": 1e-4 if name == "fp32" else 1e-3,
                          "meets_bar": e <= (1e-4 if name == "fp32" else 1e-3), "kept_rows_bit_exact": kept_ok}
            del st, rows_s
        out["rel_err_vs_fp64_oracle"] = dict(errs, sample=f"first {n_s} events of the list ({len(kept_ref)} kept "
                                                           f"rows), same noise, {args.iters} x {args.subsets} OSEM")
        out["cpu_baseline"] = cpu_baseline(cl, n1_host, n2_host, n_s, max(n_s // 4, 1000), args.iters,
                                           args.subsets, sigma, x_ref)
    return out


OTHER_WORKLOADS = {"sc": run_sc, "jsccsd": run_jsccsd, "grad": run_grad, "batched": run_batched,
                   "sparse": run_sparse, "frames": run_frames}


if __name__ == "__main__":
    main()
