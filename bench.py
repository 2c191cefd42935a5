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
    ap.add_argument("--skip-fp32", action="store_true")
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


def cpu_baseline(cl, n_sample, n_iter, n_sub, sigma, chunk=1000):
    """The numpy oracle (FP64, one thread) on a bounded sample of the same list: T built in chunks of
    `chunk` events, then one OSEM iteration over the sample's rows; extrapolated linearly per event."""
    sys.path.insert(0, HERE)
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from oracle import cone, osem
    from jscc_recon_spect_cylinder_amd import synth
    tmax, tmin, tsum = synth.energy_thresholds(cl.e0)
    with (threadpool_limits(limits=1) if threadpool_limits else _null()):
        blocks = []
        t_build = 0.0
        for lo in range(0, n_sample, chunk):
            lst = cl.list_origin[lo:min(lo + chunk, n_sample)]
            t0 = time.perf_counter()
            out = cone.backproj_list_single(None, cl.detector, cl.coor, lst, 0.0, 0.0, cl.e0, 0.02, tmax, tmin,
                                            tsum, None, None, fixed_sigma=sigma)
            t_build += time.perf_counter() - t0
            blocks.append(out["t"])
        T = np.concatenate(blocks)
        del blocks
        s = T.sum(0)
        x = np.ones(T.shape[1])
        bounds = osem.split_bounds_even(T.shape[0], n_sub)
        t0 = time.perf_counter()
        for lo, hi in bounds:
            w = osem.weight_compton(T[lo:hi], x)
            x = osem.em_plain(x, w, s)
        t_iter = time.perf_counter() - t0
    per_event = (t_build + n_iter * t_iter) / n_sample
    return {"value": 1e-6 / per_event, "unit": "Mevents/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle (1 thread), the first {n_sample} events of the same list: T build "
                      f"{t_build:.2f} s, 1 OSEM iteration ({n_sub} subsets) {t_iter:.3f} s, extrapolated linearly "
                      f"to T build + {n_iter} iterations per event", "host_cpu": _cpu_model()}


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


def _prof_read(lib, tag):
    c = ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    if lib.jscc_prof_read(tag, ctypes.byref(c), ctypes.byref(t)) != 0:
        raise RuntimeError("jscc_prof_read failed")
    return c.value, t.value


def _traffic(dtype_name):
    """HBM bytes per subset launch from the committed rocprofv3 counter summary (None when absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
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
    n1 = torch.randn(args.events, generator=g).to(dev)
    n2 = torch.randn(args.events, generator=g).to(dev)
    P = cl.coor.shape[0]

    def step(store_dtype, name):
        rows = build_compton_rows(None, det, coor, lst0.clone(), 0.0, 0.0, cl.e0, 0.02, tmax, tmin, tsum, dev,
                                  noise=(n1, n2), sigma_model="fixed", fixed_sigma=sigma, store_dtype=store_dtype)
        st = ListModeStore.from_rows(rows, name)
        s = st.backproject_ones()
        x = torch.ones_like(s)
        st.osem(x, s, args.iters, args.subsets)
        return st, x

    results = {}
    for name, sd in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        if name == "fp32" and args.skip_fp32:
            continue
        for _ in range(args.warmup):
            step(sd, name)
        torch.cuda.synchronize()
        for tag in (_lib.PROF_CONE, _lib.PROF_PROJ, _lib.PROF_PROJ_SUM):
            _prof_read(lib, tag)  # drop pairs of the warm-up
        lib.jscc_prof_enable(1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st, x = step(sd, name)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lib.jscc_prof_enable(0)
        ccnt, ctot = _prof_read(lib, _lib.PROF_CONE)
        pcnt, ptot = _prof_read(lib, _lib.PROF_PROJ)
        scnt, stot = _prof_read(lib, _lib.PROF_PROJ_SUM)
        ms_step = (t1 - t0) * 1e3 / args.steps
        # the OSEM part alone (T already built): iterations per second
        s = st.backproject_ones()
        xi = torch.ones_like(s)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        st.osem(xi, s, args.iters, args.subsets)
        torch.cuda.synchronize()
        osem_s = (time.perf_counter() - t2) / args.iters
        b = {"bf16": 2, "fp32": 4}[name]
        # algorithmic bytes of one subset launch: its rows read once (+ x read, w written)
        subset_bytes = st.n_kept / args.subsets * P * b + 2 * P * 4
        proj_ms = ptot / max(pcnt, 1)
        results[name] = {
            "ms_per_step": ms_step,
            "img": x.cpu().numpy(),
            "n_kept": st.n_kept,
            "cone_ms": ctot / max(ccnt, 1),
            "cone_Mevents_s": args.events / (ctot / max(ccnt, 1) * 1e-3) / 1e6,
            "subset_launch_ms": proj_ms,
            "subset_launches_per_step": pcnt / args.steps,
            "subset_bytes": subset_bytes,
            "subset_GBs": subset_bytes / (proj_ms * 1e-3) / 1e9,
            "colsum_ms": stot / max(scnt, 1),
            "colsum_GBs": st.n_kept * P * b / (stot / max(scnt, 1) * 1e-3) / 1e9,
            "osem_iter_s": 1.0 / osem_s,
        }
        del st, x

    bf = results["bf16"]
    rel_err = None
    if "fp32" in results:
        f = results["fp32"]["img"]
        rel_err = float(np.linalg.norm(bf["img"] - f) / np.linalg.norm(f))
    traffic = _traffic("bf16")
    roof = {"kernel": "k_proj_fused<bf16,2,4,ratio>: one OSEM subset pass T_s^T(1/T_s x)", "bound": "hbm",
            "achieved": round(bf["subset_GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bf["subset_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else int(traffic),
            "launch_ms": round(bf["subset_launch_ms"], 4), "bytes_per_launch": int(bf["subset_bytes"])}
    value = args.events / (bf["ms_per_step"] * 1e-3) / 1e6
    comp = {}
    for k, r in results.items():
        comp[k] = {kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in r.items() if kk != "img"}
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
    }
    if args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(cl, args.cpu_sample, args.iters, args.subsets, sigma)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
