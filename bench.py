#!/usr/bin/env python3
"""Benchmark: voxel-iterations/s of the registration solver on one MI355X.

Workload: the plain FP32 method on the brain-sized phantom, 182 x 218 x 182 voxels, the
reference's params.ini weights (alpha 0.004, tau 0.2, w_reg 0.2, rho_0 1, p 2, C 0.25), 32 labels.
A "step" is one solver iteration over the whole volume; `--steps` iterations (default 300, the
workload's iteration count) are timed with HIP events around the iteration loop, after
`--warmup` untimed iterations on a separate context.  Inputs are resident in HBM before timing.

Printed: one JSON line with the metric, the roofline of the dominant kernel (algorithmic bytes
per launch / its average launch time from HIP events) and a CPU baseline (the oracle, the
reference's kernels restated in C and run on this host's cores over a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
PROFILE = os.path.join(ROOT, "profiles", "solver_iteration_182x218x182.md")


def profiled_traffic():
    """HBM bytes per launch of each pass from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (its `traffic_bytes_per_launch:` line), or None."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return {k: int(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--cpu-iters", type=int, default=120,
                    help="oracle iterations for the CPU baseline sample (about 10-20 s on 16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--zc", default="0,0,0",
                    help="interior z-slices per block of pass L, pass A, pass B (0 = automatic)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before timing, run one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this solver runs one registration per GPU and this build measures one GPU; "
              f"--gpus {args.gpus} is not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    zc = tuple(int(v) for v in args.zc.split(","))
    import numpy as np
    import fastreg_amd as fr
    from fastreg_amd import phantoms

    t0 = time.time()
    pair = phantoms.brain()
    dims = pair.dims
    X, Y, Z = dims
    n = X * Y * Z
    n_int = (X - 2) * (Y - 2) * (Z - 2)
    print(f"[bench] phantom {dims} ready in {time.time() - t0:.1f}s; device {fr.device_info()}", file=sys.stderr,
          flush=True)

    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        src, dst = fr.DeviceArray(nb, np.uint8), fr.DeviceArray(nb, np.uint8)
        fr.check(fr.lib().fr_dev_memset(src.ptr, 1, nb))
        fr.check(fr.lib().fr_memcpy_dtod(dst.ptr, src.ptr, nb))
        fr.check(fr.lib().fr_dev_sync())
        src.free()
        dst.free()

    # warm-up on its own context
    pw = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=max(args.warmup, 1))
    w = fr.Solver(pw)
    w.set_chunking(*zc)
    w.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    w.run()
    w.stats()
    w.close()

    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=args.steps)
    s = fr.Solver(p)
    s.set_chunking(*zc)
    chunking = s.chunking()
    s.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    fr.check(fr.lib().fr_dev_sync())
    tw0 = time.perf_counter()
    s.run()
    st = s.stats()                      # synchronises
    wall_ms = (time.perf_counter() - tw0) * 1e3
    ms = st.solver_ms
    assert st.iterations == args.steps, st
    value = n * args.steps / (ms / 1e3)
    print(f"[bench] {args.steps} iterations: {ms:.2f} ms (events), {wall_ms:.2f} ms (wall)", file=sys.stderr,
          flush=True)

    # per-kernel durations: HIP events between the launches of iterations run in sequence on the
    # context stream (the caches as the iteration loop leaves them)
    ms_l, ms_a, ms_b = s.time_kernels(args.kernel_reps)
    # algorithmic bytes per interior voxel: pass L reads psi 12 and writes L 12; pass A reads
    # rho 4, v 12, psi_inv 12, S 4, L 12 and T 4 (gathered at psi_inv) and writes rho 4, v 12,
    # psi_inv 12; pass B reads psi 12 and v 12 (gathered at psi) and writes psi 12
    bytes_l = 24 * n_int
    bytes_a = 76 * n_int
    bytes_b = 36 * n_int
    kern = {"pass_l": (ms_l, bytes_l), "pass_a": (ms_a, bytes_a), "pass_b": (ms_b, bytes_b)}
    dom = max(kern, key=lambda k: kern[k][0])
    dms, dbytes = kern[dom]
    achieved = dbytes / (dms / 1e3) / 1e9
    prof = profiled_traffic()
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": prof.get(dom) if prof else None,
                "traffic_source": os.path.relpath(PROFILE, ROOT) if prof else None,
                "kernel_ms": {"pass_l": round(ms_l, 4), "pass_a": round(ms_a, 4), "pass_b": round(ms_b, 4)},
                "alg_bytes_per_launch": {"pass_l": bytes_l, "pass_a": bytes_a, "pass_b": bytes_b},
                "iteration_frac_88B": round(value * 88 / (HBM_PEAK_GBS * 1e9), 4)}
    fin = s.finalize(want_outputs=False)
    st2 = s.stats()
    s.close()

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        o = O.Oracle(dims, "f32")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        tc = time.perf_counter()
        o.run(O.SolverParams(max_iter=args.cpu_iters, no_labels=pair.no_labels))
        cs = time.perf_counter() - tc
        o.close()
        cpu = {"value": round(n * args.cpu_iters / cs, 1), "unit": "voxel-iterations/s", "cores": O.num_threads(),
               "kind": "port",
               "sample": f"oracle FP32 reference-order restatement, {args.cpu_iters} iterations of the same "
                         f"182x218x182 workload, OpenMP over z ({cs:.1f} s)"}

    out = {
        "metric": "voxel-iterations/s at 182x218x182 fp32",
        "value": round(value, 1),
        "unit": "voxel-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ellipsoid-shell phantom with 32 Voronoi labels; source warped by an RBF "
                "velocity exponential)",
        "config": {"workload": "plain solver, 182x218x182 fp32, params.ini weights", "dims": list(dims),
                   "iterations": args.steps, "voxels": n, "dice_all": round(st2.dice_all, 6),
                   "min_det_true": st2.min_det_true, "wall_ms": round(wall_ms, 2),
                   "z_chunk": {"pass_l": chunking[0], "pass_a": chunking[1], "pass_b": chunking[2]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
