#!/usr/bin/env python3
"""Benchmark: voxel-iterations/s of the registration solver on one MI355X.

Workload (default, `--workload plain`): the plain FP32 method on the brain-sized phantom,
182 x 218 x 182 voxels, 300 iterations, the reference's params.ini weights (alpha 0.004, tau 0.2,
w_reg 0.2, rho_0 1, p 2, C 0.25), 32 labels.  A "step" is one whole registration of that pair:
a fresh state from the inputs (already resident in HBM), the 300-iteration loop (a replayed
hipGraph), and the finalisation (both warps into output buffers, Dice and min det J), i.e. what
FastReg::operator() does on the device.  `--warmup` registrations run untimed, then `--steps`
registrations are timed on the host clock between device synchronisations; `value` is voxels x
300 x steps over that time.  The loop alone is also timed with HIP events (`loop_*`).

`--workload c1`: the same on the 64^3 blob pair, 200 iterations per registration (9 labels).

Printed: one JSON line with the metric, the roofline of the dominant kernel (algorithmic bytes
per launch / its average launch time from HIP events) and a CPU baseline (the oracle, the
reference's kernels restated in C and run on this host's cores: the median of 3 runs over a
bounded sample).
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


def host_cpu():
    """The host's CPU model and logical CPU count (what the CPU baseline ran on)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_baseline_sample(pair, iters, runs=3):
    """The oracle (FP32, the reference's operation order, OpenMP over z) on `iters` iterations of the
    pair, `runs` times: voxel-iterations/s of the median run."""
    import numpy as np
    from oracle import oracle as O

    O.build()
    dims = pair.dims
    n = dims[0] * dims[1] * dims[2]
    times = []
    for _ in range(runs):
        o = O.Oracle(dims, "f32")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        tc = time.perf_counter()
        o.run(O.SolverParams(max_iter=iters, no_labels=pair.no_labels))
        times.append(time.perf_counter() - tc)
        o.close()
    med = float(np.median(times))
    return {"value": round(n * iters / med, 1), "unit": "voxel-iterations/s", "cores": O.num_threads(),
            "kind": "port", "runs_s": [round(t, 3) for t in times], "host": host_cpu(),
            "sample": f"oracle FP32 reference-order restatement, {iters} iterations of the same {dims[0]}x{dims[1]}x"
                      f"{dims[2]} pair per run, median of {runs} runs, OpenMP over z"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed registrations (plain, c1)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed registrations before them")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--cpu-iters", type=int, default=40,
                    help="oracle iterations per CPU-baseline run at 182x218x182 (3 runs, about 4 s each on "
                         "16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--arith", choices=("reference", "fma"), default="reference",
                    help="plain, c1: arithmetic of pass A; fma times the contracted form and compares its "
                         "registration with a reference-order one (Dice, min det J, labels, fields)")
    ap.add_argument("--zc", default="0,0",
                    help="interior z-slices per block of pass A, pass B (0 = automatic)")
    ap.add_argument("--workload", choices=("plain", "c1", "batch", "precision", "adjoint"), default="plain",
                    help="plain: the single-pair headline metric at 182x218x182 (default); c1: the same on the "
                         "64^3 blob pair, 200 iterations; batch: 64 pairs of uneven sizes, "
                         "150 iterations each, batched form vs one pair at a time; precision: FP32, BF16 and "
                         "FP16 storage of v and rho against an FP64 oracle run (64^3, 300 iterations) and their "
                         "throughput at 182x218x182; adjoint: the differentiated solver at 128^3, K unrolled "
                         "iterations with checkpoints (gradient check, time ratio, device memory)")
    ap.add_argument("--adj-k", type=int, default=64, help="adjoint workload: unrolled iterations")
    ap.add_argument("--adj-ckpt", type=int, default=8, help="adjoint workload: checkpoint interval")
    ap.add_argument("--adj-dirs", type=int, default=32, help="adjoint workload: finite-difference directions")
    ap.add_argument("--cpu-adj-k", type=int, default=2,
                    help="adjoint workload: unrolled iterations of the CPU reference-path gradient sample")
    ap.add_argument("--batch-pairs", type=int, default=64)
    ap.add_argument("--batch-iters", type=int, default=150)
    ap.add_argument("--batch-sizes", choices=("mixed", "small"), default="mixed",
                    help="mixed: uniform over the five sizes; small: uniform over 48^3 and 96^3")
    ap.add_argument("--batch-storage", choices=("fp32", "bf16", "fp16"), default="fp32",
                    help="storage of v and rho in the batched leg (the one-at-a-time leg is the plain FP32 form); "
                         "with bf16 / fp16 the Dice of every pair is compared with the FP32 form's and, for the "
                         "pairs of up to 96^3 voxels, with an FP64 oracle run")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before timing, run one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this solver runs one registration per GPU and this build measures one GPU; "
              f"--gpus {args.gpus} is not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "precision":
        return precision_workload(args)
    if args.workload == "adjoint":
        return adjoint_workload(args)
    return registration_workload(args)


def registration_workload(args):
    """The plain (182x218x182, 300 iterations) or C1 (64^3, 200 iterations) registration, `--steps`
    registrations timed after `--warmup` untimed ones, inputs resident in HBM."""
    import ctypes as C
    import numpy as np
    import fastreg_amd as fr
    from fastreg_amd import phantoms

    zc = tuple(int(v) for v in args.zc.split(","))
    c1 = args.workload == "c1"
    iters = 200 if c1 else 300
    t0 = time.time()
    pair = phantoms.blobs() if c1 else phantoms.brain()
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

    L = fr.lib()
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=iters)
    s = fr.Solver(p)
    s.set_chunking(*zc)
    chunking = s.chunking()
    if args.arith == "fma":
        s.set_arith(fr.ARITH_FMA)
    s.init(pair.target, pair.target_labels, pair.source, pair.source_labels)  # inputs now resident
    outs = [fr.DeviceArray(n, np.float32), fr.DeviceArray(n, np.int16), fr.DeviceArray(n, np.float32),
            fr.DeviceArray(n, np.int16)]
    st = fr.FrStats()

    def registration():
        """one registration from the resident inputs; returns its loop time from HIP events (ms)"""
        s.reinit()
        fr.check(L.fr_solver_run(s.h))
        fr.check(L.fr_solver_finalize(s.h, *[b.ptr for b in outs]))  # waits for the outputs
        fr.check(L.fr_get_stats(s.h, C.byref(st)))
        assert st.iterations == iters, st.iterations
        return st.solver_ms

    for _ in range(max(args.warmup, 0)):
        registration()
    fr.check(L.fr_dev_sync())
    tw0 = time.perf_counter()
    loop_ms = [registration() for _ in range(args.steps)]
    fr.check(L.fr_dev_sync())
    wall_s = time.perf_counter() - tw0
    ms_step = 1e3 * wall_s / args.steps
    value = n * iters * args.steps / wall_s
    loop_med = float(np.median(loop_ms))
    print(f"[bench] {args.steps} registrations of {iters} iterations: {ms_step:.2f} ms each (wall), loop "
          f"{loop_med:.2f} ms (events, median)", file=sys.stderr, flush=True)
    dice, min_det = st.dice_all, st.min_det_true

    # per-kernel durations: HIP events between the launches of iterations run in sequence on the
    # context stream (the caches as the iteration loop leaves them)
    ms_a, ms_b = s.time_kernels(args.kernel_reps)
    # algorithmic bytes per interior voxel: pass A reads rho 4, v 12, psi_inv 12, S 4, psi 12 (for
    # -Laplacian(psi), formed in the kernel) and T 4 (sampled at psi_inv) and writes rho 4, v 12,
    # psi_inv 12; pass B reads psi 12 and v 12 (gathered at psi) and writes psi 12
    bytes_a = 76 * n_int
    bytes_b = 36 * n_int
    kern = {"pass_a": (ms_a, bytes_a), "pass_b": (ms_b, bytes_b)}
    dom = max(kern, key=lambda k: kern[k][0])
    dms, dbytes = kern[dom]
    achieved = dbytes / (dms / 1e3) / 1e9
    prof = None if c1 else profiled_traffic()
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": prof.get(dom) if prof else None,
                "traffic_source": os.path.relpath(PROFILE, ROOT) if prof else None,
                "kernel_ms": {"pass_a": round(ms_a, 4), "pass_b": round(ms_b, 4)},
                "alg_bytes_per_launch": {"pass_a": bytes_a, "pass_b": bytes_b},
                "iteration_frac_88B": round(n * 1e3 / (loop_med / iters) * 88 / (HBM_PEAK_GBS * 1e9), 4)}
    vs_ref = arith_vs_reference(s, registration, outs, st, n) if args.arith == "fma" else None
    for b in outs:
        b.free()
    s.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_sample(pair, iters if c1 else args.cpu_iters)

    name = f"{X}x{Y}x{Z}"
    out = {
        "metric": f"voxel-iterations/s at {name} fp32",
        "value": round(value, 1),
        "unit": "voxel-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": ("synthetic (seeded blob phantom with 8 Voronoi labels; source warped by an RBF velocity exponential)"
                 if c1 else "synthetic (seeded ellipsoid-shell phantom with 32 Voronoi labels; source warped by an "
                            "RBF velocity exponential)"),
        "config": {"workload": f"plain solver, {name} fp32, {iters} iterations, params.ini weights; a step is one "
                               f"registration (fresh state, loop, warps, Dice, min det J)",
                   "dims": list(dims), "iterations": iters, "voxels": n, "dice_all": round(dice, 6),
                   "min_det_true": min_det, "loop_ms_median": round(loop_med, 3),
                   "loop_us_per_iteration": round(1e3 * loop_med / iters, 2),
                   "loop_voxel_iterations_per_s": round(n * iters / (loop_med / 1e3), 1),
                   "z_chunk": {"pass_a": chunking[0], "pass_b": chunking[1]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if vs_ref is not None:
        out["config"]["arith"] = "fma (multiply-adds contracted, reciprocal in the density division)"
        out["arith_vs_reference"] = vs_ref
    print(json.dumps(out), flush=True)


def arith_vs_reference(s, registration, outs, st, n):
    """The last (contracted-arithmetic) registration of `s` against one in the reference's
    operation order from the same inputs (the order that equals the oracle bit for bit): Dice, min
    det J, warped labels that differ, and the relative L2 deviation of the final fields (psi and
    psi_inv as displacements from the identity)."""
    import numpy as np
    import fastreg_amd as fr

    def snapshot():
        state = s.get_state()
        labs = (outs[1].get().copy(), outs[3].get().copy())
        return state, labs, st.dice_all, st.min_det_true, st.iterations

    f_state, f_labs, f_dice, f_md, f_it = snapshot()
    s.set_arith(fr.ARITH_REFERENCE)
    registration()
    r_state, r_labs, r_dice, r_md, r_it = snapshot()
    X, Y, Z = s.dims
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    ident = np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float64)

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        den = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))

    fields = {"v": rel(f_state["v"][:, :3], r_state["v"][:, :3]), "rho": rel(f_state["rho"], r_state["rho"]),
              "psi_minus_id": rel(f_state["psi"][:, :3] - ident, r_state["psi"][:, :3] - ident),
              "psi_inv_minus_id": rel(f_state["psi_inv"][:, :3] - ident, r_state["psi_inv"][:, :3] - ident),
              "psi": rel(f_state["psi"][:, :3], r_state["psi"][:, :3])}
    return {"reference": {"dice_all": r_dice, "min_det_true": r_md, "iterations": r_it},
            "fma": {"dice_all": f_dice, "min_det_true": f_md, "iterations": f_it},
            "d_dice": f_dice - r_dice, "d_min_det": f_md - r_md,
            "labels_differing": {"target_o_psi_inv": int((f_labs[0] != r_labs[0]).sum()),
                                 "source_o_psi": int((f_labs[1] != r_labs[1]).sum()), "of": n},
            "final_rel_l2": {k: round(v, 9) for k, v in fields.items()}}


def precision_workload(args):
    """The 64^3 blob phantom (seed 1, 9 labels), 300 iterations with the params.ini weights and the
    energy logged every iteration, run on the GPU with v and rho stored in FP32, BF16 and FP16, and
    once by the FP64 oracle (the reference's algorithm in double).  Reported per storage: relative
    L2 error of v, rho and the displacements psi - id, psi_inv - id against FP64; the largest
    relative deviation of the energy curve; Dice (labels 1..8) and min det J next to FP64's; for the
    16-bit forms also the same errors against the FP32 form (what the storage alone changes).  Then
    the throughput of each storage at 182x218x182 (300 iterations, HIP events) and its kernel times.
    (With this phantom and step the displacements stay below 1e-3 voxel, a few FP32 ulps of the
    absolute coordinates psi holds, so FP32 itself is far from FP64 in psi and, through the floor
    of the label warp, in Dice; `label_flips_vs_fp64` counts the labels that differ from FP64's and
    how many of them that floor explains.)"""
    import numpy as np
    import fastreg_amd as fr
    from fastreg_amd import phantoms
    from oracle import oracle as O

    O.build()
    iters = 300
    pair = phantoms.blobs()
    dims = pair.dims
    n = dims[0] * dims[1] * dims[2]
    t0 = time.time()
    o = O.Oracle(dims, "f64")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=iters, verbosity=2, no_labels=pair.no_labels))
    ref = o.get_state()
    _, ref_lab1, _, ref_lab2 = o.outputs()
    ov, vx, _ = o.dice_counts(pair.no_labels)
    ref_dice = 2.0 * ov[1:].sum() / vx[1:].sum()
    ref_det = o.min_det()[0]
    ref_e = np.asarray(r["energy_hist"][:iters], np.float64)
    o.close()
    print(f"[bench] FP64 oracle, {iters} iterations at {dims}: {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    zz, yy, xx = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    ident = np.stack([xx.ravel(), yy.ravel(), zz.ravel(), np.zeros(n)], 1).astype(np.float64)

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def label_flips(p32, lab32, p64, lab64):
        """Voxels whose warped label differs from the FP64 run's, and how many of them the floor of
        the label warp explains: a component where the FP64 coordinate lies below the grid point
        and this run's does not (the FP64 lookup takes the neighbour below), and among those, how
        many hold no displacement at all in this run (psi == id exactly: every increment was below
        half an ulp of the coordinate and was rounded away)."""
        p32 = np.asarray(p32, np.float64)[:, :3]
        p64 = np.asarray(p64, np.float64)[:, :3]
        id3 = ident[:, :3]
        diff = np.asarray(lab32).reshape(-1) != np.asarray(lab64).reshape(-1)
        below = ((p64 < id3) & (p32 >= id3)).any(1)
        exact = (p32 == id3).all(1)
        return {"labels_differ": int(diff.sum()), "explained_by_floor_below_grid": int((diff & below).sum()),
                "of_which_no_displacement_held": int((diff & below & exact).sum()),
                "median_fp64_displacement_there": float(np.median(np.abs(p64 - id3).max(1)[diff])) if diff.any()
                else 0.0}

    names = {fr.STORAGE_F32: "fp32", fr.STORAGE_BF16: "bf16", fr.STORAGE_F16: "fp16"}
    errors, states = {}, {}
    for stg, nm in names.items():
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=iters, verbosity=2)
        s = fr.Solver(p)
        s.set_storage(stg)
        s.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        s.run()
        st = s.stats()
        _, eh = s.history()
        g = s.get_state()
        _, lab1, _, lab2 = s.finalize()
        st = s.stats()
        eh = np.asarray(eh[:iters], np.float64)
        errors[nm] = {
            "rel_l2_v": rel(g["v"], ref["v"]),
            "rel_l2_rho": rel(g["rho"], ref["rho"]),
            "rel_l2_psi_disp": rel(np.asarray(g["psi"], np.float64) - ident, ref["psi"] - ident),
            "rel_l2_psi_inv_disp": rel(np.asarray(g["psi_inv"], np.float64) - ident, ref["psi_inv"] - ident),
            "energy_max_rel_dev": float(np.max(np.abs(eh - ref_e) / np.abs(ref_e))),
            "dice_all": round(st.dice_all, 6),
            "min_det_true": round(st.min_det_true, 6),
            # the Dice gap to FP64, voxel by voxel: source labels o psi and target labels o psi_inv
            "label_flips_vs_fp64": {"source_o_psi": label_flips(g["psi"], lab2, ref["psi"], ref_lab2),
                                    "target_o_psi_inv": label_flips(g["psi_inv"], lab1, ref["psi_inv"], ref_lab1)},
        }
        states[nm] = g
        if nm != "fp32":  # what the storage alone changes: against the FP32 form
            f32 = states["fp32"]
            errors[nm]["vs_fp32"] = {
                "rel_l2_v": rel(g["v"], f32["v"]),
                "rel_l2_rho": rel(g["rho"], f32["rho"]),
                "rel_l2_psi_disp": rel(np.asarray(g["psi"], np.float64) - ident,
                                       np.asarray(f32["psi"], np.float64) - ident),
                "dice_all_diff": round(errors[nm]["dice_all"] - errors["fp32"]["dice_all"], 6),
            }
        s.close()
        print(f"[bench] {nm}: {errors[nm]}", file=sys.stderr, flush=True)

    # throughput at the full size
    big = phantoms.brain()
    bd = big.dims
    nb = bd[0] * bd[1] * bd[2]
    speed = {}
    for stg, nm in names.items():
        p = fr.default_params(vol_dims=bd, no_labels=big.no_labels, max_iter=300)
        s = fr.Solver(p)
        s.set_storage(stg)
        for rep in range(2):  # warm-up run, then the timed run
            s.init(big.target, big.target_labels, big.source, big.source_labels)
            s.run()
            st = s.stats()
        ka, kb = s.time_kernels(20)
        speed[nm] = {"voxel_iterations_per_s": round(nb * 300 / (st.solver_ms / 1e3), 1),
                     "ms_per_iteration": round(st.solver_ms / 300, 4),
                     "kernel_ms": {"pass_a": round(ka, 4), "pass_b": round(kb, 4)}}
        s.close()
        print(f"[bench] {nm} at {bd}: {speed[nm]}", file=sys.stderr, flush=True)

    out = {
        "metric": "storage precision of v and rho vs FP64 oracle (64^3, 300 iterations) and voxel-iterations/s "
                  "at 182x218x182",
        "value": speed["bf16"]["voxel_iterations_per_s"],
        "unit": "voxel-iterations/s (bf16 storage)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32 arithmetic; v, rho stored in fp32 / bf16 / fp16",
        "data": "synthetic (blob phantom seed 1 for the errors, brain phantom seed 2 for the speed)",
        "config": {"workload": "precision forms", "error_dims": list(dims), "iterations": iters,
                   "speed_dims": list(bd)},
        "fp64_reference": {"dice_all": round(float(ref_dice), 6), "min_det_true": round(float(ref_det), 6)},
        "errors_vs_fp64": errors,
        "speed": speed,
        "speedup_vs_fp32": {nm: round(speed[nm]["voxel_iterations_per_s"] / speed["fp32"]["voxel_iterations_per_s"], 3)
                            for nm in speed},
    }
    print(json.dumps(out), flush=True)


def adjoint_workload(args):
    """The differentiated solver on the 128^3 blob pair (the 64^3 generator with every length
    doubled, seed 5, 9 labels), K unrolled FP64 iterations (default 64) with checkpoints every
    `--adj-ckpt` (default 8), objective = SSD of the final state + 1.0 x the label-overlap
    surrogate, evaluated at a smooth random v0 (amplitude 0.5, offset 0.05: v0 = 0 would put every
    upwind branch on its kink).  Reported: device time of the forward pass (K iterations + the
    objective) and of the backward pass (segment recompute + adjoint kernels) and their ratio;
    device memory of the context at checkpoint intervals 1, ckpt and K; central finite differences
    of the forward pass along `--adj-dirs` random directions at eps 1e-4 and at eps 1e-7 against the
    adjoint's directional derivative, and for W_REG and TAU; the same at eps 1e-4 on the ramp pair
    from a small positive v0, where no branch switches (`finite_differences_without_kinks`); and,
    in the CPU leg, the reference-path
    gradient (torch autograd of the unrolled restatement, FP64, host cores) at 128^3 with
    `--cpu-adj-k` iterations against the adjoint at the same K."""
    import numpy as np
    import fastreg_amd as fr
    from fastreg_amd import phantoms

    dims = (128, 128, 128)
    X, Y, Z = dims
    n = X * Y * Z
    K, ckpt, lw = args.adj_k, args.adj_ckpt, 1.0
    alpha, w_reg, tau = (float(np.float32(x)) for x in (0.004, 0.2, 0.2))
    t0 = time.time()
    pair = phantoms.blobs(shape=(Z, Y, X), seed=5, scale=2.0)
    v0 = phantoms.smooth_velocity(dims, 5, amp=0.5, offset=0.05, sigma=4.0)
    print(f"[bench] adjoint pair {dims} ready ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    def make(k, c):
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, alpha=alpha)
        a = fr.Adjoint(p, k, c)
        a.set_images(pair.target, pair.target_labels, pair.source, pair.source_labels)
        return a

    runs = {}
    for c in sorted({1, ckpt, K}):
        a = make(K, c)
        a.forward(v0, w_reg, tau, lw)  # warm-up
        a.backward()
        fwd, bwd = [], []
        for _ in range(3):
            e = a.forward(v0, w_reg, tau, lw)
            g, gw, gt = a.backward()
            by, fm, bm = a.info()
            fwd.append(fm)
            bwd.append(bm)
        runs[c] = {"device_bytes": by, "forward_ms": round(float(np.median(fwd)), 3),
                   "backward_ms": round(float(np.median(bwd)), 3),
                   "backward_over_forward": round(float(np.median(bwd) / np.median(fwd)), 3)}
        print(f"[bench] ckpt {c}: {runs[c]}", file=sys.stderr, flush=True)
        if c == ckpt:
            main_ctx, E, G, GW, GT = a, e, g, gw, gt
        else:
            a.close()
    a = main_ctx

    def fd_check(h, v, grad, eps, ndir, seed):
        r = np.random.default_rng(seed)
        rels = []
        for _ in range(ndir):
            d = r.normal(size=(n, 3))
            fd = (h.forward(v + eps * d, w_reg, tau, lw) - h.forward(v - eps * d, w_reg, tau, lw)) / (2 * eps)
            an = float((grad * d).sum())
            rels.append(abs(fd - an) / abs(an))
        rels = np.asarray(rels)
        return {"eps": eps, "directions": ndir, "median_rel_err": float(np.median(rels)),
                "max_rel_err": float(rels.max()), "frac_below_1e-3": float(np.mean(rels < 1e-3))}

    def fd_scalars(h, v, eps, gw, gt):
        fw = (h.forward(v, w_reg + eps, tau, lw) - h.forward(v, w_reg - eps, tau, lw)) / (2 * eps)
        ft = (h.forward(v, w_reg, tau + eps, lw) - h.forward(v, w_reg, tau - eps, lw)) / (2 * eps)
        return {"w_reg": {"fd": fw, "adjoint": gw, "rel_err": abs(fw - gw) / abs(gw)},
                "tau": {"fd": ft, "adjoint": gt, "rel_err": abs(ft - gt) / abs(gt)}}

    fd = {}
    for eps in (1e-4, 1e-7):
        fd[f"v0_eps_{eps:g}"] = fd_check(a, v0, G, eps, args.adj_dirs, 55)
        print(f"[bench] FD {fd[f'v0_eps_{eps:g}']}", file=sys.stderr, flush=True)
    for eps in (1e-4, 1e-6):
        fd[f"scalars_eps_{eps:g}"] = fd_scalars(a, v0, eps, GW, GT)
    a.close()

    # the same checks where no branch switches near the evaluation point: the ramp pair from a
    # small positive v0 (every velocity component stays positive, psi ahead of and psi_inv behind
    # the grid nodes), at the full step of 1e-4
    ramp = phantoms.ramp((Z, Y, X))
    rv0 = phantoms.ramp_velocity(dims, 5)
    p = fr.default_params(vol_dims=dims, no_labels=ramp.no_labels, alpha=alpha)
    r = fr.Adjoint(p, K, ckpt)
    r.set_images(ramp.target, ramp.target_labels, ramp.source, ramp.source_labels)
    RE = r.forward(rv0, w_reg, tau, lw)
    RG, RGW, RGT = r.backward()
    fd_ramp = {"objective": RE, "v0_eps_1e-4": fd_check(r, rv0, RG, 1e-4, args.adj_dirs, 56),
               "scalars_eps_1e-4": fd_scalars(r, rv0, 1e-4, RGW, RGT)}
    print(f"[bench] FD ramp {fd_ramp['v0_eps_1e-4']}", file=sys.stderr, flush=True)
    r.close()

    cpu = None
    if not args.no_cpu:
        cpu = adjoint_cpu_baseline(dims, pair, v0, args.cpu_adj_k, alpha, w_reg, tau, lw, make)
    tot_ms = runs[ckpt]["forward_ms"] + runs[ckpt]["backward_ms"]
    no_ckpt_bytes = runs[K]["device_bytes"]
    out = {
        "metric": "differentiated solver at 128^3: voxel-iterations/s of forward + backward, adjoint/forward time, "
                  "device memory, gradient error against finite differences and the reference-path gradient",
        "value": round(n * K / (tot_ms / 1e3), 1),
        "unit": "voxel-iterations/s (forward + backward, FP64)",
        "n_gpus": 1,
        "higher_is_better": True,
        "ms_per_step": round(tot_ms, 3),
        "dtype": "f64",
        "data": "synthetic (blob phantom at 128^3, seed 5; smooth random v0)",
        "config": {"workload": f"differentiated solver, 128x128x128, K={K} unrolled iterations, checkpoint every "
                               f"{ckpt}", "dims": list(dims), "K": K, "ckpt": ckpt, "label_weight": lw},
        "objective": E,
        "grad_w_reg": GW,
        "grad_tau": GT,
        "backward_over_forward": runs[ckpt]["backward_over_forward"],
        "device_bytes": runs[ckpt]["device_bytes"],
        "device_bytes_by_ckpt": {str(c): r["device_bytes"] for c, r in runs.items()},
        "times_by_ckpt": runs,
        "memory_saving_vs_keeping_every_iteration": round(no_ckpt_bytes / runs[ckpt]["device_bytes"], 2),
        "finite_differences": fd,
        "finite_differences_without_kinks": fd_ramp,
        "vs_baseline": None,
    }
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def adjoint_cpu_baseline(dims, pair, v0, k, alpha, w_reg, tau, lw, make):
    """CPU leg of the adjoint workload: the reference-path gradient (oracle/unrolled.py, torch FP64
    autograd on the host cores) of `k` unrolled iterations on the same pair, timed, and the
    adjoint's gradient at the same k compared with it."""
    import numpy as np
    import torch
    from oracle import unrolled as U

    n = dims[0] * dims[1] * dims[2]
    t0 = time.time()
    vt = torch.tensor(v0, requires_grad=True)
    wt = torch.tensor(w_reg, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(tau, dtype=torch.float64, requires_grad=True)
    E = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels, vt, wt, tt, k,
                   alpha=alpha, no_labels=pair.no_labels, label_weight=lw)
    E.backward()
    cs = time.time() - t0
    a = make(k, 1)
    e = a.forward(v0, w_reg, tau, lw)
    g, gw, gt = a.backward()
    a.close()
    gr = vt.grad.numpy()
    res = {"value": round(n * k / cs, 1), "unit": "voxel-iterations/s (forward + backward, FP64)",
           "cores": torch.get_num_threads(), "kind": "port",
           "sample": f"torch FP64 autograd of the unrolled restatement, {dims[0]}x{dims[1]}x{dims[2]}, K={k}, "
                     f"host cores ({cs:.1f} s)",
           "adjoint_vs_reference_path_gradient": {
               "K": k, "objective_rel_diff": abs(e - E.item()) / abs(E.item()),
               "rel_l2_grad_v0": float(np.linalg.norm(g - gr) / np.linalg.norm(gr)),
               "rel_diff_grad_w_reg": abs(gw - wt.grad.item()) / abs(wt.grad.item()),
               "rel_diff_grad_tau": abs(gt - tt.grad.item()) / abs(tt.grad.item())}}
    print(f"[bench] cpu reference-path gradient: {res}", file=sys.stderr, flush=True)
    return res


BATCH_SIZES = [(48, 48, 48), (96, 96, 96), (128, 160, 128), (160, 192, 160), (182, 218, 182)]  # (X, Y, Z)


def _make_pair(job):
    i, dims = job
    from fastreg_amd import phantoms
    X, Y, Z = dims
    return i, phantoms.blobs(shape=(Z, Y, X), seed=1000 + i)


def batch_workload(args):
    """64 independent pairs of sizes drawn uniformly from BATCH_SIZES (seed 4), each the blob
    phantom with its own seed (1000 + i), 150 iterations each with the params.ini weights.  Timed
    twice on the same contexts: one pair at a time (fr_solver_run per pair, HIP events per pair)
    and as one batch (fr_batch_run: one launch per pass per iteration for all pairs, replayed
    from a hipGraph, z-chunks from the autotuner).  Both must give the same max-norm history for
    every pair.  With --batch-storage bf16 / fp16 the one-at-a-time leg stays the plain FP32 form
    and the batch stores v and rho in 16 bits; then the history check covers the first four pairs
    (run one at a time in the same storage), and the batched Dice of every pair is compared with
    the FP32 form's and, for the pairs of up to 96^3, with an FP64 oracle run's."""
    import multiprocessing as mp
    import numpy as np

    rng = np.random.default_rng(4)
    sizes = BATCH_SIZES if args.batch_sizes == "mixed" else BATCH_SIZES[:2]
    jobs = [(i, sizes[int(k)]) for i, k in enumerate(rng.integers(0, len(sizes), args.batch_pairs))]
    t0 = time.time()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        pairs = dict(pool.map(_make_pair, jobs))
    print(f"[bench] {len(jobs)} pairs generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    import fastreg_amd as fr
    from fastreg_amd import autotune
    solvers = []
    for i, dims in jobs:
        pair = pairs[i]
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=args.batch_iters)
        s = fr.Solver(p)
        s.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        solvers.append(s)
    voxels = sum(d[0] * d[1] * d[2] for _, d in jobs)

    # one pair at a time; a warm-up run of every pair first (each context captures its iteration
    # loop into a hipGraph on its first run, as the batch does on its first run)
    for s in solvers:
        s.run()
        s.stats()
        s.reinit()
    per_pair, hist = [], []
    fr.check(fr.lib().fr_dev_sync())
    tw = time.perf_counter()
    for s in solvers:
        s.run()
        st = s.stats()
        per_pair.append(st.solver_ms)
        hist.append(s.history()[0][: st.iterations].copy())
    plain_wall = (time.perf_counter() - tw) * 1e3
    plain_ms = float(sum(per_pair))
    storage = {"fp32": fr.STORAGE_F32, "bf16": fr.STORAGE_BF16, "fp16": fr.STORAGE_F16}[args.batch_storage]
    dice32 = []
    if storage != fr.STORAGE_F32:
        # Dice of the plain FP32 form, then the batched leg's storage; the first pairs are run one at a
        # time in that storage too, for the history check below
        for s in solvers:
            s.finalize(want_outputs=False)
            dice32.append(s.stats().dice_all)
            s.set_storage(storage)
        hist = hist[: min(4, len(solvers))]
        for k in range(len(hist)):
            s = solvers[k]
            s.reinit()
            s.run()
            st = s.stats()
            hist[k] = s.history()[0][: st.iterations].copy()
    for s in solvers:
        s.reinit()

    b = fr.Batch(solvers)
    tuned, _ = autotune.tune_batch(b, log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
    b.run()  # capture + first replay (warm-up)
    b.time_ms()
    for s in solvers:
        s.stats()
        s.reinit()
    tw = time.perf_counter()
    b.run()
    batch_ms = b.time_ms()
    batch_wall = (time.perf_counter() - tw) * 1e3
    for k in range(len(hist)):
        s = solvers[k]
        st = s.stats()
        assert np.array_equal(s.history()[0][: st.iterations], hist[k]), f"pair {k}: batch and single runs differ"
    b.close()

    n = len(solvers)
    dice = None
    if dice32:
        dice16 = []
        for s in solvers:
            s.finalize(want_outputs=False)
            dice16.append(s.stats().dice_all)
        dev = np.abs(np.asarray(dice16) - np.asarray(dice32))
        dice = {"batched_storage_vs_plain_fp32": {"max_abs_diff": round(float(dev.max()), 6),
                                                  "mean_abs_diff": round(float(dev.mean()), 6)}}
        dice["vs_fp64_oracle"] = _batch_dice_vs_fp64(jobs, pairs, args.batch_iters, dice16, dice32)
        print(f"[bench] Dice: {dice}", file=sys.stderr, flush=True)
    out = {
        "metric": f"pairs/s, {n} pairs of uneven sizes x {args.batch_iters} iterations",
        "value": round(n / (batch_ms / 1e3), 2),
        "unit": "pairs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32",
        "data": "synthetic (blob phantoms, one seed per pair)",
        "config": {"workload": f"batched registrations ({args.batch_sizes} sizes)", "pairs": n,
                   "iterations": args.batch_iters,
                   "sizes": {"x".join(map(str, d)): sum(1 for _, e in jobs if e == d) for d in sizes},
                   "voxels": voxels,
                   "z_chunk": {"x".join(map(str, d)): list(v) for d, v in tuned.items()}},
        "batch_ms": round(batch_ms, 2),
        "batch_wall_ms": round(batch_wall, 2),
        "batch_voxel_iterations_per_s": round(voxels * args.batch_iters / (batch_ms / 1e3), 1),
        "one_at_a_time": {"pairs_per_s": round(n / (plain_ms / 1e3), 2), "sum_ms": round(plain_ms, 2),
                          "wall_ms": round(plain_wall, 2),
                          "p50_pair_ms": round(float(np.percentile(per_pair, 50)), 3),
                          "p95_pair_ms": round(float(np.percentile(per_pair, 95)), 3)},
        "speedup": round(plain_ms / batch_ms, 3),
        "per_pair_latency_ms_batched": round(batch_ms, 2),
    }
    if dice is not None:
        out["dtype"] = f"{args.batch_storage} storage of v and rho, fp32 arithmetic (one-at-a-time leg: fp32)"
        out["dice"] = dice
    print(json.dumps(out), flush=True)


def _batch_dice_vs_fp64(jobs, pairs, iters, dice_batched, dice_fp32, max_voxels=96 ** 3):
    """Dice of the pairs of up to `max_voxels` voxels from the FP64 oracle (the reference's
    algorithm in double, same iterations and weights), against the batched leg's and the FP32
    form's Dice: per size, the largest absolute difference of each."""
    import numpy as np
    from oracle import oracle as O

    O.build()
    per = {}
    t0 = time.time()
    for k, (i, dims) in enumerate(jobs):
        if dims[0] * dims[1] * dims[2] > max_voxels:
            continue
        pair = pairs[i]
        o = O.Oracle(dims, "f64")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        o.run(O.SolverParams(max_iter=iters, no_labels=pair.no_labels))
        o.outputs()
        ov, vx, _ = o.dice_counts(pair.no_labels)
        o.close()
        d64 = 2.0 * ov[1:].sum() / vx[1:].sum()
        e = per.setdefault("x".join(map(str, dims)), {"pairs": 0, "batched_max_abs_diff": 0.0,
                                                      "fp32_max_abs_diff": 0.0})
        e["pairs"] += 1
        if sum(v["pairs"] for v in per.values()) % 8 == 0:
            print(f"[bench] FP64 oracle: {sum(v['pairs'] for v in per.values())} pairs done", file=sys.stderr,
                  flush=True)
        e["batched_max_abs_diff"] = round(float(max(e["batched_max_abs_diff"], abs(dice_batched[k] - d64))), 6)
        e["fp32_max_abs_diff"] = round(float(max(e["fp32_max_abs_diff"], abs(dice_fp32[k] - d64))), 6)
    print(f"[bench] FP64 oracle Dice of {sum(e['pairs'] for e in per.values())} pairs: {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    return per


if __name__ == "__main__":
    main()
