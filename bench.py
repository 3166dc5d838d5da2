#!/usr/bin/env python3
"""Benchmark: one per-Newton linear solve of libuipc on one MI355X.

Workload (default): the 64^3-vertex-grid FEM system (262,144 vertices,
1,250,235 stable-Neo-Hookean tets, PSD-projected, + M) plus 2,000,000 synthetic
vertex-triangle contact stencils (kappa g g^T, kappa log-uniform [1e4, 1e8]),
seeded.  One step = the whole per-Newton solve with the triplets and b already
resident in HBM:
    triplets -> upper BCOO/BSR (radix sort + segmented reduce)
    preconditioner assembly
    PCG from x0 = 0 to ||b - A x|| / ||b|| <= 1e-4 (recurrence residual)
Prints one JSON line (metric: ms per solve; PCG iterations and the SpMV's
achieved HBM bandwidth ride along).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 64] [--contacts 2000000]
                [--precond mas|diag] [--precision fp64|fp32] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--contacts", type=int, default=2_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precond", choices=["mas", "diag"], default="mas")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--mas-invert", choices=["sweep", "mfma"], default="mfma",
                    help="cluster inverse kernel: the FP64 MFMA panel sweep (default) or the column sweep")
    ap.add_argument("--mas-levels", type=lambda v: v if v == "auto" else int(v), default=0,
                    help="MAS levels applied (MASPreconditionerEngine::set_active_level_num): 0 = all, the "
                         "default; auto = the fastest level count by libuipc_amd.tune, picked before timing")
    ap.add_argument("--spmv-shape", type=lambda v: v if v == "auto" else int(v), default="auto",
                    help="SpMV launch shape: an index of mi_spmv_shapes, or auto = the fastest on this matrix by "
                         "libuipc_amd.tune.tune_spmv_shape, picked after the warm-up, before timing")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32-arm", action="store_true",
                    help="skip the FP32 arm that the default FP64 MAS run also times")
    ap.add_argument("--no-ref-arm", action="store_true",
                    help="skip the reference-precision solve (FP64 system, float MAS storage) it also times")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--producers", choices=["host", "device"], default="host",
                    help="device: write the triplets with the device Hessian producers (timed apart, outside "
                         "the solve step) instead of uploading the host generator's")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the PCG iterations directly instead of through captured HIP graphs")
    ap.add_argument("--workload", choices=["single", "batched", "learned", "coupled"], default="single",
                    help="batched: many independent FEM + contact systems of log-uniform size solved together "
                         "(mi_pcg_solve_batched), the triplets written on the device by the Hessian producers; "
                         "learned: train the learned initial guess on 16^3 systems and compare MAS-PCG from it "
                         "with MAS-PCG from x0 = 0 on held-out systems; coupled: 10,000 affine bodies + a 256^2 "
                         "cloth with bending + 500k contacts in one system (12x12 block Jacobi on the bodies, MAS "
                         "on the cloth)")
    ap.add_argument("--train-systems", type=int, default=5000)
    ap.add_argument("--test-systems", type=int, default=500)
    ap.add_argument("--ws-loss", choices=["energy", "residual"], default="energy",
                    help="training loss of the learned initial guess")
    ap.add_argument("--ws-lr", type=float, default=1e-3)
    ap.add_argument("--ws-contacts", type=int, default=20000,
                    help="contact stencils per system of the learned-guess workload (kappa_hi perturbed per system)")
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--dof-lo", type=float, default=3e3)
    ap.add_argument("--dof-hi", type=float, default=6e5)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="only run N device-to-device copies of this many MiB (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE) and print the copy bandwidth")
    return ap.parse_args()


def calibrate_copy(mb: int) -> None:
    from libuipc_amd.device import Context
    from libuipc_amd.hip_events import EventTimer

    ctx = Context(0)
    nbytes = mb << 20  # one hipMemcpyAsync + stream sync per repetition
    a = ctx.alloc(nbytes // 8, np.float64)
    b = ctx.alloc(nbytes // 8, np.float64)
    a.zero()
    ev = EventTimer(ctx.stream())
    ctx.lib.mi_memcpy_d2d(ctx.h, b.ptr, a.ptr, nbytes)
    reps = 5
    ev.start()
    for _ in range(reps):
        ctx.lib.mi_memcpy_d2d(ctx.h, b.ptr, a.ptr, nbytes)
    ms = ev.stop() / reps
    print(json.dumps({"copy_bytes": nbytes, "copy_ms": ms, "copy_GBps_read_plus_write": 2 * nbytes / ms / 1e6}))
    ctx.close()


def measured_traffic(kernel: str, args) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE / WRITE_SIZE runs scaled by a copy kernel of known size), when
    it was measured on this very workload; None otherwise."""
    if (args.grid, args.contacts, args.seed) != (64, 2_000_000, 1):
        return None
    path = os.path.join(ROOT, "profiles", f"c2_{args.precision}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == f"`{kernel}`":
            return float(cells[3]) * 1e6
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.calibrate_copy_mb > 0:
        calibrate_copy(args.calibrate_copy_mb)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the per-Newton solve of one system is not "
                                   "sharded across GPUs",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "learned":
        run_learned(args)
        return
    if args.workload == "coupled":
        run_coupled(args)
        return

    from libuipc_amd import abi, synth
    from libuipc_amd.device import Context, LinearSystem
    from libuipc_amd.hip_events import EventTimer

    t0 = time.perf_counter()
    s = synth.fem_system(args.grid, args.contacts, seed=args.seed)
    log(f"generated grid={args.grid} contacts={args.contacts}: B={s.block_rows} T={s.n_triplets} "
        f"in {time.perf_counter() - t0:.1f}s")

    ctx = Context(0)
    stream = ctx.stream()
    prec = abi.MI_FP32 if args.precision == "fp32" else abi.MI_FP64
    ls = LinearSystem(ctx, s.block_rows, prec)
    # an FP32 system takes FP32 triplet values (mi_sys_assemble_f32), except
    # from the device producers, which write doubles
    fp32_trip = prec == abi.MI_FP32 and args.producers != "device"
    rows, cols = ctx.to_device(s.rows), ctx.to_device(s.cols)
    vals = ctx.to_device(s.vals.astype(np.float32) if fp32_trip else s.vals)
    b = ctx.to_device(s.rhs)
    producers_ms = None
    if args.producers == "device":
        producers_ms = run_producers(ctx, s, args, rows, cols, vals)
    x = ctx.alloc(s.n, np.float64)
    precond = args.precond
    if precond == "mas":
        ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets, mas_precision=prec,
                              active_levels=0 if args.mas_levels == "auto" else args.mas_levels,
                              invert=abi.MI_MAS_INVERT_SWEEP if args.mas_invert == "sweep" else abi.MI_MAS_INVERT_MFMA)
    else:
        ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)

    solve_kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=args.tol, check_interval=5,
                    use_graph=0 if args.no_graph else 1)

    def step():
        ls.assemble(rows, cols, vals, stream)
        ls.assemble_preconditioners()
        return ls.solve(b, x, **solve_kw)

    if args.spmv_shape != "auto":
        ls.set_spmv_shape(args.spmv_shape)
    for w in range(args.warmup):
        st = step()
        log(f"warmup {w}: iters={st['iters']} exact={st['iters_exact']} rel_res={st['rel_res']:.3e}")
    shape_tuned = None
    if args.spmv_shape == "auto":
        from libuipc_amd.tune import tune_spmv_shape

        step()
        shape_tuned = tune_spmv_shape(ls, x)
        log(f"SpMV shape tuned: {shape_tuned['best']} of {shape_tuned['table']}")
    tuned = None
    if precond == "mas" and args.mas_levels == "auto":
        # untimed, like the warm-up: one solve per level count on this system
        from libuipc_amd.tune import tune_mas_levels

        tuned = tune_mas_levels(ls, 0, b, x, **solve_kw)
        log(f"MAS levels tuned: {tuned['best']} of {tuned['table']}")

    ev = EventTimer(stream)
    ctx.synchronize()
    times, stats = [], []
    t_wall0 = time.perf_counter()
    for k in range(args.steps):
        ev.start()
        st = step()
        ms = ev.stop()
        times.append(ms)
        stats.append(st)
        log(f"step {k}: {ms:.2f} ms iters={st['iters']} exact={st['iters_exact']} rel_res={st['rel_res']:.3e}")
    ctx.synchronize()
    wall = time.perf_counter() - t_wall0
    ms_per_step = float(np.median(times))

    # phase split of one step
    ev.start()
    ls.assemble(rows, cols, vals, stream)
    ms_assemble = ev.stop()
    ev.start()
    ls.assemble_preconditioners()
    ms_precond = ev.stop()
    ev.start()
    ls.solve(b, x, **solve_kw)
    ms_solve = ev.stop()
    # the true residual of the same solve, outside the timing (asking for it
    # changes the solve's options, which re-captures its graph)
    st = ls.solve(b, x, **solve_kw, true_residual=1)
    U = ls.nnz()

    # dominant kernel: the SpMV (+ fused x^T A x), timed with HIP events on its
    # stream.  The PCG's product is the roofline's kernel, launched as the PCG
    # launches it (on the window layout: y = A x and the per-workgroup x^T A x
    # partials on vectors in the SpMV's row order, mi_spmv_ordered without the
    # partials' sum; on the full BSR: mi_spmv with a = 1, b = 0), timed as the
    # PCG runs it: spmv_reps launches captured in one HIP graph, replayed
    # between HIP events on the library's stream.  The ABI's accumulating
    # mi_spmv_dot on the caller's vertex order (+ its one-workgroup partials
    # sum, launched eagerly) rides along.
    from libuipc_amd.hip_events import graph_ms_per_call

    lay = ls.spmv_layout()
    y = ctx.alloc(s.n, np.float64)
    d = ctx.alloc(1, np.float64)
    ordered = lay["layout"] == "window"
    if ordered:
        def pcg_product():
            ctx.lib.mi_spmv_ordered(ls.h, x.ptr, y.ptr, None, stream)
    else:
        def pcg_product():
            ctx.lib.mi_spmv(ls.h, 1.0, x.ptr, 0.0, y.ptr, stream)
    for _ in range(3):
        pcg_product()
    ctx.synchronize()
    try:
        ms_spmv = graph_ms_per_call(stream, pcg_product, args.spmv_reps)
        spmv_timing = "captured graph"
    except AssertionError:  # no stream capture (e.g. under a counter-collecting profiler): eager launches
        ctx.synchronize()
        ev.start()
        for _ in range(args.spmv_reps):
            pcg_product()
        ms_spmv = ev.stop() / args.spmv_reps
        spmv_timing = "eager launches"
    ctx.lib.mi_spmv_dot(ls.h, x.ptr, y.ptr, d.ptr, stream)
    ev.start()
    for _ in range(args.spmv_reps):
        ctx.lib.mi_spmv_dot(ls.h, x.ptr, y.ptr, d.ptr, stream)
    ms_spmv_dot = ev.stop() / args.spmv_reps
    sval = 4 if prec == abi.MI_FP32 else 8
    spmv_bytes = U * (9 * sval + 4) + (s.block_rows + 1) * 4 + 2 * s.n * 8
    gbs = spmv_bytes / (ms_spmv * 1e-3) / 1e9
    vt = "double" if prec == abi.MI_FP64 else "float"
    if ordered:
        kname = (f"spmv_win_kernel<{vt}, false, true, {lay['threads']}, {lay['window_rows']}, {lay['per']}, "
                 f"{lay['vec']}>")
    else:
        kname = f"spmv_chunk_kernel<{vt}, false>"
    # bytes the layout streams per product: words + values of its blocks, the
    # row pointers, x and y
    streamed = lay["blocks"] * (9 * sval + 4) + (s.block_rows + 1) * 4 + 2 * s.n * 8

    result = {
        "metric": "ms per Newton solve to 1e-4 rel. residual; PCG iters; SpMV HBM GB/s vs peak",
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.MI_FP64 else "f32-matrix/f64-vectors",
        "data": "synthetic (seeded generator, random jitter / kappa / contacts)",
        "config": {
            "workload": f"{args.grid}^3 grid FEM (SNH, PSD) + {args.contacts} vertex-triangle contacts; "
                        f"assemble + {precond} precond + PCG to rel. residual {args.tol:g}",
            "block_rows": s.block_rows, "dofs": s.n, "triplets": s.n_triplets, "unique_upper_blocks": U,
            "preconditioner": precond, "precision": args.precision, "criterion": "rel_residual",
            "mas_levels": (tuned["best"] if tuned else args.mas_levels) if precond == "mas" else None,
            "mas_levels_tuned": {str(k): v for k, v in tuned["table"].items()} if tuned else None,
            "mas_invert": args.mas_invert if precond == "mas" else None,
            "spmv_shape": {**lay, "tuned": shape_tuned} if shape_tuned else lay,
        },
        "pcg_iters": int(st["iters"]),
        "pcg_iters_exact": int(st["iters_exact"]),
        "rel_res": st["rel_res"],
        "rel_res_true": st["rel_res_true"],
        "phase_ms": {"assemble": round(ms_assemble, 3), "precond_assemble": round(ms_precond, 3),
                     "pcg": round(ms_solve, 3),
                     "per_iter_us": round(1e3 * ms_solve / max(1, st["iters"]), 2)},
        "wall_s_timed": round(wall, 3),
        "producers_ms": producers_ms,
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(kname, args), "algorithmic_bytes": int(spmv_bytes),
                     "avg_us": round(ms_spmv * 1e3, 2), "layout": lay["layout"], "window_rows": lay["window_rows"],
                     "streamed_bytes": int(streamed), "streamed_GBps": round(streamed / (ms_spmv * 1e-3) / 1e9, 1),
                     "spmv_dot_avg_us": round(ms_spmv_dot * 1e3, 2), "timing": spmv_timing},
    }

    if prec == abi.MI_FP64 and precond == "mas" and not args.no_fp32_arm and args.mas_levels != "auto":
        result["fp32_arm"] = fp32_arm(ctx, s, args, rows, cols, vals, b, x.numpy(), ms_per_step, solve_kw)
    if prec == abi.MI_FP64 and precond == "mas" and not args.no_ref_arm and args.mas_levels != "auto":
        result["ref_precision_arm"] = ref_precision_arm(ctx, s, args, rows, cols, vals, b, x.numpy(), ms_per_step,
                                                        solve_kw)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(s, precond, args, int(st["iters"]))
    print(json.dumps(result), flush=True)
    ls.close()
    ctx.close()


def ref_precision_arm(ctx, s, args, rows, cols, vals, b, x64, ms64, solve_kw) -> dict:
    """The same solve at the reference's own precision: FP64 triplets, matrix
    stream, vectors and dots, and the MAS preconditioner's inverses and
    multilevel R / Z stored in float with the factorisation in FP64
    (mas_preconditioner_engine.h:17-22, 241, 254: "matches GIPC").  `value`
    keeps the MAS in FP64 throughout (the plain method); this line says what
    the reference's storage types cost or save.  Same MAS level count and
    SpMV layout choice as `value`, median of the same number of steps."""
    from libuipc_amd import abi
    from libuipc_amd.device import LinearSystem
    from libuipc_amd.hip_events import EventTimer

    stream = ctx.stream()
    ls = LinearSystem(ctx, s.block_rows, abi.MI_FP64)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets, mas_precision=abi.MI_FP32,
                          active_levels=args.mas_levels)
    x = ctx.alloc(s.n, np.float64)

    def step():
        ls.assemble(rows, cols, vals, stream)
        ls.assemble_preconditioners()
        return ls.solve(b, x, **solve_kw)

    for _ in range(args.warmup):
        step()
    if args.spmv_shape == "auto":
        from libuipc_amd.tune import tune_spmv_shape

        tune_spmv_shape(ls, x)
    else:
        ls.set_spmv_shape(args.spmv_shape)
    ev = EventTimer(stream)
    ctx.synchronize()
    times = []
    for _ in range(args.steps):
        ev.start()
        step()
        times.append(ev.stop())
    st = ls.solve(b, x, **solve_kw, true_residual=1)
    xr = x.numpy()
    ms = float(np.median(times))
    ls.close()
    out = {"ms": round(ms, 3), "ratio_to_fp64": round(ms / ms64, 4), "pcg_iters": int(st["iters"]),
           "rel_res_true": st["rel_res_true"],
           "x_rel_diff_to_fp64": float(np.linalg.norm(xr - x64) / max(np.linalg.norm(x64), 1e-300)),
           "mas_levels": args.mas_levels,
           "note": "FP64 system, vectors and dots; MAS inverses and R / Z stored in float (the reference's MAS "
                   "storage types), factorisation in FP64"}
    log(f"reference-precision arm: {out}")
    return out


def fp32_arm(ctx, s, args, rows, cols, vals, b, x64, ms64, solve_kw) -> dict:
    """The precision arm on the same system: FP32 triplet values (the
    producers' blocks rounded to float, in HBM before the timing like the
    FP64 ones), FP32 matrix stream (window layout values) and FP32 MAS
    inverses / R / Z, FP64 sums of duplicates, vectors, dots and scalars;
    timed like `value` (median of the same number of steps) at the same MAS
    level count, then at the level count the tuner picks.  Reports ms, the
    assembly's share, iterations, the true residual (against the FP32-rounded
    triplets' matrix) and ||x32 - x64|| / ||x64||."""
    from libuipc_amd import abi
    from libuipc_amd.device import LinearSystem
    from libuipc_amd.hip_events import EventTimer
    from libuipc_amd.tune import tune_mas_levels

    stream = ctx.stream()
    vals32 = ctx.to_device(s.vals.astype(np.float32))
    ls = LinearSystem(ctx, s.block_rows, abi.MI_FP32)
    h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets, mas_precision=abi.MI_FP32,
                              active_levels=args.mas_levels)
    x = ctx.alloc(s.n, np.float64)

    def step():
        ls.assemble(rows, cols, vals32, stream)
        ls.assemble_preconditioners()
        return ls.solve(b, x, **solve_kw)

    def timed():
        for _ in range(args.warmup):
            step()
        shapes = None
        if args.spmv_shape == "auto":
            from libuipc_amd.tune import tune_spmv_shape

            t = tune_spmv_shape(ls, x)
            shapes = {"best": t["best"], "us": {str(k): v["us"] for k, v in t["table"].items()}}
        else:
            ls.set_spmv_shape(args.spmv_shape)
        ev = EventTimer(stream)
        ctx.synchronize()
        times = []
        for _ in range(args.steps):
            ev.start()
            step()
            times.append(ev.stop())
        ev.start()
        ls.assemble(rows, cols, vals32, stream)
        ms_asm = ev.stop()
        ls.assemble_preconditioners()
        st = ls.solve(b, x, **solve_kw, true_residual=1)
        x32 = x.numpy()
        ms = float(np.median(times))
        return {"ms": round(ms, 3), "ratio_to_fp64": round(ms / ms64, 4), "assemble_ms": round(ms_asm, 3),
                "spmv_shape": shapes, "pcg_iters": int(st["iters"]), "rel_res_true": st["rel_res_true"],
                "x_rel_diff_to_fp64": float(np.linalg.norm(x32 - x64) / max(np.linalg.norm(x64), 1e-300))}

    same = timed()
    same["mas_levels"] = args.mas_levels
    log(f"fp32 arm, same levels: {same}")
    tuned = tune_mas_levels(ls, h, b, x, **solve_kw)
    best = timed()
    best["mas_levels"] = tuned["best"]
    best["tuner_table"] = {str(k): v for k, v in tuned["table"].items()}
    log(f"fp32 arm, tuned levels: {best['ms']} ms at {tuned['best']} levels")
    ls.close()
    return {"same_levels": same, "tuned_levels": best,
            "note": "FP32 triplet values, matrix stream and MAS storage; FP64 sums of duplicates, vectors, dots "
                    "and scalars; median of the same steps"}


def run_producers(ctx, s, args, rows, cols, vals) -> dict:
    """Overwrites the triplet buffers with the device producers' output for the
    same system, times each producer with HIP events (median of 5) and checks
    the result against the host generator's triplets."""
    from libuipc_amd import synth
    from libuipc_amd.abi import check
    from libuipc_amd.hip_events import EventTimer

    p = synth.producer_inputs(args.grid, args.contacts, seed=args.seed)
    B, nt, nc = s.block_rows, len(p.tets), len(p.contact_kappa)
    dm, dtets, dX, dx = (ctx.to_device(a) for a in (p.mass, p.tets, p.rest_positions, p.positions))
    ci, cg, ck = (ctx.to_device(a) for a in (p.contact_idx, p.contact_g, p.contact_kappa)) if nc else (None,) * 3
    lib, st = ctx.lib, ctx.stream()
    o_fem, o_con = B, B + 10 * nt
    r0, c0, v0 = rows.ptr.value, cols.ptr.value, vals.ptr.value
    calls = {
        "kinetic": lambda: lib.mi_produce_kinetic(ctx.h, B, dm.ptr, 0, r0, c0, v0, st),
        "snh": lambda: lib.mi_produce_snh(ctx.h, nt, dtets.ptr, dX.ptr, dx.ptr, p.mu, p.lam, p.dt, 0,
                                          r0 + 4 * o_fem, c0 + 4 * o_fem, v0 + 72 * o_fem, st),
    }
    if nc:
        calls["contact"] = lambda: lib.mi_produce_contact(ctx.h, nc, ci.ptr, cg.ptr, ck.ptr, p.dt, 0, r0 + 4 * o_con,
                                                          c0 + 4 * o_con, v0 + 72 * o_con, st)
    ev = EventTimer(st)
    out = {}
    for name, fn in calls.items():
        times = []
        for _ in range(5):
            ev.start()
            check(fn())
            times.append(ev.stop())
        out[name] = round(float(np.median(times)), 3)
    ctx.synchronize()
    same = np.array_equal(rows.numpy(), s.rows) and np.array_equal(cols.numpy(), s.cols)
    dv = float(np.abs(vals.numpy() - s.vals).max() / np.abs(s.vals).max())
    log(f"device producers: {out} ms; structure equal to the host generator's: {same}; max rel value diff {dv:.1e}")
    if not same or dv > 1e-9:
        raise RuntimeError("device producers disagree with the host generator")
    out["triplets_per_s"] = round((B + 10 * nt + 10 * nc) / (sum(out.values()) * 1e-3), 1)
    return out


def run_batched(args) -> None:
    """Independent FEM + contact systems with DoF counts log-uniform in
    [dof_lo, dof_hi] (grid N = round((DoF / 3)^(1/3)), contacts at the density
    of 2M per 64^3 grid), all in one block-diagonal system: the device
    producers write every system's triplets at its block-row offset, one
    assembly, one MAS over the union (disconnected systems never share a
    cluster), one batched PCG with per-system scalars.  A step is assembly +
    MAS build + the batched solve, timed with HIP events."""
    from libuipc_amd import abi, synth
    from libuipc_amd.abi import check
    from libuipc_amd.device import Context, LinearSystem
    from libuipc_amd.hip_events import EventTimer

    rng = np.random.default_rng(args.seed)
    dofs = np.exp(rng.uniform(np.log(args.dof_lo), np.log(args.dof_hi), args.systems))
    grids = np.maximum(4, np.round((dofs / 3.0) ** (1.0 / 3.0)).astype(int))
    density = 2_000_000 / 64 ** 3
    contacts = np.round(density * grids.astype(np.float64) ** 3).astype(np.int64)
    ctx = Context(0)
    st = ctx.stream()
    # sizes first (cheap), so the triplet buffer is allocated once
    t0 = time.perf_counter()
    inputs, seg, toff = [], [0], [0]
    for k, g in enumerate(grids):
        p = synth.producer_inputs(int(g), int(contacts[k]), seed=args.seed * 1000 + k)
        B = len(p.mass)
        inputs.append(p)
        seg.append(seg[-1] + B)
        toff.append(toff[-1] + B + 10 * len(p.tets) + 10 * len(p.contact_kappa))
        if (k + 1) % 32 == 0:
            log(f"generated {k + 1}/{args.systems} systems' producer inputs ({time.perf_counter() - t0:.1f}s)")
    seg, toff = np.array(seg, np.int64), np.array(toff, np.int64)
    Btot, T = int(seg[-1]), int(toff[-1])
    log(f"{args.systems} systems: {3 * Btot} DoFs, {Btot} block rows, {T} triplets "
        f"({T * 80 / 1e9:.1f} GB of triplets on the device)")
    rows, cols, vals = ctx.alloc(T, np.int32), ctx.alloc(T, np.int32), ctx.alloc((T, 9), np.float64)
    r0, c0, v0 = rows.ptr.value, cols.ptr.value, vals.ptr.value
    tets_all = np.empty((sum(len(p.tets) for p in inputs), 4), np.int32)
    parts_all = np.empty(Btot, np.int32)
    t_off, p_off = 0, 0
    for k, p in enumerate(inputs):
        vo, o = int(seg[k]), int(toff[k])
        B, nt, nc = len(p.mass), len(p.tets), len(p.contact_kappa)
        dm, dt_, dX, dx = (ctx.to_device(a) for a in (p.mass, p.tets, p.rest_positions, p.positions))
        check(ctx.lib.mi_produce_kinetic(ctx.h, B, dm.ptr, vo, r0 + 4 * o, c0 + 4 * o, v0 + 72 * o, st))
        o += B
        check(ctx.lib.mi_produce_snh(ctx.h, nt, dt_.ptr, dX.ptr, dx.ptr, p.mu, p.lam, p.dt, vo, r0 + 4 * o,
                                     c0 + 4 * o, v0 + 72 * o, st))
        o += 10 * nt
        if nc:
            ci, cg, ck = (ctx.to_device(a) for a in (p.contact_idx, p.contact_g, p.contact_kappa))
            check(ctx.lib.mi_produce_contact(ctx.h, nc, ci.ptr, cg.ptr, ck.ptr, p.dt, vo, r0 + 4 * o,
                                             c0 + 4 * o, v0 + 72 * o, st))
        ctx.synchronize()  # the per-system inputs are freed next
        tets_all[t_off:t_off + nt] = p.tets + vo
        t_off += nt
        parts_all[vo:vo + B] = np.where(p.part_ids >= 0, p.part_ids + p_off, -1)
        p_off += int(p.part_ids.max()) + 1
        inputs[k] = None
    log(f"device producers done ({time.perf_counter() - t0:.1f}s)")
    b = ctx.to_device(np.random.default_rng(args.seed + 7).uniform(-1.0, 1.0, 3 * Btot))
    x = ctx.alloc(3 * Btot, np.float64)
    ls = LinearSystem(ctx, Btot)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * Btot, part_ids=parts_all, tets=tets_all,
                          build=abi.MI_MAS_BUILD_DEVICE)
    log(f"MAS hierarchy over the union built ({time.perf_counter() - t0:.1f}s)")
    seg32 = seg.astype(np.int32)
    solve_kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=args.tol, check_interval=5)

    def step():
        ls.assemble(rows, cols, vals, st)
        ls.assemble_preconditioners()
        return ls.solve_batched(seg32, b, x, **solve_kw)

    for w in range(args.warmup):
        stats = step()
        log(f"warmup {w}: iterations min/median/max "
            f"{min(s['iters_exact'] for s in stats)}/{int(np.median([s['iters_exact'] for s in stats]))}/"
            f"{max(s['iters_exact'] for s in stats)}")
    lv_tuned = None
    if args.mas_levels == "auto":
        # untimed, like the warm-up: a MAS level count per system
        from libuipc_amd.tune import tune_mas_levels_batched

        lv_tuned = tune_mas_levels_batched(ls, 0, seg32, b, x, **solve_kw)
        log(f"per-system MAS levels: {np.bincount(lv_tuned['levels']).tolist()} systems at 0.. levels; "
            f"one count for all: {lv_tuned['uniform_best']}")
    elif args.mas_levels:
        ls.mas_set_active_levels(0, args.mas_levels)
    ev = EventTimer(st)
    ctx.synchronize()
    times = []
    for k in range(args.steps):
        ev.start()
        stats = step()
        times.append(ev.stop())
        log(f"step {k}: {times[-1]:.1f} ms")
    ms = float(np.median(times))
    its = np.array([s["iters_exact"] for s in stats])
    conv = int(sum(s["converged"] for s in stats))
    ev.start()
    ls.assemble(rows, cols, vals, st)
    ms_asm = ev.stop()
    ev.start()
    ls.assemble_preconditioners()
    ms_pc = ev.stop()
    print(json.dumps({
        "metric": "ms per batched Newton solve to 1e-4 rel. residual per system; PCG iters",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded generator, device Hessian producers)",
        "config": {"workload": f"{args.systems} independent FEM + contact systems, DoFs log-uniform in "
                               f"[{args.dof_lo:g}, {args.dof_hi:g}], one batched solve",
                   "systems": args.systems, "dofs": 3 * Btot, "block_rows": Btot, "triplets": T,
                   "grid_min": int(grids.min()), "grid_max": int(grids.max()), "preconditioner": "mas (union)",
                   "mas_levels": ({"per_system_counts": np.bincount(lv_tuned["levels"]).tolist(),
                                   "uniform_best": lv_tuned["uniform_best"]} if lv_tuned else args.mas_levels)},
        "pcg_iters_exact": {"min": int(its.min()), "median": int(np.median(its)), "max": int(its.max())},
        "converged": conv, "phase_ms": {"assemble": round(ms_asm, 3), "precond_assemble": round(ms_pc, 3),
                                        "pcg": round(ms - ms_asm - ms_pc, 3)},
        "ms_per_system_mean": round(ms / args.systems, 4),
    }))


def run_coupled(args) -> None:
    """The coupled heterogeneous system: 10,000 affine bodies (12-DoF blocks,
    ortho potential) + a 256 x 256 Baraff-Witkin cloth with dihedral bending
    + 500,000 vertex-triangle contacts on the cloth, the triplets written by
    the device producers; 12x12 block Jacobi on the bodies' range, MAS on the
    cloth's (partition from mi_mesh_partition, hierarchy built on the
    device).  One step = assemble + both preconditioner assemblies + PCG to
    rel. residual --tol from x0 = 0, matrix and MAS in --precision."""
    from libuipc_amd import abi, synth
    from libuipc_amd.abi import check
    from libuipc_amd.device import Context, LinearSystem
    from libuipc_amd.hip_events import EventTimer

    n_bodies, grid, n_contacts = 10_000, 256, 500_000
    ctx = Context(0)
    st = ctx.stream()
    a = synth.abd_inputs(n_bodies, seed=args.seed)
    c = synth.cloth_inputs(grid, seed=args.seed)
    hg = synth.hinges_of(c.tris)
    cst, cg, ckap = synth.cloth_contacts(grid, n_contacts, seed=args.seed)
    Bb, V = 4 * n_bodies, grid * grid
    B = Bb + V
    nt, nh, nc = len(c.tris), len(hg), len(cst)
    T = 16 * n_bodies + V + 6 * nt + 10 * nh + 10 * nc
    rows, cols, vals = ctx.alloc(T, np.int32), ctx.alloc(T, np.int32), ctx.alloc((T, 9), np.float64)
    r0, c0, v0 = rows.ptr.value, cols.ptr.value, vals.ptr.value
    dev = {k: ctx.to_device(v) for k, v in dict(q=a.q, dy=a.dyadic, vol=a.volume, ka=a.kappa, tris=c.tris,
                                                 X=c.rest_positions, x=c.positions, hg=hg, st=cst, g=cg, kap=ckap,
                                                 m=np.full(V, 200.0 * 2e-3 / V)).items()}
    lib, off = ctx.lib, 0
    check(lib.mi_produce_abd(ctx.h, n_bodies, dev["q"].ptr, dev["dy"].ptr, dev["vol"].ptr, dev["ka"].ptr, None, a.dt,
                             0, r0, c0, v0, st))
    off += 16 * n_bodies
    check(lib.mi_produce_kinetic(ctx.h, V, dev["m"].ptr, Bb, r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, st))
    off += V
    check(lib.mi_produce_bw_shell(ctx.h, nt, dev["tris"].ptr, dev["X"].ptr, dev["x"].ptr, c.stretch, c.shear,
                                  c.strain_rate, c.dt, Bb, r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, st))
    off += 6 * nt
    check(lib.mi_produce_bending(ctx.h, nh, dev["hg"].ptr, dev["X"].ptr, dev["x"].ptr, c.bending, c.dt, Bb,
                                 r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, st))
    off += 10 * nh
    check(lib.mi_produce_contact(ctx.h, nc, dev["st"].ptr, dev["g"].ptr, dev["kap"].ptr, c.dt, Bb, r0 + 4 * off,
                                 c0 + 4 * off, v0 + 72 * off, st))
    ctx.synchronize()
    parts, _ = ctx.mesh_partition(V, 16, tris=c.tris)
    prec = abi.MI_FP32 if args.precision == "fp32" else abi.MI_FP64
    ls = LinearSystem(ctx, B, prec)
    ls.assemble(rows, cols, vals, st)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG12, 0, 3 * Bb)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 3 * Bb, 3 * V, part_ids=parts, tris=c.tris,
                          build=abi.MI_MAS_BUILD_DEVICE, mas_precision=prec)
    b = ctx.to_device(np.random.default_rng(args.seed + 11).uniform(-1.0, 1.0, 3 * B))
    x = ctx.alloc(3 * B, np.float64)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=args.tol)
    log(f"coupled system: B={B} (bodies {Bb} block rows, cloth {V}), {T} triplets, {nc} contacts")

    def step():
        ls.assemble(rows, cols, vals, st)
        ls.assemble_preconditioners()
        return ls.solve(b, x, **kw)

    for _ in range(args.warmup):
        step()
    ev = EventTimer(st)
    ctx.synchronize()
    times = []
    for k in range(args.steps):
        ev.start()
        out = step()
        times.append(ev.stop())
        log(f"step {k}: {times[-1]:.2f} ms iters={out['iters']} exact={out['iters_exact']} "
            f"rel_res={out['rel_res']:.3e}")
    ms = float(np.median(times))
    ev.start()
    ls.assemble(rows, cols, vals, st)
    ms_asm = ev.stop()
    ev.start()
    ls.assemble_preconditioners()
    ms_pc = ev.stop()
    tr = ls.solve(b, x, true_residual=1, **kw)
    print(json.dumps({
        "metric": "ms per Newton solve to 1e-4 rel. residual; PCG iters",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-matrix/f64-vectors" if prec == abi.MI_FP32 else "f64",
        "data": "synthetic (seeded generator, device Hessian producers)",
        "config": {"workload": "coupled: 10,000 affine bodies + 256^2 cloth (stretch/shear + bending) + "
                               "500k contacts; diag12 on the bodies, MAS on the cloth",
                   "block_rows": B, "dofs": 3 * B, "triplets": T, "precision": args.precision},
        "pcg_iters": out["iters"], "pcg_iters_exact": out["iters_exact"], "rel_res_true": tr["rel_res_true"],
        "phase_ms": {"assemble": round(ms_asm, 3), "precond_assemble": round(ms_pc, 3),
                     "pcg": round(ms - ms_asm - ms_pc, 3)},
    }))


def run_learned(args) -> None:
    """16^3-vertex FEM systems plus --ws-contacts contact stencils (the jitter
    drawn from [0.03, 0.08] cells, Young's modulus log-uniform in [5e4, 2e5],
    the contact stiffness log-uniform in [1e4, kappa_hi] with kappa_hi
    log-uniform in [1e6, 1e8], rhs uniform in [-1, 1]), the triplets written
    by the device producers.  The
    learned initial guess (with its coarse-level correction on the MAS's
    first coarse level) is trained one system per Adam step on
    --train-systems seeds, then MAS-PCG to 1e-4 from x0 = 0, from the
    block-Jacobi guess D^-1 b (the network with its output at zero) and from
    the learned x0 (with its fallback) are compared on --test-systems held-out
    seeds.  Each system gets one untimed solve first and the three timed
    solves rotate their order; the learned solve's time includes the
    network's inference and the residual check of its guess."""
    from libuipc_amd import abi, synth
    from libuipc_amd.abi import check
    from libuipc_amd.device import Context, LinearSystem
    from libuipc_amd.hip_events import EventTimer
    from libuipc_amd.warmstart import WarmStartNet, build_graph, coarse_of_mas

    grid = 16
    ctx = Context(0)
    st = ctx.stream()
    nc = args.ws_contacts
    first = synth.producer_inputs(grid, nc, seed=args.seed)
    B, nt = len(first.mass), len(first.tets)
    T = B + 10 * nt + 10 * nc
    rows, cols, vals = ctx.alloc(T, np.int32), ctx.alloc(T, np.int32), ctx.alloc((T, 9), np.float64)
    r0, c0, v0 = rows.ptr.value, cols.ptr.value, vals.ptr.value
    dtets = ctx.to_device(first.tets)
    ls = LinearSystem(ctx, B)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * B, part_ids=first.part_ids, tets=first.tets)
    cmap = coarse_of_mas(ls, 0)  # the coarse-level correction's nodes: the MAS level 1

    def make(seed):
        rng = np.random.default_rng(seed)
        p = synth.producer_inputs(grid, nc, seed=seed, jitter=float(rng.uniform(0.03, 0.08)),
                                  youngs=float(np.exp(rng.uniform(np.log(5e4), np.log(2e5)))),
                                  kappa_hi=float(10.0 ** rng.uniform(6.0, 8.0)))
        dm, dX, dx = ctx.to_device(p.mass), ctx.to_device(p.rest_positions), ctx.to_device(p.positions)
        check(ctx.lib.mi_produce_kinetic(ctx.h, B, dm.ptr, 0, r0, c0, v0, st))
        check(ctx.lib.mi_produce_snh(ctx.h, nt, dtets.ptr, dX.ptr, dx.ptr, p.mu, p.lam, p.dt, 0, r0 + 4 * B,
                                     c0 + 4 * B, v0 + 72 * B, st))
        if nc:
            o = B + 10 * nt
            ci, cg, ck = (ctx.to_device(a) for a in (p.contact_idx, p.contact_g, p.contact_kappa))
            check(ctx.lib.mi_produce_contact(ctx.h, nc, ci.ptr, cg.ptr, ck.ptr, p.dt, 0, r0 + 4 * o, c0 + 4 * o,
                                             v0 + 72 * o, st))
        ls.assemble(rows, cols, vals, st)
        ctx.synchronize()  # the per-system inputs are freed next
        return build_graph(ctx, ls, rng.uniform(-1.0, 1.0, 3 * B), coarse=cmap)

    net = WarmStartNet(ctx, seed=args.seed)
    t0 = time.perf_counter()
    losses = []
    for k in range(args.train_systems):
        g = make(args.seed * 100_000 + k)
        losses.append(net.loss_and_grad(ls, g, loss=args.ws_loss))
        net.adam(lr=args.ws_lr)
        if (k + 1) % 500 == 0:
            log(f"trained on {k + 1} systems: mean {args.ws_loss} loss over the last 500: "
                f"{np.mean(losses[-500:]):.4f} ({time.perf_counter() - t0:.1f}s)")
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=args.tol, true_residual=1)
    ev = EventTimer(st)
    x = ctx.alloc(3 * B, np.float64)
    arms = ("x0=0", "jacobi", "learned")
    its = {a: [] for a in arms}
    ms = {a: [] for a in arms}
    used, res0, res0_j, ok = 0, [], [], 0

    def run(arm, g):
        ev.start()
        if arm == "x0=0":
            out = ls.solve(g.b, x, **kw)
        elif arm == "jacobi":
            ctx.lib.mi_memcpy_d2d(ctx.h, x.ptr, g.x_jacobi.ptr, 8 * 3 * B)
            out = ls.solve(g.b, x, x_init=1, **kw)
        else:
            out = net.solve(ls, g, x, **kw)
        ms[arm].append(ev.stop())
        its[arm].append(out["iters_exact"])
        return out

    for k in range(args.test_systems):
        g = make(args.seed * 100_000 + 50_000 + k)
        ls.assemble_preconditioners()
        ls.solve(g.b, x, **kw)
        ctx.synchronize()
        outs = {}
        for j in range(3):
            arm = arms[(k + j) % 3]
            outs[arm] = run(arm, g)
        w = outs["learned"]
        used += w["warm_used"]
        res0.append(w["warm_rel_res0"])
        ok += int(all(o["converged"] == 1 and o["rel_res_true"] <= 2 * args.tol for o in outs.values()))
    ctx.synchronize()
    ev.start()
    for _ in range(20):
        net.forward(g)
    infer = ev.stop() / 20
    mean = lambda v: round(float(np.mean(v)), 4)  # noqa: E731
    print(json.dumps({
        "metric": "MAS-PCG iterations and ms to 1e-4 rel. residual from x0 = 0 and from the learned x0",
        "value": mean(ms["learned"]), "unit": "ms", "n_gpus": 1, "steps": args.test_systems,
        "warmup": 0, "ms_per_step": mean(ms["learned"]), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solve, bf16 network", "data": "synthetic",
        "config": {"workload": f"learned initial guess, 16^3 FEM systems + {nc} contacts "
                               "(jitter, Young's modulus, contact stiffness perturbed)",
                   "train_systems": args.train_systems, "test_systems": args.test_systems, "width": 64,
                   "layers": 3, "loss": args.ws_loss, "lr": args.ws_lr},
        "iters_mean": {a: mean(its[a]) for a in arms}, "ms_mean": {a: mean(ms[a]) for a in arms},
        "inference_ms": round(infer, 4), "guess_used": used,
        "guess_rel_res0_median": round(float(np.median(res0)), 4), "all_converged": ok,
        "train_loss_first_last_500": [mean(losses[:500]), mean(losses[-500:])],
    }))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(s, precond, args, gpu_iters):
    """The oracle (C restatement of the reference's solver logic) on this host's
    cores, the whole solve: assembly, preconditioner build, and PCG from x0 = 0
    to the same tolerance (its own iteration count), nothing extrapolated.
    Threads: this process's CPU share, at most 16 (the GPU box gives one GPU's
    process 16 CPUs); the assembly and the MAS build are single-threaded
    restatements, the PCG runs its SpMV, preconditioner and vector loops on all
    of them."""
    from oracle import oracle as O

    threads = max(1, min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1),
                         int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    t = time.perf_counter()
    r, c, v, rp = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    if precond == "mas":
        m = O.Mas(s.block_rows, s.part_ids, s.tets)
        m.assemble(r, c, v)
        pc = [dict(kind=3, blk_off=0, blk_cnt=s.block_rows, mas=m)]
    else:
        pc = [dict(kind=1, blk_off=0, blk_cnt=s.block_rows, data=O.diag3(r, c, v, 0, s.block_rows))]
    t_pc = time.perf_counter() - t
    t = time.perf_counter()
    _, res = O.pcg(s.block_rows, r, c, v, s.rhs, pc, criterion=1, tol=args.tol, threads=threads)
    t_pcg = time.perf_counter() - t
    ms = 1e3 * (t_asm + t_pc + t_pcg)
    return {"value": round(ms, 1), "unit": "ms", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "iters": int(res["iters"]), "converged": int(res["converged"]),
            "phase_s": {"assemble": round(t_asm, 3), "precond_build": round(t_pc, 3), "pcg": round(t_pcg, 3)},
            "sample": f"the whole solve of the same system by the oracle: convert_sym ({t_asm:.2f}s, 1 thread) + "
                      f"{precond} build ({t_pc:.2f}s, 1 thread) + PCG to rel. residual {args.tol:g} "
                      f"({int(res['iters'])} iterations, {t_pcg:.2f}s on {threads} threads); GPU: {gpu_iters} "
                      f"iterations"}


if __name__ == "__main__":
    main()
