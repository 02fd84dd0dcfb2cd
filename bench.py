"""Throughput of the H/C transport step on one MI355X.

Workload (default): the 2048x4096 mesh on [0,0.5]x[0,1] with seeded H, C and
divergence-free face velocities, advanced with SSP-RK3 at CFL 0.4, each step
followed by the per-cell eutectic phase.  One "step" of this benchmark is one
SSP-RK3 step over the whole mesh (3 stage launches + phase).  The metric is
cell-steps per second in mixed precision; the plain FP64 path runs the same
steps for the ratio and for the mixed-vs-FP64 error.

Prints one JSON line (rank 0).  Only one GPU is used: --gpus above 1 exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nx", type=int, default=2048)
    p.add_argument("--ny", type=int, default=4096)
    p.add_argument("--rows", type=int, default=0, help="rows per work item (0: autotune)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain 1 GiB device copy (read 1 GiB, write 1 GiB) so that "
                        "FETCH_SIZE/WRITE_SIZE counters can be scaled")
    return p.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27                      # 2^27 doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


def algorithmic_bytes(q, state_b=8, vel_b=8, phase_b=8):
    """Bytes per cell of one step and of each stage launch.
    Stage 1 reads 2 fields and writes 2; stages 2 and 3 read 4 and write 2;
    every stage reads the cell's 2 owned faces x q velocity samples; the last
    stage writes 4 phase outputs and the region byte."""
    vel = 2 * q * vel_b
    s1 = 4 * state_b + vel
    s2 = 6 * state_b + vel
    s3 = 6 * state_b + vel + 4 * phase_b + 1
    return {"step": s1 + s2 + s3, "stage1": s1, "stage2": s2, "stage3": s3}


def cpu_baseline(args):
    import oracle as O
    from mantlepar_amd import workloads as W
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    spec = W.MeshSpec(args.nx, args.ny // 4, y1=0.25, seed=2)
    H = W.enthalpy(spec)
    Cc = W.composition(spec)
    vx, vy = W.face_velocity(spec)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, table_mode=1,
                    threads=threads)
    _, mr = o.rates()
    t0 = time.perf_counter()
    o.advance(H, Cc, 1, 0.4 / mr, 3)
    el = time.perf_counter() - t0
    return {"value": spec.cells / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"1 SSP-RK3 step (FP64, per-point polynomial rebuild, long-double "
                      f"accumulation) on a {spec.nx}x{spec.ny} band (y in [0,0.25]) with the "
                      f"same cell size as the {args.nx}x{args.ny} mesh; {el:.2f} s"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by mesh strips; see the design notes).", file=sys.stderr)
        sys.exit(2)
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W

    spec = W.MeshSpec(args.nx, args.ny, seed=2)
    H = W.enthalpy(spec)
    Cc = W.composition(spec)
    vx, vy = W.face_velocity(spec)
    phase = W.config_phase()
    res = {}
    for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64")):
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=phase, precision=prec)
        t.set_face_velocity(vx, vy)
        dt = t.stable_dt(0.4)
        if args.rows:
            t.set_tile(args.rows)
            rows_used = args.rows
        else:
            best = None
            for rows in (16, 32, 64, 128):
                t.set_tile(rows)
                t.upload(H, Cc)
                ms = t.profile(2, dt)["ms_total"]
                if best is None or ms < best[0]:
                    best = (ms, rows)
            t.set_tile(best[1])
            rows_used = best[1]
        t.upload(H, Cc)
        if args.warmup:
            t.step(args.warmup, dt)
        t.upload(H, Cc)
        t.synchronize()
        t0 = time.perf_counter()
        rep = t.step(args.steps, dt)
        el = time.perf_counter() - t0
        Hf, Cf = t.download()
        # kernel durations with HIP events on the handle's stream, over the same work
        t.upload(H, Cc)
        prof = t.profile(min(args.steps, 20), dt)
        res[name] = {"seconds": el, "report": rep, "H": Hf, "C": Cf, "profile": prof,
                     "rows": rows_used}
        t.close()

    N = spec.cells
    K = args.steps
    mixed, fp64 = res["mixed"], res["fp64"]
    value = N * K / mixed["seconds"]
    fp64_value = N * K / fp64["seconds"]
    errH = float(np.max(np.abs(mixed["H"] - fp64["H"])) / np.max(np.abs(fp64["H"])))
    errC = float(np.max(np.abs(mixed["C"] - fp64["C"])) / np.max(np.abs(fp64["C"])))
    b = algorithmic_bytes(spec.q)
    prof = mixed["profile"]
    stage_bytes = N * 0.5 * (b["stage1"] + b["stage2"])
    achieved = stage_bytes / (prof["ms_stage"] * 1e-3) / 1e9
    out = {
        "metric": "H/C cell-updates/sec, mixed vs FP64 at <=1e-6 rel err; serve p50/p99 ms",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": mixed["seconds"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"single mesh {spec.nx}x{spec.ny} on [0,0.5]x[0,1], SSP-RK3 "
                               "CFL 0.4, ML-WENO (3x3 Q2 + four 2x2 Q1), local LF, per-cell "
                               "eutectic phase each step",
                   "cells": N, "q": spec.q, "precision": "mixed (FP32 reconstruction/flux "
                   "corrections, FP64 state, first-order flux, phase and sums)"},
        "fp64": {"value": fp64_value, "ms_per_step": fp64["seconds"] / K * 1e3,
                 "mixed_over_fp64": value / fp64_value},
        "error_vs_fp64": {"H": errH, "C": errC, "steps": K},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stage_sweep_kernel<1,false> (stages 1-2)",
                     "bytes_per_cell_launch": 0.5 * (b["stage1"] + b["stage2"]),
                     "ms_per_launch": prof["ms_stage"], "rows_per_item": mixed["rows"]},
        "conservation": {"total_H": mixed["report"]["total_H"], "total_C": mixed["report"]["total_C"]},
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
