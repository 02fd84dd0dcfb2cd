"""Throughput of the H/C transport step on one MI355X.

Workload (default): the 2048x4096 mesh on [0,0.5]x[0,1] with seeded H, C and
divergence-free face velocities, advanced with SSP-RK3 at CFL 0.4, each step
followed by the per-cell eutectic phase.  One "step" of this benchmark is one
SSP-RK3 step over the whole mesh (3 stages + phase; on the mixed path the
three stages run in one launch).  The metric is
cell-steps per second in mixed precision; the plain FP64 path runs the same
steps for the ratio and for the mixed-vs-FP64 error.

Prints one JSON line (rank 0).  Only one GPU is used: --gpus above 1 exits.
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=100,
                   help="untimed steps right before the timed ones (the device's first "
                        "steps after idling run slower)")
    p.add_argument("--nx", type=int, default=2048)
    p.add_argument("--ny", type=int, default=4096)
    p.add_argument("--rows", type=int, default=0, help="rows per work item (0: autotune)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--batch-steps", type=int, default=50)
    p.add_argument("--serve-requests", type=int, default=400)
    p.add_argument("--serve-rate", type=float, default=300.0, help="Poisson arrivals per second")
    p.add_argument("--serve-rates", default="300,1000,2000,4000",
                   help="arrival rates of the serving sweep (comma-separated; empty: none)")
    p.add_argument("--full-steps", type=int, default=1000,
                   help="steps of the 8192x8192 full-chip run (0: skip it)")
    p.add_argument("--limiter-steps", type=int, default=20,
                   help="steps of the FP64 bound-limiter leg (0: skip)")
    p.add_argument("--coupled-steps", type=int, default=3,
                   help="SSPRK2 steps of the coupled-law leg (0: skip)")
    p.add_argument("--coupled-batch", type=int, default=64,
                   help="meshes of the batched coupled-law leg (0: skip)")
    p.add_argument("--coupled-nx", type=int, default=1024)
    p.add_argument("--coupled-ny", type=int, default=2048)
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--tune-cache", default=os.path.join(tempfile.gettempdir(), "mantlepar_batch_tune.json"),
                   help="JSON file where the batch autotuner caches its choices")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain 1 GiB device copy (read 1 GiB, write 1 GiB) so that "
                        "FETCH_SIZE/WRITE_SIZE counters can be scaled")
    return p.parse_args()


# step32_kernel at 2048x4096, q = 5, by rows per item: (FETCH_SIZE, WRITE_SIZE)
# in KiB per launch; HBM bytes = (FETCH x 2 + WRITE) x 1024 by the copy calibration
MEASURED_STEP_LAUNCH_KIB = {8: (922321, 411511), 12: (862611, 410324), 16: (842528, 408546),
                            24: (810040, 407796), 32: (805875, 407424)}


def copy_calibration():
    import torch
    n = 1 << 27                      # 2^27 doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


_COPY_PROBE = """
import json, torch
n = 1 << 27
x = torch.ones(n, dtype=torch.float64, device="cuda")
y = torch.empty_like(x)
for _ in range(3):
    torch.mul(x, 1.0, out=y)
s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
s.record()
for _ in range(10):
    torch.mul(x, 1.0, out=y)
e.record()
torch.cuda.synchronize()
print(json.dumps(2 * n * 8 / (s.elapsed_time(e) / 10 * 1e-3) / 1e9))
"""


def practical_copy_gbs():
    """HBM bandwidth a plain device copy reaches on this GPU (1 GiB read + 1 GiB
    written, 10 launches timed with events), measured in a child process
    before this one touches the GPU: the practical ceiling beside the 8 TB/s
    published peak."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", _COPY_PROBE], capture_output=True, text=True,
                       timeout=180)
    return float(r.stdout.strip().splitlines()[-1])


# Byte width of a face-velocity sample in the algorithmic byte count: FP64
# samples for the FP64 path, FP32 for the mixed one (FP64 state, FP32
# velocity: 280 B per cell-step at q = 5).  The mixed sweep reads 16-byte
# one-sided face records, 3.2 B per sample at q = 5, so the FP32 count does
# not flatter it.
VEL_BYTES = {"fp64": 8, "mixed": 4}


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


def launch_bytes(q, one_launch):
    """Algorithmic bytes per cell of the mixed path's dominant launch (FP32
    velocity samples, then FP64 samples) and its name.  With one launch per
    step the kernel runs the three stages (the phase runs in its own kernel
    beside the next step): the three stages' state and velocity bytes, without
    the phase outputs.  Otherwise the stage-1/2 launch, mean of the two."""
    b = algorithmic_bytes(q, vel_b=VEL_BYTES["mixed"])
    b64 = algorithmic_bytes(q, vel_b=VEL_BYTES["fp64"])
    if one_launch:
        phase_out = 4 * 8 + 1
        return (b["stage1"] + b["stage2"] + b["stage3"] - phase_out,
                b64["stage1"] + b64["stage2"] + b64["stage3"] - phase_out,
                "step32_kernel<false,false> (mixed SSP-RK3 step, its three stages in one launch)")
    return (0.5 * (b["stage1"] + b["stage2"]), 0.5 * (b64["stage1"] + b64["stage2"]),
            "stage32_kernel<false,false> (mixed sweep, stages 1-2)")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(args):
    """The oracle (FP64, the reference's algorithm) on the host, timed on
    bounded samples of the headline workload: all the threads this process
    may use (capped at the 16 threads that a one-GPU share of the host gets)
    and one thread."""
    import oracle as O
    from mantlepar_amd import workloads as W
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = args.cpu_threads or min(16, avail)
    out = {"unit": "cell-steps/s", "kind": "port", "cpu_model": _cpu_model(),
           "host_cpus": os.cpu_count(), "affinity_cpus": avail}
    # the same cell size as the headline mesh: a band of its rows
    for label, nthreads, rows in (("all", threads, args.ny // 4), ("one", 1, args.ny // 32)):
        spec = W.MeshSpec(args.nx, rows, y1=rows / args.ny, seed=2)
        H = W.enthalpy(spec)
        Cc = W.composition(spec)
        vx, vy = W.face_velocity(spec)
        o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, table_mode=1,
                        threads=nthreads)
        _, mr = o.rates()
        t0 = time.perf_counter()
        o.advance(H, Cc, 1, 0.4 / mr, 3)
        el = time.perf_counter() - t0
        sample = (f"1 SSP-RK3 step (FP64, per-point polynomial rebuild, long-double "
                  f"accumulation) on a {spec.nx}x{spec.ny} band with the cell size of the "
                  f"{args.nx}x{args.ny} mesh, {nthreads} thread(s); {el:.2f} s")
        if label == "all":
            out.update({"value": spec.cells / el, "cores": nthreads, "sample": sample})
        else:
            out["one_thread"] = {"value": spec.cells / el, "cores": 1, "sample": sample}
    out["note"] = ("cores = OpenMP threads used; the host is shared, and a one-GPU share is "
                   "16 threads, so larger thread counts are not used")
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batch_inputs(count=256, seed=3):
    from mantlepar_amd import workloads as W
    specs = W.batch_specs(count, seed=seed)
    return [(s, W.enthalpy(s), W.composition(s), W.face_velocity(s)) for s in specs]


def bench_batch(args, meshes):
    """The 256-mesh batch: one launch per stage, or per step (mixed), over all
    meshes, autotuned schedule, hipGraph replay; mixed and FP64.  The profile
    times per-stage launches one by one.  The plain FP64 GPU path runs the
    same meshes one handle at a time."""
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W
    K = args.batch_steps
    cells = sum(s.cells for s, *_ in meshes)
    res = {}
    rng = np.random.default_rng(7)
    for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64"), (M.PREC_FP64, "fp64p")):
        b = M.Batch(len(meshes), phase=W.config_phase(), precision=prec)
        ids = []
        inputs = []
        for s, H, Cc, (vx, vy) in meshes:
            if name == "fp64p":    # FP64 from an FP32-sized perturbation of H and C
                H = H * (1.0 + 2.0 ** -24 * rng.standard_normal(H.shape))
                Cc = Cc * (1.0 + 2.0 ** -24 * rng.standard_normal(Cc.shape))
            inputs.append((H, Cc))
            k = b.add_mesh(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q)
            b.upload(k, H, Cc)
            b.set_face_velocity(k, vx, vy)
            b.set_dt(k, b.stable_dt(k, 0.4))
            ids.append(k)
        os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
        tune = b.autotune(2, cache_path=args.tune_cache)
        tune["per_step_launch"] = b.per_step_launch()
        b.step(2, graph=True)                        # warm-up and graph capture of 2 steps
        for (H, Cc), k in zip(inputs, ids):
            b.upload(k, H, Cc)
        # capture the timed shape and run it once as the warm-up, then time a
        # replay straight after it (no upload between: the device's first steps
        # after idling run slower)
        b.step(K, graph=True)
        t0 = time.perf_counter()
        b.step(K, graph=True)
        el = time.perf_counter() - t0
        # the states compared across precisions: K steps from the inputs
        for (H, Cc), k in zip(inputs, ids):
            b.upload(k, H, Cc)
        b.step(K, graph=True)
        prof = b.profile(min(K, 10))
        res[name] = {"seconds": el, "tune": tune, "profile": prof,
                     "states": [b.download(k) for k in ids], "info": b.info()}
        b.close()
        log(f"batch {name}: {cells * K / el:.3e} cell-steps/s, schedule {tune}")
    # plain FP64 drop-in path: one mp_transport handle per mesh, run in turn
    handles = []
    for s, H, Cc, (vx, vy) in meshes:
        t = M.Transport(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q, phase=W.config_phase(),
                        precision=M.PREC_FP64)
        t.upload(H, Cc)
        t.set_face_velocity(vx, vy)
        handles.append((t, t.stable_dt(0.4)))
    for t, dt in handles:
        t.step(1, dt)
    for (s, H, Cc, _), (t, dt) in zip(meshes, handles):
        t.upload(H, Cc)
    t0 = time.perf_counter()
    for t, dt in handles:
        t.step(K, dt, M.SCHEME_SSPRK3)
    el_plain = time.perf_counter() - t0
    for t, _ in handles:
        t.close()
    log(f"batch plain fp64 handles: {cells * K / el_plain:.3e} cell-steps/s")
    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    # Accuracy against FP64, split by the FP64 method's own conditioning: a
    # mesh is "stable" when FP64 runs from (H, C) and from (H, C)(1 + 2^-24 xi)
    # (max relative difference near 3e-7 at the start) stay within 1e-6 of
    # each other after the same steps.
    err = {"H": 0.0, "C": 0.0, "H_stable": 0.0, "C_stable": 0.0, "unstable_meshes": 0,
           "fp64_self_spread_max": 0.0}
    for (hm, cm), (h6, c6), (hp, cp) in zip(res["mixed"]["states"], res["fp64"]["states"],
                                            res["fp64p"]["states"]):
        eh, ec = rel(hm, h6), rel(cm, c6)
        spread = max(rel(hp, h6), rel(cp, c6))
        err["H"], err["C"] = max(err["H"], eh), max(err["C"], ec)
        err["fp64_self_spread_max"] = max(err["fp64_self_spread_max"], spread)
        if spread <= 1e-6:
            err["H_stable"], err["C_stable"] = max(err["H_stable"], eh), max(err["C_stable"], ec)
        else:
            err["unstable_meshes"] += 1
    mixed_v = cells * K / res["mixed"]["seconds"]
    fp64_v = cells * K / res["fp64"]["seconds"]
    plain_v = cells * K / el_plain
    return {
        "workload": f"{len(meshes)} meshes, nx in {{40..400}} step 20, ny log-uniform [5,800], "
                    f"{K} SSP-RK3 steps timed (a hipGraph replay after {K} untimed ones)",
        "meshes": len(meshes), "cells": cells, "steps": K,
        "mixed": {"value": mixed_v, "ms_per_step": res["mixed"]["seconds"] / K * 1e3,
                  "schedule": res["mixed"]["tune"], "profile": res["mixed"]["profile"]},
        "fp64": {"value": fp64_v, "ms_per_step": res["fp64"]["seconds"] / K * 1e3,
                 "schedule": res["fp64"]["tune"]},
        "plain_fp64_handles": {"value": plain_v, "ms_per_step": el_plain / K * 1e3},
        "mixed_over_plain_fp64": mixed_v / plain_v,
        "mixed_over_batched_fp64": mixed_v / fp64_v,
        "error_vs_fp64": err,
        "unit": "cell-steps/s",
    }


def _serve_once(M, W, meshes, prec, rate, n, seed=17):
    """n Poisson arrivals at `rate` per second into a fresh server; its stats."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(meshes), n)
    steps = rng.integers(20, 201, n)
    gaps = rng.exponential(1.0 / rate, n)
    srv = M.Server(max_slots=512, phase=W.config_phase(), precision=prec,
                   hbm_budget_bytes=200 * 10**9, max_epoch_steps=8)
    t_next = time.perf_counter()
    ids = []
    for i in range(n):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        s, H, Cc, (vx, vy) = meshes[picks[i]]
        ids.append(srv.submit(s.nx, s.ny, H, Cc, vx, vy, int(steps[i]), 0.4,
                              s.x0, s.x1, s.y0, s.y1, q=s.q, outputs=("H", "C")))
    srv.drain()
    st = srv.stats()
    t, w = srv.progress()
    for rid in ids:
        srv.wait(rid)
    srv.close()
    # steady throughput: the slope of the device-work curve between 20% and
    # 80% of the run's work, away from the arrival ramp and the final drain
    steady = None
    if len(t) > 4 and w[-1] > 0:
        a, b = int(np.searchsorted(w, 0.2 * w[-1])), int(np.searchsorted(w, 0.8 * w[-1]))
        if b > a and t[b] > t[a]:
            steady = float((w[b] - w[a]) / (t[b] - t[a]))
    return {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "max_ms": st["max_ms"], "cell_steps_per_s": st["cell_steps"] / st["elapsed_s"],
            "cell_steps_per_s_steady": steady,
            "completed": st["completed"], "failed": st["failed"],
            "max_resident": st["max_resident"], "epochs": st["epochs"],
            "hbm_peak_bytes": st["hbm_peak"]}


def bench_serve(args, meshes, batch=None):
    """Poisson arrivals of transport runs (meshes of the batch distribution,
    20-200 seeded SSP-RK3 steps), continuous batching under a 200 GB HBM
    budget; p50/p99 submit-to-done latency and aggregate cell-steps/s, at the
    default rate and over a sweep of rates towards saturation, where the
    throughput is set beside the batch leg's (the same meshes advanced as one
    batch)."""
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W
    n = args.serve_requests
    out = {"requests": n, "rate_per_s": args.serve_rate, "hbm_budget_bytes": 200 * 10**9,
           "workload": "Poisson arrivals; mesh drawn from the 256-mesh batch pool; "
                       "20-200 SSP-RK3 steps; CFL 0.4"}
    for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64")):
        out[name] = _serve_once(M, W, meshes, prec, args.serve_rate, n)
        log(f"serve {name}: p50 {out[name]['p50_ms']:.2f} ms p99 {out[name]['p99_ms']:.2f} ms, "
            f"{out[name]['cell_steps_per_s']:.3e} cell-steps/s")
    rates = [float(r) for r in args.serve_rates.split(",") if r]
    out["rate_sweep_requests"] = {f"{r:.0f}": max(n, int(0.5 * r)) for r in rates}
    sweep = []
    for rate in rates:
        row = {"rate_per_s": rate}
        for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64")):
            # enough requests for about half a second of arrivals, so the
            # run reaches its steady state between the ramp and the drain
            r = _serve_once(M, W, meshes, prec, rate, max(n, int(0.5 * rate)), seed=17 + int(rate))
            if batch and name in batch:
                r["of_batch_throughput"] = r["cell_steps_per_s"] / batch[name]["value"]
                if r["cell_steps_per_s_steady"]:
                    r["of_batch_throughput_steady"] = r["cell_steps_per_s_steady"] / batch[name]["value"]
            row[name] = r
            log(f"serve {name} at {rate:.0f}/s: p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms, "
                f"{r['cell_steps_per_s']:.3e} cell-steps/s (steady {r['cell_steps_per_s_steady'] or 0:.3e}), "
                f"max resident {r['max_resident']}")
        sweep.append(row)
    out["rate_sweep"] = sweep
    return out


def _threaded(fn, n, threads=16):
    """Run fn(lo, hi) over [0, n) in row blocks on a thread pool (numpy
    releases the GIL inside its loops)."""
    from concurrent.futures import ThreadPoolExecutor
    step = max(1, (n + threads - 1) // threads)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda lo: fn(lo, min(n, lo + step)), range(0, n, step)))


def bench_full_chip(args):
    """8192x8192 cells, SSP-RK3 at CFL 0.4, face velocities refreshed every 50
    steps (psi amplitude drift +-10%) plus a through-flow leaving through the
    outflow-only top segment x in [0, 0.025]; mixed and FP64.  Reports the
    stepping throughput (refresh uploads timed apart), the conservation
    balance of each run and the mixed-vs-FP64 drift of the global totals."""
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W
    spec, K, cfl, every = W.config5()
    K = args.full_steps
    nref = max(1, K // every)
    H = W.enthalpy(spec)
    Cc = W.composition(spec)
    bx, by = W.face_velocity(spec)
    tx, ty = W.throughflow_velocity(spec)
    # the through-flow is zero outside a corner block: keep only that block
    jx = np.nonzero(np.abs(tx).max(axis=(0, 2)))[0]
    ix = np.nonzero(np.abs(tx).max(axis=(0, 1)))[0]
    jy = np.nonzero(np.abs(ty).max(axis=(0, 2)))[0]
    iy = np.nonzero(np.abs(ty).max(axis=(0, 1)))[0]
    bx_blk = (slice(None), slice(jx[0], jx[-1] + 1), slice(ix[0], ix[-1] + 1))
    by_blk = (slice(None), slice(jy[0], jy[-1] + 1), slice(iy[0], iy[-1] + 1))
    tx_blk, ty_blk = tx[bx_blk].copy(), ty[by_blk].copy()
    del tx, ty
    bc, inflow = W.throughflow_boundary(spec)
    amps = W.refresh_amplitudes(nref, seed=spec.seed)
    vx, vy = np.empty_like(bx), np.empty_like(by)

    def refresh(a):
        _threaded(lambda lo, hi: np.multiply(bx[:, lo:hi], a, out=vx[:, lo:hi]), bx.shape[1])
        _threaded(lambda lo, hi: np.multiply(by[:, lo:hi], a, out=vy[:, lo:hi]), by.shape[1])
        vx[bx_blk] += tx_blk
        vy[by_blk] += ty_blk

    area = spec.dx * spec.dy
    res = {}
    rng = np.random.default_rng(11)
    for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64"), (M.PREC_FP64, "fp64p")):
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=W.config_phase(), precision=prec)
        if name == "fp64p":
            # FP64 from (H, C)(1 + 2^-24 xi): the method's own sensitivity to
            # differences of FP32 size, against which pointwise errors are read
            t.upload(H * (1.0 + 2.0 ** -24 * rng.standard_normal(H.shape)),
                     Cc * (1.0 + 2.0 ** -24 * rng.standard_normal(Cc.shape)))
        else:
            t.upload(H, Cc)
        t.set_boundary(bc["bottom"], bc["right"], bc["top"], bc["left"])
        t.set_inflow(*inflow)
        tot0 = {"total_H": float(np.sum(H)) * area, "total_C": float(np.sum(Cc)) * area}
        b_h = b_c = 0.0
        t_step = t_refresh = 0.0
        rep = tot0
        for k in range(nref):
            t0 = time.perf_counter()
            refresh(amps[k])
            t.set_face_velocity(vx, vy)
            dt = t.stable_dt(cfl)
            t.synchronize()
            t1 = time.perf_counter()
            rep = t.step(min(every, K - k * every), dt)
            t2 = time.perf_counter()
            t_refresh += t1 - t0
            t_step += t2 - t1
            b_h += rep["boundary_H"]
            b_c += rep["boundary_C"]
        Hf, Cf = t.download()
        ph = t.phase_state()
        prof = t.profile(4, dt)
        one_launch = t.layout()["one_launch_steps"]
        t.close()
        res[name] = {"t_step": t_step, "t_refresh": t_refresh, "H": Hf, "C": Cf, "T": ph["T"],
                     "phi": ph["phi"], "tot0": (tot0["total_H"], tot0["total_C"]),
                     "tot1": (rep["total_H"], rep["total_C"]), "b": (b_h, b_c), "profile": prof,
                     "one_launch": one_launch}
        log(f"full-chip {name}: {spec.cells * K / t_step:.3e} cell-steps/s "
            f"(refresh {t_refresh:.1f} s)")

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    mx, fp, fpp = res["mixed"], res["fp64"], res.pop("fp64p")
    out = {"workload": f"{spec.nx}x{spec.ny} cells, {K} SSP-RK3 steps, CFL 0.4, velocities "
                       f"refreshed every {every} steps (amplitude drift +-10%), through-flow "
                       "out of the top segment x in [0, 0.025] (outflow-only policy)",
           "cells": spec.cells, "steps": K, "unit": "cell-steps/s"}
    for name, r in res.items():
        out[name] = {
            "value": spec.cells * K / r["t_step"],
            "ms_per_step": r["t_step"] / K * 1e3,
            "refresh_s_total": r["t_refresh"],
            # one launch per step: that launch and its phase kernel; otherwise
            # a stage launch and the last stage (with the phase on FP64)
            "one_launch_steps": r["one_launch"],
            "ms_stage": r["profile"]["ms_stage"],
            "ms_phase_stage": r["profile"]["ms_phase_stage"],
            # total(end) - total(start) + outward boundary flux: zero up to rounding
            "conservation_defect_rel": {
                "H": (r["tot1"][0] - r["tot0"][0] + r["b"][0]) / r["tot0"][0],
                "C": (r["tot1"][1] - r["tot0"][1] + r["b"][1]) / r["tot0"][1]},
            "outflow": {"H": r["b"][0], "C": r["b"][1]},
        }
    out["drift_vs_fp64"] = {
        "total_H_rel": (mx["tot1"][0] - fp["tot1"][0]) / fp["tot1"][0],
        "total_C_rel": (mx["tot1"][1] - fp["tot1"][1]) / fp["tot1"][1],
        "state_max_rel": {"H": rel(mx["H"], fp["H"]), "C": rel(mx["C"], fp["C"]),
                          "T": rel(mx["T"], fp["T"]), "phi": float(np.abs(mx["phi"] - fp["phi"]).max())},
        "cells_H_diff_above_1e-6": int(np.count_nonzero(np.abs(mx["H"] - fp["H"]) > 1e-6)),
        # the same measures between FP64 runs whose initial states differ by 2^-24
        "fp64_self_spread": {"H": rel(fpp["H"], fp["H"]), "C": rel(fpp["C"], fp["C"]),
                             "T": rel(fpp["T"], fp["T"]),
                             "phi": float(np.abs(fpp["phi"] - fp["phi"]).max()),
                             "cells_H_diff_above_1e-6":
                                 int(np.count_nonzero(np.abs(fpp["H"] - fp["H"]) > 1e-6))},
    }
    launch_b, _, kernel = launch_bytes(spec.q, mx["one_launch"])
    out["roofline_launch"] = {"kernel": kernel, "ms_per_launch": mx["profile"]["ms_stage"],
                              "bytes_per_cell_launch": launch_b,
                              "achieved": spec.cells * launch_b / (mx["profile"]["ms_stage"] * 1e-3) / 1e9,
                              "unit": "GB/s"}
    return out


def bench_limiter(args, spec, H, Cc, vx, vy, phase, prec=None):
    """The sweep (FP64, or mixed with per-stage launches, as the limiter runs)
    with the bound limiter of the coupled reconstruction on every stage,
    beside the same sweep without it, on the single-mesh workload."""
    import mantlepar_amd as M
    prec = M.PREC_FP64 if prec is None else prec
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    phase=phase, precision=prec)
    t.set_face_velocity(vx, vy)
    dt = t.stable_dt(0.4)
    t.set_tile(16)
    if prec == M.PREC_MIXED:
        t.set_launch(False)
    out = {}
    for on in (False, True):
        t.set_limiter(on)
        t.upload(H, Cc)
        t.step(2, dt)
        t.upload(H, Cc)
        t.synchronize()
        t0 = time.perf_counter()
        t.step(args.limiter_steps, dt)
        el = time.perf_counter() - t0
        t.upload(H, Cc)
        prof = t.profile(min(args.limiter_steps, 10), dt)
        out["limited" if on else "unlimited"] = {
            "cell_steps_per_s": spec.cells * args.limiter_steps / el,
            "ms_stage": prof["ms_stage"], "ms_phase_stage": prof["ms_phase_stage"]}
    t.upload(H, Cc)
    tH, tC = t.limits()
    t.close()
    out["cells_limited_H"] = int(np.count_nonzero(tH < 1))
    out["cells_limited_C"] = int(np.count_nonzero(tC < 1))
    out["stage_overhead"] = out["limited"]["ms_stage"] / out["unlimited"]["ms_stage"]
    out["steps"] = args.limiter_steps
    out["points_per_cell_field"] = 86
    name = "mixed" if prec == M.PREC_MIXED else "fp64"
    log(f"limiter {name}: {out['limited']['cell_steps_per_s']:.3e} cell-steps/s, "
        f"stage x{out['stage_overhead']:.2f}")
    return out


def bench_diffusion(args, spec, H, Cc, vx, vy, phase):
    """Mixed and FP64 sweeps with the legacy preheat thermal diffusion of H
    folded into the face fluxes, beside the same sweeps without it."""
    import mantlepar_amd as M
    out = {}
    kappa = 1e-6
    for prec, name in ((M.PREC_MIXED, "mixed"), (M.PREC_FP64, "fp64")):
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=phase, precision=prec)
        t.set_face_velocity(vx, vy)
        dt = t.stable_dt(0.4)
        res = {}
        for k in (0.0, kappa):
            t.set_diffusion(k, 4, 0.9)
            t.upload(H, Cc)
            # mixed: the one-launch step; FP64: one stage launch
            res["diffusion" if k else "advection_only"] = t.profile(10, dt)["ms_stage"]
        res["stage_overhead"] = res["diffusion"] / res["advection_only"]
        out[name] = res
        t.close()
    out["kappa"] = kappa
    log(f"diffusion: mixed stage x{out['mixed']['stage_overhead']:.3f}, fp64 x{out['fp64']['stage_overhead']:.3f}")
    return out


def bench_coupled(args):
    """The coupled transport law (FP64): SSPRK2 steps with the limiter, phase-based
    fluxes, diffusion of T and the cooling source, flow samples held fixed; and
    the oracle's restatement on a bounded sample of the same workload."""
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W
    spec = W.MeshSpec(args.coupled_nx, args.coupled_ny, seed=2)
    flow = W.coupled_flow(spec)
    bc, inflow, tt, tv = W.coupled_boundary(spec)
    H, Cc = W.coupled_state(spec)
    g = M.Coupled(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                  phase=W.config_phase(hypereutectic=0),
                  params=M.default_coupled(theta=1.0, cooling=0.05, thermal_diffusivity=2e-4))
    g.set_flow(flow)
    g.set_boundary(0, bc)
    g.set_boundary(1, bc)
    g.set_inflow(*inflow)
    g.set_thermal_boundary(tt, tv)
    g.upload(H, Cc)
    g.run(1, cfl=0.5, initial=1e-3, maximum=1.0)            # warm-up
    g.upload(H, Cc)
    t0 = time.perf_counter()
    _, steps = g.run(args.coupled_steps, cfl=0.5, initial=1e-3, maximum=1.0)
    el = time.perf_counter() - t0
    g.close()
    out = {"cells": spec.cells, "steps": len(steps), "seconds": el,
           "cell_steps_per_s": spec.cells * len(steps) / el,
           "rhs_per_step": 2, "retries": sum(s["retries"] for s in steps),
           "workload": f"{spec.nx}x{spec.ny}, flow samples fixed, limiter on, kappa 2e-4, "
                       "cooling 0.05, Dirichlet bottom temperature, bottom inflow, top outflow"}
    log(f"coupled: {out['cell_steps_per_s']:.3e} cell-steps/s ({len(steps)} SSPRK2 steps)")
    if args.coupled_batch > 0:
        out["batch"] = bench_coupled_batch(args)
    if not args.no_cpu:
        try:
            import oracle as O
            sm = W.MeshSpec(64, 64, seed=2)
            sflow = W.coupled_flow(sm)
            sbc, sin_, _, stv = W.coupled_boundary(sm)
            sH, sC = W.coupled_state(sm)
            z = (np.zeros((sm.q, sm.ny, sm.nx + 1)), np.zeros((sm.q, sm.ny + 1, sm.nx)))
            o = O.Transport(W.vertex_coords(sm), sm.nx, sm.ny, *z, q=sm.q, bc=sbc,
                            inflow_H=sin_[0], inflow_C=sin_[1])
            o.set_limiter(True, 5, 4, 0.9, 0.4)
            o.set_diffusion(2e-4, 4, 0.9, types={"bottom": np.ones(sm.nx)}, values=stv)
            pm = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
            cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=1.0, cooling=0.05, kappa=2e-4)
            t0 = time.perf_counter()
            _, _, _, st = o.coupled_run(pm, cp, sflow, sH, sC, 2, cfl=0.5, initial=1e-3, maximum=1.0)
            ce = time.perf_counter() - t0
            out["cpu_port"] = {"cell_steps_per_s": sm.cells * len(st) / ce, "cores": 1,
                               "sample": f"64x64, {len(st)} steps", "seconds": ce}
        except Exception as e:   # reported, never required
            out["cpu_port"] = {"error": str(e)}
    return out


def bench_coupled_batch(args):
    """The coupled law on many small meshes (the batch size law): every handle
    advanced by mp_coupled_run_batch, one launch per kernel over all of them
    and one read-back per step attempt, against the same handles run one after
    another (mp_coupled_run each), from the same states."""
    import mantlepar_amd as M
    from mantlepar_amd import workloads as W
    specs = W.batch_specs(args.coupled_batch, seed=11)
    hs, states = [], []
    for sp in specs:
        bc, inflow, tt, tv = W.coupled_boundary(sp)
        g = M.Coupled(sp.nx, sp.ny, sp.x0, sp.x1, sp.y0, sp.y1, q=sp.q,
                      phase=W.config_phase(hypereutectic=0),
                      params=M.default_coupled(theta=1.0, cooling=0.05, thermal_diffusivity=2e-4))
        g.set_flow(W.coupled_flow(sp))
        g.set_boundary(0, bc)
        g.set_boundary(1, bc)
        g.set_inflow(*inflow)
        g.set_thermal_boundary(tt, tv)
        H, Cc = W.coupled_state(sp)
        g.upload(H, Cc)
        hs.append(g)
        states.append((H, Cc))
    tc = dict(cfl=0.5, initial=1e-3, maximum=1.0)
    nsteps = args.coupled_steps
    cells = sum(sp.cells for sp in specs)
    M.coupled_run_batch(hs, 1, tc)                          # warm-up
    for g, (H, Cc) in zip(hs, states):
        g.upload(H, Cc)
    t0 = time.perf_counter()
    res = M.coupled_run_batch(hs, nsteps, tc)
    el_b = time.perf_counter() - t0
    done_b = sum(len(r[2]) * sp.cells for r, sp in zip(res, specs))
    ends = [g.download() for g in hs]
    for g, (H, Cc) in zip(hs, states):
        g.upload(H, Cc)
    t0 = time.perf_counter()
    done_s, same = 0, True
    for k, (g, sp) in enumerate(zip(hs, specs)):
        try:
            _, rec = g.run(nsteps, **tc)
            done_s += len(rec) * sp.cells
        except M.MantleError:
            pass
    el_s = time.perf_counter() - t0
    for g, (Hb, Cb) in zip(hs, ends):
        Hs, Cs = g.download()
        same &= bool(np.array_equal(Hs, Hb) and np.array_equal(Cs, Cb))
    for g in hs:
        g.close()
    out = {"meshes": len(specs), "cells": cells, "steps": nsteps,
           "failed": sum(1 for r in res if r[0] != 0),
           "batched": {"seconds": el_b, "cell_steps_per_s": done_b / el_b},
           "one_by_one": {"seconds": el_s, "cell_steps_per_s": done_s / el_s},
           "speedup": el_s / el_b, "states_bitwise_equal": same,
           "workload": f"{len(specs)} meshes of the batch size law (seed 11), the coupled workload "
                       "of the single-mesh leg on each"}
    log(f"coupled batch: {out['batched']['cell_steps_per_s']:.3e} cell-steps/s batched, "
        f"{out['one_by_one']['cell_steps_per_s']:.3e} one by one (x{out['speedup']:.2f}), equal {same}")
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by mesh strips; see the design notes).", file=sys.stderr)
        sys.exit(2)
    try:
        copy_gbs = practical_copy_gbs()
    except Exception as e:  # informative only
        copy_gbs = None
        log(f"copy bandwidth not measured: {e}")
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
            for rows in (8, 12, 16, 24, 32):
                t.set_tile(rows)
                t.upload(H, Cc)
                ms = min(t.profile(4, dt)["ms_total"] for _ in range(3))
                if best is None or ms < best[0]:
                    best = (ms, rows)
            t.set_tile(best[1])
            rows_used = best[1]
        # W untimed steps from the initial state, then exactly K timed steps
        # straight after them: an upload between the two left the device idle
        # for the PCIe copy, and its first timed steps ran ~12% slower while it
        # came back to speed.  Both precisions run the same W + K steps.
        t.upload(H, Cc)
        if args.warmup:
            t.step(args.warmup, dt)
        # timed as up to five windows with a synchronisation between them; the
        # windows give the spread of the per-step time
        t.synchronize()
        nwin = min(5, args.steps)
        sizes = [args.steps // nwin + (1 if k < args.steps % nwin else 0) for k in range(nwin)]
        windows = []
        for n in sizes:
            t0 = time.perf_counter()
            rep = t.step(n, dt)
            windows.append((time.perf_counter() - t0) / n)
        el = sum(w * n for w, n in zip(windows, sizes))
        Hf, Cf = t.download()
        ph = t.phase_state()
        # kernel durations with HIP events on the handle's stream, over the same work
        t.upload(H, Cc)
        prof = t.profile(min(args.steps, 20), dt)
        res[name] = {"seconds": el, "report": rep, "H": Hf, "C": Cf, "T": ph["T"],
                     "phi": ph["phi"], "region": ph["region"], "profile": prof,
                     "rows": rows_used, "windows_ms": sorted(w * 1e3 for w in windows),
                     "windows_ms_in_order": [w * 1e3 for w in windows],
                     "layout": t.layout()}
        t.close()
        log(f"single {name}: {spec.cells * args.steps / el:.3e} cell-steps/s (rows {rows_used})")

    N = spec.cells
    K = args.steps
    mixed, fp64 = res["mixed"], res["fp64"]
    value = N * K / mixed["seconds"]
    fp64_value = N * K / fp64["seconds"]
    errH = float(np.max(np.abs(mixed["H"] - fp64["H"])) / np.max(np.abs(fp64["H"])))
    errC = float(np.max(np.abs(mixed["C"] - fp64["C"])) / np.max(np.abs(fp64["C"])))
    errT = float(np.max(np.abs(mixed["T"] - fp64["T"])) / np.max(np.abs(fp64["T"])))
    errphi = float(np.max(np.abs(mixed["phi"] - fp64["phi"])))
    region_diff = int(np.count_nonzero(mixed["region"] != fp64["region"]))
    b = algorithmic_bytes(spec.q, vel_b=VEL_BYTES["mixed"])
    b64 = algorithmic_bytes(spec.q, vel_b=VEL_BYTES["fp64"])
    prof = mixed["profile"]
    launch_b, launch_b64, kernel = launch_bytes(spec.q, mixed["layout"]["one_launch_steps"])
    achieved = N * launch_b / (prof["ms_stage"] * 1e-3) / 1e9
    achieved64 = N * launch_b64 / (prof["ms_stage"] * 1e-3) / 1e9
    out = {
        "metric": "H/C cell-updates/sec, mixed vs FP64 at <=1e-6 rel err; serve p50/p99 ms",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": mixed["seconds"] / K * 1e3,
        "ms_per_step_windows": {"min": mixed["windows_ms"][0],
                                "median": mixed["windows_ms"][len(mixed["windows_ms"]) // 2],
                                "max": mixed["windows_ms"][-1], "windows": len(mixed["windows_ms"]),
                                "in_order": mixed["windows_ms_in_order"]},
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
        "error_vs_fp64": {"H": errH, "C": errC, "T": errT, "phi_abs": errphi,
                          "region_cells_differing": region_diff, "steps": args.warmup + K},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kernel,
                     "bytes_per_cell_launch": launch_b,
                     "bytes_per_cell_step": b["step"],
                     "ms_per_launch": prof["ms_stage"], "ms_phase_kernel": prof["ms_phase_stage"],
                     "rows_per_item": mixed["rows"],
                     "with_fp64_samples": {"bytes_per_cell_launch": launch_b64,
                                           "achieved": achieved64,
                                           "frac": achieved64 / HBM_PEAK_GBS}},
        "conservation": {"total_H": mixed["report"]["total_H"], "total_C": mixed["report"]["total_C"]},
    }
    # HBM bytes per launch of this kernel on this workload, from two rocprofv3
    # counter runs (FETCH_SIZE x 2, WRITE_SIZE x 1 by a plain 1 GiB copy's
    # counts); the line's own run collects no counters
    kib = MEASURED_STEP_LAUNCH_KIB.get(mixed["rows"])
    if (spec.nx, spec.ny, spec.q) == (2048, 4096, 5) and mixed["layout"]["one_launch_steps"] and kib:
        out["roofline"]["traffic"] = (kib[0] * 2 + kib[1]) * 1024
        out["roofline"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of step32_kernel, "
                                             f"copy-calibrated, {mixed['rows']} rows per item "
                                             "(profiles/step_launch_2048x4096.md)")
    out["roofline"]["practical_copy_GBs"] = copy_gbs
    out["roofline"]["frac_of_practical_copy"] = achieved / copy_gbs if copy_gbs else None
    if args.limiter_steps > 0:
        out["limiter_fp64"] = bench_limiter(args, spec, H, Cc, vx, vy, phase)
        out["limiter_mixed"] = bench_limiter(args, spec, H, Cc, vx, vy, phase, M.PREC_MIXED)
        out["diffusion"] = bench_diffusion(args, spec, H, Cc, vx, vy, phase)
    if args.coupled_steps > 0:
        out["coupled_fp64"] = bench_coupled(args)
    if args.full_steps > 0:
        out["full_chip"] = bench_full_chip(args)
    if not (args.no_batch and args.no_serve):
        meshes = batch_inputs()
        if not args.no_batch:
            out["batch"] = bench_batch(args, meshes)
        if not args.no_serve:
            out["serve"] = bench_serve(args, meshes, out.get("batch"))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
