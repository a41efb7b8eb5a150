#!/usr/bin/env python3
"""Benchmark of the MI355X Lagrangian ellipsoid step (interp + advance +
collide) -- ellipsoid updates/s, mixed form vs the plain FP64 form.

Workload (default): 256^3 field from random Fourier modes (k0 = 4, u_rms = 1,
seeded), 1,000,000 spheroids with aspect ratio log-uniform in [0.2, 5] at
volume fraction 1e-3, Stokes number log-uniform in [0.1, 10], lag3, dt at
CFL 0.1, collisions on with the reference cell rule.  One step = one pass of
gather + RK2/drag/torque/quaternion/walls/wrap + cell list + contact solve +
ordered collision resolution over all particles; updates/s = Np * steps /
time of the timed steps (inputs resident in HBM, setup excluded).

value        = updates/s of the mixed form (FP32 field, gather and drag;
               FP64 positions, quaternions, contact solve and impulses)
fp64         = the same for the plain FP64 form
roofline     = the fused advance kernel (the HBM-bound gather): algorithmic
               bytes per launch / its average duration (HIP events on the
               library's stream over the timed steps) against 8 TB/s
cpu_baseline = the FP64 CPU oracle (a restatement of the reference path)
               on the host cores, one full step of the same workload

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--np N] [--n N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def advance_bytes(np_, n, field_bytes, taps, s_rw, s_par=16):
    """Algorithmic HBM bytes of one fused advance launch: particle state read
    + written and per-particle shape/response data, plus each field node
    touched at least once (6 components), the fraction touched being
    1 - exp(-Np T / N^3)."""
    touched = 1.0 - math.exp(-np_ * taps / float(n) ** 3)
    return np_ * (s_rw + s_par) + 6.0 * n ** 3 * field_bytes * touched


def run_gpu(wl, precision, steps, warmup, G, brick=0):
    """Timed steps of the captured two-step hipGraph, then a separate eager
    run with HIP-event stage timers (for the per-kernel figures)."""
    ctx = G.Context.from_workload(wl, precision=precision, graph=True, brick=brick)
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    t1 = time.perf_counter()
    st = ctx.stats()
    state = ctx.get_particles()
    ctx.close()
    tctx = G.Context.from_workload(wl, precision=precision, timing=True, brick=brick)
    tctx.step(warmup)
    tctx.reset_timers()
    tctx.step(min(steps, 10))
    tctx.sync()
    tst = tctx.stats()
    tctx.close()
    for k in ("t_advance_ms", "t_sort_ms", "t_contact_ms", "t_greedy_ms", "ntimed"):
        st[k] = tst[k]
    return (t1 - t0), st, state


def cpu_baseline(wl, threads):
    import oracle as O
    stp = O.Stepper(wl, nthreads=threads, nclosemax=max(64 * wl.np_, 1 << 20))
    t0 = time.perf_counter()
    st, err, nc, ncoll = stp.step(wl.state.copy())
    dt = time.perf_counter() - t0
    return wl.np_ / dt, dt, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--np", type=int, default=1_000_000)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--brick", type=int, default=0,
                    help="advance visiting order: bricks of BRICK^3 collision cells "
                         "(0: the library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (torch), to scale the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profile")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import gazpart_amd as G
    from gazpart_amd import workload as W

    t0 = time.perf_counter()
    wl = W.build_workload("C2", args.n, args.np, k0=4.0, seed=2024, scheme=args.scheme,
                          steps=args.steps, apr=(0.2, 5.0), phi=1e-3, st=(0.1, 10.0))
    log(f"workload built in {time.perf_counter() - t0:.1f}s: n={wl.n} np={wl.np_} "
        f"dt={wl.dtime:.3e} dpmax={wl.dpmax:.3e}")

    res = {}
    for name, prec in (("fp64", G.PREC_FP64), ("mixed", G.PREC_MIXED)):
        dt, st, state = run_gpu(wl, prec, args.steps, args.warmup, G, args.brick)
        ms = 1e3 * dt / args.steps
        res[name] = dict(seconds=dt, ms_per_step=ms, stats=st, state=state,
                         updates_per_s=wl.np_ * args.steps / dt)
        log(f"{name}: {ms:.3f} ms/step, {wl.np_ * args.steps / dt:.3e} updates/s, stats {st}")

    # mixed vs FP64 after the same steps from the same state (warmup + timed)
    a, b = res["mixed"]["state"], res["fp64"]["state"]
    relerr = {}
    for q, sl in (("x", slice(0, 3)), ("u", slice(3, 6)), ("q", slice(6, 10)),
                  ("omega", slice(10, 13))):
        relerr[q] = float(np.abs(a[sl] - b[sl]).max() / np.abs(b[sl]).max())

    # roofline of the fused advance kernel
    roof = {}
    for name, fb, srw in (("fp64", 8, 208), ("mixed", 4, 160)):
        st = res[name]["stats"]
        t_adv = st["t_advance_ms"] / max(st["ntimed"], 1) * 1e-3
        byt = advance_bytes(wl.np_, wl.n, fb, 64 if args.scheme == 3 else 125, srw)
        ach = byt / t_adv / 1e9
        roof[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                      "traffic": None, "kernel": "k_advance",
                      "bytes_per_launch": int(byt), "avg_launch_ms": round(t_adv * 1e3, 4)}
    tr_file = os.path.join(ROOT, "profiles", "traffic_k_advance_C2.json")
    if os.path.exists(tr_file):
        tr = json.load(open(tr_file))
        for name in roof:
            if name in tr:
                roof[name]["traffic"] = tr[name]

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ups, secs, err = cpu_baseline(wl, threads)
        cpu = {"value": round(ups, 1), "unit": "updates/s", "cores": threads, "kind": "port",
               "sample": f"one full step (interp+advance+cell list+LM contact+collisions) of "
                         f"the same {wl.np_} ellipsoids on the {wl.n}^3 field, FP64 CPU "
                         f"oracle with {threads} OpenMP threads, {secs:.1f}s"}
        log(f"cpu baseline: {ups:.3e} updates/s on {threads} threads ({secs:.1f}s), err {err}")

    def stage(st):
        n = max(st["ntimed"], 1)
        return {k: round(st[f"t_{k}_ms"] / n, 4) for k in ("advance", "sort", "contact", "greedy")}

    m, f = res["mixed"], res["fp64"]
    line = {
        "metric": "ellipsoid updates/s (interp+adv+collide), mixed vs FP64 at rel err <=1e-5",
        "value": round(m["updates_per_s"], 1),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 gather/drag + f64 positions/quaternion/contact",
        "data": "synthetic",
        "config": {"workload": f"C2: {wl.n}^3 random-Fourier field (k0=4, u_rms=1), "
                               f"{wl.np_} spheroids APR log-uniform [0.2,5], phi=1e-3, "
                               f"St log-uniform [0.1,10], lag{args.scheme - 0}, CFL 0.1, "
                               f"collisions on (reference cell rule)",
                   "np": wl.np_, "grid": wl.n, "interp": f"lag{args.scheme}"},
        "fp64": {"value": round(f["updates_per_s"], 1), "ms_per_step": round(f["ms_per_step"], 4),
                 "dtype": "f64", "roofline": roof["fp64"], "stage_ms": stage(f["stats"])},
        "mixed_vs_fp64_relerr": relerr,
        "speedup_mixed_over_fp64": round(m["updates_per_s"] / f["updates_per_s"], 3),
        "stage_ms": stage(m["stats"]),
        "collisions_per_step": m["stats"]["ncollision"],
        "candidate_pairs": m["stats"]["nclose"],
        "roofline": roof["mixed"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
