#!/usr/bin/env python3
"""Benchmark of the MI355X Lagrangian ellipsoid step (interp + advance +
collide) -- ellipsoid updates/s, mixed form vs the plain FP64 form.

Default workload (C2): 256^3 field from random Fourier modes (k0 = 4,
u_rms = 1, seeded), 1,000,000 spheroids with aspect ratio log-uniform in
[0.2, 5] at volume fraction 1e-3, Stokes number log-uniform in [0.1, 10],
lag3, dt at CFL 0.1, collisions on with the reference cell rule.  One step
= one pass of gather + RK2/drag/torque/quaternion/walls/wrap + cell list +
contact solve + ordered collision resolution over all particles;
updates/s = Np * steps / time of the timed steps (inputs resident in HBM,
setup excluded).

value        = updates/s of the mixed form (FP32 field, gather and drag;
               FP64 positions, quaternions, contact solve and impulses)
fp64         = the same for the plain FP64 form
roofline     = the fused advance kernel (the HBM-bound gather): algorithmic
               bytes per launch / its average duration (HIP events on the
               library's stream over timed eager steps) against 8 TB/s;
               traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE +
               WRITE_SIZE scaled by a copy calibration, when a summary for
               this configuration is committed
cpu_baseline = the FP64 CPU oracle (a restatement of the reference path,
               OpenMP over particles) on the host cores, one full step of
               the same workload (C2), or the compiled reference Fortran on
               one core (C1)

Other configurations (--config): C1 (64^3, 10k prolate spheroids), C2L4
(C2 with lag4), C3 (512^3 field, 8M tri-axial), C4 (64 ensembles on one
field, autotuned launch shapes), C5 (2M tri-axial at 5 % volume fraction,
network-seeded contact solve trained on the GPU).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C2]
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


def run_gpu(wl, precision, steps, warmup, G, groups=None, theta=None, **kw):
    """Timed steps of the captured hipGraph (8 steps, the next advance of the
    particles outside the contact solves beside those solves), then a
    separate eager run with HIP-event stage timers (per-kernel figures)."""
    def make(**extra):
        c = G.Context.from_workload(wl, precision=precision, **kw, **extra)
        if groups is not None:
            c.set_groups(groups)
        if theta is not None:
            c.seed_net_set(theta)
        return c
    ctx = make(graph=True)
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    t1 = time.perf_counter()
    st = ctx.stats()
    state = ctx.get_particles()
    st["collision_counts"] = ctx.collision_counts()
    ctx.close()
    tctx = make(timing=True)
    tctx.step(warmup)
    tctx.reset_timers()
    nt = min(steps, 10)
    nfev = nlm = ncoll = 0
    for _ in range(nt):
        tctx.step(1)
        s = tctx.stats()
        nfev += s["lm_nfev"]
        nlm += s["nlm"]
        ncoll += s["ncollision"]
    tctx.sync()
    tst = tctx.stats()
    # a caller that keeps the particles in host memory and hands them over
    # every step: upload, one step, download (pageable numpy arrays)
    st["host_handover_ms_per_step"] = None
    if groups is None:   # ensembles: the groups are set once, not handed over
        cur = tctx.get_particles()
        tctx.step(1)
        tctx.sync()
        npc = 3
        p0 = time.perf_counter()
        for _ in range(npc):
            tctx.set_particles(cur, wl.species_id)
            tctx.step(1)
            cur = tctx.get_particles()
        p1 = time.perf_counter()
        st["host_handover_ms_per_step"] = 1e3 * (p1 - p0) / npc
    tctx.close()
    for k in ("t_advance_ms", "t_sort_ms", "t_contact_ms", "t_greedy_ms", "ntimed"):
        st[k] = tst[k]
    st["per_step"] = {"lm_nfev": nfev / nt, "nlm": nlm / nt, "ncollision": ncoll / nt}
    return (t1 - t0), st, state


def cpu_baseline_port(wl, threads):
    import oracle as O
    stp = O.Stepper(wl, nthreads=threads, nclosemax=max(64 * wl.np_, 1 << 20))
    t0 = time.perf_counter()
    st, err, nc, ncoll = stp.step(wl.state.copy())
    dt = time.perf_counter() - t0
    return {"value": round(wl.np_ / dt, 1), "unit": "updates/s", "cores": threads, "kind": "port",
            "sample": f"one full step (interp+advance+cell list+LM contact+collisions) of the "
                      f"same {wl.np_} ellipsoids on the {wl.n}^3 field, FP64 CPU oracle with "
                      f"{threads} OpenMP threads, {dt:.1f}s (err {err})"}


def cpu_baseline_reference(wl, max_seconds=20.0):
    """The reference's own ADV_PARTICLE_ELLP + COLLISION (compiled Fortran,
    one rank, one core) on the same single-class workload; as many steps as
    fit in max_seconds."""
    import oracle as O
    from gazpart_amd import workload as W
    if not O.ref_available():
        return None
    sp = wl.species[0]
    R = O.Ref()
    R.setup(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, wl.ecw, wl.muw, wl.scheme,
            [[sp[W.SP_EMAJ]]], [sp[W.SP_APR]], [[sp[W.SP_RHOP]]])
    R.set_field(wl.field)
    R.set_particles(1, wl.state, np.ones(wl.np_, dtype=np.int32))
    R.collision_init()
    t0 = time.perf_counter()
    s = 0
    while s < wl.steps and time.perf_counter() - t0 < max_seconds:
        s += 1
        R.adv(s)
        R.collision(1, s, 0.0, 200 * wl.np_)
    dt = time.perf_counter() - t0
    return {"value": round(wl.np_ * s / dt, 1), "unit": "updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{s} steps of the same {wl.np_} ellipsoids on the {wl.n}^3 field through "
                      f"the reference's own Fortran ADV_PARTICLE_ELLP + COLLISION (amdflang, one "
                      f"rank, one core), {dt:.1f}s"}


def roofline(wl, res, scheme, cfg):
    roof = {}
    for name, fb, srw in (("fp64", 8, 208), ("mixed", 4, 160)):
        if name not in res:
            continue
        st = res[name]["stats"]
        t_adv = st["t_advance_ms"] / max(st["ntimed"], 1) * 1e-3
        byt = advance_bytes(wl.np_, wl.n, fb, 64 if scheme == 3 else 125, srw)
        ach = byt / t_adv / 1e9
        roof[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                      "kernel": "k_advance", "bytes_per_launch": int(byt),
                      "avg_launch_ms": round(t_adv * 1e3, 4)}
    tr_file = os.path.join(ROOT, "profiles", f"traffic_k_advance_{cfg}.json")
    if os.path.exists(tr_file):
        tr = json.load(open(tr_file))
        for name in roof:
            if name in tr:
                roof[name]["traffic"] = tr[name]
    return roof


def stage(st):
    n = max(st["ntimed"], 1)
    return {k: round(st[f"t_{k}_ms"] / n, 4) for k in ("advance", "sort", "contact", "greedy")}


QUANT = (("x", slice(0, 3)), ("u", slice(3, 6)), ("q", slice(6, 10)), ("omega", slice(10, 13)))


def relerr(a, b, ca, cb):
    """Mixed (a) against FP64 (b) after the same steps: per-quantity
    max |a - b| / max |b| over the particles that took part in no collision
    in either run (the advance alone: gather, drag, torque, quaternion,
    walls), and separately over those that did.  The reference's contact
    solve stops on ftol = 1.5e-8 with a non-zero residual, so the common
    normal it returns is fixed only to ~1e-4 and moves with 1e-10 changes of
    its inputs; collided particles carry that, whatever the precision.
    Also: particles whose number of collisions differs between the runs."""
    free = (ca == 0) & (cb == 0)
    out = {}
    for q, sl in QUANT:
        scale = np.abs(b[sl]).max()
        out[q] = float(np.abs(a[sl][:, free] - b[sl][:, free]).max() / scale)
    hit = ~free
    out["collided_particles"] = int(hit.sum())
    if hit.any():
        out["collided_max"] = {q: float(np.abs(a[sl][:, hit] - b[sl][:, hit]).max() /
                                        np.abs(b[sl]).max()) for q, sl in QUANT}
    out["collision_count_mismatch"] = int((ca != cb).sum())
    return out


def describe(cfg, wl, scheme):
    lag = f"lag{scheme}"
    if cfg == "C1":
        return (f"C1: {wl.n}^3 random-Fourier field (k0=4, u_rms=1), {wl.np_} prolate spheroids "
                f"APR 2, a=0.2 DX, St 1, {lag}, CFL 0.1, collisions on (reference cell rule)")
    if cfg in ("C2", "C2L4"):
        return (f"C2: {wl.n}^3 random-Fourier field (k0=4, u_rms=1), {wl.np_} spheroids APR "
                f"log-uniform [0.2,5], phi=1e-3, St log-uniform [0.1,10], {lag}, CFL 0.1, "
                f"collisions on (reference cell rule)")
    if cfg == "C3":
        return (f"C3: {wl.n}^3 random-Fourier field (k0=8) stored FP32 in the mixed form, "
                f"{wl.np_} tri-axial ellipsoids (a/b, a/c uniform [1,4]), phi=1e-3 (assumed), "
                f"St 1, {lag}")
    if cfg == "C4":
        return (f"C4: {len(wl.meta.get('ranges', []))} independent ensembles ({wl.np_} "
                f"ellipsoids, counts log-uniform [1e3,5e5]) on one shared {wl.n}^3 field, "
                f"per-ensemble APR [0.2,5] and St [0.1,10], one captured graph, autotuned "
                f"launch shapes")
    return (f"C5: {wl.n}^3 field, {wl.np_} tri-axial ellipsoids (a/b, a/c uniform [1,4]) at "
            f"phi=5%, St 1, {lag}, contact solve seeded by the on-GPU-trained network")


def make_workload(cfg, args, W):
    scheme = args.scheme or (4 if cfg == "C2L4" else 3)
    if cfg == "C1":
        return W.c1(np_=args.np or 10_000, n=args.n or 64, steps=100, scheme=scheme), scheme
    if cfg in ("C2", "C2L4"):
        return W.build_workload("C2", args.n or 256, args.np or 1_000_000, k0=4.0, seed=2024,
                                scheme=scheme, steps=200, apr=(0.2, 5.0), phi=1e-3,
                                st=(0.1, 10.0)), scheme
    if cfg == "C3":
        return W.c3(n=args.n or 512, np_=args.np or 8_000_000, scheme=scheme), scheme
    if cfg == "C4":
        return W.c4(n=args.n or 256), scheme
    return W.c5(n=args.n or 256, np_=args.np or 2_000_000, scheme=scheme), scheme


def stats_bench(wl, G, W, steps, cpu=True):
    """Host-timed calls (device work + result copy) of the statistics
    routines on the workload's state after 3 steps: with the workload's
    own species rows and with the particles regrouped into 2 classes x 3
    sizes; and the graph-mode step with collision statistics off and on
    every step (the 6-row grouping, so that bins stay few)."""
    import numpy as np

    def per_call(f, n=5):
        f()
        t = time.perf_counter()
        for _ in range(n):
            f()
        return round((time.perf_counter() - t) / n * 1e3, 4)

    out = {"metric": "statistics_ms_per_call", "config": {"workload": wl.name, "np": wl.np_}}
    ctx = G.Context.from_workload(wl, precision=G.PREC_MIXED, keep_fluid=True)
    ctx.step(3)
    ctx.sync()
    out["rows_own"] = ctx.nspec
    out["particle_stats_own_rows"] = per_call(lambda: ctx.particle_stats())
    out["partfluid_stats_own_rows"] = per_call(lambda: ctx.partfluid_stats())
    st = ctx.get_particles()
    ctx.close()
    rows6 = W.species_spheroid(np.full(6, float(wl.species[:, W.SP_EMAJ].mean())),
                               np.repeat([2.0, 5.0], 3), np.full(6, 1000.0))
    sid6 = (np.arange(wl.np_) % 6).astype(np.int32)

    def make6(**kw):
        c = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, scheme=wl.scheme, dpmax=wl.dpmax,
                      precision=G.PREC_MIXED, **kw)
        c.bind_field(wl.field)
        c.set_species(rows6)
        c.set_particles(st, sid6)
        return c

    c6 = make6(keep_fluid=True)
    c6.adv_particle_ellp(1)
    out["particle_stats_2x3"] = per_call(lambda: c6.particle_stats(npoly=3))
    out["partfluid_stats_2x3"] = per_call(lambda: c6.partfluid_stats(npoly=3))
    if cpu:
        # the numpy restatements on the same state, one host core
        from oracle import stats_ref as SR
        st6, fl6 = c6.get_particles(), c6.get_fluid_at_particles()
        t = time.perf_counter()
        SR.particle_stats(st6, sid6, rows6, npoly=3)
        t1 = time.perf_counter()
        SR.partfluid_stats(st6, fl6, sid6, rows6, wl.visc, wl.rhof, npoly=3)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {
            "particle_stats_2x3_ms": round((t1 - t) * 1e3, 1),
            "partfluid_stats_2x3_ms": round((t2 - t1) * 1e3, 1), "cores": 1, "kind": "port",
            "sample": f"one call of each numpy restatement on the same {wl.np_} particles "
                      f"(2 classes x 3 sizes), one core"}
    c6.close()
    for name, fout2 in (("step_ms_collision_stats_off", 0), ("step_ms_collision_stats_every_step", 1)):
        cx = make6(graph=True, overlap=1)
        if fout2:
            cx.set_collision_stats(nclass=2, npoly=3, fout2=1)
        cx.step(9)
        cx.sync()
        t = time.perf_counter()
        cx.step(steps)
        cx.sync()
        out[name] = round((time.perf_counter() - t) / steps * 1e3, 4)
        if fout2:
            out["contacts_reached_last_step"] = int(round(
                cx.collision_stats()["last"][..., 0].sum()))
        cx.close()
    return out


def dns_bench(wl, G, W, steps, warmup, cpu=True):
    """The spectral DNS supplier on the workload's grid (n^3, L, VISC),
    started from the workload's own velocity: ADV_FLUID steps alone, then
    the coupled step with the fluid evolving (fluid step, device-to-device
    bind of velocity and vorticity, one particle step in graph mode).
    Host-timed around synchronised calls; per-kernel times come from a
    kernel trace of the same command."""
    import numpy as np
    from gazpart_amd.dns import Dns

    n = wl.n
    u0 = np.stack([wl.field[c, 2:n + 2, 2:n + 2, :] for c in range(3)])
    d = Dns(n, wl.L, wl.visc, wl.dtime)
    d.set_velocity(u0)
    dt = d.cfl(wl.dtime)
    d.step(warmup)
    t = time.perf_counter()
    d.step(steps)
    fluid_ms = (time.perf_counter() - t) / steps * 1e3
    info = d.info()
    dr = Dns(n, wl.L, wl.visc, dt, reuse_spectral=True)
    dr.set_velocity(u0)
    dr.step(warmup)
    t = time.perf_counter()
    dr.step(steps)
    reuse_ms = (time.perf_counter() - t) / steps * 1e3
    dr.close()
    # algorithmic HBM bytes of one ADV_FLUID: 7 batched transforms (read +
    # write of three fields each), the velocity copy, and the element-wise
    # kernels' reads and writes (spectral fields 16 B x n^2 (n/2+1) each,
    # physical 8 B x n^3)
    rb = 3 * 8.0 * n ** 3
    cb = 3 * 16.0 * n * n * (n // 2 + 1)
    bytes_step = (7 * (rb + cb) + 2 * rb + 3 * cb + 2 * (3 * rb) + 6 * cb + 4 * cb)
    ctx = G.Context.from_workload(wl, precision=G.PREC_MIXED, graph=True, overlap=1)
    d.bind(ctx)
    ctx.step(9)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        d.step(1)
        d.bind(ctx)
        ctx.step(1)
    ctx.sync()
    coupled_ms = (time.perf_counter() - t) / steps * 1e3
    cpu_line = None
    if cpu:
        # one step of the numpy restatement on the host (FFTs on one core)
        from oracle import dns_ref as DR
        ref = DR.DnsRef(n, wl.L, wl.visc, dt)
        ref.set_velocity(u0)
        t = time.perf_counter()
        ref.step()
        sec = time.perf_counter() - t
        cpu_line = {"value": round(sec * 1e3, 1), "unit": "ms/step", "cores": 1, "kind": "port",
                    "sample": f"one ADV_FLUID step of the same {n}^3 field through the numpy "
                              f"restatement (numpy FFT, one core), {sec:.1f}s"}
    return {"metric": "dns_ms_per_step", "value": round(fluid_ms, 4), "unit": "ms",
            "cpu_baseline": cpu_line,
            "higher_is_better": False, "steps": steps, "warmup": warmup, "dtype": "f64",
            "data": "synthetic", "config": {"workload": f"{wl.name} grid {n}^3, ADV_FLUID",
                                             "dtime": dt, "visc": wl.visc},
            "reuse_spectral_ms_per_step": round(reuse_ms, 4),
            "coupled_ms_per_step": round(coupled_ms, 4),
            "coupled_particles_per_s": round(wl.np_ / (coupled_ms * 1e-3), 1),
            "eps_flu": info["eps"], "umax": [float(x) for x in info["umax"]],
            "algorithmic_bytes_per_step": bytes_step,
            "effective_GBps": round(bytes_step / (fluid_ms * 1e-3) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C2L4", "C3", "C4", "C5"])
    ap.add_argument("--np", type=int, default=0)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--scheme", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--brick", type=int, default=0,
                    help="advance visiting order: bricks of BRICK^3 collision cells "
                         "(0: the library default)")
    ap.add_argument("--lm-waves", type=int, default=0,
                    help="wavefronts of the contact-solve launch (0: the library "
                         "default, from the solve count: one per SIMD while "
                         "lanes sit idle, up to four per SIMD)")
    ap.add_argument("--pairs", type=int, default=10_000_000,
                    help="C5: solver-labelled pairs the seed network is trained on")
    ap.add_argument("--epochs", type=int, default=3, help="C5: training epochs")
    ap.add_argument("--stats", action="store_true",
                    help="time the statistics routines on the workload (particle, "
                         "particle-fluid and collision statistics) instead of the step")
    ap.add_argument("--dns", action="store_true",
                    help="time the spectral DNS field supplier on the workload's grid: "
                         "ADV_FLUID alone and coupled (fluid step, device bind, particle "
                         "step) instead of the frozen-field step")
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

    cfg = args.config
    t0 = time.perf_counter()
    wl, scheme = make_workload(cfg, args, W)
    log(f"{cfg} workload built in {time.perf_counter() - t0:.1f}s: n={wl.n} np={wl.np_} "
        f"dt={wl.dtime:.3e} dpmax={wl.dpmax:.3e}")
    groups = wl.meta.get("group") if cfg == "C4" else None
    if args.stats:
        print(json.dumps(stats_bench(wl, G, W, args.steps, not args.no_cpu)))
        return
    if args.dns:
        print(json.dumps(dns_bench(wl, G, W, args.steps, args.warmup, not args.no_cpu)))
        return

    kw = {"brick": args.brick, "lm_waves": args.lm_waves}
    extra = {}
    theta = None
    if cfg == "C4":
        from gazpart_amd.autotune import tune
        tr = tune(wl, G.PREC_MIXED, groups=groups, log=log)
        kw = {"brick": tr["best"]["brick"], "lm_waves": tr["best"]["lm_waves"]}
        extra["autotune"] = tr
    if cfg == "C5":
        from oracle import mlp_ref as M
        nctx = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime)
        nctx.seed_net_set(M.init_params(7))
        fit = nctx.seed_net_fit(args.pairs, 2024, epochs=args.epochs, batch=4096, lr=2e-3)
        ev = nctx.seed_net_eval(200_000, 77)
        theta = nctx.seed_net_get()
        nctx.close()
        log(f"seed network: fit {fit}\n  eval {ev}")
        extra["seed_net"] = {"fit": fit, "eval": ev}

    res = {}
    runs = [("fp64", G.PREC_FP64, {}), ("mixed", G.PREC_MIXED, {})]
    if cfg == "C5":
        runs = [("fp64", G.PREC_FP64, {}), ("mixed_centre_seed", G.PREC_MIXED, {}),
                ("mixed", G.PREC_MIXED, {"lm_seed": 1, "theta": theta})]
    for name, prec, ek in runs:
        dt, st, state = run_gpu(wl, prec, args.steps, args.warmup, G, groups, **kw, **ek)
        ms = 1e3 * dt / args.steps
        res[name] = dict(seconds=dt, ms_per_step=ms, stats=st, state=state,
                         updates_per_s=wl.np_ * args.steps / dt)
        shown = {k: v for k, v in st.items() if k != "collision_counts"}
        log(f"{name}: {ms:.3f} ms/step, {wl.np_ * args.steps / dt:.3e} updates/s, stats {shown}")

    roof = roofline(wl, res, scheme, cfg)
    cpu = None
    if not args.no_cpu:
        if cfg == "C1":
            cpu = cpu_baseline_reference(wl)
        if cfg in ("C1", "C2", "C2L4") and cpu is None:
            threads = args.cpu_threads or min(16, os.cpu_count() or 1)
            cpu = cpu_baseline_port(wl, threads)
        if cpu:
            log(f"cpu baseline: {cpu}")

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
        "config": {"workload": describe(cfg, wl, scheme), "np": wl.np_, "grid": wl.n,
                   "interp": f"lag{scheme}", "launch": kw,
                   "solve_waves": m["stats"]["lm_waves"]},
        "fp64": {"value": round(f["updates_per_s"], 1), "ms_per_step": round(f["ms_per_step"], 4),
                 "dtype": "f64", "roofline": roof["fp64"], "stage_ms": stage(f["stats"]),
                 "per_step": f["stats"]["per_step"]},
        "mixed_vs_fp64_relerr": relerr(m["state"], f["state"], m["stats"]["collision_counts"],
                                       f["stats"]["collision_counts"]),
        "speedup_mixed_over_fp64": round(m["updates_per_s"] / f["updates_per_s"], 3),
        "stage_ms": stage(m["stats"]),
        "per_step": m["stats"]["per_step"],
        # not the headline: particles uploaded from and downloaded to host
        # memory around every step (PCIe-inclusive), one step per call
        "host_handover": None if m["stats"]["host_handover_ms_per_step"] is None else {
            "ms_per_step": round(m["stats"]["host_handover_ms_per_step"], 3),
            "updates_per_s": round(wl.np_ * 1e3 / m["stats"]["host_handover_ms_per_step"], 1)},
        "collisions_per_step": m["stats"]["ncollision"],
        "candidate_pairs": m["stats"]["nclose"],
        "roofline": roof["mixed"],
        "cpu_baseline": cpu,
    }
    if cfg == "C5":
        c = res["mixed_centre_seed"]
        line["mixed_centre_seed"] = {"value": round(c["updates_per_s"], 1),
                                     "ms_per_step": round(c["ms_per_step"], 4),
                                     "stage_ms": stage(c["stats"]),
                                     "per_step": c["stats"]["per_step"]}
        line["seeded_vs_centre_seed_relerr"] = relerr(m["state"], c["state"],
                                                      m["stats"]["collision_counts"],
                                                      c["stats"]["collision_counts"])
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
