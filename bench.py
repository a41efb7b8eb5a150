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
roofline     = the advance (the HBM-bound gather; with field tiles the
               tiled gather and the tile mode's advance kernel together):
               algorithmic bytes per advance / its average duration (HIP
               events on the library's stream over timed eager steps)
               against 8 TB/s; traffic = HBM bytes per advance from
               rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy
               calibration, the median over the dispatches of a profiled
               run (steady state: sorted particles), when a summary for
               this configuration, kernel and visiting order is committed
cpu_baseline = the FP64 CPU oracle (a restatement of the reference path,
               OpenMP over particles) on the host cores, one full step of
               the same workload (C2), or the compiled reference Fortran on
               one core (C1)

Other configurations (--config): C1 (64^3, 10k prolate spheroids), C2L4
(C2 with lag4), C3 (512^3 field, 8M tri-axial), C4 (64 ensembles on one
field, autotuned launch shapes), C5 (2M tri-axial at 5 % volume fraction,
network-seeded contact solve trained on the GPU).

Launch shape (results do not depend on it): C4 takes the fastest of the
autotuner's table (visiting orders, solve widths, launch modes); the other
configurations tune the launch mode and the visiting order (cell order or
bricks of 4 cells), the FP64 form's order apart from the mixed form's.

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


def run_gpu(wl, precision, steps, warmup, G, groups=None, theta=None, graph=1, **kw):
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
    ctx = make(graph=graph)
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
    # the call pattern of the Fortran drop-in in resident mode: per cycle one
    # call, the previous cycle's collision pass with this cycle's advance
    # (gp_collision_advance), on a context as GPU_INIT makes it
    if groups is None:
        dctx = make(overlap=1)
        dctx.step(warmup)
        dctx.adv_particle_ellp(warmup + 1)
        dctx.collision_advance(warmup + 1)   # sizes the lists
        dctx.sync()
        d0 = time.perf_counter()
        for k in range(steps):
            dctx.collision_advance(warmup + 2 + k)
        dctx.sync()
        st["dropin_ms_per_step"] = 1e3 * (time.perf_counter() - d0) / steps
        dctx.close()
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
    for k in ("t_advance_ms", "t_sort_ms", "t_contact_ms", "t_greedy_ms", "t_solve_ms", "ntimed"):
        st[k] = tst[k]
    st["per_step"] = {"lm_nfev": nfev / nt, "nlm": nlm / nt, "ncollision": ncoll / nt}
    return (t1 - t0), st, state


def host_threads():
    """The host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where set (the host's share for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(cap)) if cap.isdigit() and int(cap) > 0 else n


def cpu_baseline_port(wl, threads):
    import oracle as O
    stp = O.Stepper(wl, nthreads=threads, nclosemax=max(64 * wl.np_, 1 << 20))
    t0 = time.perf_counter()
    st, err, nc, ncoll = stp.step(wl.state.copy())
    dt = time.perf_counter() - t0
    return {"value": round(wl.np_ / dt, 1), "unit": "updates/s", "cores": threads, "kind": "port",
            "sample": f"one full step (interp+advance+cell list+LM contact+collisions) of the "
                      f"same {wl.np_} ellipsoids on the {wl.n}^3 field, FP64 CPU oracle with "
                      f"{threads} OpenMP threads (the host cores of this process), {dt:.1f}s "
                      f"(err {err})"}


def _big_stack(fn):
    """Run fn in a thread with a 1 GiB stack: the reference's Fortran passes
    strided PART sections through stack temporaries (np doubles each)."""
    import threading
    out = {}
    old = threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=lambda: out.setdefault("r", fn()))
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    return out.get("r")


def cpu_baseline_reference(wl, max_seconds=20.0, species=None, state=None, what=None):
    """The reference's own ADV_PARTICLE_ELLP + COLLISION (compiled Fortran,
    one rank, one core) on a single-class workload; as many steps as fit in
    max_seconds.  species = (EMAJ, APR, RHOP) of the class (default: the
    workload's first row), state = the particles (default: the workload's)."""
    import oracle as O
    from gazpart_amd import workload as W
    if not O.ref_available():
        return None
    if species is None:
        sp = wl.species[0]
        species = (sp[W.SP_EMAJ], sp[W.SP_APR], sp[W.SP_RHOP])
    st = wl.state if state is None else state
    np_ = st.shape[1]

    def run():
        R = O.Ref()
        R.setup(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, wl.ecw, wl.muw, wl.scheme,
                [[species[0]]], [species[1]], [[species[2]]])
        R.set_field(wl.field)
        R.set_particles(1, st, np.ones(np_, dtype=np.int32))
        R.collision_init()
        t0 = time.perf_counter()
        s = 0
        while s < wl.steps and time.perf_counter() - t0 < max_seconds:
            s += 1
            R.adv(s)
            R.collision(1, s, 0.0, 200 * np_)
        return s, time.perf_counter() - t0

    s, dt = _big_stack(run)
    return {"value": round(np_ * s / dt, 1), "unit": "updates/s", "cores": 1,
            "kind": "reference",
            "sample": (what or f"{s} steps of the same {np_} ellipsoids on the {wl.n}^3 field") +
                      f" through the reference's own Fortran ADV_PARTICLE_ELLP + COLLISION "
                      f"(amdflang, one rank, one core), {s} step(s) in {dt:.1f}s"}


def cpu_baseline_reference_c2(wl):
    """The nearest case the reference can run beside C2: it collides
    particles of one class only, with one aspect ratio per class
    (collision.f90:76-171), so C2's per-particle aspect ratios and Stokes
    numbers are replaced by one class of spheroids at the geometric mean of
    both ranges (APR 1, St 1) and C2's volume-equivalent radius; the
    particles of C2's state in the slab x < L/16 (about 62k), one step."""
    from gazpart_amd import workload as W
    req = (1e-3 * wl.L ** 3 * 3.0 / (4.0 * math.pi * wl.np_)) ** (1.0 / 3.0)
    rhop = W.rhop_for_stokes(1.0, 2.0 * req, wl.meta["tau_flow"], wl.rhof, wl.visc)
    st = np.ascontiguousarray(wl.state[:, wl.state[0] < wl.L / 16.0])
    return cpu_baseline_reference(
        wl, max_seconds=5.0, species=(req, 1.0, rhop), state=st,
        what=f"one class (APR 1, St 1, r_eq of C2) on the C2 {wl.n}^3 field, the "
             f"{st.shape[1]} particles of C2's state with x < L/16")


def advance_kernel(launch):
    """The kernel(s) of one advance in a launch shape: the fused gather +
    advance, or with field tiles the tiled gather and the mode's advance."""
    return "k_gather_tiled+k_advance_fl" if launch.get("tile", 0) > 0 else "k_advance"


def visiting_brick(launch):
    """The brick size of a launch shape's visiting order (0: cell order)."""
    return 0 if launch.get("tile", 0) > 0 else max(launch.get("brick", 0), 0)


def roofline(wl, res, scheme, cfg, launch, launch_fp64):
    roof = {}
    for name, fb, srw in (("fp64", 8, 208), ("mixed", 4, 160)):
        if name not in res:
            continue
        kernel = advance_kernel(launch_fp64 if name == "fp64" else launch)
        st = res[name]["stats"]
        t_adv = st["t_advance_ms"] / max(st["ntimed"], 1) * 1e-3
        byt = advance_bytes(wl.np_, wl.n, fb, 64 if scheme == 3 else 125, srw)
        ach = byt / t_adv / 1e9
        roof[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                      "kernel": kernel, "bytes_per_launch": int(byt),
                      "avg_launch_ms": round(t_adv * 1e3, 4)}
    tr_file = os.path.join(ROOT, "profiles", f"traffic_k_advance_{cfg}.json")
    if os.path.exists(tr_file):
        tr = json.load(open(tr_file))
        for name in roof:
            # a committed measurement of this configuration's kernel(s) in
            # the same visiting order
            lw = launch_fp64 if name == "fp64" else launch
            if (name in tr and tr.get("kernel", "k_advance") == roof[name]["kernel"]
                    and tr.get("brick", 0) == visiting_brick(lw)):
                roof[name]["traffic"] = tr[name]
                roof[name]["traffic_source"] = tr.get("source", os.path.relpath(tr_file, ROOT))
    return roof


# Peak evaluation rate of the contact solve, had every lane of every SIMD a
# pair and the SIMDs issued without stalls: a wavefront of 64 solves takes
# VALU_PER_EVAL VALU instructions per function evaluation (SQ counters on
# the solve kernel: SQ_INSTS_VALU / evaluations per wavefront), FP64
# instructions issuing at one per 4 cycles on a SIMD (wave64 over 16 lanes)
VALU_PER_EVAL = 1567.0
SIMDS = 1024
CLOCK_HZ = 2.4e9


def contact_roofline(st, nsteps):
    """The contact-solve kernel against its issue-bound peak: evaluations/s
    achieved (the step's evaluations over the kernel's HIP-event time), the
    peak 64 lanes x SIMDs x clock / (4 x VALU_PER_EVAL), and the fraction of
    launched lanes that carried a pair."""
    n = max(st["ntimed"], 1)
    t = st["t_solve_ms"] / n * 1e-3
    if t <= 0:
        return None
    ps = st["per_step"]
    ach = ps["lm_nfev"] / t
    peak = 64.0 * SIMDS * CLOCK_HZ / (4.0 * VALU_PER_EVAL)
    return {"bound": "valu", "kernel": "k_contact", "achieved": round(ach, 1),
            "peak": round(peak, 1), "unit": "evaluations/s", "frac": round(ach / peak, 5),
            "avg_launch_ms": round(t * 1e3, 4), "evaluations_per_launch": ps["lm_nfev"],
            "solves_per_launch": ps["nlm"],
            "lanes_busy": round(ps["nlm"] / (64.0 * max(st["lm_waves"], 1)), 4),
            "longest_chain_note": "one lane per pair: the launch lasts as long as its "
                                  "slowest solve (up to 300 evaluations)"}


def stage(st):
    n = max(st["ntimed"], 1)
    return {k: round(st[f"t_{k}_ms"] / n, 4) for k in ("advance", "sort", "contact", "greedy")}


MARGIN = 1e-3     # closeness to a threshold within which a decision may differ
ACC_TOL = 1e-5    # the metric's relative-error bar


def paired(wl, G, groups, kw_a, kw_b, steps, perturb=0.0, theta_b=None):
    """Two eager runs of the workload side by side (gazpart_amd.compare):
    per-quantity relative errors of never-collided particles and of
    particles with the same collisions in both runs, the history-divergent
    count and every first divergent collision with its closeness to the
    threshold.  perturb: relative perturbation of run b's positions."""
    from gazpart_amd.compare import Paired
    a = G.Context.from_workload(wl, **kw_a)
    b = G.Context.from_workload(wl, **kw_b)
    if theta_b is not None:
        b.seed_net_set(theta_b)
    if perturb:
        st = wl.state.copy()
        g = np.random.Generator(np.random.Philox(99))
        st[0:3] *= 1.0 + perturb * g.uniform(-1.0, 1.0, st[0:3].shape)
        b.set_particles(st, wl.species_id)
    if groups is not None:
        a.set_groups(groups)
        b.set_groups(groups)
    P = Paired(a, b, wl.species, wl.species_id, wl.L)
    for _ in range(steps):
        P.step()
    rep = P.report()
    a.close()
    b.close()
    rep["first_divergent_collisions"] = rep["first_divergent_collisions"][:50]
    return rep


def accuracy_gate(mix, floor):
    """The metric's gate: never-collided particles within ACC_TOL per
    quantity; collided particles with the same history within twice the
    spread the FP64 form itself shows from inputs 1e-12 apart (or ACC_TOL,
    whichever is larger: the contact solve stops on ftol with a non-zero
    residual, so its normal moves with its inputs); every first divergent
    collision of a known kind (a decision within MARGIN of its threshold,
    the solve ending at another minimum, or the one-collision order), none
    of kind "other"."""
    nc = max(mix["never_collided"].values())
    co = max(mix["collided_same_history"].values())
    fl = max(floor["collided_same_history"].values())
    bound = max(ACC_TOL, 2.0 * fl)
    return {"tol": ACC_TOL, "never_collided_max": nc, "collided_same_history_max": co,
            "fp64_floor_collided_max": fl, "collided_bound": bound,
            "history_divergent": mix["n_history_divergent"],
            "fp64_floor_history_divergent": floor["n_history_divergent"],
            "first_divergent_kinds": mix["kinds"], "margin": MARGIN,
            "pass": bool(nc <= ACC_TOL and co <= bound and mix["kinds"]["other"] == 0)}


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
    ap.add_argument("--warmup", type=int, default=9)
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C2L4", "C3", "C4", "C5"])
    ap.add_argument("--np", type=int, default=0)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--scheme", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--acc-steps", type=int, default=20,
                    help="steps of the side-by-side accuracy runs (mixed vs FP64, FP64 vs FP64 "
                         "from inputs 1e-12 apart; 0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--brick", type=int, default=0,
                    help="advance visiting order: bricks of BRICK^3 collision cells "
                         "(0: cell order, or bricks of 4 where the autotuner finds them "
                         "faster, for each form separately)")
    ap.add_argument("--tile", type=int, default=0,
                    help="advance visiting order: field tiles of TILE^3 cells staged in LDS "
                         "(0: off; overrides --brick)")
    ap.add_argument("--lm-waves", type=int, default=0,
                    help="wavefronts of the contact-solve launch (0: the library "
                         "default, from the solve count: one per SIMD while "
                         "lanes sit idle, up to four per SIMD)")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1, 2],
                    help="steps without host synchronisation: 1 captured hipGraph replays, 2 "
                         "direct launches, 0 (default) the faster of the two on this workload")
    ap.add_argument("--no-tune", action="store_true",
                    help="no autotuner: the launch shape given by --brick, --tile, --lm-waves "
                         "and --graph (--graph 0: eager steps)")
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
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
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

    kw = {"brick": args.brick, "tile": args.tile, "lm_waves": args.lm_waves,
          "graph": args.graph}
    kw_fp64 = None   # the FP64 form's launch shape when tuned apart
    extra = {}
    theta = None
    from gazpart_amd.autotune import tune
    if cfg == "C4" and not args.no_tune:
        tr = tune(wl, G.PREC_MIXED, groups=groups, log=log)
        kw = {"brick": tr["best"]["brick"], "tile": tr["best"]["tile"],
              "lm_waves": tr["best"]["lm_waves"], "graph": tr["best"]["graph"]}
        extra["autotune"] = tr
    elif args.graph == 0 and not args.no_tune:
        # launch mode (captured graph replays or direct launches) and, unless
        # given, visiting order (cell order or bricks of 4 cells), with the
        # given solve width
        bricks = ((args.brick, 4) if args.brick <= 0 else (args.brick,)) if args.tile <= 0 else ()
        tr = tune(wl, G.PREC_MIXED, bricks=bricks,
                  tiles=(args.tile,) if args.tile > 0 else (), lm_waves=(args.lm_waves,),
                  steps=16, log=log)
        kw["graph"] = tr["best"]["graph"]
        kw["brick"] = tr["best"]["brick"]
        extra["autotune"] = tr
        if len(bricks) > 1:
            # the FP64 form's advance moves twice the bytes and can set its
            # step: its own visiting order, in the mixed form's launch mode
            trf = tune(wl, G.PREC_FP64, bricks=bricks, tiles=(), lm_waves=(args.lm_waves,),
                       graphs=(kw["graph"],), steps=8, log=log)
            kw_fp64 = dict(kw, brick=trf["best"]["brick"])
            extra["autotune_fp64"] = trf
    if cfg == "C5":
        nctx = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime)
        fit = G.train_seed_net(nctx, npairs=args.pairs, seed=2024, epochs=args.epochs)
        ev = nctx.seed_net_eval(200_000, 77)
        nctx.seed_net_gate(300, math.pi)
        ev_ungated = nctx.seed_net_eval(200_000, 77)
        theta = nctx.seed_net_get()
        nctx.close()
        log(f"seed network: fit {fit}\n  eval {ev}\n  ungated {ev_ungated}")
        extra["seed_net"] = {"fit": fit, "eval_gated": ev, "eval_ungated": ev_ungated}

    res = {}
    runs = [("fp64", G.PREC_FP64, {}), ("mixed", G.PREC_MIXED, {})]
    if cfg == "C5":
        runs = [("fp64", G.PREC_FP64, {}), ("mixed_centre_seed", G.PREC_MIXED, {}),
                ("mixed", G.PREC_MIXED, {"lm_seed": 1, "theta": theta})]
    for name, prec, ek in runs:
        k = kw_fp64 if (prec == G.PREC_FP64 and kw_fp64) else kw
        dt, st, state = run_gpu(wl, prec, args.steps, args.warmup, G, groups, **k, **ek)
        ms = 1e3 * dt / args.steps
        res[name] = dict(seconds=dt, ms_per_step=ms, stats=st, state=state,
                         updates_per_s=wl.np_ * args.steps / dt)
        shown = {k: v for k, v in st.items() if k != "collision_counts"}
        log(f"{name}: {ms:.3f} ms/step, {wl.np_ * args.steps / dt:.3e} updates/s, stats {shown}")

    roof = roofline(wl, res, scheme, cfg, kw, kw_fp64 or kw)
    cpu, cpu_ref = None, None
    if not args.no_cpu:
        if cfg == "C1":
            cpu = cpu_baseline_reference(wl)
        if cfg in ("C1", "C2", "C2L4") and cpu is None:
            cpu = cpu_baseline_port(wl, args.cpu_threads or host_threads())
        if cfg in ("C2", "C2L4"):
            cpu_ref = cpu_baseline_reference_c2(wl)
        if cpu:
            log(f"cpu baseline: {cpu}")
        if cpu_ref:
            log(f"reference Fortran beside it: {cpu_ref}")

    # accuracy: eager runs side by side (not timed)
    acc = {}
    if args.acc_steps > 0:
        t0 = time.perf_counter()
        acc["mixed_vs_fp64"] = paired(wl, G, groups, {"precision": G.PREC_FP64},
                                      {"precision": G.PREC_MIXED}, args.acc_steps)
        acc["fp64_vs_fp64_perturbed_1e-12"] = paired(
            wl, G, groups, {"precision": G.PREC_FP64}, {"precision": G.PREC_FP64},
            args.acc_steps, perturb=1e-12)
        if cfg == "C5":
            acc["seeded_vs_centre_seed"] = paired(
                wl, G, groups, {"precision": G.PREC_MIXED},
                {"precision": G.PREC_MIXED, "lm_seed": 1}, args.acc_steps, theta_b=theta)
        log(f"accuracy runs: {time.perf_counter() - t0:.1f}s")

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
                 "dtype": "f64", "launch": kw_fp64 or kw, "roofline": roof["fp64"],
                 "stage_ms": stage(f["stats"]),
                 "per_step": f["stats"]["per_step"]},
        "accuracy_gate": (accuracy_gate(acc["mixed_vs_fp64"], acc["fp64_vs_fp64_perturbed_1e-12"])
                          if acc else None),
        "accuracy": acc,
        "speedup_mixed_over_fp64": round(m["updates_per_s"] / f["updates_per_s"], 3),
        "stage_ms": stage(m["stats"]),
        "per_step": m["stats"]["per_step"],
        # the last timed step's contact solves and, of them, the pairs solved
        # ahead beside its sort and filters (pipelined steps, overlap = 1)
        "solved_ahead": {"pairs": m["stats"]["nahead"], "of_solves": m["stats"]["nlm"]},
        # not the headline: particles uploaded from and downloaded to host
        # memory around every step (PCIe-inclusive), one step per call
        "host_handover": None if m["stats"]["host_handover_ms_per_step"] is None else {
            "ms_per_step": round(m["stats"]["host_handover_ms_per_step"], 3),
            "updates_per_s": round(wl.np_ * 1e3 / m["stats"]["host_handover_ms_per_step"], 1)},
        "collisions_per_step": m["stats"]["ncollision"],
        "candidate_pairs": m["stats"]["nclose"],
        "roofline": roof["mixed"],
        "roofline_contact": contact_roofline(m["stats"], args.steps),
        "cpu_baseline": cpu,
        "reference_fortran_beside": cpu_ref,
        # the Fortran drop-in's resident mode: one gp_step(1) per cycle
        # (GPU_COLLISION after a deferred GPU_ADV_PARTICLE_ELLP)
        "dropin_resident": {"ms_per_step": round(m["stats"]["dropin_ms_per_step"], 4),
                            "updates_per_s": round(wl.np_ * 1e3 / m["stats"]["dropin_ms_per_step"], 1)}
        if m["stats"].get("dropin_ms_per_step") else None,
    }
    if cfg == "C5":
        c = res["mixed_centre_seed"]
        line["mixed_centre_seed"] = {"value": round(c["updates_per_s"], 1),
                                     "ms_per_step": round(c["ms_per_step"], 4),
                                     "stage_ms": stage(c["stats"]),
                                     "per_step": c["stats"]["per_step"]}

    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
