"""Autotuner for the launch shapes that leave every result unchanged.

Four knobs of :class:`gazpart_amd.Context` decide how the work is laid out
on the GPU, not what is computed (``tile`` changes the rounding of the
advance, nothing else):

* ``brick`` -- the advance visits particles grouped in bricks of brick^3
  collision cells (compact sets of stencils per workgroup) or, for values
  <= 0, in the reference's cell order.  The collision cell itself stays at
  the reference rule (collision_initiation.f90:50-102) because it fixes
  the candidate-pair order that the one-collision-per-particle pass follows.
* ``tile`` -- the advance visits particles in tiles of tile^3 field cells,
  one workgroup per tile that stages the field nodes its stencils reach in
  LDS and writes the fluid at its particles, then the tile mode's advance
  kernel reads them (overrides brick).  It reads less of HBM and pays off
  where a tile holds a wavefront's worth of particles or more; its results
  equal the direct path's to rounding.
* ``lm_waves`` -- the width of the contact-solve launch (one pair per lane,
  pairs dealt lane-major over the wavefronts); 0 means one wavefront per
  SIMD.
* ``graph`` -- how steps without host synchronisation are issued: 1
  replays a captured hipGraph of 8 pipelined steps, 2 launches the same
  kernels directly on the two streams.

:func:`tune` times a few steps (after a warm-up that leaves a captured
graph in place) of the caller's own workload from the same initial state
for every combination and returns the fastest, with the whole table.  The
visiting orders tried are the bricks, then the tiles.  The GPU tests check
that the knobs do not change results.
"""
from __future__ import annotations

import time

from . import Context

BRICKS = (-1, 2, 4, 8)
TILES = (6, 8)
LM_WAVES = (0, 2048, 4096)
GRAPHS = (1, 2)


def tune(wl, precision, *, steps: int = 8, warmup: int = 9, bricks=BRICKS, tiles=TILES,
         lm_waves=LM_WAVES, graphs=GRAPHS, groups=None, log=None, passes: int = 2,
         **ctx_kw) -> dict:
    """Returns {"best": {...}, "table": [...]}; each entry has brick, tile,
    lm_waves, graph and ms_per_step (steps after `warmup` steps: the first
    sizes the lists, the next 8 are one captured graph in mode 1), the
    fastest of `passes` timings taken in turn over all the candidates, so
    that one slow sample does not decide."""
    orders = [(b, 0) for b in bricks] + [(0, t) for t in tiles]
    cands = [(g, b, t, w) for g in graphs for b, t in orders for w in lm_waves]
    best_ms = {}
    for _ in range(max(1, passes)):
        for g, b, t, w in cands:
            ctx = Context.from_workload(wl, precision=precision, graph=g, brick=b, tile=t,
                                        lm_waves=w, **ctx_kw)
            if groups is not None:
                ctx.set_groups(groups)
            ctx.step(warmup)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(steps)
            ctx.sync()
            dt = time.perf_counter() - t0
            ctx.close()
            ms = round(1e3 * dt / steps, 4)
            key = (g, b, t, w)
            best_ms[key] = min(ms, best_ms.get(key, ms))
            if log:
                log(f"autotune: {{'brick': {b}, 'tile': {t}, 'lm_waves': {w}, 'graph': {g}, "
                    f"'ms_per_step': {ms}}}")
    table = [{"brick": b, "tile": t, "lm_waves": w, "graph": g, "ms_per_step": best_ms[(g, b, t, w)]}
             for g, b, t, w in cands]
    best = min(table, key=lambda r: r["ms_per_step"])
    return {"best": best, "table": table}
