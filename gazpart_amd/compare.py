"""Two runs of the same workload stepped side by side on the GPU, with the
collisions whose outcome differs listed (host side).

Used to state what the mixed form changes against the FP64 form, what a
network-seeded contact solve changes against the centre-line one, and how
far two FP64 runs from inputs 1e-12 apart drift (the contact solve's own
floor).  Each run steps eagerly (ADV_PARTICLE_ELLP then COLLISION); after
each step the resolved collisions of both runs (gp_get_contacts) are
compared.  A particle becomes history-divergent when its collision differs
between the runs (a pair resolved in one run only, or resolved in both with
contact normals more than MIN_ANGLE apart: the two solves ended at
different minima, and the impulses differ accordingly), or when it later
collides with such a particle in either run.  The first divergent collision
of every such chain (both particles still undivergent) is listed with the
decision quantities of each run on its own state (gazpart_amd.contacts):
depth, both margins, NVREL, the pair's closeness to the threshold, the
angle between the two runs' contact normals and the relative difference of
the pair's inputs, and is put in one of three classes:

  threshold  the contact decision (flag and NVREL < ZERO) differs and, in
             one run or the other, a test sits within MARGIN of its
             switching point;
  minimum    the two solves ended at different minima of their residual
             (normals more than MIN_ANGLE apart), with the contact decision
             differing or not (a pair resolved in both runs with different
             normals): the reference's solve is discontinuous in its inputs
             there, and the FP64 form started from inputs 1e-12 apart shows
             such flips too;
  order      the contact decision is the same in both runs; the pair was
             resolved in one run only because the one-collision pass had
             given one of its particles to another pair first.

Anything else (decisions that differ at the same minimum, far from every
threshold) is classed "other", which a correct pair of runs never shows.
Per-quantity errors are reported separately for particles that never
collided, and for those that collided with the same history in both runs.
"""
from __future__ import annotations

import numpy as np

from . import contacts as CT

QUANT = (("x", slice(0, 3)), ("u", slice(3, 6)), ("q", slice(6, 10)), ("omega", slice(10, 13)))
MARGIN = 1e-3       # closeness within which a threshold decision may differ
MIN_ANGLE = 1e-3    # radians between normals: different minima of the solve


def _relerr(a, b, sl, cols):
    if not cols.any():
        return 0.0
    scale = max(float(np.abs(b[sl]).max()), 1e-300)
    return float(np.abs(a[sl][:, cols] - b[sl][:, cols]).max() / scale)


def _resolved(ctx):
    h, n, r = ctx.contacts()
    return set(zip(h[r].tolist(), n[r].tolist()))


class Paired:
    """Runs a and b of one workload (species table, species ids, box L),
    already set up in contexts ctx_a and ctx_b."""

    def __init__(self, ctx_a, ctx_b, species, species_id, L):
        self.a, self.b = ctx_a, ctx_b
        self.species, self.species_id, self.L = species, species_id, L
        np_ = len(species_id)
        self.ever = np.zeros(np_, bool)
        self.tainted = np.zeros(np_, bool)
        self.roots = []
        self.step_no = 0
        self.ncoll = []

    def _quantities(self, ctx, mid, home, near):
        flag, e1, e2, cn, _ = ctx.ellipsoid_contact_detection(home, near)
        qd = CT.decision_quantities(mid, self.species, self.species_id, self.L, home, near,
                                    e1, e2, cn)
        qd["cn"] = cn
        return flag.astype(np.int32), qd

    def step(self):
        self.step_no += 1
        s = self.step_no
        mids, sets = [], []
        for ctx in (self.a, self.b):
            ctx.adv_particle_ellp(s)
            mids.append(ctx.get_particles())
            ctx.collision(s)
            sets.append(_resolved(ctx))
        ga, gb = sets
        self.ncoll.append((len(ga), len(gb)))
        diff = ga ^ gb
        # pairs resolved in both runs, by undivergent particles, whose solves
        # ended at different minima (their impulses differ)
        both = sorted(p for p in ga & gb if not (self.tainted[p[0] - 1] or self.tainted[p[1] - 1]))
        if both:
            h = np.array([p for p, _ in both], np.int32)
            n = np.array([q for _, q in both], np.int32)
            _, e1a, e2a, cna, _ = self.a.ellipsoid_contact_detection(h, n)
            _, e1b, e2b, cnb, _ = self.b.ellipsoid_contact_detection(h, n)
            ang = np.arccos(np.clip(np.abs((cna * cnb).sum(axis=1)), 0.0, 1.0))
            diff |= {both[k] for k in np.nonzero(ang > MIN_ANGLE)[0]}
        new_roots = [(p, q) for (p, q) in sorted(diff)
                     if not (self.tainted[p - 1] or self.tainted[q - 1])]
        if new_roots:
            h = np.array([p for p, _ in new_roots], np.int32)
            n = np.array([q for _, q in new_roots], np.int32)
            # flags and normals depend on positions and orientations only,
            # which the collision pass leaves as they were; NVREL takes the
            # velocities before the pass
            fa, qa = self._quantities(self.a, mids[0], h, n)
            fb, qb = self._quantities(self.b, mids[1], h, n)
            close = np.minimum(CT.closeness(qa), CT.closeness(qb))
            ca, cb = CT.contact_of(fa, qa), CT.contact_of(fb, qb)
            dot = np.clip(np.abs((qa["cn"] * qb["cn"]).sum(axis=1)), 0.0, 1.0)
            angle = np.arccos(dot)
            ia, ib = h - 1, n - 1
            din = np.zeros(len(h))
            for sl in (slice(0, 3), slice(3, 6), slice(6, 10), slice(10, 13)):
                for idx in (ia, ib):
                    da = np.abs(mids[0][sl][:, idx] - mids[1][sl][:, idx]).max(axis=0)
                    sc = np.maximum(np.abs(mids[1][sl][:, idx]).max(axis=0), 1e-300)
                    din = np.maximum(din, da / sc)
            for k, (p, q) in enumerate(new_roots):
                if ca[k] == cb[k] and angle[k] > MIN_ANGLE:
                    kind = "minimum"
                elif ca[k] == cb[k]:
                    kind = "order"
                elif close[k] <= MARGIN:
                    kind = "threshold"
                elif angle[k] > MIN_ANGLE:
                    kind = "minimum"
                else:
                    kind = "other"
                self.roots.append(dict(
                    step=s, home=p, near=q, kind=kind, resolved=((p, q) in ga, (p, q) in gb),
                    flag=(int(fa[k]), int(fb[k])),
                    depth=(float(qa["depth"][k]), float(qb["depth"][k])),
                    fm1=(float(qa["fm1"][k]), float(qb["fm1"][k])),
                    fm2=(float(qa["fm2"][k]), float(qb["fm2"][k])),
                    nvrel=(float(qa["nvrel"][k]), float(qb["nvrel"][k])),
                    closeness=float(close[k]), normal_angle=float(angle[k]),
                    input_diff=float(din[k])))
        newly = np.zeros_like(self.tainted)
        for (p, q) in diff:
            newly[p - 1] = newly[q - 1] = True
        for (p, q) in ga | gb:
            self.ever[p - 1] = self.ever[q - 1] = True
            if self.tainted[p - 1] or self.tainted[q - 1]:
                newly[p - 1] = newly[q - 1] = True
        self.tainted |= newly

    def report(self):
        sa, sb = self.a.get_particles(), self.b.get_particles()
        free = ~self.ever
        same = self.ever & ~self.tainted
        return {
            "steps": self.step_no,
            "never_collided": {q: _relerr(sa, sb, sl, free) for q, sl in QUANT},
            "collided_same_history": {q: _relerr(sa, sb, sl, same) for q, sl in QUANT},
            "n_never_collided": int(free.sum()), "n_collided_same_history": int(same.sum()),
            "n_history_divergent": int(self.tainted.sum()),
            "first_divergent_collisions": self.roots,
            "kinds": {k: sum(r["kind"] == k for r in self.roots)
                      for k in ("threshold", "minimum", "order", "other")},
            "max_closeness_threshold_kind": max(
                (r["closeness"] for r in self.roots if r["kind"] == "threshold"), default=0.0),
        }
