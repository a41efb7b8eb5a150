"""GPU steps next to the CPU oracle, for the parity tests of the bench
configurations (test helper, not collected).

Each step is checked twice:

* from identical inputs: the oracle's advance and collision pass start from
  the GPU's own state of that moment (cell lists bit-exact; contact flags
  and resolved collisions equal except for listed pairs; per-quantity
  error of the collision pass);
* as trajectories: the GPU and the oracle each carry their own state from
  the common start; a particle whose collisions differ between the two
  runs (and every particle that later collides with such a particle) is
  history-divergent and listed, the others are compared per quantity.

Decisions that differ are listed with the quantities each run decided them
on (gazpart_amd.contacts): depth along the normal, both margins, NVREL and
the pair's closeness to the threshold.
"""
from __future__ import annotations

import numpy as np

import gazpart_amd as G
import oracle as O
from gazpart_amd import contacts as CT

QUANT = {"x": slice(0, 3), "u": slice(3, 6), "q": slice(6, 10), "omega": slice(10, 13)}


def relerr(a, b, sl, cols=None):
    if cols is not None:
        a, b = a[:, cols], b[:, cols]
    if a.shape[1] == 0:
        return 0.0
    return float(np.abs(a[sl] - b[sl]).max() / max(np.abs(b[sl]).max(), 1e-300))


def oracle_quantities(wl, state, home, near):
    """The oracle's own solve outputs for a few pairs, and their decision
    quantities."""
    n = len(home)
    e1, e2, cn = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    flag = np.zeros(n, dtype=np.int32)
    _, _, b1, b2 = CT.pair_geometry(state, wl.L, home, near)
    for k in range(n):
        i, j = home[k] - 1, near[k] - 1
        f, a, b, c, _, _ = O.contact(wl.species[wl.species_id[i]], wl.species[wl.species_id[j]],
                                     b1[k], state[6:10, i], b2[k], state[6:10, j])
        flag[k], e1[k], e2[k], cn[k] = f, a, b, c
    qd = CT.decision_quantities(state, wl.species, wl.species_id, wl.L, home, near, e1, e2, cn)
    return flag, qd


def gpu_quantities(ctx, wl, state, home, near):
    flag, e1, e2, cn, nfev = ctx.ellipsoid_contact_detection(home, near)
    qd = CT.decision_quantities(state, wl.species, wl.species_id, wl.L, home, near, e1, e2, cn)
    return flag.astype(np.int32), qd, nfev


def _pairs(h, n, sel=None):
    if sel is not None:
        h, n = h[sel], n[sel]
    return set(zip(h.tolist(), n.tolist()))


class Lockstep:
    """Eager GPU steps of one workload beside the oracle."""

    def __init__(self, wl, precision, nthreads=8, nclosemax=None, **ctxkw):
        self.wl = wl
        self.ctx = G.Context.from_workload(wl, precision=precision, **ctxkw)
        self.p = O.params(wl)
        self.g = O.collision_grid(self.p, wl.dpmax)
        self.con, self.ncon = O.connectivity(self.g)
        self.nth = nthreads
        self.nmax = nclosemax or max(64 * wl.np_, 1 << 16)
        self.gpu = wl.state.copy()
        self.ora = wl.state.copy()
        self.tainted = np.zeros(wl.np_, bool)
        self.ever = np.zeros(wl.np_, bool)   # collided at some step in either run
        self.step_no = 0
        # records
        self.flag_diffs = []      # identical inputs: flags over all candidates
        self.resolved_diffs = []  # identical inputs: resolved collisions
        self.roots = []           # trajectories: first divergent collisions
        self.step_err = []        # identical inputs: per-quantity error after the step
        self.adv_err = []         # identical inputs: per-quantity error after the advance
        self.ncand = []
        self.ncoll = []
        self.nfev = []            # evaluations of the solves over every candidate
        self.solve = []           # the step's own solves: count, evaluations, fallbacks

    def _detect(self, st):
        h, n, nc, err = O.detect(self.g, self.con, self.ncon, st[0], st[1], st[2], self.nmax)
        assert err == 0, f"oracle detection error {err}"
        return h, n

    def step(self):
        wl, ctx = self.wl, self.ctx
        self.step_no += 1
        s = self.step_no
        pre = self.gpu
        # --- GPU step, eager, with the state between its halves
        ctx.adv_particle_ellp(s)
        mid = ctx.get_particles()
        h, nr = ctx.collision_detection_connectivity()
        # identical inputs: the advance
        mo, _, err = O.adv(self.p, wl.field, wl.species, pre, wl.species_id, self.nth)
        assert err == 0
        self.adv_err.append({q: relerr(mid, mo, sl) for q, sl in QUANT.items()})
        # identical inputs: the cell lists
        ho, no = self._detect(mid)
        np.testing.assert_array_equal(h, ho, err_msg=f"step {s}: HOME differs")
        np.testing.assert_array_equal(nr, no, err_msg=f"step {s}: NEAR differs")
        self.ncand.append(len(h))
        # identical inputs: flags of every candidate pair
        fg, qg, nfev = gpu_quantities(ctx, wl, mid, h, nr)
        self.nfev.append(int(nfev.sum()))
        so, fo, co, nco, err = O.interaction(self.p, wl.species, mid, wl.species_id, h, nr,
                                             self.nth)
        assert err == 0
        dif = np.nonzero(fg != fo)[0]
        if len(dif):
            fo_d, qo_d = oracle_quantities(wl, mid, h[dif], nr[dif])
            sub = {k: v[dif] for k, v in qg.items()}
            self.flag_diffs += CT.list_differences(h[dif], nr[dif], fg[dif], sub, fo_d, qo_d,
                                                   step=s)
        # the GPU's collision pass
        ctx.collision(s)
        after = ctx.get_particles()
        st = ctx.stats()
        self.solve.append({k: int(st[k]) for k in ("nlm", "lm_nfev", "seed_fallbacks")})
        gh, gn, gres = ctx.contacts()
        gset = _pairs(gh, gn, gres)
        oset = _pairs(h, nr, co.astype(bool))
        self.ncoll.append(len(gset))
        # identical inputs: resolved collisions, and the particles they touch
        bad = np.zeros(wl.np_, bool)
        rdiff = sorted(gset ^ oset)
        if rdiff:
            rh = np.array([a for a, _ in rdiff], np.int32)
            rn = np.array([b for _, b in rdiff], np.int32)
            fo_d, qo_d = oracle_quantities(wl, mid, rh, rn)
            # the GPU's solve outputs for these pairs on the same state
            sel = np.array([np.nonzero((h == a) & (nr == b))[0][0] for a, b in rdiff])
            sub = {k: v[sel] for k, v in qg.items()}
            close = np.minimum(CT.closeness(sub), CT.closeness(qo_d))
            for k, (a, b) in enumerate(rdiff):
                self.resolved_diffs.append(dict(
                    step=s, home=a, near=b, gpu=(a, b) in gset, oracle=(a, b) in oset,
                    flag=(int(fg[sel[k]]), int(fo_d[k])),
                    nvrel=(float(sub["nvrel"][k]), float(qo_d["nvrel"][k])),
                    closeness=float(close[k])))
                bad[a - 1] = bad[b - 1] = True
            # partners of the differing pairs in either run's resolved set
            for (a, b) in gset | oset:
                if bad[a - 1] or bad[b - 1]:
                    bad[a - 1] = bad[b - 1] = True
        # particles the collision pass left alone in both runs are untouched
        # (equal to the advance's output); the resolved ones carry the
        # contact solve's own rounding-level differences
        hit = np.zeros(wl.np_, bool)
        for (a, b) in gset | oset:
            hit[a - 1] = hit[b - 1] = True
        free, coll = ~hit, hit & ~bad
        self.step_err.append({
            "free": {q: relerr(after, so, sl, free) for q, sl in QUANT.items()},
            "collided": {q: relerr(after, so, sl, coll) for q, sl in QUANT.items()},
            "ncollided": int(coll.sum()), "nexcluded": int(bad.sum())})
        self.ever |= hit
        self.gpu = after
        # --- the oracle's own trajectory
        om, _, err = O.adv(self.p, wl.field, wl.species, self.ora, wl.species_id, self.nth)
        assert err == 0
        oh, on = self._detect(om)
        ot, _, oc, _, err = O.interaction(self.p, wl.species, om, wl.species_id, oh, on, self.nth)
        assert err == 0
        oset_t = _pairs(oh, on, oc.astype(bool))
        tdiff = gset ^ oset_t
        for (a, b) in sorted(tdiff):
            if not (self.tainted[a - 1] or self.tainted[b - 1]):
                self.roots.append(self._root(s, a, b, mid, om, (a, b) in gset,
                                             (a, b) in oset_t))
        newly = np.zeros(wl.np_, bool)
        for (a, b) in tdiff:
            newly[a - 1] = newly[b - 1] = True
        for (a, b) in gset | oset_t:
            if self.tainted[a - 1] or self.tainted[b - 1]:
                newly[a - 1] = newly[b - 1] = True
        self.tainted |= newly
        for (a, b) in oset_t:
            self.ever[a - 1] = self.ever[b - 1] = True
        self.ora = ot

    def _root(self, s, a, b, gmid, omid, in_g, in_o):
        """quantities of a first divergent collision, each run on its own state"""
        ha, nb = np.array([a], np.int32), np.array([b], np.int32)
        # the GPU state of this step's advance is gone by now: the GPU side
        # is recomputed by the oracle on the GPU's state (same arithmetic
        # up to rounding); the oracle side on its own state
        fg, qg = oracle_quantities(self.wl, gmid, ha, nb)
        fo, qo = oracle_quantities(self.wl, omid, ha, nb)
        return dict(step=s, home=a, near=b, gpu=in_g, oracle=in_o,
                    flag=(int(fg[0]), int(fo[0])),
                    depth=(float(qg["depth"][0]), float(qo["depth"][0])),
                    nvrel=(float(qg["nvrel"][0]), float(qo["nvrel"][0])),
                    closeness=float(min(CT.closeness(qg)[0], CT.closeness(qo)[0])))

    def trajectory_error(self):
        """per-quantity error of the particles that never collided, and of
        those that collided with the same history in both runs"""
        free = ~self.ever
        coll = self.ever & ~self.tainted
        return {"free": {q: relerr(self.gpu, self.ora, sl, free) for q, sl in QUANT.items()},
                "collided": {q: relerr(self.gpu, self.ora, sl, coll) for q, sl in QUANT.items()},
                "nfree": int(free.sum()), "ncollided": int(coll.sum()),
                "ntainted": int(self.tainted.sum())}

    def close(self):
        self.ctx.close()

