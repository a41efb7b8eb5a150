"""gazpart_amd -- MI355X (gfx950) Lagrangian ellipsoid step of GAZPART.

The compute path is the HIP library ``libgazpart_gpu.so`` (built in-tree by
``gazpart_amd._build``); this module binds its C-ABI (include/gazpart_gpu.h)
with ctypes and mirrors the reference's subroutine interface for this path:

  ===============================  =====================================
  reference (source/)              here
  ===============================  =====================================
  INTERPH  interpH.f90:26          :meth:`Context.interph`
  ADV_PARTICLE_ELLP  adv_particle_new.f90:7
                                   :meth:`Context.adv_particle_ellp`
  COLLISION  collision.f90:7       :meth:`Context.collision`
  COLLISION_DETECTION_CONNECTIVITY collision_detection_connectivity.f90:7
                                   :meth:`Context.collision_detection_connectivity`
  ellipsoid_contact_detection  ellipsoid_contact_detection.f90:1
                                   :meth:`Context.ellipsoid_contact_detection`
  ===============================  =====================================

There is no CPU fallback: if the library is missing or no GPU is visible,
constructing a :class:`Context` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgazpart_gpu.so")

GP_OK = 0
ERRORS = {
    1: "GP_ERR_NAN_DRAG", 2: "GP_ERR_NAN_OMEGA", 3: "GP_ERR_KE_COLLISION",
    4: "GP_ERR_KE_WALL", 5: "GP_ERR_OUT_OF_GRID", 6: "GP_ERR_NCLOSE_OVERFLOW",
    7: "GP_ERR_INTERP_RANGE", 8: "GP_ERR_CELL_RULE", 10: "GP_ERR_ARG", 11: "GP_ERR_STATE", 12: "GP_ERR_HIP",
    13: "GP_ERR_NOMEM", 14: "GP_ERR_IO",
}
PREC_FP64, PREC_MIXED = 0, 1

# every symbol include/gazpart_gpu.h declares
EXPORTS = (
    "gp_create", "gp_destroy", "gp_last_error_string", "gp_sync", "gp_bind_field",
    "gp_set_species", "gp_set_particles", "gp_set_groups", "gp_get_particles", "gp_get_fluid_at_particles",
    "gp_interph", "gp_interph_field", "gp_adv_particle_ellp", "gp_collision_detection", "gp_contact_batch",
    "gp_collision", "gp_step", "gp_get_stats", "gp_reset_timers", "gp_collision_grid",
    "gp_get_collision_counts", "gp_seed_net_set", "gp_seed_net_get", "gp_seed_net_forward", "gp_seed_net_loss_grad",
    "gp_seed_net_train", "gp_seed_net_pairs", "gp_seed_net_fit", "gp_seed_net_eval",
    "gp_part_file_write", "gp_part_file_info", "gp_part_file_read", "gp_field_file_write",
    "gp_field_file_read", "gp_particle_stats", "gp_set_collision_stats", "gp_get_collision_stats",
    "gp_partfluid_stats", "gp_bind_field_device", "gp_bind_field_planar", "gp_set_dtime", "gp_dns_create", "gp_dns_destroy",
    "gp_dns_error_string", "gp_dns_set_velocity", "gp_dns_get_velocity", "gp_dns_get_vorticity",
    "gp_dns_step", "gp_dns_cfl", "gp_dns_info", "gp_dns_bind", "gp_get_contacts",
    "gp_seed_net_gate", "gp_collision_advance", "gp_interph_noalloc", "gp_particle_pdf",
    "gp_lagrangian_origin", "gp_lagrangian_corr",
)
NPDF = 100
NLGR = 16
INTERPH_NOALLOC_MAX = 262144
SEED_NET_NPARAM = 5252


def seed_net_init(seed=0):
    """Initial parameters of the contact-seed network (12 -> 64 -> 64 -> 4,
    the flat layout gp_seed_net_set takes: W1, b1, W2, b2, W3, b3
    row-major): weights uniform in +-sqrt(6 / fan_in) (He), biases zero,
    from a Philox stream of the given seed."""
    g = np.random.Generator(np.random.Philox(int(seed)))
    parts = []
    for fan_out, fan_in in ((64, 12), (64, 64), (4, 64)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(g.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    theta = np.concatenate(parts).astype(np.float32)
    assert theta.size == SEED_NET_NPARAM
    return theta


def train_seed_net(ctx, npairs=2_000_000, seed=2024, epochs=3, init_seed=7):
    """Initialise and train the seed network of a context on npairs
    solver-labelled pairs (gp_seed_net_fit); returns the fit report."""
    ctx.seed_net_set(seed_net_init(init_seed))
    return ctx.seed_net_fit(npairs, seed, epochs=epochs, batch=4096, lr=2e-3)


class GpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """gp_params (include/gazpart_gpu.h)."""
    _fields_ = [("n", C.c_int), ("interp_scheme", C.c_int), ("precision", C.c_int),
                ("renorm_quat", C.c_int), ("L", C.c_double), ("visc", C.c_double),
                ("rhof", C.c_double), ("dtime", C.c_double), ("ecw", C.c_double),
                ("muw", C.c_double), ("dpmax", C.c_double), ("prefilter", C.c_int),
                ("xcd_remap", C.c_int), ("keep_fluid", C.c_int), ("timing", C.c_int),
                ("graph", C.c_int), ("brick", C.c_int), ("lm_seed", C.c_int),
                ("lm_waves", C.c_int), ("overlap", C.c_int), ("cell_rule", C.c_int),
                ("tile", C.c_int), ("reserved", C.c_int * 1)]


class Stats(C.Structure):
    """gp_stats (include/gazpart_gpu.h)."""
    _fields_ = [("nclose", C.c_int64), ("nprefilter", C.c_int64), ("ncontact", C.c_int64),
                ("ncollision", C.c_int64), ("lm_nfev", C.c_int64), ("nwall", C.c_int64),
                ("greedy_rounds", C.c_int64), ("t_advance_ms", C.c_double),
                ("t_sort_ms", C.c_double), ("t_contact_ms", C.c_double),
                ("t_greedy_ms", C.c_double), ("ntimed", C.c_int64), ("nlm", C.c_int64),
                ("seed_fallbacks", C.c_int64), ("lm_waves", C.c_int64),
                ("t_solve_ms", C.c_double), ("nahead", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SeedReport(C.Structure):
    """gp_seed_report (include/gazpart_gpu.h)."""
    _fields_ = [("npairs", C.c_int64), ("loss", C.c_double), ("label_ms", C.c_double),
                ("train_ms", C.c_double), ("nfev_ref_mean", C.c_double),
                ("nfev_net_mean", C.c_double), ("nfev_ref_max", C.c_int64),
                ("nfev_net_max", C.c_int64), ("flags_ref", C.c_int64), ("flags_net", C.c_int64),
                ("flag_agree", C.c_int64), ("fallbacks", C.c_int64), ("no_seed", C.c_int64),
                ("cn_min_dot", C.c_double), ("reserved", C.c_double * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: build it with `python -m gazpart_amd._build`")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.c_void_p, C.c_void_p
    L.gp_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
    L.gp_destroy.argtypes = [vp]
    L.gp_last_error_string.argtypes = [vp]
    L.gp_last_error_string.restype = C.c_char_p
    L.gp_sync.argtypes = [vp]
    L.gp_bind_field.argtypes = [vp, dp]
    L.gp_bind_field_device.argtypes = [vp, vp]
    L.gp_set_dtime.argtypes = [vp, C.c_double]
    L.gp_set_species.argtypes = [vp, C.c_int, dp]
    L.gp_set_particles.argtypes = [vp, C.c_int, dp, ip]
    L.gp_set_groups.argtypes = [vp, C.c_int, ip]
    L.gp_get_particles.argtypes = [vp, C.c_int, dp]
    L.gp_get_fluid_at_particles.argtypes = [vp, C.c_int, dp]
    L.gp_interph.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, C.c_int, dp]
    L.gp_interph_field.argtypes = [vp, C.c_int, dp, C.c_int, dp, dp, dp, dp]
    L.gp_interph_noalloc.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, C.c_int, dp]
    L.gp_adv_particle_ellp.argtypes = [vp, C.c_int]
    L.gp_collision_detection.argtypes = [vp, C.c_int64, ip, ip, C.POINTER(C.c_int64)]
    L.gp_contact_batch.argtypes = [vp, C.c_int64, ip, ip, vp, dp, dp, dp, ip]
    L.gp_collision.argtypes = [vp, C.c_int, C.c_double]
    L.gp_step.argtypes = [vp, C.c_int]
    L.gp_collision_advance.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.gp_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.gp_collision_grid.argtypes = [vp, ip, dp]
    L.gp_reset_timers.argtypes = [vp]
    L.gp_get_collision_counts.argtypes = [vp, C.c_int, ip, C.c_int]
    L.gp_particle_stats.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.gp_partfluid_stats.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.gp_particle_pdf.argtypes = [vp, C.c_int, dp]
    L.gp_lagrangian_origin.argtypes = [vp, C.c_int, C.c_int, dp]
    L.gp_lagrangian_corr.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
    L.gp_set_collision_stats.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.gp_get_collision_stats.argtypes = [vp, dp, dp, C.POINTER(C.c_int64)]
    cp = C.c_char_p
    L.gp_part_file_write.argtypes = [cp, C.c_int64, C.c_int, ip, dp, dp, C.c_double, dp]
    L.gp_part_file_info.argtypes = [cp, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
    L.gp_part_file_read.argtypes = [cp, C.c_int64, C.c_int, ip, dp, dp, C.POINTER(C.c_double),
                                    dp]
    L.gp_field_file_write.argtypes = [cp, C.c_int, dp]
    L.gp_field_file_read.argtypes = [cp, C.c_int, dp]
    L.gp_seed_net_set.argtypes = [vp, vp, C.c_int]
    L.gp_seed_net_get.argtypes = [vp, vp, C.c_int]
    L.gp_seed_net_forward.argtypes = [vp, C.c_int64, vp, vp]
    L.gp_seed_net_loss_grad.argtypes = [vp, C.c_int64, vp, vp, vp, C.POINTER(C.c_double)]
    L.gp_seed_net_train.argtypes = [vp, C.c_int64, vp, vp, C.c_int, C.c_int, C.c_float,
                                    C.POINTER(C.c_double)]
    L.gp_seed_net_pairs.argtypes = [vp, C.c_int64, C.c_uint64, vp, vp, vp, vp]
    L.gp_seed_net_fit.argtypes = [vp, C.c_int64, C.c_uint64, C.c_int, C.c_int, C.c_float,
                                  C.POINTER(SeedReport)]
    L.gp_seed_net_eval.argtypes = [vp, C.c_int64, C.c_uint64, C.POINTER(SeedReport)]
    L.gp_seed_net_gate.argtypes = [vp, C.c_int, C.c_double]
    L.gp_get_contacts.argtypes = [vp, C.c_int64, ip, ip, vp, C.POINTER(C.c_int64)]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Context:
    """One device context: field, species table and particles resident on
    the GPU, plus the reference-shaped entry points of the step."""

    def __init__(self, n, L, visc, rhof, dtime, *, scheme=3, precision=PREC_FP64, ecw=1.0,
                 muw=0.0, dpmax=0.0, prefilter=2, xcd_remap=0, renorm_quat=0,
                 keep_fluid=False, timing=False, graph=False, brick=0, lm_seed=0,
                 lm_waves=0, overlap=1, cell_rule=0, tile=0):
        self.lib = load_library()
        p = Params()
        p.n, p.interp_scheme, p.precision, p.renorm_quat = n, scheme, precision, renorm_quat
        p.L, p.visc, p.rhof, p.dtime, p.ecw, p.muw, p.dpmax = L, visc, rhof, dtime, ecw, muw, dpmax
        p.prefilter, p.xcd_remap = prefilter, xcd_remap
        p.keep_fluid = 1 if keep_fluid else 0
        p.timing = 1 if timing else 0
        p.graph = int(graph) if graph in (0, 1, 2) else (1 if graph else 0)
        p.brick = brick
        p.lm_seed = lm_seed
        p.lm_waves = lm_waves
        p.overlap = overlap
        p.cell_rule = cell_rule
        p.tile = tile
        self.params = p
        h = C.c_void_p()
        rc = self.lib.gp_create(C.byref(p), C.byref(h))
        if rc:
            raise GpError(rc, "gp_create failed (no GPU visible?)")
        self.h = h
        self.np = 0

    @classmethod
    def from_workload(cls, wl, **kw):
        ctx = cls(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, scheme=wl.scheme, ecw=wl.ecw,
                  muw=wl.muw, dpmax=wl.dpmax, **kw)
        ctx.bind_field(wl.field)
        ctx.set_species(wl.species)
        ctx.set_particles(wl.state, wl.species_id)
        return ctx

    def _chk(self, rc):
        if rc:
            msg = self.lib.gp_last_error_string(self.h)
            raise GpError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.gp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- data ------------------------------------------------------------
    def bind_field(self, field):
        f = np.ascontiguousarray(field, dtype=np.float64)
        n = self.params.n
        assert f.shape == (6, n + 4, n + 4, n), f.shape
        self._chk(self.lib.gp_bind_field(self.h, _ptr(f)))

    def set_species(self, rows):
        r = np.ascontiguousarray(np.atleast_2d(rows), dtype=np.float64)
        self._chk(self.lib.gp_set_species(self.h, r.shape[0], _ptr(r)))
        self.nspec = r.shape[0]

    def set_particles(self, state, species):
        st = np.ascontiguousarray(state, dtype=np.float64)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        assert st.shape[0] == 13 and sp.shape == (st.shape[1],)
        self._chk(self.lib.gp_set_particles(self.h, st.shape[1], _ptr(st), _ptr(sp)))
        self.np = st.shape[1]

    def set_groups(self, group):
        """Assign each particle (reference order) to an independent
        ensemble; ensembles share the field but never collide."""
        g = np.ascontiguousarray(group, dtype=np.int32)
        assert g.shape == (self.np,)
        self._chk(self.lib.gp_set_groups(self.h, self.np, _ptr(g)))

    def get_particles(self):
        out = np.empty((13, self.np))
        self._chk(self.lib.gp_get_particles(self.h, self.np, _ptr(out)))
        return out

    def get_fluid_at_particles(self):
        out = np.empty((6, self.np))
        self._chk(self.lib.gp_get_fluid_at_particles(self.h, self.np, _ptr(out)))
        return out

    def sync(self):
        self._chk(self.lib.gp_sync(self.h))

    def stats(self):
        s = Stats()
        self._chk(self.lib.gp_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def reset_timers(self):
        self._chk(self.lib.gp_reset_timers(self.h))

    def particle_stats(self, npoly=None, ref_norm=True):
        """STAT_PARTICLE level 1: (nspec, 41), row (J-1) npoly + IDP-1 =
        MEAN_PART(1:41, J, IDP); npoly defaults to one class of nspec sizes.
        ref_norm: the reference's normalisation by the last class's counts."""
        npoly = npoly or self.nspec
        out = np.zeros((self.nspec, 41))
        self._chk(self.lib.gp_particle_stats(self.h, self.nspec, npoly, int(bool(ref_norm)),
                                             _ptr(out)))
        return out

    def particle_pdf(self):
        """MEAN_PART_PDF of STAT_PARTICLE (stat_particle.f90:536-645) over the
        resident particles: (nspec, 12, NPDF) counts."""
        out = np.empty((self.nspec, 12, NPDF))
        self._chk(self.lib.gp_particle_pdf(self.h, self.nspec, _ptr(out)))
        return out

    def lagrangian_origin(self, slot, mean):
        """Keep u - MEAN_PART(1:3) of every particle as the origin of
        correlation slot `slot` (stat_particle.f90:760-770); mean (nspec, 3)."""
        m = np.ascontiguousarray(mean, dtype=np.float64).reshape(self.nspec, 3)
        self._chk(self.lib.gp_lagrangian_origin(self.h, slot, self.nspec, _ptr(m)))

    def lagrangian_corr(self, slot, mean, npm):
        """Per species row the sums of (u - mean) * u_t0 / npm (the
        increments of RPX/RPY/RPZ_LOC, stat_particle.f90:790-805): (nspec, 3)."""
        m = np.ascontiguousarray(mean, dtype=np.float64).reshape(self.nspec, 3)
        n = np.ascontiguousarray(npm, dtype=np.float64).reshape(self.nspec)
        out = np.empty((self.nspec, 3))
        self._chk(self.lib.gp_lagrangian_corr(self.h, slot, self.nspec, _ptr(m), _ptr(n),
                                              _ptr(out)))
        return out

    def partfluid_stats(self, npoly=None, ref_norm=True):
        """STAT_PARTFLUID: (nspec, 36), row (J-1) npoly + IDP-1 =
        MEAN_PARTFLUID(1:36, J, IDP) from the current particles and the fluid
        at them of the last advance (needs keep_fluid=True)."""
        npoly = npoly or self.nspec
        out = np.zeros((self.nspec, 36))
        self._chk(self.lib.gp_partfluid_stats(self.h, self.nspec, npoly, int(bool(ref_norm)),
                                              _ptr(out)))
        return out

    def collision_grid(self):
        nb = np.zeros(4, dtype=np.int32)
        bm = np.zeros(6)
        self._chk(self.lib.gp_collision_grid(self.h, _ptr(nb), _ptr(bm)))
        return nb, bm

    # ---- reference-shaped entry points ----------------------------------
    def interph(self, int_scheme, comp, xint, yint, zint):
        """INTERPH (interpH.f90:26): field component comp (0..5 = u v w
        wx wy wz) at the given positions; returns UINT."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (xint, yint, zint))
        out = np.empty_like(x)
        self._chk(self.lib.gp_interph(self.h, int_scheme, len(x), _ptr(x), _ptr(y), _ptr(z),
                                      comp, _ptr(out)))
        return out

    def interph_noalloc(self, int_scheme, comp, xint, yint, zint):
        """gp_interph through the library's fixed staging (no device
        allocation); at most INTERPH_NOALLOC_MAX points."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (xint, yint, zint))
        out = np.empty_like(x)
        self._chk(self.lib.gp_interph_noalloc(self.h, int_scheme, len(x), _ptr(x), _ptr(y),
                                              _ptr(z), comp, _ptr(out)))
        return out

    def interph_field(self, int_scheme, uain, xint, yint, zint):
        """INTERPH with the field component passed in (UAIN of shape
        (n+4, n+4, n) in numpy order = UAIN(1:n,-1:n+2,-1:n+2))."""
        n = self.params.n
        u = np.ascontiguousarray(uain, dtype=np.float64)
        assert u.shape == (n + 4, n + 4, n), u.shape
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (xint, yint, zint))
        out = np.empty_like(x)
        self._chk(self.lib.gp_interph_field(self.h, int_scheme, _ptr(u), len(x), _ptr(x), _ptr(y),
                                            _ptr(z), _ptr(out)))
        return out

    def adv_particle_ellp(self, ncycle=1):
        """ADV_PARTICLE_ELLP (adv_particle_new.f90:7), BOUNDARY_PARTICLE
        included."""
        self._chk(self.lib.gp_adv_particle_ellp(self.h, ncycle))

    def collision(self, ncycle=1, time=0.0):
        """COLLISION (collision.f90:7) at NPROC = 1."""
        self._chk(self.lib.gp_collision(self.h, ncycle, time))

    def collision_advance(self, ncycle=1, time=0.0, ncycle_next=None):
        """COLLISION of cycle ncycle then ADV_PARTICLE_ELLP of the next, in
        one asynchronous call (gp_collision_advance)."""
        nxt = ncycle + 1 if ncycle_next is None else ncycle_next
        self._chk(self.lib.gp_collision_advance(self.h, ncycle, time, nxt))

    def collision_counts(self, reset=False):
        """Collisions each particle took part in (reference order) since
        set_particles or the last reset."""
        out = np.empty(self.np, dtype=np.int32)
        self._chk(self.lib.gp_get_collision_counts(self.h, self.np, _ptr(out), 1 if reset else 0))
        return out

    def set_collision_stats(self, nclass=1, npoly=None, fout2=1, ncycle=1):
        """MEAN_PART_COL of collision_interaction.f90:330-453 on every
        collision pass whose cycle number is a multiple of fout2 (0: off);
        npoly defaults to one class of nspec sizes."""
        npoly = npoly or self.nspec // nclass
        self._cs_shape = (nclass, npoly, npoly, 8)
        self._chk(self.lib.gp_set_collision_stats(self.h, nclass, npoly, fout2, ncycle))

    def collision_stats(self):
        """{"last": (nclass, npoly, npoly, 8) MEAN_PART_COL(1:8, IDP1, IDP2)
        of each class at the last statistics pass, "time_sum": their sum
        over statistics passes, "nstat": that number of passes}"""
        last = np.zeros(self._cs_shape)
        tsum = np.zeros(self._cs_shape)
        n = C.c_int64(0)
        self._chk(self.lib.gp_get_collision_stats(self.h, _ptr(last), _ptr(tsum), C.byref(n)))
        return {"last": last, "time_sum": tsum, "nstat": n.value}

    def set_dtime(self, dtime):
        """DTIME of the following steps (e.g. from the fluid's CHECK_CFL)."""
        self._chk(self.lib.gp_set_dtime(self.h, float(dtime)))

    def step(self, nsteps=1):
        """nsteps x (ADV_PARTICLE_ELLP + COLLISION), asynchronous."""
        self._chk(self.lib.gp_step(self.h, nsteps))

    def collision_detection_connectivity(self, nclosemax=None):
        """COLLISION_DETECTION_CONNECTIVITY: 1-based (HOME, NEAR) in the
        reference's order for the current positions."""
        ncl = C.c_int64(0)
        if nclosemax is None:
            rc = self.lib.gp_collision_detection(self.h, 0, None, None, C.byref(ncl))
            if rc not in (0, 6):
                self._chk(rc)
            nclosemax = ncl.value
        home = np.zeros(max(nclosemax, 1), dtype=np.int32)
        near = np.zeros(max(nclosemax, 1), dtype=np.int32)
        self._chk(self.lib.gp_collision_detection(self.h, nclosemax, _ptr(home), _ptr(near),
                                                  C.byref(ncl)))
        return home[:ncl.value], near[:ncl.value]

    def ellipsoid_contact_detection(self, home, near):
        """ellipsoid_contact_detection on 1-based pairs of the resident
        particles: (flag, EPT1, EPT2, cnormal, nfev)."""
        h = np.ascontiguousarray(home, dtype=np.int32)
        nr = np.ascontiguousarray(near, dtype=np.int32)
        n = len(h)
        flag = np.zeros(n, dtype=np.uint8)
        e1, e2, cn = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        nfev = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.gp_contact_batch(self.h, n, _ptr(h), _ptr(nr), _ptr(flag), _ptr(e1),
                                            _ptr(e2), _ptr(cn), _ptr(nfev)))
        return flag, e1, e2, cn, nfev

    def contacts(self, nmax=None):
        """The approaching contacts of the last collision pass in the order
        COLLISION_INTERACTION's loop meets them: (HOME, NEAR, resolved),
        1-based; resolved marks the pairs the one-collision pass resolved."""
        n = C.c_int64(0)
        if nmax is None:
            self._chk(self.lib.gp_get_contacts(self.h, 0, None, None, None, C.byref(n)))
            nmax = n.value
        h = np.zeros(max(nmax, 1), dtype=np.int32)
        nr = np.zeros(max(nmax, 1), dtype=np.int32)
        r = np.zeros(max(nmax, 1), dtype=np.uint8)
        self._chk(self.lib.gp_get_contacts(self.h, nmax, _ptr(h), _ptr(nr), _ptr(r), C.byref(n)))
        m = min(n.value, nmax)
        return h[:m], nr[:m], r[:m].astype(bool)

    # -- contact-seed network (no reference counterpart) -----------------
    def seed_net_gate(self, maxfev=6, keep_angle=0.02):
        """Acceptance gate of the network-seeded solve (gp_seed_net_gate)."""
        self._chk(self.lib.gp_seed_net_gate(self.h, int(maxfev), float(keep_angle)))

    def seed_net_set(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == SEED_NET_NPARAM
        self._chk(self.lib.gp_seed_net_set(self.h, _ptr(t), t.size))

    def seed_net_get(self):
        t = np.empty(SEED_NET_NPARAM, dtype=np.float32)
        self._chk(self.lib.gp_seed_net_get(self.h, _ptr(t), t.size))
        return t

    def seed_net_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], 4), dtype=np.float32)
        self._chk(self.lib.gp_seed_net_forward(self.h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def seed_net_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(SEED_NET_NPARAM, dtype=np.float32)
        loss = C.c_double(0.0)
        self._chk(self.lib.gp_seed_net_loss_grad(self.h, x.shape[0], _ptr(x), _ptr(t), _ptr(g),
                                                 C.byref(loss)))
        return g, loss.value

    def seed_net_train(self, x, t, epochs=1, batch=4096, lr=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double(0.0)
        self._chk(self.lib.gp_seed_net_train(self.h, x.shape[0], _ptr(x), _ptr(t), epochs, batch,
                                             lr, C.byref(loss)))
        return loss.value

    def seed_net_pairs(self, n, seed):
        """(features (n,12), targets (n,4): normal and flag, evaluations (n,),
        poses (n,20): b1, q1, a1, b2, q2, a2) of synthetic solver-labelled
        pairs."""
        x = np.empty((n, 12), dtype=np.float32)
        t = np.empty((n, 4), dtype=np.float32)
        nf = np.empty(n, dtype=np.int32)
        poses = np.empty((n, 20))
        self._chk(self.lib.gp_seed_net_pairs(self.h, n, seed, _ptr(x), _ptr(t), _ptr(nf),
                                             _ptr(poses)))
        return x, t, nf, poses

    def seed_net_fit(self, npairs, seed, epochs=4, batch=4096, lr=2e-3):
        r = SeedReport()
        self._chk(self.lib.gp_seed_net_fit(self.h, npairs, seed, epochs, batch, lr, C.byref(r)))
        return r.as_dict()

    def seed_net_eval(self, npairs, seed):
        r = SeedReport()
        self._chk(self.lib.gp_seed_net_eval(self.h, npairs, seed, C.byref(r)))
        return r.as_dict()
