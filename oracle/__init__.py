"""ctypes bindings of the CPU FP64 oracle and of the compiled reference.

TEST INFRASTRUCTURE.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as the path being measured.

* ``lib()``  -- ``oracle/_build/liboracle.so``: the C restatement
  (gazpart_oracle.c), built by ``oracle/Makefile``.
* ``ref()``  -- ``oracle/_ref/libgazpart_ref.so``: the reference's own
  Fortran hot path compiled unmodified by ``oracle/ref/Makefile`` with a
  C-callable driver; present only where the reference sources exist.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgazpart_ref.so")
REF_SRC = "/root/reference/source"

NSTATE = 13
SP_STRIDE = 14

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    """oc_params (gazpart_oracle.h)."""
    _fields_ = [("n", C.c_int), ("interp_scheme", C.c_int), ("renorm_quat", C.c_int),
                ("pad_", C.c_int), ("L", C.c_double), ("dx", C.c_double),
                ("visc", C.c_double), ("rhof", C.c_double), ("dtime", C.c_double),
                ("ecw", C.c_double), ("muw", C.c_double)]


class CGrid(C.Structure):
    """oc_cgrid (gazpart_oracle.h)."""
    _fields_ = [("nxbox", C.c_int), ("nybox", C.c_int), ("nzbox", C.c_int),
                ("nbox", C.c_int), ("xminbox", C.c_double), ("yminbox", C.c_double),
                ("zminbox", C.c_double), ("dxbox", C.c_double), ("dybox", C.c_double),
                ("dzbox", C.c_double)]


_lib = None
_ref = None


def build() -> None:
    """Compile the C oracle (and the reference driver when the reference
    sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        L.oc_interph.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, _dp, _dp]
        L.oc_quaternion_integration.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.oc_fluid_force.argtypes = [C.POINTER(Params), _dp] + [C.c_double] * 6 + [_dp, _dp, _dp]
        L.oc_fluid_torque.argtypes = [C.POINTER(Params), _dp] + [C.c_double] * 12 + [_dp, _dp, _dp]
        L.oc_wall_test.argtypes = [C.POINTER(Params), _dp, C.c_double, C.c_double, C.c_double,
                                   _dp, _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.oc_wall_rebound.argtypes = [C.POINTER(Params), _dp] + [C.POINTER(C.c_double)] * 6 + \
            [_dp, _dp, _dp]
        L.oc_adv_particle_ellp.argtypes = [C.POINTER(Params), _dp, _dp, C.c_int, _dp, _ip,
                                           C.c_void_p, C.c_int]
        L.oc_collision_grid_init.argtypes = [C.POINTER(Params), C.c_double, C.POINTER(CGrid)]
        L.oc_box_connectivity.argtypes = [C.POINTER(CGrid), C.c_int, _ip, _ip]
        L.oc_collision_detection.argtypes = [C.POINTER(CGrid), _ip, _ip, C.c_int, _dp, _dp, _dp,
                                             C.c_int64, _ip, _ip, C.POINTER(C.c_int64)]
        L.oc_contact_detection.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_void_p,
                                           C.POINTER(C.c_int), _dp, _dp, _dp,
                                           C.POINTER(C.c_int), _dp]
        L.oc_collision_response.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oc_collision_interaction.argtypes = [C.POINTER(Params), _dp, C.c_int, _dp, _ip,
                                               C.c_int64, _ip, _ip, C.c_void_p, C.c_void_p,
                                               C.POINTER(C.c_int), C.c_int]
        L.oc_collision_interaction_stats.argtypes = [
            C.POINTER(Params), _dp, C.c_int, _dp, _ip, C.c_int64, _ip, _ip, C.c_void_p,
            C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def params(wl, renorm: int = 0) -> Params:
    return Params(n=wl.n, interp_scheme=wl.scheme, renorm_quat=renorm, pad_=0, L=wl.L,
                  dx=wl.L / (wl.n - 1), visc=wl.visc, rhof=wl.rhof, dtime=wl.dtime,
                  ecw=wl.ecw, muw=wl.muw)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# oracle functions (numpy in / out)
# ---------------------------------------------------------------------------
def interph(scheme, n, dx, comp_field, x, y, z):
    x, y, z = _c(x), _c(y), _c(z)
    out = np.zeros_like(x)
    err = lib().oc_interph(scheme, n, dx, _c(comp_field).reshape(-1), len(x), x, y, z, out)
    return out, err


def quat(q, ox, oy, oz, dt):
    q = _c(q).copy()
    lib().oc_quaternion_integration(q, ox, oy, oz, dt)
    return q


def force(p, sp, up, vp, wp, uf, vf, wf, q):
    fd, fl = np.zeros(3), np.zeros(3)
    err = lib().oc_fluid_force(C.byref(p), _c(sp), up, vp, wp, uf, vf, wf, _c(q), fd, fl)
    return fd, fl, err


def torque(p, sp, fl6, upn, ox, oy, oz, q):
    tp, tr = np.zeros(3), np.zeros(3)
    err = lib().oc_fluid_torque(C.byref(p), _c(sp), fl6[0], fl6[1], fl6[2], upn[0], upn[1],
                                upn[2], fl6[3], fl6[4], fl6[5], ox, oy, oz, _c(q), tp, tr)
    return tp, tr, err


def wall_test(p, sp, x, y, z, q):
    nrm, arm = np.zeros(3), np.zeros(3)
    dxw, hit = C.c_double(0.0), C.c_int(0)
    lib().oc_wall_test(C.byref(p), _c(sp), x, y, z, _c(q), nrm, arm, C.byref(dxw), C.byref(hit))
    return nrm, arm, dxw.value, hit.value


def wall_rebound(p, sp, uvw, omg, q, arm, nrm):
    v = [C.c_double(float(t)) for t in list(uvw) + list(omg)]
    err = lib().oc_wall_rebound(C.byref(p), _c(sp), *[C.byref(t) for t in v], _c(q), _c(arm),
                                _c(nrm))
    return np.array([t.value for t in v[:3]]), np.array([t.value for t in v[3:]]), err


def adv(p, field, spec, state, sid, nthreads=1, want_fluid=False):
    """ADV_PARTICLE_ELLP on a copy of state; returns (state, fluid, err)."""
    st = _c(state).copy()
    np_ = st.shape[1]
    fl = np.zeros((6, np_)) if want_fluid else None
    err = lib().oc_adv_particle_ellp(C.byref(p), _c(field).reshape(-1), _c(spec).reshape(-1),
                                     np_, st.reshape(-1), _c(sid, np.int32),
                                     fl.ctypes.data if want_fluid else None, nthreads)
    return st, fl, err


def collision_grid(p, dpmax):
    g = CGrid()
    lib().oc_collision_grid_init(C.byref(p), dpmax, C.byref(g))
    return g


def connectivity(g, periodic=1):
    con = np.zeros(g.nbox * 13, dtype=np.int32)
    ncon = np.zeros(g.nbox, dtype=np.int32)
    lib().oc_box_connectivity(C.byref(g), periodic, con, ncon)
    return con, ncon


def detect(g, con, ncon, x, y, z, nmax):
    home = np.zeros(max(nmax, 1), dtype=np.int32)
    near = np.zeros(max(nmax, 1), dtype=np.int32)
    nc = C.c_int64(0)
    err = lib().oc_collision_detection(C.byref(g), con, ncon, len(x), _c(x), _c(y), _c(z), nmax,
                                       home, near, C.byref(nc))
    return home[:min(nc.value, nmax)], near[:min(nc.value, nmax)], nc.value, err


def contact(spi, spj, bi, qi, bj, qj, seed=None):
    f, nf = C.c_int(0), C.c_int(0)
    e1, e2, cn, xs = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(2)
    sd = _c(seed) if seed is not None else None
    lib().oc_contact_detection(_c(spi), _c(spj), _c(bi), _c(qi), _c(bj), _c(qj),
                               sd.ctypes.data if sd is not None else None, C.byref(f), e1, e2,
                               cn, C.byref(nf), xs)
    return f.value, e1, e2, cn, nf.value, xs


def response(sp1, s1, s2, r1, r2, nrm):
    s1, s2 = _c(s1).copy(), _c(s2).copy()
    err = lib().oc_collision_response(_c(sp1), s1, s2, _c(s1[6:10]), _c(s2[6:10]), _c(r1),
                                      _c(r2), _c(nrm))
    return s1, s2, err


def interaction(p, spec, state, sid, home, near, nthreads=1):
    st = _c(state).copy()
    np_ = st.shape[1]
    nclose = len(home)
    flags = np.zeros(max(nclose, 1), dtype=np.int32)
    coll = np.zeros(max(nclose, 1), dtype=np.int32)
    nc = C.c_int(0)
    err = lib().oc_collision_interaction(C.byref(p), _c(spec).reshape(-1), np_, st.reshape(-1),
                                         _c(sid, np.int32), nclose, _c(home, np.int32),
                                         _c(near, np.int32), flags.ctypes.data, coll.ctypes.data,
                                         C.byref(nc), nthreads)
    return st, flags[:nclose], coll[:nclose], nc.value, err


def interaction_stats(p, spec, state, sid, home, near, npoly, nbin, nthreads=1):
    """interaction() with the level-1 collision statistics: also returns
    col (nbin, 8) = MEAN_PART_COL(1:8) per (species[HOME], IDP of NEAR)"""
    st = _c(state).copy()
    np_ = st.shape[1]
    nclose = len(home)
    flags = np.zeros(max(nclose, 1), dtype=np.int32)
    coll = np.zeros(max(nclose, 1), dtype=np.int32)
    col = np.zeros((nbin, 8))
    nc = C.c_int(0)
    err = lib().oc_collision_interaction_stats(
        C.byref(p), _c(spec).reshape(-1), np_, st.reshape(-1), _c(sid, np.int32), nclose,
        _c(home, np.int32), _c(near, np.int32), flags.ctypes.data, coll.ctypes.data,
        C.byref(nc), nthreads, npoly, nbin, col.reshape(-1))
    return st, flags[:nclose], coll[:nclose], nc.value, err, col


class Stepper:
    """Full oracle step (ADV_PARTICLE_ELLP then COLLISION) on one workload."""

    def __init__(self, wl, renorm=0, nthreads=1, dpmax=None, nclosemax=None):
        self.wl = wl
        self.p = params(wl, renorm)
        self.g = collision_grid(self.p, wl.dpmax if dpmax is None else dpmax)
        self.con, self.ncon = connectivity(self.g)
        self.nthreads = nthreads
        self.nmax = nclosemax or max(64 * wl.np_, 1024)
        self.home = np.zeros(self.nmax, dtype=np.int32)
        self.near = np.zeros(self.nmax, dtype=np.int32)

    def step(self, state):
        wl = self.wl
        st, _, err = adv(self.p, wl.field, wl.species, state, wl.species_id, self.nthreads)
        if err:
            return st, err, 0, 0
        home, near, nc, err = detect(self.g, self.con, self.ncon, st[0], st[1], st[2], self.nmax)
        if err:
            return st, err, nc, 0
        st, flags, coll, ncoll, err = interaction(self.p, wl.species, st, wl.species_id, home,
                                                  near, self.nthreads)
        return st, err, nc, ncoll


# ---------------------------------------------------------------------------
# compiled reference (Fortran) driver
# ---------------------------------------------------------------------------
class Ref:
    """The reference's own subroutines through oracle/ref/ref_driver.f90.

    Classes (IG) are 1-based; sizes (IDP) are 1-based; state arrays are the
    (13, np) component-major layout, which is the Fortran st(np, 13)."""

    def __init__(self):
        if not ref_available():
            raise RuntimeError("reference build missing: make -C oracle/ref")
        global _ref
        if _ref is None:
            _ref = C.CDLL(REF_PATH)
        self.L = _ref

    def setup(self, n, L, visc, rhof, dtime, ecw, muw, scheme, emaj, apr, rhop, grav=None,
              ecp=None, mup=None):
        """emaj, rhop: (nclass, nsize); apr, grav, ecp, mup: (nclass,)"""
        emaj = np.atleast_2d(np.asarray(emaj, float))
        rhop = np.atleast_2d(np.asarray(rhop, float))
        nclass, nsize = emaj.shape
        apr = np.asarray(apr, float).reshape(nclass)
        grav = np.zeros(nclass) if grav is None else np.asarray(grav, float).reshape(nclass)
        ecp = np.ones(nclass) if ecp is None else np.asarray(ecp, float).reshape(nclass)
        mup = np.zeros(nclass) if mup is None else np.asarray(mup, float).reshape(nclass)
        self.n, self.nclass, self.nsize = n, nclass, nsize
        f = self.L.ref_setup
        f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                      C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        f(n, L, visc, rhof, dtime, ecw, muw, scheme, nclass, nsize,
          _c(emaj.T), _c(apr), _c(rhop.T), _c(grav), _c(ecp), _c(mup))

    def set_field(self, field):
        f = self.L.ref_set_field
        f.argtypes = [C.c_int, _dp]
        for c in range(6):
            f(c + 1, _c(field[c]).reshape(-1))

    def set_particles(self, ig, state, idp):
        f = self.L.ref_set_particles
        f.argtypes = [C.c_int, C.c_int, _dp, _ip]
        st = _c(state)
        f(ig, st.shape[1], st.reshape(-1), _c(idp, np.int32))
        self.np_ = st.shape[1]

    def get_particles(self, ig, np_):
        f = self.L.ref_get_particles
        f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        st, fl, col = np.zeros((13, np_)), np.zeros((6, np_)), np.zeros(np_)
        f(ig, np_, st.reshape(-1), fl.reshape(-1), col)
        return st, fl, col

    def interph(self, scheme, comp, x, y, z):
        f = self.L.ref_interph
        f.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        out = np.zeros(len(x))
        f(scheme, comp + 1, len(x), _c(x), _c(y), _c(z), out)
        return out

    def force(self, ig, idp, up, vp, wp, uf, vf, wf, q):
        f = self.L.ref_force
        f.argtypes = [C.c_int, C.c_int] + [C.c_double] * 6 + [_dp, _dp, _dp]
        fd, fl = np.zeros(3), np.zeros(3)
        f(ig, idp, up, vp, wp, uf, vf, wf, _c(q), fd, fl)
        return fd, fl

    def torque(self, ig, idp, st13, fl6, ox, oy, oz, q):
        f = self.L.ref_torque
        f.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp,
                      _dp, _dp]
        tp, tr = np.zeros(3), np.zeros(3)
        f(ig, idp, _c(st13), _c(fl6), ox, oy, oz, _c(q), tp, tr)
        return tp, tr

    def quat(self, q, ox, oy, oz, dt):
        f = self.L.ref_quat
        f.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double]
        q = _c(q).copy()
        f(q, ox, oy, oz, dt)
        return q

    def wall(self, ig, idp, x, y, z, q):
        f = self.L.ref_wall
        f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                      C.POINTER(C.c_double), C.POINTER(C.c_int)]
        nrm, arm = np.zeros(3), np.zeros(3)
        dxw, hit = C.c_double(0.0), C.c_int(0)
        f(ig, idp, x, y, z, _c(q), nrm, arm, C.byref(dxw), C.byref(hit))
        return nrm, arm, dxw.value, hit.value

    def rebound(self, ig, idp, uvw, omg, q, arm, nrm):
        f = self.L.ref_rebound
        f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        u, o = _c(uvw).copy(), _c(omg).copy()
        f(ig, idp, u, o, _c(q), _c(arm), _c(nrm))
        return u, o

    def contact(self, ig, idp1, idp2, b1, q1, b2, q2):
        f = self.L.ref_contact
        f.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_int), _dp,
                      _dp, _dp]
        flag = C.c_int(0)
        e1, e2, cn = np.zeros(3), np.zeros(3), np.zeros(3)
        f(ig, idp1, idp2, _c(b1), _c(q1), _c(b2), _c(q2), C.byref(flag), e1, e2, cn)
        return flag.value, e1, e2, cn

    def response(self, ig, idp, s1, s2, r1, r2, nrm):
        f = self.L.ref_response
        f.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        a, b = _c(s1).copy(), _c(s2).copy()
        f(ig, idp, a, b, _c(r1), _c(r2), _c(nrm))
        return a, b

    def adv(self, ncycle=1):
        self.L.ref_adv.argtypes = [C.c_int]
        self.L.ref_adv(ncycle)

    def collision_init(self):
        f = self.L.ref_collision_init
        f.argtypes = [_ip, _dp]
        inum, rnum = np.zeros(4, dtype=np.int32), np.zeros(6)
        f(inum, rnum)
        return inum, rnum

    def connectivity(self, nbox):
        f = self.L.ref_connectivity
        f.argtypes = [C.c_int, _ip, _ip]
        con = np.zeros(13 * nbox, dtype=np.int32)
        ncon = np.zeros(nbox, dtype=np.int32)
        f(nbox, con, ncon)
        return con, ncon

    def detect(self, ig, nmax):
        f = self.L.ref_detect
        f.argtypes = [C.c_int, C.c_int, _ip, _ip, C.POINTER(C.c_int)]
        home, near = np.zeros(nmax, dtype=np.int32), np.zeros(nmax, dtype=np.int32)
        nc = C.c_int(0)
        f(ig, nmax, home, near, C.byref(nc))
        return home[:nc.value], near[:nc.value]

    def partfluid_stats(self, ncycle=1):
        """MEAN_PARTFLUID(1:36, J, IDP) of STAT_PARTFLUID as [J-1, IDP-1, k-1]"""
        f = self.L.ref_partfluid_stats
        f.argtypes = [C.c_int, _dp]
        out = np.zeros(36 * self.nsize * self.nclass)
        f(ncycle, out)
        return out.reshape(self.nclass, self.nsize, 36).copy()

    def cstat_on(self, fout2):
        self.L.ref_cstat_on.argtypes = [C.c_int]
        self.L.ref_cstat_on(fout2)

    def cstat_get(self):
        """MEAN_TIME_PART_COL(1:8, :, :) as [IDP1-1, IDP2-1, k-1]"""
        self.L.ref_cstat_get.argtypes = [_dp]
        out = np.zeros(8 * self.nsize * self.nsize)
        self.L.ref_cstat_get(out)
        return out.reshape(self.nsize, self.nsize, 8).transpose(1, 0, 2).copy()

    def collision(self, ig, ncycle, time, nmax):
        f = self.L.ref_collision
        f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _ip, _ip, C.POINTER(C.c_int)]
        home, near = np.zeros(nmax, dtype=np.int32), np.zeros(nmax, dtype=np.int32)
        nc = C.c_int(0)
        f(ig, ncycle, time, nmax, home, near, C.byref(nc))
        return home[:nc.value], near[:nc.value]
