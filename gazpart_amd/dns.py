"""The spectral DNS field supplier (gp_dns_*): the reference's ADV_FLUID
pseudo-spectral Navier-Stokes step on the GPU, whose velocity and spectral
vorticity feed a particle Context without a host copy.

    from gazpart_amd.dns import Dns
    d = Dns(n=256, L=2*pi, visc=0.01, dtime=1e-3)
    d.set_velocity(u)            # (3, n, n, n), x fastest ([k][j][i])
    d.step(10)                   # ADV_FLUID x 10
    d.bind(ctx)                  # velocity + vorticity -> ctx's field
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import GpError, load_library


class DnsParams(C.Structure):
    """gp_dns_params (include/gazpart_gpu.h)."""
    _fields_ = [("n", C.c_int), ("L", C.c_double), ("visc", C.c_double),
                ("dtime", C.c_double), ("forcing", C.c_int), ("seed", C.c_int),
                ("kforce_min", C.c_double), ("kforce_max", C.c_double),
                ("sigma_force", C.c_double), ("time_force", C.c_double),
                ("reuse_spectral", C.c_int), ("reserved", C.c_int * 3)]


def _lib():
    L = load_library()
    if not getattr(L, "_dns_types", False):
        vp = C.c_void_p
        L.gp_dns_create.argtypes = [C.POINTER(DnsParams), C.POINTER(C.c_void_p)]
        L.gp_dns_destroy.argtypes = [vp]
        L.gp_dns_error_string.argtypes = [vp]
        L.gp_dns_error_string.restype = C.c_char_p
        L.gp_dns_set_velocity.argtypes = [vp, vp]
        L.gp_dns_get_velocity.argtypes = [vp, vp]
        L.gp_dns_get_vorticity.argtypes = [vp, vp]
        L.gp_dns_step.argtypes = [vp, C.c_int]
        L.gp_dns_cfl.argtypes = [vp, C.c_double, C.POINTER(C.c_double)]
        L.gp_dns_info.argtypes = [vp, C.POINTER(C.c_double), vp, C.POINTER(C.c_int)]
        L.gp_dns_bind.argtypes = [vp, vp, C.c_int]
        L._dns_types = True
    return L


class Dns:
    """One DNS state on the GPU (double precision)."""

    def __init__(self, n: int, L: float, visc: float, dtime: float, forcing: dict | None = None,
                 reuse_spectral: bool = False):
        self.lib = _lib()
        p = DnsParams()
        p.n, p.L, p.visc, p.dtime = n, L, visc, dtime
        p.reuse_spectral = 1 if reuse_spectral else 0
        if forcing:
            p.forcing = 1
            p.seed = int(forcing.get("seed", -1))
            p.kforce_min = float(forcing["kmin"])
            p.kforce_max = float(forcing["kmax"])
            p.sigma_force = float(forcing["sigma"])
            p.time_force = float(forcing["tforce"])
        h = C.c_void_p()
        rc = self.lib.gp_dns_create(C.byref(p), C.byref(h))
        if rc:
            raise GpError(rc, "gp_dns_create")
        self.h = h
        self.n = n

    def _chk(self, rc):
        if rc:
            raise GpError(rc, self.lib.gp_dns_error_string(self.h).decode())

    def close(self):
        if self.h:
            self.lib.gp_dns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_velocity(self, u: np.ndarray) -> None:
        a = np.ascontiguousarray(u, dtype=np.float64)
        assert a.shape == (3, self.n, self.n, self.n)
        self._chk(self.lib.gp_dns_set_velocity(self.h, a.ctypes.data))

    def velocity(self) -> np.ndarray:
        out = np.empty((3, self.n, self.n, self.n))
        self._chk(self.lib.gp_dns_get_velocity(self.h, out.ctypes.data))
        return out

    def vorticity(self) -> np.ndarray:
        out = np.empty((3, self.n, self.n, self.n))
        self._chk(self.lib.gp_dns_get_vorticity(self.h, out.ctypes.data))
        return out

    def step(self, nsteps: int = 1) -> None:
        self._chk(self.lib.gp_dns_step(self.h, nsteps))

    def cfl(self, dtime_user: float) -> float:
        dt = C.c_double(0.0)
        self._chk(self.lib.gp_dns_cfl(self.h, dtime_user, C.byref(dt)))
        return dt.value

    def info(self) -> dict:
        eps, nf = C.c_double(0.0), C.c_int(0)
        um = np.zeros(3)
        self._chk(self.lib.gp_dns_info(self.h, C.byref(eps), um.ctypes.data, C.byref(nf)))
        return {"eps": eps.value, "umax": um, "nforce": nf.value}

    def bind(self, ctx, mesh_vorticity: bool = True) -> None:
        self._chk(self.lib.gp_dns_bind(self.h, ctx.h, 1 if mesh_vorticity else 0))
