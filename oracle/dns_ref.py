"""CPU restatement of the reference's pseudo-spectral fluid step
(ADV_FLUID, adv_fluid.f90:78-234, with NONLINEAR_TERMS, PROJ_DIVFREE, the
MESHING wavenumbers and dealiasing filter, CHECK_CFL and the STEADY
forcing of ADV_FORCING / ADD_FORCING3) -- TEST INFRASTRUCTURE: the checker
of the library's gp_dns_*, never part of the product.

The reference transforms with P3DFFT (not in this image, so the reference's
fluid step cannot be compiled here); numpy's FFT stands in with the same
conventions: forward r2c unnormalised and multiplied by FACTOR = 1/N^3
afterwards, backward c2r unnormalised.  Arrays are [k][j][i] (x fastest in
memory, as the Fortran (i, j, k)); spectral arrays [k][j][i] with i in
0..N/2 (the halved x axis, as FSTART(1):FEND(1)).

Pinned by the mathematics (tests/test_dns.py): a 2D Taylor-Green vortex and
an ABC (Beltrami) flow are exact Navier-Stokes solutions whose nonlinear
term the projection removes, so they decay exactly as exp(-nu k^2 t); a
shear wave likewise; the projection leaves k . U = 0; the filter's three
bands; the step is second order in time on a nonlinear flow; and the
Ornstein-Uhlenbeck forcing's random generator reproduces the Park-Miller
minimal standard's published check value.
"""
from __future__ import annotations

import math

import numpy as np

ZERO = 1.0e-33          # mod_dns_dim.f90:136
INFINITY = 1.0e33       # mod_dns_dim.f90:137
RTRUNC = 8.0 / 9.0      # meshing.f90:38


def wavenumbers(n: int, L: float):
    """KX (i = 0..N/2), KY, KZ (j, k = 0..N-1; indices above N/2 negative),
    meshing.f90:96-121 with KFIRST = 2 pi / L (init_run.f90:128)."""
    kf = 2.0 * math.pi / L
    kx = np.arange(n // 2 + 1) * kf
    kyz = np.array([(j if j <= n // 2 else j - n) * kf for j in range(n)])
    return kx, kyz, kyz


def dealias_filter(n: int, L: float) -> np.ndarray:
    """FILTER of meshing.f90:151-180, ALIASCONTROL = 2: 1 inside RTRUNC, a
    cos^4 taper up to the last wavenumber KLAST = pi N / L, 0 beyond."""
    kx, ky, kz = wavenumbers(n, L)
    klast = math.pi / (L / n)
    k2s = ((kz[:, None, None] / klast) ** 2 + (ky[None, :, None] / klast) ** 2 +
           (kx[None, None, :] / klast) ** 2)
    f = np.ones_like(k2s)
    band = (k2s > RTRUNC ** 2) & (k2s <= 1.0)
    f[band] = np.cos((math.pi / 2.0) * (np.sqrt(k2s[band]) - RTRUNC) / (1.0 - RTRUNC)) ** 4
    f[k2s > 1.0] = 0.0
    return f


def ftran(a: np.ndarray) -> np.ndarray:
    """P3DFFT_FTRAN_R2C followed by the * FACTOR of adv_fluid.f90:92-94"""
    n = a.shape[-1]
    return np.fft.rfftn(a, axes=(-3, -2, -1)) * (1.0 / float(n ** 3))


def btran(c: np.ndarray, n: int) -> np.ndarray:
    """P3DFFT_BTRAN_C2R (unnormalised)"""
    return np.fft.irfftn(c, s=(n, n, n), axes=(-3, -2, -1), norm="forward")


class DnsRef:
    """State: physical u (3, N, N, N) and spectral U (3, N, N, N/2+1)."""

    def __init__(self, n: int, L: float, visc: float, dtime: float, forcing: "OuForcing | None" = None):
        self.n, self.L, self.visc, self.dtime = n, L, visc, dtime
        kx, ky, kz = wavenumbers(n, L)
        self.k = (kx[None, None, :], ky[None, :, None], kz[:, None, None])
        k2 = self.k[0] ** 2 + self.k[1] ** 2 + self.k[2] ** 2
        self.kappa2 = k2
        self.kappa2_inf = np.where(k2 <= ZERO, INFINITY, k2)
        self.filter = dealias_filter(n, L)
        self.forcing = forcing
        self.u = None
        self.U = None
        self.eps = 0.0

    def set_velocity(self, u: np.ndarray) -> None:
        self.u = np.array(u, dtype=np.float64)
        self.U = ftran(self.u)

    def vorticity(self, U: np.ndarray | None = None) -> np.ndarray:
        """i k x U as NONLINEAR_TERMS forms it (nonlinear_terms.f90:104-112)"""
        U = self.U if U is None else U
        kx, ky, kz = self.k
        i = 1j
        return np.stack([i * U[2] * ky - i * U[1] * kz,
                         i * U[0] * kz - i * U[2] * kx,
                         i * U[1] * kx - i * U[0] * ky])

    def nonlinear(self, u: np.ndarray, U: np.ndarray, rhs: np.ndarray, irk: int) -> None:
        """NONLINEAR_TERMS (nonlinear_terms.f90:96-206): rhs += P(u x omega)"""
        z = btran(self.vorticity(U), self.n)
        if irk == 1:
            self.eps = self.visc * float((z[0] ** 2).sum() + (z[1] ** 2).sum() + (z[2] ** 2).sum())
        t = np.stack([u[1] * z[2] - u[2] * z[1],
                      u[2] * z[0] - u[0] * z[2],
                      u[0] * z[1] - u[1] * z[0]])
        T = ftran(t)
        kx, ky, kz = self.k
        k2 = self.kappa2_inf
        rhs[0] += (ky ** 2 + kz ** 2) / k2 * T[0] - kx * ky / k2 * T[1] - kx * kz / k2 * T[2]
        rhs[1] += -ky * kx / k2 * T[0] + (kx ** 2 + kz ** 2) / k2 * T[1] - ky * kz / k2 * T[2]
        rhs[2] += -kz * kx / k2 * T[0] - kz * ky / k2 * T[1] + (kx ** 2 + ky ** 2) / k2 * T[2]

    def project(self, U: np.ndarray) -> None:
        """PROJ_DIVFREE (proj_divfree.f90:46-60)"""
        kx, ky, kz = self.k
        k2 = self.kappa2_inf
        div = kx * U[0] + ky * U[1] + kz * U[2]
        U[0] -= kx / k2 * div
        U[1] -= ky / k2 * div
        U[2] -= kz / k2 * div

    def step(self) -> None:
        """ADV_FLUID: integrating-factor two-stage Runge-Kutta
        (adv_fluid.f90:78-234)."""
        n, dt = self.n, self.dtime
        U = ftran(self.u)
        if self.forcing is not None:
            self.forcing.advance(dt)
        e = np.exp(-self.visc * self.kappa2 * dt)
        rhs = np.zeros_like(U)
        if self.forcing is not None:
            self.forcing.add(rhs)
        self.nonlinear(self.u, U, rhs, 1)
        unp1 = (U + dt / 2 * rhs) * e
        U = (U + dt * rhs) * e
        u = btran(U.copy(), n)
        rhs = np.zeros_like(U)
        if self.forcing is not None:
            self.forcing.add(rhs)
        self.nonlinear(u, U, rhs, 2)
        U = unp1 + dt / 2 * rhs
        self.project(U)
        U *= self.filter
        self.U = U
        self.u = btran(U.copy(), n)

    def cfl(self, dtime_user: float) -> float:
        """CHECK_CFL (check_cfl.f90:61-110) with NPROC = 1: the new DTIME"""
        n, L = self.n, self.L
        dx = L / float(n - 1)
        dts = []
        for c in range(3):
            m = float(np.abs(self.u[c]).max())
            dts.append(0.05 * dx / m if m > ZERO else INFINITY)
        dt_c = min(INFINITY, *dts)
        dt_f = min(INFINITY, 1.0 * dx ** 2 / self.visc)
        return min(dt_c, dt_f, dtime_user)


# ---------------------------------------------------------------------------
# STEADY forcing: an Ornstein-Uhlenbeck process on the modes with
# KFORCE_MIN <= |k| <= KFORCE_MAX (adv_forcing.f90:33-125), driven by the
# reference's RAN1 (Park-Miller minimal standard with a Bays-Durham shuffle)
# and Box-Muller GAUSS.
# ---------------------------------------------------------------------------
class Ran1:
    IA, IM, IQ, IR, NTAB = 16807, 2147483647, 127773, 2836, 32

    def __init__(self):
        self.iv = [0] * self.NTAB
        self.iy = 0
        self.iset = 0
        self.gset = 0.0

    def _lcg(self, idum: int) -> int:
        k = idum // self.IQ
        idum = self.IA * (idum - k * self.IQ) - self.IR * k
        return idum + self.IM if idum < 0 else idum

    def ran1(self, idum: int) -> tuple[float, int]:
        ndiv = 1 + (self.IM - 1) // self.NTAB
        am = 1.0 / self.IM
        rnmx = 1.0 - 1.2e-7
        if idum <= 0 or self.iy == 0:
            idum = max(-idum, 1)
            for j in range(self.NTAB + 8, 0, -1):
                idum = self._lcg(idum)
                if j <= self.NTAB:
                    self.iv[j - 1] = idum
            self.iy = self.iv[0]
        idum = self._lcg(idum)
        j = self.iy // ndiv
        self.iy = self.iv[j]
        self.iv[j] = idum
        return min(am * self.iy, rnmx), idum

    def gauss(self, idum: int) -> tuple[float, int]:
        """polar Box-Muller, one value cached"""
        if self.iset == 0:
            while True:
                r, idum = self.ran1(idum)
                v1 = 2.0 * r - 1.0
                r, idum = self.ran1(idum)
                v2 = 2.0 * r - 1.0
                rsq = v1 * v1 + v2 * v2
                if 0.0 < rsq < 1.0:
                    break
            fac = math.sqrt(-2.0 * math.log(rsq) / rsq)
            self.gset = v1 * fac
            self.iset = 1
            return v2 * fac, idum
        self.iset = 0
        return self.gset, idum


class OuForcing:
    """Forcing modes (i, j, k) in the reference's loop order, their complex
    amplitudes per component, and the generator state."""

    def __init__(self, n: int, L: float, kmin: float, kmax: float, sigma: float, tforce: float,
                 seed: int, f0: np.ndarray | None = None):
        kx, ky, kz = wavenumbers(n, L)
        modes = []
        for k in range(n):
            for j in range(n):
                for i in range(n // 2 + 1):
                    kf = (kx[i] ** 2 + ky[j] ** 2 + kz[k] ** 2) ** 0.5
                    if kmin <= kf <= kmax:
                        modes.append((i, j, k, kf))
        self.modes = modes
        self.kmax, self.sigma, self.tforce = kmax, sigma, tforce
        self.f = np.zeros((3, len(modes)), dtype=np.complex128) if f0 is None else f0.copy()
        self.rng = Ran1()
        self.idum = seed

    def advance(self, dt: float) -> None:
        alpha = dt / self.tforce
        beta = (2.0 * self.sigma ** 2 * dt / self.tforce) ** 0.5
        for nf, (_, _, _, kf) in enumerate(self.modes):
            xampf = (self.kmax - kf) / 0.20 / self.kmax
            xampl = (math.exp(2.0 * xampf) - 1.0) / (math.exp(2.0 * xampf) + 1.0)
            for c in range(3):
                g, self.idum = self.rng.gauss(self.idum)
                re = self.f[c, nf].real * (1.0 - alpha) + g * xampl * beta
                g, self.idum = self.rng.gauss(self.idum)
                im = self.f[c, nf].imag * (1.0 - alpha) + g * xampl * beta
                self.f[c, nf] = complex(re, im)

    def add(self, rhs: np.ndarray) -> None:
        """ADD_FORCING3 (add_forcing3.f90:40-58)"""
        for nf, (i, j, k, _) in enumerate(self.modes):
            rhs[:, k, j, i] += self.f[:, nf]
