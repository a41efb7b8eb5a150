"""The synthetic frozen field of gazpart_amd.workload (no reference
counterpart; the configurations ask for a divergence-free random-Fourier
velocity with E(k) ~ k^4 exp(-2 (k/k0)^2) and u_rms = 1), checked against
what the mathematics fixes (CPU only):

* divergence: the velocity on its periodic transform grid has k . u_hat = 0
  to rounding (spectral derivative of the stored samples);
* u_rms = sqrt(<|u|^2> / 3) = 1;
* spectrum: shell-summed energy follows k^4 exp(-2 (k/k0)^2) within the
  sampling error of random Gaussian amplitudes (chi-square, 2 x 3 degrees
  of freedom per mode);
* vorticity: the stored vorticity is the curl of the stored velocity on the
  reference mesh DX = L/(N-1) (meshing.f90:43), checked by 4th-order central
  differences in index space (ghost planes in y, z; index period N in x);
* ghost planes: j, k in {-1, 0, N+1, N+2} repeat {N-1, N, 1, 2}.
"""
import math

import numpy as np
import pytest

from gazpart_amd import workload as W


@pytest.fixture(scope="module")
def fld64():
    return W.random_fourier_field(64, 4.0, 11)


def interior(f, n):
    # [c, k, j, i] with k, j the interior planes
    return f[:, 2:n + 2, 2:n + 2, :]


def test_divergence_free_spectrally(fld64):
    f, _ = fld64
    n = 64
    u = interior(f, n)[:3]
    uh = [np.fft.fftn(u[c], axes=(0, 1, 2)) for c in range(3)]    # axes (z, y, x)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    div = kx * uh[0] + ky * uh[1] + kz * uh[2]
    mag = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) * np.sqrt(sum(np.abs(h) ** 2 for h in uh))
    assert np.abs(div).max() <= 1e-12 * mag.max()
    # and not trivially: a single component alone is not divergence-free
    assert np.abs(kx * uh[0]).max() > 1e-3 * mag.max()


def test_urms_is_one_and_umax(fld64):
    f, umax = fld64
    u = interior(f, 64)[:3]
    urms = math.sqrt(float((u * u).sum(axis=0).mean()) / 3.0)
    assert abs(urms - 1.0) < 1e-12
    assert abs(umax - math.sqrt(float((u * u).sum(axis=0).max()))) < 1e-12


def test_spectrum_shape():
    n, k0 = 64, 4.0
    f, _ = W.random_fourier_field(n, k0, 5)
    u = interior(f, n)[:3]
    e = sum(np.abs(np.fft.fftn(u[c], axes=(0, 1, 2))) ** 2 for c in range(3)) / n ** 6
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    shell = np.rint(kk).astype(int)
    # energy per mode ~ amp^2 = k^2 exp(-2 (k/k0)^2) (times the projection's
    # 2/3 of a Gaussian vector), so shell energy ~ N_modes(k) k^2 exp(...),
    # i.e. ~ k^4 exp(-2 (k/k0)^2) for a filled shell
    ratios = []
    for s in range(2, 13):
        m = shell == s
        model = (kk[m] ** 2 * np.exp(-2.0 * (kk[m] / k0) ** 2)).sum()
        meas = e[m].sum()
        nmodes = m.sum() // 2          # complex-conjugate pairs
        ratios.append((meas / model, 1.0 / math.sqrt(3.0 * nmodes)))
    r = np.array([a for a, _ in ratios])
    sig = np.array([b for _, b in ratios])
    c = np.median(r)
    assert (np.abs(r / c - 1.0) <= 4.0 * sig + 0.02).all(), (r / c, sig)
    # a k^2 spectrum instead would be far off
    r2 = []
    for s in range(2, 13):
        m = shell == s
        r2.append(e[m].sum() / (np.exp(-2.0 * (kk[m] / k0) ** 2)).sum())
    r2 = np.array(r2)
    assert (r2 / np.median(r2)).max() > 3.0


def test_vorticity_is_curl_on_reference_mesh():
    n, L = 128, 2.0 * math.pi
    f, _ = W.random_fourier_field(n, 2.0, 3)
    dx = L / (n - 1)
    u, v, w = f[0], f[1], f[2]                  # [k, j, i], k/j with 2 ghosts each side

    def d(a, axis):
        # 4th-order central difference on interior points
        if axis == 2:                            # x: index-periodic with period n
            p1, m1 = np.roll(a, -1, axis=2), np.roll(a, 1, axis=2)
            p2, m2 = np.roll(a, -2, axis=2), np.roll(a, 2, axis=2)
            out = (8.0 * (p1 - m1) - (p2 - m2)) / (12.0 * dx)
            return out[2:n + 2, 2:n + 2, :]
        sl = lambda o: (slice(2 + o, n + 2 + o), slice(2, n + 2), slice(None)) if axis == 0 \
            else (slice(2, n + 2), slice(2 + o, n + 2 + o), slice(None))
        return (8.0 * (a[sl(1)] - a[sl(-1)]) - (a[sl(2)] - a[sl(-2)])) / (12.0 * dx)

    wx = d(w, 1) - d(v, 0)
    wy = d(u, 0) - d(w, 2)
    wz = d(v, 2) - d(u, 1)
    for c, fd in ((3, wx), (4, wy), (5, wz)):
        st = f[c, 2:n + 2, 2:n + 2, :]
        assert np.abs(fd - st).max() <= 2e-3 * np.abs(st).max(), c
    # a vorticity with the wrong mesh factor (n/(n-1) off) would not pass
    st = f[3, 2:n + 2, 2:n + 2, :] * (n / (n - 1.0)) ** 2
    assert np.abs(wx - st).max() > 2e-3 * np.abs(st).max()


def test_ghost_planes(fld64):
    f, _ = fld64
    n = 64
    for dst, src in ((0, n), (1, n + 1), (n + 2, 2), (n + 3, 3)):
        np.testing.assert_array_equal(f[:, :, dst, :], f[:, :, src, :])
        np.testing.assert_array_equal(f[:, dst, :, :], f[:, src, :, :])
