"""Probe formation of the oracle (stemlib.cpp:1952-2191) against closed forms."""
import numpy as np

from qstem_amd import StemConfig


def cfg(**kw):
    base = dict(nx=64, ny=64, resolutionX=0.1, resolutionY=0.1, v0=200.0, slices=1, sliceThickness=1.0,
                potNx=64, potNy=64, alpha=20.0)
    base.update(kw)
    return StemConfig(**base)


def test_probe_normalised(oracle):
    for kw in ({}, dict(Cs=1e4, df0=-50.0), dict(gaussFlag=1, gaussScale=0.1), dict(aAIS=3.0), dict(ismoth=1)):
        psi = oracle.probe(cfg(**kw))
        assert abs(np.sum(np.abs(psi) ** 2) - 64 * 64) < 1e-8


def test_aberration_free_probe_is_real_and_peaked(oracle):
    """Cs = df = 0: psi(k) is a real non-negative aperture times the shift phase, so the
    real-space probe peaks at the window centre with value sum(aperture)*scale."""
    c = cfg()
    psi = oracle.probe(c)
    ix, iy = np.unravel_index(np.argmax(np.abs(psi)), psi.shape)
    assert (ix, iy) == (32, 32)
    assert abs(psi[32, 32].imag) < 1e-9 * abs(psi[32, 32])
    wl = 12.26 / np.sqrt(200e3 + 200 ** 2 * 0.9788)  # probe's own wavelength formula
    k = np.fft.fftfreq(64, d=0.1)
    k2 = k[:, None] ** 2 + k[None, :] ** 2
    ap = (k2 <= (np.sin(0.02) / wl) ** 2).astype(float)
    pk = np.fft.ifft2(ap) * ap.size  # unnormalised backward transform, centred at 0
    norm = np.sqrt(64 * 64 / np.sum(np.abs(pk) ** 2))
    assert abs(abs(psi[32, 32]) - abs(pk[0, 0]) * norm) < 1e-9


def test_chi_closed_form(oracle):
    """Defocus + Cs phase: psi(k) = exp(-i chi) with chi = pi/lambda (df t^2 + Cs t^4 / 2), t = lambda k."""
    c = cfg(Cs=5e4, df0=-120.0, alpha=30.0)
    psi = oracle.probe(c)
    Psi = np.fft.fft2(psi)
    wl = 12.26 / np.sqrt(200e3 + 200 ** 2 * 0.9788)
    kx = np.fft.fftfreq(64, d=0.1)
    k2 = kx[:, None] ** 2 + kx[None, :] ** 2
    inside = k2 <= (np.sin(0.03) / wl) ** 2
    t2 = k2 * wl ** 2
    chi = 2 * np.pi / wl * (t2 * c.df0 / 2 + t2 ** 2 * c.Cs / 4)
    # centring shift x0 = ax - (nx/2)*resX, with resX stored as float (MULS float_tt)
    r32 = float(np.float32(0.1))
    x0 = 64 * r32 - float(np.float32(32 * np.float32(0.1)))
    kx = np.fft.fftfreq(64, d=r32)
    k2 = kx[:, None] ** 2 + kx[None, :] ** 2
    inside = k2 <= (np.sin(0.03) / wl) ** 2
    t2 = k2 * wl ** 2
    chi = 2 * np.pi / wl * (t2 * c.df0 / 2 + t2 ** 2 * c.Cs / 4)
    shift = np.exp(-2j * np.pi * x0 * (kx[:, None] + kx[None, :]))
    ratio = Psi[inside] / (np.exp(-1j * chi[inside]) * shift[inside])
    assert np.allclose(ratio, ratio[0], rtol=1e-9, atol=1e-9)
    assert np.all(np.abs(Psi[~inside]) < 1e-9 * np.abs(Psi).max())


def test_probe_at_position_is_a_shift(oracle):
    """qo_probe_at: moving the centre by whole pixels (dx, dy) is a pure phase ramp in k,
    so the real-space probe is the centred one rolled by (dx, dy) pixels; the centre
    itself reproduces qo_probe.  A wrong sign of the shift term fails the roll."""
    c = cfg(Cs=2e4, df0=-40.0)
    psi0 = oracle.probe(c)
    cx = float(np.float32(32) * np.float32(0.1))
    assert np.array_equal(oracle.probe(c, at=(cx, cx)), psi0)
    dx = float(np.float32(0.1))  # the pixel size as the config stores it
    shifted = oracle.probe(c, at=(cx + 5 * dx, cx - 3 * dx))
    assert np.abs(shifted - np.roll(psi0, (5, -3), axis=(0, 1))).max() < 1e-9
    ix, iy = np.unravel_index(np.argmax(np.abs(shifted)), shifted.shape)
    i0, j0 = np.unravel_index(np.argmax(np.abs(psi0)), psi0.shape)
    assert (ix - i0, iy - j0) == (5, -3)
