"""Pins the oracle's multislice loop (stemlib.cpp:2375-2534) with physics invariants."""
import numpy as np

from qstem_amd import Detector, StemConfig


def cfg(**kw):
    base = dict(nx=64, ny=64, resolutionX=0.08, resolutionY=0.08, v0=200.0, slices=5, sliceThickness=2.0,
                potNx=80, potNy=72, alpha=15.0, df0=-80.0, bandlimittrans=0,
                detectors=[Detector(0.0, 1000.0, "all"), Detector(70.0, 200.0, "ADF")])
    base.update(kw)
    return StemConfig(**base)


def test_empty_specimen_equals_defocus_change(oracle):
    """t = 1: S slices of dz propagate the probe by D = S*dz, i.e. probe(df) -> probe(df + D*l_w/l_p)
    (chi's defocus term and the Fresnel propagator share the sign convention)."""
    c = cfg()
    trans = np.ones((c.slices, c.potNx, c.potNy), complex)
    psi0 = oracle.probe(c)
    ex, _, _ = oracle.multislice(c, trans, np.array([[3, 5]]), psi0)
    lw = oracle.wavelength(200.0)
    lp = 12.26 / np.sqrt(200e3 + 200 ** 2 * 0.9788)
    D = c.slices * np.float32(c.sliceThickness)
    c2 = cfg(df0=float(np.float32(c.df0) + D * lw / lp))
    want = oracle.probe(c2)
    assert np.allclose(ex[0], want, atol=1e-6 * np.abs(want).max())


def test_full_detector_is_parseval_total(oracle):
    c = cfg(bandlimittrans=1)
    rng = np.random.default_rng(0)
    V = rng.uniform(0, 20, (c.slices, c.potNx, c.potNy)).astype(np.float32)
    trans = oracle.make_trans(c, V)
    bins, nb = oracle.record_bins(c.slices)
    ex, det, _ = oracle.multislice(c, trans, np.array([[0, 0], [10, 4]]), oracle.probe(c), bins=bins, nbins=nb)
    total = np.sum(np.abs(ex) ** 2, axis=(1, 2)) / (c.nx * c.ny)
    assert np.allclose(det[nb - 1, 0] / np.float32(c.electronScale), total, rtol=1e-12)
    assert np.all(total < 1.0) and np.all(total > 0.5)  # band limit removes some intensity


def test_literal_sizing_adf_is_empty(oracle):
    """Wave grid = lateral supercell (dx = 15.62/128 A at 200 kV): the 2/3 aperture ends at 68.5 mrad,
    so a 70-200 mrad ring receives nothing."""
    c = cfg(nx=128, ny=128, resolutionX=15.62 / 128, resolutionY=15.62 / 128, potNx=160, potNy=160)
    kx2, ky2, k2max = oracle.k_tables(c)
    theta_max = np.sqrt(k2max) * oracle.wavelength(200.0)
    assert abs(theta_max - 0.0685) < 5e-4
    rng = np.random.default_rng(1)
    V = rng.uniform(0, 30, (c.slices, 160, 160)).astype(np.float32)
    bins, nb = oracle.record_bins(c.slices)
    _, det, _ = oracle.multislice(c, oracle.make_trans(c, V), np.array([[7, 9]]), oracle.probe(c), bins=bins,
                                  nbins=nb)
    assert np.all(det[:, 1] == 0.0) and np.all(det[:, 0] > 0)


def test_bins_last_write_wins(oracle):
    """Default interval: bin 0 holds the value after slice S-2, bin 1 after the last slice (Q4)."""
    c = cfg()
    rng = np.random.default_rng(2)
    V = rng.uniform(0, 25, (c.slices, c.potNx, c.potNy)).astype(np.float32)
    trans = oracle.make_trans(c, V)
    psi0 = oracle.probe(c)
    bins, nb = oracle.record_bins(c.slices)
    assert nb == 2 and list(bins) == [-1, -1, -1, 0, 1]
    _, det, _ = oracle.multislice(c, trans, np.array([[1, 2]]), psi0, bins=bins, nbins=nb)
    b4 = np.array([-1, -1, -1, 0])
    _, det4, _ = oracle.multislice(c, trans[:4], np.array([[1, 2]]), psi0, bins=b4, nbins=1)
    assert np.allclose(det[0], det4[0], rtol=1e-12)


def test_transmission_unit_modulus_and_band(oracle):
    c = cfg()
    rng = np.random.default_rng(3)
    V = rng.uniform(0, 40, (2, c.potNx, c.potNy)).astype(np.float32)
    t = oracle.make_trans(c, V)
    assert np.allclose(np.abs(t), 1.0)
    sig = (1 + 200.0 / 511.0) * oracle.wavelength(200.0)  # gamma*lambda (stemlib.cpp:2271-2274)
    assert np.allclose(np.angle(t), np.angle(np.exp(1j * sig * V.astype(float))), atol=1e-6)
    c.bandlimittrans = 1
    tb = oracle.make_trans(c, V)
    T = np.fft.fft2(tb, axes=(1, 2))
    kx = np.fft.fftfreq(c.potNx, d=np.float32(c.resolutionX))
    ky = np.fft.fftfreq(c.potNy, d=np.float32(c.resolutionY))
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    kmax = 2 / 3 * min(1 / (2 * c.resolutionX), 1 / (2 * c.resolutionY))
    assert np.all(np.abs(T[:, k2 > kmax ** 2 * 1.001]) < 1e-9 * np.abs(T).max())
    inner = k2 < kmax ** 2 * 0.999
    assert np.allclose(T[:, inner], np.fft.fft2(t, axes=(1, 2))[:, inner], rtol=1e-9, atol=1e-9)
