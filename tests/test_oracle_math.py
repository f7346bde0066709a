"""Pins the oracle's numerics against independent mathematics (CPU only)."""
import ctypes

import numpy as np
import pytest


def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 16, 17, 31, 64])
def test_fft1d_brute_force(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        assert np.allclose(oracle.fft1(x, sign), brute_dft(x, sign), atol=1e-10 * max(n, 1))


def test_fft2d_conventions(oracle):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((24, 40)) + 1j * rng.standard_normal((24, 40))
    f = oracle.fft2(a, -1)
    assert np.allclose(f, np.fft.fft2(a), atol=1e-9)  # FFTW_FORWARD = numpy forward, unnormalised
    assert np.allclose(oracle.fft2(f, +1) / a.size, a, atol=1e-12)  # backward unnormalised
    assert np.isclose(np.sum(np.abs(f) ** 2), a.size * np.sum(np.abs(a) ** 2))  # Parseval
    imp = np.zeros((16, 16), complex)
    imp[3, 5] = 1
    k = np.arange(16)
    assert np.allclose(oracle.fft2(imp, -1), np.exp(-2j * np.pi * (3 * k[:, None] + 5 * k[None, :]) / 16))


def test_wavelength_closed_form(oracle):
    # lambda = h c / sqrt(E (2 m c^2 + E)), stem3/stemutil.cpp:714-725
    assert abs(oracle.wavelength(200.0) - 0.0250793474) < 1e-9
    assert abs(oracle.wavelength(300.0) - 0.0196874946) < 1e-9
    hc, mc2 = 12.398419843320026, 510.99895  # CODATA 2018 (keV*A, keV)
    for e in (80.0, 200.0, 300.0):
        assert abs(oracle.wavelength(e) / (hc / np.sqrt(e * (2 * mc2 + e))) - 1) < 1e-5


def _spline(oracle, x, y):
    lib = oracle.load()
    n = len(x)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    P = ctypes.POINTER(ctypes.c_double)
    x = np.ascontiguousarray(x, float)
    y = np.ascontiguousarray(y, float)
    lib.qo_splinh(x.ctypes.data_as(P), y.ctypes.data_as(P), b.ctypes.data_as(P), c.ctypes.data_as(P),
                  d.ctypes.data_as(P), n)
    lib.qo_seval.argtypes = [P] * 5 + [ctypes.c_int, ctypes.c_double]
    return lambda t: lib.qo_seval(x.ctypes.data_as(P), y.ctypes.data_as(P), b.ctypes.data_as(P),
                                  c.ctypes.data_as(P), d.ctypes.data_as(P), n, float(t))


def test_akima_spline_properties(oracle):
    x = np.linspace(0, 6, 25) ** 1.3
    ev = _spline(oracle, x, 3 * x - 1)  # Akima reproduces straight lines exactly
    for t in np.linspace(0, x[-2], 40):
        assert abs(ev(t) - (3 * t - 1)) < 1e-9
    y = np.exp(-x)
    ev = _spline(oracle, x, y)
    for xi, yi in zip(x[:-1], y[:-1]):  # interpolates the nodes
        assert abs(ev(xi) - yi) < 1e-12
    t = np.linspace(0.05, x[-3], 57)
    assert np.max(np.abs([ev(v) for v in t] - np.exp(-t))) < 5e-3  # smooth data, coarse grid


def test_rez_table_sanity():
    """f_e(0) of the Rez table follows Mott-Bethe ordering and the grid is increasing."""
    from qstem_amd.multislice import rez_table
    t = rez_table()
    assert np.all(np.diff(t[0, :120]) > 0)
    f0 = t[1:, 0]
    assert f0[7] < f0[21] < f0[37]  # O < Ti < Sr
    assert np.all(t[1:, 120:] == 0)
