"""Shared parameter sets and synthetic inputs for the parity tests."""
import numpy as np

from qstem_amd import Detector, StemConfig


def smooth_potential(nslice, nx, ny, seed, vmax=25.0, sigma_px=2.0):
    """Seeded smooth non-negative projected potential [nslice][nx][ny] (V*Angstrom units)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nslice, nx, ny))
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.fftfreq(ny)[None, :]
    g = np.exp(-2 * (np.pi * sigma_px) ** 2 * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(noise, axes=(1, 2)) * g, axes=(1, 2)))
    f -= f.min(axis=(1, 2), keepdims=True)
    f /= f.max(axis=(1, 2), keepdims=True)
    return (vmax * f).astype(np.float32)


def small_config(nx=128, ny=128, pad=32, slices=6, dx=0.05, v0=200.0, bandlimit=1, compat=0, shifted=True):
    dets = [Detector(0.0, 10.0, "BF"), Detector(11.0, 22.0, "ABF"), Detector(70.0, 200.0, "ADF")]
    if shifted:
        dets.append(Detector(20.0, 60.0, "DPC", 3.0, -2.0))
    return StemConfig(
        nx=nx, ny=ny, resolutionX=dx, resolutionY=dx, v0=v0, slices=slices, sliceThickness=1.9525,
        potNx=nx + pad, potNy=ny + pad, by=(ny + pad) * dx * 1.5, bandlimittrans=bandlimit,
        Cs=1.0e4, df0=-50.0, alpha=25.0, detectors=dets, compat=compat,
        aberrations={"a33": 20.0, "phi33": 0.3},
    )


def probe_positions(cfg, n, seed=0):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, cfg.potNx - cfg.nx + 1, n)
    py = rng.integers(0, cfg.potNy - cfg.ny + 1, n)
    return np.stack([px, py], axis=1).astype(np.int32)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
