"""Accuracy of the GPU adjoint's gradients as numbers: relative L2 error of the
FP32 GPU gradients against the oracle's FP64 adjoint (qo_adjoint, itself pinned by
FP64 central differences in tests/test_oracle_adjoint.py).

* dL/dV and dL/d(df0, Cs) on the 4D-STEM adjoint configuration's shape: 256^2 wave,
  dx = 0.05 A, 200 kV, seeded Gaussian-filtered random potentials (sigma 0.5 A, max
  phase 0.3 rad), band-limited transmission, L2 loss on diffraction intensities
  against a perturbed specimen (1.05 V, defocus -45 A); reduced to 16 slices and a
  4x3 scan so the FP64 oracle runs in seconds.
* dL/d(x, y) of 50 atoms (O/Ti/Sr, random positions): the FP64 chain
  dL/dx_a = sum_v dL/dV_v dV_v/dx_a with the oracle's dL/dV and dV/dx_a from FP64
  central differences of the oracle's potential (make_slices_d, h = 1e-6 A).

Measured errors are printed; the bounds sit above them with a margin (the target
for every output is 1e-4 relative L2)."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _random_potential(S, pot, dx, seed, v0=200.0, max_phase=0.3):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((S, pot, pot))
    k = np.fft.fftfreq(pot, d=dx)
    g = np.exp(-2 * (np.pi * 0.5) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    V = np.real(np.fft.ifft2(np.fft.fft2(noise, axes=(1, 2)) * g, axes=(1, 2)))
    sigma = (1 + v0 / 511.0) * 12.3984244 / np.sqrt(v0 * (2 * 510.99906 + v0))
    return (V * (max_phase / sigma) / np.abs(V).max()).astype(np.float32)


def test_gradient_accuracy_4dstem_shape(oracle):
    from qstem_amd import Multislice, StemConfig
    n, dx, S = 256, 0.05, 16
    pot = n + 32
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=S, sliceThickness=2.0,
                     potNx=pot, potNy=pot, bandlimittrans=1, Cs=1.0e4, df0=-50.0, alpha=20.0, electronScale=1.0)
    V = _random_potential(S, pot, dx, seed=7)
    ix, iy = np.meshgrid(np.arange(4) * 10, np.arange(3) * 10, indexing="ij")
    ipos = np.stack([ix.ravel(), iy.ravel()], 1).astype(np.int32)
    # target: the perturbed specimen and defocus, FP64
    ex, _, _ = oracle.multislice(cfg, oracle.make_trans_d(cfg, 1.05 * V.astype(np.float64)), ipos,
                                 oracle.probe_d(cfg, -45.0, 1.0e4))
    target = (np.abs(np.fft.fft2(ex, axes=(1, 2)) / (n * n)) ** 2).astype(np.float32)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        g = ms.adjoint(ipos, target)
    psi0 = oracle.probe_d(cfg, float(np.float32(cfg.df0)), float(np.float32(cfg.Cs)))
    loss, gV, gpsi = oracle.adjoint(cfg, V.astype(np.float64), ipos, psi0, target.astype(np.float64))
    g_df, g_cs = oracle.probe_param_gradients(cfg, gpsi)
    e = {"loss": abs(g["loss"] / loss - 1), "grad_V": rel_l2(g["grad_V"], gV),
         "grad_df0": abs(g["grad_df0"] / g_df - 1), "grad_Cs": abs(g["grad_Cs"] / g_cs - 1)}
    print("relative errors vs FP64 adjoint: " + ", ".join("%s %.2e" % kv for kv in e.items()))
    assert e["loss"] < 1e-4 and e["grad_V"] < 1e-4
    assert e["grad_df0"] < 1e-4 and e["grad_Cs"] < 1e-4


def test_atom_gradient_accuracy(oracle):
    from qstem_amd import Multislice, StemConfig
    n, dx, S, R = 128, 0.05, 6, 2.5
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=S, sliceThickness=2.0,
                     potNx=n + 24, potNy=n + 24, bandlimittrans=1, Cs=2.0e4, df0=-60.0, alpha=20.0,
                     electronScale=1.0)
    rng = np.random.default_rng(21)
    na = 50
    ext = (n + 24) * dx
    xyz = np.column_stack([rng.uniform(1.0, ext - 1.0, na), rng.uniform(1.0, ext - 1.0, na),
                           rng.uniform(0.5, S * 2.0 - 0.5, na)])
    Z = rng.choice([8, 22, 38], na)
    atoms = np.column_stack([xyz, np.full(na, 0.5), np.ones(na), np.zeros(na), Z])
    atoms = atoms[np.argsort(atoms[:, 2], kind="stable")]
    ipos = np.array([[0, 0], [12, 8], [24, 24], [6, 18]], np.int32)
    target = np.zeros((len(ipos), n, n), np.float32)
    with Multislice(cfg) as ms:
        V32 = ms.build_potential(atoms, atomRadius=R, return_V=True)
        ms.make_probe()
        ms.adjoint(ipos, target)
        g_gpu = ms.atom_gradients(atoms)[:, :2]
    gO = oracle.lut_geometry(R, cfg.resolutionX, cfg.sliceThickness)
    keys = sorted(set(int(z) for z in atoms[:, 6]))
    luts = oracle.atom_luts(gO, keys, [0.5] * len(keys), cfg.sliceThickness)
    kinds = [keys.index(int(z)) for z in atoms[:, 6]]
    Zi = atoms[:, 6].astype(int)
    pos0 = atoms[:, :3].astype(np.float32).astype(np.float64)

    def Vd(pos):
        return oracle.make_slices_d(cfg, gO, luts, kinds, pos, Zi, 0.0, 0.0, atomRadius=R)

    V0 = Vd(pos0)
    assert rel_l2(V0, V32) < 2e-5
    psi0 = oracle.probe_d(cfg, float(np.float32(cfg.df0)), float(np.float32(cfg.Cs)))
    _, gV, _ = oracle.adjoint(cfg, V0, ipos, psi0, target.astype(np.float64))
    h = 1e-6
    g_ref = np.zeros((na, 2))
    for ia in range(na):
        for d in (0, 1):
            pp, pm = pos0.copy(), pos0.copy()
            pp[ia, d] += h
            pm[ia, d] -= h
            g_ref[ia, d] = np.sum(gV * (Vd(pp) - Vd(pm))) / (2 * h)
    err = rel_l2(g_gpu, g_ref)
    print("atom gradients (%d atoms): relative L2 error vs FP64 chain %.2e" % (na, err))
    assert err < 1e-4


@pytest.mark.parametrize("n,slices,nprobe,pad", [(512, 4, 3, 32), (1024, 3, 1, 0), (2048, 2, 1, 0)])
def test_gradient_accuracy_large_sizes(oracle, n, slices, nprobe, pad):
    """dL/dV, dL/d(df0, Cs) of the kept column and row sweep shapes at 512^2 (16-column
    tiles, 512-thread blocks), 1024^2 (8-column tiles, {16,8,8} over 64 lanes) and
    2048^2 (block-wide {8,8,8,4}) against the FP64 adjoint: an indexing or
    synchronisation defect of the reverse sweeps would show here even where the forward
    parity tests pass.  Probe incident waves over a smooth strong phase object (the
    plane-wave weak-phase case is FP32-conditioned, as the FD tests note).  At 512^2
    three probes' windows overlap (dL/dt by atomics); at 1024^2 and 2048^2 one window
    covers the whole slice (dL/dt stored by its single writer)."""
    from qstem_amd import Multislice, StemConfig
    dx = 0.025 if n >= 1024 else 0.03
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=slices, sliceThickness=2.0,
                     potNx=n + pad, potNy=n + pad, bandlimittrans=1, Cs=1.0e4, df0=-40.0, alpha=25.0,
                     electronScale=1.0)
    rng = np.random.default_rng(n)
    V = _random_potential(slices, n + pad, dx, seed=n + 1, max_phase=1.5)
    ipos = rng.integers(0, pad + 1, (nprobe, 2)).astype(np.int32)
    ex, _, _ = oracle.multislice(cfg, oracle.make_trans_d(cfg, 0.9 * V.astype(np.float64)), ipos,
                                 oracle.probe_d(cfg, -40.0, 1.0e4))
    target = (np.abs(np.fft.fft2(ex, axes=(1, 2)) / (n * n)) ** 2).astype(np.float32)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        g = ms.adjoint(ipos, target)
    psi0 = oracle.probe_d(cfg, float(np.float32(cfg.df0)), float(np.float32(cfg.Cs)))
    loss, gV, gpsi = oracle.adjoint(cfg, V.astype(np.float64), ipos, psi0, target.astype(np.float64))
    g_df, g_cs = oracle.probe_param_gradients(cfg, gpsi)
    e = {"loss": abs(g["loss"] / loss - 1), "grad_V": rel_l2(g["grad_V"], gV),
         "grad_df0": abs(g["grad_df0"] / g_df - 1), "grad_Cs": abs(g["grad_Cs"] / g_cs - 1)}
    print("%d^2: relative errors vs FP64 adjoint: " % n + ", ".join("%s %.2e" % kv for kv in e.items()))
    assert max(e.values()) < 1e-4
