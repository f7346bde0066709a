"""The oracle's FP64 adjoint (qo_adjoint) pinned by FP64 central differences of
the oracle's own loss, h = 1e-6 for the potential: the loss it returns equals
diffraction_loss, and its dL/dV and probe-parameter gradients match directional
finite differences to 1e-6 relative (FP64 throughout), with and without the
transmission band limit.  The GPU gradients are then compared with it in
relative L2 (tests/test_gpu_gradient_accuracy.py)."""
import numpy as np
import pytest

from cases import probe_positions, small_config, smooth_potential


def _case(oracle, bandlimit):
    cfg = small_config(nx=64, ny=64, pad=16, slices=3, dx=0.05, bandlimit=bandlimit)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=9, vmax=20.0).astype(np.float64)
    ipos = probe_positions(cfg, 2, seed=3)
    psi0 = oracle.probe_d(cfg, float(cfg.df0), float(cfg.Cs))
    ex, _, _ = oracle.multislice(cfg, oracle.make_trans_d(cfg, 1.1 * V), ipos, psi0)
    target = np.abs(np.fft.fft2(ex, axes=(1, 2)) / (cfg.nx * cfg.ny)) ** 2
    return cfg, V, ipos, psi0, target


def _L(oracle, cfg, V, ipos, psi0, target):
    return oracle.diffraction_loss(cfg, oracle.make_trans_d(cfg, V), ipos, psi0, target)[0]


@pytest.mark.parametrize("bandlimit", [1, 0])
def test_adjoint_dV_matches_central_differences(oracle, bandlimit):
    cfg, V, ipos, psi0, target = _case(oracle, bandlimit)
    loss, gV, _ = oracle.adjoint(cfg, V, ipos, psi0, target)
    L0 = _L(oracle, cfg, V, ipos, psi0, target)
    assert L0 > 0 and abs(loss / L0 - 1) < 1e-12
    rng = np.random.default_rng(4)
    h = 1e-6
    for trial in range(3):
        D = rng.standard_normal(V.shape)
        if trial == 2:  # a direction on one slice and a small patch: an index or sign error shows
            D = np.zeros(V.shape)
            D[1, 20:30, 25:35] = rng.standard_normal((10, 10))
        fd = (_L(oracle, cfg, V + h * D, ipos, psi0, target) - _L(oracle, cfg, V - h * D, ipos, psi0, target)) / (2 * h)
        ad = float(np.sum(gV * D))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)


def test_adjoint_probe_gradients_match_central_differences(oracle):
    cfg, V, ipos, psi0, target = _case(oracle, 1)
    _, _, gpsi = oracle.adjoint(cfg, V, ipos, psi0, target)
    g_df, g_cs = oracle.probe_param_gradients(cfg, gpsi)
    for name, g, h in (("df0", g_df, 1e-3), ("Cs", g_cs, 1.0)):
        args = {"df0": float(cfg.df0), "Cs": float(cfg.Cs)}
        args[name] += h
        Lp = _L(oracle, cfg, V, ipos, oracle.probe_d(cfg, **args), target)
        args[name] -= 2 * h
        Lm = _L(oracle, cfg, V, ipos, oracle.probe_d(cfg, **args), target)
        fd = (Lp - Lm) / (2 * h)
        assert abs(g - fd) <= 1e-6 * abs(fd), (name, g, fd)
