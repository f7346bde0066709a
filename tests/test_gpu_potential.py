"""GPU potential builder (make3DSlices + getAtomPotential3D) against the oracle."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def perovskite_atoms(nc=(2, 2, 2), a=3.905, dw=(0.6214, 0.4390, 0.7323), jitter=0.0, seed=0):
    """SrTiO3-type cell (stem3/debug_data/SrTiO3.cfg:13-24) replicated, neutral atoms, z-sorted."""
    basis = [(0, 0, 0, dw[0], 38), (0.5, 0.5, 0.5, dw[1], 22), (0, 0.5, 0.5, dw[2], 8),
             (0.5, 0, 0.5, dw[2], 8), (0.5, 0.5, 0, dw[2], 8)]
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nc[0]):
        for j in range(nc[1]):
            for k in range(nc[2]):
                for fx, fy, fz, B, Z in basis:
                    d = rng.normal(0, jitter, 3) if jitter else np.zeros(3)
                    out.append(((i + fx) * a + d[0], (j + fy) * a + d[1], (k + fz) * a + d[2], B, 1.0, 0.0, Z))
    arr = np.array(out)
    return arr[np.argsort(arr[:, 2], kind="stable")]


def _cfg(nx, pad, dx, slices, dz):
    from cases import small_config
    cfg = small_config(nx=nx, ny=nx, pad=pad, slices=slices, dx=dx)
    cfg.sliceThickness = dz
    return cfg


@pytest.mark.parametrize("periodic,pad,slices,dz", [(0, 32, 8, 3.905 / 2), (1, 32, 8, 3.905 / 2),
                                                  (0, 42, 20, 0.5)])
def test_potential_matches_oracle(oracle, periodic, pad, slices, dz):
    """Scatter (periodic) and gather (non-periodic) forms; the last case has a grid (170)
    and a slice count (20) that are not multiples of the gather's 16-pixel tiles and
    16-slice blocks."""
    from qstem_amd import Multislice
    cfg = _cfg(128, pad, 0.05, slices, dz)
    atoms = perovskite_atoms((2, 2, 2), jitter=0.05, seed=3)
    offx, offy = -0.7, -0.4
    with Multislice(cfg) as ms:
        Vg = ms.build_potential(atoms, atomRadius=5.0, potOffsetX=offx, potOffsetY=offy, czOffset=0.3,
                                periodicXY=periodic, return_V=True)
    g = oracle.lut_geometry(5.0, cfg.resolutionX, cfg.sliceThickness)
    keys = sorted(set((int(a[6]), float(np.float32(a[3]))) for a in atoms))
    luts = oracle.atom_luts(g, [k[0] for k in keys], [k[1] for k in keys], cfg.sliceThickness)
    kinds = [keys.index((int(a[6]), float(np.float32(a[3])))) for a in atoms]
    Vo = oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offy, czOffset=0.3,
                            atomRadius=5.0, periodic=periodic)
    assert Vo.max() > 0
    assert rel_l2(Vg, Vo) < 2e-5


def test_potential_then_multislice_runs():
    """Potential -> transmission -> probe -> multislice, all on the GPU, finite and normalised."""
    from qstem_amd import Multislice
    cfg = _cfg(128, 32, 0.05, 8, 3.905 / 2)
    atoms = perovskite_atoms((2, 2, 2))
    with Multislice(cfg) as ms:
        ms.build_potential(atoms, potOffsetX=-0.5, potOffsetY=-0.5)
        ms.make_probe()
        out = ms.run_stem(np.array([[0, 0], [16, 16]]), want_intensity=True)
    assert np.all(np.isfinite(out["det"]))
    assert np.all(out["intensity"] > 0.9) and np.all(out["intensity"] <= 1.0 + 1e-5)


def test_charged_atoms_rejected():
    from qstem_amd import Multislice, QstemError
    cfg = _cfg(64, 16, 0.05, 4, 1.0)
    atoms = perovskite_atoms((1, 1, 1))
    atoms[0, 5] = 2.0
    with Multislice(cfg) as ms:
        with pytest.raises(QstemError):
            ms.build_potential(atoms)
