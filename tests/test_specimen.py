"""Specimen input: .cfg reader and supercell/Einstein model (CPU)."""
import os

import numpy as np

from qstem_amd.specimen import build_supercell, einstein_sigma, perovskite_cell, read_cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_srtio3_cfg():
    c = read_cfg(os.path.join(GOLD, "SrTiO3.cfg"))
    assert np.allclose(c.Mm, np.eye(3) * 3.905)
    assert list(c.Z) == [38, 22, 8, 8, 8]
    assert np.allclose(c.dw, [0.6214, 0.4390, 0.7323, 0.7323, 0.7323])
    assert np.allclose(c.q, [2, 4, -2, -2, -2])
    assert np.allclose(c.frac[1], [0.5, 0.5, 0.5])


def test_supercell_geometry_and_einstein_statistics():
    cell = perovskite_cell()
    atoms, ext = build_supercell(cell, 3, 4, 5)
    assert atoms.shape == (5 * 60, 7) and np.allclose(ext, (3 * 3.905, 4 * 3.905, 5 * 3.905))
    assert np.all(np.diff(atoms[:, 2]) >= 0)
    a1, _ = build_supercell(cell, 10, 10, 10, tds=True, temperature=300.0, seed=4)
    a0, _ = build_supercell(cell, 10, 10, 10)
    # same multiset of lattice sites: match by rounding to the undisplaced site
    sr0 = a0[a0[:, 6] == 38][:, :3]
    sr1 = a1[a1[:, 6] == 38][:, :3]
    d = sr1 - np.round(sr1 / 3.905) * 3.905
    sig = einstein_sigma(0.6214, 300.0)
    assert abs(d.std() / sig - 1) < 0.05 and abs(d.mean()) < 3 * sig / np.sqrt(d.size)
    assert len(sr0) == len(sr1) == 1000
    b1, _ = build_supercell(cell, 2, 2, 2, tds=True, seed=4)
    b2, _ = build_supercell(cell, 2, 2, 2, tds=True, seed=4)
    assert np.array_equal(b1, b2)  # seeded, reproducible


def test_einstein_sigma_matches_sigma_0p08():
    """sigma = 0.08 A per axis at 300 K corresponds to B = 24 pi^2 sigma^2 in the reference's model."""
    assert abs(einstein_sigma(24 * np.pi ** 2 * 0.08 ** 2) - 0.08) < 1e-12
