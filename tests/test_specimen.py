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


def test_amorphous_carbon_density_and_min_distance():
    from qstem_amd.specimen import amorphous_positions
    box = (12.0, 11.0, 10.0)
    p = amorphous_positions(box, density_gcm3=2.0, rmin=1.4, seed=3)
    n_expected = 2.0 / 12.011 * 6.02214076e23 * 1e-24 * 12 * 11 * 10
    assert len(p) == round(n_expected)
    assert np.all(p >= 0) and np.all(p < np.array(box))
    d = p[:, None, :] - p[None, :, :]
    d[..., :2] -= np.array(box[:2]) * np.round(d[..., :2] / np.array(box[:2]))  # periodic in x, y
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, 99)
    assert r.min() >= 1.4
    # seeded: same seed, same positions
    assert np.array_equal(p, amorphous_positions(box, 2.0, 1.4, seed=3))


def test_gold_sphere_and_specimen_composition():
    from qstem_amd.specimen import carbon_gold_specimen, fcc_sphere, gaussian_phonons
    s = fcc_sphere((0, 0, 0), 10.0, 4.08)
    # FCC nearest-neighbour distance a/sqrt(2); count ~ sphere volume / (a^3/4)
    d = np.sqrt(((s[:, None] - s[None]) ** 2).sum(-1))
    np.fill_diagonal(d, 99)
    assert abs(d.min() - 4.08 / np.sqrt(2)) < 1e-9
    assert abs(len(s) / (4 / 3 * np.pi * 1000 / (4.08 ** 3 / 4)) - 1) < 0.15
    atoms = carbon_gold_specimen(box=(30.0, 30.0, 30.0), sphere_radius=8.0, seed=2)
    au = atoms[atoms[:, 6] == 79]
    c = atoms[atoms[:, 6] == 6]
    assert len(au) == len(fcc_sphere((15, 15, 15), 8.0, 4.08))
    assert np.all(np.sqrt(((c[:, :3] - 15.0) ** 2).sum(-1)) > 8.0)  # carbon removed inside the sphere
    assert np.all(np.diff(atoms[:, 2]) >= 0)
    u = gaussian_phonons(atoms, sigma=0.08, seed=5, sort=False)[:, :3] - atoms[:, :3]
    n = u.size
    assert abs(u.std() - 0.08) < 4 * 0.08 / np.sqrt(2 * n)   # sample std of n normals
    assert abs(u.mean()) < 4 * 0.08 / np.sqrt(n)
    assert np.array_equal(gaussian_phonons(atoms, 0.08, seed=5), gaussian_phonons(atoms, 0.08, seed=5))
    ph = gaussian_phonons(atoms, 0.08, seed=5)
    assert np.all(np.diff(ph[:, 2]) >= 0) and not np.array_equal(ph, gaussian_phonons(atoms, 0.08, seed=6))
