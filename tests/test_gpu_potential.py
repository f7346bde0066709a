"""GPU potential builder (make3DSlices + getAtomPotential3D + getAtomPotentialOffset3D) against the oracle."""
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


def _oracle_V(oracle, cfg, atoms, R, offx, offy, czOffset, periodic, compat_q5=1, norm_fix=0):
    g = oracle.lut_geometry(R, cfg.resolutionX, cfg.sliceThickness)
    keys = sorted(set((int(a[6]), float(np.float32(a[3]))) for a in atoms))
    Zs, Bs = [k[0] for k in keys], [k[1] for k in keys]
    luts = oracle.atom_luts(g, Zs, Bs, cfg.sliceThickness, norm_fix=bool(norm_fix))
    offs = None
    if np.any(atoms[:, 5]):
        offs = oracle.ion_luts(g, Zs, Bs, cfg.sliceThickness, compat_q5=bool(compat_q5), norm_fix=bool(norm_fix),
                               resY=cfg.resolutionY, atomRadius=R)
    kinds = [keys.index((int(a[6]), float(np.float32(a[3])))) for a in atoms]
    return oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offy, czOffset=czOffset,
                              atomRadius=R, periodic=periodic, offs_luts=offs)


@pytest.mark.parametrize("periodic,pad,slices,dz,R", [(0, 32, 8, 3.905 / 2, 5.0), (1, 32, 8, 3.905 / 2, 5.0),
                                                    (0, 42, 20, 0.5, 5.0), (1, 42, 6, 3.905 / 2, 1.5),
                                                    (1, 42, 6, 3.905 / 2, 4.0)])
def test_potential_matches_oracle(oracle, periodic, pad, slices, dz, R):
    """Gather form, non-periodic and periodic boxes.  The third case has a grid (170)
    and a slice count (20) that are not multiples of the gather's 16-pixel tiles and
    16-slice blocks; the last two periodic boxes with smaller radii on that grid and
    atoms within one radius of both edges, so that wrapped pixel boxes split into two
    tile ranges (R = 1.5 A) or, 160 of 170 pixels wide, have both wrapped pieces in
    one tile (R = 4 A, the atom at x = 0.07 A)."""
    from qstem_amd import Multislice
    cfg = _cfg(128, pad, 0.05, slices, dz)
    atoms = perovskite_atoms((2, 2, 2), jitter=0.05, seed=3)
    offx, offy = -0.7, -0.4
    if periodic and R < 5:
        ext = cfg.potNx * cfg.resolutionX
        atoms[:5, 0] = [0.3, ext - 0.4, 0.75, ext - 1.0, 0.07]
        atoms[:5, 1] = [ext - 0.2, 0.5, 1.2, ext - 0.7, 0.07]
        offx = offy = 0.0
    with Multislice(cfg) as ms:
        Vg = ms.build_potential(atoms, atomRadius=R, potOffsetX=offx, potOffsetY=offy, czOffset=0.3,
                                periodicXY=periodic, return_V=True)
    Vo = _oracle_V(oracle, cfg, atoms, R, offx, offy, 0.3, periodic)
    assert Vo.max() > 0
    assert rel_l2(Vg, Vo) < 2e-5


@pytest.mark.parametrize("periodic,compat_q5,norm_fix", [(0, 1, 0), (0, 0, 0), (1, 1, 1), (0, 0, 1)])
def test_charged_potential_matches_oracle(oracle, periodic, compat_q5, norm_fix):
    """Ion offsets (getAtomPotentialOffset3D) with the charges of the reference's
    SrTiO3.cfg (Sr +2, Ti +4, O -2), with and without quirk Q5, and the LUT
    normalisation switch."""
    from qstem_amd import Multislice
    cfg = _cfg(128, 32, 0.05, 8, 3.905 / 2)
    atoms = perovskite_atoms((2, 2, 2), jitter=0.05, seed=4)
    atoms[:, 5] = np.where(atoms[:, 6] == 38, 2.0, np.where(atoms[:, 6] == 22, 4.0, -2.0))
    with Multislice(cfg) as ms:
        Vg = ms.build_potential(atoms, atomRadius=5.0, potOffsetX=-0.7, potOffsetY=-0.4, czOffset=0.3,
                                periodicXY=periodic, compat_q5=compat_q5, lut_norm_fix=norm_fix, return_V=True)
    Vo = _oracle_V(oracle, cfg, atoms, 5.0, -0.7, -0.4, 0.3, periodic, compat_q5=compat_q5, norm_fix=norm_fix)
    neutral = atoms.copy()
    neutral[:, 5] = 0
    Vn = _oracle_V(oracle, cfg, neutral, 5.0, -0.7, -0.4, 0.3, periodic, norm_fix=norm_fix)
    assert rel_l2(Vo, Vn) > 0.05  # the charges change the potential
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


def test_charged_atoms_need_the_ion_table():
    from qstem_amd import Multislice, QstemError
    cfg = _cfg(64, 16, 0.05, 4, 1.0)
    atoms = perovskite_atoms((1, 1, 1))
    atoms[0, 5] = 2.0
    with Multislice(cfg) as ms:
        ms._have_ion_table = True  # skip the wrapper's automatic qs_set_ion_scattering_table
        with pytest.raises(QstemError, match="ion"):
            ms.build_potential(atoms)
