"""Specimen input for the potential builder: .cfg reader, supercell
replication and Einstein frozen-phonon displacements.

Restates (for the NCELL path, orthogonal or not, zero crystal tilt):
  readCFGCellParams  libs/fileio_fftw3.cpp:721-779
  readNextCFGAtom    libs/fileio_fftw3.cpp:908-989
  replicateUnitCell  libs/fileio_fftw3.cpp:1188-1306  (without vacancy/occupancy sampling)
  readUnitCell       libs/fileio_fftw3.cpp:1313-1657  (fractional -> Cartesian, box origin)
  phononDisplacement libs/fileio_fftw3.cpp:586-614    (Einstein model)
The reference seeds its Gaussian generator from the clock (fileio_fftw3.cpp:496);
here displacements come from a seeded numpy Generator so runs are reproducible.
"""
import re
from dataclasses import dataclass

import numpy as np

# element symbols in the order of fileio_fftw3.cpp:72-79 (Z = index + 1)
ELEMENTS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
            "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm "
            "Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No "
            "Lr").split()


def z_number(symbol: str) -> int:
    s = symbol.strip()
    s = s[:2] if len(s) > 1 and s[1].islower() else s[:1]
    return ELEMENTS.index(s) + 1


@dataclass
class UnitCell:
    Mm: np.ndarray          # 3x3 cell vectors (rows), Angstrom
    frac: np.ndarray        # (n, 3) fractional positions
    dw: np.ndarray          # (n,) Debye-Waller B
    occ: np.ndarray
    q: np.ndarray
    Z: np.ndarray


def read_cfg(path: str) -> UnitCell:
    """Extended-CFG reader: 'Number of particles =', 'A =', 'H0(i,j) =', '.NO_VELOCITY.',
    'entry_count =' then blocks of mass / symbol / data lines (fileio_fftw3.cpp:721-989)."""
    text = open(path).read().splitlines()

    def param(title, default=None):
        for line in text:
            line = line.split("#")[0]
            if title in line:
                return line.split(title, 1)[1].split()[0]
        return default

    ncoord = int(param("Number of particles =", "0"))
    scale = float(param("A =", "1.0"))
    Mm = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            v = param("H0(%d,%d) =" % (i + 1, j + 1))
            if v is not None:
                Mm[i, j] = float(v)
    Mm *= scale
    novel = any(".NO_VELOCITY." in l for l in text)
    entry = int(param("entry_count =", "3"))
    if not novel:
        entry += 3
    off = 3 * (0 if novel else 1)
    start = 0
    for i, l in enumerate(text):
        if "entry_count" in l:
            start = i + 1
    frac, dw, occ, q, Z = [], [], [], [], []
    mass, element = 28.0, 1
    i = start
    while i < len(text) and len(frac) < ncoord:
        line = text[i].split("#")[0].strip()
        i += 1
        if not line:
            continue
        toks = line.split()
        if len(toks) == 1 and re.match(r"^[0-9.]+$", toks[0]) and float(toks[0]) >= 1.0:
            mass = float(toks[0])
            element = z_number(text[i].strip())
            i += 1
            continue
        vals = [float(t) for t in toks[:entry]]
        frac.append(vals[:3])
        dw.append(vals[3 + off] if entry > 3 + off else 0.45 * 28.0 / mass)
        occ.append(vals[4 + off] if entry > 4 + off else 1.0)
        q.append(vals[5 + off] if entry > 5 + off else 0.0)
        Z.append(element)
    return UnitCell(Mm, np.array(frac, float), np.array(dw, float), np.array(occ, float), np.array(q, float),
                    np.array(Z, int))


def einstein_sigma(dw, temperature=300.0):
    """Per-axis displacement sigma of phononDisplacement's Einstein branch:
    sqrt(T/300) * sqrt(B/(8 pi^2)) / sqrt(3) (fileio_fftw3.cpp:495-498, 590-594)."""
    return np.sqrt(temperature / 300.0) * np.sqrt(np.asarray(dw) / (8 * np.pi ** 2)) / np.sqrt(3.0)


def build_supercell(cell: UnitCell, ncx: int, ncy: int, ncz: int, tds: bool = False, temperature: float = 300.0,
                    seed: int = 0, xOffset: float = 0.0, yOffset: float = 0.0):
    """Replicated model in Cartesian Angstrom as (n, 7) [x, y, z, dw, occ, q, Z], z-sorted, plus the
    supercell extents (ax, by, c).  Box origin at the minimum corner as readUnitCell places it."""
    n = len(cell.Z)
    ic = np.stack(np.meshgrid(np.arange(ncx), np.arange(ncy), np.arange(ncz), indexing="ij"), -1).reshape(-1, 3)
    frac = (cell.frac[None, :, :] + ic[:, None, :]).reshape(-1, 3)
    rep = lambda a: np.tile(a, len(ic))
    dw, occ, q, Z = rep(cell.dw), rep(cell.occ), rep(cell.q), rep(cell.Z)
    if tds:
        rng = np.random.default_rng(seed)
        u = rng.standard_normal((len(frac), 3)) * einstein_sigma(dw, temperature)[:, None]
        frac = frac + u @ np.linalg.inv(cell.Mm)  # displacement converted to fractional (MmInv)
    xyz = frac @ cell.Mm
    corners = np.array([[i, j, k] for i in (0, ncx) for j in (0, ncy) for k in (0, ncz)], float) @ cell.Mm
    lo, hi = corners.min(0), corners.max(0)
    xyz = xyz - lo
    xyz[:, 0] += xOffset
    xyz[:, 1] += yOffset
    atoms = np.column_stack([xyz, dw, occ, q, Z])
    atoms = atoms[np.argsort(atoms[:, 2], kind="stable")]
    return atoms, tuple(hi - lo)


def perovskite_cell(a=3.905, dw=(0.6214, 0.4390, 0.7323), Zs=(38, 22, 8), charges=(0.0, 0.0, 0.0)) -> UnitCell:
    """ABO3 cubic perovskite with the basis of stem3/debug_data/SrTiO3.cfg:13-24."""
    frac = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], float)
    return UnitCell(np.eye(3) * a, frac, np.array([dw[0], dw[1], dw[2], dw[2], dw[2]]), np.ones(5),
                    np.array([charges[0], charges[1], charges[2], charges[2], charges[2]]),
                    np.array([Zs[0], Zs[1], Zs[2], Zs[2], Zs[2]]))


# ---------------------------------------------------------------------------------
# Amorphous-carbon slab with an embedded gold nanoparticle, and Gaussian frozen
# phonons.  The reference builds amorphous models with its separate gbmaker tool
# (random insertion with a minimum distance, gbmaker.cpp:1090-1134); this is an
# independent seeded generator of the same kind of specimen.

CARBON_MASS = 12.011        # g/mol
AVOGADRO = 6.02214076e23


def amorphous_positions(box, density_gcm3=2.0, rmin=1.4, mass=CARBON_MASS, seed=1, max_tries=200):
    """Random sequential insertion of carbon atoms into an orthogonal box
    (Angstrom) at the given mass density, no two closer than rmin (periodic in x
    and y, open in z).  A cell list of side >= rmin keeps each test local.
    Returns an (n, 3) array; n = round(density * volume / mass)."""
    box = np.asarray(box, dtype=np.float64)
    n = int(round(density_gcm3 / mass * AVOGADRO * 1e-24 * float(np.prod(box))))
    rng = np.random.default_rng(seed)
    ncell = np.maximum((box // rmin).astype(int), 1)
    side = box / ncell
    cells = {}
    pos = np.empty((n, 3))
    r2 = rmin * rmin
    k = 0
    while k < n:
        for _ in range(max_tries):
            p = rng.random(3) * box
            c = (p // side).astype(int)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        cz = c[2] + dz
                        if cz < 0 or cz >= ncell[2]:
                            continue
                        key = ((c[0] + dx) % ncell[0], (c[1] + dy) % ncell[1], cz)
                        for j in cells.get(key, ()):
                            d = pos[j] - p
                            d[:2] -= box[:2] * np.round(d[:2] / box[:2])
                            if d @ d < r2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[k] = p
                cells.setdefault(tuple(c), []).append(k)
                k += 1
                break
        else:
            raise RuntimeError("amorphous_positions: no free site after %d tries (density too high)" % max_tries)
    return pos


def fcc_sphere(center, radius, a=4.08):
    """Sites of an FCC lattice (cube edge a, one site at the centre) within radius."""
    m = int(np.ceil(radius / a)) + 1
    g = np.arange(-m, m + 1)
    corners = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    sites = (corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    keep = np.einsum("ij,ij->i", sites, sites) <= radius * radius
    return sites[keep] + np.asarray(center, dtype=np.float64)


def carbon_gold_specimen(box=(80.0, 80.0, 100.0), density_gcm3=2.0, rmin=1.4, sphere_radius=20.0, a_au=4.08,
                         seed=1, dw_c=0.5, dw_au=0.6):
    """Amorphous carbon slab with a gold FCC sphere at its centre (carbon removed
    inside the sphere).  Returns (n, 7) atoms x, y, z, dw, occ, q, Z for
    qs_build_potential, sorted by z."""
    box = np.asarray(box, dtype=np.float64)
    c = amorphous_positions(box, density_gcm3, rmin, seed=seed)
    centre = box / 2
    d = c - centre
    c = c[np.einsum("ij,ij->i", d, d) > sphere_radius * sphere_radius]
    au = fcc_sphere(centre, sphere_radius, a_au)
    atoms = np.zeros((len(c) + len(au), 7))
    atoms[:len(c), :3] = c
    atoms[:len(c), 3] = dw_c
    atoms[:len(c), 6] = 6
    atoms[len(c):, :3] = au
    atoms[len(c):, 3] = dw_au
    atoms[len(c):, 6] = 79
    atoms[:, 4] = 1.0
    return atoms[np.argsort(atoms[:, 2], kind="stable")]


def gaussian_phonons(atoms, sigma=0.08, seed=1, sort=True):
    """One frozen-phonon configuration: every atom displaced by N(0, sigma^2) per
    axis (Angstrom), from a seeded generator; rows re-sorted by z unless sort is False."""
    rng = np.random.default_rng(seed)
    out = np.array(atoms, dtype=np.float64, copy=True)
    out[:, :3] += rng.normal(0.0, sigma, size=(len(out), 3))
    return out[np.argsort(out[:, 2], kind="stable")] if sort else out
