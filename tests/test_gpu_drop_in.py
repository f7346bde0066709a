"""The stem3-compatible driver (qstem_amd/bin/qstem3) end to end against the FP64 oracle.

qstem3 reads tests/golden/stem_small.qsc (STO 4x4x3 cells, 128x128, 6 slices,
outputs every 2 slices, ADF/BF/shifted detectors), builds potential, probe and
all scan positions on the GPU and writes .img files.  The oracle recomputes the
same run from the same atoms in FP64 (LUT, scatter, transmission, probe,
multislice) and every STEM image must agree to 1e-4 relative L2.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_qstem3_images_match_oracle(oracle, tmp_path):
    from qstem_amd import Detector, StemConfig
    from qstem_amd.imgio import read_img
    from qstem_amd.specimen import build_supercell, read_cfg
    for f in ("stem_small.qsc", "STO_neutral.cfg"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    exe = os.path.join(ROOT, "qstem_amd", "bin", "qstem3")
    r = subprocess.run([exe, "stem_small.qsc"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out_small"
    names = sorted(os.listdir(out))
    assert names == sorted(["%s_%d.img" % (d, t) for d in ("ADF", "BF", "DPC") for t in range(3)] +
                           ["ADF.img", "BF.img", "DPC.img"]), names

    # the same run in the oracle
    f = np.float32
    dx = f(0.05)
    cell = read_cfg(os.path.join(GOLD, "STO_neutral.cfg"))
    atoms, ext = build_supercell(cell, 4, 4, 3)
    x0, x1 = f(5.8575), f(9.7625)
    npx = int(f(x1 - x0) / dx)
    stop = f(x0 + dx * npx)
    pot = 128 + npx
    dets = [Detector(70.0, 200.0, "ADF"), Detector(0.0, 10.0, "BF"), Detector(20.0, 60.0, "DPC", 2.0, -1.0)]
    wl = oracle.wavelength(200.0)
    cfg = StemConfig(nx=128, ny=128, resolutionX=0.05, resolutionY=0.05, v0=200.0, slices=6, sliceThickness=1.9525,
                     potNx=pot, potNy=pot, by=float(f(ext[1])), bandlimittrans=1, Cs=float(f(f(0.001) * 1e7)),
                     df0=float(f(10.0 * f(-5.0))), alpha=20.0, detectors=dets, electronScale=1.0, compat=3)
    g = oracle.lut_geometry(3.0, cfg.resolutionX, cfg.sliceThickness)
    keys = []
    for a in atoms:  # LUT per Z, B of the lowest atom (compat Q6, as the driver uses)
        if int(a[6]) not in [k[0] for k in keys]:
            keys.append((int(a[6]), float(f(a[3]))))
    luts = oracle.atom_luts(g, [k[0] for k in keys], [k[1] for k in keys], cfg.sliceThickness)
    kinds = [[k[0] for k in keys].index(int(a[6])) for a in atoms]
    offx = float(f(x0 - 0.5 * 128 * dx))
    V = oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offx, atomRadius=3.0)
    trans = oracle.make_trans(cfg, V)
    from qstem_amd import scan_offsets
    ipos = scan_offsets(x0, stop, 4, x0, stop, 3, dx, dx, pot, pot, 128, 128)
    bins, nb = oracle.record_bins(6, outputInterval=2)
    _, det, _ = oracle.multislice(cfg, trans, ipos, oracle.probe(cfg), bins=bins, nbins=nb)
    for ti in range(nb):
        for di, d in enumerate(dets):
            name = "%s_%d.img" % (d.name, ti) if ti < nb - 1 else "%s.img" % d.name
            img = read_img(str(out / name))
            assert img.data.shape == (4, 3) and img.comment == "STEM image"
            want = det[ti, di].reshape(4, 3)
            assert rel_l2(img.data, want) < 1e-4, (name, rel_l2(img.data, want))
            t_expect = (ti + 1) * 2 * f(1.9525) if ti < nb - 1 else 6 * f(1.9525)
            assert abs(img.t - t_expect) < 1e-4
            assert img.params[0] == 1.0 and len(img.params) == 2 + 12
            assert np.allclose(img.params[2:], (img.data.astype(np.float32) ** 2).ravel(), rtol=1e-6)
