"""qstem3 (the stem3-compatible driver) in STEM, TEM and CBED mode and with
"sequence: r p" lines, against the FP64 oracle.

The STEM/TEM/CBED runs use the reference's own stem3/debug_data/stem_STO_4x4.qsc
(tests/golden/stem_STO_4x4.qsc: charged SrTiO3 9x9x20 cells cut to its Cube, 400x400
wave at 0.0625 A, 40 slices of 1.9525 A with outputs every 8, four annular
detectors, no band limit of the transmission) with three keys changed so that the
FP64 oracle finishes in seconds and the run is deterministic: "tds: no" (the
reference's phonon stream is time-seeded) and a 2x2 scan over 1 A.  The oracle
recomputes each run from the same atoms: neutral and ion-offset LUTs (Q5), the
potential, transmission, probe or plane wave, and the multislice in FP64, run in
pieces up to each output slice where the mode writes a file there.

Bin and file indices follow the reference's bookkeeping: with 40 slices and an
interval of 8, collectIntensity's bins are t = slice/8 for slices 0..38 and t = 5
for slice 39, so CBED's diff_4.img holds slice 38 (Q4); interimWave writes wave_t.img
at slices 7, 15, 23, 31 and 39, the last as wave_4.img.
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
EXE = os.path.join(ROOT, "qstem_amd", "bin", "qstem3")
f = np.float32
TOL = 1e-4


def _write_qsc(src, dst, **keys):
    """Copy a .qsc file with the lines of the given keys replaced (key=value -> "key: value")."""
    repl = dict(keys)
    out, seen = [], set()
    for line in open(src).read().splitlines():
        for k, v in repl.items():
            if line.startswith(k + ":"):
                line = "%s: %s" % (k, v)
                seen.add(k)
        out.append(line)
    out += ["%s: %s" % (k, v) for k, v in repl.items() if k not in seen]
    open(dst, "w").write("\n".join(out) + "\n")


def _run_qstem3(tmp_path, qsc_src, cfg_name, **keys):
    shutil.copy(os.path.join(GOLD, cfg_name), tmp_path / cfg_name)
    _write_qsc(os.path.join(GOLD, qsc_src), tmp_path / "run.qsc", **keys)
    print("qstem3", qsc_src, keys, flush=True)
    r = subprocess.run([EXE, "run.qsc"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


# -- the reduced stem_STO_4x4.qsc ------------------------------------------------------
NX, DX, SLICES, DZ, R = 400, 0.0625, 40, 1.9525, 5.0
X0, Y0, X1, Y1 = 11.67, 11.45, 12.67, 12.45
STO_KEYS = dict(tds="no", scan_x_stop=X1, scan_x_pixels=2, scan_y_stop=Y1, scan_y_pixels=2)


def _sto_oracle(oracle, potN, offx, offy, detectors=()):
    """Config, transmission functions [40][potN][potN] of the driver's potential window."""
    from qstem_amd import StemConfig
    from qstem_amd.specimen import build_supercell, read_cfg
    atoms, ext = build_supercell(read_cfg(os.path.join(GOLD, "SrTiO3.cfg")), 9, 9, 20)
    cfg = StemConfig(nx=NX, ny=NX, resolutionX=DX, resolutionY=DX, v0=200.0, slices=SLICES, sliceThickness=DZ,
                     potNx=potN, potNy=potN, by=float(f(35.145)), bandlimittrans=0, Cs=float(f(f(0.05) * 1e7)),
                     df0=float(f(10.0 * f(-13.7))), alpha=15.0, ismoth=1, Cc=1e7, dE_E=0.0, electronScale=1.0,
                     compat=3, detectors=list(detectors))
    g = oracle.lut_geometry(R, cfg.resolutionX, cfg.sliceThickness)
    keys = sorted(set((int(a[6]), float(f(a[3]))) for a in atoms))
    Zs, Bs = [k[0] for k in keys], [k[1] for k in keys]
    luts = oracle.atom_luts(g, Zs, Bs, cfg.sliceThickness)
    offs = oracle.ion_luts(g, Zs, Bs, cfg.sliceThickness, compat_q5=True, resY=cfg.resolutionY, atomRadius=R)
    kinds = [keys.index((int(a[6]), float(f(a[3])))) for a in atoms]
    V = oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offy, czOffset=0.976,
                           atomRadius=R, offs_luts=offs)
    return cfg, oracle.make_trans(cfg, V)


def _pieces(oracle, cfg, trans, psi0, ends):
    """Oracle waves and diffraction patterns after slices ends[i] - 1, run piece by piece."""
    psi, a, out = psi0, 0, []
    for b in ends:
        psi, _, diff = oracle.multislice(cfg, trans[a:b], np.zeros((1, 2), np.int32), psi, want_diff=True)
        psi = psi[0]
        out.append((psi.copy(), diff[0].copy()))
        a = b
    return out


def test_qstem3_stem_sto_4x4(oracle, tmp_path):
    from qstem_amd import Detector, scan_offsets
    from qstem_amd.imgio import read_img
    _run_qstem3(tmp_path, "stem_STO_4x4.qsc", "SrTiO3.cfg", **STO_KEYS)
    out = tmp_path / "STO_4x4"
    npx = int((f(X1) - f(X0)) / f(DX))
    npy = int((f(Y1) - f(Y0)) / f(DX))
    assert npx == npy
    sx1, sy1 = f(f(X0) + f(DX) * npx), f(f(Y0) + f(DX) * npy)
    offx = float(f(np.float64(f(X0)) - 0.5 * NX * np.float64(f(DX))))
    offy = float(f(np.float64(f(Y0)) - 0.5 * NX * np.float64(f(DX))))
    dets = [Detector(70.0, 133.0, "detector1"), Detector(50.0, 70.0, "detector2"), Detector(30.0, 50.0, "detector3"),
            Detector(0.0, 40.0, "detector4")]
    cfg, trans = _sto_oracle(oracle, NX + npx, offx, offy, dets)
    ipos = scan_offsets(X0, sx1, 2, Y0, sy1, 2, DX, DX, cfg.potNx, cfg.potNy, NX, NX)
    bins, nb = oracle.record_bins(SLICES, outputInterval=8)
    assert nb == 6
    _, det, _ = oracle.multislice(cfg, trans, ipos, oracle.probe(cfg), bins=bins, nbins=nb)
    worst = 0.0
    for t in range(nb):
        for di, d in enumerate(dets):
            name = "%s_%d.img" % (d.name, t) if t < nb - 1 else "%s.img" % d.name
            img = read_img(str(out / name))
            assert img.data.shape == (2, 2) and img.comment == "STEM image"
            e = rel_l2(img.data, det[t, di].reshape(2, 2))
            worst = max(worst, e)
            assert e < TOL, (name, e)
            t_expect = (t + 1) * 8 * f(DZ) if t < nb - 1 else SLICES * f(DZ)
            assert abs(img.t - t_expect) < 1e-4
    print("STO 4x4 STEM: 24 detector images, worst relative L2 vs FP64 oracle %.2e" % worst)


def test_qstem3_cbed_sto_4x4(oracle, tmp_path):
    """CBED: probe at (scan_x_start, scan_y_start), diff_t.img at the last slice of each
    bin, diffAvg_1.img (the final pattern, written as diff.img and renamed)."""
    from qstem_amd.imgio import read_img
    _run_qstem3(tmp_path, "stem_STO_4x4.qsc", "SrTiO3.cfg", mode="CBED", **STO_KEYS)
    out = tmp_path / "STO_4x4"
    offx = f(np.float64(f(X0)) - 0.5 * NX * np.float64(f(DX)))
    offy = f(np.float64(f(Y0)) - 0.5 * NX * np.float64(f(DX)))
    cfg, trans = _sto_oracle(oracle, NX, float(offx), float(offy))
    psi0 = oracle.probe(cfg, at=(float(f(X0) - offx), float(f(Y0) - offy)))
    ends = [8, 16, 24, 32, 39, 40]  # diff_t.img: slices 7, 15, 23, 31, 38 (Q4), 39
    res = _pieces(oracle, cfg, trans, psi0, ends)
    worst = 0.0
    for t, (b, (_, diff)) in enumerate(zip(ends, res)):
        img = read_img(str(out / ("diff_%d.img" % t)))
        assert img.comment == "Diffraction Pattern" and img.data.shape == (NX, NX)
        assert abs(img.t - (b - 1) * f(DZ)) < 1e-3  # the wave's thickness before that slice
        assert abs(img.dx - 1.0 / (NX * f(DX))) < 1e-9
        e = rel_l2(img.data, diff)
        worst = max(worst, e)
        assert e < TOL, (t, e)
    avg = read_img(str(out / "diffAvg_1.img"))
    assert avg.comment == "DEBUG Wave intensity"
    assert rel_l2(avg.data, res[-1][1]) < TOL
    assert not os.path.exists(out / "diff_6.img")
    print("STO 4x4 CBED: 6 diffraction patterns, worst relative L2 vs FP64 oracle %.2e" % worst)


def test_qstem3_tem_sto_4x4(oracle, tmp_path):
    """TEM: plane wave on the window centred on the model, wave_t.img at slices 7, 15,
    23, 31 and 39, the exit wave wave.img, and after the slab loop diffAvg_1.img and
    waveIntensity.img = (nx ny)^2 |psi|^2."""
    from qstem_amd.imgio import read_img
    _run_qstem3(tmp_path, "stem_STO_4x4.qsc", "SrTiO3.cfg", mode="TEM", **STO_KEYS)
    out = tmp_path / "STO_4x4"
    cx = f(f(35.145) / f(2.0))  # scan_x_start defaults to ax/2 in TEM mode
    off = f(np.float64(cx) - 0.5 * np.float64(f(NX * f(DX))))
    cfg, trans = _sto_oracle(oracle, NX, float(off), float(off))
    ends = [8, 16, 24, 32, 40]
    res = _pieces(oracle, cfg, trans, np.ones((NX, NX), np.complex128), ends)
    worst = 0.0
    for t, (b, (psi, _)) in enumerate(zip(ends, res)):
        img = read_img(str(out / ("wave_%d.img" % t)))
        assert img.comment == "Wave Function" and img.data.dtype == np.complex64
        assert abs(img.t - b * f(DZ)) < 1e-3 and len(img.params) == 9
        e = rel_l2(img.data, psi)
        worst = max(worst, e)
        assert e < TOL, (t, e)
    assert not os.path.exists(out / "wave_5.img")
    exitw = read_img(str(out / "wave.img"))
    assert exitw.comment == "Exit face wave function for no TDS"
    assert rel_l2(exitw.data, res[-1][0]) < TOL
    assert rel_l2(read_img(str(out / "diffAvg_1.img")).data, res[-1][1]) < TOL
    inten = read_img(str(out / "waveIntensity.img")).data
    assert rel_l2(inten, float(NX * NX) ** 2 * np.abs(res[-1][0]) ** 2) < TOL
    print("STO 4x4 TEM: 5 waves, worst relative L2 vs FP64 oracle %.2e" % worst)


def test_qstem3_sequence_repeats(oracle, tmp_path):
    """stem_small.qsc with two sequence lines, "2 1" then "1 2": the first runs the slab
    twice (runMulsSTEM's mRepeat loop: bins 0..2 are recorded during the second pass,
    at slice indices 2s), the second restarts from the probe and runs two slabs whose
    slice indices (6..17) all fall in the last bin, which it overwrites."""
    from qstem_amd import Detector, StemConfig, scan_offsets
    from qstem_amd.imgio import read_img
    from qstem_amd.specimen import build_supercell, read_cfg
    shutil.copy(os.path.join(GOLD, "STO_neutral.cfg"), tmp_path / "STO_neutral.cfg")
    text = open(os.path.join(GOLD, "stem_small.qsc")).read().replace("sequence: 1 1", "sequence: 2 1\nsequence: 1 2")
    (tmp_path / "run.qsc").write_text(text)
    r = subprocess.run([EXE, "run.qsc"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out_small"
    dx = f(0.05)
    atoms, ext = build_supercell(read_cfg(os.path.join(GOLD, "STO_neutral.cfg")), 4, 4, 3)
    x0, x1 = f(5.8575), f(9.7625)
    npx = int(f(x1 - x0) / dx)
    stop = f(x0 + dx * npx)
    pot = 128 + npx
    dets = [Detector(70.0, 200.0, "ADF"), Detector(0.0, 10.0, "BF"), Detector(20.0, 60.0, "DPC", 2.0, -1.0)]
    cfg = StemConfig(nx=128, ny=128, resolutionX=0.05, resolutionY=0.05, v0=200.0, slices=6, sliceThickness=1.9525,
                     potNx=pot, potNy=pot, by=float(f(ext[1])), bandlimittrans=1, Cs=float(f(f(0.001) * 1e7)),
                     df0=float(f(10.0 * f(-5.0))), alpha=20.0, detectors=dets, electronScale=1.0, compat=3)
    g = oracle.lut_geometry(3.0, cfg.resolutionX, cfg.sliceThickness)
    keys = []
    for a in atoms:
        if int(a[6]) not in [k[0] for k in keys]:
            keys.append((int(a[6]), float(f(a[3]))))
    luts = oracle.atom_luts(g, [k[0] for k in keys], [k[1] for k in keys], cfg.sliceThickness)
    kinds = [[k[0] for k in keys].index(int(a[6])) for a in atoms]
    offx = float(f(x0 - 0.5 * 128 * dx))
    V = oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offx, atomRadius=3.0)
    trans = oracle.make_trans(cfg, V)
    trans2 = np.concatenate([trans, trans])
    ipos = scan_offsets(x0, stop, 4, x0, stop, 3, dx, dx, pot, pot, 128, 128)
    b1, nb = oracle.record_bins(6, outputInterval=2, mulsRepeat1=2)
    assert nb == 4 and list(b1) == [-1, -1, -1, -1, -1, -1, 0, 1, 2, -1, -1, 3]
    _, det1, _ = oracle.multislice(cfg, trans2, ipos, oracle.probe(cfg), bins=b1, nbins=nb)
    b2 = np.full(12, -1, np.int32)
    b2[11] = 3
    _, det2, _ = oracle.multislice(cfg, trans2, ipos, oracle.probe(cfg), bins=b2, nbins=nb)
    want = det1.copy()
    want[3] = det2[3]
    for t in range(nb):
        for di, d in enumerate(dets):
            name = "%s_%d.img" % (d.name, t) if t < nb - 1 else "%s.img" % d.name
            img = read_img(str(out / name))
            e = rel_l2(img.data, want[t, di].reshape(4, 3))
            assert e < TOL, (name, e)



def _small_oracle(oracle, pot, offx, offy, dets=()):
    """stem_small.qsc's specimen (STO_neutral.cfg 4x4x3 cells, 6 slices, 128^2 at 0.05 A,
    radius 3 A, band-limited transmission): config and transmission functions."""
    from qstem_amd import StemConfig
    from qstem_amd.specimen import build_supercell, read_cfg
    atoms, ext = build_supercell(read_cfg(os.path.join(GOLD, "STO_neutral.cfg")), 4, 4, 3)
    cfg = StemConfig(nx=128, ny=128, resolutionX=0.05, resolutionY=0.05, v0=200.0, slices=6, sliceThickness=1.9525,
                     potNx=pot, potNy=pot, by=float(f(ext[1])), bandlimittrans=1, Cs=float(f(f(0.001) * 1e7)),
                     df0=float(f(10.0 * f(-5.0))), alpha=20.0, detectors=list(dets), electronScale=1.0, compat=3)
    g = oracle.lut_geometry(3.0, cfg.resolutionX, cfg.sliceThickness)
    keys = []
    for a in atoms:
        if int(a[6]) not in [k[0] for k in keys]:
            keys.append((int(a[6]), float(f(a[3]))))
    luts = oracle.atom_luts(g, [k[0] for k in keys], [k[1] for k in keys], cfg.sliceThickness)
    kinds = [[k[0] for k in keys].index(int(a[6])) for a in atoms]
    V = oracle.make_slices(cfg, g, luts, kinds, [tuple(a) for a in atoms], offx, offy, atomRadius=3.0)
    return cfg, oracle.make_trans(cfg, V), ext


def _run_small(tmp_path, mode, extra):
    """stem_small.qsc in another mode; extra "key: value" lines replace the file's line
    of that key (readparam reads the first match) or are appended."""
    shutil.copy(os.path.join(GOLD, "STO_neutral.cfg"), tmp_path / "STO_neutral.cfg")
    keys = dict(line.split(":", 1) for line in extra)
    keys["mode"] = " " + mode
    _write_qsc(os.path.join(GOLD, "stem_small.qsc"), tmp_path / "run.qsc", **{k: v.strip() for k, v in keys.items()})
    r = subprocess.run([EXE, "run.qsc"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return tmp_path / "out_small"


def test_qstem3_tem_tilted_beam(oracle, tmp_path):
    """TEM with beam tilt and tilt back (doTEM, stem3.cpp:2403-2420, 2494-2512): the
    incident wave exp(i (ktx x + kty y)), kt = 2 pi sin(tilt)/lambda, with the
    reference's y = resolutionY (ix - nx/2), and the exit wave.img multiplied by cos
    (real part) and sin (imaginary part) of the back-tilt phase, as the reference
    writes them; wave_t.img at slices 1, 3, 5 before the back tilt."""
    from qstem_amd.imgio import read_img
    out = _run_small(tmp_path, "TEM", ["Beam tilt X: 0.2 deg", "Beam tilt Y: 0.1 deg", "Tilt back: yes"])
    cx = f(f(f(4 * 3.905)) / f(2.0))
    off = f(np.float64(cx) - 0.5 * np.float64(f(128 * f(0.05))))
    cfg, trans, _ = _small_oracle(oracle, 128, float(off), float(off))
    lam = oracle.wavelength(200.0)
    pi = 3.1415926535897  # the driver's constant, as stem3.cpp's
    tx, ty = f(np.float64(f(0.2)) * pi / 180.0), f(np.float64(f(0.1)) * pi / 180.0)  # "deg" values in float
    ktx, kty = 2 * pi * np.sin(np.float64(tx)) / lam, 2 * pi * np.sin(np.float64(ty)) / lam
    ix = np.arange(128)
    xq = (f(0.05) * (ix - 64).astype(np.float32)).astype(np.float64)
    phase = ktx * xq[:, None] + kty * xq[:, None] * np.ones((1, 128))  # y from ix, as the reference
    psi0 = np.cos(phase).astype(np.float32) + 1j * np.sin(phase).astype(np.float32)
    res = _pieces(oracle, cfg, trans, psi0.astype(np.complex128), [2, 4, 6])
    for t, (psi, _) in enumerate(res):
        w = read_img(str(out / ("wave_%d.img" % t)))
        assert rel_l2(w.data, psi) < TOL, (t, rel_l2(w.data, psi))
    back = -ktx * xq[:, None] - kty * xq[:, None] * np.ones((1, 128))
    exit_ref = res[-1][0].real * np.cos(back) + 1j * res[-1][0].imag * np.sin(back)
    w = read_img(str(out / "wave.img"))
    assert rel_l2(w.data, exit_ref) < TOL


def test_qstem3_cbed_save_level_2(oracle, tmp_path):
    """CBED with save level 2: per slice, collectIntensity's diff_t.img of the k-space
    pattern is overwritten by the second collectIntensity call on the real-space wave
    (stemlib.cpp:2476-2482), so diff_t.img ends as fftshift(|psi|^2)/sqrt(N) of the
    wave after the bin's last slice; wave_t.img holds that wave."""
    from qstem_amd.imgio import read_img
    out = _run_small(tmp_path, "CBED", ["save level: 2"])
    x0 = f(5.8575)
    off = f(np.float64(x0) - 0.5 * 128 * np.float64(f(0.05)))
    cfg, trans, _ = _small_oracle(oracle, 128, float(off), float(off))
    psi0 = oracle.probe(cfg, at=(float(x0 - off), float(x0 - off)))
    res = _pieces(oracle, cfg, trans, psi0, [2, 4, 5, 6])  # bins: slices 1, 3, 4 (Q4), 5
    for t, k in ((0, 0), (1, 1), (2, 3)):  # interimWave: slices 1, 3, 5 (slice 5 is t = 5 // 2 = 2)
        w = read_img(str(out / ("wave_%d.img" % t)))
        assert rel_l2(w.data, res[k][0]) < TOL, (t, rel_l2(w.data, res[k][0]))
    for t in range(4):
        psi = res[t][0]
        d = read_img(str(out / ("diff_%d.img" % t))).data
        real_int = np.fft.fftshift(np.abs(psi) ** 2) / 128.0
        # an intensity carries twice the wave's relative error
        assert rel_l2(d, real_int) < 2 * TOL, (t, rel_l2(d, real_int))
