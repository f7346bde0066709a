"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerance: relative L2 <= 1e-4 for every output (the FP32 target stated for
this path); the probe and transmission functions, which involve one or two
transforms only, are held to 2e-5.
"""
import numpy as np
import pytest

from cases import probe_positions, rel_l2, small_config, smooth_potential

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _ctx(cfg):
    from qstem_amd import Multislice
    return Multislice(cfg)


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_probe_matches_oracle(oracle, n):
    cfg = small_config(nx=n, ny=n, pad=16, dx=0.03)
    with _ctx(cfg) as ms:
        g = ms.make_probe()
    o = oracle.probe(cfg)
    assert abs(np.sum(np.abs(g) ** 2) / (n * n) - 1) < 1e-5
    assert rel_l2(g, o) < 2e-5


@pytest.mark.parametrize("n,opts", [(128, dict(gaussFlag=1, gaussScale=0.1)), (256, dict(aAIS=3.0)),
                                    (256, dict(ismoth=1, Cc=1e7, dE_E=2e-6)), (400, dict(aAIS=4.0, ismoth=1)),
                                    (128, dict(gaussFlag=1, gaussScale=0.02, aAIS=2.5))])
def test_probe_options_match_oracle(oracle, n, opts):
    """The probe's real-space Gaussian envelope (gaussian:), the AIS aperture with its
    cosine edge (AIS aperture:), the smoothed aperture rim (smooth:) and the chromatic
    defocus offset Cc dE/E, at power-of-two and 400^2 (row-transform) grids, and at a
    centre moved by whole and fractional pixels (qs_make_probe_at)."""
    cfg = small_config(nx=n, ny=n, pad=16, dx=0.04)
    for k, v in opts.items():
        setattr(cfg, k, v)
    cx = float(np.float32(n // 2) * np.float32(0.04))
    with _ctx(cfg) as ms:
        g = ms.make_probe()
        g_at = ms.make_probe(at=(cx + 3 * 0.04, cx - 1.37))
    assert abs(np.sum(np.abs(g) ** 2) / (n * n) - 1) < 1e-5
    assert rel_l2(g, oracle.probe(cfg)) < 2e-5
    assert rel_l2(g_at, oracle.probe(cfg, at=(cx + 3 * 0.04, cx - 1.37))) < 2e-5


@pytest.mark.parametrize("pot", [(160, 160), (128, 128), (200, 136), (256, 128), (128, 512), (1024, 2048),
                                 (2048, 1024), (2056, 136)])
def test_transmission_matches_oracle(oracle, pot):
    """Power-of-two grids take the band-compact sweeps, the others row transforms and
    transposes: mixed-radix LDS passes for lengths with prime factors 2, 3, 5, 7 (160,
    200), Bluestein through a power of two up to 4096 (136 = 8*17), and beyond that
    Bluestein through the LDS passes (2056 = 8*257: length 4116 = 4*3*7^3)."""
    cfg = small_config(nx=128, ny=128, pad=0)
    cfg.potNx, cfg.potNy = pot
    V = smooth_potential(3, pot[0], pot[1], seed=3)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t = ms.get_trans()
    o = oracle.make_trans(cfg, V)
    assert rel_l2(t, o) < 2e-5


@pytest.mark.parametrize("pad", [0, 32, 128])
def test_transmission_quirk_q1(oracle, pad):
    """Q1 with by = 1.5 potSizeY: the negative-ky half of the band reaches past
    potNy/3 (every negative ky is kept), which the band-compact sweeps cannot hold;
    a rough potential (0.5 px filter) puts weight on those columns."""
    cfg = small_config(nx=128, ny=128, pad=pad, compat=1)
    V = smooth_potential(2, cfg.potNx, cfg.potNy, seed=4, sigma_px=0.5)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t = ms.get_trans()
    assert rel_l2(t, oracle.make_trans(cfg, V)) < 2e-5


@pytest.mark.parametrize("shape", [(128, 128), (256, 256), (64, 128), (512, 512)])
def test_multislice_matches_oracle(oracle, shape):
    nx, ny = shape
    slices = 4 if nx == 512 else 6
    cfg = small_config(nx=nx, ny=ny, pad=24, slices=slices, dx=0.03 if nx >= 256 else 0.05)
    V = smooth_potential(slices, cfg.potNx, cfg.potNy, seed=11)
    ipos = probe_positions(cfg, 3, seed=5)
    bins, nbins = oracle.record_bins(slices, outputInterval=2)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t32 = ms.get_trans()
        psi0 = ms.make_probe()
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_diff=True, want_exit=True, want_intensity=True)
    # oracle runs on the same FP32 transmission and probe so the comparison isolates the multislice
    ex, det, diff = oracle.multislice(cfg, t32.astype(np.complex128), ipos, psi0.astype(np.complex128),
                                      bins=bins, nbins=nbins, want_diff=True)
    assert rel_l2(out["exit"], ex) < TOL
    assert rel_l2(out["diff"], diff) < TOL
    for b in range(nbins):
        for d in range(det.shape[1]):
            assert rel_l2(out["det"][b, d], det[b, d]) < TOL, (b, d)
    norm = np.sum(np.abs(ex) ** 2, axis=(1, 2)) / (nx * ny)
    assert np.max(np.abs(out["intensity"] - norm) / norm) < TOL


def test_full_chain_against_fp64_oracle(oracle):
    """Probe, transmission and multislice all from the oracle in FP64 vs all on the GPU in FP32."""
    cfg = small_config(nx=256, ny=256, pad=40, slices=10, dx=0.03, v0=300.0)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=21)
    ipos = probe_positions(cfg, 4, seed=9)
    bins, nbins = oracle.record_bins(cfg.slices)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True)
    ex, det, _ = oracle.multislice(cfg, oracle.make_trans(cfg, V), ipos, oracle.probe(cfg), bins=bins, nbins=nbins)
    assert rel_l2(out["exit"], ex) < TOL
    for b in range(nbins):
        for d in range(det.shape[1]):
            if np.any(det[b, d] != 0):
                assert rel_l2(out["det"][b, d], det[b, d]) < TOL, (b, d)


def test_batch_sizes_agree():
    """Per-probe launches (batch 1) and batched launches give the same numbers."""
    cfg = small_config(nx=128, ny=128, pad=24, slices=5)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=2)
    ipos = probe_positions(cfg, 7, seed=1)
    res = []
    for batch in (1, 3, 0):
        with _ctx(cfg) as ms:
            ms.set_potential(V)
            ms.make_probe()
            ms.set_batch(batch)
            res.append(ms.run_stem(ipos, want_exit=True))
    for r in res[1:]:
        # detector sums use LDS float atomics inside a block: order-dependent rounding only
        assert rel_l2(r["det"], res[0]["det"]) < 1e-6
        assert np.array_equal(r["exit"], res[0]["exit"])


def test_continue_equals_one_run(oracle):
    """Two calls of 3 slices with start=1 equal one call of 6 slices (slab continuation, cellDiv > 1)."""
    cfg = small_config(nx=128, ny=128, pad=16, slices=6)
    V = smooth_potential(6, cfg.potNx, cfg.potNy, seed=8)
    ipos = probe_positions(cfg, 2, seed=3)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        full = ms.run_stem(ipos, want_exit=True)["exit"]
    with _ctx(cfg) as ms:
        ms.set_potential(V[:3])
        ms.make_probe()
        ms.run_stem(ipos, nslice=3, want_det=False)
        ms.set_potential(V[3:])
        part = ms.run_stem(ipos, nslice=3, start=1, want_exit=True, want_det=False)["exit"]
    assert rel_l2(part, full) < 1e-6


def test_graph_replay_equals_eager():
    """The HIP-graph replay of the slice loop equals eager launches, on repeated
    replays with eager work in between and with new window offsets each call."""
    cfg = small_config(nx=256, ny=256, pad=24, slices=5, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=12)
    ipos_a = probe_positions(cfg, 6, seed=2)
    ipos_b = probe_positions(cfg, 6, seed=3)
    bins = [-1, 0, -1, -1, 1]
    kw = dict(bins=np.array(bins, np.int32), nbins=2, want_diff=True, want_exit=True, want_intensity=True)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        ms.set_batch(4)  # two probe groups per call
        ref_a = ms.run_stem(ipos_a, **kw)
        ref_b = ms.run_stem(ipos_b, **kw)
        ms.set_graph(True)
        outs = []
        for ip in (ipos_a, ipos_b, ipos_a, ipos_b):
            outs.append(ms.run_stem(ip, **kw))
            ms.get_trans()  # eager work between replays
        ms.set_graph(False)
        again = ms.run_stem(ipos_a, **kw)
    for out, ref in zip(outs, (ref_a, ref_b, ref_a, ref_b)):
        assert np.array_equal(out["exit"], ref["exit"])
        assert rel_l2(out["det"], ref["det"]) < 1e-6
        assert rel_l2(out["diff"], ref["diff"]) < 1e-6
        assert np.allclose(out["intensity"], ref["intensity"], rtol=1e-6)
    assert np.array_equal(again["exit"], ref_a["exit"])


def test_autotune_keeps_results():
    """The autotuner only changes launch shapes: results after tuning equal the defaults."""
    cfg = small_config(nx=256, ny=256, pad=24, slices=4, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=13)
    ipos = probe_positions(cfg, 40, seed=4)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        ref = ms.run_stem(ipos, want_exit=True)
        best = ms.autotune(ipos)
        assert best["batch"] in (16, 32, 40) and best["rate"] > 0
        # 0 = library default (kept unless another shape is measurably faster)
        assert best["rows_per_wave"] in (0, 1, 2, 4, 8) and best["tiles_per_block"] in (0, 1, 2, 4)
        assert ms.get_tuning() == (best["rows_per_wave"], best["tiles_per_block"], best["batch"])
        out = ms.run_stem(ipos, want_exit=True)
    assert np.array_equal(out["exit"], ref["exit"])
    assert rel_l2(out["det"], ref["det"]) < 1e-6


@pytest.mark.parametrize("shape", [(400, 400), (400, 160), (160, 400), (384, 384), (320, 448), (600, 768), (300, 441),
                                   (1280, 640), (96, 160), (200, 128), (112, 98), (136, 88)])
def test_generic_grid_sizes_match_oracle(oracle, shape):
    """Wave grids that are not powers of two (the reference's stem_STO_4x4.qsc uses
    400 x 400): sizes whose prime factors are 2, 3, 5, 7 take the fused mixed-radix
    path (radix 2, 3, 4, 5, 7, 8 passes: 400 = 4*4*5*5, 96 = 8*4*3, 160 = 8*4*5, 200 =
    8*5*5, 112 = 4*4*7, 98 = 2*7*7), others (136 = 8*17, 88 = 8*11) the unfused path with
    Bluestein rows: probe, exit waves, every detector image (with a shifted detector),
    diffraction patterns and intensities against the oracle, and a run continued over a
    second call equal to one call over all slices.  300, 320, 384, 400, 441, 448, 600,
    640, 768 and 1280 run the compiled plans (wavefront-resident kernels) on their axes,
    160 the runtime plan."""
    nx, ny = shape
    print("generic grid", shape, flush=True)
    cfg = small_config(nx=nx, ny=ny, pad=24, slices=4, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=12)
    V[2:] = V[:2]  # a continued call restarts at slice 0: slices 2, 3 repeat 0, 1
    ipos = probe_positions(cfg, 3, seed=6)
    bins, nbins = oracle.record_bins(cfg.slices, outputInterval=2)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t32 = ms.get_trans()
        psi0 = ms.make_probe()
        ms.set_batch(2)  # two probe groups
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_diff=True, want_exit=True, want_intensity=True)
        ms.run_stem(ipos, nslice=2, bins=bins[:2], nbins=nbins, want_det=False)
        cont = ms.run_stem(ipos, nslice=2, start=1, bins=bins[2:], nbins=nbins, want_det=False, want_exit=True)
    assert rel_l2(psi0, oracle.probe(cfg)) < 2e-5
    assert rel_l2(t32, oracle.make_trans(cfg, V)) < 2e-5
    ex, det, diff = oracle.multislice(cfg, t32.astype(np.complex128), ipos, psi0.astype(np.complex128),
                                      bins=bins, nbins=nbins, want_diff=True)
    assert rel_l2(out["exit"], ex) < TOL
    assert rel_l2(out["diff"], diff) < TOL
    checked = 0
    for b in range(nbins):
        for d in range(det.shape[1]):
            if np.any(det[b, d] != 0):
                assert rel_l2(out["det"][b, d], det[b, d]) < TOL, (b, d)
                checked += 1
    assert checked >= 2 * nbins
    norm = np.sum(np.abs(ex) ** 2, axis=(1, 2)) / (nx * ny)
    assert np.max(np.abs(out["intensity"] - norm) / norm) < TOL
    assert rel_l2(cont["exit"], out["exit"]) < 1e-6


@pytest.mark.parametrize("n,ndet", [(400, 4), (400, 6), (288, 4), (300, 4), (320, 4), (360, 4), (384, 6), (441, 4),
                                    (448, 4), (500, 4), (540, 4), (600, 4), (640, 4), (768, 4), (900, 4), (1280, 4),
                                    (1350, 4), (1536, 4), (1800, 6)])
def test_compiled_plans_equal_runtime_plans(oracle, monkeypatch, n, ndet):
    """Lengths with a compiled plan (the MixPlan table of qs_mix.hpp: 400 = 20 x 20, the
    others three or four passes such as 384 = 24 x 8 x 2 or 1200 = 30 x 10 x 2 x 2;
    wavefront-resident kernels, the default); QSTEM_COMPILED_PLANS=0 at context
    creation keeps the runtime-planned
    kernels.  Both agree to FP32 rounding on every output, over 300 probes in groups of
    128 (the last group partial, rows and columns not multiples of a block's), with up
    to four detectors (the column kernel's small variant) and with six (two shifted)."""
    from qstem_amd import Detector
    cfg = small_config(nx=n, ny=n, pad=40, slices=3, dx=0.03)
    if ndet == 6:
        cfg.detectors = cfg.detectors + [Detector(5.0, 15.0, "D5"), Detector(30.0, 90.0, "D6", -1.5, 2.0)]
    assert len(cfg.detectors) == ndet
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=21)
    ipos = probe_positions(cfg, 300, seed=8)
    bins, nbins = oracle.record_bins(cfg.slices, outputInterval=1)
    res = []
    for flag in ("1", "0"):
        monkeypatch.setenv("QSTEM_COMPILED_PLANS", flag)
        print("compiled plans", flag, n, ndet, flush=True)
        with _ctx(cfg) as ms:
            ms.set_potential(V)
            ms.make_probe()
            ms.set_batch(128)
            res.append(ms.run_stem(ipos, bins=bins, nbins=nbins, want_diff=True, want_exit=True,
                                   want_intensity=True))
    a, b = res
    assert rel_l2(a["exit"], b["exit"]) < 2e-6
    assert rel_l2(a["diff"], b["diff"]) < 4e-6
    for k in range(nbins):
        for d in range(a["det"].shape[1]):
            assert rel_l2(a["det"][k, d], b["det"][k, d]) < 4e-6, (k, d)
    assert np.max(np.abs(a["intensity"] - b["intensity"]) / b["intensity"]) < 4e-6


@pytest.mark.parametrize("n", [96, 88])
def test_generic_grid_refusals(n):
    """The adjoint, uneven batches and the no-allocation run need the power-of-two
    sweeps, and the autotuner the power-of-two or the mixed-radix path (88 = 8*11 has
    neither): they refuse other grids with an error, before any launch."""
    from qstem_amd import QstemError, run_groups
    cfg = small_config(nx=n, ny=n, pad=8, slices=2, dx=0.05)
    with _ctx(cfg) as ms:
        ms.set_potential(smooth_potential(2, cfg.potNx, cfg.potNy, seed=3))
        ms.make_probe()
        ipos = np.zeros((1, 2), np.int32)
        calls = [lambda: ms.adjoint(ipos, np.zeros((1, n, n), np.float32)),
                 lambda: ms.run_stem_noalloc(ipos),
                 lambda: run_groups([(ms, ipos, None)])]
        if n == 88:
            calls.append(lambda: ms.autotune(ipos))
        for call in calls:
            with pytest.raises(QstemError, match="power"):
                call()


_TUNE_SCRIPT = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
from qstem_amd import Multislice
from cases import probe_positions, small_config, smooth_potential
cfg = small_config(nx=256, ny=256, pad=32, slices=8, dx=0.03)
with Multislice(cfg) as ms:
    ms.set_potential(smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=4))
    ms.make_probe()
    t0 = time.time()
    best = ms.autotune(probe_positions(cfg, 64, seed=2))
    best["seconds"] = time.time() - t0
print(json.dumps(best))
"""


def test_autotune_disk_cache(tmp_path):
    """QSTEM_TUNE_CACHE names a JSON file: the first process measures and writes it; a
    second process reads it and applies the stored launch shape without measuring."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = tmp_path / "tune.json"
    env = dict(os.environ, QSTEM_TUNE_CACHE=str(path))
    runs = []
    for i in range(2):
        print("autotune process %d" % i, flush=True)
        r = subprocess.run([sys.executable, "-c", _TUNE_SCRIPT, root], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
        if i == 0:
            entries = json.load(open(path))["qstem_amd_autotune"]
            assert len(entries) == 1 and entries[0]["nx"] == 256 and entries[0]["batch"] == runs[0]["batch"]
    a, b = runs
    assert (a["batch"], a["rows_per_wave"], a["tiles_per_block"]) == (b["batch"], b["rows_per_wave"],
                                                                       b["tiles_per_block"])
    assert abs(a["rate"] - b["rate"]) <= 1e-6 * a["rate"] + 0.1 and b["seconds"] < 0.5 * a["seconds"]


def test_autotune_mixed_radix_path(oracle):
    """On a 400 x 400 grid the autotuner chooses group size, rows per block, columns per
    tile and the radix plans; the tuned run agrees with the default one and with the
    oracle, and qs_set_tuning refuses shapes that do not fit."""
    from qstem_amd import QstemError
    cfg = small_config(nx=400, ny=400, pad=24, slices=4, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=5)
    ipos = probe_positions(cfg, 6, seed=4)
    bins, nbins = oracle.record_bins(cfg.slices, outputInterval=2)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        psi0 = ms.make_probe()
        t32 = ms.get_trans()
        ref = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True)
        best = ms.autotune(ipos)
        print("mixed-radix tuning", best)
        tuned = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True)
        with pytest.raises(QstemError):
            ms.set_tuning(3, 0)
        with pytest.raises(QstemError):
            ms.set_tuning(32, 0)
        ms.set_tuning(2, 4)
        other = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True)
    assert rel_l2(tuned["exit"], ref["exit"]) < 1e-5 and rel_l2(other["exit"], ref["exit"]) < 1e-5
    assert rel_l2(tuned["det"], ref["det"]) < 1e-5
    ex, det, _ = oracle.multislice(cfg, t32.astype(np.complex128), ipos, psi0.astype(np.complex128), bins=bins,
                                   nbins=nbins)
    assert rel_l2(tuned["exit"], ex) < TOL
