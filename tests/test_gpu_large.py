"""GPU parity at the large transform sizes (1024 one-wavefront transforms;
2048 and 4096 block-wide transforms) against the FP64 oracle.

Tolerance: relative L2 <= 1e-4 for exit waves and detector images, 2e-5 for the
probe.  Windows of 2048 and 4096 use potN = N (a single window at the origin,
the plane-wave TEM geometry), since row transforms of the transmission
function are limited to 4096 points (Bluestein for other lengths needs
2*potN-1 <= 4096).
"""
import numpy as np
import pytest

from cases import probe_positions, rel_l2, small_config, smooth_potential

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _ctx(cfg):
    from qstem_amd import Multislice
    return Multislice(cfg)


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_probe_large(oracle, n):
    cfg = small_config(nx=n, ny=n, pad=0, dx=0.02)
    with _ctx(cfg) as ms:
        g = ms.make_probe()
    o = oracle.probe(cfg)
    assert abs(np.sum(np.abs(g) ** 2) / (n * n) - 1) < 1e-5
    assert rel_l2(g, o) < 2e-5


@pytest.mark.parametrize("shape,pad,nprobe", [((1024, 1024), 8, 2), ((1024, 2048), 0, 1), ((2048, 1024), 0, 1),
                                             ((2048, 2048), 0, 1), ((4096, 4096), 0, 1)])
def test_multislice_large(oracle, shape, pad, nprobe):
    nx, ny = shape
    slices = 2
    print("large multislice", shape, "pad", pad, flush=True)
    cfg = small_config(nx=nx, ny=ny, pad=pad, slices=slices, dx=0.02, v0=300.0)
    V = smooth_potential(slices, cfg.potNx, cfg.potNy, seed=17, sigma_px=3.0)
    ipos = probe_positions(cfg, nprobe, seed=6)
    bins, nbins = oracle.record_bins(slices)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t32 = ms.get_trans()
        psi0 = ms.make_probe()
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True, want_intensity=True)
    ex, det, _ = oracle.multislice(cfg, t32.astype(np.complex128), ipos, psi0.astype(np.complex128),
                                   bins=bins, nbins=nbins)
    assert rel_l2(out["exit"], ex) < TOL
    for b in range(nbins):
        for d in range(det.shape[1]):
            if np.any(det[b, d] != 0):
                assert rel_l2(out["det"][b, d], det[b, d]) < TOL, (b, d)
    norm = np.sum(np.abs(ex) ** 2, axis=(1, 2)) / (nx * ny)
    assert np.max(np.abs(out["intensity"] - norm) / norm) < TOL


@pytest.mark.parametrize("n,slices", [(1024, 64), (2048, 16)])
def test_multislice_many_slices(oracle, n, slices):
    """Error growth over many slices at the sizes whose transforms take their twiddles
    from the table (products of two table values): exit wave and detector sums against
    the FP64 oracle after 64 slices at 1024^2 and 16 at 2048^2."""
    print("many-slice multislice", n, slices, flush=True)
    cfg = small_config(nx=n, ny=n, pad=0, slices=slices, dx=0.02, v0=300.0)
    V = smooth_potential(slices, cfg.potNx, cfg.potNy, seed=23, sigma_px=3.0)
    ipos = probe_positions(cfg, 1, seed=3)
    bins, nbins = oracle.record_bins(slices)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        t32 = ms.get_trans()
        psi0 = ms.make_probe()
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True)
    ex, det, _ = oracle.multislice(cfg, t32.astype(np.complex128), ipos, psi0.astype(np.complex128),
                                   bins=bins, nbins=nbins)
    err = rel_l2(out["exit"], ex)
    print("exit-wave rel L2", err, flush=True)
    assert err < TOL
    for b in range(nbins):
        for d in range(det.shape[1]):
            if np.any(det[b, d] != 0):
                assert rel_l2(out["det"][b, d], det[b, d]) < TOL, (b, d)


@pytest.mark.parametrize("n,nprobe,pad,slices", [(512, 32, 64, 6), (1024, 4, 0, 4), (2048, 1, 0, 4), (4096, 1, 0, 3)])
def test_exit_waves_repeat_bit_for_bit(n, nprobe, pad, slices):
    """Repeated runs give identical exit waves: the slice sweeps have no atomics or
    order-dependent sums, so any difference is a race.  (Wide stores whose data registers
    were overwritten before the store had read them made results differ between runs.)
    Two probe groups on the two HIP streams at 512^2 and 1024^2, one probe otherwise."""
    cfg = small_config(nx=n, ny=n, pad=pad, slices=slices, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=n)
    ipos = probe_positions(cfg, nprobe, seed=3) if pad else np.zeros((nprobe, 2), np.int32)
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        if nprobe > 1:
            ms.set_batch(nprobe // 2)  # two probe groups, one per stream
        runs = [ms.run_stem(ipos, want_exit=True)["exit"] for _ in range(4)]
    assert np.all(np.isfinite(runs[0])) and np.abs(runs[0]).max() > 0
    for r, ex in enumerate(runs[1:], 1):
        assert np.array_equal(ex, runs[0]), (r, float(np.abs(ex - runs[0]).max()))
