"""The FP32 CPU multislice (oracle/qs_cpu32.c, the benchmark's CPU baseline)
against the FP64 oracle: exit waves and every detector image to 1e-4 relative
L2, the same bar as the GPU path, so that the baseline times a correct port."""
import numpy as np
import pytest

from cases import probe_positions, rel_l2, small_config, smooth_potential


@pytest.mark.parametrize("shape,slices", [((128, 128), 6), ((256, 256), 3), ((64, 128), 4), ((400, 400), 2),
                                          ((90, 150), 3), ((48, 120), 2)])
def test_cpu32_matches_oracle(oracle, shape, slices):
    """Powers of two, and the mixed-radix sizes (radix 3 and 5 passes, and row and
    column blocks narrower than 16 and 32 at the edge: 400 = 4*4*5*5, 90 = 2*3*3*5,
    150 = 2*3*5*5, 48 = 4*4*3, 120 = 4*2*3*5)."""
    nx, ny = shape
    cfg = small_config(nx=nx, ny=ny, pad=24, slices=slices, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=5)
    trans = oracle.make_trans(cfg, V)
    psi0 = oracle.probe(cfg)
    ipos = probe_positions(cfg, 5, seed=6)
    bins, nbins = oracle.record_bins(cfg.slices, outputInterval=2)
    ex, det = oracle.cpu32_multislice(cfg, trans, psi0, ipos, bins=bins, nbins=nbins, want_exit=True, nthreads=2)
    ex_o, det_o, _ = oracle.multislice(cfg, trans, ipos, psi0, bins=bins, nbins=nbins)
    assert rel_l2(ex, ex_o) < 1e-4
    checked = 0
    for b in range(nbins):
        for d in range(det_o.shape[1]):
            if np.any(det_o[b, d]):
                assert rel_l2(det[b, d], det_o[b, d]) < 1e-4, (b, d)
                checked += 1
    assert checked >= nbins


def test_cpu32_threads_agree(oracle):
    """One probe per OpenMP thread: the result does not depend on the thread count."""
    cfg = small_config(nx=128, ny=128, pad=16, slices=3, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=7)
    trans = oracle.make_trans(cfg, V)
    psi0 = oracle.probe(cfg)
    ipos = probe_positions(cfg, 7, seed=8)
    bins, nbins = oracle.record_bins(cfg.slices)
    a = oracle.cpu32_multislice(cfg, trans, psi0, ipos, bins=bins, nbins=nbins, want_exit=True, nthreads=1)
    b = oracle.cpu32_multislice(cfg, trans, psi0, ipos, bins=bins, nbins=nbins, want_exit=True, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
