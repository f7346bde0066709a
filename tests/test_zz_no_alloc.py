"""qs_run_stem_noalloc (every device buffer in the library's static scratch array)
against qs_run_stem, through the C ABI: two calls in a row, the second with fewer
probes, each compared with qs_run_stem on the same inputs and with the FP64
oracle.  Tolerance: 1e-5 relative L2 against qs_run_stem (same kernels; only the
order of the detector sums' float atomics may differ), 1e-4 against the oracle."""
import numpy as np
import pytest

from cases import probe_positions, rel_l2, small_config, smooth_potential

pytestmark = pytest.mark.gpu


def _compare(ms, ipos, bins, nbins, ref_det=None):
    a = ms.run_stem(ipos, bins=bins, nbins=nbins, want_intensity=True)
    b = ms.run_stem_noalloc(ipos, bins=bins, nbins=nbins, want_intensity=True)
    for k in range(nbins):
        for d in range(a["det"].shape[1]):
            if np.any(a["det"][k, d]):
                assert rel_l2(b["det"][k, d], a["det"][k, d]) < 1e-5, (k, d)
    assert rel_l2(b["intensity"], a["intensity"]) < 1e-5
    return b


@pytest.mark.parametrize("n,probes", [(256, (40, 7)), (512, (30, 5))])
def test_noalloc_matches_run_stem(oracle, n, probes):
    from qstem_amd import Multislice
    cfg = small_config(nx=n, ny=n, pad=24, slices=4, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=11)
    bins, nbins = oracle.record_bins(cfg.slices)
    trans = oracle.make_trans(cfg, V)
    psi0 = oracle.probe(cfg)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        for call, npr in enumerate(probes):
            ipos = probe_positions(cfg, npr, seed=20 + call)
            print("noalloc: n=%d call %d probes %d" % (n, call, npr), flush=True)
            got = _compare(ms, ipos, bins, nbins)
            _, det, _ = oracle.multislice(cfg, trans, ipos, psi0, bins=bins, nbins=nbins)
            for k in range(nbins):
                for d in range(det.shape[1]):
                    if np.any(det[k, d]):
                        assert rel_l2(got["det"][k, d], det[k, d]) < 1e-4, (call, k, d)


def test_noalloc_rejects_more_than_the_scratch():
    """512^2 waves need ~1.44 MB of state per probe: 60 probes exceed the 64 MiB
    array, which the host rejects before any launch."""
    from qstem_amd import Multislice, QstemError
    cfg = small_config(nx=512, ny=512, pad=8, slices=2, dx=0.03)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=12)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        with pytest.raises(QstemError, match="scratch"):
            ms.run_stem_noalloc(probe_positions(cfg, 60, seed=3))
