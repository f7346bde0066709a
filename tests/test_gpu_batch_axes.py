"""GPU: frozen-phonon configurations as a batch axis and uneven batches of
several window sizes over one shared transmission function.

Every batched result is checked against the same items run one configuration /
one window size at a time through the plain path (bit-identical exit waves;
detector sums to 1e-6, as they add in a launch-shape dependent order), and
samples against the FP64 oracle at the 1e-4 relative-L2 target.
"""
import numpy as np
import pytest

from cases import probe_positions, rel_l2, small_config, smooth_potential

pytestmark = pytest.mark.gpu


def _ctx(cfg):
    from qstem_amd import Multislice
    return Multislice(cfg)


def test_phonon_configurations_batch(oracle):
    cfg = small_config(nx=128, ny=128, pad=32, slices=5)
    Vs = [smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=40 + k) for k in range(3)]
    ipos = probe_positions(cfg, 7, seed=8)
    config = np.array([0, 2, 1, 2, 0, 1, 2], np.int32)
    bins, nbins = oracle.record_bins(cfg.slices, outputInterval=2)
    with _ctx(cfg) as ms:
        ms.set_configs(3)
        for k in range(3):
            ms.select_config(k)
            ms.set_potential(Vs[k])
        ms.make_probe()
        ms.set_batch(3)  # probe groups straddle configurations
        out = ms.run_stem(ipos, bins=bins, nbins=nbins, want_exit=True, config=config)
        ms.select_config(2)
        t2 = ms.get_trans()
        psi0 = ms.make_probe()
    for k in range(3):
        sel = np.nonzero(config == k)[0]
        with _ctx(cfg) as ms1:
            ms1.set_potential(Vs[k])
            ms1.make_probe()
            ref = ms1.run_stem(ipos[sel], bins=bins, nbins=nbins, want_exit=True)
        assert np.array_equal(out["exit"][sel], ref["exit"]), k
        assert rel_l2(out["det"][:, :, sel], ref["det"]) < 1e-6, k
    sel = np.nonzero(config == 2)[0]
    ex, det, _ = oracle.multislice(cfg, t2.astype(np.complex128), ipos[sel], psi0.astype(np.complex128),
                                   bins=bins, nbins=nbins)
    assert rel_l2(out["exit"][sel], ex) < 1e-4


def test_uneven_window_sizes_share_one_potential(oracle):
    from qstem_amd import run_groups
    cfg_a = small_config(nx=128, ny=128, pad=192, slices=4, dx=0.03)  # potN 320
    cfg_b = small_config(nx=256, ny=256, pad=64, slices=4, dx=0.03)   # potN 320
    V = [smooth_potential(4, 320, 320, seed=50 + k) for k in range(2)]
    ipa = probe_positions(cfg_a, 5, seed=1)
    ipb = probe_positions(cfg_b, 3, seed=2)
    cfa = np.array([0, 1, 1, 0, 1], np.int32)
    cfb = np.array([1, 0, 1], np.int32)
    bins, nbins = oracle.record_bins(4)
    with _ctx(cfg_a) as A, _ctx(cfg_b) as B:
        A.set_configs(2)
        for k in range(2):
            A.select_config(k)
            A.set_potential(V[k])
        B.share_trans(A)
        A.make_probe()
        psi_b = B.make_probe()
        B.set_batch(2)
        outs = run_groups([(A, ipa, cfa), (B, ipb, cfb)], bins=bins, nbins=nbins, want_intensity=True)
        ref_a = A.run_stem(ipa, bins=bins, nbins=nbins, config=cfa, want_intensity=True)
        ref_b = B.run_stem(ipb, bins=bins, nbins=nbins, config=cfb, want_intensity=True, want_exit=True)
        A.select_config(1)
        t1 = A.get_trans()
    assert rel_l2(outs[0]["det"], ref_a["det"]) < 1e-6
    assert rel_l2(outs[1]["det"], ref_b["det"]) < 1e-6
    assert np.allclose(outs[0]["intensity"], ref_a["intensity"], rtol=1e-6)
    assert np.allclose(outs[1]["intensity"], ref_b["intensity"], rtol=1e-6)
    sel = np.nonzero(cfb == 1)[0]
    ex, det, _ = oracle.multislice(cfg_b, t1.astype(np.complex128), ipb[sel], psi_b.astype(np.complex128),
                                   bins=bins, nbins=nbins)
    assert rel_l2(ref_b["exit"][sel], ex) < 1e-4
    for d in range(det.shape[1]):
        if np.any(det[-1, d] != 0):
            assert rel_l2(outs[1]["det"][-1, d, sel], det[-1, d]) < 1e-4, d


def test_share_trans_rejects_other_grids():
    from qstem_amd import QstemError
    cfg_a = small_config(nx=128, ny=128, pad=64)
    cfg_b = small_config(nx=128, ny=128, pad=32)
    with _ctx(cfg_a) as A, _ctx(cfg_b) as B:
        with pytest.raises(QstemError):
            B.share_trans(A)


def test_two_streams_equal_one_stream():
    """Probe groups alternating over two HIP streams (qs_set_streams 2) give the same
    results as one stream, eager and as a HIP graph, for run_stem with every output
    and for run_groups; an invalid stream count is rejected."""
    from qstem_amd import QstemError, run_groups
    cfg = small_config(nx=128, ny=128, pad=40, slices=5)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=60)
    ipos = probe_positions(cfg, 11, seed=9)
    bins = np.array([0, -1, 1, -1, 2], np.int32)
    kw = dict(bins=bins, nbins=3, want_diff=True, want_exit=True, want_intensity=True)
    res = {}
    with _ctx(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        ms.set_batch(3)  # four probe groups: two per stream, the last one partial
        for ns in (1, 2):
            ms.set_streams(ns)
            assert ms.get_streams() == ns
            for graph in (False, True, True):
                ms.set_graph(graph)
                res[(ns, graph, len(res))] = ms.run_stem(ipos, **kw)
            ms.set_graph(False)
        with pytest.raises(QstemError):
            ms.set_streams(3)
    ref = res[(1, False, 0)]
    for key, out in res.items():
        assert np.array_equal(out["exit"], ref["exit"]), key
        assert rel_l2(out["det"], ref["det"]) < 1e-6, key
        assert rel_l2(out["diff"], ref["diff"]) < 1e-6, key
        assert np.allclose(out["intensity"], ref["intensity"], rtol=1e-6), key

    cfg_a = small_config(nx=128, ny=128, pad=64, slices=3, dx=0.03)   # potN 192
    cfg_b = small_config(nx=64, ny=64, pad=128, slices=3, dx=0.03)    # potN 192
    Vg = smooth_potential(3, 192, 192, seed=61)
    ipa = probe_positions(cfg_a, 6, seed=3)
    ipb = probe_positions(cfg_b, 9, seed=4)
    gb, gn = np.array([0, -1, 1], np.int32), 2
    outs = {}
    for ns in (1, 2):
        with _ctx(cfg_a) as A, _ctx(cfg_b) as B:
            A.set_potential(Vg)
            B.share_trans(A)
            A.make_probe()
            B.make_probe()
            A.set_streams(ns)
            B.set_batch(4)
            outs[ns] = run_groups([(A, ipa, None), (B, ipb, None)], bins=gb, nbins=gn, want_intensity=True)
    for g in range(2):
        assert rel_l2(outs[2][g]["det"], outs[1][g]["det"]) < 1e-6, g
        assert np.allclose(outs[2][g]["intensity"], outs[1][g]["intensity"], rtol=1e-6), g
