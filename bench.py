"""Headline benchmark: batched FP32 STEM multislice on one MI355X.

Workload (the batched plain-STEM configuration): SrTiO3-type perovskite
a = 3.905 A, 12x12 cells lateral, 40 cells thick -> 80 slices of 1.9525 A;
512x512 wave at dx = 0.030 A; 300 kV, alpha = 25 mrad, Cs = 1 um,
defocus = -5 nm; 64x64 scan positions over one unit cell; 256 probes per
launch; detectors BF 0-10, ABF 11-22, ADF 70-200 mrad.

The projected potential is built on the GPU from the atoms of the 12x12x40
supercell (neutral Sr, Ti, O with the Debye-Waller factors of the reference's
SrTiO3.cfg; qs_build_potential), and its build time is reported in setup_s.

A "step" is one launch group: 256 probes advanced through all 80 slices
(20480 probe-slices).  --steps 16 covers the whole 64x64 scan.

value = probe-slices/s over the timed steps (inputs resident in HBM).
Also reported: the per-probe GPU loop (same kernels, one probe per launch
sequence), the batched/per-probe ratio, the adjoint/forward time ratio of the
4D-STEM adjoint workload, the detector error against the FP64 oracle on a
sample of probes, the roofline of the slice-step kernel pair, the plain CPU
FP32 multislice (oracle/qs_cpu32.c, OpenMP over probes) timed on the host cores
on a sample of this workload and on the CPU configuration (128^2, 20 slices,
16x16 scan), and, as extra keys, the uneven frozen-phonon batch, the reference's
own example configuration (stem_STO_4x4.qsc: charged SrTiO3, 400^2 wave, 20x20
scan, 3 phonon configurations) and the 4096^2 TEM workload with its checkpointed
adjoint.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "probe-slices/s batched vs per-probe FP32 at ≤1e-4 rel err; adjoint/fwd time"
HBM_PEAK_GBS = 8000.0


def perovskite_config():
    from qstem_amd import Detector, StemConfig, scan_offsets
    a = 3.905
    dx = 0.030
    nx = 512
    ncell = 12
    ax = ncell * a
    # scan one unit cell in the middle of the 12x12 supercell
    x0 = np.float32(6 * a)
    x1 = np.float32(7 * a)
    f32 = np.float32
    rng_px = int(f32(x1 - x0) / f32(dx))  # stem3.cpp:1439-1444
    pot = nx + rng_px
    cfg = StemConfig(
        nx=nx, ny=nx, resolutionX=dx, resolutionY=dx, v0=300.0, slices=80, sliceThickness=a / 2,
        potNx=pot, potNy=pot, by=ax, bandlimittrans=1, Cs=1.0e-3 * 1e7, df0=-5.0 * 10.0, alpha=25.0,
        detectors=[Detector(0.0, 10.0, "BF"), Detector(11.0, 22.0, "ABF"), Detector(70.0, 200.0, "ADF")],
    )
    stopx = f32(x0 + f32(dx) * f32(rng_px))
    ipos = scan_offsets(x0, stopx, 64, x0, stopx, 64, dx, dx, pot, pot, nx, nx)
    potoff = float(x0 - 0.5 * nx * dx)
    return cfg, ipos, potoff


def uneven_workload(nconfig=16, steps=3, streams=0):
    """Uneven batch with frozen phonons: seeded amorphous carbon slab 80 x 80 x 100 A
    at 2.0 g/cm^3 (1.4 A minimum separation) with a 2 nm gold FCC sphere (a = 4.08 A)
    at its centre; nconfig Gaussian phonon configurations (sigma = 0.08 A); 100 slices
    of 1 A; 200 kV, 20 mrad; dx = 0.08 A on a 1088^2 potential grid; windows 256^2,
    512^2 and 1024^2 in the ratio 1/2 : 1/3 : 1/6 (6, 4 and 2 positions per
    configuration), all advanced by one qs_run_groups call (streams: 0 = library
    default, else 1 or 2 HIP streams).  Returns the record."""
    import torch
    from qstem_amd import Detector, Multislice, StemConfig, record_bins, run_groups
    from qstem_amd.specimen import carbon_gold_specimen, gaussian_phonons
    t0 = time.perf_counter()
    atoms = carbon_gold_specimen(seed=1)
    t_spec = time.perf_counter() - t0
    dx, potN, nslice = 0.08, 1088, 100
    dets = [Detector(0.0, 10.0, "BF"), Detector(40.0, 100.0, "ADF")]

    def make(n):
        return Multislice(StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=nslice,
                                     sliceThickness=1.0, potNx=potN, potNy=potN, bandlimittrans=1, Cs=0.0, df0=0.0,
                                     alpha=20.0, detectors=dets))
    big = make(1024)
    big.set_configs(nconfig)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(nconfig):
        big.select_config(k)
        big.build_potential(gaussian_phonons(atoms, 0.08, seed=k + 1), tds=1)
    torch.cuda.synchronize()
    t_pot = time.perf_counter() - t0
    ctxs = {1024: big, 512: make(512), 256: make(256)}
    for n in (512, 256):
        ctxs[n].share_trans(big)
    if streams:
        for ms in ctxs.values():
            ms.set_streams(streams)
    per_cfg = {256: 6, 512: 4, 1024: 2}
    groups = []
    for n, ms in ctxs.items():
        ms.make_probe()
        span = potN - n
        g = np.linspace(0, span, per_cfg[n]).astype(np.int32)
        pos = np.stack([g, g[::-1]], 1)
        ipos = np.tile(pos, (nconfig, 1))
        config = np.repeat(np.arange(nconfig, dtype=np.int32), per_cfg[n])
        groups.append((ms, ipos, config))
    bins, nbins = record_bins(nslice)
    run_groups(groups, bins=bins, nbins=nbins)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = run_groups(groups, bins=bins, nbins=nbins)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    items = sum(len(g[1]) for g in groups)
    # the same items one window size at a time (each context's own qs_run_stem)
    t0 = time.perf_counter()
    for ms, ipos, config in groups:
        ms.run_stem(ipos, bins=bins, nbins=nbins, config=config)
    torch.cuda.synchronize()
    el_seq = time.perf_counter() - t0
    trans_bytes = nconfig * nslice * potN * potN * 8
    return {"metric": "item-slices/s, uneven batch (256^2/512^2/1024^2) x %d phonon configurations" % nconfig,
            "value": round(items * nslice / el, 1), "unit": "item-slices/s", "items": items,
            "items_per_size": {str(n): len(g[1]) for n, g in zip(ctxs, groups)}, "slices": nslice,
            "streams": big.get_streams(),
            "ms_per_batch": round(el * 1e3, 2), "sequential_sizes_ms": round(el_seq * 1e3, 2),
            "atoms": int(len(atoms)), "specimen_s": round(t_spec, 2),
            "potential_build_s_per_config": round(t_pot / nconfig, 3),
            "transmission_GB": round(trans_bytes / 1e9, 2),
            "adf_mean_per_size": {str(n): float(np.mean(o["det"][-1, 1])) for n, o in zip(ctxs, outs)}}


def reference_config_workload(runs=3, cpu_probes=64):
    """The reference's own example, stem3/debug_data/stem_STO_4x4.qsc, through the
    library as qstem3 runs it: charged SrTiO3 (Sr +2, Ti +4, O -2; Debye-Waller factors
    of SrTiO3.cfg) 9 x 9 x 20 cells, 400^2 wave at 0.0625 A (mixed-radix path), 200 kV,
    40 slices of 1.9525 A, 20 x 20 scan over [11.67, 19.71] x [11.45, 19.32] A, four
    annular detectors, outputs every 8 slices, Einstein phonons at 300 K, `runs`
    configurations, each built on the GPU (ion-offset LUTs) and run over all 400
    positions.  Returns probe-slices/s of the multislice (device time) and the wall
    time per configuration including the potential build, and the plain CPU FP32
    multislice (oracle/qs_cpu32.c, one probe per OpenMP thread as doSTEM) on the last
    configuration's transmission functions for cpu_probes of the scan positions."""
    import torch
    from qstem_amd import Detector, Multislice, StemConfig, record_bins, scan_offsets
    from qstem_amd.specimen import build_supercell, perovskite_cell
    f = np.float32
    nx, dx, nslice = 400, 0.0625, 40
    x0, x1, y0, y1 = 11.67, 19.71, 11.45, 19.32
    npx = int((f(x1) - f(x0)) / f(dx))
    npy = int((f(y1) - f(y0)) / f(dx))
    sx1, sy1 = f(f(x0) + f(dx) * npx), f(f(y0) + f(dx) * npy)
    dets = [Detector(70.0, 133.0, "detector1"), Detector(50.0, 70.0, "detector2"), Detector(30.0, 50.0, "detector3"),
            Detector(0.0, 40.0, "detector4")]
    cfg = StemConfig(nx=nx, ny=nx, resolutionX=dx, resolutionY=dx, v0=200.0, slices=nslice, sliceThickness=1.9525,
                     potNx=nx + npx, potNy=nx + npy, by=float(f(35.145)), bandlimittrans=0, Cs=float(f(f(0.05) * 1e7)),
                     df0=float(f(10.0 * f(-13.7))), alpha=15.0, ismoth=1, Cc=1e7, electronScale=1.0, compat=3,
                     detectors=dets)
    cell = perovskite_cell(charges=(2.0, 4.0, -2.0))
    ipos = scan_offsets(x0, sx1, 20, y0, sy1, 20, dx, dx, cfg.potNx, cfg.potNy, nx, nx)
    offx = float(f(np.float64(f(x0)) - 0.5 * nx * np.float64(f(dx))))
    offy = float(f(np.float64(f(y0)) - 0.5 * nx * np.float64(f(dx))))
    bins, nbins = record_bins(nslice, outputInterval=8)
    configs = [build_supercell(cell, 9, 9, 20, tds=True, temperature=300.0, seed=k + 1)[0] for k in range(runs + 1)]
    with Multislice(cfg) as ms:
        def run(atoms):
            ms.build_potential(atoms, atomRadius=5.0, potOffsetX=offx, potOffsetY=offy, czOffset=0.976, tds=1)
            ms.make_probe()
            out = ms.run_stem(ipos, bins=bins, nbins=nbins)
            return ms.last_run_ms, out
        run(configs[0])  # warm-up
        torch.cuda.synchronize()
        dev_ms = []
        t0 = time.perf_counter()
        for atoms in configs[1:]:
            ms_dev, out = run(atoms)
            dev_ms.append(ms_dev)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / runs
        cpu = None
        if cpu_probes > 0:
            trans = ms.get_trans().astype(np.complex128)
            psi0 = ms.make_probe().astype(np.complex128)
            model, ncpu = cpu_info()
            threads = min(ncpu, cpu_probes)
            sel = ipos[np.linspace(0, len(ipos) - 1, cpu_probes).astype(int)]
            rate, cwall, cdet = cpu32_rate(cfg, trans, psi0, sel, bins, nbins, threads)
            ref = ms.run_stem(sel, bins=bins, nbins=nbins)["det"]
            err = float(np.max([np.linalg.norm(cdet[b, d] - ref[b, d]) / max(np.linalg.norm(ref[b, d]), 1e-30)
                                for b in range(nbins) for d in range(len(dets))]))
            cpu = {"value": round(rate, 1), "unit": "probe-slices/s", "cores": threads, "kind": "port",
                   "sample": "%d of the 400 positions x 40 slices (%.2f s wall)" % (cpu_probes, cwall),
                   "cpu_model": model, "max_rel_err_gpu_vs_cpu": err}
    adf = out["det"][2, 0]
    return {"metric": "probe-slices/s, the reference's stem_STO_4x4.qsc (400^2, 20x20 scan, 40 slices)",
            "value": round(len(ipos) * nslice / (np.mean(dev_ms) * 1e-3), 1), "unit": "probe-slices/s",
            "runs": runs, "multislice_ms_per_run": round(float(np.mean(dev_ms)), 2),
            "wall_s_per_run_with_potential": round(wall, 3), "atoms": int(len(configs[0])),
            "adf_70_133_mean_t46.86": round(float(adf.mean()), 4), "cpu_baseline": cpu}


def tem_workload(nslice=200, streams=0):
    """Large TEM with the checkpointed adjoint: plane wave on a 4096^2 grid at
    dx = 390.5/4096 A (100 x 100 perovskite cells), 300 kV, 200 slices of a/2
    (synthetic projected potential of the two perovskite slice types, as in the
    batched workload); forward FP32 with the exit wave, then the adjoint of
    L = sum (I - target)^2 with the library's checkpoint interval (every slice kept
    when the states fit in half the free device memory), and with checkpoints every
    ceil(sqrt(slices)) slices (least memory); streams: 0 = library default, else 1 or 2
    HIP streams (with two, the band limit of dL/dt overlaps the next slice's reverse
    sweeps).  Reports device times, their ratio and the library's peak device memory
    per phase."""
    import torch
    from qstem_amd import Multislice, StemConfig, device_memory
    n, a = 4096, 3.905
    dx = 100 * a / n
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=300.0, slices=nslice, sliceThickness=a / 2,
                     potNx=n, potNy=n, bandlimittrans=1, Cs=0.0, df0=0.0, alpha=20.0, electronScale=1.0,
                     detectors=[])
    x = np.arange(n) * dx
    planes = {0: [(0.0, 0.0, 38), (0.5, 0.5, 8)], 1: [(0.5, 0.5, 22), (0.5, 0.0, 8), (0.0, 0.5, 8)]}
    V = np.empty((nslice, n, n), np.float32)
    for sl, atoms in planes.items():
        P = np.zeros((n, n), np.float32)
        for fx, fy, Z in atoms:
            ddx = (((x - fx * a + a / 2) % a) - a / 2).astype(np.float32)
            ddy = (((x - fy * a + a / 2) % a) - a / 2).astype(np.float32)
            P += np.float32(3.0 * Z ** 0.7) * np.exp(-(ddx[:, None] ** 2 + ddy[None, :] ** 2) / np.float32(2 * 0.25 ** 2))
        V[sl::2] = P
    ipos = np.zeros((1, 2), np.int32)
    with Multislice(cfg) as ms:
        if streams:
            ms.set_streams(streams)
        device_memory(reset_peak=True)
        t0 = time.perf_counter()
        ms.set_potential(V)
        t_pot = time.perf_counter() - t0
        del V
        ms.set_incident_wave(np.ones((n, n), np.complex64))
        _, peak_setup = device_memory(reset_peak=True)
        ms.run_stem(ipos, want_det=False, want_exit=True)  # warm-up
        fw = []
        for _ in range(2):
            ex = ms.run_stem(ipos, want_det=False, want_exit=True)["exit"]
            fw.append(ms.last_run_ms)
        _, peak_fwd = device_memory(reset_peak=True)
        I = np.abs(np.fft.fft2(ex[0].astype(np.complex128)) / (n * n)) ** 2
        target = (0.9 * I).astype(np.float32)[None]
        torch.cuda.synchronize()

        def adjoint_runs(k):
            wall = []
            for _ in range(2):
                t0 = time.perf_counter()
                o = ms.adjoint(ipos, target, ckpt_every=k, want_grad_V=True, want_grad_probe=False)
                torch.cuda.synchronize()
                wall.append(time.perf_counter() - t0)
            return o, ms.last_run_ms, min(wall), ms.last_ckpt_interval

        out, adm, adw, k_def = adjoint_runs(0)
        _, peak_adj = device_memory(reset_peak=True)
        # the same adjoint in its two halves (forward pass kept, then the reverse pass)
        ms.forward_save(ipos, target)
        fsm = ms.last_run_ms
        ms.adjoint_saved(want_grad_V=True, want_grad_probe=False)
        rvm = ms.last_run_ms
        _, adm_sq, _, k_sq = adjoint_runs(int(np.ceil(np.sqrt(nslice))))
        ms_streams = ms.get_streams()
        ms.forward_save(ipos, target, ckpt_every=k_sq)
        fsm_sq = ms.last_run_ms
        ms.adjoint_saved(want_grad_V=True, want_grad_probe=False)
        rvm_sq = ms.last_run_ms
    f_ms = min(fw)
    return {"workload": "TEM plane wave 4096^2, perovskite 100x100 cells, %d slices, 300 kV" % nslice,
            "forward_ms": round(f_ms, 2), "adjoint_ms": round(adm, 2), "adjoint_over_forward": round(adm / f_ms, 3),
            "forward_save_ms": round(fsm, 2), "reverse_ms": round(rvm, 2),
            "reverse_over_forward": round(rvm / f_ms, 3),
            "adjoint_wall_s": round(adw, 3), "checkpoint_every": k_def, "streams": ms_streams, "loss": out["loss"],
            "sqrt_checkpoints": {"checkpoint_every": k_sq, "adjoint_ms": round(adm_sq, 2),
                                 "adjoint_over_forward": round(adm_sq / f_ms, 3), "forward_save_ms": round(fsm_sq, 2),
                                 "reverse_ms": round(rvm_sq, 2), "reverse_over_forward": round(rvm_sq / f_ms, 3)},
            "exit_norm": float(np.sum(np.abs(ex[0]) ** 2) / (n * n)),
            "peak_device_GB": {"potential_setup": round(peak_setup / 1e9, 2), "forward": round(peak_fwd / 1e9, 2),
                               "adjoint": round(peak_adj / 1e9, 2)},
            "potential_set_s": round(t_pot, 2)}


def adjoint_workload(steps=3):
    """4D-STEM adjoint workload: 32x32 scan, 256x256 wave (dx = 0.05 A), 200 kV, 20 slices of
    seeded Gaussian-filtered random potentials (sigma 0.5 A, max phase 0.3 rad); loss = L2
    between simulated and target diffraction intensities.  Returns the timing record."""
    import torch
    from qstem_amd import Multislice, StemConfig
    n, dx, S = 256, 0.05, 20
    step_px = 4
    pot = n + 32 * step_px
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=S, sliceThickness=2.0,
                     potNx=pot, potNy=pot, bandlimittrans=1, Cs=1.0e4, df0=-50.0, alpha=20.0, electronScale=1.0)
    rng = np.random.default_rng(7)
    noise = rng.standard_normal((S, pot, pot))
    k = np.fft.fftfreq(pot, d=dx)
    g = np.exp(-2 * (np.pi * 0.5) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    V = np.real(np.fft.ifft2(np.fft.fft2(noise, axes=(1, 2)) * g, axes=(1, 2)))
    sigma = (1 + 200.0 / 511.0) * 12.3984244 / np.sqrt(200.0 * (2 * 510.99906 + 200.0))
    V = (V * (0.3 / sigma) / np.abs(V).max()).astype(np.float32)
    ix, iy = np.meshgrid(np.arange(32) * step_px, np.arange(32) * step_px, indexing="ij")
    ipos = np.stack([ix.ravel(), iy.ravel()], 1).astype(np.int32)
    ms = Multislice(cfg)
    ms.set_potential(1.05 * V)
    ms.make_probe()
    d = ms.run_stem(ipos, want_det=False, want_diff=True)["diff"]
    target = (np.fft.ifftshift(d, axes=(1, 2)) / float(n * n) ** 1.5).astype(np.float32)
    ms.set_potential(V)
    ms.make_probe()
    ms.run_stem(ipos, want_det=False, want_diff=True)
    ms.adjoint(ipos, target)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ms.run_stem(ipos, want_det=False, want_diff=True)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / steps
    res, kused = {}, {}
    k_sq = int(np.ceil(np.sqrt(S)))
    for k_ck in (0, k_sq):  # the library's interval, then ceil(sqrt(S)) (least memory)
        t0 = time.perf_counter()
        for _ in range(steps):
            out = ms.adjoint(ipos, target, ckpt_every=k_ck)
        torch.cuda.synchronize()
        res[k_ck] = (time.perf_counter() - t0) / steps
        kused[k_ck] = ms.last_ckpt_interval
    ms.close()
    return {"workload": "4D-STEM 32x32 scan, 256x256, 20 slices, L2 on diffraction intensities",
            "forward_ms": round(fwd * 1e3, 2), "adjoint_ms": round(res[0] * 1e3, 2),
            "adjoint_over_forward": round(res[0] / fwd, 3), "ckpt_every": kused[0],
            "adjoint_ms_sqrt_ckpt": round(res[k_sq] * 1e3, 2), "sqrt_ckpt_every": k_sq,
            "adjoint_over_forward_sqrt_ckpt": round(res[k_sq] / fwd, 3),
            "note": "adjoint = one call returning loss, dL/dV, dL/d(df, Cs): forward with checkpoints, "
                    "segment recompute, reverse sweeps"}


def cpu_config():
    """The CPU configuration: perovskite 4x4 cells lateral, 10 cells thick, 2 slices per
    cell (20 slices), 128^2 wave at dx = 0.040 A (the ADF ring inside the 2/3 band
    limit), 200 kV, 20 mrad, Cs = 0, df = 0, 16x16 scan over one unit cell, ADF
    70-200 mrad.  Returns (cfg, ipos, potential offset)."""
    from qstem_amd import Detector, StemConfig, scan_offsets
    a, dx, nx = 3.905, 0.040, 128
    f32 = np.float32
    x0, x1 = f32(2 * a), f32(3 * a)
    rng_px = int(f32(x1 - x0) / f32(dx))
    pot = nx + rng_px
    cfg = StemConfig(nx=nx, ny=nx, resolutionX=dx, resolutionY=dx, v0=200.0, slices=20, sliceThickness=a / 2,
                     potNx=pot, potNy=pot, by=4 * a, bandlimittrans=1, Cs=0.0, df0=0.0, alpha=20.0,
                     detectors=[Detector(70.0, 200.0, "ADF")])
    stopx = f32(x0 + f32(dx) * f32(rng_px))
    ipos = scan_offsets(x0, stopx, 16, x0, stopx, 16, dx, dx, pot, pot, nx, nx)
    return cfg, ipos, float(x0 - 0.5 * nx * dx)


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu32_rate(cfg, trans, psi0, ipos, bins, nbins, threads):
    """Plain CPU FP32 multislice (one probe per OpenMP thread) over ipos: probe-slices/s."""
    from oracle import oracle
    oracle.cpu32_multislice(cfg, trans, psi0, ipos[:threads], bins=bins, nbins=nbins, nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    _, det = oracle.cpu32_multislice(cfg, trans, psi0, ipos, bins=bins, nbins=nbins, nthreads=threads)
    wall = time.perf_counter() - t0
    return len(ipos) * trans.shape[0] / wall, wall, det


def _oracle_worker(args):
    cfg_d, trans_path, probe_path, ipos, bins, nbins = args
    sys.path.insert(0, ROOT)
    from oracle import oracle
    from qstem_amd import Detector, StemConfig
    dets = [Detector(**d) for d in cfg_d.pop("detectors")]
    cfg = StemConfig(detectors=dets, **cfg_d)
    trans = np.load(trans_path, mmap_mode="r").astype(np.complex128)
    psi0 = np.load(probe_path).astype(np.complex128)
    t0 = time.perf_counter()
    _, det, _ = oracle.multislice(cfg, trans, ipos, psi0, bins=bins, nbins=nbins)
    return det, time.perf_counter() - t0


def oracle_sample(cfg, trans, psi0, ipos, bins, nbins, nsample, workdir):
    """FP64 oracle on a bounded sample (the parity check of the timed path): one probe
    through all slices per worker process."""
    import dataclasses
    from multiprocessing import get_context
    os.makedirs(workdir, exist_ok=True)
    tp = os.path.join(workdir, "trans.npy")
    pp = os.path.join(workdir, "probe.npy")
    np.save(tp, trans)
    np.save(pp, psi0)
    cfg_d = dataclasses.asdict(cfg)
    jobs = [(dict(cfg_d), tp, pp, ipos[i:i + 1], bins, nbins) for i in range(nsample)]
    t0 = time.perf_counter()
    with get_context("spawn").Pool(nsample) as pool:
        res = pool.map(_oracle_worker, jobs)
    wall = time.perf_counter() - t0
    dets = np.concatenate([r[0] for r in res], axis=2)
    return dets, wall, [r[1] for r in res]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256, help="probes per qs_run_stem call (one step)")
    ap.add_argument("--group", type=int, default=0, help="probes advanced together inside the library (0 = auto)")
    ap.add_argument("--loop-probes", type=int, default=8, help="probes for the per-probe loop baseline")
    ap.add_argument("--cpu-sample", type=int, default=32, help="probes for the FP32 CPU baseline (0 = skip)")
    ap.add_argument("--oracle-sample", type=int, default=4, help="probes checked against the FP64 oracle (0 = skip)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the uneven-batch and TEM records")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches only")
    ap.add_argument("--tem-slices", type=int, default=200, help="slices of the tem workload")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "uneven", "tem", "reference", "copy", "copyrate"],
                    help="batched: the headline line; uneven: frozen-phonon uneven batch; tem: 4096^2 plane wave "
                         "with the checkpointed adjoint; reference: stem_STO_4x4.qsc's 400^2 scan (mixed-radix "
                         "path; --cpu-sample positions for the CPU leg, 0 = skip) (own JSON lines); copy: a plain "
                         "512 MiB device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters); copyrate: plain "
                         "device copy rates")
    ap.add_argument("--rows-per-wave", type=int, default=0, help="with --group: row-sweep steps per wavefront")
    ap.add_argument("--tiles-per-block", type=int, default=0, help="with --group: column tiles per block")
    ap.add_argument("--streams", type=int, default=0, choices=[0, 1, 2],
                    help="HIP streams for the probe groups of the timed runs (0 = library default)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); got --gpus %d" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    if a.workload == "reference":
        print(json.dumps(reference_config_workload(cpu_probes=a.cpu_sample)), flush=True)
        return
    if a.workload == "uneven":
        print(json.dumps(uneven_workload(streams=a.streams)), flush=True)
        return
    if a.workload == "copy":  # counter calibration: a plain 512 MiB device-to-device copy, 3 times
        import torch
        x = torch.ones(2 ** 27, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_per_copy": 2 ** 29}), flush=True)
        return
    if a.workload == "copyrate":
        # plain device-to-device copy rate (bytes read + written per second) at 512 MiB
        # and at the 184.5 MB of one 128-probe 512^2 state, and an in-place pass (read and
        # write the same buffer, as the sweeps do) over the 184.5 MB, which the 256 MB
        # Infinity Cache holds
        import torch
        x = torch.ones(2 ** 27, device="cuda")
        y = torch.empty_like(x)
        rates = {}
        state_bytes = 128 * 512 * 352 * 8
        for name, nbytes, op in (("copy_536870912", 2 ** 29, "copy"), ("copy_%d" % state_bytes, state_bytes, "copy"),
                                 ("inplace_%d" % state_bytes, state_bytes, "inplace")):
            xs, ys = x[: nbytes // 4], y[: nbytes // 4]
            run = (lambda: ys.copy_(xs)) if op == "copy" else (lambda: xs.mul_(1.0))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            for _ in range(3):
                run()
            e0.record()
            for _ in range(20):
                run()
            e1.record()
            torch.cuda.synchronize()
            rates[name] = round(2 * nbytes * 20 / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
        print(json.dumps({"workload": "copyrate", "GBps_read_plus_write": rates}), flush=True)
        return
    if a.workload == "tem":
        print(json.dumps(tem_workload(nslice=a.tem_slices, streams=a.streams)), flush=True)
        return

    from qstem_amd import Multislice, record_bins
    from qstem_amd.specimen import build_supercell, perovskite_cell
    import torch

    cfg, ipos_all, potoff = perovskite_config()
    P = a.batch
    nslice = cfg.slices
    bins, nbins = record_bins(nslice)
    atoms, _ = build_supercell(perovskite_cell(), 12, 12, 40)
    ms = Multislice(cfg)
    ms.build_potential(atoms[:64], potOffsetX=potoff, potOffsetY=potoff)  # warm-up (LUT kernels, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms.build_potential(atoms, potOffsetX=potoff, potOffsetY=potoff)
    torch.cuda.synchronize()
    t_pot = time.perf_counter() - t0
    t0 = time.perf_counter()
    psi0 = ms.make_probe()
    t_probe = time.perf_counter() - t0
    if a.streams:
        ms.set_streams(a.streams)
    streams = ms.get_streams()
    ms.set_batch(a.group)
    ms.set_tuning(a.rows_per_wave, a.tiles_per_block)
    tuning = None
    if a.group == 0:  # autotuner picks group size and sweep launch shapes (not timed)
        t0 = time.perf_counter()
        tuning = ms.autotune(ipos_all[:512])
        tuning["seconds"] = round(time.perf_counter() - t0, 3)

    nbatches = (ipos_all.shape[0] + P - 1) // P

    def step(i):
        sl = ipos_all[(i % nbatches) * P:((i % nbatches) + 1) * P]
        return ms.run_stem(sl, bins=bins, nbins=nbins)

    def timed(graph, profile):
        """K steps between two device synchronisations; eager launches with per-launch
        HIP events (kernel durations for the roofline), or the captured HIP graph."""
        ms.set_graph(graph)
        for i in range(a.warmup if not graph else max(a.warmup, 1)):  # graph: capture on the first call
            step(i)
        torch.cuda.synchronize()
        if profile:
            ms.set_profiling(True)
        t0 = time.perf_counter()
        for i in range(a.steps):
            step(i)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        kt = ms.kernel_times() if profile else None
        ms.set_profiling(False)
        ms.set_graph(False)
        return t1 - t0, kt

    # kernel durations for the roofline: eager launches on one stream, so that no other
    # launch overlaps the timed ones (with two streams, two groups' sweeps share the GPU)
    # and the library's default group size for one stream
    tuned = ms.get_tuning()
    ms.set_streams(1)
    ms.set_batch(0)
    _, (ms_rows, ms_cols, n_rows, n_cols) = timed(graph=False, profile=True)
    ms.set_streams(streams)
    ms.set_batch(tuned[2])
    el_eager, _ = timed(graph=False, profile=False)
    el_graph, _ = timed(graph=True, profile=False) if not a.no_graph else (float("inf"), None)
    elapsed = min(el_eager, el_graph)
    probe_slices = a.steps * P * nslice
    value = probe_slices / elapsed

    # per-probe GPU loop: same kernels, one probe per launch sequence on one stream
    # (library default launch shapes, which suit one probe per launch)
    ms.set_streams(1)
    ms.set_batch(1)
    ms.set_tuning(0, 0)
    sl = ipos_all[:a.loop_probes]
    ms.run_stem(sl[:1], bins=bins, nbins=nbins)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop_out = ms.run_stem(sl, bins=bins, nbins=nbins)
    torch.cuda.synchronize()
    loop_rate = a.loop_probes * nslice / (time.perf_counter() - t0)
    ms.set_tuning(tuned[0], tuned[1])
    ms.set_batch(tuned[2])
    ms.set_streams(streams)

    # roofline of the slice step (row sweep + column sweep, one logical kernel in two launches)
    N2 = cfg.nx * cfg.ny
    W = cfg.potNx * cfg.potNy
    ndet = len(cfg.detectors)
    t_step_ms = ms_rows / max(n_rows, 1) + ms_cols / max(n_cols, 1)
    G = n_rows and round(P * a.steps * nslice / n_rows)  # probes per kernel launch inside the library
    # algorithmic bytes: wave in + out, the union transmission window read once per launch, detector sums
    bytes_per_probe_slice = 16.0 * N2 + 8.0 * W / G + 4.0 * ndet
    achieved = bytes_per_probe_slice * G / (t_step_ms * 1e-3) / 1e9
    # HBM traffic needs the profiler's counters, which this process cannot read about
    # itself: null here, measured with rocprofv3 --pmc in a separate run (profiles/)
    traffic = None

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "probe-slices/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 chi)",
        "data": "synthetic: perovskite supercell atoms (neutral Sr/Ti/O, DW of SrTiO3.cfg), potential built on "
                "the GPU",
        "config": {"workload": "batched plain STEM, perovskite 12x12x40 cells, 80 slices, 512x512 wave, "
                               "dx=0.030 A, 300 kV, 25 mrad, Cs=1 um, df=-5 nm, 64x64 scan, BF/ABF/ADF",
                   "probes_per_call": P, "probes_per_kernel_launch": tuned[2] or G, "streams": streams,
                   "slices": nslice, "nx": cfg.nx, "potN": cfg.potNx},
        "launch_modes": {"eager_events": round(probe_slices / el_eager, 1),
                         "hip_graph": round(probe_slices / el_graph, 1) if el_graph < float("inf") else None,
                         "value_from": "hip_graph" if el_graph < el_eager else "eager_events"},
        "per_probe_loop": {"value": round(loop_rate, 1), "unit": "probe-slices/s", "probes": a.loop_probes},
        "batched_over_per_probe": round(value / loop_rate, 2),
        "kernel_ms": {"rows_sweep": round(ms_rows / max(n_rows, 1), 4), "cols_sweep": round(ms_cols / max(n_cols, 1), 4)},
        "setup_s": {"potential_and_transmission": round(t_pot, 3), "atoms": int(len(atoms)),
                    "probe": round(t_probe, 4)},
        "tuning": tuning,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "slice step = k_sweep_rows<512,1> + k_sweep_cols<512>",
                     "launches": "eager, one stream, %d probes per launch (HIP events per launch)" % G,
                     "bytes_per_probe_slice": bytes_per_probe_slice},
    }

    if not a.no_adjoint:
        result["adjoint"] = adjoint_workload()

    trans = ms.get_trans()
    if a.oracle_sample > 0:
        ns = min(a.oracle_sample, len(os.sched_getaffinity(0)), 16)
        import shutil
        import tempfile
        tmpdir = tempfile.mkdtemp(prefix="qs_bench_")
        try:
            det_o, _, _ = oracle_sample(cfg, trans, psi0, ipos_all[:ns], bins, nbins, ns, tmpdir)
        finally:
            shutil.rmtree(tmpdir, ignore_errors=True)
        det_g = ms.run_stem(ipos_all[:ns], bins=bins, nbins=nbins)["det"]
        errs = []
        for b in range(nbins):
            for d in range(ndet):
                ref = det_o[b, d]
                if np.any(ref != 0):
                    errs.append(float(np.linalg.norm(det_g[b, d] - ref) / np.linalg.norm(ref)))
        result["max_rel_err_vs_oracle"] = max(errs) if errs else None
        result["oracle_sample_probes"] = ns

    if a.cpu_sample > 0:
        model, ncpu = cpu_info()
        threads = min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu)))
        nsamp = max(a.cpu_sample, threads)
        rate, wall, det_c = cpu32_rate(cfg, trans, psi0, ipos_all[:nsamp], bins, nbins, threads)
        det_g = ms.run_stem(ipos_all[:nsamp], bins=bins, nbins=nbins)["det"]
        cerr = max(float(np.linalg.norm(det_g[b, d] - det_c[b, d]) / np.linalg.norm(det_c[b, d]))
                   for b in range(nbins) for d in range(ndet) if np.any(det_c[b, d]))
        # the CPU configuration (128^2, 20 slices, 16x16 scan), whole scan
        c1, ip1, off1 = cpu_config()
        m1 = Multislice(c1)
        m1.build_potential(build_supercell(perovskite_cell(), 4, 4, 10)[0], potOffsetX=off1, potOffsetY=off1)
        psi1 = m1.make_probe()
        b1, nb1 = record_bins(c1.slices)
        rate1, wall1, _ = cpu32_rate(c1, m1.get_trans(), psi1, ip1, b1, nb1, threads)
        m1.run_stem(ip1, bins=b1, nbins=nb1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m1.run_stem(ip1, bins=b1, nbins=nb1)
        torch.cuda.synchronize()
        gpu1 = len(ip1) * c1.slices / (time.perf_counter() - t0)
        m1.close()
        result["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "probe-slices/s", "cores": threads, "kind": "port",
            "sample": "%d probes x %d slices of this workload (%.1f s wall), plain FP32 multislice "
                      "(oracle/qs_cpu32.c: own Stockham FFT, one probe per OpenMP thread as doSTEM)" % (nsamp, nslice, wall),
            "cpu_model": model, "host_cpus": ncpu, "max_rel_err_gpu_vs_cpu": cerr,
            "cpu_config": {"workload": "128^2, 20 slices, 16x16 scan, perovskite 4x4x10 cells, 200 kV, ADF",
                           "value": round(rate1, 1), "wall_s": round(wall1, 3),
                           "gpu_same_config": round(gpu1, 1), "unit": "probe-slices/s"}}

    if not a.no_extra:
        ms.close()
        del trans
        result["uneven"] = uneven_workload(steps=2)
        result["reference_config"] = reference_config_workload()
        result["tem"] = tem_workload(nslice=a.tem_slices)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
