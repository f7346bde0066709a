"""Adjoint multislice on the GPU against FP64 central finite differences of the oracle.

Loss: L = sum_p sum_k (|Psi_p(k) P(k)|^2/(nx ny)^2 - target_p(k))^2 after the last slice.
Gradients checked: dL/dV along random directions, dL/d(defocus), dL/d(Cs), dL/d(atom x, y);
FD step h = 1e-6 on the FP64 oracle model (double-precision inputs).  Tolerance: 1e-3
relative for directional derivatives (the GPU adjoint runs in FP32).
"""
import numpy as np
import pytest

from cases import probe_positions, smooth_potential

pytestmark = pytest.mark.gpu


def _cfg(bandlimit=1, nx=128, pad=24, slices=6):
    from qstem_amd import StemConfig
    return StemConfig(nx=nx, ny=nx, resolutionX=0.05, resolutionY=0.05, v0=200.0, slices=slices,
                      sliceThickness=1.9525, potNx=nx + pad, potNy=nx + pad, bandlimittrans=bandlimit, Cs=2.0e4,
                      df0=-60.0, alpha=20.0, electronScale=1.0)


def _target(oracle, cfg, V, ipos):
    """Intensities of a perturbed specimen and defocus: a non-trivial loss landscape."""
    c2 = _cfg(cfg.bandlimittrans, cfg.nx, cfg.potNx - cfg.nx, cfg.slices)
    _, I = oracle.diffraction_loss(c2, oracle.make_trans_d(c2, 1.05 * V.astype(np.float64)), ipos,
                                   oracle.probe_d(c2, -50.0, 2.0e4), 0.0)
    return I.astype(np.float32)


@pytest.mark.parametrize("bandlimit,pad", [(1, 24), (0, 24), (1, 128)])
def test_adjoint_loss_and_gradients_fd(oracle, bandlimit, pad):
    """pad 24: 152^2 transmission grid (band limit by row transforms and transposes);
    pad 128: 256^2 (band limit by band-compact sweeps, gT from three windows)."""
    from qstem_amd import Multislice
    cfg = _cfg(bandlimit, pad=pad)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=31, vmax=20.0)
    ipos = probe_positions(cfg, 3, seed=4)
    target = _target(oracle, cfg, V, ipos)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        out = ms.adjoint(ipos, target)
    Vd = V.astype(np.float64)
    psi0 = oracle.probe_d(cfg, float(np.float32(cfg.df0)), float(np.float32(cfg.Cs)))

    def L(Vx, df=float(np.float32(cfg.df0)), Cs=float(np.float32(cfg.Cs))):
        return oracle.diffraction_loss(cfg, oracle.make_trans_d(cfg, Vx), ipos, oracle.probe_d(cfg, df, Cs),
                                       target)[0]

    L0 = L(Vd)
    assert abs(out["loss"] / L0 - 1) < 1e-4, (out["loss"], L0)
    h = 1e-6
    rng = np.random.default_rng(7)
    for trial in range(3):
        D = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=100 + trial, vmax=1.0).astype(np.float64)
        D *= rng.choice([-1.0, 1.0], size=(cfg.slices, 1, 1))
        fd = (L(Vd + h * D) - L(Vd - h * D)) / (2 * h)
        ad = float(np.sum(out["grad_V"].astype(np.float64) * D))
        assert abs(ad - fd) <= 1e-3 * abs(fd), (trial, ad, fd)
    df0, Cs = float(np.float32(cfg.df0)), float(np.float32(cfg.Cs))
    fd_df = (L(Vd, df0 + h) - L(Vd, df0 - h)) / (2 * h)
    assert abs(out["grad_df0"] - fd_df) <= 1e-3 * abs(fd_df), (out["grad_df0"], fd_df)
    hc = 1e-6 * Cs
    fd_cs = (L(Vd, df0, Cs + hc) - L(Vd, df0, Cs - hc)) / (2 * hc)
    assert abs(out["grad_Cs"] - fd_cs) <= 1e-3 * abs(fd_cs), (out["grad_Cs"], fd_cs)


def test_checkpoint_interval_does_not_change_gradients():
    from qstem_amd import Multislice
    cfg = _cfg(1, slices=7)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=5, vmax=20.0)
    ipos = probe_positions(cfg, 2, seed=2)
    rng = np.random.default_rng(0)
    target = rng.uniform(0, 1e-4, (2, cfg.nx, cfg.ny)).astype(np.float32)
    res = []
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        for k in (1, 2, 0, 7, 9):
            res.append(ms.adjoint(ipos, target, ckpt_every=k))
            # 0: the library's choice, every slice kept when the states fit in memory
            assert ms.last_ckpt_interval == (min(k, cfg.slices) if k > 0 else 1), k
    for r in res[1:]:
        # the last column sweep sums the loss with LDS float atomics: order-dependent rounding
        assert abs(r["loss"] - res[0]["loss"]) <= 1e-6 * abs(res[0]["loss"])
        g0 = res[0]["grad_V"]
        assert np.linalg.norm(r["grad_V"] - g0) <= 1e-6 * np.linalg.norm(g0)  # atomics: order-dependent rounding
        assert abs(r["grad_df0"] - res[0]["grad_df0"]) <= 1e-6 * abs(res[0]["grad_df0"])


def test_atom_gradients_check_the_build():
    """qs_atom_gradients reuses the LUTs and kinds of the last qs_build_potential: a
    different radius, atom count or charge, or a potential set otherwise since, is
    rejected on the host before any launch."""
    from qstem_amd import Multislice, QstemError
    cfg = _cfg(1, nx=64, pad=16, slices=3)
    cfg.sliceThickness = 2.0
    atoms = np.array([[2.0, 2.0, 1.0, 0.5, 1.0, 0.0, 22], [2.8, 1.7, 3.1, 0.5, 1.0, 0.0, 8]])
    with Multislice(cfg) as ms:
        V = ms.build_potential(atoms, atomRadius=2.5, return_V=True)
        ms.make_probe()
        ms.adjoint(np.array([[0, 0]], np.int32), np.zeros((1, 64, 64), np.float32))
        assert ms.atom_gradients(atoms).shape == (2, 3)
        with pytest.raises(QstemError, match="parameters"):
            ms.atom_gradients(atoms, atomRadius=5.0)
        with pytest.raises(QstemError, match="count"):
            ms.atom_gradients(atoms[:1])
        q = atoms.copy()
        q[0, 5] = 1.0
        with pytest.raises(QstemError, match="charge"):
            ms.atom_gradients(q)
        ms.set_potential(V)
        with pytest.raises(QstemError, match="qs_build_potential"):
            ms.atom_gradients(atoms)


def test_atom_position_gradients_fd(oracle):
    from qstem_amd import Multislice
    cfg = _cfg(1, nx=128, pad=24, slices=4)
    cfg.sliceThickness = 2.0
    rng = np.random.default_rng(11)
    n_atoms = 6
    xyz = np.column_stack([rng.uniform(2.5, 5.0, n_atoms), rng.uniform(2.5, 5.0, n_atoms),
                           rng.uniform(0.5, 7.5, n_atoms)])
    Z = np.array([8, 22, 38, 8, 22, 38])
    atoms = np.column_stack([xyz, np.full(n_atoms, 0.5), np.ones(n_atoms), np.zeros(n_atoms), Z])
    atoms = atoms[np.argsort(atoms[:, 2], kind="stable")]
    ipos = np.array([[0, 0], [12, 8], [24, 24]], np.int32)
    R = 2.5
    with Multislice(cfg) as ms:
        V32 = ms.build_potential(atoms, atomRadius=R, return_V=True)
        ms.make_probe()
        target = np.zeros((3, cfg.nx, cfg.ny), np.float32)
        out = ms.adjoint(ipos, target)
        g = ms.atom_gradients(atoms, atomRadius=R)
    gO = oracle.lut_geometry(R, cfg.resolutionX, cfg.sliceThickness)
    keys = sorted(set(int(z) for z in atoms[:, 6]))
    luts = oracle.atom_luts(gO, keys, [0.5] * len(keys), cfg.sliceThickness)
    kinds = [keys.index(int(z)) for z in atoms[:, 6]]
    psi0 = oracle.probe(cfg)

    def L(pos):
        V = oracle.make_slices_d(cfg, gO, luts, kinds, pos, atoms[:, 6].astype(int), 0.0, 0.0, atomRadius=R)
        return oracle.diffraction_loss(cfg, oracle.make_trans_d(cfg, V), ipos, psi0, target)[0]

    pos0 = atoms[:, :3].astype(np.float32).astype(np.float64)
    V0 = oracle.make_slices_d(cfg, gO, luts, kinds, pos0, atoms[:, 6].astype(int), 0.0, 0.0, atomRadius=R)
    assert np.linalg.norm(V0 - V32) <= 1e-4 * np.linalg.norm(V0)
    h = 1e-6
    for ia in range(n_atoms):
        for d in (0, 1):
            pp = pos0.copy()
            pm = pos0.copy()
            pp[ia, d] += h
            pm[ia, d] -= h
            fd = (L(pp) - L(pm)) / (2 * h)
            assert abs(g[ia, d] - fd) <= 2e-3 * abs(fd) + 1e-6 * np.abs(g).max(), (ia, d, g[ia, d], fd)
        assert g[ia, 2] == 0.0


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_torch_autograd_function(oracle, device):
    """torch.autograd.Function wrapper: gradients equal the C-ABI adjoint's; a gradient
    step lowers the loss.  With V and the target in GPU memory the library reads them in
    place and writes dL/dV into a tensor on that device."""
    import torch
    from qstem_amd import Multislice
    from qstem_amd.autograd import multislice_loss
    cfg = _cfg(1, slices=4)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=3, vmax=20.0)
    ipos = probe_positions(cfg, 2, seed=1)
    target_np = _target(oracle, cfg, V, ipos)
    target = torch.tensor(target_np, device=device) if device == "cuda" else target_np
    with Multislice(cfg) as ms:
        Vt = torch.tensor(V, requires_grad=True, device=device)
        df = torch.tensor(float(np.float32(cfg.df0)), dtype=torch.float64, requires_grad=True)
        cs = torch.tensor(float(np.float32(cfg.Cs)), dtype=torch.float64, requires_grad=True)
        loss = multislice_loss(Vt, df, cs, ms, ipos, target)
        loss.backward()
        ms.set_potential(V)
        ms.make_probe()
        ref = ms.adjoint(ipos, target_np)
        assert abs(float(loss.detach()) - ref["loss"]) <= 1e-6 * ref["loss"]
        assert Vt.grad.device == Vt.device
        assert np.allclose(Vt.grad.cpu().numpy(), ref["grad_V"], rtol=1e-5, atol=1e-6 * np.abs(ref["grad_V"]).max())
        assert abs(float(df.grad) - ref["grad_df0"]) <= 1e-5 * abs(ref["grad_df0"])
        with torch.no_grad():
            step = 0.1 * float(loss) / float((Vt.grad ** 2).sum())
            V2 = (Vt - step * Vt.grad).detach()
        loss2 = multislice_loss(V2, df.detach(), cs.detach(), ms, ipos, target)
        assert float(loss2) < float(loss)
        # only V needs a gradient: the reverse pass skips the probe gradients
        V3 = torch.tensor(V, requires_grad=True, device=device)
        multislice_loss(V3, df.detach(), cs.detach(), ms, ipos, target).backward()
        assert np.allclose(V3.grad.cpu().numpy(), ref["grad_V"], rtol=1e-5, atol=1e-6 * np.abs(ref["grad_V"]).max())


def test_forward_save_then_adjoint_saved(oracle):
    """The two halves give qs_adjoint's loss and gradients, also with a qs_run_stem in
    between; the kept forward pass is consumed by the reverse pass and dropped when the
    potential changes."""
    from qstem_amd import Multislice, QstemError
    cfg = _cfg(1, slices=5)
    V = smooth_potential(cfg.slices, cfg.potNx, cfg.potNy, seed=9, vmax=20.0)
    ipos = probe_positions(cfg, 3, seed=2)
    target = _target(oracle, cfg, V, ipos)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        ref = ms.adjoint(ipos, target, ckpt_every=2)
        loss = ms.forward_save(ipos, target, ckpt_every=2)
        # same kernels in the same order; float atomics may reorder the last bits
        assert abs(loss - ref["loss"]) <= 1e-6 * ref["loss"]
        ms.set_graph(True)  # a captured run in between leaves the kept pass alone
        ms.run_stem(probe_positions(cfg, 2, seed=7), want_det=False, want_exit=True)
        out = ms.adjoint_saved()
        assert np.allclose(out["grad_V"], ref["grad_V"], rtol=1e-5, atol=1e-6 * np.abs(ref["grad_V"]).max())
        assert abs(out["grad_df0"] - ref["grad_df0"]) <= 1e-5 * abs(ref["grad_df0"])
        assert abs(out["grad_Cs"] - ref["grad_Cs"]) <= 1e-5 * abs(ref["grad_Cs"])
        with pytest.raises(QstemError, match="no saved forward pass"):
            ms.adjoint_saved()
        ms.forward_save(ipos, target)
        ms.set_potential(V)
        with pytest.raises(QstemError, match="no saved forward pass"):
            ms.adjoint_saved(want_grad_probe=False)


@pytest.mark.parametrize("nx", [512, 1024, 2048, 4096])
def test_adjoint_plane_wave_large(oracle, nx):
    """Plane-wave (TEM) adjoint at the large transform sizes: loss and a directional
    dL/dV against FP64 central differences of the oracle (2 slices, checkpoint every slice)."""
    from qstem_amd import Multislice, StemConfig
    print("plane-wave adjoint", nx, flush=True)
    cfg = StemConfig(nx=nx, ny=nx, resolutionX=0.05, resolutionY=0.05, v0=300.0, slices=2, sliceThickness=1.9525,
                     potNx=nx, potNy=nx, bandlimittrans=1, Cs=0.0, df0=0.0, alpha=20.0, electronScale=1.0)
    # a strong phase object (up to ~2.5 rad per slice): dL/dV is Im(g conj(u)) of nearly
    # parallel factors for a weak one, which FP32 resolves poorly at these sizes
    V = smooth_potential(2, nx, nx, seed=61, vmax=80.0, sigma_px=4.0)
    ipos = np.zeros((1, 2), np.int32)
    psi = np.ones((nx, nx), np.complex64)
    psi_d = psi.astype(np.complex128)
    # target 0 (L = sum I^2): with a plane wave the unscattered beam holds almost all
    # intensity, and a target close to it would leave the loss a difference of nearly
    # equal FP32 numbers; without a target the check is well conditioned
    target = np.zeros((1, nx, nx), np.float32)
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.set_incident_wave(psi)
        out = ms.adjoint(ipos, target, ckpt_every=1, want_grad_probe=False)

    def L(Vx):
        return oracle.diffraction_loss(cfg, oracle.make_trans_d(cfg, Vx), ipos, psi_d, target)[0]

    Vd = V.astype(np.float64)
    L0 = L(Vd)
    assert abs(out["loss"] / L0 - 1) < 1e-4, (out["loss"], L0)
    h = 1e-6
    from scipy.ndimage import gaussian_filter
    g = out["grad_V"].astype(np.float64)
    # along the (smoothed) gradient itself: well conditioned, 1e-3
    Dg = np.stack([gaussian_filter(g[k], 3.0, mode="wrap") for k in range(2)])
    Dg /= np.abs(Dg).max()
    fd = (L(Vd + h * Dg) - L(Vd - h * Dg)) / (2 * h)
    ad = float(np.sum(g * Dg))
    assert abs(ad - fd) <= 1e-3 * abs(fd), (ad, fd)
    # a random smooth direction nearly orthogonal to the gradient: the inner product is
    # a small residue of the FP32 gradient, so its tolerance is wider (5e-3)
    D = smooth_potential(2, nx, nx, seed=62, vmax=1.0, sigma_px=4.0).astype(np.float64)
    fd = (L(Vd + h * D) - L(Vd - h * D)) / (2 * h)
    ad = float(np.sum(g * D))
    print("random direction: adjoint", ad, "central difference", fd, flush=True)
    assert abs(ad - fd) <= 5e-3 * abs(fd), (ad, fd)


@pytest.mark.parametrize("nx,slices", [(256, 5), (4096, 3)])
def test_band_limit_overlap_keeps_gradients(nx, slices):
    """With two streams the band limit of dL/dt of slice s runs beside the reverse
    sweeps of slice s-1 (two dL/dt buffers in turn); one stream runs them in order.
    Loss and dL/dV agree (dL/dt adds in an order set by atomics only where windows
    overlap; here one window covers the slice, so they are equal to rounding).  At
    4096^2 the two modes also take the two variants of the reverse column sweep
    (registers staged directly beside the overlap, through LDS without it)."""
    from qstem_amd import Multislice, StemConfig
    cfg = StemConfig(nx=nx, ny=nx, resolutionX=0.05, resolutionY=0.05, v0=300.0, slices=slices,
                     sliceThickness=1.9525, potNx=nx, potNy=nx, bandlimittrans=1, Cs=0.0, df0=0.0, alpha=20.0,
                     electronScale=1.0)
    V = smooth_potential(slices, nx, nx, seed=63, vmax=40.0, sigma_px=4.0)
    ipos = np.zeros((1, 2), np.int32)
    target = np.zeros((1, nx, nx), np.float32)
    outs = {}
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.set_incident_wave(np.ones((nx, nx), np.complex64))
        for ns in (1, 2):
            ms.set_streams(ns)
            for k in (1, 2):  # every slice kept, and a recomputed segment
                outs[(ns, k)] = ms.adjoint(ipos, target, ckpt_every=k, want_grad_probe=False)
    ref = outs[(1, 1)]
    for key, o in outs.items():
        assert abs(o["loss"] / ref["loss"] - 1) < 1e-6, key
        g0 = ref["grad_V"]
        assert np.linalg.norm(o["grad_V"] - g0) <= 1e-6 * np.linalg.norm(g0), key


def test_atoms_autograd_function():
    """atoms_multislice_loss: dL/d(x, y) through torch autograd equals qs_atom_gradients
    after qs_adjoint on the same build, and a small step along -grad lowers the loss
    (the target is the pattern of the atoms shifted by 0.05 A)."""
    import torch
    from qstem_amd import Multislice
    from qstem_amd.autograd import atoms_multislice_loss
    cfg = _cfg(1, nx=128, pad=24, slices=4)
    cfg.sliceThickness = 2.0
    rng = np.random.default_rng(5)
    n_atoms = 8
    xyz = np.column_stack([rng.uniform(2.0, 5.5, n_atoms), rng.uniform(2.0, 5.5, n_atoms),
                           rng.uniform(0.5, 7.5, n_atoms)])
    Z = np.array([8, 22, 38, 8, 22, 38, 8, 22])
    atoms = np.column_stack([xyz, np.full(n_atoms, 0.5), np.ones(n_atoms), np.zeros(n_atoms), Z])
    atoms = atoms[np.argsort(atoms[:, 2], kind="stable")]
    ipos = np.array([[0, 0], [12, 8], [24, 24]], np.int32)
    R = 2.5
    with Multislice(cfg) as ms:
        shifted = atoms.copy()
        shifted[:, 0] += 0.05
        ms.build_potential(shifted, atomRadius=R)
        ms.make_probe()
        ex = ms.run_stem(ipos, want_exit=True, want_det=False)["exit"]
        target = (np.abs(np.fft.fft2(ex.astype(np.complex128), axes=(1, 2)) / (128 * 128)) ** 2).astype(np.float32)
        xy = torch.tensor(atoms[:, :2], dtype=torch.float64, requires_grad=True)
        loss = atoms_multislice_loss(xy, ms, atoms, ipos, target, atomRadius=R)
        loss.backward()
        ms.build_potential(atoms, atomRadius=R)
        ms.make_probe()
        ref = ms.adjoint(ipos, target)
        g_ref = ms.atom_gradients(atoms)[:, :2]
        assert abs(float(loss.detach()) - ref["loss"]) <= 1e-6 * ref["loss"]
        assert np.allclose(xy.grad.numpy(), g_ref, rtol=1e-5, atol=1e-6 * np.abs(g_ref).max())
        step = 0.2 * float(loss.detach()) / float((xy.grad ** 2).sum())
        xy2 = (xy - step * xy.grad).detach()
        loss2 = atoms_multislice_loss(xy2, ms, atoms, ipos, target, atomRadius=R)
        assert float(loss2) < float(loss.detach())


@pytest.mark.parametrize("n,bandlimit", [(2048, 0), (2048, 1), (1024, 0), (512, 1)])
def test_adjoint_repeats_agree(n, bandlimit):
    """Repeats of forward_save + adjoint_saved on one whole-slice window give the same
    gradients: with one probe every dL/dV element has one writer and no atomics, so dL/dV
    repeats bit for bit; dL/d(df0, Cs) end in an FP64 reduction whose atomics may order
    differently.  (Wide stores whose data registers were overwritten before the store read
    them made dL/dV differ between repeats by up to 1.5e-2 at 2048^2.)"""
    from qstem_amd import Multislice, StemConfig
    dx = 0.025
    cfg = StemConfig(nx=n, ny=n, resolutionX=dx, resolutionY=dx, v0=200.0, slices=2, sliceThickness=2.0,
                     potNx=n, potNy=n, bandlimittrans=bandlimit, Cs=1.0e4, df0=-40.0, alpha=25.0,
                     electronScale=1.0)
    V = smooth_potential(cfg.slices, n, n, seed=n + 1, vmax=20.0)
    target = np.random.default_rng(n).random((1, n, n)).astype(np.float32) * 1e-9
    ipos = np.zeros((1, 2), np.int32)
    runs = []
    with Multislice(cfg) as ms:
        ms.set_potential(V)
        ms.make_probe()
        for _ in range(6):
            ms.forward_save(ipos, target)
            runs.append(ms.adjoint_saved(want_grad_V=True, want_grad_probe=True))
    g0 = runs[0]
    assert np.all(np.isfinite(g0["grad_V"])) and np.linalg.norm(g0["grad_V"]) > 0
    for r, g in enumerate(runs[1:], 1):
        assert np.array_equal(g["grad_V"], g0["grad_V"]), (r, float(np.abs(g["grad_V"] - g0["grad_V"]).max()))
        for k in ("grad_df0", "grad_Cs"):
            assert abs(g[k] - g0[k]) <= 1e-12 * abs(g0[k]) + 1e-30, (r, k, g[k], g0[k])
