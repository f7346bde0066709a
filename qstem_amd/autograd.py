"""torch.autograd.Function over the adjoint multislice (qs_adjoint).

    loss = MultisliceLoss.apply(V, df0, Cs, ms, ipos, target)
    loss.backward()   # V.grad, df0.grad, Cs.grad

    loss = atoms_multislice_loss(xy, ms, atoms, ipos, target, atomRadius=R)
    loss.backward()   # xy.grad: dL/d(x, y) per atom

V is the projected potential [nslice][potNx][potNy]; a V in GPU memory is read
by the library in place and dL/dV is written into a tensor on V's device (no
copy through host memory); a CPU V goes through host copies.  df0 and Cs are
0-d tensors in Angstrom.  The forward pass runs the forward
half of the adjoint (qs_forward_save: the loss, checkpoints kept by the
context); the backward pass runs the reverse half (qs_adjoint_saved) for the
inputs that need gradients.  Between the two, the context may run other
qs_run_stem calls but must not change its potential, probe or optics.
"""
import numpy as np
import torch

from .multislice import Multislice, QstemError


class MultisliceLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, df0, Cs, ms: Multislice, ipos, target):
        cfg = ms.cfg
        changed = (float(df0) != float(np.float32(cfg.df0))) or (float(Cs) != float(np.float32(cfg.Cs)))
        if changed:
            cfg.df0 = float(df0)
            cfg.Cs = float(Cs)
            ms._check(ms._lib.qs_set_optics(ms._h, cfg.to_c()), "qs_set_optics")
        ms.set_potential(V.detach() if V.is_cuda else V.detach().to(torch.float32).numpy())
        ms.make_probe()
        tgt = target
        if torch.is_tensor(target) and not target.is_cuda:
            tgt = target.detach().to(torch.float32).numpy()
        loss = ms.forward_save(np.asarray(ipos), tgt)
        ctx.ms = ms
        ctx.v_device, ctx.v_dtype = V.device, V.dtype
        return torch.tensor(loss, dtype=torch.float64, device=V.device)

    @staticmethod
    def backward(ctx, grad_out):
        need_v, need_df, need_cs = ctx.needs_input_grad[:3]
        on_device = ctx.v_device.type == "cuda"
        out = ctx.ms.adjoint_saved(want_grad_V=need_v, want_grad_probe=need_df or need_cs,
                                   device=ctx.v_device if on_device else None)
        go = grad_out.detach().to("cpu", torch.float64)
        gV = gdf = gcs = None
        if need_v:
            g = out["grad_V"] if on_device else torch.from_numpy(out["grad_V"])
            gV = g.to(ctx.v_device, ctx.v_dtype) * grad_out.detach().to(ctx.v_device, ctx.v_dtype)
        if need_df:
            gdf = torch.tensor(out["grad_df0"], dtype=torch.float64) * go
        if need_cs:
            gcs = torch.tensor(out["grad_Cs"], dtype=torch.float64) * go
        return gV, gdf, gcs, None, None, None


def multislice_loss(V, df0, Cs, ms, ipos, target):
    """Differentiable sum over probes and pixels of (I - target)^2."""
    if not isinstance(ms, Multislice):
        raise QstemError("ms must be a qstem_amd.Multislice")
    return MultisliceLoss.apply(V, df0, Cs, ms, ipos, target)


class AtomsMultisliceLoss(torch.autograd.Function):
    """The same loss as a function of the atom positions: the forward pass builds the
    potential of the atoms at xy (qs_build_potential, make3DSlices' LUT potential) and
    runs qs_forward_save; the backward pass runs qs_adjoint_saved for dL/dV and
    qs_atom_gradients for dL/d(x, y) (the potential samples z by nearest slice, so z
    gets no gradient)."""

    @staticmethod
    def forward(ctx, xy, ms: Multislice, atoms, ipos, target, pot):
        a = np.array(atoms, dtype=np.float64).reshape(-1, 7)
        a[:, :2] = xy.detach().to("cpu", torch.float64).numpy()
        ms.build_potential(a, **pot)
        ms.make_probe()
        tgt = target
        if torch.is_tensor(target) and not target.is_cuda:
            tgt = target.detach().to(torch.float32).numpy()
        loss = ms.forward_save(np.asarray(ipos), tgt)
        ctx.ms, ctx.atoms, ctx.pot = ms, a, pot
        ctx.xy_device, ctx.xy_dtype = xy.device, xy.dtype
        return torch.tensor(loss, dtype=torch.float64, device=xy.device)

    @staticmethod
    def backward(ctx, grad_out):
        ctx.ms.adjoint_saved(want_grad_V=True, want_grad_probe=False)
        g = ctx.ms.atom_gradients(ctx.atoms, **ctx.pot)[:, :2]
        gxy = torch.from_numpy(np.ascontiguousarray(g, dtype=np.float64)).to(ctx.xy_device, ctx.xy_dtype)
        return gxy * grad_out.detach().to(ctx.xy_device, ctx.xy_dtype), None, None, None, None, None


def atoms_multislice_loss(xy, ms, atoms, ipos, target, **pot):
    """Differentiable loss in the atoms' lateral positions xy [natom][2] (Angstrom);
    atoms: the (n, 7) rows x, y, z, dw, occ, q, Z whose x, y are replaced by xy; pot:
    build_potential's keyword arguments (atomRadius, offsets, ...)."""
    if not isinstance(ms, Multislice):
        raise QstemError("ms must be a qstem_amd.Multislice")
    return AtomsMultisliceLoss.apply(xy, ms, atoms, ipos, target, dict(pot))
