"""torch.autograd.Function over the adjoint multislice (qs_adjoint).

    loss = MultisliceLoss.apply(V, df0, Cs, ms, ipos, target)
    loss.backward()   # V.grad, df0.grad, Cs.grad

V is the projected potential [nslice][potNx][potNy] (any device; it is
moved through host memory to the library, which owns its device buffers),
df0 and Cs are 0-d tensors in Angstrom.  The forward pass evaluates the loss
and the gradients in one adjoint call (the backward pass then only scales
them), which costs one forward plus one reverse sweep of the multislice.
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
        ms.set_potential(V.detach().to("cpu", torch.float32).numpy())
        ms.make_probe()
        tgt = target.detach().to("cpu", torch.float32).numpy() if torch.is_tensor(target) else target
        out = ms.adjoint(np.asarray(ipos), tgt, want_grad_V=True, want_grad_probe=True)
        ctx.save_for_backward(torch.from_numpy(out["grad_V"]).to(V.device),
                              torch.tensor(out["grad_df0"], dtype=torch.float64),
                              torch.tensor(out["grad_Cs"], dtype=torch.float64))
        return torch.tensor(out["loss"], dtype=torch.float64, device=V.device)

    @staticmethod
    def backward(ctx, grad_out):
        gV, gdf, gcs = ctx.saved_tensors
        go = grad_out.to(torch.float64)
        return (gV * go.to(gV.device, gV.dtype), (gdf * go.cpu()).to(torch.float64),
                (gcs * go.cpu()).to(torch.float64), None, None, None)


def multislice_loss(V, df0, Cs, ms, ipos, target):
    """Differentiable sum over probes and pixels of (I - target)^2."""
    if not isinstance(ms, Multislice):
        raise QstemError("ms must be a qstem_amd.Multislice")
    return MultisliceLoss.apply(V, df0, Cs, ms, ipos, target)
