"""torch.autograd.Function around the GSE-k step (test harness).

Forward: E(xyzq) from gse_gsek_step (energy only).  Backward: one adjoint
sweep, gse_gsek_adjoint, which recomputes the spread and k-pass from xyzq and
returns dE/dr (exact, cube support) and dE/dq, already multiplied by the
upstream gradient on the device.  The gradient with respect to the (n, 4)
xyzq tensor is [dE/dx, dE/dy, dE/dz, dE/dq] per particle.
"""
import torch


class GSEEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyzq, gsek):
        o = gsek.step(xyzq.detach().contiguous(), force=False, potential=False)
        ctx.gsek = gsek
        ctx.save_for_backward(xyzq)
        return o["energy"].reshape(()).clone()

    @staticmethod
    def backward(ctx, grad_e):
        (xyzq,) = ctx.saved_tensors
        a = ctx.gsek.adjoint(xyzq.detach().contiguous(), grad_energy=grad_e.reshape(1),
                             want_virial=False)
        g = torch.empty_like(xyzq)
        g[:, :3] = a["dEdr"].T.to(xyzq.dtype)
        g[:, 3] = a["dEdq"].to(xyzq.dtype)
        return g, None


def energy(xyzq, gsek):
    return GSEEnergy.apply(xyzq, gsek)
