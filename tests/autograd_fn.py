"""torch.autograd.Function over the explicit-matrix solve (test harness).

forward: B = (I - diag(rho) F)^-1 E by the mixed-precision gather, refined
to an FP64 residual of 1e-12; backward: the adjoint solve of the C ABI
(rad_lin_adjoint), dL/dE = lambda, dL/drho = lambda o (F B), with the
visibility (F) held fixed.  Tensors are float64 (n, nrhs) on the CPU; the
solves run on the GPU through the library."""
import torch


class RadiositySolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lin, rho, E):
        B, _ = lin.gather(rho.detach().cpu().numpy(), E.detach().cpu().numpy(), prec="f16", tol=1e-12,
                          max_outer=60, sweeps=6)
        Bt = torch.from_numpy(B)
        ctx.lin = lin
        ctx.save_for_backward(rho, Bt)
        return Bt

    @staticmethod
    def backward(ctx, g):
        rho, B = ctx.saved_tensors
        dE, drho, _ = ctx.lin.adjoint(rho.detach().cpu().numpy(), B.numpy(), g.detach().cpu().numpy(),
                                      tol=1e-12, max_outer=60, sweeps=6)
        return None, torch.from_numpy(drho), torch.from_numpy(dE)
