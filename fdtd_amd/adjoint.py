"""torch.autograd binding of the discrete adjoint eps gradient.

``eps_objective(eps, make_grid, box, nsteps, dft)`` returns the objective
J = sum |F|^2 of DFT monitor ``dft`` as a torch scalar whose gradient with
respect to the design permittivities ``eps`` (shape (k, j, i) of the design
box) comes from fdtd_adjoint_eps_gradient.  The forward call already runs the
adjoint (the reverse sweep needs the forward pass's checkpoints), so backward
only scales the stored gradient.  A thin layer for checks and optimisation
loops; the arithmetic is the library's.  No counterpart in the reference.
"""
import torch

from .grid import Grid  # noqa: F401  (make_grid returns one)


class _EpsObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, make_grid, box, nsteps, dft, comp_mask, checkpoint_every, history_mb):
        g = make_grid()
        try:
            g.set_design(box, eps.detach().to("cpu", torch.float64).numpy())
            J, grad = g.adjoint_eps_gradient(nsteps, dft, comp_mask, checkpoint_every, history_mb)
        finally:
            g.close()
        ctx.save_for_backward(torch.from_numpy(grad).to(device=eps.device, dtype=eps.dtype))
        return torch.tensor(J, dtype=eps.dtype, device=eps.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None, None, None, None, None


def eps_objective(eps, make_grid, box, nsteps, dft, comp_mask=4, checkpoint_every=100, history_mb=4096.0):
    """J(eps) with d J / d eps through the adjoint.  make_grid() must return a
    fresh Grid (step 0) with everything but the design set up."""
    return _EpsObjective.apply(eps, make_grid, list(box), int(nsteps), int(dft), int(comp_mask),
                               int(checkpoint_every), float(history_mb))
