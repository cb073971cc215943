"""Adjoint gradients of a per-element material parameter field (no counterpart in the reference, whose
solvers have no adjoint; the per-element parameters follow ew/sections/base/sectionbase.py:86-128).

For an objective J(U) evaluated at equilibrium R(U, theta) = 0 (R = P + PExt, the Newton residual):

    K lambda = dJ/dU  on the free dofs, lambda = 0 on Dirichlet dofs      (K = dF_int/dU, symmetric)
    dJ/dtheta_e = lambda^T dR/dtheta_e                                     (ewfe_adjoint_param_grad)

``EquilibriumDisplacement`` wraps this as a torch.autograd.Function: forward runs the Newton solve for the
given parameter field and returns the converged displacement dof vector; backward returns dJ/dtheta.
Supported: C3D8TL + Neo-Hooke Wa (theta = mu_e) and C3D8 + linear elasticity (theta = E_e).
"""

from __future__ import annotations

import numpy as np

from ._lib import check, lib, ptr


def parameter_gradient(solver, model, dJdU, rtol=1e-12, maxit=100000):
    """dJ/dtheta_e (device tensor, element order) at the converged state of `solver` / `model`."""
    import torch

    asm = solver.assembly
    dm = solver.theDofManager
    U = model.U
    # tangent at the converged state (the hyperelastic / elastic tangent depends on U only)
    P, F, zero = dm.constructDofVector(), dm.constructDofVector(), dm.constructDofVector()
    K = asm.newValues("fp64")
    asm.evaluate(U, zero, P, F, K)
    rows = solver._allRows32(solver._activeDirichlets) if solver._activeDirichlets else None
    if rows is not None:
        asm.dirichletRows(rows, K)
    b = dJdU.detach().to(torch.float64).contiguous().clone()
    if rows is not None:
        b[rows.long()] = 0.0
    lam = dm.constructDofVector()
    # the forward solve's linear solver: the multigrid-preconditioned CG where the solver chose it
    amg = getattr(solver, "_linsolver", "pcg") == "amg"
    it, rel = asm.pcg(K, b, lam, rows, rtol=rtol, maxit=maxit, check_every=4 if amg else 32,
                      precond="amg" if amg else "jacobi")
    grad = torch.empty(len(asm.block), dtype=torch.float64, device=asm.device)
    mp = asm.matParams
    check(lib().ewfe_adjoint_param_grad(asm._plan, asm.matKind, ptr(mp), ptr(asm.elParams), ptr(asm.X), ptr(U),
                                        ptr(lam), ptr(grad), asm._stream()))
    return grad, lam, (it, rel)


class EquilibriumDisplacement:
    """``U = EquilibriumDisplacement.apply(theta, run)`` with ``run(theta_numpy) -> (solver, model)`` performing
    the converged forward solve; differentiable w.r.t. theta (float64 tensor, one value per element)."""

    _fn = None

    @classmethod
    def apply(cls, theta, run):
        if cls._fn is None:
            import torch

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, runner):
                    solver, model = runner(th.detach().cpu().numpy())
                    ctx.solver, ctx.model = solver, model
                    return model.U.detach().clone()

                @staticmethod
                def backward(ctx, gU):
                    grad, _, _ = parameter_gradient(ctx.solver, ctx.model, gU)
                    return grad.to(gU.device), None

            cls._fn = _F
        return cls._fn.apply(theta, run)


def update_element_parameters(assembly, theta):
    """Replace the per-element parameter field of an existing assembly (device copy, element order)."""
    import torch

    t = torch.as_tensor(np.ascontiguousarray(theta, dtype=np.float64), device=assembly.device)
    if assembly.elParams is None or assembly.elParams.numel() != t.numel():
        assembly.elParams = t
    else:
        assembly.elParams.copy_(t)
    assembly.block.elParams = np.asarray(theta, dtype=np.float64)
