"""Host logic of the autograd wrappers (no GPU): a GMRES solve that stops at
its iteration cap above the tolerance must raise in the forward and in the
backward pass instead of returning an iterate as the solution."""
import numpy as np
import pytest
import torch

from mlfma_amd import autograd as AG


class _CappedPlan:
    """Stands in for a Plan whose solves hit maxit (status NOT_CONVERGED) on
    the adjoint only when forward_ok is set."""

    def __init__(self, n=4, forward_ok=False):
        self.n = n
        self.forward_ok = forward_ok

    def solve(self, b, adjoint=False, **opts):
        ok = self.forward_ok and not adjoint
        rep = {"iterations": opts.get("maxit", 3), "status": "OK" if ok else "ERR_NOT_CONVERGED",
               "resid_est": 1e-9 if ok else 0.4, "resid_true": 1e-9 if ok else 0.4, "seconds": 0.0}
        return np.asarray(b, dtype=np.complex128) * 2.0, rep


def test_forward_solve_not_converged_raises():
    b = torch.ones(4, dtype=torch.complex128, requires_grad=True)
    with pytest.raises(RuntimeError, match="forward GMRES solve did not converge"):
        AG.solve(_CappedPlan(), b, maxit=3)


def test_adjoint_solve_not_converged_raises():
    b = torch.ones(4, dtype=torch.complex128, requires_grad=True)
    x = AG.solve(_CappedPlan(forward_ok=True), b, maxit=3)
    assert torch.allclose(x.detach(), 2 * torch.ones(4, dtype=torch.complex128))
    with pytest.raises(RuntimeError, match="adjoint GMRES solve did not converge"):
        x.abs().pow(2).sum().backward()
