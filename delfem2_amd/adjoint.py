"""Material sensitivities of the static tet solve by the adjoint method.

No reference counterpart (the reference has no adjoint).  With per-element
moduli E_e and K(E) = sum_e E_e K_e^1 (K_e^1 the element stiffness at unit
modulus, fem_solidlinear.cpp:409-435 at the Lame pair of E=1), u = K^-1 f and
an objective J(u):

    dJ/dE_e = -lam_e^T K_e^1 u_e,   K lam = dJ/du  (fixed DOFs zero)

The adjoint solve reuses the forward ILU(0) factors (K is symmetric); the
per-element contraction is the HIP kernel ``dfm2h_adjoint_grad_tet``.  For
the compliance C = f.u, dJ/du = f and lam = u.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from . import solver as S
from .problem import DeviceSystem


def unit_lame(nu: np.ndarray):
    """(myu, lambda) at E = 1, same association as workload.lame."""
    nu = np.asarray(nu, np.float64)
    return 1.0 / (2.0 * (1.0 + nu)), nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


class ElasticSystem(DeviceSystem):
    """A DeviceSystem whose element moduli are a tensor input.

    ``state`` changes whenever the device matrix, right-hand side and factors
    are rebuilt, so a backward pass can tell whether they still hold the
    forward solve's system."""

    def __init__(self, prob, device: int = 0, nu=None):
        super().__init__(prob, device)
        nu = prob.extra.get("nu") if nu is None else nu
        m1, l1 = unit_lame(nu if np.ndim(nu) else np.full(prob.ne, float(nu)))
        self.mu1 = torch.from_numpy(np.ascontiguousarray(m1)).to(self.dev)
        self.lam1 = torch.from_numpy(np.ascontiguousarray(l1)).to(self.dev)
        self.state = 0

    def set_modulus(self, E: torch.Tensor) -> None:
        E = E.detach().to(self.dev, torch.float64).contiguous()
        self.myu = (E * self.mu1).contiguous()
        self.lam = (E * self.lam1).contiguous()

    def assemble(self, disp=None) -> bool:
        self.state += 1
        return super().assemble(disp)

    def factor(self) -> bool:
        self.state += 1
        return super().factor()

    def build(self, E: torch.Tensor) -> None:
        """K(E), f and the ILU(0) factors of K(E) on the device."""
        self.set_modulus(E)
        if not self.assemble():
            raise RuntimeError("element edge missing from the matrix pattern")
        if not self.factor():
            raise RuntimeError("ILU(0) met more than 10 singular pivots")

    def adjoint_grad(self, u: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
        """g_e = -lam_e^T K_e^1 u_e for every element (device kernel)."""
        g = torch.empty(self.prob.ne, dtype=torch.float64, device=self.dev)
        N.call("dfm2h_adjoint_grad_tet", self.mesh.h, S._dptr(self.xyz), S._dptr(self.mu1),
               S._dptr(self.lam1), S._dptr(u.contiguous()), S._dptr(lam.contiguous()),
               S._dptr(g))
        return g


def _converged(hist, tol: float, what: str) -> None:
    """Raises when a Solve_PCG history stops short of the tolerance."""
    if len(hist) > 1 and not hist[-1] < tol * hist[0]:
        raise RuntimeError(f"{what} did not converge: |r|/|r0| = {hist[-1] / hist[0]:.3e} "
                           f"after {len(hist) - 1} iterations (tol {tol:.1e})")


class ElasticTetSolve(torch.autograd.Function):
    """u = K(E)^-1 f on the GPU; backward by one adjoint PCG solve with the
    forward factors and the per-element contraction kernel.  Each solve must
    reach the tolerance (else RuntimeError).  When the system was rebuilt
    for other moduli between forward and backward (another load case, a line
    search), backward rebuilds K(E) and its factors for the forward E first."""

    @staticmethod
    def forward(ctx, E: torch.Tensor, sys: ElasticSystem, tol: float = 1e-10,
                max_nitr: int = 20000):
        sys.build(E)
        u, hist, _ = sys.solve(tol, max_nitr)
        _converged(hist, tol, "forward solve")
        ctx.sys = sys
        ctx.tol = tol
        ctx.max_nitr = max_nitr
        ctx.hist = hist
        ctx.state = sys.state
        ctx.f = sys.b.clone()
        ctx.save_for_backward(u, E.detach().clone())
        return u

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        u, E = ctx.saved_tensors
        sys: ElasticSystem = ctx.sys
        if sys.state != ctx.state:  # K, f or the factors belong to other moduli now
            sys.build(E)
            ctx.state = sys.state
        r = grad_u.detach().to(torch.float64).contiguous().clone()
        S.setRHS_Zero(r, sys.flag, 0)
        if torch.equal(r, ctx.f):
            # the adjoint system is the forward one (compliance: dJ/du = f), and
            # the solver is deterministic: its solution would be u, bit for bit
            lam = u
        else:
            lam = torch.zeros_like(r)
            hist = S.Solve_PCG(r, lam, ctx.tol, ctx.max_nitr, sys.A, sys.ilu)
            _converged(hist, ctx.tol, "adjoint solve")
        return sys.adjoint_grad(u, lam), None, None, None


def compliance_and_grad(sys: ElasticSystem, rho: torch.Tensor, p: float = 3.0,
                        tol: float = 1e-10):
    """SIMP compliance C = f.u with E = rho^p, and dC/drho by autograd."""
    rho = rho.detach().clone().requires_grad_(True)
    E = rho ** p
    u = ElasticTetSolve.apply(E, sys, tol)
    C = torch.dot(sys.b, u)
    C.backward()
    return C.detach(), rho.grad.detach(), u.detach()
