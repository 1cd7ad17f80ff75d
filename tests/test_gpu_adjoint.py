"""Adjoint sensitivities: the contraction kernel against the oracle (bit for
bit), the whole gradient against a CPU adjoint built from the oracle, and
against central finite differences (h = 1e-6) on seeded elements.

FD method: for element k and s = +-h, with the base solution u,
    C(rho + s e_k) - C(rho) = f . du,   K(rho + s e_k) du = r0 - dK_s u,
    r0 = f - K(rho) u,  dK_s = ((rho_k + s)^p - rho_k^p) K_k^1
(an exact identity).  K(rho + s e_k) is re-assembled on the GPU and du solved
to 1e-12 with the base factors as preconditioner; dK_s u is the exact local
12-entry element term, so neither two nearly equal compliances nor
f - K(rho+s) u (which loses ~6 digits) is ever formed.  Tolerance: 1e-5
relative (the target).
"""
import numpy as np
import pytest
import torch

import oracle
from delfem2_amd import solver as S
from delfem2_amd import workload as W
from delfem2_amd.adjoint import ElasticSystem, ElasticTetSolve, compliance_and_grad, unit_lame

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_contraction_kernel_bit_exact(gpu):
    prob = W.cube(4, sigma=0.5, seed=3, ordering="color4")
    sys = ElasticSystem(prob)
    u = W.uniform(1, prob.ndof) - 0.5
    lam = W.uniform(2, prob.ndof) - 0.5
    g = sys.adjoint_grad(torch.from_numpy(u).cuda(), torch.from_numpy(lam).cuda()).cpu().numpy()
    m1, l1 = unit_lame(prob.extra["nu"])
    gc = oracle.adjoint_grad(m1, l1, prob.xyz, prob.tet, u, lam)
    assert np.array_equal(bits(g), bits(gc))


def test_compliance_adjoint_reuses_forward_solution(gpu):
    """For the compliance the adjoint system is the forward one; the backward
    pass then takes lam = u.  A second, explicit solve of that system gives
    the same bits (the solver is deterministic), so the gradient is the one
    the explicit adjoint solve would give (up to how torch forms 3 rho^2)."""
    prob = W.cantilever(8, 4, 4, seed=7, ordering="color4")
    sys = ElasticSystem(prob)
    rho = torch.from_numpy(prob.extra["density"]).cuda()
    C, grad, u = compliance_and_grad(sys, rho, tol=1e-10)
    r = sys.b.clone()
    lam = torch.zeros_like(r)
    S.Solve_PCG(r, lam, 1e-10, 20000, sys.A, sys.ilu)
    assert torch.equal(lam, u)
    g = (3.0 * rho ** 2 * sys.adjoint_grad(u, lam)).cpu().numpy()
    assert np.allclose(g, grad.cpu().numpy(), rtol=1e-14, atol=0.0)


def cpu_compliance_grad(prob, rho, p=3.0):
    """CPU adjoint: oracle assembly / ILU / PCG (tol 1e-12), lam = u."""
    E = rho ** p
    m1, l1 = unit_lame(prob.extra["nu"])
    ind, ps = oracle.psup(prob.tet, prob.np_)
    C = oracle.System(ind, ps)
    b = np.zeros(prob.ndof)
    C.assemble(b, E * m1, E * l1, prob.rho, prob.g, prob.xyz, prob.tet)
    b += prob.f_ext
    C.set_fixed_bc(prob.flag)
    oracle.set_rhs_zero(b, prob.flag)
    C.factor()
    u, _ = C.solve_pcg(b.copy(), 1e-12, 20000)
    g = oracle.adjoint_grad(m1, l1, prob.xyz, prob.tet, u, u)
    return float(b @ u), p * rho ** (p - 1) * g, u


def element_stiffness_unit(prob, k):
    """K_k^1 (12 x 12, node-major) from the oracle's element matrix."""
    m1, l1 = unit_lame(prob.extra["nu"])
    nodes = prob.tet[4 * k:4 * k + 4].astype(np.int64)
    P = prob.xyz.reshape(-1, 3)[nodes]
    em = np.zeros(144)
    er = np.zeros(12)
    oracle.lib().orc_emat_tet(em, er, m1[k], l1[k], np.ascontiguousarray(P.ravel()),
                              np.zeros(12))
    return em.reshape(4, 4, 3, 3).transpose(0, 2, 1, 3).reshape(12, 12), nodes


def fd_grad(sys, rho, elems, h=1e-6, p=3.0, tol=1e-12, weights=None):
    """Central differences in rho_k of J_w = w.u for each w in `weights`
    (default: [f], the compliance), by the perturbation identity; one du
    per (element, sign) serves every w.  Returns an array [len(weights),
    len(elems)]."""
    prob = sys.prob
    rho_t = torch.from_numpy(rho).cuda()
    sys.set_modulus(rho_t ** p)
    sys.assemble()
    sys.factor()
    u, _, _ = sys.solve(tol, 20000)
    r0 = sys.b.clone()
    ws = [sys.b.clone()] if weights is None else list(weights)
    sys.A.MatVec(r0, -1.0, u, 1.0)
    u_h = u.cpu().numpy()
    flag = prob.flag
    out = []
    for k in elems:
        Ke, nodes = element_stiffness_unit(prob, k)
        dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        Ku = Ke @ u_h[dofs]
        dc = []
        for s in (h, -h):
            r2 = rho_t.clone()
            r2[k] += s
            sys.set_modulus(r2 ** p)
            sys.assemble()          # K(rho + s e_k); f unchanged; factors kept from the base
            dE = (rho[k] + s) ** p - rho[k] ** p
            r = r0.clone()
            upd = torch.from_numpy(np.where(flag[dofs] != 0, 0.0, -dE * Ku)).cuda()
            r[torch.from_numpy(dofs).cuda()] += upd
            du = torch.zeros_like(r)
            S.Solve_PCG(r, du, tol, 20000, sys.A, sys.ilu)
            dc.append([torch.dot(w, du).item() for w in ws])
        out.append([(a - b) / (2 * h) for a, b in zip(dc[0], dc[1])])
    sys.set_modulus(rho_t ** p)
    sys.assemble()
    return np.array(out).T


def weighted_objective_grad(sys, rho, w, p=3.0, tol=1e-12):
    """J = w.u through ElasticTetSolve: dJ/du = w != f, so backward runs the
    adjoint PCG solve K lam = w with the forward factors."""
    rho = rho.detach().clone().requires_grad_(True)
    u = ElasticTetSolve.apply(rho ** p, sys, tol)
    J = torch.dot(w, u)
    J.backward()
    return J.detach(), rho.grad.detach(), u.detach()


def cpu_weighted_grad(prob, rho, w, p=3.0):
    """CPU adjoint of J = w.u: oracle assembly / ILU / PCG (tol 1e-12) for u
    and for K lam = w (fixed DOFs zero)."""
    E = rho ** p
    m1, l1 = unit_lame(prob.extra["nu"])
    ind, ps = oracle.psup(prob.tet, prob.np_)
    C = oracle.System(ind, ps)
    b = np.zeros(prob.ndof)
    C.assemble(b, E * m1, E * l1, prob.rho, prob.g, prob.xyz, prob.tet)
    b += prob.f_ext
    C.set_fixed_bc(prob.flag)
    oracle.set_rhs_zero(b, prob.flag)
    C.factor()
    u, _ = C.solve_pcg(b.copy(), 1e-12, 20000)
    rw = np.where(prob.flag != 0, 0.0, w)
    lam, _ = C.solve_pcg(rw, 1e-12, 20000)
    g = oracle.adjoint_grad(m1, l1, prob.xyz, prob.tet, u, lam)
    return float(w @ u), p * rho ** (p - 1) * g


def test_compliance_gradient_vs_cpu_adjoint_and_fd(gpu):
    prob = W.cantilever(12, 6, 6, seed=21, ordering="color4")
    rho = prob.extra["density"]
    sys = ElasticSystem(prob)
    C, grad, u = compliance_and_grad(sys, torch.from_numpy(rho).cuda(), tol=1e-12)
    grad = grad.cpu().numpy()
    Cc, gc, uc = cpu_compliance_grad(prob, rho)
    assert abs(C.item() - Cc) <= 1e-10 * abs(Cc)
    assert np.linalg.norm(grad - gc) / np.linalg.norm(gc) <= 1e-8
    assert np.all(grad < 0)  # stiffer elements lower the compliance
    elems = (W.uniform(5, 8) * prob.ne).astype(int)
    fd = fd_grad(sys, rho, elems)[0]
    rel = np.abs(fd - grad[elems]) / np.abs(grad[elems])
    print("FD rel err", rel.max())
    assert rel.max() <= 1e-5


def test_weighted_objective_runs_adjoint_solve(gpu):
    """J = w.u with a seeded w (dJ/du = w != f): the backward pass solves
    K lam = w with the forward factors.  Gradient vs the CPU adjoint (<= 1e-8
    rel-L2) and vs central FD (<= 1e-5 rel) on 8 seeded elements."""
    prob = W.cantilever(12, 6, 6, seed=22, ordering="color4")
    rho = prob.extra["density"]
    w_h = W.uniform(31, prob.ndof) - 0.5
    sys = ElasticSystem(prob)
    w = torch.from_numpy(w_h).cuda()
    J, grad, u = weighted_objective_grad(sys, torch.from_numpy(rho).cuda(), w)
    grad = grad.cpu().numpy()
    Jc, gc = cpu_weighted_grad(prob, rho, w_h)
    assert abs(J.item() - Jc) <= 1e-9 * abs(Jc)
    assert np.linalg.norm(grad - gc) / np.linalg.norm(gc) <= 1e-8
    elems = (W.uniform(6, 8) * prob.ne).astype(int)
    fd = fd_grad(sys, rho, elems, weights=[w])[0]
    rel = np.abs(fd - grad[elems]) / np.abs(grad[elems])
    print("weighted FD rel err", rel.max())
    assert rel.max() <= 1e-5


def test_backward_after_the_system_was_rebuilt(gpu):
    """A second forward solve for other moduli between a forward and its
    backward (two load cases on one system): the first backward rebuilds
    K(E1) and its factors and returns the gradient an undisturbed backward
    gives, bit for bit."""
    prob = W.cantilever(8, 4, 4, seed=23, ordering="color4")
    rho1 = torch.from_numpy(prob.extra["density"]).cuda()
    rho2 = rho1.flip(0).contiguous()
    w = torch.from_numpy(W.uniform(32, prob.ndof) - 0.5).cuda()
    _, g_ref, _ = weighted_objective_grad(ElasticSystem(prob), rho1, w)
    sys = ElasticSystem(prob)
    r1 = rho1.clone().requires_grad_(True)
    u1 = ElasticTetSolve.apply(r1 ** 3.0, sys, 1e-12)
    u2 = ElasticTetSolve.apply(rho2 ** 3.0, sys, 1e-12)  # rebuilds sys for E2
    assert not torch.equal(u1, u2)
    torch.dot(w, u1).backward()
    assert torch.equal(r1.grad, g_ref)


def test_c4_cantilever_96x48x48(gpu):
    """The full differentiation workload: 1,327,104 tets, FP64 forward +
    adjoint, dC/drho for all elements, FD on 64 seeded elements; and the
    weighted objective J = w.u (a real adjoint solve) on the same elements
    (one FD du per element and sign serves both objectives)."""
    prob = W.cantilever(96, 48, 48, seed=20261204, ordering="color4")
    rho = prob.extra["density"]
    sys = ElasticSystem(prob)
    torch.cuda.synchronize()
    C, grad, u = compliance_and_grad(sys, torch.from_numpy(rho).cuda(), tol=1e-12)
    grad = grad.cpu().numpy()
    assert grad.shape == (prob.ne,) and np.all(np.isfinite(grad)) and np.all(grad < 0)
    w = torch.from_numpy(W.uniform(33, prob.ndof) - 0.5).cuda()
    _, gw, _ = weighted_objective_grad(sys, torch.from_numpy(rho).cuda(), w)
    gw = gw.cpu().numpy()
    elems = (W.uniform(20261204, 64) * prob.ne).astype(int)
    fd, fdw = fd_grad(sys, rho, elems, weights=[sys.b.clone(), w])
    rel = np.abs(fd - grad[elems]) / np.abs(grad[elems])
    relw = np.abs(fdw - gw[elems]) / np.abs(gw[elems])
    print(f"C={C.item():.6e} FD max rel err over 64 elements {rel.max():.3e}, "
          f"J=w.u {relw.max():.3e}")
    assert rel.max() <= 1e-5
    assert relw.max() <= 1e-5
