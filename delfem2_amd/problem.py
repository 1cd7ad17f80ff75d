"""End-to-end drivers over the device API: the canonical call sequence of
examples_oldgl_glfw/618_Fem3Tet/main.cpp:234-319 (pattern, assembly, fixed
BCs, ILU copy + decompose, Solve_PCG) for a ``workload.Problem``.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import solver as S
from .workload import Problem


class DeviceSystem:
    """Pattern, matrix, ILU(0) and device-resident inputs of one Problem."""

    def __init__(self, prob: Problem, device: int = 0, fill_level: int = 0):
        S.init(device)
        self.prob = prob
        dev = torch.device("cuda", device)
        self.dev = dev
        t0 = time.perf_counter()
        self.mesh = S.TetMesh(prob.tet, prob.np_)
        ind, arr = S.JArray_PSuP_MeshElem_Device(self.mesh)
        self.psup_ind, self.psup = ind, arr
        self.A = S.CMatrixSparse()
        self.A.Initialize(prob.np_, 3, True)
        self.A.SetPattern(ind, arr)
        self.ilu = S.CPreconditionerILU()
        if fill_level == 0:
            self.ilu.SetPattern0(self.A)
        else:
            self.ilu.Initialize_ILUk(self.A, fill_level)
        self.xyz = torch.from_numpy(prob.xyz).to(dev)
        self.myu = torch.from_numpy(prob.myu).to(dev)
        self.lam = torch.from_numpy(prob.lam).to(dev)
        self.flag = torch.from_numpy(prob.flag).to(dev)
        self.f_ext = None if prob.f_ext is None else torch.from_numpy(prob.f_ext).to(dev)
        self.b = torch.zeros(prob.ndof, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        self.t_setup = time.perf_counter() - t0

    @property
    def ndof(self) -> int:
        return self.prob.ndof

    def assemble(self, disp: torch.Tensor | None = None) -> bool:
        """setZero, MergeLinSys (per-element Lame), SetFixedBC, setRHS_Zero."""
        self.A.setZero()
        self.b.zero_()
        ok = S.MergeLinSys_SolidLinear_Static_MeshTet3D(self.A, self.b, self.myu, self.lam,
                                                         self.prob.rho, self.prob.g, self.xyz,
                                                         self.mesh, disp)
        if self.f_ext is not None:
            self.b.add_(self.f_ext)
        self.A.SetFixedBC(self.flag)
        S.setRHS_Zero(self.b, self.flag, 0)
        return ok

    def factor(self) -> bool:
        self.ilu.CopyValue(self.A)
        return self.ilu.Decompose()

    def solve(self, tol: float = 1e-8, max_nitr: int = 5000):
        r = self.b.clone()
        x = torch.zeros_like(self.b)
        hist = S.Solve_PCG(r, x, tol, max_nitr, self.A, self.ilu)
        return x, hist, r

    def solve_mixed(self, tol: float = 1e-8, delta: float = 1e-1, max_outer: int = 200,
                    max_inner: int = 20000):
        x = torch.zeros_like(self.b)
        hist, n_inner = S.Solve_PCG_Mixed(self.b, x, tol, self.A, self.ilu, delta,
                                          max_outer, max_inner)
        return x, hist, n_inner


class BatchSystem(DeviceSystem):
    """The union system of a ``workload.Batch`` plus its batched solver."""

    def __init__(self, bat, device: int = 0):
        super().__init__(bat.union, device)
        self.batch = bat
        self.solver = None

    def factor(self) -> bool:
        ok = super().factor()
        if self.solver is None:
            self.solver = S.BatchSolver(self.A, self.ilu, self.batch.node_offset)
        return ok

    def solve_batch(self, tol: float = 1e-8, fp32: bool = False, delta: float = 1e-1,
                    max_nitr: int = 20000):
        x = torch.zeros_like(self.b)
        it, rr, st = self.solver.solve(self.b, x, tol, fp32, delta, max_nitr)
        return x, it, rr, st


class DynamicSystem(DeviceSystem):
    """Implicit time stepping of a Problem on the device (body force only).

    ``scheme="newmark"``: the step of
    examples_oldgl_glfw/618_Fem3Tet/main.cpp:361-416 --
    MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D (M + beta dt^2 K and the
    Newmark residual), SetFixedBC, setRHS_Zero, ILU CopyValue + Decompose,
    Solve_PCG for the acceleration increment x, then
    u += dt v + (0.5 dt^2) a + (dt^2 beta) x, v += (dt gamma) x + dt a,
    a += x on the free DOFs (XPlusAYBZCW / XPlusAYBZ / XPlusAY).

    ``scheme="beuler"``: the step of
    examples_oldgl_glfw/603_Fem3DStiffwarp/main.cpp:135-178 --
    MergeLinSys_SolidLinear_BEuler_MeshTet3D (K + lumped M / dt^2 and
    (f - K (u + dt v)) / dt), the same BC / ILU steps, a solve for the
    velocity increment dv, then u += dt dv + dt v, v += dv.  The example
    solves with Solve_PBiCGStab; the matrix is symmetric positive definite,
    so this driver uses the fused Solve_PCG (the solution agrees to the
    solver tolerance, the iterates do not).

    ``scheme="stiffwarp"``: the step of
    examples_oldgl_glfw/603_Fem3DStiffwarp/main.cpp:180-227 -- node
    rotations (RotationAtMeshPoints), MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D
    (R_i K_e R_i^T blocks, not symmetric), BCs, ILU(0), Solve_PBiCGStab for
    dv, then the backward-Euler updates.
    """

    def __init__(self, prob: Problem, scheme: str = "newmark", dt: float = 0.03,
                 gamma_newmark: float = 0.6, beta_newmark: float = 0.36, device: int = 0,
                 fill_level: int = 0):
        if scheme not in ("newmark", "beuler", "stiffwarp"):
            raise ValueError(f"unknown scheme {scheme!r}")
        super().__init__(prob, device, fill_level)
        self.scheme = scheme
        self.dt = float(dt)
        self.gamma = float(gamma_newmark)
        self.beta = float(beta_newmark)
        self.u = torch.zeros_like(self.b)
        self.v = torch.zeros_like(self.b)
        self.a = torch.zeros_like(self.b)
        self.R = torch.zeros(prob.np_ * 9, dtype=torch.float64, device=self.dev)

    def assemble(self, disp=None) -> bool:
        self.A.setZero()
        self.b.zero_()
        p = self.prob
        if self.scheme == "newmark":
            ok = S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
                self.A, self.b, self.myu, self.lam, p.rho, p.g, self.dt, self.gamma, self.beta,
                self.xyz, self.mesh, self.u, self.v, self.a)
        elif self.scheme == "stiffwarp":
            S.RotationAtMeshPoints(self.A, self.xyz, self.u, self.R)
            ok = S.MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(
                self.A, self.b, self.myu, self.lam, p.rho, p.g, self.dt, self.xyz, self.mesh,
                self.u, self.v, self.R)
        else:
            ok = S.MergeLinSys_SolidLinear_BEuler_MeshTet3D(
                self.A, self.b, self.myu, self.lam, p.rho, p.g, self.dt, self.xyz, self.mesh,
                self.u, self.v)
        self.A.SetFixedBC(self.flag)
        S.setRHS_Zero(self.b, self.flag, 0)
        return ok

    def step(self, tol: float = 1e-4, max_nitr: int = 1000) -> int:
        """One time step; returns the PCG iteration count."""
        if not self.assemble():
            raise RuntimeError("an element edge is missing from the matrix pattern")
        if not self.factor():
            raise RuntimeError("ILU decomposition failed")
        x = torch.zeros_like(self.b)
        if self.scheme == "stiffwarp":  # history without a start entry
            hist = [0.0] + S.Solve_PBiCGStab(self.b, x, tol, max_nitr, self.A, self.ilu)
        else:
            hist = S.Solve_PCG(self.b, x, tol, max_nitr, self.A, self.ilu)
        dt = self.dt
        if self.scheme == "newmark":
            S.XPlusAYBZCW(self.u, self.flag, dt, self.v, 0.5 * dt * dt, self.a,
                          dt * dt * self.beta, x)
            S.XPlusAYBZ(self.v, self.flag, dt * self.gamma, x, dt, self.a)
            S.XPlusAY(self.a, self.flag, 1.0, x)
        else:
            S.XPlusAYBZ(self.u, self.flag, dt, x, dt, self.v)
            S.XPlusAY(self.v, self.flag, 1.0, x)
        return len(hist) - 1
