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
