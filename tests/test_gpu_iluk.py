"""ILU(k) (Initialize_ILUk, ls_ilu_block_sparse.cpp:849-999) on the device
against the oracle (itself pinned to the reference's fixtures): fill pattern
identical, factors and preconditioner application bit-exact (FP64), fused
PCG solution within 1e-7 (iterations +-1), mixed within 1e-6."""
import numpy as np
import pytest
import torch

import oracle
from delfem2_amd import solver as S
from delfem2_amd import workload as W
from delfem2_amd.problem import DeviceSystem

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def cpu(prob, fill):
    ind, ps = oracle.psup(prob.tet, prob.np_)
    C = oracle.System(ind, ps)
    b = np.zeros(prob.ndof)
    C.assemble(b, prob.myu, prob.lam, prob.rho, prob.g, prob.xyz, prob.tet)
    if prob.f_ext is not None:
        b += prob.f_ext
    C.set_fixed_bc(prob.flag)
    oracle.set_rhs_zero(b, prob.flag)
    C.set_fill_level(fill)
    assert C.factor()
    C.b = b
    return C


CASES = {
    "c6_k1": (lambda: W.cube(6, sigma=0.5, seed=12), 1),
    "c6_color4_k2": (lambda: W.cube(6, sigma=0.5, seed=12, ordering="color4"), 2),
    "c10_k1": (lambda: W.cube(10), 1),
    "c3_full": (lambda: W.cube(3, sigma=0.5, seed=99), -1),
    "cantilever_k2": (lambda: W.cantilever(8, 4, 4, seed=7, ordering="color4"), 2),
}


@pytest.mark.parametrize("case", list(CASES))
def test_iluk_bit_exact_and_pcg(gpu, case):
    build, fill = CASES[case]
    prob = build()
    print(f"[iluk {case}] ndof={prob.ndof} fill={fill}", flush=True)
    C = cpu(prob, fill)
    G = DeviceSystem(prob, fill_level=fill)
    assert G.assemble()
    assert G.factor()
    assert np.array_equal(G.ilu.colind(), C.l_colind)
    rp, di, vc, vd = G.ilu.download()
    assert np.array_equal(rp, C.l_rowptr) and np.array_equal(di, C.l_diaind)
    assert np.array_equal(bits(vc), bits(C.l_vcrs)) and np.array_equal(bits(vd), bits(C.l_vdia))
    v = G.b.clone()
    G.ilu.SolvePrecond(v)
    pc = C.b.copy()
    C.solve_precond(pc)
    assert np.array_equal(bits(v.cpu().numpy()), bits(pc))
    xc, hc = C.solve_pcg(C.b.copy(), 1e-8, 5000)
    x, hist, _ = G.solve(1e-8, 5000)
    assert abs(len(hist) - len(hc)) <= 1
    assert np.linalg.norm(x.cpu().numpy() - xc) / np.linalg.norm(xc) <= 1e-7
    xe, _ = C.solve_pcg(C.b.copy(), 1e-12, 5000)
    xm, hm, _ = G.solve_mixed(1e-8)
    assert hm[-1] / hm[0] < 1e-8
    assert np.linalg.norm(xm.cpu().numpy() - xe) / np.linalg.norm(xe) <= 1e-6


def test_iluk_64cube(gpu):
    """ILU(1) at the headline size: fewer iterations than ILU(0), FP64 true
    residual below the tolerance for both solvers."""
    prob = W.cube(64, sigma=0.5, ordering="color4")
    res = {}
    for fill in (0, 1):
        G = DeviceSystem(prob, fill_level=fill)
        G.assemble()
        assert G.factor()
        x, hist, _ = G.solve(1e-8, 20000)
        xm, hm, n_inner = G.solve_mixed(1e-8)
        for xx in (x, xm):
            r = G.b.clone()
            G.A.MatVec(r, -1.0, xx, 1.0)
            assert (torch.linalg.norm(r) / torch.linalg.norm(G.b)).item() < 1.5e-8
        res[fill] = (len(hist) - 1, n_inner, G.ilu.levels())
    print("[iluk 64^3] fill -> (FP64 iterations, mixed inner, levels):", res, flush=True)
    assert res[1][0] < res[0][0]
