"""Time-stepping drivers on the device through the C ABI, against the
oracle (itself pinned to the reference: tests/test_oracle_dynamics.py).

Bit-exact: Newmark-beta, backward-Euler and stiffness-warped assembly (A and
b, per-element Lame pairs, nonzero u / v / a; warped with the oracle's
rotations), XPlusAYBZ and XPlusAYBZCW.  Tolerance: node rotations (device
trigonometry) within 1e-12; BiCGStab and the stepped state (dot products in
another order): iteration counts within 1 of the oracle's, x / u / v / a
within rel-L2 1e-7 at solver tolerance 1e-10.
"""
import numpy as np
import pytest
import torch

import oracle
from delfem2_amd import solver as S
from delfem2_amd import workload as W
from delfem2_amd.problem import DynamicSystem

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda").to(dtype)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("scheme", ["newmark", "beuler"])
@pytest.mark.parametrize("order", ["natural", "color4"])
def test_dynamic_assembly_bit_exact(gpu, scheme, order):
    p = W.cube(5, sigma=0.5, seed=31, ordering=order)
    print(f"[dyn asm] {scheme} {order} ndof={p.ndof}", flush=True)
    u, v, a = ((W.uniform(k, p.ndof) - 0.5) * s for k, s in ((41, 1e-2), (42, 1e-1), (43, 1.0)))
    g = np.array([0.1, -0.4, -2.0])
    rho, dt, beta, gamma = 1.2, 0.025, 0.3, 0.55
    ind, ps = oracle.psup(p.tet, p.np_)
    C = oracle.System(ind, ps)
    bc = np.zeros(p.ndof)
    G = DynamicSystem(p, scheme)
    G.A.setZero()
    bd = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    if scheme == "newmark":
        assert C.assemble_newmark(bc, p.myu, p.lam, rho, g, dt, beta, p.xyz, p.tet, u, v, a) == 0
        assert S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
            G.A, bd, G.myu, G.lam, rho, g, dt, gamma, beta, G.xyz, G.mesh, dev(u), dev(v), dev(a))
    else:
        assert C.assemble_beuler(bc, p.myu, p.lam, rho, g, dt, p.xyz, p.tet, u, v) == 0
        assert S.MergeLinSys_SolidLinear_BEuler_MeshTet3D(G.A, bd, G.myu, G.lam, rho, g, dt,
                                                          G.xyz, G.mesh, dev(u), dev(v))
    vc, vd = G.A.download()
    assert np.array_equal(bits(bd.cpu().numpy()), bits(bc))
    assert np.array_equal(bits(vc), bits(C.vcrs))
    assert np.array_equal(bits(vd), bits(C.vdia))
    # scalar Lame pair (the reference signature) = per-element arrays of that pair
    q = W.cube(3, seed=2, ordering=order)
    H = DynamicSystem(q, scheme)
    for per_element in (True, False):
        H.A.setZero()
        b = torch.zeros(q.ndof, dtype=torch.float64, device="cuda")
        m = H.myu if per_element else float(q.myu[0])
        l_ = H.lam if per_element else float(q.lam[0])
        z = torch.zeros_like(b)
        if scheme == "newmark":
            S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(H.A, b, m, l_, rho, g, dt, gamma, beta,
                                                            H.xyz, H.mesh, z, z, z)
        else:
            S.MergeLinSys_SolidLinear_BEuler_MeshTet3D(H.A, b, m, l_, rho, g, dt, H.xyz, H.mesh,
                                                       z, z)
        out = (b.cpu().numpy(),) + tuple(H.A.download())
        if per_element:
            ref = out
        else:
            assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(out, ref))


def test_xplus_bit_exact(gpu):
    S.init(0)
    n = 1001
    X = W.uniform(1, n) - 0.5
    Y, Z, Wv = (W.uniform(k, n) - 0.5 for k in (2, 3, 4))
    flag = (np.arange(n) % 5 == 0).astype(np.int32)
    Xc = X.copy()
    oracle.xplusaybzcw(Xc, flag, 0.3, Y, -1.7, Z, 2.5, Wv)
    Xd = dev(X)
    S.XPlusAYBZCW(Xd, dev(flag, torch.int32), 0.3, dev(Y), -1.7, dev(Z), 2.5, dev(Wv))
    assert np.array_equal(bits(Xd.cpu().numpy()), bits(Xc))
    Xc = X.copy()
    oracle.xplusaybz(Xc, flag, -0.6, Y, 1.1, Z)
    Xd = dev(X)
    S.XPlusAYBZ(Xd, dev(flag, torch.int32), -0.6, dev(Y), 1.1, dev(Z))
    assert np.array_equal(bits(Xd.cpu().numpy()), bits(Xc))
    S.XPlusAYBZ(Xd[:0], dev(flag, torch.int32)[:0], 1.0, Xd[:0], 1.0, Xd[:0])  # n = 0


@pytest.mark.parametrize("scheme", ["newmark", "beuler"])
def test_stepping_vs_oracle(gpu, scheme):
    p = W.cube(8, sigma=0.5, seed=9, ordering="color4")
    p.g = np.array([0.0, -0.5, -3.0])
    print(f"[dyn steps] {scheme} ndof={p.ndof}", flush=True)
    G = DynamicSystem(p, scheme, dt=0.03, gamma_newmark=0.6, beta_newmark=0.36)
    C = oracle.DynamicRun(p, scheme, dt=0.03, gamma_newmark=0.6, beta_newmark=0.36)
    for k in range(5):
        i_d = G.step(1e-10, 2000)
        i_c = C.step(1e-10, 2000)
        assert abs(i_d - i_c) <= 1, (k, i_d, i_c)
    for dvec, cvec in ((G.u, C.u), (G.v, C.v), (G.a, C.a)):
        assert rel(dvec.cpu().numpy(), cvec) <= 1e-7
    assert np.abs(C.u).max() > 1e-4  # the body moved
    # fixed DOFs never move
    fixed = p.flag != 0
    assert np.all(G.u.cpu().numpy()[fixed] == 0.0)


def _rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def test_rotations_vs_oracle(gpu):
    """Device RotationAtMeshPoints vs the oracle (host libm trigonometry):
    agreement to 1e-12; a rigid rotation Q gives Q at every node."""
    p = W.cube(6, sigma=0.5, seed=3, ordering="color4")
    print(f"[rot] ndof={p.ndof}", flush=True)
    G = DynamicSystem(p, "stiffwarp")
    X = p.xyz.reshape(-1, 3)
    for u in ((W.uniform(71, p.ndof) - 0.5) * 0.2,
              (((X @ _rot([0.2, 1.0, -0.3], 1.1).T) - X)).reshape(-1)):
        Rd = torch.zeros(p.np_ * 9, dtype=torch.float64, device="cuda")
        S.RotationAtMeshPoints(G.A, G.xyz, dev(u), Rd)
        Rc = oracle.rotation_at_points(p.xyz, u, G.psup_ind, G.psup)
        assert np.abs(Rd.cpu().numpy() - Rc).max() < 1e-12
    assert np.abs(Rc.reshape(-1, 3, 3) - _rot([0.2, 1.0, -0.3], 1.1)[None]).max() < 1e-12


@pytest.mark.parametrize("order", ["natural", "color4"])
def test_stiffwarp_assembly_bit_exact(gpu, order):
    """The warped assembly with the oracle's rotations uploaded: A and b
    bit-exact (the rotations themselves agree to rounding, above)."""
    p = W.cube(5, sigma=0.5, seed=33, ordering=order)
    u, v = ((W.uniform(k, p.ndof) - 0.5) * s for k, s in ((44, 5e-2), (45, 1e-1)))
    g = np.array([0.1, -0.4, -2.0])
    rho, dt = 1.2, 0.025
    ind, ps = oracle.psup(p.tet, p.np_)
    R = oracle.rotation_at_points(p.xyz, u, ind, ps)
    C = oracle.System(ind, ps)
    bc = np.zeros(p.ndof)
    assert C.assemble_stiffwarp(bc, p.myu, p.lam, rho, g, dt, p.xyz, p.tet, u, v, R) == 0
    G = DynamicSystem(p, "stiffwarp")
    G.A.setZero()
    bd = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    assert S.MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(G.A, bd, G.myu, G.lam, rho, g, dt, G.xyz,
                                                         G.mesh, dev(u), dev(v), dev(R))
    vc, vd = G.A.download()
    assert np.array_equal(bits(bd.cpu().numpy()), bits(bc))
    assert np.array_equal(bits(vc), bits(C.vcrs))
    assert np.array_equal(bits(vd), bits(C.vdia))


@pytest.mark.parametrize("order", ["natural", "color4"])
def test_listed_assembly_equals_walking_form(gpu, order, monkeypatch):
    """Every kind of the gather assembly: the listed form (element lists
    built once per matrix and mesh, per-tet geometry records) gives the bits
    of the walking form (DFM2H_ASSEMBLY_WALK=1) on the first call (lists
    built), a repeated call (lists reused) and a second mesh handle of the
    same tets (lists rebuilt)."""
    p = W.cube(4, sigma=0.5, seed=35, ordering=order)
    u, v, a = ((W.uniform(k, p.ndof) - 0.5) * s for k, s in ((46, 1e-2), (47, 1e-1), (48, 1.0)))
    R = W.uniform(49, p.np_ * 9) - 0.5
    g = [0.1, -0.4, -2.0]
    rho, dt, beta, gamma = 1.2, 0.025, 0.3, 0.55
    G = DynamicSystem(p, "newmark")
    meshes = [G.mesh, S.TetMesh(p.tet, p.np_)]
    du, dv, da, dR = dev(u), dev(v), dev(a), dev(R)

    def run(kind, mesh):
        G.A.setZero()
        b = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
        if kind == "static":
            ok = S.MergeLinSys_SolidLinear_Static_MeshTet3D(G.A, b, G.myu, G.lam, rho, g, G.xyz,
                                                            mesh, du)
        elif kind == "newmark":
            ok = S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
                G.A, b, G.myu, G.lam, rho, g, dt, gamma, beta, G.xyz, mesh, du, dv, da)
        elif kind == "beuler":
            ok = S.MergeLinSys_SolidLinear_BEuler_MeshTet3D(G.A, b, G.myu, G.lam, rho, g, dt,
                                                            G.xyz, mesh, du, dv)
        else:
            ok = S.MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(G.A, b, G.myu, G.lam, rho, g, dt,
                                                               G.xyz, mesh, du, dv, dR)
        assert ok
        return (b.cpu().numpy(),) + tuple(G.A.download())

    for kind in ("static", "newmark", "beuler", "stiffwarp"):
        print(f"[asm forms] {kind} {order}", flush=True)
        monkeypatch.setenv("DFM2H_ASSEMBLY_WALK", "1")
        ref = run(kind, meshes[0])
        monkeypatch.delenv("DFM2H_ASSEMBLY_WALK")
        for mesh in (meshes[0], meshes[0], meshes[1]):
            out = run(kind, mesh)
            assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(out, ref)), kind


def test_listed_assembly_many_elements_per_node(gpu, monkeypatch):
    """A fan of 40 tets around one node (more than a row's 32 staged
    records): the listed assembly's per-block path gives the walking form's
    bits, static and Newmark-beta."""
    nring = 20
    ang = 2 * np.pi * np.arange(nring) / nring
    xyz = np.concatenate([[0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, -1.0],
                          np.stack([np.cos(ang), np.sin(ang), 0.1 * np.sin(3 * ang)], 1).ravel()])
    tets = []
    for k in range(nring):
        a_, b_ = 3 + k, 3 + (k + 1) % nring
        tets += [[0, 1, a_, b_], [0, 2, b_, a_]]
    tet = np.array(tets, np.uint32)
    npt = xyz.size // 3
    ind, ps = oracle.psup(tet, npt)
    A = S.CMatrixSparse()
    A.Initialize(npt, 3, True)
    A.SetPattern(ind, ps)
    mesh = S.TetMesh(tet, npt)
    ne = len(tet)
    myu = dev(1.0 + W.uniform(51, ne))
    lam = dev(0.5 + W.uniform(52, ne))
    du, dv, da = (dev((W.uniform(k, 3 * npt) - 0.5) * 0.1) for k in (53, 54, 55))
    dx = dev(xyz)
    g = [0.1, -0.4, -2.0]

    def run(kind):
        A.setZero()
        b = torch.zeros(3 * npt, dtype=torch.float64, device="cuda")
        if kind == "static":
            ok = S.MergeLinSys_SolidLinear_Static_MeshTet3D(A, b, myu, lam, 1.1, g, dx, mesh, du)
        else:
            ok = S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(A, b, myu, lam, 1.1, g, 0.02, 0.6,
                                                                 0.3, dx, mesh, du, dv, da)
        assert ok
        return (b.cpu().numpy(),) + tuple(A.download())

    for kind in ("static", "newmark"):
        monkeypatch.setenv("DFM2H_ASSEMBLY_WALK", "1")
        ref = run(kind)
        monkeypatch.delenv("DFM2H_ASSEMBLY_WALK")
        out = run(kind)
        assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(out, ref)), kind
        assert np.abs(ref[2]).max() > 0


def test_newmark_assembly_64cube_vs_reference(gpu):
    """The Newmark-beta assembly at the bench's size (64^3, colour-4 order,
    uniform Lame pair, seeded u / v / a): A and b hash to the compiled
    reference driver's SHA-256."""
    import hashlib
    import json
    import os
    meta = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    m = meta["cases"]["dyn_newmark_c64"]

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

    p = W.cube(64, seed=5, ordering="color4")
    print(f"[newmark 64] ndof={p.ndof}", flush=True)
    u, v, a = ((W.uniform(k, p.ndof) - 0.5) * s for k, s in ((61, 1e-2), (62, 1e-1), (63, 1.0)))
    G = DynamicSystem(p, "newmark")
    G.A.setZero()
    bd = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    assert S.MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
        G.A, bd, float(p.myu[0]), float(p.lam[0]), p.rho, m["g"], m["dt"], m["gamma"], m["beta"],
        G.xyz, G.mesh, dev(u), dev(v), dev(a))
    vc, vd = G.A.download()
    assert sha(vc) == m["assembly"]["vcrs"] and sha(vd) == m["assembly"]["vdia"]
    assert sha(bd.cpu().numpy()) == m["assembly"]["b"]


def test_bicgstab_vs_oracle(gpu):
    """Device Solve_PBiCGStab on the warped (non-symmetric) system vs the
    oracle's: iteration counts within 1, x rel-L2 <= 1e-7, true residual."""
    p = W.cube(8, sigma=0.5, seed=6, ordering="color4")
    X = p.xyz.reshape(-1, 3)
    u = (((X @ _rot([1.0, 0.2, -0.4], 0.9).T) - X) * 0.5).reshape(-1)
    v = (W.uniform(9, p.ndof) - 0.5) * 0.1
    g = np.array([0.0, 0.0, -3.0])
    ind, ps = oracle.psup(p.tet, p.np_)
    R = oracle.rotation_at_points(p.xyz, u, ind, ps)
    C = oracle.System(ind, ps)
    b = np.zeros(p.ndof)
    C.assemble_stiffwarp(b, p.myu, p.lam, 1.0, g, 0.03, p.xyz, p.tet, u, v, R)
    C.set_fixed_bc(p.flag)
    oracle.set_rhs_zero(b, p.flag)
    assert C.factor()
    xc, hc = C.solve_pbicgstab(b.copy(), 1e-10, 1000)
    G = DynamicSystem(p, "stiffwarp")
    G.A.setZero()
    bd = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    S.MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(G.A, bd, G.myu, G.lam, 1.0, g, 0.03, G.xyz,
                                                  G.mesh, dev(u), dev(v), dev(R))
    G.A.SetFixedBC(G.flag)
    S.setRHS_Zero(bd, G.flag, 0)
    assert np.array_equal(bits(bd.cpu().numpy()), bits(b))
    assert G.factor()
    r = bd.clone()
    x = torch.zeros_like(bd)
    hd = S.Solve_PBiCGStab(r, x, 1e-10, 1000, G.A, G.ilu)
    print(f"[bicgstab] iterations device {len(hd)} oracle {len(hc)}", flush=True)
    assert abs(len(hd) - len(hc)) <= 1 and hd[-1] < 1e-10
    assert rel(x.cpu().numpy(), xc) <= 1e-7
    res = bd.clone()
    G.A.MatVec(res, -1.0, x, 1.0)
    assert torch.linalg.norm(res).item() < 1e-9 * torch.linalg.norm(bd).item()
    # the history entry by entry (dot products in another order: rounding)
    k = min(len(hd), len(hc))
    assert np.allclose(hd[:k], hc[:k], rtol=1e-6, atol=0.0)
    # the iteration cap: max_niter - 1 iterations, one entry each, as the oracle
    r = bd.clone()
    h4 = S.Solve_PBiCGStab(r, x, 1e-30, 4, G.A, G.ilu)
    _, hc4 = C.solve_pbicgstab(b.copy(), 1e-30, 4)
    assert len(h4) == len(hc4) == 3
    assert np.allclose(h4, hc4, rtol=1e-9, atol=0.0)
    # zero right-hand side: x = 0 and the history is |b| alone
    z = torch.zeros_like(bd)
    hz = S.Solve_PBiCGStab(z, x, 1e-10, 1000, G.A, G.ilu)
    assert hz == [0.0] and torch.count_nonzero(x).item() == 0


def test_stiffwarp_stepping_vs_oracle(gpu):
    p = W.cube(6, sigma=0.5, seed=9, ordering="color4")
    p.g = np.array([0.0, -2.0, -3.0])
    print(f"[stiffwarp steps] ndof={p.ndof}", flush=True)
    G = DynamicSystem(p, "stiffwarp", dt=0.03)
    Cr = oracle.DynamicRun(p, "stiffwarp", dt=0.03)
    for k in range(5):
        i_d = G.step(1e-10, 1000)
        i_c = Cr.step(1e-10, 1000)
        assert abs(i_d - i_c) <= 1, (k, i_d, i_c)
    for dvec, cvec in ((G.u, Cr.u), (G.v, Cr.v)):
        assert rel(dvec.cpu().numpy(), cvec) <= 1e-7
    assert np.abs(Cr.u).max() > 1e-4
