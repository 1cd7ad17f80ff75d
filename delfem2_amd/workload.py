"""Synthetic inputs for the tet elasticity solve: voxel tet meshes, node
orderings, a portable seeded RNG and the material / load / clamp arrays.

The mesh follows gridvoxel.cpp:194-271 (MeshTet3D_VoxelGrid): nodes x-major,
``ip = ix*(ny+1)*(nz+1) + iy*(nz+1) + iz``, six Kuhn tets per voxel taken from
the voxel's corner map.  Everything here is host-side numpy set-up; no solver
arithmetic happens here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# corner k of a voxel is offset (dx, dy, dz); gridvoxel.cpp:228-237
_CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.int64)
# the six tets of a voxel as corner indices; gridvoxel.cpp:238-267
_SPLIT = np.array([[0, 1, 2, 6], [0, 1, 6, 5], [0, 4, 5, 6],
                   [0, 2, 3, 6], [0, 3, 7, 6], [0, 4, 6, 7]], np.int64)


def mesh_tet_voxel(nx: int, ny: int, nz: int):
    """Integer node coordinates (float64, [np*3]) and tets (uint32, [ne*4])."""
    mx, my, mz = nx + 1, ny + 1, nz + 1
    g = np.indices((mx, my, mz)).reshape(3, -1).T
    xyz = np.ascontiguousarray(g.astype(np.float64).reshape(-1))
    v = np.indices((nx, ny, nz)).reshape(3, -1).T
    base = v[:, 0] * (my * mz) + v[:, 1] * mz + v[:, 2]
    off = _CORNER[:, 0] * (my * mz) + _CORNER[:, 1] * mz + _CORNER[:, 2]
    corners = base[:, None] + off[None, :]
    tet = corners[:, _SPLIT].reshape(-1).astype(np.uint32)
    return xyz, tet


def color8_permutation(nx: int, ny: int, nz: int):
    """new->old node map that groups nodes by (ix&1, iy&1, iz&1) parity,
    stable inside a colour.  Every Kuhn-split edge is a 0/1 offset in each
    axis, so no two nodes of one colour are neighbours and the ILU(0)
    dependency depth drops from nx+ny+nz+1 levels to 8."""
    g = np.indices((nx + 1, ny + 1, nz + 1)).reshape(3, -1)
    color = (g[0] & 1) * 4 + (g[1] & 1) * 2 + (g[2] & 1)
    return np.argsort(color, kind="stable").astype(np.int64)


def color4_permutation(nx: int, ny: int, nz: int):
    """new->old node map grouping nodes by (ix+iy+iz) mod 4, stable inside a
    colour.  Every Kuhn-split edge offset has a coordinate sum of 1, 2 or 3
    (or its negative), so this is a valid 4-colouring: 4 dependency levels
    per triangular sweep instead of 8."""
    g = np.indices((nx + 1, ny + 1, nz + 1)).reshape(3, -1)
    color = (g[0] + g[1] + g[2]) % 4
    return np.argsort(color, kind="stable").astype(np.int64)


def permute_mesh(xyz: np.ndarray, tet: np.ndarray, new2old: np.ndarray):
    """Relabel nodes: node ``k`` of the result is node ``new2old[k]`` of the input."""
    old2new = np.empty_like(new2old)
    old2new[new2old] = np.arange(new2old.size)
    xyz2 = np.ascontiguousarray(xyz.reshape(-1, 3)[new2old].reshape(-1))
    tet2 = np.ascontiguousarray(old2new[tet.astype(np.int64)].astype(np.uint32))
    return xyz2, tet2, old2new


# ----------------------------------------------------------------------------
# portable RNG: SplitMix64 counter stream -> 53-bit uniforms -> Box-Muller

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs ``offset .. offset+n-1`` of the SplitMix64 stream for ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    """N(0,1) by Box-Muller on consecutive uniform pairs (cosine branch)."""
    u = uniform(seed, 2 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def lame(E: np.ndarray, nu: np.ndarray):
    """(myu, lambda) of isotropic linear elasticity, one fixed association:
    myu = E/(2(1+nu)), lambda = E*nu/((1+nu)(1-2nu))."""
    E = np.asarray(E, np.float64)
    nu = np.asarray(nu, np.float64)
    myu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return np.ascontiguousarray(myu), np.ascontiguousarray(lam)


# ----------------------------------------------------------------------------


@dataclass
class Problem:
    """A static linear-elastic tet problem ready for assembly."""
    name: str
    xyz: np.ndarray          # float64 [np*3], physical coordinates
    tet: np.ndarray          # uint32 [ne*4]
    myu: np.ndarray          # float64 [ne]
    lam: np.ndarray          # float64 [ne]
    flag: np.ndarray         # int32 [np*3], 1 = fixed DOF
    rho: float = 1.0
    g: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -1e-3]))
    f_ext: np.ndarray | None = None  # extra nodal load added to b (float64 [np*3])
    grid: tuple = (0, 0, 0)
    ordering: str = "natural"
    extra: dict = field(default_factory=dict)

    @property
    def np_(self) -> int:
        return self.xyz.size // 3

    @property
    def ne(self) -> int:
        return self.tet.size // 4

    @property
    def ndof(self) -> int:
        return self.xyz.size


def multicolor_permutation(tet: np.ndarray, np_: int):
    """new->old node map of the library's multicolour ordering for any tet
    mesh (dfm2h_multicolor_order on the mesh's one-ring pattern): nodes
    colour by colour, no two neighbours sharing a colour -- the ILU(0) sweeps'
    level count is the colour count.  Returns (new2old, ncolor)."""
    from . import solver
    ind, ps = solver.JArray_PSuP_MeshElem(tet, np_)
    return solver.MultiColorOrder(ind, ps)


def _order(xyz, tet, nx, ny, nz, ordering):
    if ordering == "natural":
        return xyz, tet, None
    if ordering in ("color8", "color4", "multicolor"):
        if ordering == "multicolor":
            n2o = multicolor_permutation(tet, xyz.size // 3)[0]
        else:
            n2o = (color8_permutation if ordering == "color8" else color4_permutation)(nx, ny, nz)
        xyz2, tet2, _ = permute_mesh(xyz, tet, n2o)
        return xyz2, tet2, n2o
    raise ValueError(ordering)


def cube(n: int, sigma: float = 0.0, seed: int = 20261203, nu=0.3,
         nu_range=None, ordering: str = "natural", scale_div: bool = True) -> Problem:
    """Unit cube of n^3 hexes x 6 tets, bottom face z=0 clamped, body force
    rho*g with g=(0,0,-1e-3).  E_e = exp(sigma*z_e) with z_e ~ N(0,1) from the
    portable RNG (E=1 when sigma=0); nu scalar or U[nu_range) per element."""
    xyz_i, tet = mesh_tet_voxel(n, n, n)
    iz = xyz_i[2::3].copy()
    xyz = xyz_i / float(n)
    ne = tet.size // 4
    E = np.exp(sigma * normal(seed, ne)) if sigma != 0.0 else np.ones(ne)
    if nu_range is not None:
        lo, hi = nu_range
        nuv = lo + (hi - lo) * uniform(seed + 1, ne)
    else:
        nuv = np.full(ne, float(nu))
    myu, lam = lame(E, nuv)
    flag = np.zeros(xyz.size, np.int32)
    flag.reshape(-1, 3)[iz == 0] = 1
    xyz, tet, n2o = _order(xyz, tet, n, n, n, ordering)
    if n2o is not None:
        flag = np.ascontiguousarray(flag.reshape(-1, 3)[n2o].reshape(-1))
    p = Problem(f"cube{n}", xyz, tet, myu, lam, flag, grid=(n, n, n), ordering=ordering)
    p.extra.update(E=E, nu=nuv, new2old=n2o, sigma=sigma, seed=seed)
    return p


def cantilever(nx: int = 96, ny: int = 48, nz: int = 48, seed: int = 20261204,
               ordering: str = "natural") -> Problem:
    """SIMP cantilever: unit-aspect hexes h=1/ny, x=0 clamped, densities
    rho_e ~ U(0.2, 1.0), E_e = rho_e^3, nu=0.3; a total load (0,0,-1) spread
    evenly over the central 3x3 nodes of the x=max face; no body force."""
    xyz_i, tet = mesh_tet_voxel(nx, ny, nz)
    ix, iy, iz = xyz_i[0::3].copy(), xyz_i[1::3].copy(), xyz_i[2::3].copy()
    xyz = xyz_i / float(ny)
    ne = tet.size // 4
    dens = 0.2 + 0.8 * uniform(seed, ne)
    E = dens ** 3
    myu, lam = lame(E, np.full(ne, 0.3))
    flag = np.zeros(xyz.size, np.int32)
    flag.reshape(-1, 3)[ix == 0] = 1
    f = np.zeros(xyz.size)
    cy, cz = ny // 2, nz // 2
    patch = (ix == nx) & (np.abs(iy - cy) <= 1) & (np.abs(iz - cz) <= 1)
    f.reshape(-1, 3)[patch, 2] = -1.0 / patch.sum()
    xyz, tet, n2o = _order(xyz, tet, nx, ny, nz, ordering)
    if n2o is not None:
        flag = np.ascontiguousarray(flag.reshape(-1, 3)[n2o].reshape(-1))
        f = np.ascontiguousarray(f.reshape(-1, 3)[n2o].reshape(-1))
    p = Problem(f"cantilever{nx}x{ny}x{nz}", xyz, tet, myu, lam, flag,
                g=np.zeros(3), f_ext=f, grid=(nx, ny, nz), ordering=ordering)
    p.extra.update(density=dens, E=E, nu=np.full(ne, 0.3), new2old=n2o, seed=seed)
    return p


def batch_grids(n_mesh: int = 512, lo: int = 4, hi: int = 24, seed: int = 20261205):
    """Per-mesh (nx, ny, nz) drawn uniformly from {lo..hi}^3."""
    u = uniform(seed, 3 * n_mesh).reshape(n_mesh, 3)
    return (lo + np.floor(u * (hi - lo + 1))).astype(np.int64)


def tet_volumes(xyz: np.ndarray, tet: np.ndarray) -> np.ndarray:
    """Signed tet volumes, operand for operand as femutil.cpp:202-212
    (TetVolume3D): ((a - b) + c) * (1/6) with the cofactor products in that
    order."""
    P = xyz.reshape(-1, 3)[tet.reshape(-1, 4).astype(np.int64)]
    d1, d2, d3 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]
    a = d1[:, 0] * (d2[:, 1] * d3[:, 2] - d3[:, 1] * d2[:, 2])
    b = d1[:, 1] * (d2[:, 0] * d3[:, 2] - d3[:, 0] * d2[:, 2])
    c = d1[:, 2] * (d2[:, 0] * d3[:, 1] - d3[:, 0] * d2[:, 1])
    return ((a - b) + c) * (1.0 / 6.0)


def body_force_load(xyz: np.ndarray, tet: np.ndarray, rho: float, g) -> np.ndarray:
    """Nodal load of a uniform body force rho*g: every element adds
    vol*rho*g*0.25 to each of its four nodes, elements in ascending order
    (the right-hand side fem_solidlinear.h:192-237 builds with zero
    displacement, summed in the same order)."""
    vol = tet_volumes(xyz, tet)
    b = np.zeros(xyz.size)
    nodes = tet.reshape(-1).astype(np.int64)
    for d in range(3):
        e = np.repeat(vol * rho * float(g[d]) * 0.25, 4)
        np.add.at(b, nodes * 3 + d, e)
    return b


@dataclass
class Batch:
    """Independent meshes stacked into one block-diagonal union problem.

    ``union.f_ext`` carries every mesh's body-force load (the union itself
    has g = 0); mesh m owns nodes ``node_offset[m]:node_offset[m+1]`` and
    elements ``elem_offset[m]:elem_offset[m+1]``."""
    union: Problem
    node_offset: np.ndarray   # uint32 [n+1]
    elem_offset: np.ndarray   # int64 [n+1]
    grids: np.ndarray         # int64 [n, 3]
    g: np.ndarray             # float64 [n, 3]
    index: np.ndarray         # which meshes of the full set these are

    @property
    def nmesh(self) -> int:
        return self.grids.shape[0]

    def mesh(self, m: int) -> Problem:
        """Mesh m as a standalone problem (body force rho*g_m, no f_ext)."""
        U = self.union
        n0, n1 = int(self.node_offset[m]), int(self.node_offset[m + 1])
        e0, e1 = int(self.elem_offset[m]), int(self.elem_offset[m + 1])
        tet = np.ascontiguousarray((U.tet[4 * e0:4 * e1].astype(np.int64) - n0).astype(np.uint32))
        return Problem(f"mesh{int(self.index[m])}", np.ascontiguousarray(U.xyz[3 * n0:3 * n1]), tet,
                       np.ascontiguousarray(U.myu[e0:e1]), np.ascontiguousarray(U.lam[e0:e1]),
                       np.ascontiguousarray(U.flag[3 * n0:3 * n1]), rho=U.rho,
                       g=self.g[m].copy(), grid=tuple(int(v) for v in self.grids[m]),
                       ordering=U.ordering)

    def mesh_slice(self, m: int) -> slice:
        return slice(3 * int(self.node_offset[m]), 3 * int(self.node_offset[m + 1]))


def batch(n_mesh: int = 512, lo: int = 4, hi: int = 24, seed: int = 20261205,
          sigma: float = 0.5, nu: float = 0.3, ordering: str = "color4",
          select=None) -> Batch:
    """``n_mesh`` unit-cube meshes with (nx, ny, nz) hexes drawn uniformly from
    {lo..hi}^3 (6 tets per hex), z=0 face clamped, E_e = exp(sigma*z_e)
    (lognormal), nu fixed, rho = 1 and a per-mesh body force g_m uniform in
    [-1e-3, 1e-3]^3.  Each mesh is ordered on its own (``ordering``).
    ``select`` keeps a subset of the meshes (same definitions)."""
    grids_all = batch_grids(n_mesh, lo, hi, seed)
    g_all = (2.0 * uniform(seed + 1, 3 * n_mesh) - 1.0).reshape(n_mesh, 3) * 1e-3
    index = np.arange(n_mesh) if select is None else np.asarray(select, np.int64)
    xyzs, tets, myus, lams, flags, fs = [], [], [], [], [], []
    noff, eoff = [0], [0]
    for m in index:
        nx, ny, nz = (int(v) for v in grids_all[m])
        xyz_i, tet = mesh_tet_voxel(nx, ny, nz)
        iz = xyz_i[2::3].copy()
        xyz = xyz_i.reshape(-1, 3) / np.array([nx, ny, nz], np.float64)
        xyz = np.ascontiguousarray(xyz.reshape(-1))
        ne = tet.size // 4
        E = np.exp(sigma * normal(seed + 1000003 * (int(m) + 1), ne))
        myu, lam = lame(E, np.full(ne, float(nu)))
        flag = np.zeros(xyz.size, np.int32)
        flag.reshape(-1, 3)[iz == 0] = 1
        xyz, tet, n2o = _order(xyz, tet, nx, ny, nz, ordering)
        if n2o is not None:
            flag = np.ascontiguousarray(flag.reshape(-1, 3)[n2o].reshape(-1))
        f = body_force_load(xyz, tet, 1.0, g_all[m])
        xyzs.append(xyz)
        tets.append(tet.astype(np.int64) + noff[-1])
        myus.append(myu)
        lams.append(lam)
        flags.append(flag)
        fs.append(f)
        noff.append(noff[-1] + xyz.size // 3)
        eoff.append(eoff[-1] + ne)
    union = Problem(f"batch{len(index)}", np.concatenate(xyzs),
                    np.ascontiguousarray(np.concatenate(tets).astype(np.uint32)),
                    np.concatenate(myus), np.concatenate(lams), np.concatenate(flags),
                    rho=1.0, g=np.zeros(3), f_ext=np.concatenate(fs), ordering=ordering)
    union.extra.update(seed=seed, sigma=sigma)
    return Batch(union, np.array(noff, np.uint32), np.array(eoff, np.int64),
                 grids_all[index].copy(), g_all[index].copy(), index)
