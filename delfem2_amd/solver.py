"""Python mirror of DelFEM2's solve-path API over the HIP library.

Same names and argument meaning as the reference C++ API so that callers
(and tests) read like the reference's own code:

* ``CMatrixSparse``      -- ls_block_sparse.h:24-204 (3x3 blocks, diagonal kept apart)
* ``CPreconditionerILU`` -- ls_ilu_block_sparse.h:21-60
* ``MergeLinSys_SolidLinear_Static_MeshTet3D`` -- fem_solidlinear.h:192-237
* ``MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D`` / ``..._BEuler_...`` /
  ``MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D`` -- fem_solidlinear.h:331-530
* ``RotationAtMeshPoints`` -- 603_Fem3DStiffwarp/main.cpp:80-106
* ``Solve_PBiCGStab`` -- vecxitrsol.h:350-437
* ``XPlusAYBZ / XPlusAYBZCW`` -- vecxitrsol.cpp:312-342
* ``Solve_PCG``          -- lsitrsol.h:80-148 (fused on the device)
* ``Dot / AddScaledVec / ScaleAndAddVec / setRHS_Zero / XPlusAY`` --
  view_vectorx.h:61-109, vecxitrsol.cpp:226-296

Vectors are torch tensors on the GPU (float64, contiguous): torch is only the
device allocator and stream here; every computation is a HIP kernel of
libdfm2_hip.so.  The reference's ``bool`` returns are kept (Decompose, Merge),
other failures raise ``DFM2HError``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N

_initialised = False


def init(device: int = 0) -> None:
    """Binds the library to ``cuda:device`` and to torch's current stream."""
    global _initialised
    N.load()
    torch.cuda.set_device(device)
    N.call("dfm2h_init", device)
    N.call("dfm2h_set_stream", C.c_void_p(torch.cuda.current_stream(device).cuda_stream))
    _initialised = True


def _ensure():
    if not _initialised:
        init(torch.cuda.current_device() if torch.cuda.is_available() else 0)


def _dptr(t: torch.Tensor, dtype=torch.float64) -> C.c_void_p:
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _hptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def synchronize() -> None:
    N.call("dfm2h_synchronize")


def JArray_PSuP_MeshElem(tet: np.ndarray, np_: int):
    """Sorted one-ring pattern of a tet mesh (msh_topology_uniform.cpp:301-316
    + jagarray.cpp:72-92), built on the host: (psup_ind, psup)."""
    tet = np.ascontiguousarray(tet, np.uint32)
    ind = np.zeros(np_ + 1, np.uint32)
    n = C.c_size_t(0)
    N.call("dfm2h_psup_tet", _hptr(tet), tet.size // 4, np_, _hptr(ind), None, 0, C.byref(n))
    arr = np.zeros(max(n.value, 1), np.uint32)
    N.call("dfm2h_psup_tet", _hptr(tet), tet.size // 4, np_, _hptr(ind), _hptr(arr), arr.size,
           C.byref(n))
    return ind, arr[: n.value]


def JArray_PSuP_MeshElem_Device(mesh: "TetMesh"):
    """The same pattern built on the device from a TetMesh handle (any number
    of elements around a node): (psup_ind, psup) on the host."""
    ind = np.zeros(mesh.np_ + 1, np.uint32)
    n = C.c_size_t(0)
    N.call("dfm2h_psup_tet_device", mesh.h, _hptr(ind), None, 0, C.byref(n))
    arr = np.zeros(max(n.value, 1), np.uint32)
    N.call("dfm2h_psup_tet_device", mesh.h, _hptr(ind), _hptr(arr), arr.size, C.byref(n))
    return ind, arr[: n.value]


def MultiColorOrder(psup_ind: np.ndarray, psup: np.ndarray):
    """Node relabelling for the level schedule (no reference counterpart):
    new2old with the nodes colour by colour, no two neighbours sharing a
    colour, and the colour count -- the ILU(0) sweeps' level bound in the
    relabelled mesh (`workload.permute_mesh` applies it).  The pattern must
    be symmetric without self entries, as `JArray_PSuP_MeshElem` builds it."""
    ind = np.ascontiguousarray(psup_ind, np.uint32)
    ps = np.ascontiguousarray(psup, np.uint32)
    n = ind.size - 1
    new2old = np.zeros(max(n, 1), np.uint32)
    nc = C.c_uint(0)
    N.call("dfm2h_multicolor_order", n, _hptr(ind), _hptr(ps) if ps.size else None,
           _hptr(new2old), C.byref(nc))
    return new2old[:n].astype(np.int64), int(nc.value)


class CMatrixSparse:
    """3x3 block-CSR matrix with separate diagonal blocks, values in HBM."""

    def __init__(self):
        _ensure()
        self.h = None
        self.nrowblk_ = 0
        self.nrowdim_ = 0

    def Initialize(self, nblk: int, len_: int, is_dia: bool) -> None:
        h = C.c_void_p()
        N.call("dfm2h_bsr_create", C.byref(h), int(nblk), int(len_), int(bool(is_dia)))
        self._free()
        self.h = h
        self.nrowblk_ = int(nblk)
        self.nrowdim_ = int(len_)

    def SetPattern(self, colind: np.ndarray, rowptr: np.ndarray) -> None:
        colind = np.ascontiguousarray(colind, np.uint32)
        rowptr = np.ascontiguousarray(rowptr, np.uint32)
        N.call("dfm2h_bsr_set_pattern", self.h, _hptr(colind), colind.size, _hptr(rowptr),
               rowptr.size)
        self.col_ind_ = colind.copy()
        self.row_ptr_ = rowptr.copy()

    def setZero(self) -> bool:
        N.call("dfm2h_bsr_set_zero", self.h)
        return True

    def Merge(self, aIpRow, aIpCol, emat) -> bool:
        """ls_block_sparse.h:260-310 Merge: adds emat[nrow][ncol][3][3] (host)
        at the blocks (aIpRow[i], aIpCol[j]), the diagonal block taking
        emat[i][i]; staged in call order and applied on the device before the
        values are next used.  False when a block is missing from the pattern
        (the blocks before it are added, as in the reference)."""
        ir = np.ascontiguousarray(aIpRow, np.uint32).ravel()
        ic = np.ascontiguousarray(aIpCol, np.uint32).ravel()
        em = np.ascontiguousarray(emat, np.float64)
        if em.size != ir.size * ic.size * 9:
            raise ValueError("emat must hold nrow x ncol 3x3 blocks")
        st = N.load().dfm2h_bsr_merge(self.h, ir.size, _hptr(ir), ic.size, _hptr(ic), _hptr(em))
        if st == N.ERR_PATTERN:
            return False
        N.check(st)
        return True

    def values_changed(self) -> None:
        """Declares that the device value arrays were written directly."""
        N.call("dfm2h_bsr_values_changed", self.h)

    @property
    def ncrs(self) -> int:
        n = C.c_size_t(0)
        N.call("dfm2h_bsr_info", self.h, None, C.byref(n), None, None)
        return n.value

    def MatVec(self, y: torch.Tensor, alpha: float, x: torch.Tensor, beta: float) -> None:
        N.call("dfm2h_bsr_matvec", self.h, _dptr(y), float(alpha), _dptr(x), float(beta))

    def SetFixedBC(self, bc_flag: torch.Tensor) -> None:
        N.call("dfm2h_bsr_set_fixed_bc", self.h, _dptr(bc_flag, torch.int32))

    def upload(self, val_crs: np.ndarray | None, val_dia: np.ndarray | None) -> None:
        vc = None if val_crs is None else np.ascontiguousarray(val_crs, np.float64)
        vd = None if val_dia is None else np.ascontiguousarray(val_dia, np.float64)
        N.call("dfm2h_bsr_upload", self.h, None if vc is None else _hptr(vc),
               None if vd is None else _hptr(vd))

    def download(self):
        vc = np.zeros(self.ncrs * 9)
        vd = np.zeros(self.nrowblk_ * 9)
        N.call("dfm2h_bsr_download", self.h, _hptr(vc), _hptr(vd))
        return vc, vd

    def _free(self):
        if getattr(self, "h", None):
            N.load().dfm2h_bsr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class TetMesh:
    """Tet connectivity on the device plus the element-around-node lists that
    the gather assembly walks (built once per mesh)."""

    def __init__(self, tet: np.ndarray, np_: int):
        _ensure()
        tet = np.ascontiguousarray(tet, np.uint32)
        h = C.c_void_p()
        N.call("dfm2h_tetmesh_create", C.byref(h), _hptr(tet), tet.size // 4, int(np_))
        self.h = h
        self.ntet = tet.size // 4
        self.np_ = int(np_)

    def __del__(self):
        try:
            if self.h:
                N.load().dfm2h_tetmesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


def MergeLinSys_SolidLinear_Static_MeshTet3D(mat_A: CMatrixSparse, vec_b: torch.Tensor,
                                              myu, lam, rho: float, g, aXYZ: torch.Tensor,
                                              mesh: TetMesh, aDisp: torch.Tensor | None = None):
    """fem_solidlinear.h:192-237 on the device.  ``myu``/``lam`` are scalars
    (the reference signature) or per-element float64 GPU tensors.  Returns
    True, or False when an element edge is missing from the pattern (the
    reference's Merge result)."""
    g = np.ascontiguousarray(np.asarray(g, np.float64).reshape(3))
    if isinstance(myu, torch.Tensor):
        pm, pl, sm, sl = _dptr(myu), _dptr(lam), 0.0, 0.0
    else:
        pm, pl, sm, sl = None, None, float(myu), float(lam)
    st = N.load().dfm2h_merge_linsys_solidlinear_static_tet(
        mat_A.h, _dptr(vec_b), pm, pl, sm, sl, float(rho), _hptr(g), _dptr(aXYZ), mesh.h,
        None if aDisp is None else _dptr(aDisp))
    if st == N.ERR_PATTERN:
        return False
    N.check(st)
    return True


def _lame(myu, lam):
    if isinstance(myu, torch.Tensor):
        return _dptr(myu), _dptr(lam), 0.0, 0.0
    return None, None, float(myu), float(lam)


def _opt(t):
    return None if t is None else _dptr(t)


def MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(mat_A: CMatrixSparse, vec_b: torch.Tensor,
                                                  myu, lam, rho: float, g, dt: float,
                                                  gamma_newmark: float, beta_newmark: float,
                                                  aXYZ: torch.Tensor, mesh: TetMesh,
                                                  aVal: torch.Tensor, aVelo: torch.Tensor,
                                                  aAcc: torch.Tensor) -> bool:
    """fem_solidlinear.h:331-379 on the device (is_initial_iter = true, as the
    reference driver passes): M + beta dt^2 K into A, the Newmark residual into
    vec_b.  ``myu``/``lam`` scalars or per-element tensors."""
    g = np.ascontiguousarray(np.asarray(g, np.float64).reshape(3))
    pm, pl, sm, sl = _lame(myu, lam)
    st = N.load().dfm2h_merge_linsys_solidlinear_newmark_tet(
        mat_A.h, _dptr(vec_b), pm, pl, sm, sl, float(rho), _hptr(g), float(dt),
        float(gamma_newmark), float(beta_newmark), _dptr(aXYZ), mesh.h, _opt(aVal), _opt(aVelo),
        _opt(aAcc))
    if st == N.ERR_PATTERN:
        return False
    N.check(st)
    return True


def MergeLinSys_SolidLinear_BEuler_MeshTet3D(mat_A: CMatrixSparse, vec_b: torch.Tensor, myu,
                                             lam, rho: float, g, dt: float,
                                             aXYZ: torch.Tensor, mesh: TetMesh,
                                             aDisp: torch.Tensor,
                                             aVelo: torch.Tensor) -> bool:
    """fem_solidlinear.h:381-443 on the device: K + lumped M / dt^2 into A,
    (f - K (u + dt v)) / dt into vec_b."""
    g = np.ascontiguousarray(np.asarray(g, np.float64).reshape(3))
    pm, pl, sm, sl = _lame(myu, lam)
    st = N.load().dfm2h_merge_linsys_solidlinear_beuler_tet(
        mat_A.h, _dptr(vec_b), pm, pl, sm, sl, float(rho), _hptr(g), float(dt), _dptr(aXYZ),
        mesh.h, _opt(aDisp), _opt(aVelo))
    if st == N.ERR_PATTERN:
        return False
    N.check(st)
    return True


class CPreconditionerILU:
    """Block ILU(0) with level-scheduled factorisation and sweeps."""

    def __init__(self):
        _ensure()
        h = C.c_void_p()
        N.call("dfm2h_ilu_create", C.byref(h))
        self.h = h
        self.n_singular = 0

    def SetPattern0(self, m: CMatrixSparse) -> None:
        N.call("dfm2h_ilu_set_pattern0", self.h, m.h)
        self.nblk = m.nrowblk_
        self.ncrs = m.ncrs

    def Initialize_ILUk(self, m: CMatrixSparse, fill_level: int) -> None:
        """ls_ilu_block_sparse.cpp:849-999: ILU(k) pattern (0: SetPattern0,
        -1: every fill)."""
        N.call("dfm2h_ilu_set_pattern_k", self.h, m.h, int(fill_level))
        self.nblk = m.nrowblk_
        self.ncrs = self._ncrs()

    def _ncrs(self) -> int:
        n = C.c_size_t(0)
        N.call("dfm2h_ilu_info", self.h, None, C.byref(n), None)
        return n.value

    def colind(self) -> np.ndarray:
        """Row starts of the ILU pattern (A's for ILU(0))."""
        ci = np.zeros(self.nblk + 1, np.uint32)
        N.call("dfm2h_ilu_info", self.h, None, None, _hptr(ci))
        return ci

    def CopyValue(self, m: CMatrixSparse) -> None:
        N.call("dfm2h_ilu_copy_value", self.h, m.h)

    def Decompose(self) -> bool:
        ok = C.c_int(0)
        ns = C.c_int(0)
        N.call("dfm2h_ilu_decompose", self.h, C.byref(ok), C.byref(ns))
        self.n_singular = ns.value
        return bool(ok.value)

    def SolvePrecond(self, vec: torch.Tensor) -> None:
        N.call("dfm2h_ilu_solve_precond", self.h, _dptr(vec))

    def ForwardSubstitution(self, vec: torch.Tensor) -> None:
        N.call("dfm2h_ilu_forward", self.h, _dptr(vec))

    def BackwardSubstitution(self, vec: torch.Tensor) -> None:
        N.call("dfm2h_ilu_backward", self.h, _dptr(vec))

    def levels(self):
        f = C.c_uint(0)
        b = C.c_uint(0)
        N.call("dfm2h_ilu_levels", self.h, C.byref(f), C.byref(b))
        return f.value, b.value

    def set_mixed_options(self, fp64_vectors: bool = False, update_period: int = 100) -> None:
        """Solve_PCG_Mixed options: Krylov vector precision (default FP32;
        the matrix and factors are streamed in FP32 either way) and a forced
        reliable update after `update_period` inner iterations without one
        (0: off; default 100)."""
        N.call("dfm2h_pcg_set_mixed_options", self.h, int(bool(fp64_vectors)),
               int(update_period))

    def mixed_fell_back(self) -> bool:
        """Whether the last Solve_PCG_Mixed stalled (500 inner iterations
        without a reliable update that halved the best FP64 residual) and was
        finished by the FP64 solver on the correction (no reference
        counterpart)."""
        f = C.c_int(0)
        N.call("dfm2h_pcg_mixed_fell_back", self.h, C.byref(f))
        return bool(f.value)

    def set_graph_chunks(self, large: int = 16, small: int = 2) -> None:
        """Iterations per HIP-graph launch of the fused solvers: small chunks
        at the start and near the predicted end, large ones in between (no
        reference counterpart; changes no result)."""
        N.call("dfm2h_pcg_set_graph_chunks", self.h, int(large), int(small))

    def set_scalar_fold(self, mode: int = -1) -> None:
        """Beta step of the fused solvers as a separate one-workgroup kernel
        (0), folded into the p update (1), or by system size (-1, the
        default); no reference counterpart, changes no result."""
        N.call("dfm2h_pcg_set_scalar_fold", self.h, int(mode))

    def set_sweep_shape(self, fp32: bool, shape: int) -> None:
        """Sweep launch shape of the fused solvers (-1 = autotune)."""
        N.call("dfm2h_pcg_set_shape", self.h, int(bool(fp32)), int(shape))

    def sweep_shape(self, fp32: bool):
        """(shape, name, autotuner ms per sweep pair of every shape)."""
        v = C.c_int(-1)
        name = C.c_char_p()
        ms = np.zeros(16)
        n = C.c_size_t(0)
        N.call("dfm2h_pcg_get_shape", self.h, int(bool(fp32)), C.byref(v), C.byref(name),
               _hptr(ms), ms.size, C.byref(n))
        return v.value, name.value.decode(), ms[: n.value]

    def download(self):
        rp = np.zeros(max(self.ncrs, 1), np.uint32)
        di = np.zeros(self.nblk, np.uint32)
        vc = np.zeros(max(self.ncrs * 9, 1))
        vd = np.zeros(self.nblk * 9)
        N.call("dfm2h_ilu_download", self.h, _hptr(rp), _hptr(di), _hptr(vc), _hptr(vd))
        return rp[: self.ncrs], di, vc[: self.ncrs * 9], vd

    def __del__(self):
        try:
            if self.h:
                N.load().dfm2h_ilu_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---- view_vectorx.h operator set -------------------------------------------

def Dot(x: torch.Tensor, y: torch.Tensor) -> float:
    out = C.c_double(0.0)
    N.call("dfm2h_vec_dot", _dptr(x), _dptr(y), x.numel(), C.byref(out))
    return out.value


def AddScaledVec(y: torch.Tensor, alpha: float, x: torch.Tensor) -> None:
    N.call("dfm2h_vec_add_scaled", _dptr(y), float(alpha), _dptr(x), y.numel())


def ScaleAndAddVec(y: torch.Tensor, beta: float, x: torch.Tensor) -> None:
    N.call("dfm2h_vec_scale_and_add", _dptr(y), float(beta), _dptr(x), y.numel())


def AddMatVec(lhs: torch.Tensor, scale_lhs: float, scale_rhs: float, mat: CMatrixSparse,
              rhs: torch.Tensor) -> None:
    mat.MatVec(lhs, scale_rhs, rhs, scale_lhs)


def setRHS_Zero(vec_b: torch.Tensor, aBCFlag: torch.Tensor, iflag_nonzero: int = 0) -> None:
    N.call("dfm2h_set_rhs_zero", _dptr(vec_b), _dptr(aBCFlag, torch.int32), vec_b.numel(),
           int(iflag_nonzero))


def XPlusAY(X: torch.Tensor, aBCFlag: torch.Tensor, alpha: float, Y: torch.Tensor) -> None:
    N.call("dfm2h_xplusay", _dptr(X), X.numel(), _dptr(aBCFlag, torch.int32), float(alpha),
           _dptr(Y))


def RotationAtMeshPoints(mat_A: CMatrixSparse, aXYZ: torch.Tensor,
                         aDisp: torch.Tensor | None, aR: torch.Tensor) -> None:
    """603_Fem3DStiffwarp/main.cpp:80-106 over the matrix's one-ring
    pattern: aR[9*np] receives each node's polar-decomposition rotation."""
    N.call("dfm2h_rotation_at_points", mat_A.h, _dptr(aXYZ), _opt(aDisp), _dptr(aR))


def MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(mat_A: CMatrixSparse, vec_b: torch.Tensor,
                                                myu, lam, rho: float, g, dt: float,
                                                aXYZ: torch.Tensor, mesh: TetMesh,
                                                aDisp: torch.Tensor, aVelo: torch.Tensor,
                                                aR: torch.Tensor) -> bool:
    """fem_solidlinear.h:445-530 on the device (rotations aR[9*np])."""
    g = np.ascontiguousarray(np.asarray(g, np.float64).reshape(3))
    pm, pl, sm, sl = _lame(myu, lam)
    st = N.load().dfm2h_merge_linsys_solidstiffwarp_beuler_tet(
        mat_A.h, _dptr(vec_b), pm, pl, sm, sl, float(rho), _hptr(g), float(dt), _dptr(aXYZ),
        mesh.h, _opt(aDisp), _opt(aVelo), _dptr(aR))
    if st == N.ERR_PATTERN:
        return False
    N.check(st)
    return True


def Solve_PBiCGStab(r_vec: torch.Tensor, x_vec: torch.Tensor, conv_ratio_tol: float,
                    max_niter: int, mat: CMatrixSparse, ilu: CPreconditionerILU) -> list[float]:
    """vecxitrsol.h:350-437 on device vectors: r_vec holds b and receives the
    residual, x_vec the solution; returns the relative-residual history."""
    cap = int(max_niter) + 2
    hist = np.zeros(cap)
    nh = C.c_size_t(0)
    N.call("dfm2h_solve_pbicgstab", mat.h, ilu.h, _dptr(r_vec), _dptr(x_vec),
           float(conv_ratio_tol), int(max_niter), _hptr(hist), cap, C.byref(nh))
    return hist[: nh.value].tolist()


def XPlusAYBZ(X: torch.Tensor, aBCFlag: torch.Tensor, alpha: float, Y: torch.Tensor,
              beta: float, Z: torch.Tensor) -> None:
    """vecxitrsol.cpp:312-325"""
    N.call("dfm2h_xplusaybz", _dptr(X), X.numel(), _dptr(aBCFlag, torch.int32), float(alpha),
           _dptr(Y), float(beta), _dptr(Z))


def XPlusAYBZCW(X: torch.Tensor, aBCFlag: torch.Tensor, alpha: float, Y: torch.Tensor,
                beta: float, Z: torch.Tensor, gamma: float, W: torch.Tensor) -> None:
    """vecxitrsol.cpp:327-342"""
    N.call("dfm2h_xplusaybzcw", _dptr(X), X.numel(), _dptr(aBCFlag, torch.int32), float(alpha),
           _dptr(Y), float(beta), _dptr(Z), float(gamma), _dptr(W))


def Solve_PCG(r_vec: torch.Tensor, x_vec: torch.Tensor, conv_ratio_tol: float, max_nitr: int,
              mat: CMatrixSparse, ilu: CPreconditionerILU) -> list[float]:
    """lsitrsol.h:80-148, fused on the device.  r_vec holds b and receives the
    final residual, x_vec the solution.  Returns the residual-norm history
    (absolute norms, entry 0 = |b|)."""
    cap = int(max_nitr) + 2
    hist = np.zeros(cap)
    nh = C.c_size_t(0)
    N.call("dfm2h_solve_pcg", mat.h, ilu.h, _dptr(r_vec), _dptr(x_vec), float(conv_ratio_tol),
           int(max_nitr), _hptr(hist), cap, C.byref(nh))
    return hist[: nh.value].tolist()


def pcg_kernel_times(mat: CMatrixSparse, ilu: CPreconditionerILU, fp32: bool,
                     nrep: int = 20) -> dict:
    """Average ms per launch of each fused-iteration kernel class (HIP events
    on the solver stream), after a solve of that precision."""
    ms = np.zeros(8)
    N.call("dfm2h_pcg_kernel_times", mat.h, ilu.h, int(bool(fp32)), int(nrep), _hptr(ms))
    out = dict(zip(["spmv", "update", "fwd_sweep", "bwd_sweep", "pupdate", "alpha_step",
                    "spmv_back_to_back"], ms[:7].tolist()))
    out["folded"] = bool(ms[7] != 0.0)
    return out


def measure_copy_gbs(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Read + write rate (GB/s) of a plain 16-byte-vector device copy of
    `nbytes` bytes on the solver stream: the achievable bandwidth the kernels'
    rates are compared with."""
    out = np.zeros(1)
    N.call("dfm2h_measure_copy", int(nbytes), int(reps), _hptr(out))
    return float(out[0])


def Solve_PCG_Mixed(b_vec: torch.Tensor, x_vec: torch.Tensor, conv_ratio_tol: float,
                    mat: CMatrixSparse, ilu: CPreconditionerILU, delta: float = 1e-1,
                    max_outer: int = 200, max_inner: int = 20000):
    """FP32 ILU(0)-PCG inside FP64 refinement by reliable updates (no
    reference counterpart): refine whenever the FP32 residual fell by `delta`;
    when the refinement stalls, the FP64 solver finishes (ilu.mixed_fell_back()).
    Returns (FP64 residual-norm history, FP32 iterations plus, after a
    fallback, the FP64 solver's)."""
    cap = int(max_outer) + int(max_inner) + 4  # room for a fallback's FP64 history
    hist = np.zeros(cap)
    nh = C.c_size_t(0)
    ni = C.c_size_t(0)
    N.call("dfm2h_solve_pcg_mixed", mat.h, ilu.h, _dptr(b_vec), _dptr(x_vec),
           float(conv_ratio_tol), float(delta), int(max_outer), int(max_inner), _hptr(hist),
           cap, C.byref(nh), C.byref(ni))
    return hist[: nh.value].tolist(), ni.value


class BatchSolver:
    """Many independent meshes solved together (no reference counterpart).

    ``mat`` / ``ilu`` hold the union system (assembled, BCs set, decomposed
    with the ordinary calls); mesh m owns nodes ``node_offset[m] ..
    node_offset[m+1]-1``.  Each mesh runs its own ILU(0)-PCG (lsitrsol.h:80-148)
    inside shared kernels; converged meshes drop out.
    """

    def __init__(self, mat: CMatrixSparse, ilu: CPreconditionerILU, node_offset):
        _ensure()
        off = np.ascontiguousarray(node_offset, np.uint32)
        h = C.c_void_p()
        N.call("dfm2h_batch_create", C.byref(h), mat.h, ilu.h, _hptr(off), off.size - 1)
        self.h = h
        self.nmesh = off.size - 1
        self._keep = (mat, ilu)  # the batch refers to both

    def info(self) -> dict:
        nm, ns = C.c_size_t(0), C.c_size_t(0)
        f, b = C.c_uint(0), C.c_uint(0)
        N.call("dfm2h_batch_info", self.h, C.byref(nm), C.byref(ns), C.byref(f), C.byref(b))
        return {"nmesh": nm.value, "nslices": ns.value, "levels": (f.value, b.value)}

    def solve(self, b: torch.Tensor, x: torch.Tensor, tol: float, fp32: bool = False,
              delta: float = 1e-1, max_nitr: int = 20000):
        """Returns (iterations, relative residuals, state) per mesh; state is 1
        converged, 0 iteration cap, -1 NaN."""
        it = np.zeros(self.nmesh, np.uint32)
        rr = np.zeros(self.nmesh)
        st = np.zeros(self.nmesh, np.int32)
        N.call("dfm2h_batch_solve", self.h, _dptr(b), _dptr(x), float(tol), int(bool(fp32)),
               float(delta), int(max_nitr), _hptr(it), _hptr(rr), _hptr(st))
        return it, rr, st

    def fell_back(self) -> np.ndarray:
        """Per mesh: whether the last FP32-form solve stalled on it and
        finished it with the FP64 form (no reference counterpart)."""
        f = np.zeros(self.nmesh, np.int32)
        N.call("dfm2h_batch_fell_back", self.h, _hptr(f))
        return f.astype(bool)

    def tune(self, fp32: bool):
        """Runs the launch-shape autotuner; returns (best variant, ms per
        iteration of every variant)."""
        best = C.c_int(-1)
        ms = np.zeros(16)
        nv = C.c_size_t(0)
        N.call("dfm2h_batch_tune", self.h, int(bool(fp32)), C.byref(best), _hptr(ms), ms.size,
               C.byref(nv))
        return best.value, ms[: nv.value]

    def set_variant(self, fp32: bool, variant: int) -> None:
        N.call("dfm2h_batch_set_variant", self.h, int(bool(fp32)), int(variant))

    def set_update_period(self, period: int) -> None:
        """FP32 form: forced reliable update after `period` inner iterations
        without one (0: off; default 25)."""
        N.call("dfm2h_batch_set_update_period", self.h, int(period))

    def variant(self, fp32: bool):
        v = C.c_int(-1)
        name = C.c_char_p()
        N.call("dfm2h_batch_get_variant", self.h, int(bool(fp32)), C.byref(v), C.byref(name))
        return v.value, name.value.decode()

    def __del__(self):
        try:
            if self.h:
                N.load().dfm2h_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
