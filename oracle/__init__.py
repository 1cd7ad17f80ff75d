"""CPU oracle for the tet-FEM / BSR / ILU(0) / PCG path -- TEST INFRASTRUCTURE.

Two ctypes-loaded libraries, both built by ``oracle/Makefile``:

* ``liboracle.so`` -- the C restatement in ``oracle.c`` (the checker);
* ``_ref/libdfm2ref.so`` -- the reference's own headers compiled through
  ``ref_shim.cpp`` (present only where the reference tree was available at
  build time; the built file travels with the tree).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module.  The product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libdfm2ref.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_mesh_tet_voxel.argtypes = [C.c_uint, C.c_uint, C.c_uint, _f64p, _u32p]
        L.orc_psup.argtypes = [_u32p, C.c_size_t, C.c_uint, C.c_size_t, _u32p, _vp]
        L.orc_psup.restype = C.c_size_t
        L.orc_tet_volume.argtypes = [_f64p] * 4
        L.orc_tet_volume.restype = C.c_double
        L.orc_tet_dldx.argtypes = [_f64p] * 5
        L.orc_emat_tet.argtypes = [_f64p, _f64p, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_assemble_tet_static.argtypes = [C.c_uint, _u32p, _u32p, _f64p, _f64p, _f64p,
                                              _f64p, _f64p, C.c_double, _f64p, _f64p, _u32p,
                                              C.c_size_t, _vp]
        L.orc_assemble_tet_static.restype = C.c_int
        L.orc_set_fixed_bc.argtypes = [C.c_uint, _u32p, _u32p, _f64p, _f64p, _i32p]
        L.orc_set_rhs_zero.argtypes = [_f64p, _i32p, C.c_size_t, C.c_int]
        L.orc_matvec.argtypes = [C.c_uint, _u32p, _u32p, _f64p, _f64p, _f64p, C.c_double,
                                 _f64p, C.c_double]
        L.orc_ilu_set_pattern0.argtypes = [C.c_uint, _u32p, _u32p, _u32p, _u32p]
        L.orc_ilu_pattern_k.argtypes = [C.c_uint, _u32p, _u32p, C.c_int, _u32p, _vp, _vp,
                                        C.c_size_t]
        L.orc_ilu_pattern_k.restype = C.c_size_t
        L.orc_ilu_copy_value.argtypes = [C.c_uint, _u32p, _u32p, _f64p, _f64p, _u32p, _u32p,
                                         _f64p, _f64p]
        L.orc_ilu_decompose.argtypes = [C.c_uint, _u32p, _u32p, _u32p, _f64p, _f64p,
                                        C.POINTER(C.c_int)]
        L.orc_ilu_decompose.restype = C.c_int
        L.orc_ilu_forward.argtypes = [C.c_uint, _u32p, _u32p, _u32p, _f64p, _f64p, _f64p]
        L.orc_ilu_backward.argtypes = [C.c_uint, _u32p, _u32p, _u32p, _f64p, _f64p]
        L.orc_dot.argtypes = [_f64p, _f64p, C.c_size_t]
        L.orc_dot.restype = C.c_double
        L.orc_solve_pcg.argtypes = [_f64p, _f64p, C.c_double, C.c_uint, C.c_uint, _u32p, _u32p,
                                    _f64p, _f64p, _u32p, _u32p, _u32p, _f64p, _f64p, _f64p,
                                    C.c_size_t]
        L.orc_solve_pcg.restype = C.c_size_t
        L.orc_xplusay.argtypes = [_f64p, C.c_size_t, _i32p, C.c_double, _f64p]
        L.orc_adjoint_grad_tet.argtypes = [_f64p, _f64p, _f64p, _f64p, _u32p, C.c_size_t,
                                           _f64p, _f64p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref():
    """The compiled reference (raises if it was never built here)."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF):
            raise FileNotFoundError(_REF)
        R = C.CDLL(_REF)
        R.ref_mesh_tet_voxel.argtypes = [C.c_uint, C.c_uint, C.c_uint, _f64p, _u32p]
        R.ref_psup.argtypes = [_u32p, C.c_size_t, C.c_size_t, _u32p, _vp]
        R.ref_psup.restype = C.c_size_t
        R.ref_create.argtypes = [C.c_uint, _u32p, _u32p]
        R.ref_create.restype = _vp
        R.ref_destroy.argtypes = [_vp]
        R.ref_get_matrix.argtypes = [_vp, _f64p, _f64p]
        R.ref_set_matrix.argtypes = [_vp, _f64p, _f64p]
        R.ref_assemble_uniform.argtypes = [_vp, _f64p, C.c_double, C.c_double, C.c_double,
                                           _f64p, _f64p, C.c_size_t, _u32p, C.c_size_t, _f64p]
        R.ref_assemble_per_element.argtypes = [_vp, _f64p, _f64p, _f64p, C.c_double, _f64p,
                                               _f64p, _u32p, C.c_size_t, _f64p]
        R.ref_assemble_per_element.restype = C.c_int
        R.ref_emat_tet.argtypes = [_f64p, _f64p, C.c_double, C.c_double, _f64p, _f64p]
        R.ref_tet_dldx.argtypes = [_f64p] * 5
        R.ref_set_fixed_bc.argtypes = [_vp, _i32p]
        R.ref_set_rhs_zero.argtypes = [_f64p, _i32p, C.c_size_t, C.c_int]
        R.ref_matvec.argtypes = [_vp, _f64p, C.c_double, _f64p, C.c_double]
        R.ref_ilu_factor.argtypes = [_vp]
        R.ref_ilu_factor.restype = C.c_int
        R.ref_ilu_init_k.argtypes = [_vp, C.c_int, _vp]
        R.ref_ilu_init_k.restype = C.c_size_t
        R.ref_ilu_get.argtypes = [_vp, _vp, _vp, _vp, _vp]
        R.ref_ilu_forward.argtypes = [_vp, _f64p]
        R.ref_ilu_backward.argtypes = [_vp, _f64p]
        R.ref_ilu_solve_precond.argtypes = [_vp, _f64p]
        R.ref_solve_pcg.argtypes = [_vp, _f64p, _f64p, C.c_double, C.c_uint, _f64p, C.c_size_t]
        R.ref_solve_pcg.restype = C.c_size_t
        R.ref_xplusay.argtypes = [_f64p, C.c_size_t, _i32p, C.c_double, _f64p]
        _ref = R
    return _ref


# ----------------------------------------------------------------------------
# numpy-level helpers over liboracle.so


def mesh_tet_voxel(nx: int, ny: int, nz: int):
    """Integer-coordinate voxel tet mesh (gridvoxel.cpp:194-271)."""
    np_ = (nx + 1) * (ny + 1) * (nz + 1)
    xyz = np.zeros(np_ * 3, np.float64)
    tet = np.zeros(nx * ny * nz * 24, np.uint32)
    lib().orc_mesh_tet_voxel(nx, ny, nz, xyz, tet)
    return xyz, tet


def psup(tet: np.ndarray, np_: int):
    """Sorted one-ring pattern (msh_topology_uniform.cpp:301-316 + JArray_Sort)."""
    tet = np.ascontiguousarray(tet, np.uint32)
    ind = np.zeros(np_ + 1, np.uint32)
    n = lib().orc_psup(tet, tet.size // 4, 4, np_, ind, None)
    arr = np.zeros(max(int(n), 1), np.uint32)
    lib().orc_psup(tet, tet.size // 4, 4, np_, ind, arr.ctypes.data)
    return ind, arr[:n]


class System:
    """CPU system: BSR matrix A (+ rhs) and its ILU(0) factor, in numpy arrays."""

    def __init__(self, colind: np.ndarray, rowptr: np.ndarray):
        self.nblk = int(colind.size - 1)
        self.colind = np.ascontiguousarray(colind, np.uint32)
        self.rowptr = np.ascontiguousarray(rowptr, np.uint32)
        ncrs = int(self.colind[-1])
        self.vcrs = np.zeros(ncrs * 9, np.float64)
        self.vdia = np.zeros(self.nblk * 9, np.float64)
        self.l_rowptr = np.zeros(max(ncrs, 1), np.uint32)
        self.l_diaind = np.zeros(self.nblk, np.uint32)
        lib().orc_ilu_set_pattern0(self.nblk, self.colind, self.rowptr, self.l_rowptr,
                                   self.l_diaind)
        self.l_rowptr = self.l_rowptr[:ncrs].copy() if ncrs else self.l_rowptr
        self.l_colind = self.colind  # the ILU(0) pattern has A's rows
        self.l_vcrs = np.zeros(ncrs * 9, np.float64)
        self.l_vdia = np.zeros(self.nblk * 9, np.float64)
        self.n_singular = 0

    def set_fill_level(self, lev_fill: int) -> None:
        """Initialize_ILUk (ls_ilu_block_sparse.cpp:849-999): the ILU(k)
        pattern (lev_fill 0: A's pattern; -1: every fill)."""
        if lev_fill == 0:
            ncrs = int(self.colind[-1])
            self.l_rowptr = np.zeros(max(ncrs, 1), np.uint32)
            self.l_diaind = np.zeros(self.nblk, np.uint32)
            lib().orc_ilu_set_pattern0(self.nblk, self.colind, self.rowptr, self.l_rowptr,
                                       self.l_diaind)
            self.l_rowptr = self.l_rowptr[:ncrs].copy() if ncrs else self.l_rowptr
            self.l_colind = self.colind
        else:
            ci = np.zeros(self.nblk + 1, np.uint32)
            n = lib().orc_ilu_pattern_k(self.nblk, self.colind, self.rowptr, int(lev_fill), ci,
                                        None, None, 0)
            rp = np.zeros(max(n, 1), np.uint32)
            di = np.zeros(self.nblk, np.uint32)
            lib().orc_ilu_pattern_k(self.nblk, self.colind, self.rowptr, int(lev_fill), ci,
                                    rp.ctypes.data, di.ctypes.data, rp.size)
            self.l_colind, self.l_rowptr, self.l_diaind = ci, rp[:n].copy(), di
            ncrs = n
        self.l_vcrs = np.zeros(ncrs * 9, np.float64)
        self.l_vdia = np.zeros(self.nblk * 9, np.float64)

    @property
    def ndof(self) -> int:
        return self.nblk * 3

    def assemble(self, b, myu_e, lambda_e, rho, g, xyz, tet, disp=None):
        d = None if disp is None else np.ascontiguousarray(disp, np.float64)
        st = lib().orc_assemble_tet_static(
            self.nblk, self.colind, self.rowptr, self.vcrs, self.vdia, b,
            np.ascontiguousarray(myu_e, np.float64), np.ascontiguousarray(lambda_e, np.float64),
            float(rho), np.ascontiguousarray(g, np.float64), np.ascontiguousarray(xyz, np.float64),
            np.ascontiguousarray(tet, np.uint32), tet.size // 4,
            None if d is None else d.ctypes.data)
        return st

    def set_fixed_bc(self, flag):
        lib().orc_set_fixed_bc(self.nblk, self.colind, self.rowptr, self.vcrs, self.vdia,
                               np.ascontiguousarray(flag, np.int32))

    def matvec(self, y, alpha, x, beta):
        lib().orc_matvec(self.nblk, self.colind, self.rowptr, self.vcrs, self.vdia, y,
                         float(alpha), x, float(beta))

    def factor(self) -> bool:
        lib().orc_ilu_copy_value(self.nblk, self.l_colind, self.l_rowptr, self.l_vcrs,
                                 self.l_vdia, self.colind, self.rowptr, self.vcrs, self.vdia)
        ns = C.c_int(0)
        ok = lib().orc_ilu_decompose(self.nblk, self.l_colind, self.l_rowptr, self.l_diaind,
                                     self.l_vcrs, self.l_vdia, C.byref(ns))
        self.n_singular = ns.value
        return bool(ok)

    def forward(self, vec):
        lib().orc_ilu_forward(self.nblk, self.l_colind, self.l_rowptr, self.l_diaind,
                              self.l_vcrs, self.l_vdia, vec)

    def backward(self, vec):
        lib().orc_ilu_backward(self.nblk, self.l_colind, self.l_rowptr, self.l_diaind,
                               self.l_vcrs, vec)

    def solve_precond(self, vec):
        self.forward(vec)
        self.backward(vec)

    def solve_pcg(self, r, tol, max_nitr):
        """Returns (x, history); r is overwritten with the final residual."""
        x = np.zeros(self.ndof, np.float64)
        cap = max_nitr + 2
        hist = np.zeros(cap, np.float64)
        n = lib().orc_solve_pcg(r, x, float(tol), int(max_nitr), self.nblk, self.colind,
                                self.rowptr, self.vcrs, self.vdia, self.l_colind, self.l_rowptr,
                                self.l_diaind, self.l_vcrs, self.l_vdia, hist, cap)
        return x, hist[:n].copy()


class RefSystem:
    """The same system on the compiled reference (CMatrixSparse + CPreconditionerILU)."""

    def __init__(self, colind, rowptr):
        self.colind = np.ascontiguousarray(colind, np.uint32)
        self.rowptr = np.ascontiguousarray(rowptr, np.uint32)
        self.nblk = int(self.colind.size - 1)
        self.h = ref().ref_create(self.nblk, self.colind, self.rowptr)

    def __del__(self):
        if getattr(self, "h", None) and _ref is not None:
            _ref.ref_destroy(self.h)
            self.h = None

    @property
    def ndof(self):
        return self.nblk * 3

    def matrix(self):
        ncrs = int(self.colind[-1])
        vcrs = np.zeros(ncrs * 9)
        vdia = np.zeros(self.nblk * 9)
        ref().ref_get_matrix(self.h, vcrs, vdia)
        return vcrs, vdia

    def set_matrix(self, vcrs, vdia):
        ref().ref_set_matrix(self.h, np.ascontiguousarray(vcrs), np.ascontiguousarray(vdia))

    def assemble(self, b, myu_e, lambda_e, rho, g, xyz, tet, disp=None):
        if disp is None:
            disp = np.zeros(self.ndof)
        return ref().ref_assemble_per_element(
            self.h, b, np.ascontiguousarray(myu_e, np.float64),
            np.ascontiguousarray(lambda_e, np.float64), float(rho),
            np.ascontiguousarray(g, np.float64), np.ascontiguousarray(xyz, np.float64),
            np.ascontiguousarray(tet, np.uint32), tet.size // 4,
            np.ascontiguousarray(disp, np.float64))

    def assemble_uniform(self, b, myu, lam, rho, g, xyz, tet, disp=None):
        if disp is None:
            disp = np.zeros(self.ndof)
        ref().ref_assemble_uniform(self.h, b, float(myu), float(lam), float(rho),
                                   np.ascontiguousarray(g, np.float64),
                                   np.ascontiguousarray(xyz, np.float64), xyz.size // 3,
                                   np.ascontiguousarray(tet, np.uint32), tet.size // 4,
                                   np.ascontiguousarray(disp, np.float64))

    def set_fixed_bc(self, flag):
        ref().ref_set_fixed_bc(self.h, np.ascontiguousarray(flag, np.int32))

    def matvec(self, y, alpha, x, beta):
        ref().ref_matvec(self.h, y, float(alpha), x, float(beta))

    def factor(self) -> bool:
        return bool(ref().ref_ilu_factor(self.h))

    def set_fill_level(self, lev_fill: int):
        """Initialize_ILUk on the reference; returns the ILU colind."""
        ci = np.zeros(self.nblk + 1, np.uint32)
        self.l_ncrs = int(ref().ref_ilu_init_k(self.h, int(lev_fill), ci.ctypes.data))
        self.l_colind = ci
        return ci

    def ilu(self):
        ncrs = getattr(self, "l_ncrs", int(self.colind[-1]))
        rp = np.zeros(max(ncrs, 1), np.uint32)
        di = np.zeros(self.nblk, np.uint32)
        vc = np.zeros(max(ncrs * 9, 1))
        vd = np.zeros(self.nblk * 9)
        ref().ref_ilu_get(self.h, rp.ctypes.data, di.ctypes.data, vc.ctypes.data, vd.ctypes.data)
        return rp[:ncrs], di, vc[:ncrs * 9], vd

    def forward(self, vec):
        ref().ref_ilu_forward(self.h, vec)

    def backward(self, vec):
        ref().ref_ilu_backward(self.h, vec)

    def solve_precond(self, vec):
        ref().ref_ilu_solve_precond(self.h, vec)

    def solve_pcg(self, r, tol, max_nitr):
        x = np.zeros(self.ndof)
        cap = max_nitr + 2
        hist = np.zeros(cap)
        n = ref().ref_solve_pcg(self.h, r, x, float(tol), int(max_nitr), hist, cap)
        return x, hist[:n].copy()


def set_rhs_zero(b, flag, iflag_nonzero=0):
    lib().orc_set_rhs_zero(b, np.ascontiguousarray(flag, np.int32), b.size, int(iflag_nonzero))


def xplusay(X, flag, alpha, Y):
    lib().orc_xplusay(X, X.size, np.ascontiguousarray(flag, np.int32), float(alpha), Y)


def adjoint_grad(mu1_e, lambda1_e, xyz, tet, u, lam):
    ne = tet.size // 4
    g = np.zeros(ne)
    lib().orc_adjoint_grad_tet(g, np.ascontiguousarray(mu1_e, np.float64),
                               np.ascontiguousarray(lambda1_e, np.float64),
                               np.ascontiguousarray(xyz, np.float64),
                               np.ascontiguousarray(tet, np.uint32), ne,
                               np.ascontiguousarray(u, np.float64),
                               np.ascontiguousarray(lam, np.float64))
    return g
