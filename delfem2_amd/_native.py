"""ctypes bindings of ``libdfm2_hip.so`` (declared in include/dfm2_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DFM2H_LIBRARY", os.path.join(HERE, "libdfm2_hip.so"))
HEADER = os.path.join(os.path.dirname(HERE), "include", "dfm2_hip.h")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_PATTERN = 3
ERR_SINGULAR = 4
ERR_NOT_CONVERGED = 5
ERR_NAN = 6
ERR_STATE = 7


class DFM2HError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dfm2h status {code}] {msg}")
        self.code = code


_vp = C.c_void_p
_sz = C.c_size_t
_d = C.c_double
_u = C.c_uint
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pu = C.POINTER(C.c_uint)
_pi = C.POINTER(C.c_int)
_psz = C.POINTER(C.c_size_t)
_pvp = C.POINTER(C.c_void_p)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "dfm2h_init": [_i],
    "dfm2h_set_stream": [_vp],
    "dfm2h_get_stream": [],
    "dfm2h_synchronize": [],
    "dfm2h_last_error": [],
    "dfm2h_malloc": [_pvp, _sz],
    "dfm2h_free": [_vp],
    "dfm2h_memcpy_htod": [_vp, _vp, _sz],
    "dfm2h_memcpy_dtoh": [_vp, _vp, _sz],
    "dfm2h_memcpy_dtod": [_vp, _vp, _sz],
    "dfm2h_psup_tet": [_vp, _sz, _sz, _vp, _vp, _sz, _psz],
    "dfm2h_multicolor_order": [_sz, _vp, _vp, _vp, _vp],
    "dfm2h_bsr_create": [_pvp, _sz, _u, _i],
    "dfm2h_bsr_destroy": [_vp],
    "dfm2h_bsr_set_pattern": [_vp, _vp, _sz, _vp, _sz],
    "dfm2h_bsr_set_zero": [_vp],
    "dfm2h_bsr_upload": [_vp, _vp, _vp],
    "dfm2h_bsr_download": [_vp, _vp, _vp],
    "dfm2h_bsr_info": [_vp, _psz, _psz, _pvp, _pvp],
    "dfm2h_bsr_values_changed": [_vp],
    "dfm2h_bsr_merge": [_vp, _u, _vp, _u, _vp, _vp],
    "dfm2h_bsr_matvec": [_vp, _vp, _d, _vp, _d],
    "dfm2h_bsr_set_fixed_bc": [_vp, _vp],
    "dfm2h_set_rhs_zero": [_vp, _vp, _sz, _i],
    "dfm2h_xplusay": [_vp, _sz, _vp, _d, _vp],
    "dfm2h_tetmesh_create": [_pvp, _vp, _sz, _sz],
    "dfm2h_psup_tet_device": [_vp, _vp, _vp, _sz, _psz],
    "dfm2h_tetmesh_destroy": [_vp],
    "dfm2h_merge_linsys_solidlinear_static_tet": [_vp, _vp, _vp, _vp, _d, _d, _d, _vp, _vp,
                                                  _vp, _vp],
    "dfm2h_merge_linsys_solidlinear_newmark_tet": [_vp, _vp, _vp, _vp, _d, _d, _d, _vp, _d,
                                                   _d, _d, _vp, _vp, _vp, _vp, _vp],
    "dfm2h_merge_linsys_solidlinear_beuler_tet": [_vp, _vp, _vp, _vp, _d, _d, _d, _vp, _d,
                                                  _vp, _vp, _vp, _vp],
    "dfm2h_merge_linsys_solidstiffwarp_beuler_tet": [_vp, _vp, _vp, _vp, _d, _d, _d, _vp,
                                                     _d, _vp, _vp, _vp, _vp, _vp],
    "dfm2h_rotation_at_points": [_vp, _vp, _vp, _vp],
    "dfm2h_xplusaybz": [_vp, _sz, _vp, _d, _vp, _d, _vp],
    "dfm2h_solve_pbicgstab": [_vp, _vp, _vp, _vp, _d, _u, _vp, _sz, _psz],
    "dfm2h_xplusaybzcw": [_vp, _sz, _vp, _d, _vp, _d, _vp, _d, _vp],
    "dfm2h_ilu_create": [_pvp],
    "dfm2h_ilu_destroy": [_vp],
    "dfm2h_ilu_set_pattern0": [_vp, _vp],
    "dfm2h_ilu_set_pattern_k": [_vp, _vp, _i],
    "dfm2h_ilu_info": [_vp, _psz, _psz, _vp],
    "dfm2h_ilu_copy_value": [_vp, _vp],
    "dfm2h_ilu_decompose": [_vp, _pi, _pi],
    "dfm2h_ilu_solve_precond": [_vp, _vp],
    "dfm2h_ilu_forward": [_vp, _vp],
    "dfm2h_ilu_backward": [_vp, _vp],
    "dfm2h_ilu_download": [_vp, _vp, _vp, _vp, _vp],
    "dfm2h_ilu_levels": [_vp, _pu, _pu],
    "dfm2h_vec_dot": [_vp, _vp, _sz, _pd],
    "dfm2h_vec_dot_noalloc": [_vp, _vp, _sz, _pd],
    "dfm2h_vec_add_scaled": [_vp, _d, _vp, _sz],
    "dfm2h_vec_scale_and_add": [_vp, _d, _vp, _sz],
    "dfm2h_vec_copy": [_vp, _vp, _sz],
    "dfm2h_vec_set_zero": [_vp, _sz],
    "dfm2h_solve_pcg": [_vp, _vp, _vp, _vp, _d, _u, _vp, _sz, _psz],
    "dfm2h_solve_pcg_mixed": [_vp, _vp, _vp, _vp, _d, _d, _u, _u, _vp, _sz, _psz, _psz],
    "dfm2h_adjoint_grad_tet": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dfm2h_pcg_kernel_times": [_vp, _vp, _i, _u, _vp],
    "dfm2h_measure_copy": [_sz, _u, _vp],
    "dfm2h_pcg_set_shape": [_vp, _i, _i],
    "dfm2h_pcg_set_mixed_options": [_vp, _i, _u],
    "dfm2h_pcg_mixed_fell_back": [_vp, _pi],
    "dfm2h_pcg_set_graph_chunks": [_vp, _u, _u],
    "dfm2h_pcg_set_scalar_fold": [_vp, _i],
    "dfm2h_pcg_get_shape": [_vp, _i, _pi, C.POINTER(C.c_char_p), _vp, _sz, _psz],
    "dfm2h_batch_create": [_pvp, _vp, _vp, _vp, _sz],
    "dfm2h_batch_destroy": [_vp],
    "dfm2h_batch_info": [_vp, _psz, _psz, _pu, _pu],
    "dfm2h_batch_solve": [_vp, _vp, _vp, _d, _i, _d, _u, _vp, _vp, _vp],
    "dfm2h_batch_tune": [_vp, _i, _pi, _vp, _sz, _psz],
    "dfm2h_batch_set_variant": [_vp, _i, _i],
    "dfm2h_batch_set_update_period": [_vp, _u],
    "dfm2h_batch_get_variant": [_vp, _i, _pi, C.POINTER(C.c_char_p)],
    "dfm2h_batch_fell_back": [_vp, _vp],
}
_RESTYPE = {"dfm2h_last_error": C.c_char_p, "dfm2h_get_stream": _vp}

_lib = None


def header_symbols() -> list[str]:
    """Every function the C header declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dfm2h_[a-z0-9_]+)\s*\(", txt)))


def load(path: str = LIB_PATH):
    """Loads the library (no device call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m delfem2_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(path)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    _lib = L
    return L


def check(status: int) -> None:
    if status != OK:
        msg = load().dfm2h_last_error()
        raise DFM2HError(status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
