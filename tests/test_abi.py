"""The C-ABI library loads on a machine without a GPU and exports every
entry point the header declares; host-only entry points run here."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from delfem2_amd import _native as N
from delfem2_amd import workload as W


def test_library_exports_every_header_symbol():
    L = N.load()
    declared = N.header_symbols()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    # the Python bindings cover the whole header, nothing more
    assert sorted(N.SIGNATURES) == declared


def test_status_codes_match_header():
    txt = open(N.HEADER).read()
    for name, val in [("OK", 0), ("ERR_INVALID", 1), ("ERR_HIP", 2), ("ERR_PATTERN", 3),
                      ("ERR_SINGULAR", 4), ("ERR_NOT_CONVERGED", 5), ("ERR_NAN", 6),
                      ("ERR_STATE", 7)]:
        assert f"#define DFM2H_{name} {val}" in txt
        assert getattr(N, name) == val


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 5), (10, 10, 10)])
def test_host_pattern_matches_oracle(dims):
    xyz, tet = W.mesh_tet_voxel(*dims)
    np_ = xyz.size // 3
    ind = np.zeros(np_ + 1, np.uint32)
    n = C.c_size_t(0)
    N.call("dfm2h_psup_tet", C.c_void_p(tet.ctypes.data), tet.size // 4, np_,
           C.c_void_p(ind.ctypes.data), None, 0, C.byref(n))
    arr = np.zeros(n.value, np.uint32)
    N.call("dfm2h_psup_tet", C.c_void_p(tet.ctypes.data), tet.size // 4, np_,
           C.c_void_p(ind.ctypes.data), C.c_void_p(arr.ctypes.data), arr.size, C.byref(n))
    ind2, arr2 = oracle.psup(tet, np_)
    assert np.array_equal(ind, ind2) and np.array_equal(arr, arr2)


def test_host_pattern_rejects_bad_input():
    tet = np.array([0, 1, 2, 9], np.uint32)
    ind = np.zeros(5, np.uint32)
    st = N.load().dfm2h_psup_tet(C.c_void_p(tet.ctypes.data), 1, 4, C.c_void_p(ind.ctypes.data),
                                 None, 0, None)
    assert st == N.ERR_INVALID
    assert b"out of range" in N.load().dfm2h_last_error()
    # too small a capacity is an error, not an overflow
    xyz, tet = W.mesh_tet_voxel(1, 1, 1)
    ind = np.zeros(9, np.uint32)
    arr = np.zeros(3, np.uint32)
    st = N.load().dfm2h_psup_tet(C.c_void_p(tet.ctypes.data), 6, 8, C.c_void_p(ind.ctypes.data),
                                 C.c_void_p(arr.ctypes.data), arr.size, None)
    assert st == N.ERR_INVALID


def test_null_handles_are_errors_not_crashes():
    L = N.load()
    assert L.dfm2h_bsr_set_zero(None) == N.ERR_INVALID
    assert L.dfm2h_ilu_solve_precond(None, None) == N.ERR_INVALID
    assert L.dfm2h_solve_pcg(None, None, None, None, 1e-8, 10, None, 0, None) == N.ERR_INVALID
    out = C.c_void_p()
    assert L.dfm2h_bsr_create(C.byref(out), 10, 2, 1) == N.ERR_INVALID  # only 3x3 blocks


def test_product_never_imports_oracle():
    root = os.path.join(os.path.dirname(N.HERE), "delfem2_amd")
    for dp, _, files in os.walk(root):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dp, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
                assert "liboracle" not in txt and "libdfm2ref" not in txt, f
