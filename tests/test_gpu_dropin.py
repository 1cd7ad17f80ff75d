"""Drop-in binaries built from tests/cpp against the reference's headers:
the reference's own Solve_PCG template on the device types of
include/delfem2_hip.hpp (tests/cpp/dropin_pcg.cpp), the reference's facade
beside its device twin (facade.cpp), and the reference's time-stepping
sequences beside the same calls on device types (dynamics.cpp)."""
import json
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "dropin_pcg")

pytestmark = pytest.mark.gpu


def test_reference_template_on_device_types(gpu):
    if not os.path.exists(BIN):
        pytest.skip("drop-in binary not built (needs the reference headers at build time)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["template_iterations"] == 44  # the reference's count on this problem
    assert abs(out["fused_iterations"] - 44) <= 1
    assert out["x_rel_diff"] < 1e-9


FACADE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "facade")


def test_facade_against_reference_facade(gpu):
    """ls_solver_block_sparse_ilu.h's LinearSystemSolver_BlockSparseILU
    (reference, host) and its device twin driven by the same caller code:
    merged values and rhs bit-identical; Solve_PcgIlu x rel-L2 < 1e-9;
    Solve_Cg x rel-L2 < 1e-8 and the same iteration count (+-1)."""
    if not os.path.exists(FACADE):
        pytest.skip("facade binary not built (needs the reference headers at build time)")
    r = subprocess.run([FACADE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["same_values"] and out["same_rhs"]
    assert out["pcg_x_rel"] < 1e-9 and out["cg_x_rel"] < 1e-8
    assert abs(out["cg_iterations_ref"] - out["cg_iterations_dev"]) <= 1


DYN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "dynamics")


def test_time_stepping_against_reference(gpu):
    """The reference's Newmark-beta step (618_Fem3Tet) and stiffness-warping
    step (603_Fem3DStiffwarp), compiled from its headers on the host, beside
    the same sequences on the delfem2::hip types: first Newmark assembly
    bit-identical; per-step iterations within 1; u, v (, a) after five steps
    rel-L2 < 1e-7 (solver tolerance 1e-10)."""
    if not os.path.exists(DYN):
        pytest.skip("dynamics binary not built (needs the reference headers at build time)")
    r = subprocess.run([DYN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["same_values"] and out["same_rhs"]
    for k in ("newmark_u_rel", "newmark_v_rel", "newmark_a_rel", "stiffwarp_u_rel",
              "stiffwarp_v_rel"):
        assert out[k] < 1e-7, k


MERGE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "dropin_merge")


def test_reference_element_loops_on_device_matrix(gpu):
    """The reference's MergeLinSys_* templates (static, Newmark-beta,
    backward Euler, stiffness-warped), unchanged, on
    delfem2::hip::CMatrixSparse (Merge found by ADL) and on the reference's
    host CMatrixSparse<double>: A and b bit-identical, and the device bulk
    assembly of each kind bit-identical to both; a host flag array passed to
    SetFixedBC is refused with DFM2H_ERR_INVALID."""
    if not os.path.exists(MERGE):
        pytest.skip("merge binary not built (needs the reference headers at build time)")
    r = subprocess.run([MERGE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for kind in ("static", "newmark", "beuler", "stiffwarp"):
        assert out[f"{kind}_template_same"] and out[f"{kind}_bulk_same"], kind
    assert out["host_pointer_status"] == 1
