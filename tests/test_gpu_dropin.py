"""The reference's own Solve_PCG template, compiled unchanged against the
device types of include/delfem2_hip.hpp, runs on the GPU and agrees with
the fused device overload (tests/cpp/dropin_pcg.cpp)."""
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
