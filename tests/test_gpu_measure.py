"""Measurement entries of the library: the plain-copy rate the kernels'
bandwidth is compared with, and the per-kernel timing after a solve."""
import math

import pytest
import torch

from delfem2_amd import solver as S
from delfem2_amd import workload as W
from delfem2_amd.problem import DeviceSystem

pytestmark = pytest.mark.gpu


def test_copy_rate_is_a_plausible_bandwidth():
    gbs = S.measure_copy_gbs(64 << 20, 3)
    assert math.isfinite(gbs) and 100.0 < gbs < 20000.0
    with pytest.raises(S.N.DFM2HError):
        S.measure_copy_gbs(8, 3)
    with pytest.raises(S.N.DFM2HError):
        S.measure_copy_gbs(1 << 20, 0)


def test_kernel_times_after_a_solve():
    dev = DeviceSystem(W.cube(8, sigma=0.5, seed=3))
    dev.assemble()
    assert dev.factor()
    b0 = dev.b.clone()
    x, hist, _ = dev.solve(1e-8)
    assert hist[-1] < 1e-8 * hist[0]
    x0 = x.clone()
    t = S.pcg_kernel_times(dev.A, dev.ilu, fp32=False, nrep=3)
    assert set(t) >= {"spmv", "update", "fwd_sweep", "bwd_sweep", "pupdate"}
    assert all(math.isfinite(v) and v >= 0.0 for v in t.values())
    assert t["spmv"] > 0.0
    # nothing the caller owns is touched by the timing
    assert torch.equal(b0, dev.b) and torch.equal(x0, x)
