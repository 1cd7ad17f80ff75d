"""The allocation-free dot (dfm2h_vec_dot_noalloc) against dfm2h_vec_dot,
through the C ABI: bit-identical results on two calls in a row, the second
with a smaller input (the scratch keeps the first call's partials; the
second must not read them); device free memory unchanged by the calls; host
pointers refused as by every device argument."""
import ctypes as C

import numpy as np
import pytest
import torch

from delfem2_amd import _native as N
from delfem2_amd import workload as W

pytestmark = pytest.mark.gpu


def _dot(name, x, y, n):
    out = C.c_double(0.0)
    N.check(getattr(N.load(), name)(C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), n,
                                    C.byref(out)))
    return out.value


def test_noalloc_dot_matches_dot(gpu):
    n1, n2 = 3_000_001, 70_001  # 2048 workgroups, then 274: fewer partials than before
    x = torch.from_numpy(W.uniform(71, n1) - 0.5).cuda()
    y = torch.from_numpy(W.uniform(72, n1) - 0.5).cuda()
    ref1 = _dot("dfm2h_vec_dot", x, y, n1)  # loads the module, allocates its own partials
    ref2 = _dot("dfm2h_vec_dot", x, y, n2)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    got1 = _dot("dfm2h_vec_dot_noalloc", x, y, n1)
    got2 = _dot("dfm2h_vec_dot_noalloc", x, y, n2)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert np.float64(got1).view(np.uint64) == np.float64(ref1).view(np.uint64)
    assert np.float64(got2).view(np.uint64) == np.float64(ref2).view(np.uint64)
    # the values themselves (sequential numpy sum, rounding tolerance)
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    assert got1 == pytest.approx(float(xh @ yh), rel=1e-12)
    assert got2 == pytest.approx(float(xh[:n2] @ yh[:n2]), rel=1e-12)
    # edge cases: n = 0 gives 0 without a launch; host memory is refused
    assert _dot("dfm2h_vec_dot_noalloc", x, y, 0) == 0.0
    xh2 = np.ones(8)
    out = C.c_double()
    st = N.load().dfm2h_vec_dot_noalloc(xh2.ctypes.data, xh2.ctypes.data, 8, C.byref(out))
    assert st == N.ERR_INVALID
