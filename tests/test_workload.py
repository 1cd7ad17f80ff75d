"""Synthetic inputs: mesh generator equals the restated reference generator,
orderings are valid colourings, the RNG is the documented SplitMix64."""
import numpy as np
import pytest

import oracle
from delfem2_amd import workload as W


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (7, 5, 3)])
def test_mesh_matches_oracle(dims):
    a = W.mesh_tet_voxel(*dims)
    b = oracle.mesh_tet_voxel(*dims)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n", [2, 3, 6])
def test_color8_is_a_colouring(n):
    p = W.cube(n, ordering="color8")
    ind, ps = oracle.psup(p.tet, p.np_)
    n2o = p.extra["new2old"]
    g = np.indices((n + 1,) * 3).reshape(3, -1)
    color = ((g[0] & 1) * 4 + (g[1] & 1) * 2 + (g[2] & 1))[n2o]
    assert np.all(np.diff(color) >= 0)  # colour-contiguous
    for i in range(p.np_):
        assert not np.any(color[ps[ind[i]:ind[i + 1]]] == color[i])
    # same clamped set and same geometry, relabelled
    q = W.cube(n)
    assert np.array_equal(p.xyz.reshape(-1, 3), q.xyz.reshape(-1, 3)[n2o])
    assert np.array_equal(p.flag.reshape(-1, 3), q.flag.reshape(-1, 3)[n2o])


def test_splitmix64_known_values():
    # SplitMix64 with seed 0: the published first outputs of the generator
    z = W.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = W.uniform(1, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    z = W.normal(2, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_lame_parameters():
    myu, lam = W.lame(np.array([1.0, 2.0]), np.array([0.3, 0.25]))
    assert np.allclose(myu, [1 / 2.6, 2 / 2.5])
    assert np.allclose(lam, [0.3 / (1.3 * 0.4), 0.5 / (1.25 * 0.5)])


def test_cube_config_sizes():
    p = W.cube(10)
    assert (p.np_, p.ne, p.ndof) == (1331, 6000, 3993)
    assert p.flag.sum() == 121 * 3
    c = W.cantilever(8, 4, 4)
    assert np.isclose(c.f_ext.sum(), -1.0) and (c.f_ext != 0).sum() == 9
    g = W.batch_grids(512)
    assert g.min() >= 4 and g.max() <= 24 and g.shape == (512, 3)
