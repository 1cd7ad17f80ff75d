"""Multicolour node ordering (dfm2h_multicolor_order).  It has no reference
counterpart -- the reference factorises in the caller's node order -- so it
is pinned by what defines it: a valid colouring (no two neighbours share a
colour), new2old a bijection that lists the colours in order with ascending
old indices inside each, an ILU(0) level count of the relabelled pattern
equal to the colour count (brute force over the rows), the least possible
count (4, a tet being a 4-clique) on Kuhn-split voxel grids, determinism,
and refusal of patterns that are not symmetric."""
import numpy as np
import pytest

from delfem2_amd import solver as S
from delfem2_amd import workload as W


def delaunay_mesh(n=3000, seed=3):
    from scipy.spatial import Delaunay
    pts = W.uniform(seed, 3 * n).reshape(-1, 3)
    return Delaunay(pts).simplices.astype(np.uint32).ravel(), n


def colours_of(new2old, ncolor, ind, ps):
    """Splits the order into the fewest runs of ascending old indices with no
    two neighbours in one run (a new run starts only where it must) and
    checks that ncolor runs suffice: a colouring with a neighbour pair inside
    one class, or classes out of order, needs more.  Returns the run of every
    old node."""
    n = new2old.size
    col = np.full(n, -1, np.int64)
    c = 0
    for k, p in enumerate(new2old):
        if k > 0:
            nb = ps[ind[p]:ind[p + 1]]
            if p < new2old[k - 1] or (col[nb] == c).any():
                c += 1
        col[p] = c
    assert c + 1 <= ncolor
    return col


def levels(ind, ps, new2old):
    """Forward dependency levels of ILU(0) on the relabelled pattern."""
    n = new2old.size
    old2new = np.empty(n, np.int64)
    old2new[new2old] = np.arange(n)
    lev = np.zeros(n, np.int64)
    for k in range(n):
        p = new2old[k]
        lower = old2new[ps[ind[p]:ind[p + 1]]]
        lower = lower[lower < k]
        lev[k] = 1 + lev[lower].max() if lower.size else 0
    return int(lev.max()) + 1 if n else 0


def check(tet, np_):
    ind, ps = S.JArray_PSuP_MeshElem(tet, np_)
    ind, ps = ind.astype(np.int64), ps.astype(np.int64)
    n2o, nc = S.MultiColorOrder(ind, ps)
    assert np.array_equal(np.sort(n2o), np.arange(np_))
    colours_of(n2o, nc, ind, ps)  # at most nc independent ascending runs
    assert levels(ind, ps, n2o) == nc
    n2o2, nc2 = S.MultiColorOrder(ind, ps)
    assert nc2 == nc and np.array_equal(n2o2, n2o)
    return nc, int(np.diff(ind).max())


@pytest.mark.parametrize("n", [1, 2, 5, 12])
def test_voxel_grids_take_four_colours(n):
    xyz, tet = W.mesh_tet_voxel(n, n, n)
    nc, _ = check(tet, xyz.size // 3)
    assert nc == 4


def test_voxel_64_cube_four_colours():
    """At 64^3 (274,625 nodes) as well: the C2 cube in the caller's natural
    order relabelled to 4 sweep levels instead of 193."""
    xyz, tet = W.mesh_tet_voxel(64, 64, 64)
    ind, ps = S.JArray_PSuP_MeshElem(tet, xyz.size // 3)
    n2o, nc = S.MultiColorOrder(ind, ps)
    assert nc == 4 and np.array_equal(np.sort(n2o), np.arange(xyz.size // 3))


def test_unstructured_mesh():
    tet, np_ = delaunay_mesh()
    nc, dmax = check(tet, np_)
    assert 4 <= nc <= dmax + 1


def test_isolated_nodes_and_empty_pattern():
    # node 5 belongs to no element
    tet = np.array([0, 1, 2, 3, 1, 2, 3, 4], np.uint32)
    nc, _ = check(tet, 6)
    assert nc == 4
    n2o, nc = S.MultiColorOrder(np.zeros(4, np.uint32), np.zeros(0, np.uint32))
    assert nc == 1 and np.array_equal(n2o, np.arange(3))


def test_bad_patterns_refused():
    ind = np.array([0, 1, 2], np.uint32)
    with pytest.raises(Exception):  # 0 lists 1, 1 lists itself
        S.MultiColorOrder(ind, np.array([1, 1], np.uint32))
    with pytest.raises(Exception):  # not symmetric
        S.MultiColorOrder(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32))
    with pytest.raises(Exception):  # out of range
        S.MultiColorOrder(ind, np.array([1, 7], np.uint32))
    with pytest.raises(Exception):  # repeated entry
        S.MultiColorOrder(np.array([0, 2, 4], np.uint32), np.array([1, 1, 0, 0], np.uint32))
