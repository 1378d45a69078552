"""The device octree bit-exact at the configuration sizes and on the
reference's own head mesh.

For each mesh: leaf depth, cube side and corner against the oracle tree;
per level the sorted Morton keys and FMMLevelSetNumSR counts
(octree.c:890-919) against the oracle; the element permutation against the
oracle; and the M2L pair list against I4 restricted to the level set
(octree.c:497-543, computed in numpy by tests/treeref.py) as a set of
(target key, source key) with the offset code of each pair, in offset-major
order (the layout the translation kernels bucket by).

Meshes: the 20,480-triangle sphere at s = 5 (the differentiation workload's
reference tree), the 60,500-triangle perturbed ellipsoid at depths 5 and 6
(the batched-frequency workload), the 81,920-triangle sphere at depth 6,
the 327,680-triangle sphere at depth 7, and Head_20kHz.obj at s = 5 as the
reference's octree driver builds it (testOctree:20-39: srcBoxes with s = 5
and the level sets)."""
import os

import numpy as np
import pytest

import fmbem_gpu_c_amd as F
import oracle as O
import treeref as R
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _check_tree(name, mesh, s=5, depth=0):
    V, Fc = mesh.arrays()
    t = F.Tree(mesh, s=s, depth=depth)
    ot = O.Tree(V, Fc, s=s, depth=depth)
    lmin, lmax, D, pmin = t.info()
    assert lmax == ot.lmax, (name, lmax, ot.lmax)
    oD, opmin = ot.geom()
    assert D == oD and np.array_equal(pmin, opmin), (name, D, oD, pmin, opmin)
    np.testing.assert_array_equal(t.elem_box()[1], ot.perm())
    npairs = []
    for l in range(2, lmax + 1):
        keys = t.level(l)
        np.testing.assert_array_equal(keys, ot.keys(l))
        num = t.num_sr(l)
        np.testing.assert_array_equal(num, ot.num_sr(l))
        tg, sr, off = t.pairs(l)
        assert len(tg) == int(num.sum())
        assert np.all(np.diff(off) >= 0)  # offset-major
        got = R.pair_codes(keys[tg], keys[sr])
        wt, ws, wo = R.m2l_pairs(keys, l)
        want = R.pair_codes(wt, ws)
        og, ow = np.argsort(got), np.argsort(want)
        np.testing.assert_array_equal(got[og], want[ow])
        np.testing.assert_array_equal(off[og], wo[ow])
        npairs.append(len(tg))
    print("%s: lmax %d, D %.6f, boxes %s, pairs %s" % (
        name, lmax, D, [len(t.level(l)) for l in range(2, lmax + 1)], npairs), flush=True)
    return t, lmax


@pytest.mark.parametrize("name,make,s,depth", [
    ("sphere-l5 s=5", lambda: F.Mesh.icosphere(5, 0.1), 5, 0),
    ("ellipsoid-60k depth 5", lambda: F.Mesh.ellipsoid(55, 0.08, 0.07, 0.10, 2, 8, 0.05, 7), 5, 5),
    ("ellipsoid-60k depth 6", lambda: F.Mesh.ellipsoid(55, 0.08, 0.07, 0.10, 2, 8, 0.05, 7), 5, 6),
    ("sphere-l6 depth 6", lambda: F.Mesh.icosphere(6, 0.3), 5, 6),
    ("sphere-l7 depth 7", lambda: F.Mesh.icosphere(7, 0.5), 5, 7),
])
def test_tree_bit_exact_at_config_sizes(name, make, s, depth):
    _check_tree(name, make(), s, depth)


def test_head_mesh_tree_as_the_reference_driver_builds_it():
    """Head_20kHz.obj (21,102 vertices, 42,200 triangles) at s = 5, as
    testOctree does; also s = 3 and s = 12 for other leaf depths."""
    m = F.Mesh.from_obj(os.path.join(GOLDEN, "Head_20kHz.obj"))
    V, Fc = m.arrays()
    assert V.shape == (21102, 3) and Fc.shape == (42200, 3)
    depths = {}
    for s in (5, 3, 12):
        t, lmax = _check_tree("head s=%d" % s, m, s=s)
        depths[s] = lmax
    assert depths[3] >= depths[5] >= depths[12]
