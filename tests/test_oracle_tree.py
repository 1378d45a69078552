"""Oracle octree (fo_tree.c, restating octree.c) against brute-force
geometric definitions: Morton digits, neighbours, the interaction list I4
(children of the parent's neighbourhood minus the own neighbourhood,
octree.c:517-543), the leaf depth of deterLmax (octree.c:635-679) as the
smallest level whose fullest box holds <= s centroids, and truncNum."""
import itertools

import numpy as np
import pytest

import oracle as O


def _grid(key, l):
    c = O.box_center(key, l)
    return tuple(int(v) for v in np.floor(c * (1 << l)))


def test_morton_digit_order():
    # x is the most significant bit of each octal digit (octree.c:132)
    l = 3
    s = np.array([0.7, 0.2, 0.9])  # grid (5, 1, 7) at level 3
    key = O.morton_key(s, l)
    t = (5, 1, 7)
    want = 0
    for i in range(l - 1, -1, -1):
        want = want * 8 + 4 * ((t[0] >> i) & 1) + 2 * ((t[1] >> i) & 1) + ((t[2] >> i) & 1)
    assert key == want
    assert _grid(key, l) == t


@pytest.mark.parametrize("l", [2, 3, 4])
def test_neighbors_brute_force(l):
    rng = np.random.default_rng(l)
    N = 1 << l
    for key in rng.choice(N ** 3, 20, replace=False):
        g = _grid(int(key), l)
        want = set()
        for d in itertools.product((-1, 0, 1), repeat=3):
            q = tuple(a + b for a, b in zip(g, d))
            if d != (0, 0, 0) and all(0 <= v < N for v in q):
                want.add(O.morton_key((np.array(q) + 0.5) / N, l))
        assert set(O.neighbors(int(key), l).tolist()) == want


@pytest.mark.parametrize("l", [2, 3, 4])
def test_I4_brute_force(l):
    N = 1 << l
    rng = np.random.default_rng(10 + l)
    for key in rng.choice(N ** 3, 12, replace=False):
        g = np.array(_grid(int(key), l))
        want = set()
        for q in itertools.product(range(N), repeat=3):
            q = np.array(q)
            if np.max(np.abs(q - g)) <= 1:
                continue
            if np.max(np.abs(q // 2 - g // 2)) <= 1:
                want.add(O.morton_key((q + 0.5) / N, l))
        assert set(O.I4(int(key), l).tolist()) == want


def test_I4_counts_and_offsets():
    l = 4
    N = 1 << l
    interior = O.morton_key(np.array([6.5, 6.5, 9.5]) / N, l)
    assert len(O.I4(interior, l)) == 189
    offs, d2 = set(), set()
    for key in range(N ** 3):
        g = np.array(_grid(key, l))
        for s in O.I4(key, l):
            o = tuple(g - np.array(_grid(int(s), l)))
            offs.add(o)
            d2.add(int(np.dot(o, o)))
    assert len(offs) == 316
    assert len(d2) == 15
    dirs = {tuple(np.round(np.array(o) / np.linalg.norm(o), 12)) for o in offs}
    assert len(dirs) == 290


def _brute_lmax(pts, s):
    for l in range(1, 11):
        keys = [O.morton_key(p, l) for p in pts]
        if np.max(np.bincount(keys)) <= s:
            return l
    return 11


@pytest.mark.parametrize("seed,s", [(0, 1), (1, 3), (2, 5), (3, 10)])
def test_deter_lmax_brute_force(seed, s):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.01, 0.99, (300, 3))
    assert O.deter_lmax(pts, s) == max(_brute_lmax(pts, s), 0)


def test_tree_levels_of_sphere_mesh():
    import fmbem_gpu_c_amd as F
    m = F.Mesh.icosphere(3, 0.1)
    V, Fc = m.arrays()
    t = O.Tree(V, Fc, s=5)
    D, pmin = t.geom()
    c = V[Fc].mean(1)
    lo, hi = c.min(0), c.max(0)
    assert D == pytest.approx((hi - lo).max() + 2 * min(0.01, 1e-6 * (hi - lo).min()), rel=1e-12)
    for l in range(2, t.lmax + 1):
        keys = t.keys(l)
        want = sorted({O.morton_key((p - pmin) / D, l) for p in c})
        assert keys.tolist() == want
    # every leaf box holds at most s elements, the level above does not
    L = t.lmax
    cnt = np.bincount([O.morton_key((p - pmin) / D, L) for p in c])
    assert cnt.max() <= 5
    cnt = np.bincount([O.morton_key((p - pmin) / D, L - 1) for p in c])
    assert cnt.max() > 5


def test_trunc_num_closed_form():
    # octree.c:777-796
    def ref(k, eps, sigma, a):
        p_lo = 1 - np.log(eps * (1 - 1 / sigma) ** 1.5) / np.log(sigma)
        p_hi = k * a + (3 * np.log(1 / eps)) ** 1.5 / 2 * (k * a) ** (1 / 3)
        return int(np.ceil((p_lo ** 4 + p_hi ** 4) ** 0.25))
    for k, a in [(18.32, 0.05), (18.32, 0.025), (183.2, 0.25), (183.2, 0.0078), (91.59, 0.15)]:
        assert O.trunc_num(k, 0.05, 1.5, a) == ref(k, 0.05, 1.5, a)
    # the per-level orders quoted for the 327,680-triangle, 10 kHz sphere
    D = 1.0
    assert [O.trunc_num(2 * np.pi * 1e4 / 343.0, 0.05, 1.5, D / 2 ** l) for l in range(2, 8)] == \
        [94, 62, 42, 31, 23, 18]


def test_deter_lmax_sentinel_for_coincident_points():
    """More than s points in one level-10 box: the reference's deterLmax
    returns 11 (octree.c:662-665), which the device build maps to
    FMB_ERR_DEPTH; s points or fewer still give a depth."""
    rng = np.random.default_rng(2)
    pts = rng.uniform(0, 1, (50, 3))
    dup = np.vstack([pts, np.repeat(pts[:1], 5, axis=0)])  # 6 copies of one point
    assert O.deter_lmax(dup, 5) == 11
    assert O.deter_lmax(dup, 6) <= 10
    ok = np.vstack([pts, np.repeat(pts[:1], 4, axis=0)])  # 5 copies
    assert O.deter_lmax(ok, 5) <= 10


def test_numpy_index_sets_match_oracle():
    """tests/treeref.py (used for the full-size device tree checks) against
    the oracle: Morton round trip and box centres, I4 of sampled boxes at
    levels 2-6, and the pair set of a random level set = I4 restricted to
    it, with offset codes."""
    import treeref as R
    rng = np.random.default_rng(9)
    for l in range(2, 7):
        keys = rng.choice(8 ** l, size=min(40, 8 ** l), replace=False)
        ijk = R.morton_decode(keys, l)
        np.testing.assert_array_equal(R.morton_encode(ijk, l), keys)
        for k, c in zip(keys[:10], ijk[:10]):
            np.testing.assert_allclose(O.box_center(int(k), l), (c + 0.5) / 2 ** l, rtol=0,
                                       atol=1e-15)
            np.testing.assert_array_equal(R.i4_keys(int(k), l), np.sort(O.I4(int(k), l)))
    l = 4
    keys = np.sort(rng.choice(8 ** l, size=900, replace=False))
    tg, sr, oc = R.m2l_pairs(keys, l)
    kset = set(keys.tolist())
    want = {(int(b), int(s)) for b in keys for s in O.I4(int(b), l) if int(s) in kset}
    assert set(zip(tg.tolist(), sr.tolist())) == want
    d = R.morton_decode(tg, l) - R.morton_decode(sr, l)
    np.testing.assert_array_equal((d[:, 0] + 3) * 49 + (d[:, 1] + 3) * 7 + d[:, 2] + 3, oc)


def test_reduce_matches_generation_at_lower_order():
    """reduceSparseCoaxTransMat / reduceSparseRotMat (translation.cu:22-56)
    as exported by the library, on the oracle's matrices: main.c's case
    ((S|R), k = 13.1, t = 2.3, p = 5 reduced to 2, main.c:27-38) and wider
    ones equal the matrices generated at the lower order; the layout is
    checked entry by entry against the oracle's block offsets."""
    import fmbem_gpu_c_amd as F
    for kind, k, t, p1, p2 in [(O.COAX_SR, 13.1, 2.3, 5, 2), (O.COAX_SR, 18.3, 0.1, 16, 9),
                               (O.COAX_RR, 91.6, 0.05, 30, 17)]:
        big = O.coax_sparse(kind, p1, k, t)
        got = F.reduce_coax_mat(big, p1, p2)
        want = O.coax_sparse(kind, p2, k, t)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        for m in range(p2):
            a, b = O.lib().fo_coax_block(p1, m), O.lib().fo_coax_block(p2, m)
            A = big[a:a + (p1 - m) ** 2].reshape(p1 - m, p1 - m)
            B = got[b:b + (p2 - m) ** 2].reshape(p2 - m, p2 - m)
            np.testing.assert_array_equal(A[:p2 - m, :p2 - m], B)
    r = O.rot_sparse(12, 0.3, 1.1, -0.7)
    np.testing.assert_array_equal(F.reduce_rot_mat(r, 12, 7), O.rot_sparse(7, 0.3, 1.1, -0.7))
    with pytest.raises(F.FmbError):
        F.reduce_coax_mat(big, p1, p1 + 1)
