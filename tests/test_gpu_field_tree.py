"""Field points through the tree (fmb_field_eval_fmm; field.hip,
k_field_tree) against the oracle's treecode (fo_bem_field_fmm: the same
tree, orders and separation rule, FP64 expansions, so 1e-10 relative), and
against the direct representation formula (fmb_field_eval) within the
truncation error of the plan's orders."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber

pytestmark = pytest.mark.gpu


def _points(rng, a):
    """32 points on a circle of 5 radii, 48 between 1.01 and 2 radii, 16
    inside the sphere of radius a."""
    th = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    far = 5 * a * np.stack([np.cos(th), np.sin(th), 0.2 * np.sin(3 * th)], 1)
    d = rng.standard_normal((64, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    near = d[:48] * a * rng.uniform(1.01, 2.0, 48)[:, None]
    inside = d[48:] * a * rng.uniform(0.2, 0.9, 16)[:, None]
    return np.concatenate([far, near, inside])


def _admittance(rng, n):
    return 0.01 * (1 + 0.5 * rng.uniform(-1, 1, n)) + 0.005j * rng.uniform(-1, 1, n)


def _compare(pl, ob, x, pts, dirs, tol_direct, tol_oracle=1e-10):
    xt = torch.tensor(x, device="cuda")
    got = pl.field(xt, pts, dirs, method="fmm").cpu().numpy()
    want = ob.field(x, pts, dirs, fmm=True)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    direct = pl.field(xt, pts, dirs).cpu().numpy()
    etr = np.abs(got - direct).max() / np.abs(direct).max()
    print("field through the tree: vs oracle treecode %.2e, vs direct %.2e" % (err, etr),
          flush=True)
    assert err < tol_oracle
    assert etr < tol_direct


@pytest.mark.parametrize("robin", [False, True])
def test_field_tree_matches_oracle(sphere_l3, robin):
    """Config-1 sphere at 1 kHz with the default orders; three right-hand
    sides in passes of two (nrhs_max = 2), incident plane waves added."""
    m, V, Fc = sphere_l3
    k = wavenumber(1000.0)
    n = Fc.shape[0]
    rng = np.random.default_rng(31)
    sig = 1j * k * _admittance(rng, n) if robin else None
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, robin=None if sig is None else torch.tensor(sig, device="cuda"),
                nrhs_max=2)
    ob = O.Bem(V, Fc, k, sigma=sig)
    ob.fmm_setup(s=5)
    L, op = ob.orders()
    np.testing.assert_array_equal(pl.orders()[2:L + 1], op[2:L + 1])
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    dirs = [[0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.0, 1.0, 0.0]]
    _compare(pl, ob, x, _points(rng, 0.1), dirs, 1e-4)


@pytest.mark.parametrize("orders", [[62, 44, 30, 20], [33, 28, 24, 8]])
def test_field_tree_high_orders_match_oracle(orders):
    """Level-2 orders above 64 lanes' worth of columns are not reached here;
    orders 62 and 33 run the 64-lane groups, 20, 24, 28, 30 the 32-lane
    groups and 8 the 16-lane groups (5,120 triangles, 16 kHz, depth 5)."""
    m = F.Mesh.icosphere(4, 0.1)
    V, Fc = m.arrays()
    t = F.Tree(m, depth=5)
    k = wavenumber(16000.0)
    pl = F.Plan(t, k, nrhs_max=2, p_rule=F.PRULE_EXPLICIT, p_levels=[0, 0] + orders)
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=5, plev=dict(zip(range(2, 6), orders)))
    rng = np.random.default_rng(32)
    x = rng.standard_normal((2, Fc.shape[0])) + 1j * rng.standard_normal((2, Fc.shape[0]))
    _compare(pl, ob, x, _points(rng, 0.1), None, 1e-3 if orders[-1] < 10 else 1e-5)


def test_field_tree_order_above_64():
    """A level-2 order of 94 (two columns per lane) against the oracle's
    treecode at points off a 1,280-triangle sphere at 16 kHz.  Depth 4 makes
    the leaf boxes about one panel wide: panels reach out of their boxes and
    the expansions converge more slowly at the separation limit, which sets
    the distance to the direct sum, not the orders."""
    m = F.Mesh.icosphere(3, 0.1)
    V, Fc = m.arrays()
    t = F.Tree(m, depth=4)
    k = wavenumber(16000.0)
    orders = [94, 62, 44]
    pl = F.Plan(t, k, nrhs_max=1, p_rule=F.PRULE_EXPLICIT, p_levels=[0, 0] + orders)
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=4, plev=dict(zip(range(2, 5), orders)))
    rng = np.random.default_rng(33)
    x = rng.standard_normal((1, Fc.shape[0])) + 1j * rng.standard_normal((1, Fc.shape[0]))
    _compare(pl, ob, x, _points(rng, 0.1), None, 1e-4)


def test_field_tree_config4():
    """The differentiation case: 20,480 triangles, r = 0.1 m, 4 kHz, seeded
    admittance, the 32 objective points on the 0.5 m circle plus points
    near the surface, from the solved surface pressure of a plane wave."""
    m = F.Mesh.icosphere(5, 0.1)
    V, Fc = m.arrays()
    n = Fc.shape[0]
    assert n == 20480
    k = wavenumber(4000.0)
    rng = np.random.default_rng(34)
    sig = 1j * k * _admittance(rng, n)
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, robin=torch.tensor(sig, device="cuda"), nrhs_max=1)
    b = pl.rhs_plane_wave(np.array([[0.0, 0.0, 1.0]]))
    x, info = pl.gmres(b, restart=30, tol=1e-8)
    assert info.relres[0] <= 1e-8
    ob = O.Bem(V, Fc, k, sigma=sig)
    ob.fmm_setup(s=5)
    _compare(pl, ob, x.cpu().numpy(), _points(rng, 0.1), [[0.0, 0.0, 1.0]], 1e-4)


def test_field_tree_chunks(sphere_l3):
    """40,000 points (two passes of the level-2 sums, 32,768 points each at
    most), 9 right-hand sides (two launches of 8) in upward passes of 4
    (nrhs_max), against the oracle's treecode and the direct sum (which
    runs unsliced above 1,024 points)."""
    m, V, Fc = sphere_l3
    k = wavenumber(1000.0)
    n = Fc.shape[0]
    rng = np.random.default_rng(35)
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, nrhs_max=4)
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(s=5)
    d = rng.standard_normal((40000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * rng.uniform(0.05, 1.0, 40000)[:, None]
    x = rng.standard_normal((9, n)) + 1j * rng.standard_normal((9, n))
    dirs = rng.standard_normal((9, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    _compare(pl, ob, x, pts, dirs, 1e-4)


def test_field_tree_points_on_box_faces(sphere_l3):
    """Points on the faces, edges and corners of level-3..5 boxes (pmin + D
    i / 2^l), where the box index of a point depends on the last bit of
    (y - pmin) / D: the box-major levels, the per-point traversal and the
    oracle must classify every point alike, or a box is counted twice or
    missed."""
    m, V, Fc = sphere_l3
    k = wavenumber(1000.0)
    n = Fc.shape[0]
    rng = np.random.default_rng(36)
    t = F.Tree(m, s=5)
    lmin, lmax, D, pmin = t.info()
    assert lmax == 4
    # orders 30 and 26 run box-major, 12 (level 4) per point
    pl = F.Plan(t, k, nrhs_max=1, p_rule=F.PRULE_EXPLICIT, p_levels=[0, 0, 30, 26, 12])
    orders = pl.orders()
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(s=5, plev={l: int(orders[l]) for l in range(2, lmax + 1)})
    pts = []
    for l in (3, 4, 5):
        g = rng.integers(-2, 2 ** l + 3, size=(200, 3)).astype(np.float64)
        pts.append(np.asarray(pmin) + D * g / 2.0 ** l)
    pts = np.concatenate(pts)
    # keep points 3 mm or more off the 0.1 m sphere, so off its panels (the
    # representation formula's close rule covers near points, not points on
    # a panel)
    pts = pts[np.abs(np.linalg.norm(pts, axis=1) - 0.1) > 0.003]
    x = rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))
    # a point on a face belongs to the same box on the device and in the
    # oracle only if both trees have the same cube to the last bit; a box
    # counted twice or missed would show against the direct sum either way
    oD, opmin = O.Tree(V, Fc, s=5).geom()
    same = oD == D and np.array_equal(np.asarray(pmin), opmin)
    print("cube bit-identical to the oracle's:", same, flush=True)
    _compare(pl, ob, x, pts, None, 1e-3, 1e-10 if same else 1e-3)
