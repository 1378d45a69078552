"""Transposed FMM pass and adjoint gradients on the GPU.

* fmb_matvec_transpose is the exact transpose of the FMM operator fmb_matvec
  applies: y^T (A x) = (A^T y)^T x to rounding (FP64 1e-12, FP32 pass 1e-5),
  for the near and far parts separately, on trees that exercise every
  translation mode; and it approaches the oracle's dense A^T like the
  forward pass approaches the dense A.
* fmb_adjoint_grad against the oracle's dense FP64 gradient (the formula
  pinned by central differences in test_oracle_adjoint.py; FMM-level
  agreement), against central finite differences of J computed with the
  same FMM operator along random directions (1e-5 relative, step 1e-6 in
  the admittance), and
  through the torch.autograd wrapper.
"""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber
from test_oracle_adjoint import admittance, circle_points, dense_adjoint_gradient

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _bilinear(a, b):
    return np.sum(a * b)


@pytest.mark.parametrize("freq,orders,robin", [
    (1000.0, None, False), (1000.0, None, True),
    (4000.0, [0, 0, 33, 28, 24, 8], False), (4000.0, [0, 0, 6, 12, 20, 17], True),
    (16000.0, [0, 0, 62, 44, 30, 20], True)])
def test_transpose_is_exact_transpose(sphere_l3, freq, orders, robin):
    """Explicit orders on the 5,120-triangle sphere at depth 5 run every LDS
    tile shape and the global-workspace mode, with pin != pout in M2M/L2L."""
    k = wavenumber(freq)
    kw = dict(nrhs_max=3, precision=F.FMB_MIXED)
    if orders is not None:
        m = F.Mesh.icosphere(4, 0.1)
        kw.update(p_rule=F.PRULE_EXPLICIT, p_levels=orders)
        t = F.Tree(m, depth=5)
    else:
        m = sphere_l3[0]
        t = F.Tree(m, s=5)
    n = m.arrays()[1].shape[0]
    sig = torch.tensor(1j * k * admittance(n, 2), device="cuda") if robin else None
    pl = F.Plan(t, k, robin=sig, **kw)
    x, y = _rand((3, n), 1), _rand((3, n), 2)
    xt, yt = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-12), (F.FMB_MIXED, 1e-5)):
        for parts in (1, 2, 3):
            Ax = pl.matvec_parts(xt, parts, prec=prec).cpu().numpy()
            Aty = pl.matvec_t(yt, parts=parts, prec=prec).cpu().numpy()
            for r in range(3):
                lhs, rhs = _bilinear(y[r], Ax[r]), _bilinear(Aty[r], x[r])
                scale = np.linalg.norm(y[r]) * np.linalg.norm(Ax[r])
                assert abs(lhs - rhs) < tol * scale, (prec, parts, r, lhs, rhs)


def test_transpose_vs_dense_oracle(sphere_l3):
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    k = wavenumber(1000.0)
    sig = 1j * k * admittance(n, 3)
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, robin=torch.tensor(sig, device="cuda"))
    A = O.Bem(V, Fc, k, sigma=sig).dense_matrix()
    y = _rand(n, 5)
    got = pl.matvec_t(torch.tensor(y, device="cuda")).cpu().numpy()
    want = A.T @ y
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-4
    fwd = pl.matvec(torch.tensor(y, device="cuda")).cpu().numpy()
    assert np.linalg.norm(fwd - A @ y) / np.linalg.norm(A @ y) < 1e-4


def test_gmres_transpose(sphere_l3):
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    k = wavenumber(1000.0)
    pl = F.Plan(F.Tree(m, s=5), k, nrhs_max=2, precision=F.FMB_MIXED)
    b = torch.tensor(_rand((2, n), 8), device="cuda")
    lam, info = pl.gmres_t(b, tol=1e-10)
    assert max(info.relres[0], info.relres[1]) <= 1e-10
    r = b - pl.matvec_t(lam)
    for i in range(2):
        assert (torch.linalg.norm(r[i]) / torch.linalg.norm(b[i])).item() < 2e-10


def _fmm_objective(pl, k, Y, pts, d, tol=1e-13):
    pl.retune(k, robin=torch.tensor(1j * k * Y, device="cuda"))
    x, _ = pl.gmres(pl.rhs_plane_wave(d), tol=tol)
    p = pl.field(x, pts, d).cpu().numpy()[0]
    return np.sum(np.abs(p) ** 2)


@pytest.mark.parametrize("precision", [F.FMB_FP64, F.FMB_MIXED])
def test_adjoint_gradient(sphere_l3, precision):
    """Sphere of 1,280 triangles at 4 kHz with a random admittance."""
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    k = wavenumber(4000.0)
    Y = admittance(n, 4)
    pts = circle_points()
    d = np.array([[0.0, 0.0, 1.0]])
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, robin=torch.tensor(1j * k * Y, device="cuda"), precision=precision)
    x, _ = pl.gmres(pl.rhs_plane_wave(d), tol=1e-12)
    J, gs, info = pl.adjoint_grad(x, pts, d, tol=1e-12)
    assert info.relres[0] <= 1e-12
    gs = gs.cpu().numpy()
    # the oracle's dense FP64 gradient
    Jd, gd, _ = dense_adjoint_gradient(V, Fc, k, 1j * k * Y, pts, d[0])
    assert abs(J - Jd) / Jd < 1e-4
    assert np.linalg.norm(gs - gd) / np.linalg.norm(gd) < 1e-3
    # central differences of the FMM objective along random complex
    # directions (entries of unit size), step 1e-6 in Y: dJ = Re(conj(g) dY)
    gY = -1j * k * gs
    h = 1e-6
    for seed in (6, 7):
        dY = _rand(n, seed)
        fd = (_fmm_objective(pl, k, Y + h * dY, pts, d)
              - _fmm_objective(pl, k, Y - h * dY, pts, d)) / (2 * h)
        lin = np.real(np.sum(np.conj(gY) * dY))
        assert abs(lin - fd) <= 1e-5 * abs(fd), (seed, lin, fd)


def test_autograd_admittance_objective(sphere_l3):
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    k = wavenumber(4000.0)
    Yn = admittance(n, 9)
    pts = circle_points()
    d = np.array([[0.6, 0.0, 0.8]])
    pl = F.Plan(F.Tree(m, s=5), k, robin=torch.tensor(1j * k * Yn, device="cuda"))
    Y = torch.tensor(Yn, device="cuda", requires_grad=True)
    J = F.admittance_objective(Y, pl, pts, d, tol=1e-13)
    J.backward()
    g = Y.grad.cpu().numpy()
    # directional derivative along a random complex direction, step 1e-6
    dY = _rand(n, 10)
    h = 1e-6
    fd = (_fmm_objective(pl, k, Yn + h * dY, pts, d)
          - _fmm_objective(pl, k, Yn - h * dY, pts, d)) / (2 * h)
    lin = np.real(np.sum(np.conj(g) * dY))
    assert abs(lin - fd) <= 1e-5 * abs(fd), (lin, fd)
    assert abs(J.item() - _fmm_objective(pl, k, Yn, pts, d)) < 1e-10 * J.item()


def _objective_on_mesh(V, Fc, k, Y, pts, d, s=5, tol=1e-13):
    """J with the FMM operator of a tree and plan built on the mesh (V, Fc)."""
    m = F.Mesh.from_arrays(V, Fc)
    pl = F.Plan(F.Tree(m, s=s), k, robin=torch.tensor(1j * k * Y, device="cuda"))
    x, _ = pl.gmres(pl.rhs_plane_wave(d), tol=tol)
    p = pl.field(x, pts, d).cpu().numpy()[0]
    return np.sum(np.abs(p) ** 2)


def _generic_rotation(V, Fc, gap=2e-6):
    """A seeded rotation of the mesh after which no centroid coordinate lies
    within `gap` of a box plane of its octree at any level: the icosphere's
    symmetric centroids sit exactly on the midplanes, and a displacement
    would move them between boxes (a discontinuity of the FMM operator, not
    a derivative)."""
    for seed in range(50):
        Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
        W = V @ Q.T
        c = W[Fc].mean(axis=1)
        ot = O.Tree(W, Fc, s=5)
        D, pmin = ot.geom()
        ok = True
        for l in range(2, ot.lmax + 1):
            hb = D / 2 ** l
            rel = (c - np.asarray(pmin)) / hb
            near = np.round(rel)
            inner = (near > 0) & (near < 2 ** l)  # the cube's faces hold the fixed extremes
            if (np.abs(rel - near) * hb)[inner].min() < gap:
                ok = False
                break
        if ok:
            return W
    raise AssertionError("no generic rotation found")


def test_vertex_gradients(sphere_l3):
    """dJ/dV and dJ/d(vertex normal displacement) on the 1,280-triangle
    sphere (rotated generically) at 4 kHz with a random admittance: the
    directional derivative along a random normal displacement field
    phi_v N_v against central differences (step 1e-6 m) of J with the FMM
    operator rebuilt on the displaced mesh (1e-5 relative), and against those
    of the oracle's dense objective (FMM-level agreement); dJ_dvn =
    dJ_dV . N_v; FP32 Arnoldi (mixed plan) gives the same gradient."""
    _, V0, Fc = sphere_l3
    V = _generic_rotation(V0, Fc)
    m = F.Mesh.from_arrays(V, Fc)
    n, nv = Fc.shape[0], V.shape[0]
    k = wavenumber(4000.0)
    Y = admittance(n, 12)
    pts = circle_points()
    d = np.array([[0.0, 0.0, 1.0]])
    t = F.Tree(m, s=5)
    Nv = t.vertex_normals()
    np.testing.assert_allclose(np.linalg.norm(Nv, axis=1), 1.0, atol=1e-12)
    assert np.all(np.sum(Nv * V, axis=1) > 0)  # outward on the sphere
    grads = {}
    for prec in (F.FMB_FP64, F.FMB_MIXED):
        pl = F.Plan(t, k, robin=torch.tensor(1j * k * Y, device="cuda"), precision=prec)
        x, _ = pl.gmres(pl.rhs_plane_wave(d), tol=1e-12)
        J, gs, info, dV, dvn = pl.adjoint_grad(x, pts, d, tol=1e-12, vertices=True)
        grads[prec] = (J, dV.cpu().numpy(), dvn.cpu().numpy())
    J, dV, dvn = grads[F.FMB_FP64]
    np.testing.assert_allclose(dvn, np.sum(dV * Nv, axis=1), rtol=1e-12, atol=1e-14 * np.abs(dvn).max())
    _, dV32, dvn32 = grads[F.FMB_MIXED]
    assert np.linalg.norm(dvn32 - dvn) / np.linalg.norm(dvn) < 1e-6
    # displacements vanish on the elements whose centroids bound the tree's
    # cube, so the rebuilt tree keeps its boxes (the adjoint differentiates
    # the FMM operator of a fixed tree)
    cen = V[Fc].mean(axis=1)
    fixed = np.zeros(nv, bool)
    for a in range(3):
        for ext in (cen[:, a].min(), cen[:, a].max()):
            fixed[Fc[np.abs(cen[:, a] - ext) < 1e-4].ravel()] = True
    phi = np.random.default_rng(13).uniform(-1, 1, nv) * ~fixed
    h = 1e-6
    Jp = _objective_on_mesh(V + h * phi[:, None] * Nv, Fc, k, Y, pts, d)
    Jm = _objective_on_mesh(V - h * phi[:, None] * Nv, Fc, k, Y, pts, d)
    fd = (Jp - Jm) / (2 * h)
    lin = np.sum(dvn * phi)
    print("vertex-normal directional derivative: adjoint %.10e  fd %.10e" % (lin, fd))
    assert abs(lin - fd) <= 1e-5 * abs(fd), (lin, fd)
    # a tangential-plus-normal direction through dJ_dV
    dir3 = np.random.default_rng(14).uniform(-1, 1, (nv, 3)) * ~fixed[:, None]
    Jp = _objective_on_mesh(V + h * dir3, Fc, k, Y, pts, d)
    Jm = _objective_on_mesh(V - h * dir3, Fc, k, Y, pts, d)
    fd3 = (Jp - Jm) / (2 * h)
    lin3 = np.sum(dV * dir3)
    assert abs(lin3 - fd3) <= 1e-5 * abs(fd3), (lin3, fd3)
    # the oracle's dense objective, same displacement (step 1e-6 m)
    sig = 1j * k * Y

    def dense_J(Vd):
        b = O.Bem(Vd, Fc, k, sigma=sig)
        xd = np.linalg.solve(b.dense_matrix(), b.rhs_plane(d[0]))
        return np.sum(np.abs(b.field(xd, pts, dirs=d)[0]) ** 2)

    fdd = (dense_J(V + h * phi[:, None] * Nv) - dense_J(V - h * phi[:, None] * Nv)) / (2 * h)
    print("dense-oracle directional derivative %.10e" % fdd)
    assert abs(lin - fdd) <= 2e-3 * abs(fdd), (lin, fdd)


def _movable(t, V, Fc, gap=2e-6):
    """Vertices that may move without any element changing its octree box
    or the tree its cube: those of elements whose centroids are at least
    `gap` from every box plane and not among the extreme centroids."""
    _, L, D, pmin = t.info()
    c = V[Fc].mean(axis=1)
    bad = np.zeros(len(Fc), bool)
    for l in range(2, L + 1):
        hb = D / 2 ** l
        rel = (c - np.asarray(pmin)) / hb
        bad |= (np.abs(rel - np.round(rel)) * hb < gap).any(axis=1)
    for a in range(3):
        for ext in (c[:, a].min(), c[:, a].max()):
            bad |= np.abs(c[:, a] - ext) < 1e-4
    fixed = np.zeros(len(V), bool)
    fixed[Fc[bad].ravel()] = True
    return ~fixed


def test_adjoint_config4_sphere_20480():
    """The differentiation configuration: rigid-sphere mesh r = 0.1 m,
    icosphere level 5 (20,480 triangles), 4 kHz, plane wave along +z,
    per-triangle admittance Y = 0.01 (1 + 0.5 U(-1,1)) + 0.005i U(-1,1)
    (seeded), J = sum |p|^2 at 32 points on a 0.5 m circle.  dJ/dY and
    dJ/d(vertex normal displacement) in FP64 and mixed form against central
    differences (step 1e-6) of the FMM objective: 1e-5 relative."""
    import time
    m = F.Mesh.icosphere(5, 0.1)
    V, Fc = m.arrays()
    n = len(Fc)
    k = wavenumber(4000.0)
    Y = admittance(n, 2024)
    pts = circle_points()
    d = np.array([[0.0, 0.0, 1.0]])
    t = F.Tree(m, s=5)
    Nv = t.vertex_normals()
    out = {}
    for prec in (F.FMB_FP64, F.FMB_MIXED):
        pl = F.Plan(t, k, robin=torch.tensor(1j * k * Y, device="cuda"), precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, info = pl.gmres(pl.rhs_plane_wave(d), tol=1e-12)
        t1 = time.perf_counter()
        J, gs, ainfo, dV, dvn = pl.adjoint_grad(x, pts, d, tol=1e-12, vertices=True)
        t2 = time.perf_counter()
        print("config 4 %s: forward solve %.3fs (%d its), adjoint + gradients %.3fs (%d its), J %.12e"
              % ("fp64" if prec == F.FMB_FP64 else "mixed", t1 - t0, info.iters, t2 - t1,
                 ainfo.iters, J))
        out[prec] = (J, (-1j * k * gs).cpu().numpy(), dvn.cpu().numpy(), pl)
    J, gY, dvn, pl = out[F.FMB_FP64]
    _, gY32, dvn32, _ = out[F.FMB_MIXED]
    assert np.linalg.norm(gY32 - gY) / np.linalg.norm(gY) < 1e-6
    assert np.linalg.norm(dvn32 - dvn) / np.linalg.norm(dvn) < 1e-6
    h = 1e-6
    # admittance: random complex direction
    dY = _rand(n, 77)
    fd = (_fmm_objective(pl, k, Y + h * dY, pts, d) - _fmm_objective(pl, k, Y - h * dY, pts, d)) / (2 * h)
    lin = np.real(np.sum(np.conj(gY) * dY))
    print("config 4 dJ/dY directional: adjoint %.10e fd %.10e" % (lin, fd))
    assert abs(lin - fd) <= 1e-5 * abs(fd)
    # vertex normal displacement: random field on the movable vertices
    phi = np.random.default_rng(78).uniform(-1, 1, len(V)) * _movable(t, V, Fc)
    Jp = _objective_on_mesh(V + h * phi[:, None] * Nv, Fc, k, Y, pts, d)
    Jm = _objective_on_mesh(V - h * phi[:, None] * Nv, Fc, k, Y, pts, d)
    fdv = (Jp - Jm) / (2 * h)
    linv = np.sum(dvn * phi)
    print("config 4 dJ/dvn directional: adjoint %.10e fd %.10e" % (linv, fdv))
    assert abs(linv - fdv) <= 1e-5 * abs(fdv)
