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
