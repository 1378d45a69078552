"""Refinement of the FP64 operator around GMRES on a lower-order inner plan
(fmb_gmres_refined) on the GPU.

The 5,120-triangle sphere (radius 0.1 m, depth 5) at 8 kHz, three plane
waves, rigid and with a Robin admittance.  The outer plan has the default
truncation (eps 0.05); the inner plan is built on the same tree with eps 0.2
or 0.5, so its leaf order is lower.  Checks:
* the residual b - A x recomputed here with the outer plan's FP64 pass meets
  the tolerance and equals the reported one;
* the solution matches an FP64 GMRES solve of the outer plan (1e-7);
* one outer FP64 pass per cycle (a zero first guess needs none), and a
  converged first guess returns after a single check;
* plans that differ in tree, wavenumber or Robin coefficients are refused
  before any pass runs;
* the transposed form (fmb_gmres_refined_transpose) against the transposed
  FP64 pass and fmb_gmres_transpose;
* the batch over frequencies (fmb_gmres_batch_refined, graph and eager)
  against each plan's FP64 residual and individual refined solves;
* adjoint gradients with the adjoint solve refined (fmb_adjoint_grad_refined)
  against those of fmb_adjoint_grad."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
from conftest import wavenumber
from test_oracle_adjoint import admittance, circle_points

pytestmark = pytest.mark.gpu

FREQ = 8000.0


def _dirs(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


@pytest.fixture(scope="module")
def sphere():
    m = F.Mesh.icosphere(4, 0.1)
    return m, F.Tree(m, depth=5)


@pytest.mark.parametrize("eps_in,robin", [(0.2, False), (0.5, False), (0.2, True)])
def test_refined_solve(sphere, eps_in, robin):
    m, t = sphere
    k = wavenumber(FREQ)
    n = m.arrays()[1].shape[0]
    sig = torch.tensor(1j * k * admittance(n, 2), device="cuda") if robin else None
    outer = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED, nrhs_max=3)
    inner = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED, nrhs_max=3, eps=eps_in)
    po, pi = outer.orders(), inner.orders()
    print("orders outer", po[2:6].tolist(), "inner", pi[2:6].tolist(), flush=True)
    assert pi[5] < po[5]
    b = outer.rhs_plane_wave(_dirs(3))
    tol = 1e-9
    x, info = outer.gmres_refined(inner, b, tol=tol, inner_tol=1e-3)
    print("cycles", info.restarts, "inner steps", info.iters, "outer fp64", info.matvecs_fp64,
          "inner passes", info.matvecs_work, flush=True)
    r = b - outer.matvec(x, prec=F.FMB_FP64)
    rel = (torch.linalg.vector_norm(r, dim=1) / torch.linalg.vector_norm(b, dim=1)).cpu().numpy()
    print("true relres", rel, flush=True)
    assert (rel <= tol).all(), rel
    np.testing.assert_allclose(rel, [info.relres[i] for i in range(3)], rtol=1e-4, atol=0)
    assert info.restarts >= 2 and info.matvecs_fp64 == info.restarts
    outer.set_work_precision(F.FMB_FP64)
    xr, _ = outer.gmres(b, tol=1e-11)
    err = (torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr)).item()
    print("vs FP64 GMRES", err, flush=True)
    assert err < 1e-7, err
    # a converged first guess: one outer check, no cycle
    x2, info2 = outer.gmres_refined(inner, b, x0=x, tol=tol, inner_tol=1e-3)
    assert info2.restarts == 0 and info2.matvecs_fp64 == 1 and info2.iters == 0
    assert torch.equal(x2, x)


def test_refined_rejects_mismatched_plans(sphere):
    m, t = sphere
    k = wavenumber(FREQ)
    n = m.arrays()[1].shape[0]
    outer = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=1)
    b = outer.rhs_plane_wave(_dirs(1))
    sig = torch.tensor(1j * k * admittance(n, 2), device="cuda")
    others = {
        "tree": F.Plan(F.Tree(m, depth=4), k, nrhs_max=1, eps=0.2),
        "wavenumber": F.Plan(t, 1.01 * k, nrhs_max=1, eps=0.2),
        "robin": F.Plan(t, k, robin=sig, nrhs_max=1, eps=0.2),
    }
    for name, inner in others.items():
        with pytest.raises(F.FmbError, match="inner plan|Robin"):
            outer.gmres_refined(inner, b)
        print("refused:", name, flush=True)
    robin_outer = F.Plan(t, k, robin=sig, nrhs_max=1)
    robin_other = F.Plan(t, k, robin=2 * sig, nrhs_max=1, eps=0.2)
    with pytest.raises(F.FmbError, match="Robin"):
        robin_outer.gmres_refined(robin_other, b)
    inner = F.Plan(t, k, nrhs_max=1, eps=0.2)
    for it in (0.0, 1.0, -1e-3):
        with pytest.raises(F.FmbError, match="invalid"):
            outer.gmres_refined(inner, b, inner_tol=it)
    # more right-hand sides than the inner plan holds
    outer3 = F.Plan(t, k, nrhs_max=3)
    b3 = outer3.rhs_plane_wave(_dirs(3))
    with pytest.raises(F.FmbError, match="nrhs_max"):
        outer3.gmres_refined(inner, b3)


def test_refined_transpose(sphere):
    """A^T x = b with a Robin admittance (A^T differs from A): the residual
    recomputed with the outer plan's transposed FP64 pass, and the solution
    against fmb_gmres_transpose with FP64 Arnoldi."""
    m, t = sphere
    k = wavenumber(FREQ)
    n = m.arrays()[1].shape[0]
    sig = torch.tensor(1j * k * admittance(n, 2), device="cuda")
    outer = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED, nrhs_max=2)
    inner = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED, nrhs_max=2, eps=0.5)
    rng = np.random.default_rng(7)
    b = torch.tensor(rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)), device="cuda")
    tol = 1e-9
    x, info = outer.gmres_refined_t(inner, b, tol=tol, inner_tol=1e-3)
    print("cycles", info.restarts, "inner steps", info.iters, "outer fp64", info.matvecs_fp64,
          flush=True)
    r = b - outer.matvec_t(x, prec=F.FMB_FP64)
    rel = (torch.linalg.vector_norm(r, dim=1) / torch.linalg.vector_norm(b, dim=1)).cpu().numpy()
    print("true relres", rel, flush=True)
    assert (rel <= tol).all(), rel
    assert info.matvecs_fp64 == info.restarts >= 2
    outer.set_work_precision(F.FMB_FP64)
    xr, _ = outer.gmres_t(b, tol=1e-11)
    err = (torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr)).item()
    print("vs FP64 transposed GMRES", err, flush=True)
    assert err < 1e-7, err
    # the forward solve of the same system is a different solution
    xf, _ = outer.gmres_refined(inner, b, tol=tol, inner_tol=1e-3)
    assert (torch.linalg.vector_norm(xf - x) / torch.linalg.vector_norm(x)).item() > 1e-3


@pytest.mark.parametrize("graph", [True, False])
def test_batch_refined_matches_individual(graph):
    """fmb_gmres_batch_refined over 1, 4 and 9 kHz (tree depths 3, 4, 5 on one
    ellipsoid mesh, 2, 3 and 1 plane waves, inner plans at eps 0.3): every
    plan's residual recomputed with its own FP64 pass, and the solutions
    against individual fmb_gmres_refined solves."""
    m = F.Mesh.ellipsoid(12, 0.08, 0.07, 0.10, seed=3)
    n = m.arrays()[1].shape[0]
    nr = [2, 3, 1]
    outers, inners, bs, singles = [], [], [], []
    dirs = _dirs(sum(nr))
    c = 0
    for f, dpt, r in zip((1000.0, 4000.0, 9000.0), (3, 4, 5), nr):
        t = F.Tree(m, depth=dpt)
        k = wavenumber(f)
        o = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=r)
        i = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=r, eps=0.3)
        b = o.rhs_plane_wave(dirs[c:c + r])
        c += r
        x, _ = o.gmres_refined(i, b, tol=1e-9, inner_tol=1e-3)
        outers.append(o); inners.append(i); bs.append(b); singles.append(x)
    b = torch.cat(bs, 0).contiguous()
    x, infos = F.gmres_batch_refined(outers, inners, nr, b, graph=graph, tol=1e-9, inner_tol=1e-3)
    assert len(infos) == 3
    c = 0
    for p, (o, r) in enumerate(zip(outers, nr)):
        xp = x[c:c + r].contiguous()
        res = bs[p] - o.matvec(xp, prec=F.FMB_FP64)
        rel = (torch.linalg.vector_norm(res, dim=1) / torch.linalg.vector_norm(bs[p], dim=1)).cpu().numpy()
        err = (torch.linalg.vector_norm(xp - singles[p]) / torch.linalg.vector_norm(singles[p])).item()
        print("plan", p, "cycles", infos[p].restarts, "outer fp64", infos[p].matvecs_fp64,
              "relres", rel, "vs single %.2e" % err, flush=True)
        assert (rel <= 1e-9).all(), (p, rel)
        assert 1 <= infos[p].matvecs_fp64 <= infos[p].restarts
        assert err < 1e-7, (p, err)
        c += r


def test_refined_adjoint_gradient(sphere):
    """J, dJ/dsigma, dJ/dV and dJ/dvn with the adjoint solve refined around an
    eps 0.5 inner plan against fmb_adjoint_grad (both adjoint solves to
    1e-11): 1e-8 relative; J does not depend on the adjoint solve."""
    m, t = sphere
    k = wavenumber(FREQ)
    n = m.arrays()[1].shape[0]
    sig = torch.tensor(1j * k * admittance(n, 12), device="cuda")
    pl = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED)
    inner = F.Plan(t, k, robin=sig, precision=F.FMB_MIXED, eps=0.5)
    d = np.array([[0.0, 0.0, 1.0]])
    x, _ = pl.gmres(pl.rhs_plane_wave(d), tol=1e-11)
    pts = circle_points()
    J0, g0, i0, dV0, dvn0 = pl.adjoint_grad(x, pts, d, tol=1e-11, vertices=True)
    J1, g1, i1, dV1, dvn1 = pl.adjoint_grad(x, pts, d, tol=1e-11, vertices=True, inner=inner)
    print("adjoint: plain %d steps, refined %d cycles %d inner steps" % (i0.iters, i1.restarts, i1.iters),
          flush=True)
    assert i1.restarts >= 2
    assert abs(J1 - J0) <= 1e-13 * abs(J0)
    for a, b in ((g1, g0), (dV1, dV0), (dvn1, dvn0)):
        err = (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()
        print("gradient rel diff %.2e" % err, flush=True)
        assert err < 1e-8, err
    with pytest.raises(F.FmbError, match="inner plan"):
        pl.adjoint_grad(x, pts, d, inner=F.Plan(t, 1.01 * k, robin=sig, eps=0.5))
