"""Frequency sweeps on the GPU: a plan re-targeted with fmb_plan_retune
against a plan created at the new wavenumber (same operator up to the
summation order of the atomics: FP64 pass 1e-13, FP32 pass 1e-5 relative)
and against the FP64 oracle (1e-9); warm starts of GMRES across
neighbouring frequencies."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber

pytestmark = pytest.mark.gpu


def _robin(n, k, seed):
    rng = np.random.default_rng(seed)
    Y = 0.01 * (1 + 0.5 * rng.uniform(-1, 1, n)) + 0.005j * rng.uniform(-1, 1, n)
    return 1j * k * Y


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,robin,k_max_hz", [
    (F.FMB_FP64, False, 0.0), (F.FMB_MIXED, False, 0.0),
    (F.FMB_FP64, True, 0.0), (F.FMB_MIXED, True, 5000.0)])
def test_retune_matches_fresh_plan(sphere_l3, precision, robin, k_max_hz):
    """1 kHz -> 4 kHz (orders grow: the rotation tables are rebuilt unless
    k_max covers them) -> 700 Hz (orders shrink: the tables are kept)."""
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    t = F.Tree(m, s=5)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    xt = torch.tensor(x, device="cuda")
    kw = dict(precision=precision, nrhs_max=2, k_max=wavenumber(k_max_hz))

    def sig(k):
        return torch.tensor(_robin(n, k, 11), device="cuda") if robin else None

    k0 = wavenumber(1000.0)
    pl = F.Plan(t, k0, robin=sig(k0), **kw)
    precs = [F.FMB_FP64] + ([F.FMB_MIXED] if precision == F.FMB_MIXED else [])
    for f in (4000.0, 700.0):
        k = wavenumber(f)
        pl.retune(k, robin=sig(k))
        assert pl.k == k
        fresh = F.Plan(t, k, robin=sig(k), **kw)
        np.testing.assert_array_equal(pl.orders(), fresh.orders())
        for pr in precs:
            y1 = pl.matvec(xt, prec=pr).cpu().numpy()
            y2 = fresh.matvec(xt, prec=pr).cpu().numpy()
            tol = 1e-13 if pr == F.FMB_FP64 else 1e-5
            assert _rel(y1, y2) < tol, (f, pr, _rel(y1, y2))
        if not robin and f == 4000.0:
            ob = O.Bem(V, Fc, k)
            ob.fmm_setup(s=5)
            yo = ob.fmm_matvec(x)
            y1 = pl.matvec(xt, prec=F.FMB_FP64).cpu().numpy()
            assert _rel(y1, yo) < 1e-9
        del fresh


def test_failed_retune_leaves_no_operator(sphere_l3):
    m, V, Fc = sphere_l3
    t = F.Tree(m, s=5)
    k = wavenumber(1000.0)
    pl = F.Plan(t, k)
    x = torch.ones(Fc.shape[0], dtype=torch.complex128, device="cuda")
    y0 = pl.matvec(x).cpu().numpy()
    with pytest.raises(F.FmbError):
        pl.retune(5000.0)  # level-2 order above 200
    assert pl.k == 0.0
    with pytest.raises(F.FmbError):
        pl.matvec(x)
    pl.retune(k)
    assert _rel(pl.matvec(x).cpu().numpy(), y0) < 1e-13


def test_sweep_warm_start(sphere_l3):
    """Neighbouring frequencies 1000 -> 1060 Hz, two plane waves: the
    phase-carrying guess starts closer to the solution than the previous
    solution itself and saves Arnoldi steps over a cold start; all three
    starts reach the same solution."""
    m, V, Fc = sphere_l3
    t = F.Tree(m, s=5)
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    k1, k2 = wavenumber(1000.0), wavenumber(1060.0)
    pl = F.Plan(t, k1, nrhs_max=2)
    x1, _ = pl.gmres(pl.rhs_plane_wave(dirs), tol=1e-10)
    pl.retune(k2)
    b = pl.rhs_plane_wave(dirs)
    bn = b.cpu().numpy()

    def res0(x0):
        r = bn - pl.matvec(x0).cpu().numpy()
        return max(np.linalg.norm(r[i]) / np.linalg.norm(bn[i]) for i in range(2))

    guess = pl.sweep_guess(k1, x1, dirs)
    r_prev, r_guess = res0(x1), res0(guess)
    assert r_guess < r_prev < 1.0, (r_guess, r_prev)
    xc, ic = pl.gmres(b, tol=1e-10)
    xg, ig = pl.gmres(b, x0=guess, tol=1e-10)
    assert ig.iters < ic.iters, (ig.iters, ic.iters)
    assert max(ig.relres[0], ig.relres[1]) <= 1e-10
    assert _rel(xg.cpu().numpy(), xc.cpu().numpy()) < 1e-8
    # dirs=None leaves the solution unchanged
    same = pl.sweep_guess(k1, x1)
    assert torch.equal(same, x1)
