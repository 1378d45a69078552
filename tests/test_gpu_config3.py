"""The batched-frequency workload at full size.

The perturbed ellipsoid (semi-axes 0.08 / 0.07 / 0.10 m, spherical-harmonic
perturbation of degrees 2-8 at 5% RMS, geodesic frequency 55: 60,500
triangles) at 16 frequencies log-spaced over 1-16 kHz, each with its own
plan: octree depth 5 below 4 kHz and 6 from there, truncNum orders per level
(uneven depths and orders), tile shapes autotuned, 8 plane waves per
frequency on a Fibonacci sphere of directions.  The reference solves one
wavenumber per genOctree / OprL (translation.cu:3007, 3412), so the batch has
no reference counterpart; its parts do:

* one frequency's FP64 batch matvec against the oracle FMM on the same tree
  and orders (1e-9 relative), inside an eager batch pass and the replays of
  the captured graph;
* the FP32 batch pass against the FP64 one (1e-5 relative);
* fmb_gmres_batch (graph-captured, mixed precision, tolerance 1e-6) against
  each frequency's own fmb_gmres: every plan's FP64 true residual <= 1e-6 and
  the solutions within 2e-5 relative (both meet the tolerance; plans retire
  at different steps, so the iterates differ at the tolerance's scale)."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber

pytestmark = pytest.mark.gpu

FREQS = [1000.0 * 16.0 ** (i / 15.0) for i in range(16)]
NRHS = 8


def depth_of(f):
    return 5 if f < 4000.0 else 6


def fibonacci_dirs(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


@pytest.fixture(scope="module")
def workload():
    m = F.Mesh.ellipsoid(55, 0.08, 0.07, 0.10, 2, 8, 0.05, seed=7)
    V, Fc = m.arrays()
    assert Fc.shape[0] == 60500
    trees = {d: F.Tree(m, depth=d) for d in (5, 6)}
    plans = [F.Plan(trees[depth_of(f)], wavenumber(f), precision=F.FMB_MIXED, nrhs_max=NRHS,
                    autotune=1) for f in FREQS]
    return m, V, Fc, trees, plans


def test_uneven_depths_and_orders(workload):
    m, V, Fc, trees, plans = workload
    orders = [tuple(p.orders()[2:depth_of(f) + 1].tolist()) for f, p in zip(FREQS, plans)]
    print("orders per frequency:", orders, flush=True)
    assert {len(o) for o in orders} == {4, 5}
    assert len(set(orders)) >= 12
    assert orders[-1][0] > orders[0][0]
    for p in plans:
        for l in range(2, p.tree.info()[1] + 1):
            mode, nt, mt, nw, ms = p.tile_shape(0, l, F.FMB_MIXED)
            assert mode in (1, 2, 3) and ms > 0


def test_batch_matvec_against_oracle_and_fp32(workload):
    m, V, Fc, trees, plans = workload
    n = Fc.shape[0]
    rng = np.random.default_rng(17)
    x = rng.standard_normal((NRHS * len(plans), n)) + 1j * rng.standard_normal((NRHS * len(plans), n))
    xt = torch.tensor(x, device="cuda")
    nr = [NRHS] * len(plans)
    e64 = F.matvec_batch(plans, nr, xt, graph=False, prec=F.FMB_FP64)[0].cpu().numpy()
    g64 = F.matvec_batch(plans, nr, xt, graph=True, prec=F.FMB_FP64, reps=2).cpu().numpy()
    g32 = F.matvec_batch(plans, nr, xt, graph=True, prec=F.FMB_MIXED, reps=2).cpu().numpy()
    for rep in range(2):
        assert np.linalg.norm(g64[rep] - e64) / np.linalg.norm(e64) < 1e-13
        err32 = np.linalg.norm(g32[rep] - e64) / np.linalg.norm(e64)
        print("replay %d: FP32 batch vs FP64 %.2e" % (rep, err32), flush=True)
        assert err32 < 1e-5
    # one frequency (2.4 kHz, depth 5) against the oracle FMM on two of its
    # right-hand sides
    i = 5
    f, pl = FREQS[i], plans[i]
    L = depth_of(f)
    ob = O.Bem(V, Fc, wavenumber(f))
    ob.fmm_setup(depth=L, plev={l: int(pl.orders()[l]) for l in range(2, L + 1)})
    assert ob.orders()[1][2:L + 1].tolist() == pl.orders()[2:L + 1].tolist()
    cols = slice(i * NRHS, i * NRHS + 2)
    want = ob.fmm_matvec(x[cols])
    for name, got in (("eager", e64[cols]), ("graph", g64[0][cols])):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print("%.0f Hz %s batch vs oracle FMM %.2e" % (f, name, err), flush=True)
        assert err < 1e-9


def test_batch_solve_matches_individual_solves(workload):
    m, V, Fc, trees, plans = workload
    dirs = fibonacci_dirs(NRHS)
    for p in plans:
        p.set_work_precision(F.FMB_MIXED)
    b = torch.cat([p.rhs_plane_wave(dirs) for p in plans], 0).contiguous()
    torch.cuda.synchronize()
    xb, infos = F.gmres_batch(plans, [NRHS] * len(plans), b, graph=True, tol=1e-6, maxit=3000)
    for p, (f, info) in enumerate(zip(FREQS, infos)):
        rr = max(info.relres[r] for r in range(NRHS))
        assert rr <= 1e-6, (f, rr)
    xb = xb.cpu().numpy()
    worst = 0.0
    for p, (f, pl) in enumerate(zip(FREQS, plans)):
        xs, info = pl.gmres(b[p * NRHS:(p + 1) * NRHS].contiguous(), tol=1e-6, maxit=3000)
        xs = xs.cpu().numpy()
        err = np.linalg.norm(xb[p * NRHS:(p + 1) * NRHS] - xs) / np.linalg.norm(xs)
        worst = max(worst, err)
        print("%.0f Hz: batch vs individual %.2e (iters %d)" % (f, err, info.iters), flush=True)
    assert worst < 2e-5
