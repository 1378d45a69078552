"""Batched solves over frequencies on the GPU (fmb_gmres_batch).

Three plans on one perturbed-ellipsoid mesh at 1, 4 and 9 kHz (different
tree depths and per-level orders) with 2, 3 and 1 plane waves, solved in
lock step, with the FMM passes replayed from one hipGraph (one branch per
plan) and eagerly.  Every replay feeds the Krylov process with eager work
and host synchronisations between replays; the solutions are checked
against individual fmb_gmres solves of each plan (1e-7 relative, tolerance
1e-9) and between graph and eager runs."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
from conftest import wavenumber

pytestmark = pytest.mark.gpu


def fibonacci_dirs(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def _plans(precision, nr):
    m = F.Mesh.ellipsoid(12, 0.08, 0.07, 0.10, seed=3)
    out = []
    for f, dpt, r in zip((1000.0, 4000.0, 9000.0), (3, 4, 5), nr):
        t = F.Tree(m, depth=dpt)
        out.append(F.Plan(t, wavenumber(f), precision=precision, nrhs_max=r))
    return m, out


@pytest.mark.parametrize("precision", [F.FMB_FP64, F.FMB_MIXED])
def test_batch_matvec_graph_replays(precision):
    """Three replays of one captured graph (eager copies and host
    synchronisations between them) and the eager batch pass against each
    plan's own fmb_matvec: FP64 1e-13, FP32 pass 1e-6 relative."""
    nr = [2, 3, 1]
    m, plans = _plans(precision, nr)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(5)
    x = torch.tensor(rng.standard_normal((6, n)) + 1j * rng.standard_normal((6, n)), device="cuda")
    precs = [F.FMB_FP64] + ([F.FMB_MIXED] if precision == F.FMB_MIXED else [])
    for prec in precs:
        want = torch.cat([pl.matvec(x[o:o + r].contiguous(), prec=prec)
                          for pl, o, r in zip(plans, (0, 2, 5), nr)], 0).cpu().numpy()
        tol = 1e-13 if prec == F.FMB_FP64 else 1e-6
        eager = F.matvec_batch(plans, nr, x, graph=False, prec=prec, reps=2).cpu().numpy()
        graph = F.matvec_batch(plans, nr, x, graph=True, prec=prec, reps=3).cpu().numpy()
        for name, ys in (("eager", eager), ("graph", graph)):
            for rep, y in enumerate(ys):
                err = np.linalg.norm(y - want) / np.linalg.norm(want)
                print(name, "prec", prec, "replay", rep, "rel err %.2e" % err, flush=True)
                assert err < tol, (name, prec, rep, err)


@pytest.mark.parametrize("precision", [F.FMB_FP64, F.FMB_MIXED])
def test_batch_matches_individual_solves(precision):
    m = F.Mesh.ellipsoid(12, 0.08, 0.07, 0.10, seed=3)
    n = m.arrays()[1].shape[0]
    freqs, depths, nr = (1000.0, 4000.0, 9000.0), (3, 4, 5), [2, 3, 1]
    plans, bs, singles = [], [], []
    dirs = fibonacci_dirs(6)
    c = 0
    for f, dpt, r in zip(freqs, depths, nr):
        t = F.Tree(m, depth=dpt)
        pl = F.Plan(t, wavenumber(f), precision=precision, nrhs_max=r)
        d = dirs[c:c + r]
        c += r
        b = pl.rhs_plane_wave(d)
        x, info = pl.gmres(b, tol=1e-9)
        plans.append((t, pl))
        bs.append(b)
        singles.append(x)
    orders = [pl.orders()[2:t.info()[1] + 1].tolist() for t, pl in plans]
    assert len({len(o) for o in orders}) == 3  # uneven depths
    b = torch.cat(bs, 0).contiguous()
    xs = torch.cat(singles, 0).cpu().numpy()
    res = {}
    for graph in (True, False):
        x, infos = F.gmres_batch([pl for _, pl in plans], nr, b, graph=graph, tol=1e-9)
        assert len(infos) == 3
        for p, info in enumerate(infos):
            assert max(info.relres[i] for i in range(nr[p])) <= 1e-9
        res[graph] = x.cpu().numpy()
        for i in range(sum(nr)):
            err = np.linalg.norm(res[graph][i] - xs[i]) / np.linalg.norm(xs[i])
            assert err < 1e-7, (graph, i, err)
    assert np.linalg.norm(res[True] - res[False]) / np.linalg.norm(res[False]) < 1e-7
