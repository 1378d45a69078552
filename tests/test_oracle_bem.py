"""Oracle Burton-Miller BEM (fo_bem.c): the FMM matvec against the dense
O(N^2) matvec with the same quadrature, the dense solve against the rigid-
and impedance-sphere Mie series (scipy), the oracle's Mie series against
scipy golden values, and GMRES against numpy.linalg.solve."""
import os

import numpy as np
import pytest
import scipy.special as sp

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import GOLDEN, wavenumber

G = np.load(os.path.join(GOLDEN, "special.npz"))


@pytest.fixture(scope="module")
def sphere2():
    return F.Mesh.icosphere(2, 0.1).arrays()


@pytest.fixture(scope="module")
def sphere3():
    return F.Mesh.icosphere(3, 0.1).arrays()


def test_mie_matches_scipy_golden():
    for ka, key in ((G["mie_ka1"], "mie1"), (G["mie_ka2"], "mie2")):
        got = O.mie_rigid(float(ka), G["mie_ct"], nterms=40 if key == "mie1" else 70)
        np.testing.assert_allclose(got, G[key], rtol=1e-10)


@pytest.mark.parametrize("plev,tol", [({2: 16, 3: 14}, 3e-6), ({2: 22, 3: 20}, 1e-7)])
def test_fmm_matches_dense(sphere3, plev, tol):
    V, Fc = sphere3
    k = wavenumber(1000.0)
    b = O.Bem(V, Fc, k)
    b.fmm_setup(depth=3, plev=plev)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(b.n) + 1j * rng.standard_normal(b.n)
    yd = b.dense_matvec(x)[0]
    yf = b.fmm_matvec(x)[0]
    assert np.linalg.norm(yf - yd) / np.linalg.norm(yd) < tol


def test_fmm_error_decreases_with_order(sphere3):
    V, Fc = sphere3
    k = wavenumber(1000.0)
    x = np.random.default_rng(2).standard_normal(len(Fc)) + 0j
    yd = O.Bem(V, Fc, k).dense_matvec(x)[0]
    errs = []
    for p in (10, 14, 18, 22):
        b = O.Bem(V, Fc, k)
        b.fmm_setup(depth=3, plev={2: p, 3: p})
        errs.append(np.linalg.norm(b.fmm_matvec(x)[0] - yd) / np.linalg.norm(yd))
    assert all(a > 1.5 * c for a, c in zip(errs, errs[1:])), errs


def _mie_impedance(ka, k, a, sigma, ct, nterms):
    """Surface pressure on a sphere with dp/dr = sigma p, plane wave e^{ikz}:
    sum (2n+1) i^n P_n (i / (k a^2)) / (k h_n'(ka) - sigma h_n(ka))."""
    n = np.arange(nterms + 1)
    h = sp.spherical_jn(n, ka) + 1j * sp.spherical_yn(n, ka)
    hp = sp.spherical_jn(n, ka, True) + 1j * sp.spherical_yn(n, ka, True)
    c = (2 * n + 1) * 1j ** n * (1j / (k * a * a)) / (k * hp - sigma * h)
    P = np.array([sp.eval_legendre(q, ct) for q in n])
    return (c[:, None] * P).sum(0)


@pytest.mark.parametrize("Y", [0.0, 0.3 + 0.2j])
def test_dense_solution_converges_to_mie(sphere2, sphere3, Y):
    errs = []
    for V, Fc in (sphere2, sphere3):
        k = wavenumber(1000.0)
        sig = 1j * k * Y * np.ones(len(Fc))
        b = O.Bem(V, Fc, k, sigma=sig if Y else None)
        A = b.dense_matrix()
        u = np.linalg.solve(A, b.rhs_plane([0, 0, 1.0]))
        c, _, _ = b.centroids()
        ct = c[:, 2] / np.linalg.norm(c, axis=1)
        ref = _mie_impedance(k * 0.1, k, 0.1, 1j * k * Y, ct, 40)
        errs.append(np.linalg.norm(u - ref) / np.linalg.norm(ref))
    # rigid: 1.5% -> 0.5%; impedance: 3.4% -> 1.5% (first order in h)
    assert errs[1] < (0.01 if Y == 0 else 0.02) and errs[1] < 0.6 * errs[0], errs


def test_gmres_matches_direct_solve(sphere2):
    V, Fc = sphere2
    k = wavenumber(1500.0)
    b = O.Bem(V, Fc, k)
    A = b.dense_matrix()
    rhs = b.rhs_plane([0.6, 0.0, 0.8])
    u = np.linalg.solve(A, rhs)
    x, it, rr, rc = b.gmres(rhs, use_fmm=False, m=10, tol=1e-10)
    assert rc == 0 and rr <= 1e-10
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-8
    # restarts happened (m=10) and the true residual is what was reported
    assert np.linalg.norm(A @ x - rhs) / np.linalg.norm(rhs) == pytest.approx(rr, rel=1e-3)


def mie_exterior_scattered(k, a, r, ct, nterms=40, sigma=0.0):
    """Scattered pressure of a unit plane wave e^{ikz} on a sphere of radius a
    with dp/dr = sigma p (sigma = 0: rigid) at distance r, angle acos(ct):
    sum (2n+1) i^n c_n h_n(kr) P_n(ct), h_n = j_n + i y_n (e^{ikr}), with
    c_n = -(k j_n'(ka) - sigma j_n(ka)) / (k h_n'(ka) - sigma h_n(ka))."""
    n = np.arange(nterms + 1)
    ja, jpa = sp.spherical_jn(n, k * a), sp.spherical_jn(n, k * a, derivative=True)
    ha = ja + 1j * sp.spherical_yn(n, k * a)
    hp = jpa + 1j * sp.spherical_yn(n, k * a, derivative=True)
    h = sp.spherical_jn(n, k * r) + 1j * sp.spherical_yn(n, k * r)
    c = (2 * n + 1) * (1j ** n) * (-(k * jpa - sigma * ja) / (k * hp - sigma * ha))
    P = np.array([sp.eval_legendre(q, ct) for q in n])
    return ((c * h)[:, None] * P).sum(0)


@pytest.mark.parametrize("Y", [0.0, 0.3 + 0.2j])
def test_field_points_converge_to_exterior_mie(sphere2, sphere3, Y):
    """The representation formula at 32 points on an r = 0.5 m circle (the
    differentiation config's objective points) from the dense solve on a
    rigid (Y = 0) or impedance sphere (dp/dn = ikY p, the sigma_j G term of
    the formula): the scattered field approaches the exterior Mie series
    under mesh refinement, and the incident wave is added exactly."""
    k = wavenumber(1000.0)
    th = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    pts = np.stack([0.5 * np.sin(th), np.zeros_like(th), 0.5 * np.cos(th)], 1)
    exact = mie_exterior_scattered(k, 0.1, 0.5, np.cos(th), sigma=1j * k * Y)
    errs = []
    for V, Fc in (sphere2, sphere3):
        sig = 1j * k * Y * np.ones(len(Fc))
        b = O.Bem(V, Fc, k, sigma=sig if Y else None)
        A = b.dense_matrix()
        u = np.linalg.solve(A, b.rhs_plane([0.0, 0.0, 1.0]))
        ps = b.field(u, pts)[0]
        errs.append(np.linalg.norm(ps - exact) / np.linalg.norm(exact))
        pt = b.field(u, pts, dirs=[[0.0, 0.0, 1.0]])[0]
        np.testing.assert_allclose(pt - ps, np.exp(1j * k * pts[:, 2]), rtol=0, atol=1e-12)
    print("field errors vs Mie", Y, errs)
    assert errs[1] < 0.02, errs
    assert errs[1] < 0.6 * errs[0], errs


def _field_points(rng):
    """32 points on an r = 0.5 m circle, 48 between 1.01 and 2 radii off the
    0.1 m sphere, 16 inside it."""
    th = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    far = np.stack([0.5 * np.cos(th), 0.5 * np.sin(th), 0.1 * np.sin(3 * th)], 1)
    d = rng.standard_normal((64, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    near = d[:48] * rng.uniform(0.101, 0.2, 48)[:, None]
    inside = d[48:] * rng.uniform(0.02, 0.09, 16)[:, None]
    return np.concatenate([far, near, inside])


@pytest.mark.parametrize("robin", [False, True])
def test_field_through_tree_matches_direct(sphere3, robin):
    """The treecode field (fo_bem_field_fmm) against the direct
    representation formula (fo_bem_field): the same panels next to each
    point, expansions for the rest, so the difference is the truncation
    error of the orders; far points see only level-2 expansions."""
    V, Fc = sphere3
    k = wavenumber(1000.0)
    rng = np.random.default_rng(5)
    n = len(Fc)
    sig = 1j * k * (0.01 * (1 + 0.5 * rng.uniform(-1, 1, n))
                    + 0.005j * rng.uniform(-1, 1, n)) if robin else None
    b = O.Bem(V, Fc, k, sigma=sig)
    b.fmm_setup(depth=3, plev={2: 20, 3: 18})
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    pts = _field_points(rng)
    dirs = [[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]]
    pd = b.field(x, pts, dirs)
    pf = b.field(x, pts, dirs, fmm=True)
    e = np.abs(pf - pd) / np.abs(pd).max(axis=1, keepdims=True)
    assert e[:, :32].max() < 1e-12
    assert e.max() < 1e-5


def test_field_through_tree_converges_with_order(sphere3):
    """A wrong sign, index or scale in the expansion part would stall the
    error; with the orders raised it falls by 10x per step."""
    V, Fc = sphere3
    k = wavenumber(1000.0)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(len(Fc)) + 1j * rng.standard_normal(len(Fc))
    pts = _field_points(rng)[32:]
    b = O.Bem(V, Fc, k)
    pd = b.field(x, pts)
    errs = []
    for p in (12, 20, 28):
        b = O.Bem(V, Fc, k)
        b.fmm_setup(depth=3, plev={2: p, 3: p - 2})
        errs.append(np.abs(b.field(x, pts, fmm=True) - pd).max() / np.abs(pd).max())
    assert all(a > 10 * c for a, c in zip(errs, errs[1:])), errs


@pytest.mark.parametrize("ka", [1.8, 27.5])
def test_exterior_mie_series_meets_surface_series(ka):
    """The exterior series (oracle.mie_rigid_field, scipy Bessel functions)
    at r = a against the oracle's surface series (fo_mie_rigid, Wronskian
    form): the same scattering coefficients; and far out it tends to the
    incident wave plus an outgoing e^{ikr}/r term."""
    k, a = ka / 0.1, 0.1
    ct = np.linspace(-1, 1, 41)
    st = np.sqrt(1 - ct ** 2)
    pts = a * np.stack([st, 0 * st, ct], 1)
    got = O.mie_rigid_field(k, a, pts, [0.0, 0.0, 1.0])
    want = O.mie_rigid(ka, ct, nterms=int(ka) + 40)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)
    r = np.array([1.0, 2.0]) * 400 * ka * ka * a
    far = np.stack([0 * r, r * 0.6, r * 0.8], 1)
    ps = O.mie_rigid_field(k, a, far, [0.0, 0.0, 1.0]) - np.exp(1j * k * far[:, 2])
    f = ps * r * np.exp(-1j * k * r)
    assert abs(f[0] - f[1]) < 0.02 * abs(f[1])
