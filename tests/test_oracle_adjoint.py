"""The adjoint gradient formula, pinned on the oracle's dense operator.

J = sum_q |p(y_q)|^2 at 32 points on an r = 0.5 m circle, p from the dense
Burton-Miller solve on a 320-triangle sphere with per-element Robin
coefficients sigma = i k Y.  The gradient assembled from the adjoint formula
(the same formula fmb_adjoint_grad evaluates with FMM operators),
  dJ/dRe sigma_j + i dJ/dIm sigma_j = 2 conj(-x_j ((C^T lam)_j + (Sf^T u)_j)),
  u = conj(p),  A^T lam = F^T u,  C = dA/dsigma's column factor,
must match central finite differences of J in Re sigma_j and Im sigma_j.
C and Sf come from the oracle by differences of operators that are affine
in sigma (exact up to rounding)."""
import numpy as np
import pytest

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber


def circle_points(r=0.5, n=32):
    th = np.linspace(0, 2 * np.pi, n, endpoint=False)
    return np.stack([r * np.sin(th), np.zeros_like(th), r * np.cos(th)], 1)


def admittance(n, seed):
    rng = np.random.default_rng(seed)
    return 0.01 * (1 + 0.5 * rng.uniform(-1, 1, n)) + 0.005j * rng.uniform(-1, 1, n)


def dense_adjoint_gradient(V, Fc, k, sig, pts, d):
    """(J, dJ/dsigma) from the oracle's dense matrices; also returns x."""
    n = len(Fc)
    b0 = O.Bem(V, Fc, k, sigma=sig)
    A = b0.dense_matrix()
    C = O.Bem(V, Fc, k, sigma=sig + 1.0).dense_matrix() - A
    rhs = b0.rhs_plane(d)
    x = np.linalg.solve(A, rhs)
    eye = np.eye(n)
    Fm = b0.field(eye, pts).T                        # F[q][j]
    F0 = O.Bem(V, Fc, k).field(eye, pts).T           # sigma = 0: Kf
    F1 = O.Bem(V, Fc, k, sigma=np.ones(n, complex)).field(eye, pts).T
    Sf = F0 - F1
    p = b0.field(x, pts, dirs=[d])[0]
    u = np.conj(p)
    lam = np.linalg.solve(A.T, Fm.T @ u)
    w = -x * (C.T @ lam + Sf.T @ u)
    return np.sum(np.abs(p) ** 2), 2 * np.conj(w), x


def objective(V, Fc, k, sig, pts, d):
    b = O.Bem(V, Fc, k, sigma=sig)
    x = np.linalg.solve(b.dense_matrix(), b.rhs_plane(d))
    p = b.field(x, pts, dirs=[d])[0]
    return np.sum(np.abs(p) ** 2)


def test_adjoint_formula_matches_central_differences():
    V, Fc = F.Mesh.icosphere(2, 0.1).arrays()
    k = wavenumber(4000.0)
    n = len(Fc)
    sig = 1j * k * admittance(n, 4)
    pts = circle_points()
    d = np.array([0.0, 0.0, 1.0])
    J, g, _ = dense_adjoint_gradient(V, Fc, k, sig, pts, d)
    assert J > 0
    h = 1e-6 * k  # step 1e-6 in the admittance Y (sigma = i k Y)
    rng = np.random.default_rng(1)
    for j in rng.choice(n, 4, replace=False):
        e = np.zeros(n, complex)
        e[j] = h
        dre = (objective(V, Fc, k, sig + e, pts, d) - objective(V, Fc, k, sig - e, pts, d)) / (2 * h)
        dim = (objective(V, Fc, k, sig + 1j * e, pts, d)
               - objective(V, Fc, k, sig - 1j * e, pts, d)) / (2 * h)
        fd = dre + 1j * dim
        assert abs(g[j] - fd) <= 1e-6 * abs(fd), (j, g[j], fd)


def test_planted_sign_error_fails_the_check():
    """Dropping the field-point term Sf^T u (a plausible slip) is caught by
    the same finite differences."""
    V, Fc = F.Mesh.icosphere(2, 0.1).arrays()
    k = wavenumber(4000.0)
    n = len(Fc)
    sig = 1j * k * admittance(n, 4)
    pts = circle_points()
    d = np.array([0.0, 0.0, 1.0])
    b0 = O.Bem(V, Fc, k, sigma=sig)
    A = b0.dense_matrix()
    C = O.Bem(V, Fc, k, sigma=sig + 1.0).dense_matrix() - A
    x = np.linalg.solve(A, b0.rhs_plane(d))
    Fm = b0.field(np.eye(n), pts).T
    u = np.conj(b0.field(x, pts, dirs=[d])[0])
    lam = np.linalg.solve(A.T, Fm.T @ u)
    g_wrong = 2 * np.conj(-x * (C.T @ lam))
    j = 7
    h = 1e-6 * k  # step 1e-6 in the admittance Y (sigma = i k Y)
    e = np.zeros(n, complex)
    e[j] = h
    dre = (objective(V, Fc, k, sig + e, pts, d) - objective(V, Fc, k, sig - e, pts, d)) / (2 * h)
    dim = (objective(V, Fc, k, sig + 1j * e, pts, d) - objective(V, Fc, k, sig - 1j * e, pts, d)) / (2 * h)
    fd = dre + 1j * dim
    assert abs(g_wrong[j] - fd) > 1e-3 * abs(fd)
