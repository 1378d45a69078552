"""Oracle quadrature and panel integrals (fo_quad.c).  Pins: Gauss-Legendre
nodes (numpy golden), polynomial exactness of the triangle rules, the kernel
formulas of integral.cu:72-78, 391-415, 602-622 against finite differences,
the centroid self terms against a k = 0 closed form and an independent
sub-triangle Duffy rule (the reference's own split, integral.cu:202-249),
the hypersingular term against the reference-style theta rule
(integral.cu:666-751), and closed-surface identities."""
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "special.npz"))
TRI = np.array([[0.0, 0.0, 0.0], [0.013, 0.001, 0.0], [0.004, 0.011, 0.0]])


@pytest.mark.parametrize("n", [3, 8, 24, 64])
def test_gauss_legendre_matches_numpy(n):
    x, w = O.gauss_legendre(n)
    np.testing.assert_allclose(x, G["gl%d_x" % n], atol=1e-14)
    np.testing.assert_allclose(w, G["gl%d_w" % n], atol=1e-14)


def test_gauss3_closed_form():
    x, w = O.gauss_legendre(3)
    np.testing.assert_allclose(x, [-math.sqrt(0.6), 0, math.sqrt(0.6)], atol=1e-15)
    np.testing.assert_allclose(w, [5 / 9, 8 / 9, 5 / 9], atol=1e-15)


def _mono_integral(rule, order, a, b):
    B, w = O.tri_rule(rule, order)
    # reference triangle (1,0),(0,1),(0,0): point = b0*v0 + b1*v1 + b2*v2
    x, y = B[:, 0], B[:, 1]
    return 0.5 * np.sum(w * x ** a * y ** b)


@pytest.mark.parametrize("rule,order,deg", [(O.RULE_GAUSS7, 3, 5), (O.RULE_COLLAPSED, 3, 4),
                                            (O.RULE_COLLAPSED, 8, 14)])
def test_triangle_rule_exactness(rule, order, deg):
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert _mono_integral(rule, order, a, b) == pytest.approx(exact, rel=1e-12, abs=1e-15)


def test_kernels_match_finite_differences():
    k = 40.0
    x = np.array([0.01, 0.02, -0.005])
    y = np.array([-0.01, 0.004, 0.012])
    nx = np.array([0.3, -0.4, 0.866])
    nx /= np.linalg.norm(nx)
    ny = np.array([-0.6, 0.1, 0.79])
    ny /= np.linalg.norm(ny)
    K = O.kernels(k, x, nx, y, ny)
    h = 1e-7
    g = lambda xx, yy: O.kernels(k, xx, nx, yy, ny)[0]
    dny = (g(x, y + h * ny) - g(x, y - h * ny)) / (2 * h)
    dnx = (g(x + h * nx, y) - g(x - h * nx, y)) / (2 * h)
    assert abs(K[1] - dny) < 1e-6 * abs(dny)
    assert abs(K[2] - dnx) < 1e-6 * abs(dnx)
    gy = lambda xx: O.kernels(k, xx, nx, y, ny)[1]
    d2 = (gy(x + h * nx) - gy(x - h * nx)) / (2 * h)
    assert abs(K[3] - d2) < 1e-5 * abs(d2)


def _self_k0_closed(tri):
    """int_T 1/(4 pi |x - c|) dS from the centroid: (1/4pi) sum_e h asinh(s/h)."""
    c = tri.mean(0)
    tot = 0.0
    for e in range(3):
        P, Q = tri[e], tri[(e + 1) % 3]
        L = np.linalg.norm(Q - P)
        ev = (Q - P) / L
        tp = np.dot(c - P, ev)
        h = np.linalg.norm(P + tp * ev - c)
        tot += h * (math.asinh((L - tp) / h) - math.asinh(-tp / h))
    return tot / (4 * math.pi)


def test_self_single_layer_k0_closed_form():
    S = O.panel_self(0.0, TRI, nedge=24)[0]
    assert S.real == pytest.approx(_self_k0_closed(TRI), rel=1e-12)
    assert abs(S.imag) < 1e-18


def _subtri_duffy(k, tri, which, order):
    """Split at the centroid into three triangles with the centroid as the
    collapsed (singular) vertex (integral.cu:202-249), high-order rule."""
    c = tri.mean(0)
    n = np.cross(tri[0] - tri[2], tri[1] - tri[2])
    n /= np.linalg.norm(n)
    tot = 0.0
    for e in range(3):
        sub = np.array([tri[e], tri[(e + 1) % 3], c])
        tot += O.panel_regular(k, sub, c, n, rule=O.RULE_COLLAPSED, order=order)[which]
    return tot


@pytest.mark.parametrize("k", [18.3, 91.6, 183.2])
def test_self_single_layer_vs_duffy(k):
    S = O.panel_self(k, TRI, nedge=32)[0]
    ref = _subtri_duffy(k, TRI, 0, 40)
    assert abs(S - ref) < 1e-9 * abs(ref)


def test_hypersingular_k0_vs_reference_theta_rule():
    T = O.panel_self(0.0, TRI, nedge=32)[3]
    ref = O.panel_hyper_ref(0.0, TRI, 20000)
    assert abs(T - ref) < 2e-4 * abs(T)


@pytest.mark.parametrize("k", [18.3, 91.6])
def test_hypersingular_k_part_vs_duffy(k):
    """T(k) - T(0) has an O(1/r) kernel: brute force converges."""
    dT = O.panel_self(k, TRI, nedge=32)[3] - O.panel_self(0.0, TRI, nedge=32)[3]
    c = TRI.mean(0)
    n = np.cross(TRI[0] - TRI[2], TRI[1] - TRI[2])
    n /= np.linalg.norm(n)
    tot = 0.0
    for e in range(3):
        sub = np.array([TRI[e], TRI[(e + 1) % 3], c])
        B, w = O.tri_rule(O.RULE_COLLAPSED, 48)
        area = 0.5 * np.linalg.norm(np.cross(sub[0] - sub[2], sub[1] - sub[2]))
        for q in range(len(w)):
            y = B[q] @ sub
            tot += w[q] * area * (O.kernels(k, c, n, y, n)[3] - O.kernels(1e-9, c, n, y, n)[3])
    assert abs(dT - tot) < 1e-6 * abs(dT)


def test_flat_panel_double_layer_self_terms_vanish():
    v = O.panel_self(50.0, TRI)
    assert v[1] == 0 and v[2] == 0


def test_closed_surface_identities():
    """k -> 0 on a closed sphere: sum_j K_ij = -1/2 (Gauss) and sum_j T_ij = 0
    (the double layer of a constant is piecewise constant)."""
    import fmbem_gpu_c_amd as F
    m = F.Mesh.icosphere(3, 0.1)
    V, Fc = m.arrays()
    k = 1e-6
    b0 = O.Bem(V, Fc, k, alpha=0.0)
    b1 = O.Bem(V, Fc, k, alpha=1.0)
    rows = [0, 37, 400, 1000]
    for i in rows:
        r0 = b0.dense_row(i)
        r1 = b1.dense_row(i)
        # A(alpha=0) row = 1/2 - K_i.  ->  sum = 1/2 + 1/2
        assert abs(r0.sum() - 1.0) < 5e-5
        # A(1) - A(0) = -T_i: the hypersingular row sum vanishes up to
        # discretisation error, relative to the self term
        T = -(r1 - r0)
        assert abs(T.sum()) < 1e-3 * abs(T[i])
