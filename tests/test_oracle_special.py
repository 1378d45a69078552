"""Oracle special functions (fo_special.c) against scipy golden vectors and
identities.  Pins: j_n, y_n (scipy), Y_n^m convention (scipy with the
(-1)^m Condon-Shortley factor removed for m >= 0: sphHrmc times GSL Plm,
translation.cu:153-169), orthonormality, the Wronskian, and the normal
derivative formula of dirDev_R (integral.cu:284-331) against central finite
differences."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "special.npz"))


def test_sph_jn_matches_scipy():
    for i, x in enumerate(G["xs"]):
        j = O.sph_jn(120, x)
        ref = G["jn"][i]
        # envelope: relative where |j| is not tiny next to its neighbours
        env = np.maximum.accumulate(np.abs(ref)[::-1])[::-1]
        env = np.maximum(env, 1e-280)  # below: denormal range, scipy returns 0
        assert np.max(np.abs(j - ref) / env) < 1e-12, x


def test_sph_yn_matches_scipy():
    for i, x in enumerate(G["xs"]):
        y = O.sph_yn(120, x)
        ref = G["yn"][i]
        ok = np.isfinite(ref) & (np.abs(ref) < 1e300)
        assert np.max(np.abs(y[ok] - ref[ok]) / np.abs(ref[ok])) < 1e-12, x


@pytest.mark.parametrize("x", [0.3, 2.0, 17.0, 140.0])
def test_wronskian(x):
    j, y = O.sph_jn(80, x), O.sph_yn(80, x)
    n = np.arange(1, 81)
    w = j[n] * y[n - 1] - j[n - 1] * y[n]
    ok = np.abs(y[n]) < 1e250
    np.testing.assert_allclose(w[ok] * x * x, 1.0, rtol=1e-10)


def test_ylm_convention_matches_scipy():
    nY = 24
    for q in range(len(G["th"])):
        Y = O.ylm(nY, G["th"][q], G["ph"][q])
        for n in range(nY + 1):
            for m in range(-n, n + 1):
                eps = (-1) ** m if m >= 0 else 1
                assert abs(Y[n * n + n + m] - eps * G["Y"][q, n * n + n + m]) < 1e-12


def test_ylm_orthonormal():
    nq = 30
    x, w = np.polynomial.legendre.leggauss(nq)
    nph = 64
    ph = np.arange(nph) * 2 * np.pi / nph
    Ys = np.array([O.ylm(12, np.arccos(xx), pp) for xx in x for pp in ph])
    W = np.repeat(w, nph) * 2 * np.pi / nph
    M = (Ys.conj().T * W) @ Ys
    np.testing.assert_allclose(M, np.eye(M.shape[0]), atol=1e-12)


@pytest.mark.parametrize("k", [3.0, 18.3, 91.6])
def test_rbasis_normal_derivative_fd(k):
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal(3) * 0.4 / k
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        h = 1e-6 / k
        d = O.rbasis_dn(10, k, v, nrm)
        fd = (O.rbasis(10, k, v + h * nrm) - O.rbasis(10, k, v - h * nrm)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.max(np.abs(d - fd)) < 1e-7 * scale


def test_sbasis_is_h_times_y():
    v = np.array([0.3, -0.2, 0.5])
    k = 7.0
    r = np.linalg.norm(v)
    S = O.sbasis(6, k, v)
    Y = O.ylm(6, np.arccos(v[2] / r), np.arctan2(v[1], v[0]))
    h = O.sph_jn(6, k * r) + 1j * O.sph_yn(6, k * r)
    for n in range(7):
        np.testing.assert_allclose(S[n * n:(n + 1) ** 2], h[n] * Y[n * n:(n + 1) ** 2], rtol=1e-13)


def test_ab_coefficients_closed_form():
    # translation.cu:189-212
    assert O.lib().fo_acoef(3, 1) == pytest.approx(np.sqrt(5 * 3 / (7 * 9)))
    assert O.lib().fo_bcoef(3, 1) == pytest.approx(np.sqrt(1 * 2 / (5 * 7)))
    assert O.lib().fo_bcoef(3, -2) == pytest.approx(-np.sqrt(4 * 5 / (5 * 7)))
    assert O.lib().fo_bcoef(2, 3) == 0.0
