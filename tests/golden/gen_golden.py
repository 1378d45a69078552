"""Writes tests/golden/special.npz: golden vectors from scipy/numpy for the
oracle's special functions, Gauss rules and the rigid-sphere Mie series.

scipy evaluates these with its own algorithms (Cephes/AMOS/specfun), so the
vectors are independent of the oracle's recurrences.  Run: python gen_golden.py
"""
import os

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def mie_rigid_scipy(ka, ct, nterms):
    n = np.arange(nterms + 1)
    h = sp.spherical_jn(n, ka) + 1j * sp.spherical_yn(n, ka)
    hp = sp.spherical_jn(n, ka, derivative=True) + 1j * sp.spherical_yn(n, ka, derivative=True)
    cn = (2 * n + 1) * (1j ** n) * 1j / (ka ** 2 * hp)
    P = np.array([sp.eval_legendre(q, ct) for q in n])
    return (cn[:, None] * P).sum(0)


def main():
    rng = np.random.default_rng(20260101)
    xs = np.array([1e-3, 0.05, 0.37, 0.9, 1.83, 3.14159, 7.3, 27.5, 91.6, 183.2, 250.0])
    nmax = 120
    n = np.arange(nmax + 1)
    jn = np.array([sp.spherical_jn(n, x) for x in xs])
    yn = np.array([sp.spherical_yn(n, x) for x in xs])
    # Y_n^m with the Condon-Shortley phase (scipy), n <= 24
    th = rng.uniform(0, np.pi, 6)
    ph = rng.uniform(-np.pi, np.pi, 6)
    nY = 24
    Y = np.zeros((6, (nY + 1) ** 2), dtype=np.complex128)
    for q in range(6):
        for nn in range(nY + 1):
            for m in range(-nn, nn + 1):
                if hasattr(sp, "sph_harm_y"):
                    Y[q, nn * nn + nn + m] = sp.sph_harm_y(nn, m, th[q], ph[q])
                else:
                    Y[q, nn * nn + nn + m] = sp.sph_harm(m, nn, ph[q], th[q])
    gl = {}
    for g in (3, 8, 24, 64):
        x, w = np.polynomial.legendre.leggauss(g)
        gl["gl%d_x" % g] = x
        gl["gl%d_w" % g] = w
    ct = np.cos(np.linspace(0, np.pi, 64))
    mie1 = mie_rigid_scipy(2 * np.pi * 1000.0 / 343.0 * 0.1, ct, 40)
    mie2 = mie_rigid_scipy(2 * np.pi * 5000.0 / 343.0 * 0.3, ct, 70)
    np.savez(os.path.join(HERE, "special.npz"), xs=xs, jn=jn, yn=yn, th=th, ph=ph, Y=Y,
             mie_ct=ct, mie_ka1=2 * np.pi * 1000.0 / 343.0 * 0.1, mie1=mie1,
             mie_ka2=2 * np.pi * 5000.0 / 343.0 * 0.3, mie2=mie2, **gl)


if __name__ == "__main__":
    main()
