"""Writes tests/golden/coax_mp.npz: coaxial translation matrices (R|R) and
(S|R) computed with 50-digit arithmetic (mpmath), for a few |m| blocks.

The values follow the coaxial recurrences of the reference
(translation.cu:555-584 initial column (-1)^n' sqrt(2n'+1) f_n'(kt),
translation.cu:689-711 sectorial, translation.cu:713-745 advancing) evaluated
exactly enough that every printed digit is correct: starting values come
from mpmath's Bessel functions, and the recurrences lose far fewer than the
30 guard digits.  The FP64 oracle and the device generator are measured
against these.  Entries with n' < n use the symmetry
(E|F)^m_{n'n} = (-1)^{n+n'} (E|F)^m_{nn'}, as both FP64 codes do.

Run: python gen_coax_mp.py
"""
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 50

CASES = [
    # kind, p, k, t, blocks m
    ("SR", 94, mp.mpf("183.2"), mp.mpf("0.5"), [0, 1, 20, 47, 80, 93]),
    ("SR", 30, mp.mpf("91.6"), mp.mpf("0.45"), [0, 5, 17, 29]),
    ("RR", 94, mp.mpf("183.2"), mp.mpf("0.108"), [0, 1, 47, 93]),
]


def sph(kind, n, x):
    f = mp.sqrt(mp.pi / (2 * x))
    j = f * mp.besselj(n + mp.mpf(1) / 2, x)
    if kind == "RR":
        return mp.mpc(j, 0)
    y = f * mp.bessely(n + mp.mpf(1) / 2, x)
    return mp.mpc(j, y)


def acoef(n, m):
    if n < abs(m):
        return mp.mpf(0)
    n, m = mp.mpf(n), mp.mpf(m)
    return mp.sqrt((n + 1 + m) * (n + 1 - m) / ((2 * n + 1) * (2 * n + 3)))


def bcoef(n, m):
    nf, mf = mp.mpf(n), mp.mpf(m)
    if m >= 0 and n >= m:
        return mp.sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)))
    if -n <= m < 0:
        return -mp.sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)))
    return mp.mpf(0)


def blocks(kind, p, k, t, ms):
    nmax = 2 * p - 2
    x = k * t
    # sectorial values S[m][n'] = E(m, n', m)
    col = [(-1) ** q * mp.sqrt(2 * q + 1) * sph(kind, q, x) for q in range(nmax + 1)]
    S = {0: col}
    for m in range(0, max(ms)):
        den = bcoef(m + 1, -m - 1)
        cur, nx = S[m], {}
        for q in range(m + 1, nmax - m):
            nx[q] = (bcoef(q, -m - 1) / den) * cur[q - 1] - (bcoef(q + 1, m) / den) * cur[q + 1]
        S[m + 1] = nx
    out = {}
    for m in ms:
        ld = p - m
        E = {m: {q: S[m][q] for q in range(m, nmax - m + 1)}}
        for n in range(m, p - 1):
            an = acoef(n, m)
            nxt = {}
            for q in range(m, nmax - n):
                v0 = E[n][q - 1] if q - 1 >= m else 0
                v1 = E[n][q + 1]
                v2 = E[n - 1][q] if n - 1 >= m else 0
                nxt[q] = (acoef(q - 1, m) * v0 - acoef(q, m) * v1 + acoef(n - 1, m) * v2) / an
            E[n + 1] = nxt
        B = np.zeros((ld, ld), dtype=np.complex128)
        for q in range(m, p):
            for n in range(m, p):
                v = E[n][q] if q >= n else (-1) ** (n + q) * E[q][n]
                B[q - m, n - m] = complex(v)
        out[m] = B
    return out


def main():
    arrays = {}
    for kind, p, k, t, ms in CASES:
        bl = blocks(kind, p, k, t, ms)
        tag = "%s_p%d_kt%g" % (kind, p, float(k * t))
        arrays[tag + "_meta"] = np.array([0 if kind == "RR" else 1, p, float(k), float(t)])
        arrays[tag + "_ms"] = np.array(ms)
        for m, B in bl.items():
            arrays["%s_m%d" % (tag, m)] = B
        print(tag, "done")
    np.savez_compressed(os.path.join(HERE, "coax_mp.npz"), **arrays)


if __name__ == "__main__":
    main()
