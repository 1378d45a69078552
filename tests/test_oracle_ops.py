"""Oracle rotate-coaxial-rotate operators (fo_ops.c).  Pins: orthogonality of
every rotation block up to p = 120, identity at beta = 0, the closed-form
Wigner d^1, the expansion of R_n^m about a rotated frame, agreement with the
reference's own rotation recurrence (sparseRotMatGen, translation.cu:
1171-1237) where that recurrence is still accurate (p <= 12), the (S|R) and
(R|R) addition theorems evaluated directly, the coaxial symmetry
(E|F)_{n'n} = (-1)^{n+n'} (E|F)_{nn'}, the reference invariant
reduce(p1 -> p2) == generate(p2) (main.c:29-38), and the m <-> -m symmetry
of the rotation blocks behind the even/odd split of the translations."""
import numpy as np
import pytest

import oracle as O


def _blocks(H, p):
    out = []
    for n in range(p):
        off = n * (4 * n * n - 1) // 3
        ld = 2 * n + 1
        out.append(H[off:off + ld * ld].reshape(ld, ld).T)  # [m'][m]
    return out


@pytest.mark.parametrize("p", [20, 60, 120])
@pytest.mark.parametrize("beta", [0.2, np.arccos(1 / np.sqrt(3)), np.pi / 2, 2.4, np.pi])
def test_rotation_blocks_orthogonal(p, beta):
    for n, B in enumerate(_blocks(O.rot_H(p, beta), p)):
        assert np.max(np.abs(B @ B.T - np.eye(2 * n + 1))) < 1e-12


def test_rotation_at_zero_angle():
    """H(0) = diag((-1)^m) in this normalisation (the reference recurrence
    gives the same: H_n^{1,1}(0) = -b_{n+1}^0 / b_{n+1}^0); forward and back
    rotation in the RCR then compose to the identity."""
    for n, B in enumerate(_blocks(O.rot_H(15, 0.0), 15)):
        np.testing.assert_allclose(B, np.diag((-1.0) ** np.arange(-n, n + 1)), atol=1e-14)
        np.testing.assert_allclose(B @ B, np.eye(2 * n + 1), atol=1e-14)
    np.testing.assert_allclose(O.rot_H(15, 0.0), O.rot_H_ref(15, 0.0), atol=1e-14)


def test_wigner_d1_closed_form():
    b = 0.7
    B = _blocks(O.rot_H(3, b), 3)[1]
    c, s = np.cos(b), np.sin(b)
    d1 = np.array([[(1 + c) / 2, s / np.sqrt(2), (1 - c) / 2],
                   [-s / np.sqrt(2), c, s / np.sqrt(2)],
                   [(1 - c) / 2, -s / np.sqrt(2), (1 + c) / 2]])  # d^1_{m'm}, m' = -1..1
    eps = lambda m: (-1) ** m if m >= 0 else 1
    H = np.array([[eps(mp) * (-1) ** m * eps(m) * d1[mp + 1, m + 1] for m in (-1, 0, 1)]
                  for mp in (-1, 0, 1)])
    np.testing.assert_allclose(B, H, atol=1e-15)


@pytest.mark.parametrize("beta", [0.3, np.arccos(1 / np.sqrt(3)), np.pi / 2, 2.5])
def test_rotation_even_odd_split(beta):
    """The symmetries the matrix-core translations rely on: H_n^{m'm} =
    H_n^{-m',-m} and H_n symmetric (checked to p = 60, also against the
    reference's recurrence at p = 10).  In the coordinates e_m = (z_m +
    z_{-m}) / 2 (e_0 = z_0), o_m = (z_m - z_{-m}) / 2, y = H_n z then splits
    into an even block (E_{m'0} = H^{m'0}, E_{m'm} = H^{m'm} + H^{m',-m}) and
    an odd block (O_{m'm} = H^{m'm} - H^{m',-m}), recombined as y_{+-m} =
    e_m +- o_m: checked here against the full product on random vectors."""
    rng = np.random.default_rng(11)
    # (the reference's recurrence carries about 1e-11 of rounding at p = 10)
    for H, p, tol in ((O.rot_H(60, beta), 60, 1e-13), (O.rot_H_ref(10, beta), 10, 1e-10)):
        for n, B in enumerate(_blocks(H, p)):
            np.testing.assert_allclose(B, B[::-1, ::-1], atol=tol)
            np.testing.assert_allclose(B, B.T, atol=tol)
    for n, B in enumerate(_blocks(O.rot_H(40, beta), 40)):
        h = lambda mp, m: B[mp + n, m + n]
        E = np.array([[h(i, k) + h(i, -k) if k else h(i, 0) for k in range(n + 1)]
                      for i in range(n + 1)])
        Od = np.array([[h(i, k) - h(i, -k) for k in range(1, n + 1)] for i in range(1, n + 1)])
        z = rng.standard_normal(2 * n + 1) + 1j * rng.standard_normal(2 * n + 1)
        zp, zm = z[n:], z[n::-1]          # z_{m}, z_{-m}, m = 0..n
        e = (zp + zm) / 2
        e[0] = z[n]
        ye = E @ e
        yo = Od @ ((zp - zm) / 2)[1:] if n else np.zeros(0)
        y = np.empty(2 * n + 1, complex)
        y[n] = ye[0]
        y[n + 1:] = ye[1:] + yo
        y[:n] = (ye[1:] - yo)[::-1]
        np.testing.assert_allclose(y, B @ z, atol=1e-12 * max(1.0, np.abs(z).max()))


@pytest.mark.parametrize("beta", [0.4, np.arccos(1 / np.sqrt(3)), 2.6])
def test_matches_reference_recurrence_at_small_p(beta):
    np.testing.assert_allclose(O.rot_H(12, beta), O.rot_H_ref(12, beta), atol=1e-9)


def test_reference_recurrence_unstable_at_large_p():
    """The reference's recurrence in m drifts away from orthogonality
    (documented defect); the degree recurrence used instead does not."""
    bad = max(np.max(np.abs(B @ B.T - np.eye(B.shape[0])))
              for B in _blocks(O.rot_H_ref(30, 0.955), 30))
    assert bad > 1e-2


def test_rotated_frame_expansion():
    """sum_m T^{m'm}(alpha,beta,gamma) C^m R^m(x) evaluated in the rotated
    frame reproduces the field: RCR with t along its own axis is a pure
    coaxial shift."""
    rng = np.random.default_rng(2)
    k, p = 11.0, 14
    C = rng.standard_normal(p * p) + 1j * rng.standard_normal(p * p)
    t = np.array([0.05, -0.03, 0.08])
    D = O.translate(O.COAX_RR, p, p, k, t, C)
    for _ in range(4):
        dx = rng.standard_normal(3) * 0.01
        f = np.sum(C * O.rbasis(p - 1, k, t + dx))
        g = np.sum(D * O.rbasis(p - 1, k, dx))
        assert abs(f - g) < 1e-12 * abs(f)


@pytest.mark.parametrize("k,t,p", [(18.3, 0.05, 16), (18.3, 0.1, 24), (91.6, 0.3, 40),
                                   (183.2, 0.5, 60)])
def test_SR_addition_theorem(k, t, p):
    rng = np.random.default_rng(int(k))
    p0 = 5
    C = np.zeros(p * p, complex)
    C[:p0 * p0] = rng.standard_normal(p0 * p0) + 1j * rng.standard_normal(p0 * p0)
    tv = rng.standard_normal(3)
    tv *= t / np.linalg.norm(tv)
    D = O.translate(O.COAX_SR, p, p, k, tv, C)
    for _ in range(4):
        dx = rng.standard_normal(3)
        dx *= 0.15 * t / np.linalg.norm(dx)
        f = np.sum(C[:p0 * p0] * O.sbasis(p0 - 1, k, tv + dx))
        g = np.sum(D * O.rbasis(p - 1, k, dx))
        assert abs(f - g) < 1e-9 * abs(f)


def _coax_blocks(mat, p):
    out = []
    for m in range(p):
        off = O.lib().fo_coax_block(p, m)
        ld = p - m
        out.append(mat[off:off + ld * ld].reshape(ld, ld).T)  # [n'][n]
    return out


@pytest.mark.parametrize("kind", [O.COAX_RR, O.COAX_SR])
def test_coax_symmetry_and_reduction(kind):
    p1, p2 = 30, 17
    a = O.coax_sparse(kind, p1, 91.6, 0.07)
    b = O.coax_sparse(kind, p2, 91.6, 0.07)
    for m, (A, B) in enumerate(zip(_coax_blocks(a, p1), _coax_blocks(b, p2))):
        n = np.arange(m, p1)
        sgn = (-1.0) ** (n[:, None] + n[None, :])
        np.testing.assert_allclose(A, sgn * A.T, rtol=1e-13, atol=0)
        sub = A[:p2 - m, :p2 - m]
        scale = np.abs(sub).max(axis=0)
        assert np.max(np.abs(sub - B) / scale) < 1e-11


def test_coax_initial_column():
    # (-1)^n' sqrt(2n'+1) j_n'(kt) / h_n'(kt) (translation.cu:580, 943-945)
    p, k, t = 12, 30.0, 0.2
    for kind in (O.COAX_RR, O.COAX_SR):
        col0 = _coax_blocks(O.coax_sparse(kind, p, k, t), p)[0][:, 0]
        n = np.arange(p)
        f = O.sph_jn(p - 1, k * t) + (1j * O.sph_yn(p - 1, k * t) if kind else 0)
        np.testing.assert_allclose(col0, (-1.0) ** n * np.sqrt(2 * n + 1) * f, rtol=1e-12)


def _mp_cases():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "coax_mp.npz"))
    for tag in sorted(k[:-5] for k in z.files if k.endswith("_meta")):
        kind, p, k, t = z[tag + "_meta"]
        yield tag, int(kind), int(p), float(k), float(t), {
            int(m): z["%s_m%d" % (tag, m)] for m in z[tag + "_ms"]}


def coax_mp_error(got_sparse, p, blocks):
    """Largest |got - exact| over |m| blocks, relative to the largest exact
    entry of each column: (err over |m| <= p/2, err over |m| > p/2)."""
    lo = hi = 0.0
    for m, B in blocks.items():
        off = O.lib().fo_coax_block(p, m)
        ld = p - m
        G = got_sparse[off:off + ld * ld].reshape(ld, ld).T
        e = float(np.max(np.abs(G - B) / np.abs(B).max(axis=0)[None, :]))
        if 2 * m <= p:
            lo = max(lo, e)
        else:
            hi = max(hi, e)
    return lo, hi


def test_coax_vs_50_digit_values():
    """Coaxial matrices against the same recurrences run in 50-digit
    arithmetic (tests/golden/gen_coax_mp.py).  Blocks |m| <= p/2 hold to
    1e-10 of their column maximum (4e-11 seen at m = 47 of (R|R) p = 94); the sectorial recurrence loses digits
    towards |m| = p-1 (1.5e-7 at m = 93 of (S|R) p = 94, kt = 91.6: a
    property of the recurrence in FP64, shared by the device code), bounded
    here at 1e-6."""
    for tag, kind, p, k, t, blocks in _mp_cases():
        lo, hi = coax_mp_error(O.coax_sparse(kind, p, k, t), p, blocks)
        assert lo < 1e-10, (tag, lo)
        assert hi < 1e-6, (tag, hi)
