"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same inputs.  Tolerances: tree/index work bit-exact; FP64 operators 1e-11
relative; FP64 FMM matvec 1e-9 relative to the oracle FMM (same algorithm,
different summation order); mixed path 1e-5 relative to FP64."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber

pytestmark = pytest.mark.gpu


def _oracle_tree(V, Fc, s=5, depth=0):
    return O.Tree(V, Fc, s=s, depth=depth)


def test_tree_matches_oracle(sphere_l3):
    m, V, Fc = sphere_l3
    t = F.Tree(m, s=5)
    ot = _oracle_tree(V, Fc, s=5)
    lmin, lmax, D, pmin = t.info()
    assert lmax == ot.lmax
    oD, opmin = ot.geom()
    assert D == pytest.approx(oD, rel=1e-15)
    np.testing.assert_allclose(pmin, opmin, rtol=0, atol=1e-15)
    for l in range(2, lmax + 1):
        np.testing.assert_array_equal(t.level(l), ot.keys(l))
        np.testing.assert_array_equal(t.num_sr(l), ot.num_sr(l))
        keys = t.level(l)
        tg, sr, off = t.pairs(l)
        got = set(zip(keys[tg].tolist(), keys[sr].tolist()))
        want = set()
        kset = set(keys.tolist())
        for b in keys:
            for s in O.I4(int(b), l):
                if int(s) in kset:
                    want.add((int(b), int(s)))
        assert got == want
    box, perm = t.elem_box()
    np.testing.assert_array_equal(perm, ot.perm())


@pytest.mark.parametrize("p", [8, 17, 40, 94])
def test_rot_mats_match_oracle(p):
    rng = np.random.default_rng(p)
    ang = np.column_stack([rng.uniform(-np.pi, np.pi, 5), rng.uniform(0, np.pi, 5),
                           rng.uniform(-np.pi, np.pi, 5)])
    got = F.gen_rot_mats(ang, p)
    for i in range(5):
        want = O.rot_sparse(p, *ang[i])
        err = np.max(np.abs(got[i] - want)) / np.max(np.abs(want))
        assert err < 1e-11, (i, err)


def _coax_cols(mat, p):
    """|m|-blocks of a sparse coaxial matrix as a list of (p-m)x(p-m) arrays [row n', col n]."""
    out = []
    for m in range(p):
        off = O.lib().fo_coax_block(p, m)
        ld = p - m
        out.append(mat[off:off + ld * ld].reshape(ld, ld).T)
    return out


@pytest.mark.parametrize("kind,p,k,t", [(O.COAX_RR, 12, 18.3, 0.02), (O.COAX_SR, 16, 18.3, 0.1),
                                        (O.COAX_SR, 30, 91.6, 0.3), (O.COAX_RR, 30, 91.6, 0.05)])
def test_coax_mats_match_oracle(kind, p, k, t):
    """Entries agree to 1e-8 of the largest entry of their column (the
    columns span many decades; the translation uses whole columns).  Both
    sides run the same FP64 recurrences, the device with fused multiply-adds
    and the host without; the sectorial and advancing recurrences amplify
    those last-bit differences (a few 1e-9 at kt = 41, p = 30), still far
    below the FP32 pass's rounding."""
    got = F.gen_coax_mats(kind, k, [t, 1.5 * t], p)
    for i, tt in enumerate([t, 1.5 * t]):
        want = O.coax_sparse(kind, p, k, tt)
        err = 0.0
        for bg, bw in zip(_coax_cols(got[i], p), _coax_cols(want, p)):
            scale = np.abs(bw).max(axis=0)
            err = max(err, float(np.max(np.abs(bg - bw) / scale[None, :])))
        assert err < 1e-8, (i, tt, err)


def test_coax_mats_vs_50_digit_values():
    """Device (R|R), (S|R) matrices up to p = 94 against the recurrences run
    in 50-digit arithmetic (tests/golden/coax_mp.npz), with the oracle's
    bounds: 1e-10 of the column maximum for |m| <= p/2, 1e-6 towards
    |m| = p-1 where the FP64 sectorial recurrence itself loses digits."""
    from test_oracle_ops import _mp_cases, coax_mp_error
    for tag, kind, p, k, t, blocks in _mp_cases():
        got = F.gen_coax_mats(kind, k, [t], p)[0]
        lo, hi = coax_mp_error(got, p, blocks)
        print("coax %s: |m|<=p/2 %.2e, tail %.2e" % (tag, lo, hi), flush=True)
        assert lo < 1e-10, (tag, lo)
        assert hi < 1e-6, (tag, hi)


def _config1_pair(sphere_l3, precision=F.FMB_FP64, nrhs=1):
    m, V, Fc = sphere_l3
    k = wavenumber(1000.0)
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, precision=precision, nrhs_max=nrhs)
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(s=5)
    return t, pl, ob, k


def test_matvec_fp64_matches_oracle(sphere_l3):
    t, pl, ob, k = _config1_pair(sphere_l3)
    L, op = ob.orders()
    np.testing.assert_array_equal(pl.orders()[2:L + 1], op[2:L + 1])
    rng = np.random.default_rng(7)
    n = ob.n
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = pl.matvec(torch.tensor(x, device="cuda")).cpu().numpy()
    yo = ob.fmm_matvec(x)[0]
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    assert err < 1e-9, err
    yd = ob.dense_matvec(x)[0]
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 1e-4


def test_matvec_mixed_vs_fp64(sphere_l3):
    t, pl, ob, k = _config1_pair(sphere_l3, precision=F.FMB_MIXED, nrhs=2)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((2, ob.n)) + 1j * rng.standard_normal((2, ob.n))
    xt = torch.tensor(x, device="cuda")
    y64 = pl.matvec(xt, prec=F.FMB_FP64).cpu().numpy()
    y32 = pl.matvec(xt, prec=F.FMB_MIXED).cpu().numpy()
    yo = ob.fmm_matvec(x)
    assert np.linalg.norm(y64 - yo) / np.linalg.norm(yo) < 1e-9
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) < 1e-5


@pytest.mark.parametrize("precision", [F.FMB_FP64, F.FMB_MIXED])
def test_gmres_config1(sphere_l3, precision):
    t, pl, ob, k = _config1_pair(sphere_l3, precision=precision)
    b = pl.rhs_plane_wave([0.0, 0.0, 1.0])
    np.testing.assert_allclose(b[0].cpu().numpy(), ob.rhs_plane([0, 0, 1.0]), rtol=1e-13, atol=1e-13)
    x, info = pl.gmres(b, restart=30, tol=1e-6)
    assert info.relres[0] <= 1e-6
    u = x[0].cpu().numpy()
    xo, it, rr, rc = ob.gmres(ob.rhs_plane([0, 0, 1.0]), use_fmm=True)
    assert rc == 0
    assert np.linalg.norm(u - xo) / np.linalg.norm(xo) < 1e-5
    c, nrm, area = ob.centroids()
    ct = c[:, 2] / np.linalg.norm(c, axis=1)
    mie = O.mie_rigid(k * 0.1, ct)
    assert np.linalg.norm(u - mie) / np.linalg.norm(mie) < 0.02


def _far_field(pl, x, prec):
    return pl.matvec_parts(x, 2, prec=prec).cpu().numpy()


@pytest.mark.parametrize("freq,orders", [(4000.0, [0, 0, 6, 12, 20, 17]),
                                         (4000.0, [0, 0, 33, 28, 24, 8]),
                                         (4000.0, [0, 0, 47, 40, 31, 25]),
                                         (16000.0, [0, 0, 62, 44, 30, 20])])
def test_rcr_matrix_cores_match_valu(freq, orders):
    """The matrix-core translations (rcr_kernels.h: M2L, M2M, L2L) against the
    one-translation-per-workgroup VALU kernels on the same plan: far field
    only (P2M .. L2P), three right-hand sides so batches straddle box pairs,
    level orders rising and falling so M2M / L2L run with pin != pout, and
    orders chosen to run every LDS tile shape and the global-workspace mode
    of both precisions.  FP64: 1e-12 relative (summation order only); FP32
    pass: 1e-5 relative to the FP32 VALU pass on the same operators."""
    m = F.Mesh.icosphere(4, 0.1)
    k = wavenumber(freq)
    t = F.Tree(m, depth=5)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=3, p_rule=F.PRULE_EXPLICIT, p_levels=orders)
    pv = F.Plan(t, k, rcr_path=F.RCR_VALU, **kw)
    pm = F.Plan(t, k, rcr_path=F.RCR_AUTO, **kw)
    rng = np.random.default_rng(11)
    n = m.arrays()[1].shape[0]
    x = torch.tensor(rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n)), device="cuda")
    yv = _far_field(pv, x, F.FMB_FP64)
    ym = _far_field(pm, x, F.FMB_FP64)
    e64 = np.linalg.norm(ym - yv) / np.linalg.norm(yv)
    yv32 = _far_field(pv, x, F.FMB_MIXED)
    ym32 = _far_field(pm, x, F.FMB_MIXED)
    e32 = np.linalg.norm(ym32 - yv32) / np.linalg.norm(yv32)
    e32_64 = np.linalg.norm(ym32 - yv) / np.linalg.norm(yv)
    print("f %g orders %s: fp64 %.2e fp32 %.2e (vs fp64 %.2e)"
          % (freq, orders[2:], e64, e32, e32_64), flush=True)
    assert e64 < 1e-12
    assert e32 < 1e-5


@pytest.mark.parametrize("robin", [False, True])
def test_field_points_match_oracle(sphere_l3, robin):
    """Pressure at 32 points on an r = 0.5 m circle plus 8 points just off the
    surface (close-panel rule) from random surface data, device (FP64
    direct sum) against the oracle's representation formula; rigid and
    Robin (sigma = i k Y) surfaces, incident plane waves added."""
    m, V, Fc = sphere_l3
    k = wavenumber(1000.0)
    n = Fc.shape[0]
    rng = np.random.default_rng(21)
    sig = None
    if robin:
        Y = 0.01 * (1 + 0.5 * rng.uniform(-1, 1, n)) + 0.005j * rng.uniform(-1, 1, n)
        sig = 1j * k * Y
    t = F.Tree(m, s=5)
    pl = F.Plan(t, k, robin=None if sig is None else torch.tensor(sig, device="cuda"), nrhs_max=2)
    ob = O.Bem(V, Fc, k, sigma=sig)
    th = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    pts = np.stack([0.5 * np.sin(th), np.zeros_like(th), 0.5 * np.cos(th)], 1)
    c = V[Fc].mean(1)[:8]
    pts = np.vstack([pts, c * 1.02])
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    dirs = [[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]]
    got = pl.field(torch.tensor(x, device="cuda"), pts, dirs).cpu().numpy()
    want = ob.field(x, pts, dirs)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("field robin=%s rel err %.2e" % (robin, err), flush=True)
    assert err < 1e-12


def test_autotuned_shapes_give_the_same_operator():
    """Tile shapes chosen by timing (autotune) against the fixed rule on the
    5,120-triangle sphere with orders that make the rule pick both LDS and
    global-workspace batches: the same operator (FP64 pass 1e-13, FP32 pass
    1e-5 relative; summation order only), every shape reported with its
    time, and a second plan with the same shapes served from the cache."""
    m = F.Mesh.icosphere(4, 0.1)
    k = wavenumber(4000.0)
    t = F.Tree(m, depth=5)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=3, p_rule=F.PRULE_EXPLICIT,
              p_levels=[0, 0, 62, 44, 30, 20])
    ref = F.Plan(t, k, **kw)
    tuned = F.Plan(t, k, autotune=1, **kw)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(31)
    x = torch.tensor(rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n)), device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-13), (F.FMB_MIXED, 1e-5)):
        a = ref.matvec(x, prec=prec).cpu().numpy()
        b = tuned.matvec(x, prec=prec).cpu().numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < tol
        for kind in range(3):
            for l in range(2 if kind == 0 else 3, 6):
                mode, nt, mt, nw, ms = tuned.tile_shape(kind, l, prec)
                assert mode in (1, 2, 3, 4) and nt in (1, 2) and mt in (2, 3, 4, 6) and ms > 0
                assert nw in (4, 8)
                assert ref.tile_shape(kind, l, prec)[3:] == (4, 0.0)
    again = F.Plan(t, k, autotune=1, **kw)
    for kind in range(3):
        for l in range(2 if kind == 0 else 3, 6):
            assert again.tile_shape(kind, l, F.FMB_MIXED)[:4] == tuned.tile_shape(kind, l, F.FMB_MIXED)[:4]


def test_config1_depth3_order8_vs_oracle_and_mie(sphere_l3):
    """The plain-method configuration as specified: 1,280 flat triangles,
    r = 0.1 m, 1 kHz, plane wave along +z, Burton-Miller (alpha = i/k),
    7-point rule with analytic singular self terms, octree depth 3 with
    order 8 at the leaves (16 at level 2, truncNum's), GMRES(30), FP64.
    The solution against the oracle's FMM solve of the same operator
    (1e-8 at tolerance 1e-10) and the surface pressure against the Mie
    series at 64 surface points (2%)."""
    m, V, Fc = sphere_l3
    n = Fc.shape[0]
    k = wavenumber(1000.0)
    t = F.Tree(m, depth=3)
    pl = F.Plan(t, k, p_rule=F.PRULE_EXPLICIT, p_levels=[0, 0, 16, 8])
    assert pl.orders()[2:4].tolist() == [16, 8]
    b = pl.rhs_plane_wave([0.0, 0.0, 1.0])
    x, info = pl.gmres(b, restart=30, tol=1e-10)
    u = x[0].cpu().numpy()
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=3, plev={2: 16, 3: 8})
    xo, it, rr, rc = ob.gmres(ob.rhs_plane([0, 0, 1.0]), use_fmm=True, m=30, tol=1e-10)
    assert rc == 0
    assert np.linalg.norm(u - xo) / np.linalg.norm(xo) < 1e-8
    idx = np.linspace(0, n - 1, 64).astype(int)
    c, _, _ = ob.centroids()
    ct = c[idx, 2] / np.linalg.norm(c[idx], axis=1)
    mie = O.mie_rigid(k * 0.1, ct)
    assert np.linalg.norm(u[idx] - mie) / np.linalg.norm(mie) < 0.02


@pytest.mark.parametrize("freq", [4000.0, 12000.0])
def test_concurrent_levels_match_serial(freq):
    """The coarse M2L levels and the L2L chain on a second stream beside
    the M2L of levels L-2 and L-1, then the finest M2L alone (the default),
    against all levels in order on one stream: the 20,480-triangle sphere at
    depth 6, FP64 pass 1e-13, FP32 pass 1e-6 relative (summation order
    only), twice in a row."""
    m = F.Mesh.icosphere(5, 0.1)
    t = F.Tree(m, depth=6)
    k = wavenumber(freq)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=2)
    ser = F.Plan(t, k, serial_levels=1, **kw)
    con = F.Plan(t, k, **kw)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(41)
    x = torch.tensor(rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)), device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-13), (F.FMB_MIXED, 1e-6)):
        a = ser.matvec(x, prec=prec).cpu().numpy()
        for _ in range(2):
            b = con.matvec(x, prec=prec).cpu().numpy()
            assert np.linalg.norm(a - b) / np.linalg.norm(a) < tol


@pytest.mark.parametrize("freq", [4000.0, 12000.0])
def test_eight_wave_shapes_match_four_wave(freq):
    """Every matrix-core translation with 8 waves per workgroup (the fixed
    tile rule otherwise: LDS batches at low orders, global-workspace batches
    at high ones) against 4 waves: the 20,480-triangle sphere at depth 6,
    forward and transposed passes, FP64 1e-13 and FP32 1e-6 relative (the
    unit schedule changes only the order of the target sums)."""
    m = F.Mesh.icosphere(5, 0.1)
    t = F.Tree(m, depth=6)
    k = wavenumber(freq)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=2, serial_levels=1)
    p4 = F.Plan(t, k, **kw)
    p8 = F.Plan(t, k, waves=8, **kw)
    _, L, _, _ = t.info()
    for kind in range(3):
        for l in range(2 if kind == 0 else 3, L + 1):
            for prec in (F.FMB_FP64, F.FMB_MIXED):
                assert p4.tile_shape(kind, l, prec)[3] == 4
                assert p8.tile_shape(kind, l, prec)[3] == 8
    modes = {p8.tile_shape(0, l, F.FMB_FP64)[0] for l in range(2, L + 1)}
    print("f", freq, "orders", p8.orders()[2:L + 1].tolist(), "M2L modes", sorted(modes), flush=True)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(43)
    x = torch.tensor(rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)), device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-13), (F.FMB_MIXED, 1e-6)):
        a = p4.matvec(x, prec=prec).cpu().numpy()
        b = p8.matvec(x, prec=prec).cpu().numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < tol, (prec, np.linalg.norm(a - b) / np.linalg.norm(a))
    a = p4.matvec_t(x).cpu().numpy()
    b = p8.matvec_t(x).cpu().numpy()
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-13


def test_waves_parameter_is_checked():
    m = F.Mesh.icosphere(2, 0.1)
    t = F.Tree(m, depth=2)
    with pytest.raises(F.FmbError):
        F.Plan(t, wavenumber(1000.0), waves=6)


@pytest.mark.parametrize("freq", [4000.0, 12000.0])
def test_split_bf16_rotations(freq):
    """Split-BF16 rotations in the FP32 pass (LDS-mode translations; the
    global-workspace ones stay FP32): the 20,480-triangle sphere at depth 6.
    Each operand is hi + lo in bf16 and a product keeps three terms, so the
    pass meets the FP64 one to about 1e-5 (bound 1e-4) and the FP32 pass to
    the same; the transposed pass matches its transpose identity."""
    m = F.Mesh.icosphere(5, 0.1)
    t = F.Tree(m, depth=6)
    k = wavenumber(freq)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=2)
    p32 = F.Plan(t, k, **kw)
    pbf = F.Plan(t, k, split_bf16=1, **kw)
    _, L, _, _ = t.info()
    print("f", freq, "orders", pbf.orders()[2:L + 1].tolist(),
          "M2L modes", [pbf.tile_shape(0, l, F.FMB_MIXED)[0] for l in range(2, L + 1)], flush=True)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(47)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    y = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    xt, yt = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
    a64 = p32.matvec(xt, prec=F.FMB_FP64).cpu().numpy()
    a32 = p32.matvec(xt, prec=F.FMB_MIXED).cpu().numpy()
    abf = pbf.matvec(xt, prec=F.FMB_MIXED).cpu().numpy()
    e64 = np.linalg.norm(abf - a64) / np.linalg.norm(a64)
    e32 = np.linalg.norm(abf - a32) / np.linalg.norm(a32)
    print("split-BF16 vs FP64 %.2e, vs FP32 %.2e (FP32 vs FP64 %.2e)"
          % (e64, e32, np.linalg.norm(a32 - a64) / np.linalg.norm(a64)), flush=True)
    assert e64 < 1e-4 and e32 < 1e-4
    assert e64 > 1e-8  # the split path really ran
    aty = pbf.matvec_t(yt, prec=F.FMB_MIXED).cpu().numpy()
    for r in range(2):
        lhs, rhs = np.sum(y[r] * abf[r]), np.sum(aty[r] * x[r])
        assert abs(lhs - rhs) < 1e-4 * np.linalg.norm(y[r]) * np.linalg.norm(abf[r])
    with pytest.raises(F.FmbError):
        F.Plan(t, k, split_bf16=1, precision=F.FMB_FP64)


def test_split_bf16_retune_and_graph():
    """A split-BF16 plan retuned to a higher frequency (orders grow, so the
    rotation tables and their split copies are rebuilt) against a fresh
    split-BF16 plan there (1e-6 relative: same operators, order of the
    target sums only), and a graph-captured batch of two split-BF16 plans
    replayed three times against their eager matvecs."""
    m = F.Mesh.icosphere(4, 0.1)
    t = F.Tree(m, depth=5)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=2, split_bf16=1)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(53)
    x = torch.tensor(rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)), device="cuda")
    pl = F.Plan(t, wavenumber(2000.0), **kw)
    k = wavenumber(9000.0)
    pl.retune(k)
    fresh = F.Plan(t, k, **kw)
    assert any(fresh.tile_shape_bf(0, l, F.FMB_MIXED)[4] for l in range(2, 6))
    a = pl.matvec(x, prec=F.FMB_MIXED).cpu().numpy()
    b = fresh.matvec(x, prec=F.FMB_MIXED).cpu().numpy()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    p2 = F.Plan(t, wavenumber(5000.0), **kw)
    xb = torch.cat([x, x], 0).contiguous()
    want = torch.cat([fresh.matvec(x, prec=F.FMB_MIXED), p2.matvec(x, prec=F.FMB_MIXED)], 0).cpu().numpy()
    got = F.matvec_batch([fresh, p2], [2, 2], xb, graph=True, prec=F.FMB_MIXED, reps=3).cpu().numpy()
    for rep, y in enumerate(got):
        err = np.linalg.norm(y - want) / np.linalg.norm(want)
        print("graph replay", rep, "rel err %.2e" % err, flush=True)
        assert err < 1e-6


@pytest.mark.parametrize("freq", [4000.0, 12000.0])
def test_operators_in_lds_match(freq):
    """Translations with the bucket's operator blocks held in LDS on a
    persistent grid (ldso = 1: FP32 orders up to about 20, FP64 up to 16)
    against the default tile shapes on the 20,480-triangle sphere at depth
    6, forward and transposed passes: FP64 1e-13, FP32 1e-6 relative (the
    same arithmetic, summation order of the target atomics only)."""
    m = F.Mesh.icosphere(5, 0.1)
    t = F.Tree(m, depth=6)
    k = wavenumber(freq)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=3)
    ref = F.Plan(t, k, **kw)
    opl = F.Plan(t, k, ldso=1, **kw)
    _, L, _, _ = t.info()
    modes = {prec: [opl.tile_shape(kind, l, prec)[0] for kind in range(3)
                    for l in range(2 if kind == 0 else 3, L + 1)]
             for prec in (F.FMB_FP64, F.FMB_MIXED)}
    print("f", freq, "orders", opl.orders()[2:L + 1].tolist(), "modes", modes, flush=True)
    assert 4 in modes[F.FMB_MIXED]
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(67)
    x = torch.tensor(rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n)), device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-13), (F.FMB_MIXED, 1e-6)):
        for rep in range(2):
            a = ref.matvec(x, prec=prec).cpu().numpy()
            b = opl.matvec(x, prec=prec).cpu().numpy()
            e = np.linalg.norm(a - b) / np.linalg.norm(a)
            assert e < tol, (prec, rep, e)
        at = ref.matvec_t(x, prec=prec).cpu().numpy()
        bt = opl.matvec_t(x, prec=prec).cpu().numpy()
        assert np.linalg.norm(at - bt) / np.linalg.norm(at) < tol


def test_operators_in_lds_match_oracle():
    """The LDS-operator translations against the oracle FMM: the
    5,120-triangle sphere at depth 5, 4 kHz, orders [20, 18, 16, 14] so that
    every level runs that mode in FP32 (and levels up to order 16 in FP64);
    FP64 pass 1e-9, FP32 pass 1e-5."""
    m = F.Mesh.icosphere(4, 0.1)
    V, Fc = m.arrays()
    t = F.Tree(m, depth=5)
    k = wavenumber(4000.0)
    orders = [20, 18, 16, 14]
    pl = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=2, ldso=1, p_rule=F.PRULE_EXPLICIT,
                p_levels=[0, 0] + orders)
    assert all(pl.tile_shape(0, l, F.FMB_MIXED)[0] == 4 for l in range(2, 6))
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=5, plev=dict(zip(range(2, 6), orders)))
    rng = np.random.default_rng(71)
    x = rng.standard_normal((2, Fc.shape[0])) + 1j * rng.standard_normal((2, Fc.shape[0]))
    want = ob.fmm_matvec(x)
    xt = torch.tensor(x, device="cuda")
    e64 = np.linalg.norm(pl.matvec(xt, prec=F.FMB_FP64).cpu().numpy() - want) / np.linalg.norm(want)
    e32 = np.linalg.norm(pl.matvec(xt, prec=F.FMB_MIXED).cpu().numpy() - want) / np.linalg.norm(want)
    print("LDS operators vs oracle: FP64 %.2e, FP32 %.2e" % (e64, e32), flush=True)
    assert e64 < 1e-9 and e32 < 1e-5


@pytest.mark.parametrize("tile", [1, 7, 64])
def test_tile_major_m2l_order_matches(tile):
    """M2L work in tile-major order on every level (the pairs of `tile`
    consecutive target boxes for every offset, each XCD taking a contiguous
    range; m2l_tile_min_mb = 0 lifts the size rule) against the offset-major
    order on the 20,480-triangle sphere at depth 6, 12 kHz: the same
    translations in another order, so FP64 1e-13 and FP32 1e-6 relative,
    forward and transposed, twice in a row."""
    m = F.Mesh.icosphere(5, 0.1)
    t = F.Tree(m, depth=6)
    k = wavenumber(12000.0)
    kw = dict(precision=F.FMB_MIXED, nrhs_max=3)
    ref = F.Plan(t, k, m2l_tile=0, **kw)
    til = F.Plan(t, k, m2l_tile=tile, m2l_tile_min_mb=0, **kw)
    n = m.arrays()[1].shape[0]
    rng = np.random.default_rng(73)
    x = torch.tensor(rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n)), device="cuda")
    for prec, tol in ((F.FMB_FP64, 1e-13), (F.FMB_MIXED, 1e-6)):
        a = ref.matvec(x, prec=prec).cpu().numpy()
        for _ in range(2):
            b = til.matvec(x, prec=prec).cpu().numpy()
            assert np.linalg.norm(a - b) / np.linalg.norm(a) < tol
        at = ref.matvec_t(x, prec=prec).cpu().numpy()
        bt = til.matvec_t(x, prec=prec).cpu().numpy()
        assert np.linalg.norm(at - bt) / np.linalg.norm(at) < tol
