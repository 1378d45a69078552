"""Oracle parity of the translation modes the large workloads run.

The 327,680-triangle workload runs its coarse M2L levels at orders 94, 62
and 42, where the translation batch no longer fits in LDS and the kernels
use the global-workspace mode (FP64 from order 34, FP32 from order 48).
Here the same modes run on the 5,120-triangle sphere at depth 5 with
explicit orders and are compared with the oracle's FMM (translation.cu
3461-3759 restated in FP64) on the same tree and orders:

* far field only (fmb_matvec_parts parts = 2: P2M .. L2P, the part the
  reference's OprL computes) and the full operator (parts = 3);
* FP64 pass: 1e-9 relative (same algorithm, different summation order);
* FP32 pass: 1e-5 relative;
* orders [62, 44, 30, 20] and [94, 62, 44, 30] at levels 2-5, each level's
  tile mode asserted, so that order 94 runs in global-workspace mode in both
  precisions, order 62 likewise, and order 44 in FP64 global-workspace and
  FP32 LDS mode.

The smallest orders (1 to 5, where the even/odd rotation blocks degenerate)
and each workspace mode requested for every translation (rcr_mode) run the
same comparison.  Also the device coaxial generator on the
reference's own truncation check (main.c:27-38: (S|R), k = 13.1, t = 2.3,
order 5 reduced to 2)."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F
import oracle as O
from conftest import wavenumber

pytestmark = pytest.mark.gpu

LDS, GWS = 1, 2


@pytest.fixture(scope="module")
def sphere4():
    m = F.Mesh.icosphere(4, 0.1)
    V, Fc = m.arrays()
    return m, V, Fc, F.Tree(m, depth=5)


@pytest.mark.parametrize("orders,modes64,modes32", [
    ([62, 44, 30, 20], [GWS, GWS, LDS, LDS], [GWS, LDS, LDS, LDS]),
    ([94, 62, 44, 30], [GWS, GWS, GWS, LDS], [GWS, GWS, LDS, LDS]),
])
def test_global_workspace_translations_match_oracle(sphere4, orders, modes64, modes32):
    m, V, Fc, t = sphere4
    k = wavenumber(16000.0)
    plev = dict(zip(range(2, 6), orders))
    pl = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=2, p_rule=F.PRULE_EXPLICIT,
                p_levels=[0, 0] + orders)
    assert pl.orders()[2:6].tolist() == orders
    for l in range(2, 6):
        assert pl.tile_shape(0, l, F.FMB_FP64)[0] == modes64[l - 2], l
        assert pl.tile_shape(0, l, F.FMB_MIXED)[0] == modes32[l - 2], l
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=5, plev=plev)
    L, op = ob.orders()
    assert L == 5 and op[2:6].tolist() == orders
    rng = np.random.default_rng(61)
    n = Fc.shape[0]
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    xt = torch.tensor(x, device="cuda")
    full = ob.fmm_matvec(x)
    far = full - ob.near_matvec(x)
    for parts, want in ((2, far), (3, full)):
        y64 = pl.matvec_parts(xt, parts, prec=F.FMB_FP64).cpu().numpy()
        y32 = pl.matvec_parts(xt, parts, prec=F.FMB_MIXED).cpu().numpy()
        e64 = np.linalg.norm(y64 - want) / np.linalg.norm(want)
        e32 = np.linalg.norm(y32 - want) / np.linalg.norm(want)
        print("orders %s parts %d: FP64 %.2e, FP32 %.2e vs oracle" % (orders, parts, e64, e32),
              flush=True)
        assert e64 < 1e-9, (parts, e64)
        assert e32 < 1e-5, (parts, e32)


@pytest.mark.parametrize("orders,ldso", [([5, 3, 2, 1], 0), ([3, 3, 1, 1], 0), ([5, 3, 2, 1], 1)])
def test_low_orders_match_oracle(sphere4, orders, ldso):
    """The even/odd rotation blocks at their smallest: order 1 (degree 0
    only, no odd block), source and target orders differing by one or two
    (M2M and L2L between levels), equal orders; LDS mode, and with ldso = 1
    the operators-in-LDS mode.  Far field and full operator against the
    oracle on the same tree and orders, forward; FP64 1e-9, FP32 1e-5."""
    m, V, Fc, t = sphere4
    k = wavenumber(2000.0)
    plev = dict(zip(range(2, 6), orders))
    pl = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=3, p_rule=F.PRULE_EXPLICIT,
                p_levels=[0, 0] + orders, ldso=ldso)
    assert pl.orders()[2:6].tolist() == orders
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=5, plev=plev)
    rng = np.random.default_rng(62)
    n = Fc.shape[0]
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    xt = torch.tensor(x, device="cuda")
    full = ob.fmm_matvec(x)
    far = full - ob.near_matvec(x)
    for parts, want in ((2, far), (3, full)):
        y64 = pl.matvec_parts(xt, parts, prec=F.FMB_FP64).cpu().numpy()
        y32 = pl.matvec_parts(xt, parts, prec=F.FMB_MIXED).cpu().numpy()
        e64 = np.linalg.norm(y64 - want) / np.linalg.norm(want)
        e32 = np.linalg.norm(y32 - want) / np.linalg.norm(want)
        print("orders %s ldso %d parts %d: FP64 %.2e, FP32 %.2e vs oracle"
              % (orders, ldso, parts, e64, e32), flush=True)
        assert e64 < 1e-9, (parts, e64)
        assert e32 < 1e-5, (parts, e32)


@pytest.mark.parametrize("mode", [F.RCR_MODE_LDS, F.RCR_MODE_GWS, F.RCR_MODE_LDSP,
                                  F.RCR_MODE_LDSO])
def test_every_translation_mode_matches_oracle(sphere4, mode):
    """Each workspace mode of the matrix-core translations requested for
    every translation (rcr_mode; LDSO where batch and operators fit in LDS,
    the rule's mode elsewhere) at orders [20, 14, 10, 7], 4 kHz, with the
    mode asserted per level: far field and full operator against the oracle,
    forward.  FP64 1e-9, FP32 1e-5."""
    m, V, Fc, t = sphere4
    k = wavenumber(4000.0)
    orders = [20, 14, 10, 7]
    plev = dict(zip(range(2, 6), orders))
    pl = F.Plan(t, k, precision=F.FMB_MIXED, nrhs_max=2, p_rule=F.PRULE_EXPLICIT,
                p_levels=[0, 0] + orders, rcr_mode=mode)
    for prec in (F.FMB_FP64, F.FMB_MIXED):
        for kind in range(3):
            for l in range(2 if kind == 0 else 3, 6):
                got = pl.tile_shape(kind, l, prec)[0]
                assert got == mode or (mode == F.RCR_MODE_LDSO and got == F.RCR_MODE_LDS), \
                    (prec, kind, l, got)
    ob = O.Bem(V, Fc, k)
    ob.fmm_setup(depth=5, plev=plev)
    rng = np.random.default_rng(63)
    n = Fc.shape[0]
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    xt = torch.tensor(x, device="cuda")
    full = ob.fmm_matvec(x)
    far = full - ob.near_matvec(x)
    for parts, want in ((2, far), (3, full)):
        y64 = pl.matvec_parts(xt, parts, prec=F.FMB_FP64).cpu().numpy()
        y32 = pl.matvec_parts(xt, parts, prec=F.FMB_MIXED).cpu().numpy()
        e64 = np.linalg.norm(y64 - want) / np.linalg.norm(want)
        e32 = np.linalg.norm(y32 - want) / np.linalg.norm(want)
        print("mode %d parts %d: FP64 %.2e, FP32 %.2e vs oracle" % (mode, parts, e64, e32),
              flush=True)
        assert e64 < 1e-9, (parts, e64)
        assert e32 < 1e-5, (parts, e32)


def test_device_coax_truncation_as_main_c():
    """main.c:27-38: (S|R) at k = 13.1, t = 2.3 generated at order 5 and
    reduced to order 2 equals the order-2 matrix (device FP64 generator,
    host reduction), and both equal the oracle."""
    big = F.gen_coax_mats(F.COAX_SR, 13.1, [2.3], 5)[0]
    small = F.gen_coax_mats(F.COAX_SR, 13.1, [2.3], 2)[0]
    red = F.reduce_coax_mat(big, 5, 2)
    assert red.shape == small.shape == (F.lib().fmb_coax_size(2),)
    np.testing.assert_allclose(red, small, rtol=1e-13, atol=1e-13 * np.abs(small).max())
    want = O.coax_sparse(O.COAX_SR, 2, 13.1, 2.3)
    np.testing.assert_allclose(small, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    rot = F.gen_rot_mats([[0.4, 1.2, -0.3]], 9)[0]
    np.testing.assert_array_equal(F.reduce_rot_mat(rot, 9, 4), F.gen_rot_mats([[0.4, 1.2, -0.3]], 4)[0])
