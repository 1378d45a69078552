"""The ctypes shim checks tensor sizes before a pointer reaches the C ABI
(which sees no sizes): a vector of the wrong element count, a batch whose
rows do not match its right-hand-side counts, or more right-hand sides than
a plan holds raise FmbError(-1) on the host, before any device access.  No
GPU is needed: the plans here are stand-ins carrying only the sizes."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F


def _plan(n=10, nrhs_max=2):
    p = F.Plan.__new__(F.Plan)
    p.h, p.tree, p.n, p.nrhs_max = None, None, n, nrhs_max
    return p


def _z(*shape):
    return torch.zeros(shape, dtype=torch.complex128)


def _refused(fn, *a, **kw):
    with pytest.raises(F.FmbError) as e:
        fn(*a, **kw)
    assert e.value.code == -1
    return str(e.value)


def test_vector_length_is_checked():
    p = _plan()
    for bad in (_z(11), _z(2, 9), _z(10, 10, 1), _z(0, 10)):
        _refused(p.matvec, bad)
        _refused(p.matvec_parts, bad, 3)
        _refused(p.matvec_t, bad)
        _refused(p.gmres, bad)
        _refused(p.gmres_t, bad)
        _refused(p.field, bad, np.zeros((1, 3)))
        _refused(p.field, bad, np.zeros((1, 3)), method="fmm")
        _refused(p.adjoint_grad, bad, np.zeros((1, 3)))
        _refused(p.sweep_guess, 1.0, bad)


def test_output_and_guess_must_match_the_input():
    p = _plan()
    assert "differs" in _refused(p.matvec, _z(2, 10), _z(1, 10))
    assert "differs" in _refused(p.matvec, _z(10), _z(2, 10))
    assert "differs" in _refused(p.gmres, _z(2, 10), x0=_z(10))
    assert "(1, 11)" in _refused(p.gmres_t, _z(10), x0=_z(1, 11))


def test_right_hand_sides_above_nrhs_max():
    p = _plan(nrhs_max=2)
    assert "nrhs_max" in _refused(p.matvec, _z(3, 10))
    assert "nrhs_max" in _refused(p.gmres, _z(3, 10))
    _refused(p.rhs_plane_wave, np.zeros((3, 3)))
    _refused(p.rhs_plane_wave, np.zeros((2, 2)))


def test_refined_pairs_and_robin():
    p, q = _plan(10), _plan(12)
    assert "inner plan" in _refused(p.gmres_refined, q, _z(10))
    assert "inner plan" in _refused(p.adjoint_grad, _z(10), np.zeros((1, 3)), inner=q)
    _refused(p.gmres_refined, _plan(10, nrhs_max=1), _z(2, 10))
    _refused(p.retune, 1.0, robin=_z(9))
    _refused(F.admittance_objective, _z(11), p, np.zeros((1, 3)), [[0, 0, 1.0]])


def test_batch_rows_are_checked():
    a, b = _plan(10, 2), _plan(10, 3)
    _refused(F.matvec_batch, [a, b], [2, 3], _z(4, 10))
    _refused(F.matvec_batch, [a, b], [2, 3], _z(5, 11))
    _refused(F.matvec_batch, [a, b], [3, 3], _z(6, 10))  # above a's nrhs_max
    _refused(F.matvec_batch, [a, b], [2], _z(2, 10))
    _refused(F.matvec_batch, [a, _plan(12)], [1, 1], _z(2, 10))
    _refused(F.matvec_batch, [a], [1], _z(10), reps=0)
    _refused(F.gmres_batch, [a, b], [1, 1], _z(3, 10))
    _refused(F.gmres_batch, [a, b], [1, 1], _z(2, 10), x0=_z(3, 10))
    _refused(F.gmres_batch_refined, [a, b], [a], [1, 1], _z(2, 10))
    _refused(F.gmres_batch_refined, [a, b], [a, _plan(12)], [1, 1], _z(2, 10))


def test_field_method_is_checked():
    assert "method" in _refused(_plan().field, _z(10), np.zeros((1, 3)), method="tree")
