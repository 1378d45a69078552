"""fmb_gen_coax_mats_noalloc, the allocation-free variant of
fmb_gen_coax_mats (its recurrence table in the library's fixed device scratch
array), against fmb_gen_coax_mats on the same inputs: two calls in a row,
the second with fewer and lower-order matrices, both (S|R) and (R|R),
bit-identical (the same kernels on the same inputs).  Device free memory
does not change across the calls, and requests beyond the scratch size are
refused with FMB_ERR_RANGE."""
import numpy as np
import pytest
import torch

import fmbem_gpu_c_amd as F

pytestmark = pytest.mark.gpu


def _check(kind, k, t, p):
    want = F.gen_coax_mats(kind, k, t, p)
    td = torch.tensor(np.asarray(t, dtype=np.float64), device="cuda")
    got = F.gen_coax_mats_noalloc(kind, k, td, p)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    np.testing.assert_array_equal(got, want)
    return got


def test_noalloc_matches_allocating_entry_point():
    cap = F.lib().fmb_coax_scratch_bytes()
    assert cap == 8 << 20
    # first call: 12 (S|R) matrices of order 40 (12 * 40 * 79 * 16 B = 607 KB)
    t1 = np.linspace(0.05, 0.3, 12)
    _check(F.COAX_SR, 91.6, t1, 40)
    # warm the allocator-free path, then measure free memory around two calls
    # (every tensor is created before the first reading: torch's own
    # allocations would change the free memory too)
    td = torch.tensor(t1, device="cuda")
    td3 = td[:3].contiguous()
    out = torch.empty((12, F.lib().fmb_coax_size(40)), dtype=torch.complex128, device="cuda")
    out3 = torch.empty((3, F.lib().fmb_coax_size(9)), dtype=torch.complex128, device="cuda")
    F.gen_coax_mats_noalloc(F.COAX_SR, 91.6, td, 40, out=out)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    F.gen_coax_mats_noalloc(F.COAX_SR, 91.6, td, 40, out=out)
    F.gen_coax_mats_noalloc(F.COAX_RR, 18.3, td3, 9, out=out3)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    np.testing.assert_array_equal(out3.cpu().numpy(), F.gen_coax_mats(F.COAX_RR, 18.3, t1[:3], 9))
    # second call: fewer, lower-order matrices; the scratch keeps the first
    # call's larger table, which must not leak into this result
    _check(F.COAX_SR, 91.6, t1[:3], 17)
    _check(F.COAX_RR, 18.3, [0.02, 0.031], 12)
    # the reference's own truncation case (main.c:27-38) through this path
    big = _check(F.COAX_SR, 13.1, [2.3], 5)[0]
    small = _check(F.COAX_SR, 13.1, [2.3], 2)[0]
    np.testing.assert_allclose(F.reduce_coax_mat(big, 5, 2), small, rtol=1e-13, atol=1e-15)


def test_noalloc_refuses_requests_beyond_its_scratch():
    # 40 matrices of order 94: 40 * 94 * 187 * 16 B = 11.2 MB > 8 MiB
    td = torch.linspace(0.1, 0.5, 40, dtype=torch.float64, device="cuda")
    with pytest.raises(F.FmbError) as e:
        F.gen_coax_mats_noalloc(F.COAX_SR, 183.2, td, 94)
    assert e.value.code == -7 and "scratch" in str(e.value)
    # 29 of them fit (29 * 94 * 187 * 16 B = 8.16 MB): accepted and correct
    got = F.gen_coax_mats_noalloc(F.COAX_SR, 183.2, td[:29].contiguous(), 94)
    torch.cuda.synchronize()
    want = F.gen_coax_mats(F.COAX_SR, 183.2, td[:29].cpu().numpy(), 94)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
