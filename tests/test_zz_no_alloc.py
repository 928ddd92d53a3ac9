"""GPU check of the allocation-free forward, mbxg_ad_forward_noalloc (mbx_amd/csrc/adjoint.hip):
called through the C ABI, its energies and gradients must equal mbxg_ad_forward's (same kernels,
a caller's workspace) bit for bit, on two calls in a row with the second one smaller (the scratch
keeps the first call's contents), for water dimers and trimers.  Sizes beyond the scratch limit
are refused with an error before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu

FAMILIES = {0: ("dimers", 6), 1: ("trimers", 9)}  # MBXG_AD_MBPOL_2B, MBXG_AD_MBPOL_3B
MAX_NMERS = 8192


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _inputs(fam, n, seed):
    L = engine.lib()
    nat = FAMILIES[fam][1]
    if fam == 0:
        x = configs.water_dimers(n, seed, (2.4, 7.0))
    else:
        x = configs.water_trimers(n, seed, (2.5, 5.5))
    coef = np.zeros(1163)
    theta = np.zeros(10)
    engine._check(L.mbxg_ad_defaults(fam, coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    ncoef = 1153 if fam == 0 else 1163
    dev = "cuda:0"
    return (torch.from_numpy(np.ascontiguousarray(x.reshape(n, 3 * nat))).to(dev),
            torch.from_numpy(coef[:ncoef].copy()).to(dev), torch.from_numpy(theta).to(dev))


def _with_workspace(fam, xyz, coef, theta, stream):
    L = engine.lib()
    n = xyz.shape[0]
    nbytes = L.mbxg_ad_workspace_bytes(fam, n)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=xyz.device)
    e = torch.empty(n, dtype=torch.float64, device=xyz.device)
    g = torch.empty_like(xyz)
    engine._check(L.mbxg_ad_forward(fam, n, _p(xyz), _p(coef), _p(theta), _p(e), _p(g), _p(ws), nbytes, stream))
    return e, g


def _no_alloc(fam, xyz, coef, theta, stream):
    L = engine.lib()
    n = xyz.shape[0]
    e = torch.full((n,), float("nan"), dtype=torch.float64, device=xyz.device)
    g = torch.full_like(xyz, float("nan"))
    engine._check(L.mbxg_ad_forward_noalloc(fam, n, _p(xyz), _p(coef), _p(theta), _p(e), _p(g), stream))
    return e, g


@pytest.mark.parametrize("fam", sorted(FAMILIES))
def test_no_alloc_forward_equals_workspace_forward(fam):
    torch.cuda.init()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    name = FAMILIES[fam][0]
    for n, seed in ((MAX_NMERS - 37, 11 + fam), (1000 + 77 * fam, 23 + fam)):  # second call smaller
        print("no-alloc forward: %d %s (seed %d)" % (n, name, seed), flush=True)
        xyz, coef, theta = _inputs(fam, n, seed)
        e_ref, g_ref = _with_workspace(fam, xyz, coef, theta, stream)
        e, g = _no_alloc(fam, xyz, coef, theta, stream)
        torch.cuda.synchronize()
        e_ref, g_ref, e, g = (t.cpu().numpy() for t in (e_ref, g_ref, e, g))
        assert np.isfinite(e).all() and np.isfinite(g).all()
        assert np.count_nonzero(e) > n // 10  # the inputs exercise the polynomial
        np.testing.assert_array_equal(e, e_ref)
        np.testing.assert_array_equal(g, g_ref)


def test_no_alloc_refuses_beyond_scratch():
    L = engine.lib()
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    rc = L.mbxg_ad_forward_noalloc(1, MAX_NMERS + 1, _p(dummy), _p(dummy), _p(dummy), _p(dummy), _p(dummy), None)
    assert rc != 0
    assert b"scratch limit" in L.mbxg_last_error()
    rc = L.mbxg_ad_forward_noalloc(7, 10, _p(dummy), _p(dummy), _p(dummy), _p(dummy), _p(dummy), None)
    assert rc != 0
    # zero n-mers: nothing to do
    assert L.mbxg_ad_forward_noalloc(0, 0, None, None, None, None, None, None) == 0
