"""CPU checks of the differentiable n-mer path's host side: the C ABI's host-only entry points
(mbxg_ad_shape, mbxg_ad_defaults, mbxg_ad_workspace_bytes), the plugin's argument checks, and
the oracle's parameterised MB-pol holders that the GPU tests difference."""
import os

import numpy as np
import pytest

import oracle
from mbx_amd import configs, engine

needs_lib = pytest.mark.skipif(not os.path.exists(engine.LIB_PATH), reason="libmbxg.so not built")


@needs_lib
def test_shapes_and_defaults():
    from mbx_amd import autograd as ad

    assert ad.shape("mbpol_2b") == (6, 1153, 10)
    assert ad.shape("mbpol_3b") == (9, 1163, 10)
    for fam in ("mbpol_2b", "mbpol_3b"):
        coef, theta = ad.defaults(fam, device="cpu")
        np.testing.assert_array_equal(coef.numpy(), oracle.table(fam).coef)
        np.testing.assert_array_equal(theta.numpy(), oracle.MBPOL_THETA[fam])
    with pytest.raises(ValueError):
        ad.shape("ch4h2o_2b")
    lib = engine.lib()
    assert lib.mbxg_ad_workspace_bytes(0, 4096) > lib.mbxg_ad_workspace_bytes(0, 64) > 0
    assert lib.mbxg_ad_workspace_bytes(7, 64) == 0
    assert lib.mbxg_ad_shape(7, None, None, None) != 0


@needs_lib
def test_argument_checks():
    import torch

    from mbx_amd import autograd as ad

    coef, theta = ad.defaults("mbpol_2b", device="cpu")
    with pytest.raises(TypeError):  # host tensors are refused before any device call
        ad.NmerEnergy.apply(torch.zeros(2, 18, dtype=torch.float64), coef, theta, "mbpol_2b")


def test_oracle_parameters_default_to_reference():
    d = configs.water_dimers(64, 1, (2.4, 7.0))
    t = configs.water_trimers(64, 2, (2.5, 5.0))
    for fam, x in (("mbpol_2b", d), ("mbpol_3b", t)):
        e0, g0 = oracle.mbpol_nmers(fam, x)
        e1, g1 = oracle.mbpol_nmers(fam, x, theta=oracle.MBPOL_THETA[fam])
        np.testing.assert_array_equal(e0, e1)
        np.testing.assert_array_equal(g0, g1)


@pytest.mark.parametrize("fam", ["mbpol_2b", "mbpol_3b"])
def test_oracle_gradients_with_parameters(fam):
    """With perturbed nonlinear parameters (the M-site gammas included), the oracle's analytic
    coordinate gradient still matches five-point differences of its energy."""
    rng = np.random.default_rng(5)
    x = (configs.water_dimers(4, 3, (2.6, 6.0)) if fam == "mbpol_2b" else configs.water_trimers(4, 3, (2.6, 4.4)))
    theta = oracle.MBPOL_THETA[fam] * rng.uniform(0.9, 1.1, 10)
    e, g = oracle.mbpol_nmers(fam, x, theta=theta)
    h = 1e-4
    for i in range(len(x)):
        for j in range(x.shape[1]):
            def en(d):
                y = x[i:i + 1].copy()
                y[0, j] += d
                return oracle.mbpol_nmers(fam, y, theta=theta, grad=False)[0][0]
            fd = (-en(2 * h) + 8 * en(h) - 8 * en(-h) + en(-2 * h)) / (12 * h)
            assert fd == pytest.approx(g[i, j], rel=1e-6, abs=1e-7)
