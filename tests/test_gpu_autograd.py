"""GPU parity of the differentiable n-mer path: mbx_amd.autograd.NmerEnergy over the C ABI
(mbxg_ad_forward / mbxg_ad_backward, mbx_amd/csrc/adjoint.hip).

Pins:
  * forward E and dE/dx against the oracle (which follows x2b-v9x.cpp / x3b-v2x.cpp) and, where
    oracle/_ref is built, against the compiled reference's own analytic gradients (ref_x2b,
    ref_x3b); with perturbed coefficients and nonlinear parameters against the oracle given the
    same values.  Energies to 1e-10 relative (or 2e-16 of the sum of |monomial terms| for
    strongly cancelling n-mers), gradients to 1e-8 kcal/mol/A.
  * backward: torch.autograd.gradcheck (central differences of the GPU forward, every input
    entry, including all coefficients) on a few n-mers; and on the full workload of 4096 dimers
    and 4096 trimers, an energy + force loss against perturbed targets, differentiated along
    random directions in coordinates and coefficients and along each nonlinear parameter, with
    five-point central differences (h = 1e-5) of the loss evaluated by the oracle on the CPU.  The
    loss is quadratic in the coefficients, so that difference is exact up to rounding.
  * Euler's identity: sum_k a_k dE/da_k = E (energy linear in the coefficients).
  * dE/dtheta has no counterpart in the reference; its pins are the central differences.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from mbx_amd import autograd as ad
from mbx_amd import configs

pytestmark = pytest.mark.gpu

FAMS = ("mbpol_2b", "mbpol_3b")
NAT = {"mbpol_2b": 6, "mbpol_3b": 9}


def sample(family, n, seed):
    if family == "mbpol_2b":
        return configs.water_dimers(n, seed, (2.4, 7.0))
    return configs.water_trimers(n, seed, (2.5, 5.5))


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def forward(family, x, coef=None, theta=None):
    c0, t0 = ad.defaults(family)
    coef = c0 if coef is None else dev(coef)
    theta = t0 if theta is None else dev(theta)
    e, g = ad.NmerEnergy.apply(dev(x), coef, theta, family)
    torch.cuda.synchronize()
    return e.cpu().numpy(), g.cpu().numpy()


def check_energies(e, e_ref, scale):
    err = np.abs(e - e_ref)
    bar = np.maximum(1e-10 * np.abs(e_ref), 2e-16 * scale)
    assert np.all(err <= bar), (err / np.maximum(bar, 1e-300)).max()


@pytest.mark.parametrize("family", FAMS)
def test_defaults_match_oracle(family):
    coef, theta = ad.defaults(family)
    np.testing.assert_array_equal(coef.cpu().numpy(), oracle.table(family).coef)
    np.testing.assert_array_equal(theta.cpu().numpy(), oracle.MBPOL_THETA[family])


@pytest.mark.parametrize("family", FAMS)
def test_forward_matches_oracle(family):
    x = sample(family, 700, 11)
    e, g = forward(family, x)
    e_ref, g_ref = oracle.mbpol_nmers(family, x)
    scale, _ = oracle.mbpol_nmers(family, x, a=np.abs(oracle.table(family).coef), grad=False)
    assert np.count_nonzero(e_ref) > 200
    check_energies(e, e_ref, np.abs(scale))
    np.testing.assert_allclose(g, g_ref, rtol=0, atol=1e-8)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("family", FAMS)
def test_forward_matches_reference_gradients(family):
    """The compiled reference's x2b_v9x::eval / x3b_v2x::eval, one n-mer per call."""
    x = sample(family, 300, 12)
    e, g = forward(family, x)
    L = oracle.ref()
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: a.ctypes.data_as(dp)
    for i in range(len(x)):
        w = [np.ascontiguousarray(x[i, 9 * m:9 * m + 9]) for m in range(NAT[family] // 3)]
        gr = [np.zeros(9) for _ in w]
        vir = np.zeros(9)
        if family == "mbpol_2b":
            er = L.ref_x2b(p(w[0]), p(w[1]), 1, p(gr[0]), p(gr[1]), p(vir))
        else:
            er = L.ref_x3b(p(w[0]), p(w[1]), p(w[2]), 1, p(gr[0]), p(gr[1]), p(gr[2]), p(vir))
        assert e[i] == pytest.approx(er, rel=1e-10, abs=1e-11)
        np.testing.assert_allclose(g[i], np.concatenate(gr), rtol=0, atol=1e-8)


@pytest.mark.parametrize("family", FAMS)
def test_forward_custom_parameters(family):
    rng = np.random.default_rng(13)
    x = sample(family, 300, 13)
    coef = oracle.table(family).coef * rng.uniform(0.5, 1.5, oracle.table(family).ncoef)
    theta = oracle.MBPOL_THETA[family] * rng.uniform(0.95, 1.05, 10)
    e, g = forward(family, x, coef, theta)
    e_ref, g_ref = oracle.mbpol_nmers(family, x, a=coef, theta=theta)
    scale, _ = oracle.mbpol_nmers(family, x, a=np.abs(coef), theta=theta, grad=False)
    check_energies(e, e_ref, np.abs(scale))
    np.testing.assert_allclose(g, g_ref, rtol=0, atol=1e-8)


def test_zero_and_ragged_sizes():
    for family in FAMS:
        for n in (0, 1, 63, 65):
            x = sample(family, max(n, 1), 14)[:n]
            coef, theta = ad.defaults(family)
            xt = dev(x.reshape(n, 3 * NAT[family])).requires_grad_()
            e, g = ad.NmerEnergy.apply(xt, coef.requires_grad_(), theta.requires_grad_(), family)
            (e.sum() + (g * g).sum()).backward()
            assert xt.grad.shape == xt.shape and torch.isfinite(xt.grad).all()
            if n == 0:
                assert float(coef.grad.abs().sum()) == 0.0 and float(theta.grad.abs().sum()) == 0.0
            else:
                e_ref, _ = oracle.mbpol_nmers(family, x)
                scale, _ = oracle.mbpol_nmers(family, x, a=np.abs(oracle.table(family).coef), grad=False)
                check_energies(e.detach().cpu().numpy(), e_ref, np.abs(scale))


def close_nmers(family, n, seed):
    """n-mers inside the polynomial range, some in the switching region."""
    if family == "mbpol_2b":
        return configs.water_dimers(n, seed, (2.6, 6.3))
    return configs.water_trimers(n, seed, (2.6, 4.4))


@pytest.mark.parametrize("family", FAMS)
def test_gradcheck(family):
    """Every entry of the Jacobian of (E, dE/dx) by central differences of the GPU forward."""
    x = dev(close_nmers(family, 3, 15)).requires_grad_()
    coef, theta = ad.defaults(family)
    coef.requires_grad_()
    theta.requires_grad_()
    f = lambda xx, cc, tt: ad.NmerEnergy.apply(xx, cc, tt, family)
    assert torch.autograd.gradcheck(f, (x, coef, theta), eps=1e-5, atol=1e-6, rtol=1e-5, nondet_tol=1e-12)


@pytest.mark.parametrize("family", FAMS)
def test_euler_identity(family):
    x = dev(close_nmers(family, 200, 16))
    coef, theta = ad.defaults(family)
    coef.requires_grad_()
    e, _ = ad.NmerEnergy.apply(x, coef, theta, family)
    (da,) = torch.autograd.grad(e.sum(), coef)
    lhs = float((da * coef.detach()).sum())
    assert lhs == pytest.approx(float(e.detach().sum()), rel=1e-10)


class LossC4:
    """Energy + force MSE over the dimers and trimers of the workload, against targets perturbed
    by N(0, 0.1) kcal/mol and N(0, 0.1) kcal/mol/A from the reference parameters' values."""

    def __init__(self, nd, nt, seed=4096):
        self.x = {"mbpol_2b": configs.water_dimers(nd, seed), "mbpol_3b": configs.water_trimers(nt, seed + 1)}
        rng = np.random.default_rng(seed + 2)
        self.ref = {}
        for fam, x in self.x.items():
            e, g = oracle.mbpol_nmers(fam, x)
            self.ref[fam] = (e + rng.normal(0, 0.1, e.shape), -g + rng.normal(0, 0.1, g.shape))

    def cpu(self, params):
        """The loss by the oracle; params = {family: (coef, theta, x)}."""
        L = 0.0
        for fam, (coef, theta, x) in params.items():
            e, g = oracle.mbpol_nmers(fam, x, a=coef, theta=theta)
            er, fr = self.ref[fam]
            L += np.mean((e - er) ** 2) + np.mean((-g - fr) ** 2)
        return L

    def gpu(self, params):
        L = 0.0
        for fam, (coef, theta, x) in params.items():
            e, f = ad.energy_forces(x, coef, theta, fam)
            er, fr = (dev(v) for v in self.ref[fam])
            L = L + ad.energy_force_loss(e, f, er, fr)
        return L


def test_c4_loss_gradients_against_differences():
    """Full workload: 4096 dimers (r_OO ~ U(2.4, 9) A) and 4096 trimers (O-O ~ U(2.5, 7) A)."""
    c4 = LossC4(4096, 4096)
    base = {fam: (oracle.table(fam).coef.copy(), oracle.MBPOL_THETA[fam].copy(), c4.x[fam]) for fam in FAMS}
    tens = {fam: tuple(dev(v).requires_grad_() for v in base[fam]) for fam in FAMS}
    L = c4.gpu(tens)
    L.backward()
    L0 = c4.cpu(base)
    assert float(L.detach()) == pytest.approx(L0, rel=1e-9)
    grads = {fam: [t.grad.cpu().numpy() for t in tens[fam]] for fam in FAMS}
    h = 1e-5
    rng = np.random.default_rng(77)

    def five_point(fam, slot, direction, step):
        def at(t):
            p = list(base[fam])  # only the perturbed family's terms of the loss change
            p[slot] = base[fam][slot] + t * direction
            return c4.cpu({fam: p})
        return (-at(2 * step) + 8 * at(step) - 8 * at(-step) + at(-2 * step)) / (12 * step)

    def check(an, fam, slot, direction, rel):
        """Five-point central difference at h (truncation O(h^4): the loss's third derivatives
        along the Coulomb-like decay rates are large, from the closest intermolecular H-H
        contacts).  Its rounding noise (the oracle's loss carries ~1e-13 relative rounding) is
        measured as the spread of the same difference at 1.5 h and 2.25 h, where truncation is
        still negligible."""
        fd = [five_point(fam, slot, direction, h * f) for f in (1.0, 1.5, 2.25)]
        noise = max(3 * (max(fd) - min(fd)), 2e-13 * c4.cpu({fam: base[fam]}) / h)
        assert an == pytest.approx(fd[0], rel=rel, abs=noise), (fam, slot, fd)

    for fam in FAMS:
        gc, gt, gx = grads[fam]
        # coefficients: L is quadratic in them, so the difference is exact up to rounding
        check(float(gc @ (d := rng.normal(size=gc.shape))), fam, 0, d, 1e-8)
        # each nonlinear parameter
        for i in range(10):
            check(gt[i], fam, 1, np.eye(10)[i], 2e-6)
        # coordinates, along two random directions
        for _ in range(2):
            check(float((gx * (d := rng.normal(size=gx.shape))).sum()), fam, 2, d, 2e-6)
