"""PyTorch plugin: MB-pol 2B / 3B energies of independent water dimers and trimers as a
differentiable function of the coordinates, the linear polynomial coefficients and the nonlinear
parameters.

The reference computes these energies in x2o::x2b_v9x::eval (potential/2b/x2b-v9x.cpp:1514) and
x2o::x3b_v2x::eval (potential/3b/x3b-v2x.cpp:1403) with the fitted coefficients and parameters
compiled in; its Python plugin (plugins/python/mbx/mbx.py) exposes energies and gradients of a
whole system only.  Here

    E, G = NmerEnergy.apply(xyz, coef, theta, family)      # G = dE/dxyz, per n-mer

is a torch.autograd.Function whose forward and backward run the HIP kernels of
mbx_amd/csrc/adjoint.hip through the C ABI (include/mbxg.h, mbxg_ad_*).  Because G is itself
an output with a gradient, losses in forces (F = -G) differentiate correctly: the backward is
the vector-Jacobian product of (xyz, coef, theta) -> (E, G), computed by a hand-written adjoint
(forward-over-reverse) kernel, not by taping.  There is no CPU fallback: the library and a GPU
are required.

    family   atoms per n-mer       coefficients   theta (10 nonlinear parameters)
    mbpol_2b 6  (O H H O H H)      1153           k_HH_intra k_OH_intra k_HH_coul k_OH_coul k_OO_coul
                                                  k_XH_main k_XO_main k_XX_main in_plane_gamma
                                                  out_of_plane_gamma   (x2b-v9x.cpp:1330-1345)
    mbpol_3b 9  (O H H x 3)        1163           kHH_intra kOH_intra kHH kOH kOO dHH_intra dOH_intra
                                                  dHH dOH dOO          (x3b-v2x.cpp:1256-1271)

Coordinates are n-mer-local, shape (n, 3*atoms) or (n, atoms, 3), float64 on a CUDA (HIP) device.
"""
import ctypes

import numpy as np
import torch

from . import engine

FAMILY = {"mbpol_2b": 0, "mbpol_3b": 1}


def _family(name):
    if name not in FAMILY:
        raise ValueError("unknown family %r (have %s)" % (name, ", ".join(FAMILY)))
    return FAMILY[name]


def shape(family):
    """(atoms per n-mer, number of coefficients, number of nonlinear parameters)."""
    a, c, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    engine._check(engine.lib().mbxg_ad_shape(_family(family), ctypes.byref(a), ctypes.byref(c), ctypes.byref(t)))
    return a.value, c.value, t.value


def defaults(family, device="cuda"):
    """The reference's fitted coefficients and nonlinear parameters, as float64 tensors."""
    _, nc, nt = shape(family)
    coef, theta = np.zeros(nc), np.zeros(nt)
    engine._check(engine.lib().mbxg_ad_defaults(_family(family), engine._d(coef), engine._d(theta)))
    return torch.tensor(coef, device=device), torch.tensor(theta, device=device)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_inputs(xyz, coef, theta, family):
    nat, nc, nt = shape(family)
    for name, t in (("xyz", xyz), ("coef", coef), ("theta", theta)):
        if not t.is_cuda or t.dtype != torch.float64:
            raise TypeError("%s must be a float64 tensor on a GPU" % name)
    if xyz.numel() % (3 * nat) or (xyz.dim() > 1 and xyz.shape[0] * 3 * nat != xyz.numel()):
        raise ValueError("xyz must hold %d coordinates per n-mer" % (3 * nat))
    if coef.numel() != nc or theta.numel() != nt:
        raise ValueError("%s takes %d coefficients and %d parameters" % (family, nc, nt))
    return xyz.numel() // (3 * nat)


def _workspace(fam, n, device):
    nbytes = engine.lib().mbxg_ad_workspace_bytes(fam, n)
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device), nbytes


class NmerEnergy(torch.autograd.Function):
    """(xyz, coef, theta) -> (E [n], dE/dxyz [same shape as xyz])."""

    @staticmethod
    def forward(ctx, xyz, coef, theta, family):
        n = _check_inputs(xyz, coef, theta, family)
        fam = _family(family)
        xyz, coef, theta = xyz.contiguous(), coef.contiguous(), theta.contiguous()
        energy = torch.empty(n, dtype=torch.float64, device=xyz.device)
        grad = torch.empty_like(xyz)
        ws, nbytes = _workspace(fam, n, xyz.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(xyz.device).cuda_stream)
        engine._check(engine.lib().mbxg_ad_forward(fam, n, _ptr(xyz), _ptr(coef), _ptr(theta), _ptr(energy),
                                                   _ptr(grad), _ptr(ws), nbytes, stream))
        ctx.save_for_backward(xyz, coef, theta)
        ctx.fam, ctx.n = fam, n
        ctx.ws, ctx.nbytes = ws, nbytes  # holds the polynomial's value and gradient for the backward
        return energy, grad

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g_energy, g_grad):
        xyz, coef, theta = ctx.saved_tensors
        n, fam = ctx.n, ctx.fam
        if g_energy is None:
            g_energy = torch.zeros(n, dtype=torch.float64, device=xyz.device)
        g_energy = g_energy.contiguous()
        g_grad = g_grad.contiguous() if g_grad is not None else None
        d_xyz, d_coef, d_theta = torch.empty_like(xyz), torch.empty_like(coef), torch.empty_like(theta)
        ws, nbytes = ctx.ws, ctx.nbytes  # read-only for the polynomial part: repeated backwards may share it
        stream = ctypes.c_void_p(torch.cuda.current_stream(xyz.device).cuda_stream)
        engine._check(engine.lib().mbxg_ad_backward(fam, n, _ptr(xyz), _ptr(coef), _ptr(theta), _ptr(g_energy),
                                                    _ptr(g_grad), _ptr(d_xyz), _ptr(d_coef), _ptr(d_theta),
                                                    _ptr(ws), nbytes, 1, stream))
        return d_xyz, d_coef, d_theta, None


def energy_forces(xyz, coef, theta, family):
    """Per-n-mer energies (kcal/mol) and forces F = -dE/dxyz (kcal/mol/A)."""
    e, g = NmerEnergy.apply(xyz, coef, theta, family)
    return e, -g


def energy_force_loss(energy, forces, energy_ref, forces_ref, w_energy=1.0, w_forces=1.0):
    """Mean squared energy error plus mean squared force-component error."""
    return w_energy * torch.mean((energy - energy_ref) ** 2) + w_forces * torch.mean((forces - forces_ref) ** 2)
