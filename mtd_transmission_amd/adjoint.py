"""Adjoint-iterative reconstruction (the differentiation form).

The Born forward model of the library (``mtd_forward_spectra``: the
simulator's Ewald-cap sampling of the centred contrast spectrum, calcHolo in
Simul_holo/src/fonctions.cpp:109-140) and its hand-written adjoint
(``mtd_adjoint_spectra``) sit behind a ``torch.autograd.Function``, so any
torch loss of the predicted spectra differentiates through the gfx950
kernels. ``reconstruct_tv`` uses it to solve

    min_{dn >= 0}  1/2 ||A dn - y||_W^2 / kappa  +  lam * TV_eps(dn)

for a real index contrast dn = n - n0 on the V^3 grid (positivity n >= n0),
where y are measured per-angle Born spectra (``Context.field_to_spectrum``
output), W are per-sample data weights (``Context.cap_weights``:
1/(q^2 * hits of the sample's voxel), under which A^H W A = Tp^6 times the
projection on the measured voxels, so every measured frequency converges at
the same rate; W = 1 gives plain least squares), kappa = ||A^H W A|| puts the
data term in index units, and TV_eps is the isotropic total variation with
forward differences smoothed by eps. The solver is FISTA with the projection
on dn >= 0. The missing cone the plain method leaves empty is filled by the
positivity and TV priors.

There is no CPU path: the forward and adjoint run through libmtd.so.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch


class BornSpectra(torch.autograd.Function):
    """S = A(dn) for a real contrast volume dn (V^3 float32, on the GPU) ->
    complex64 spectra [A][C][C]. Backward: Re(A^H grad_S), the vector-Jacobian
    product of a complex-linear map applied to a real input (torch passes the
    conjugate-Wirtinger gradient, 2 (S - y) for ||S - y||^2)."""

    @staticmethod
    def forward(ctx, dn, mtd, g, posspec, A=None):
        # posspec None: the context's geometry plan (Context.born_plan) of A angles
        A = posspec.numel() // 2 if posspec is not None else A
        C = g.C
        S = torch.empty((A, C, C), dtype=torch.complex64, device=dn.device)
        mtd.forward_spectra(g, dn.detach().contiguous(), None, posspec, S)
        ctx.mtd, ctx.g, ctx.posspec = mtd, g, posspec
        return S

    @staticmethod
    def backward(ctx, gS):
        V = ctx.g.dim_final
        re = torch.empty((V, V, V), dtype=torch.float32, device=gS.device)
        im = torch.empty_like(re)
        ctx.mtd.adjoint_spectra(ctx.g, gS.contiguous(), ctx.posspec, re, im)
        return re, None, None, None, None


def born_spectra(dn, mtd, g, posspec, A=None):
    """Differentiable A(dn); posspec None: the geometry plan of A angles."""
    return BornSpectra.apply(dn, mtd, g, posspec, A)


def tv(dn, eps):
    """Isotropic total variation, forward differences (zero across the far
    faces), smoothed: sum sqrt(dx^2 + dy^2 + dz^2 + eps^2) - eps."""
    dz = torch.diff(dn, dim=0, append=dn[-1:])
    dy = torch.diff(dn, dim=1, append=dn[:, -1:])
    dx = torch.diff(dn, dim=2, append=dn[:, :, -1:])
    return (torch.sqrt(dx * dx + dy * dy + dz * dz + eps * eps) - eps).sum()


def data_term(S, y, weights):
    """1/2 sum W |S - y|^2 (W = 1 when weights is None)."""
    r2 = torch.view_as_real(S - y).pow(2)
    if weights is not None:
        r2 = r2 * weights.unsqueeze(-1)
    return 0.5 * r2.sum()


def normal_norm(mtd, g, posspec, weights=None, iters=20, seed=0):
    """kappa = ||A^H W A|| restricted to real volumes, by power iteration on
    Re(A^H W A). The operator is tiny in SI units (Tp^3 ~ 1e-21 m^3), so norms
    are taken in float64: float32 squares of its outputs underflow. With the
    cap weights, kappa = Tp^6 V^3 exactly (the centred DFT's norm squared)."""
    V = g.dim_final
    A = posspec.numel() // 2
    mtd.born_plan(g, posspec)
    posspec = None
    gen = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn((V, V, V), device="cuda", generator=gen)
    v /= v.norm()
    lam = 0.0
    for _ in range(iters):
        v = v.requires_grad_(True)
        S = born_spectra(v, mtd, g, posspec, A)
        (w,) = torch.autograd.grad(data_term(S, 0, weights), v)
        w = w.double()
        lam = float(w.norm())
        v = (w / lam).float().detach()
    return lam


@dataclass
class TVResult:
    dn: torch.Tensor
    losses: list
    kappa: float
    iters: int
    seconds: float


def reconstruct_tv(mtd, g, posspec, y, n_iter=100, lam=1e-3, eps=1e-3, dn0=None,
                   positivity=True, kappa=None, weights=None, record=False, monitor=None,
                   fused=False, graph=False):
    """FISTA on 1/2 ||A dn - y||_W^2 / kappa + lam TV_eps(dn), dn >= 0.

    y: complex64 [A][C][C] measured spectra (field_to_spectrum frame).
    weights: float [A][C][C] (Context.cap_weights) or None.
    dn0: start volume (e.g. the plain reconstruction's real part), zeros
    otherwise. The step is 1/L with L = 1 + 12 lam / eps (the data term's
    Hessian has norm 1 after the kappa scaling; 12 bounds ||grad||^2 in
    3-D). Each iteration takes the data gradient by autograd through
    BornSpectra (GPU forward + adjoint, the output gradient W (S - y) formed
    directly), then one fused library step (TV gradient, gradient step,
    projection, momentum). record: the objective at
    each iterate z. monitor(k, x): called after iteration k (1-based) with
    the current estimate. fused: the same iterations run by the library
    (mtd_tv_solve: no autograd, identical iterates); graph: and captured as
    one HIP graph. record and monitor need the step-by-step loop. Returns a
    TVResult."""
    V = g.dim_final
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if kappa is None:
        kappa = normal_norm(mtd, g, posspec, weights)
    # one geometry plan for all the forward / adjoint calls of the loop
    A = posspec.numel() // 2
    mtd.born_plan(g, posspec)
    # lam and eps as the library receives them (float32), so that the step
    # is the one mtd_tv_solve derives from them
    lam, eps = float(np.float32(lam)), float(np.float32(eps))
    step = 1.0 / (1.0 + 12.0 * lam / eps)
    x = torch.zeros((V, V, V), device="cuda") if dn0 is None else dn0.detach().clone()
    if fused:
        if record or monitor is not None:
            raise ValueError("record and monitor need the step-by-step loop (fused=False)")
        mtd.tv_solve(g, y, weights, x, n_iter, lam, eps, 1.0 / kappa, positivity, graph)
        torch.cuda.synchronize()
        return TVResult(x, [], kappa, n_iter, time.perf_counter() - t0)
    if positivity:
        x.clamp_(min=0.0)
    z = x.clone()
    zn = torch.empty_like(x)
    t = 1.0
    losses = []
    for k in range(1, n_iter + 1):
        zg = z.detach().requires_grad_(True)
        S = born_spectra(zg, mtd, g, None, A)
        # d/dS* of 1/2 sum W |S - y|^2 is W (S - y): handed to autograd as the
        # output gradient, so the backward is BornSpectra's adjoint alone
        with torch.no_grad():
            gS = S - y
            if weights is not None:
                gS *= weights
            if record:
                ld = data_term(S, y, weights)
                losses.append(float(ld) / kappa + lam * float(tv(z, eps)))
        (gd,) = torch.autograd.grad(S, zg, grad_outputs=gS)
        tn = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
        mtd.tv_fista_step(z, gd, x, zn, 1.0 / kappa, lam, eps, step, (t - 1.0) / tn,
                          positivity)
        z, zn = zn, z
        t = tn
        if monitor is not None:
            monitor(k, x)
    torch.cuda.synchronize()
    return TVResult(x, losses, kappa, n_iter, time.perf_counter() - t0)
