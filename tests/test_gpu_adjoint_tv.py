"""The differentiation form on the GPU: the autograd gradient through the
hand-written adjoint against the FP64 oracle gradient (64^3, 60 angles), and
the TV + positivity reconstruction against the plain method on noisy
Born-model data (index RMSE against the phantom)."""
import math

import numpy as np
import pytest

from cases import rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, N0, NA = 6.33e-7, 1.333, 1.2


def _say(msg):
    print(msg, flush=True)


def _geoms(nx, V):
    import mtd_transmission_amd as m
    N = 8 * nx
    pix = nx * LAM / (NA * N)
    cx, cy = N // 2 - round(0.3 * N), N // 2 + round(0.3 * N)
    return (ref.derive_geom(N, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1),
            m.geom_derive(N, nx, V, cx, cy, NA, N0, LAM, pix))


def _fermat(A, nx, frac=0.95):
    j = np.arange(A)
    r = frac * nx * np.sqrt(j / A)
    th = j * np.pi * (3 - math.sqrt(5))
    return np.concatenate([nx + np.rint(r * np.cos(th)),
                           nx + np.rint(r * np.sin(th))]).astype(np.int32)


def _ellipsoids(V, seed, k=4):
    rng = np.random.default_rng(seed)
    zz, yy, xx = (np.mgrid[0:V, 0:V, 0:V] - V // 2).astype(np.float64)
    dn = np.zeros((V, V, V))
    for _ in range(k):
        c = rng.uniform(-V / 6, V / 6, size=3)
        ax = rng.uniform(V / 14, V / 7, size=3)
        inside = ((xx - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 \
            + ((zz - c[2]) / ax[2]) ** 2 <= 1
        dn = np.maximum(dn, inside * rng.uniform(0.03, 0.09))
    return dn.astype(np.float32)


def test_autograd_gradient_matches_fp64(gpu_ctx):
    """d/dn 1/2 ||A n - y||^2 by torch autograd through BornSpectra (GPU
    forward + adjoint) against Re(A^H (A n - y)) of the FP64 oracle."""
    import torch
    from mtd_transmission_amd import adjoint as adj
    nx, V, A = 16, 64, 60
    _say("autograd gradient 64^3 x 60 angles")
    g, gl = _geoms(nx, V)
    pos = _fermat(A, nx)
    n = _ellipsoids(V, 4)
    rng = np.random.default_rng(2)
    y = (2e-12 * (rng.normal(size=(A, g.C, g.C)) + 1j * rng.normal(size=(A, g.C, g.C))))
    want = ref.born_adjoint(ref.born_forward(n.astype(np.float64), pos, g, "f64") - y, pos, g).real
    tn = torch.from_numpy(n).cuda().requires_grad_(True)
    ty = torch.from_numpy(y.astype(np.complex64)).cuda()
    tpos = torch.from_numpy(pos).cuda()
    S = adj.born_spectra(tn, gpu_ctx, gl, tpos)
    loss = 0.5 * torch.view_as_real(S - ty).pow(2).sum()
    (grad,) = torch.autograd.grad(loss, tn)
    got = grad.double().cpu().numpy()
    e = rel_l2(got, want)
    _say(f"  gradient rel-L2 vs f64 {e:.3e}")
    assert e < 1e-5


def test_cap_weights(gpu_ctx):
    """Weights against the oracle (FP32, rel 1e-6), and the property they
    exist for: ||A^H W A|| = Tp^6 V^3 (power iteration vs the closed form)."""
    import torch
    from mtd_transmission_amd import adjoint as adj
    nx, V, A = 16, 64, 60
    g, gl = _geoms(nx, V)
    pos = _fermat(A, nx)
    pos[A // 3], pos[A + A // 3] = 0, 0  # a skipped angle
    want = ref.cap_weights(pos, g)
    tpos = torch.from_numpy(pos).cuda()
    w = torch.empty((A, g.C, g.C), device="cuda")
    gpu_ctx.cap_weights(gl, tpos, w)
    got = w.cpu().numpy()
    assert np.array_equal(got == 0, want == 0)
    nz = want != 0
    err = float(np.max(np.abs(got[nz] / want[nz] - 1)))
    kw = adj.normal_norm(gpu_ctx, gl, tpos, w)
    kc = ref.tomo_pixel(g) ** 6 * V ** 3
    _say(f"  weights max rel err {err:.2e}; kappa {kw:.6e} closed form {kc:.6e}")
    assert err < 1e-6
    assert kw == pytest.approx(kc, rel=1e-4)


def test_fused_tv_step_matches_torch(gpu_ctx):
    """The fused library step against the same step written with torch
    autograd (TV by torch.diff), on a random volume with a random data
    gradient; positivity on and off."""
    import torch
    from mtd_transmission_amd import adjoint as adj
    V = 48
    gen = torch.Generator(device="cuda").manual_seed(4)
    z = 0.05 * torch.rand((V, V, V), device="cuda", generator=gen) - 0.01
    gd = torch.randn((V, V, V), device="cuda", generator=gen)
    x0 = 0.05 * torch.rand((V, V, V), device="cuda", generator=gen)
    sd, lam, eps, step, beta = 0.3, 2e-3, 1e-3, 0.05, 0.4
    for pos in (True, False):
        zg = z.clone().requires_grad_(True)
        (gtv,) = torch.autograd.grad(adj.tv(zg, eps), zg)
        xr = z - step * (sd * gd + lam * gtv)
        if pos:
            xr = xr.clamp(min=0.0)
        zr = xr + beta * (xr - x0)
        x = x0.clone()
        zn = torch.empty_like(z)
        gpu_ctx.tv_fista_step(z, gd, x, zn, sd, lam, eps, step, beta, pos)
        torch.cuda.synchronize()
        ex, ez = rel_l2(x.cpu().numpy(), xr.cpu().numpy()), rel_l2(zn.cpu().numpy(), zr.cpu().numpy())
        _say(f"  fused step positivity={pos}: x rel-L2 {ex:.2e}, z rel-L2 {ez:.2e}")
        assert ex < 1e-5 and ez < 1e-5


def test_tv_beats_plain_on_born_data(gpu_ctx):
    """Noisy (1 %) Born-model spectra of a 4-ellipsoid phantom: FISTA with TV
    and positivity from the plain start lowers the index RMSE by >= 30 %."""
    import torch
    from mtd_transmission_amd import adjoint as adj
    nx, V, A = 16, 64, 60
    _, gl = _geoms(nx, V)
    pos = torch.from_numpy(_fermat(A, nx)).cuda()
    dn = torch.from_numpy(_ellipsoids(V, 3)).cuda()
    C = 2 * nx
    y = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.forward_spectra(gl, dn, None, pos, y)
    gen = torch.Generator(device="cuda").manual_seed(1)
    sig = 0.01 * float(y.abs().double().pow(2).mean().sqrt())
    y = y + sig * torch.complex(torch.randn(y.shape, device="cuda", generator=gen),
                                torch.randn(y.shape, device="cuda", generator=gen))
    vol = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.empty((V, V, V), dtype=torch.int32, device="cuda")
    gpu_ctx.ewald_scatter(gl, y, pos, vol, cnt)
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    gpu_ctx.finalize(gl, vol, cnt, re, im, None)
    r = adj.reconstruct_tv(gpu_ctx, gl, pos, y, n_iter=100, lam=1e-4, eps=1e-3, dn0=re,
                           record=True)
    rmse = lambda x: float((x.double() - dn.double()).pow(2).mean().sqrt())
    rp, rt = rmse(re), rmse(r.dn)
    _say(f"  index RMSE plain {rp:.5f}  adjoint-TV {rt:.5f}  ({r.seconds * 1e3:.0f} ms, "
         f"loss {r.losses[0]:.3g} -> {r.losses[-1]:.3g})")
    assert all(np.isfinite(r.losses))
    assert r.losses[-1] < r.losses[0]
    assert float(r.dn.min()) >= 0.0
    assert rt <= 0.7 * rp


@pytest.mark.parametrize("weighted", [True, False])
def test_fused_solver_matches_loop(gpu_ctx, weighted):
    """mtd_tv_solve, eagerly and captured as one HIP graph, gives the
    step-by-step loop's iterate bit for bit (64^3, 60 angles, 12 iterations,
    noisy Born data, plain start)."""
    import torch
    from mtd_transmission_amd import adjoint as adj
    nx, V, A = 16, 64, 60
    _say(f"fused solver vs loop, weighted={weighted}")
    _, gl = _geoms(nx, V)
    pos = torch.from_numpy(_fermat(A, nx)).cuda()
    dn = torch.from_numpy(_ellipsoids(V, 5)).cuda()
    C = 2 * nx
    y = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.forward_spectra(gl, dn, None, pos, y)
    gen = torch.Generator(device="cuda").manual_seed(3)
    sig = 0.01 * float(y.abs().double().pow(2).mean().sqrt())
    y = y + sig * torch.complex(torch.randn(y.shape, device="cuda", generator=gen),
                                torch.randn(y.shape, device="cuda", generator=gen))
    vol = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.empty((V, V, V), dtype=torch.int32, device="cuda")
    gpu_ctx.ewald_scatter(gl, y, pos, vol, cnt)
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    gpu_ctx.finalize(gl, vol, cnt, re, im, None)
    W = None
    if weighted:
        W = torch.empty((A, C, C), device="cuda")
        gpu_ctx.cap_weights(gl, pos, W)
    kw = dict(n_iter=12, lam=1e-3, eps=1e-2, dn0=re, weights=W)
    kappa = adj.normal_norm(gpu_ctx, gl, pos, W)
    loop = adj.reconstruct_tv(gpu_ctx, gl, pos, y, kappa=kappa, **kw)
    fused = adj.reconstruct_tv(gpu_ctx, gl, pos, y, kappa=kappa, fused=True, **kw)
    graph = adj.reconstruct_tv(gpu_ctx, gl, pos, y, kappa=kappa, fused=True, graph=True, **kw)
    again = adj.reconstruct_tv(gpu_ctx, gl, pos, y, kappa=kappa, fused=True, graph=True, **kw)
    _say(f"  loop {loop.seconds * 1e3:.1f} ms, fused {fused.seconds * 1e3:.1f} ms, "
         f"graph {graph.seconds * 1e3:.1f} ms; max |x| {float(loop.dn.abs().max()):.4f}")
    assert float(loop.dn.abs().max()) > 0
    assert torch.equal(fused.dn, loop.dn)
    assert torch.equal(graph.dn, loop.dn)
    assert torch.equal(again.dn, loop.dn)
