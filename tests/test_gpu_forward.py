"""Born forward model (Simul_holo's calcHolo over all angles) and its adjoint
on the GPU, through the C ABI, against the oracle.

Bars: zero bins (outside the pupil, skipped angles) exactly zero; spectra and
adjoint volumes within rel-L2 1e-5 of the FP32 oracle, FP64 error printed;
the GPU pair satisfies <A x, y> = <x, A^H y> to 1e-5 at the differentiation
workload's full size (400 angles, C = 128, V = 256); simulate -> reconstruct
gives back the contrast on the hit voxels."""
import math

import numpy as np
import pytest

from cases import rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

TOL = 1e-5
LAM, N0, NA = 6.33e-7, 1.333, 1.2


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _empty(shape, dtype):
    import torch
    return torch.empty(shape, dtype=dtype, device="cuda")


def _say(msg):
    print(msg, flush=True)


def _geoms(nx, V, N=None):
    import mtd_transmission_amd as m
    N = N or 4 * 2 * nx
    pix = nx * LAM / (NA * N)
    cx, cy = N // 2 - round(0.3 * N), N // 2 + round(0.3 * N)
    g = ref.derive_geom(N, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    gl = m.geom_derive(N, nx, V, cx, cy, NA, N0, LAM, pix)
    return g, gl


def _ring_pos(nx, A, frac=0.9, skip_corner=True):
    t = np.arange(A) * 2 * math.pi / A
    px = nx + np.round(frac * nx * np.cos(t)).astype(np.int64)
    py = nx + np.round(frac * nx * np.sin(t)).astype(np.int64)
    if skip_corner:  # (0, 0): |fm0|^2 = 2 NXMAX^2 > R^2, a skipped angle
        px[A // 2], py[A // 2] = 0, 0
    return np.concatenate([px, py])


def _phantom(V, seed, cplx=False):
    rng = np.random.default_rng(seed)
    zz, yy, xx = (np.mgrid[0:V, 0:V, 0:V] - V // 2).astype(np.float64)
    dn = np.zeros((V, V, V))
    for _ in range(3):
        c = rng.uniform(-V / 6, V / 6, size=3)
        r = rng.uniform(V / 16, V / 8)
        dn += rng.uniform(0.02, 0.06) * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2
                                                 + (zz - c[2]) ** 2) / (2 * r * r))
    im = 0.1 * dn * rng.uniform(0.5, 1.0) if cplx else None
    return dn.astype(np.float32), (None if im is None else im.astype(np.float32))


SIZES = {"small": (16, 64, 24), "mid": (32, 128, 60)}


@pytest.mark.parametrize("name,cplx", [("small", False), ("small", True), ("mid", False)])
def test_forward_spectra(gpu_ctx, name, cplx):
    import torch
    nx, V, A = SIZES[name]
    _say(f"forward_spectra case={name} cplx={cplx}")
    g, gl = _geoms(nx, V)
    pos = _ring_pos(nx, A)
    dn, dni = _phantom(V, 5, cplx)
    x = dn.astype(np.complex128) + (0 if dni is None else 1j * dni.astype(np.float64))
    S64 = ref.born_forward(x, pos, g, "f64")
    S32 = ref.born_forward(x.astype(np.complex64), pos, g, "f32")
    C = g.C
    out = _empty((A, C, C), torch.complex64)
    gpu_ctx.forward_spectra(gl, _t(dn), None if dni is None else _t(dni),
                            _t(pos.astype(np.int32)), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    e32, e64 = rel_l2(got, S32), rel_l2(got, S64)
    _say(f"  spectra rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert np.all(got[S64 == 0] == 0)
    assert not got[A // 2].any()  # the skipped angle
    assert e32 < TOL


def test_specular_outside_crop(gpu_ctx):
    """Speculars a caller places outside the crop (x or y = -1 or C + 1, not
    skipped since |fm0| < R): bins whose voxel would leave the volume
    laterally are 0 (the reference indexes out of bounds there), the others
    equal the oracle's, for the forward model and its adjoint."""
    import torch
    nx, V = 16, 64
    C = 2 * nx
    _say("forward/adjoint with speculars outside the crop")
    g, gl = _geoms(nx, V)
    pos = np.array([-1, C + 1, nx, nx, nx, nx, -1, nx])  # x of 4 angles, then y
    A = 4
    dn, _ = _phantom(V, 9)
    S32 = ref.born_forward(dn.astype(np.complex64), pos, g, "f32")
    S64 = ref.born_forward(dn.astype(np.complex128), pos, g, "f64")
    assert (S64[0] == 0).sum() > (S64[3] == 0).sum()  # lateral exits happen
    out = _empty((A, C, C), torch.complex64)
    tpos = _t(pos.astype(np.int32))
    gpu_ctx.forward_spectra(gl, _t(dn), None, tpos, out)
    rng = np.random.default_rng(4)
    y = (rng.normal(size=(A, C, C)) + 1j * rng.normal(size=(A, C, C))).astype(np.complex64)
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    gpu_ctx.adjoint_spectra(gl, _t(y), tpos, re, im)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(got[S64 == 0] == 0)
    e32 = rel_l2(got, S32)
    adj = re.cpu().numpy() + 1j * im.cpu().numpy()
    e64 = rel_l2(adj, ref.born_adjoint(y.astype(np.complex128), pos, g))
    _say(f"  forward rel-L2 vs f32 {e32:.3e}; adjoint vs f64 {e64:.3e}")
    assert e32 < TOL and e64 < TOL


@pytest.mark.parametrize("name", ["small", "mid"])
def test_adjoint_spectra(gpu_ctx, name):
    import torch
    nx, V, A = SIZES[name]
    _say(f"adjoint_spectra case={name}")
    g, gl = _geoms(nx, V)
    pos = _ring_pos(nx, A)
    C = g.C
    rng = np.random.default_rng(12)
    y = (rng.normal(size=(A, C, C)) + 1j * rng.normal(size=(A, C, C))).astype(np.complex64)
    want = ref.born_adjoint(y.astype(np.complex128), pos, g)
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    gpu_ctx.adjoint_spectra(gl, _t(y), _t(pos.astype(np.int32)), re, im)
    torch.cuda.synchronize()
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    e64 = rel_l2(got, want)
    _say(f"  adjoint rel-L2 vs f64 {e64:.3e}")
    assert e64 < TOL


@pytest.mark.parametrize("nx,V,A", [(16, 64, 24), (8, 32, 10), (32, 128, 60), (64, 256, 400),
                                    (256, 1024, 32)])
def test_gpu_pair_is_adjoint(gpu_ctx, nx, V, A):
    """<A x, y> = <x, A^H y> with the GPU forward and adjoint (complex x);
    volumes 32^3 and 128^3 take the balanced column passes (4 x 8 and
    8 x 16 points on all of a column's threads) both ways; 256^3 is the
    differentiation workload (400 angles, 128^2 spectra), 1024^3 the large
    configuration's 512^2 spectra and 1024^3 volume (its forward and inverse
    1024-point passes)."""
    import torch
    _say(f"adjointness nx={nx} V={V} A={A}")
    g, gl = _geoms(nx, V)
    pos = _t(_ring_pos(nx, A, 0.95).astype(np.int32))
    C = 2 * nx
    gen = torch.Generator(device="cuda").manual_seed(21)
    xr = torch.randn((V, V, V), device="cuda", generator=gen)
    xi = torch.randn((V, V, V), device="cuda", generator=gen)
    y = torch.randn((A, C, C), dtype=torch.complex64, device="cuda", generator=gen)
    Ax = _empty((A, C, C), torch.complex64)
    gpu_ctx.forward_spectra(gl, xr, xi, pos, Ax)
    ar = _empty((V, V, V), torch.float32)
    ai = _empty((V, V, V), torch.float32)
    gpu_ctx.adjoint_spectra(gl, y, pos, ar, ai)
    torch.cuda.synchronize()
    lhs = torch.vdot(y.ravel().to(torch.complex128), Ax.ravel().to(torch.complex128)).item()
    # <x, A^H y> = sum conj(A^H y) x, accumulated in float64 plane by plane
    rhs = 0j
    for z in range(V):
        ar_, ai_ = ar[z].double(), ai[z].double()
        xr_, xi_ = xr[z].double(), xi[z].double()
        rhs += complex(float((ar_ * xr_ + ai_ * xi_).sum()), float((ar_ * xi_ - ai_ * xr_).sum()))
    _say(f"  <Ax,y> {lhs:.6e}  <x,A^H y> {rhs:.6e}")
    assert abs(lhs - rhs) <= TOL * abs(lhs)


def test_simulate_then_reconstruct(gpu_ctx):
    """GPU forward spectra -> GPU Ewald scatter + finalize = the contrast
    projected on the hit voxels (oracle, FP64)."""
    import torch
    nx, V, A = 16, 64, 40
    _say("simulate -> reconstruct")
    g, gl = _geoms(nx, V)
    pos = _ring_pos(nx, A, 0.9, skip_corner=False)
    dn, _ = _phantom(V, 6)
    S = _empty((A, g.C, g.C), torch.complex64)
    tpos = _t(pos.astype(np.int32))
    gpu_ctx.forward_spectra(gl, _t(dn), None, tpos, S)
    vol = _empty((V, V, V), torch.complex64)
    cnt = _empty((V, V, V), torch.int32)
    gpu_ctx.ewald_scatter(gl, S, tpos, vol, cnt)
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    gpu_ctx.finalize(gl, vol, cnt, re, im, None)
    torch.cuda.synchronize()
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    TF = ref.contrast_spectrum(dn.astype(np.complex128), g, "f64")
    masked = np.where(cnt.cpu().numpy() > 0, TF, 0)
    want = ref.fftshift3(np.fft.ifftn(ref.fftshift3(masked), norm="forward")) * g.delta_f ** 3
    e = rel_l2(got, want)
    _say(f"  round trip rel-L2 vs f64 projection {e:.3e}")
    assert e < TOL


def test_adjoint_of_no_angles_is_zero(gpu_ctx):
    import torch
    g, gl = _geoms(16, 64)
    V = 64
    re = torch.full((V, V, V), 7.0, device="cuda")
    im = torch.full((V, V, V), 7.0, device="cuda")
    pos = _empty((0,), torch.int32)
    gpu_ctx.adjoint_spectra(gl, _empty((0, 32, 32), torch.complex64), pos, re, im)
    torch.cuda.synchronize()
    assert not re.any() and not im.any()


def test_geometry_plan_reuse(gpu_ctx):
    """Forward, adjoint and weights through a geometry plan (posspec None)
    are bit-identical to the per-call geometry; a plan of another angle
    count is refused."""
    import torch
    import mtd_transmission_amd as m
    nx, V, A = 16, 64, 24
    g, gl = _geoms(nx, V)
    pos = _t(_ring_pos(nx, A).astype(np.int32))
    dn, _ = _phantom(V, 8)
    tdn = _t(dn)
    C = g.C
    S0 = _empty((A, C, C), torch.complex64)
    gpu_ctx.forward_spectra(gl, tdn, None, pos, S0)
    r0, i0 = _empty((V, V, V), torch.float32), _empty((V, V, V), torch.float32)
    gpu_ctx.adjoint_spectra(gl, S0, pos, r0, i0)
    w0 = _empty((A, C, C), torch.float32)
    gpu_ctx.cap_weights(gl, pos, w0)
    gpu_ctx.born_plan(gl, pos)
    S1 = _empty((A, C, C), torch.complex64)
    gpu_ctx.forward_spectra(gl, tdn, None, None, S1)
    r1, i1 = _empty((V, V, V), torch.float32), _empty((V, V, V), torch.float32)
    gpu_ctx.adjoint_spectra(gl, S0, None, r1, i1)
    w1 = _empty((A, C, C), torch.float32)
    gpu_ctx.cap_weights(gl, None, w1)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(S0), torch.view_as_real(S1))
    assert torch.equal(r0, r1) and torch.equal(i0, i1)
    assert torch.equal(w0, w1)
    with pytest.raises(m.MTDError):
        gpu_ctx.forward_spectra(gl, tdn, None, None, _empty((A - 1, C, C), torch.complex64))
