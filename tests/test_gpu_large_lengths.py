"""The large configuration's transform lengths: 2048^2 holograms with crop 512
(its preprocessing against the FP32 oracle), and the 1024^3 finalisation
(3-D inverse FFT with the centring signs and constants) against its closed
form for a spectrum of a few normalised deltas, each a plane wave of known
frequency, compared at 200 000 sampled voxels."""
import math

import numpy as np
import pytest

from cases import rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, N0, NA = 6.33e-7, 1.333, 1.2


def _say(msg):
    print(msg, flush=True)


def test_preprocess_2048_crop_512(gpu_ctx):
    import torch
    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth
    n, nx, V, A = 2048, 256, 1024, 2
    _say("preprocess 2048^2 holograms, crop 512")
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    ph = synth.make_phantom(77, 4)
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph, seed=78)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    C = 2 * nx
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    fields = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.preprocess(gl, holos, mask, pos, fields)
    torch.cuda.synchronize()
    f32, p32 = ref.preprocess(holos.cpu().numpy(), go, mask, "f32")
    assert np.array_equal(pos.cpu().numpy(), p32)
    e = rel_l2(fields.cpu().numpy(), f32)
    _say(f"  fields rel-L2 vs f32 {e:.3e}")
    assert e < 1e-5


@pytest.mark.parametrize("dtype,nx,cx,cy", [("u8", 64, 700, 1400), ("u16", 320, 1020, 1030),
                                            ("f32", 300, 333, 1700), ("u16", 512, 600, 1500)])
def test_holo_to_crop_2048(gpu_ctx, dtype, nx, cx, cy):
    """The 2048 row pass (each row's even and odd samples as one 1024-point
    complex transform) for every input type, with crops up to N/4 (the
    separation's first pass kept in registers) and wider (staged in the
    output slots; crop 1024 reaches bins past N/2). Fringes of the carrier at
    (cx, cy) with a seeded speckle and noise; crops within rel-L2 1e-5 of the
    FP32 oracle and within the FP64 bound of the FP64 oracle."""
    import torch
    import mtd_transmission_amd as m
    from lengths import fp64_bound
    n, A, C = 2048, 2, 2 * nx
    _say(f"holo_to_crop N=2048 {dtype} crop={C} carrier=({cx},{cy})")
    rng = np.random.default_rng(C + cx)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    fr = 2 * np.pi * ((cx - n // 2) * xx + (cy - n // 2) * yy) / n
    top = {"u8": 250.0, "u16": 60000.0, "f32": 4000.0}[dtype]
    holos = []
    for a in range(A):
        ph = rng.uniform(-0.5, 0.5, (n // 64, n // 64)).repeat(64, 0).repeat(64, 1)
        h = 0.5 + 0.35 * np.cos(fr + ph + a) + 0.05 * rng.standard_normal((n, n))
        holos.append(np.clip(h, 0, 1) * top)
    h = np.stack(holos)
    h = {"u8": h.astype(np.uint8), "u16": h.astype(np.uint16),
         "f32": h.astype(np.float32)}[dtype]
    pix = 1e-7
    gl = m.geom_derive(n, nx, 2 * C, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, 2 * C, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    crops = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.holo_to_crop(gl, torch.from_numpy(h).cuda(), crops)
    torch.cuda.synchronize()
    got = crops.cpu().numpy()
    exp32 = np.stack([ref.holo_to_crop(x, go, "f32") for x in h])
    exp64 = np.stack([ref.holo_to_crop(x, go, "f64") for x in h])
    e32, e64 = rel_l2(got, exp32), rel_l2(got, exp64)
    _say(f"  rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert e32 < 1e-5 and e64 <= fp64_bound(exp32, exp64)


def test_finalize_1024_deltas_closed_form(gpu_ctx):
    import torch
    import mtd_transmission_amd as m
    n, nx, V = 2048, 256, 1024
    _say("finalize 1024^3: three normalised deltas")
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    # (x, y, z) voxels of the centred spectrum, values, counts
    deltas = [((512 + 37, 512 - 101, 512 + 3), 6.0 - 3.0j, 3),
              ((512 - 250, 512 + 7, 512 - 64), -1.5 + 2.5j, 1),
              ((512, 512, 512), 0.5 + 0.25j, 2)]
    vol = torch.zeros((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.zeros((V, V, V), dtype=torch.int32, device="cuda")
    for (x, y, z), v, c in deltas:
        vol[z, y, x] = v
        cnt[z, y, x] = c
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    gpu_ctx.finalize(gl, vol, cnt, re, im, None)
    torch.cuda.synchronize()
    del vol, cnt
    rng = np.random.default_rng(9)
    idx = rng.integers(0, V ** 3, size=200_000)
    ti = torch.from_numpy(idx).cuda()
    got = (re.view(-1)[ti].double() + 1j * im.view(-1)[ti].double()).cpu().numpy()
    z, rem = np.divmod(idx, V * V)
    y, x = np.divmod(rem, V)
    kv = 2 * math.pi / go.lam
    cte = -1 / (2 * go.n0 * kv * kv)
    want = np.zeros(idx.size, complex)
    for (ux, uy, uz), v, c in deltas:
        k = (ux - V // 2, uy - V // 2, uz - V // 2)
        ph = 2 * np.pi * (k[0] * (x - V // 2) + k[1] * (y - V // 2) + k[2] * (z - V // 2)) / V
        want += (v / c) * np.exp(1j * ph)
    want *= cte * go.delta_f ** 3
    e = rel_l2(got, want)
    _say(f"  index rel-L2 vs closed form at 200000 voxels {e:.3e}")
    assert e < 1e-5


def test_scatter_1024_bit_exact_at_hits(gpu_ctx):
    """Ewald accumulation into a 1024^3 spectrum + counts (crop 512, 12
    angles on a Fermat spiral, random spectra): at every voxel the oracle's
    sparse form hits, values and counts bit-exact; no count and no value
    anywhere else (totals and nonzero voxels agree)."""
    import torch
    import mtd_transmission_amd as m
    n, nx, V, A = 2048, 256, 1024, 12
    _say("Ewald accumulation at 1024^3, 12 angles")
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    C = 2 * nx
    j = np.arange(A)
    r = 0.95 * nx * np.sqrt(j / A)
    th = j * np.pi * (3 - math.sqrt(5))
    pos = np.concatenate([nx + np.rint(r * np.cos(th)),
                          nx + np.rint(r * np.sin(th))]).astype(np.int64)
    rng = np.random.default_rng(10)
    S = (rng.normal(size=(A, C, C)) + 1j * rng.normal(size=(A, C, C))).astype(np.complex64)
    vox, sums, c = ref.ewald_scatter_sparse(S, pos, go, "f32")
    vol = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.empty((V, V, V), dtype=torch.int32, device="cuda")
    gpu_ctx.ewald_scatter(gl, torch.from_numpy(S).cuda(),
                          torch.from_numpy(pos.astype(np.int32)).cuda(), vol, cnt)
    torch.cuda.synchronize()
    tv = torch.from_numpy(vox).cuda()
    gv = vol.view(-1)[tv].cpu().numpy()
    gc = cnt.view(-1)[tv].cpu().numpy()
    assert np.array_equal(gc, c)
    assert np.array_equal(gv.view(np.uint32), sums.view(np.uint32))
    total = int(cnt.sum(dtype=torch.int64))
    nz = int((cnt != 0).sum())
    nzv = int((torch.view_as_real(vol) != 0).any(-1).sum())
    _say(f"  {vox.size} hit voxels, {int(c.sum())} hits; GPU total {total}, "
         f"nonzero counts {nz}, nonzero values {nzv}")
    assert total == int(c.sum()) and nz == vox.size
    assert nzv == int(np.count_nonzero(sums))


def test_forward_1024_point_object_closed_form(gpu_ctx):
    """Born forward model at 1024^3 with crop 512: a point contrast at
    centred offset d gives on every cap sample w(fd) Tp^3 c exp(-2 pi i u.d/V),
    u = (fd - fm0, koz), and 0 off the pupil (the closed form the oracle's
    forward model is pinned to); 4 angles, one of them off the sphere."""
    import torch
    import mtd_transmission_amd as m
    n, nx, V = 2048, 256, 1024
    _say("forward model at 1024^3: point contrast")
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    C, R = 2 * nx, go.rayon
    d = np.array([37, -55, 12])
    cval = 0.05
    dn = torch.zeros((V, V, V), device="cuda")
    dn[V // 2 + d[2], V // 2 + d[1], V // 2 + d[0]] = cval
    spec = [(nx, nx), (nx + 140, nx - 90), (nx - 200, nx + 60), (0, 0)]
    A = len(spec)
    pos = np.array([p[0] for p in spec] + [p[1] for p in spec], dtype=np.int64)
    S = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.forward_spectra(gl, dn, None, torch.from_numpy(pos.astype(np.int32)).cuda(), S)
    torch.cuda.synchronize()
    del dn
    got = S.cpu().numpy().astype(np.complex128)
    kv = 2 * 3.14159 / go.lam
    tp3 = ref.tomo_pixel(go) ** 3
    fy, fx = np.mgrid[-nx:nx, -nx:nx]
    pupil = fx * fx + fy * fy < nx * nx
    h = np.sqrt(np.maximum(R * R - fx * fx - fy * fy, 0.0))
    want = np.zeros_like(got)
    for a in range(A):
        fxm0, fym0 = int(pos[a]) - nx, int(pos[a + A]) - nx
        if ref.round_half_away(R * R - fxm0 ** 2 - fym0 ** 2) < 0:
            continue  # off the sphere: a skipped angle, all zero
        fzm0 = math.sqrt(R * R - fxm0 ** 2 - fym0 ** 2)
        koz = ref.round_half_away(h - fzm0)
        ud = (fx - fxm0) * d[0] + (fy - fym0) * d[1] + koz * d[2]
        w = np.where(pupil, (-1j * kv / np.where(pupil, h / R, 1.0)) * tp3 * cval
                     * np.exp(-2j * np.pi * ud / V), 0)
        want[a] = w
    assert not got[3].any()
    assert np.all(got[:, ~pupil] == 0)
    e = rel_l2(got, want)
    _say(f"  spectra rel-L2 vs closed form {e:.3e}")
    assert e < 1e-5
