"""The plain method at the transform lengths the other parity cases do not
reach, against the FP32 oracle: 768^2 holograms (24 x 32 two-step FFT) with
crop 192 and a 384^3 volume, 2048^2 holograms with crop 128 and 256^3, and
1024^2 holograms with crop 320 and a 640^3 volume. A few angles each (seeded
ellipsoid phantom, Fermat spiral, Poisson noise); speculars bit-exact, index
volume within rel-L2 1e-5 of the FP32 oracle and within the FP64 bound
(twice the FP32 oracle's own distance, at least 1e-5) of the FP64 oracle."""
import numpy as np
import pytest

import lengths as L
from cases import rel_l2
from lengths import fp64_bound
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, N0, NA = 6.33e-7, 1.333, 1.2
TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n,nx,V,A", [(768, 96, 384, 5), (2048, 64, 256, 3),
                                      (1024, 160, 640, 4)])
def test_plain_at_size(gpu_ctx, n, nx, V, A):
    import torch
    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth
    _say(f"plain method N={n} crop={2 * nx} V={V} A={A}")
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    ph = synth.scale_phantom(synth.make_phantom(40 + n, 3), 0.45 * n * pix / 33.8e-6)
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph, seed=n + 1)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    gpu_ctx.reconstruct_plain(gl, holos, mask, re, im, pos)
    torch.cuda.synchronize()
    h = holos.cpu().numpy()
    f32, p32 = ref.preprocess(h, go, mask, "f32")
    assert np.array_equal(pos.cpu().numpy(), p32)
    idx32 = ref.reconstruct(f32, p32, go, "f32")[0]
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    e32 = rel_l2(got, idx32)
    del idx32
    f64, p64 = ref.preprocess(h, go, mask, "f64")
    if V <= 384:
        idx32 = ref.reconstruct(f32, p32, go, "f32")[0]
        idx64 = ref.reconstruct(f64, p64, go, "f64")[0]
        e64, bound = rel_l2(got, idx64), fp64_bound(idx32, idx64)
    else:  # the oracle's exact sum at sampled voxels
        sel = np.random.default_rng(V).integers(0, V ** 3, size=400)
        o32 = L.index_at(f32, p32, go, "f32", sel)
        o64 = L.index_at(f64, p64, go, "f64", sel)
        e64, bound = rel_l2(got.ravel()[sel], o64), fp64_bound(o32, o64)
    _say(f"  index rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e} (bound {bound:.2e})")
    assert e32 < TOL and e64 <= bound
