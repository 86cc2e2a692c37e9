"""Parity at the benchmark's full size: 600 holograms 1024^2 uint16 (seeded
5-12-ellipsoid phantom, Fermat spiral to 0.95 NA, Poisson noise), crop 256^2,
512^3 volume, through the C ABI.

Preprocessing is independent per hologram, so the oracle (FP32) checks a
sample of 40 of the 600 holograms the GPU processed in one stack: speculars
bit-exact, fields within rel-L2 1e-5. The Ewald accumulation of all 600
angles' spectra into the 512^3 spectrum and counts is bit-exact against the
oracle's, the hit total equals the closed form sum of P over non-skipped
angles, the finalisation of that spectrum is within rel-L2 1e-5 of the
oracle's, and the fused plain method (the demodulation's column transform in
the hologram column pass, one row kernel from the row inverse DFT to the
forward row FFT, normalisation in the gather, spectra pre-weighted in the
forward pass) is within rel-L2 1e-5 of the oracle's finalised index too."""
import numpy as np
import pytest

from cases import rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, N0, NA = 6.33e-7, 1.333, 1.2
N, NX, V, A, SEED = 1024, 128, 512, 600, 2024
TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def full(gpu_ctx):
    import torch
    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth
    pix = NX * LAM / (NA * N)
    cx, cy = N // 2 - round(0.3 * N), N // 2 + round(0.3 * N)
    rng = np.random.default_rng(SEED)
    ph = synth.make_phantom(SEED, int(rng.integers(5, 13)))
    kx, ky = synth.scan_fermat(A, NX, 0.95)
    _say("full size: synthesizing 600 x 1024^2 uint16 holograms")
    holos, mask = synth.simulate_torch(N, NX, cx, cy, pix, LAM, N0, kx, ky, ph, seed=SEED + 1)
    gl = m.geom_derive(N, NX, V, cx, cy, NA, N0, LAM, pix)
    go = ref.derive_geom(N, NX, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    C = 2 * NX
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    fields = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.preprocess(gl, holos, mask, pos, fields)
    spectra = torch.empty_like(fields)
    gpu_ctx.field_to_spectrum(gl, fields, pos, spectra)
    torch.cuda.synchronize()
    return dict(gl=gl, go=go, holos=holos, mask=mask, pos=pos, fields=fields, spectra=spectra)


def test_full_preprocess_sample(full):
    """40 of the 600 holograms: speculars bit-exact, fields rel-L2 1e-5."""
    idx = np.arange(0, A, 15)
    _say(f"full size: oracle preprocessing of {idx.size} holograms")
    import torch
    # uint16 tensors have no CUDA indexing: gather through an int16 view
    h = full["holos"].view(torch.int16)[torch.from_numpy(idx).cuda()].cpu().numpy()
    h = h.view(np.uint16)
    f32, p32 = ref.preprocess(h, full["go"], full["mask"], "f32")
    pos = full["pos"].cpu().numpy()
    n = idx.size
    assert np.array_equal(pos[idx], p32[:n]) and np.array_equal(pos[A + idx], p32[n:])
    got = full["fields"][idx].cpu().numpy()
    e = rel_l2(got, f32)
    _say(f"  fields rel-L2 vs f32 {e:.3e}")
    assert e < TOL


def test_full_scatter_finalize_and_fused(full, gpu_ctx):
    import torch
    gl, go = full["gl"], full["go"]
    pos = full["pos"].cpu().numpy().astype(np.int64)
    S = full["spectra"].cpu().numpy()
    vol = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.empty((V, V, V), dtype=torch.int32, device="cuda")
    gpu_ctx.ewald_scatter(gl, full["spectra"], full["pos"], vol, cnt)
    torch.cuda.synchronize()
    _say("full size: oracle Ewald accumulation of 600 angles into 512^3")
    vol32, cnt32 = ref.ewald_scatter(S, pos, go, "f32")
    gc = cnt.cpu().numpy().ravel()
    assert np.array_equal(gc, cnt32)
    # closed form: every non-skipped angle hits P voxels
    _, _, pupil, _, _ = ref.pupil_tables(go)
    R = go.rayon
    fm = pos[:A] - NX, pos[A:] - NX
    live = ref.round_half_away(R * R - fm[0] ** 2 - fm[1] ** 2) >= 0
    assert int(gc.sum()) == int(live.sum()) * int(pupil.sum())
    gv = vol.cpu().numpy().ravel()
    assert np.array_equal(gv.view(np.uint32), vol32.view(np.uint32))
    _say(f"  hits {int(gc.sum())} ({int(live.sum())} angles x P {int(pupil.sum())}), "
         "values bit-exact")
    del gv, gc
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    gpu_ctx.finalize(gl, vol, cnt, re, im, None)
    torch.cuda.synchronize()
    _say("full size: oracle finalisation of the 512^3 spectrum")
    idx32, _ = ref.finalize(vol32, cnt32, go, "f32")
    del vol32, cnt32
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    e = rel_l2(got, idx32)
    _say(f"  index rel-L2 vs f32 {e:.3e}")
    assert e < TOL
    del got
    # the fused plain method (one call from the holograms) against the same
    # oracle index and against the stage chain
    re2 = torch.empty_like(re)
    im2 = torch.empty_like(re)
    pos2 = torch.empty_like(full["pos"])
    gpu_ctx.reconstruct_plain(gl, full["holos"], full["mask"], re2, im2, pos2)
    torch.cuda.synchronize()
    assert torch.equal(pos2, full["pos"])
    got2 = re2.cpu().numpy() + 1j * im2.cpu().numpy()
    e2 = rel_l2(got2, idx32)
    del idx32
    chain = re.cpu().numpy() + 1j * im.cpu().numpy()
    ec = rel_l2(got2, chain)
    _say(f"  fused plain method: index rel-L2 vs f32 {e2:.3e}, vs the stage chain {ec:.3e}")
    assert e2 < TOL and ec < TOL
