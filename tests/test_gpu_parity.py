"""Parity of the gfx950 path (through the C ABI) against the oracle.

Bars: posspec and hit counts bit-exact; the Ewald accumulation bit-exact
against the FP32 oracle on identical spectra; complex crops, fields, spectra
and the index volume within rel-L2 1e-5 of the FP32 oracle (the reference FP32
path), and within twice the FP32 oracle's own distance from the FP64 oracle
(at least 1e-5) of the FP64 oracle, which each test also prints.
"""
import numpy as np
import pytest

from lengths import fp64_bound
from cases import make_case, rel_l2
from oracle import ref

TOL = 1e-5

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _empty(shape, dtype):
    import torch
    return torch.empty(shape, dtype=dtype, device="cuda")


def _say(msg):
    print(msg, flush=True)


def oracle_front(case, prec):
    """Oracle crops, posspec, fields."""
    fields, pos, crops = ref.preprocess(case.acq.holos, case.g, case.acq.mask, prec,
                                        return_crops=True)
    return crops, pos, fields


def oracle_spectra(case, fields, pos, prec):
    A = fields.shape[0]
    return np.stack([ref.field_to_spectrum(fields[a], int(pos[a]), int(pos[a + A]),
                                           case.g, prec) for a in range(A)])


@pytest.mark.parametrize("name", ["tiny", "small", "u16", "cfg1"])
def test_holo_to_crop(gpu_ctx, name):
    import torch
    case = make_case(name)
    _say(f"holo_to_crop case={name} holos={case.acq.holos.shape} {case.acq.holos.dtype}")
    g = case.lib_geom()
    A, C = case.acq.holos.shape[0], g.C
    crops = _empty((A, C, C), torch.complex64)
    gpu_ctx.holo_to_crop(g, _t(case.acq.holos), crops)
    torch.cuda.synchronize()
    got = crops.cpu().numpy()
    exp32 = np.stack([ref.holo_to_crop(h, case.g, "f32") for h in case.acq.holos])
    exp64 = np.stack([ref.holo_to_crop(h, case.g, "f64") for h in case.acq.holos])
    e32, e64 = rel_l2(got, exp32), rel_l2(got, exp64)
    _say(f"  rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}  (f32 oracle vs f64 {rel_l2(exp32, exp64):.3e})")
    assert e32 < TOL and e64 <= fp64_bound(exp32, exp64)


@pytest.mark.parametrize("name", ["nomask", "small", "u16", "cfg1"])
def test_crop_to_field(gpu_ctx, name):
    import torch
    case = make_case(name)
    _say(f"crop_to_field case={name}")
    g = case.lib_geom()
    crops32, pos32, fields32 = oracle_front(case, "f32")
    _, pos64, fields64 = oracle_front(case, "f64")
    A, C = crops32.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.crop_to_field(g, _t(crops32), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    gp = pos.cpu().numpy()
    assert np.array_equal(gp, pos32), (gp, pos32)
    assert np.array_equal(gp, pos64)
    got = fields.cpu().numpy()
    e32, e64 = rel_l2(got, fields32), rel_l2(got, fields64)
    _say(f"  fields rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert e32 < TOL and e64 <= fp64_bound(fields32, fields64)


@pytest.mark.parametrize("name", ["small", "u16"])
@pytest.mark.parametrize("born,c_aber", [(0, 1), (1, 0)])
def test_crop_to_field_modes(gpu_ctx, name, born, c_aber):
    """Rytov (BORN 0) and the C_ABER 0 branch (main.cpp:311-341)."""
    import torch
    case = make_case(name, born=born, c_aber=c_aber)
    _say(f"crop_to_field modes case={name} born={born} c_aber={c_aber}")
    g = case.lib_geom()
    crops32, pos32, fields32 = oracle_front(case, "f32")
    A, C = crops32.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.crop_to_field(g, _t(crops32), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), pos32)
    e32 = rel_l2(fields.cpu().numpy(), fields32)
    _say(f"  rel-L2 vs f32 {e32:.3e}")
    assert e32 < TOL


@pytest.mark.parametrize("name,born", [("small", 1), ("small", 0), ("u16", 0), ("cfg1", 1),
                                       ("cfg1", 0)])
def test_crop_to_field_volkov(gpu_ctx, name, born):
    """DEROUL 1 / VOLKOV 1: the wrapped phase is unwrapped
    (deroul_volkov4_total_sym_paire) before the fit (main.cpp:260-277). With
    Rytov (BORN 0) the unwrapped phase is the field's imaginary part."""
    import torch
    case = make_case(name, born=born, deroul=1)
    _say(f"crop_to_field volkov case={name} born={born}")
    g = case.lib_geom()
    crops32, pos32, fields32 = oracle_front(case, "f32")
    _, _, fields64 = oracle_front(case, "f64")
    A, C = crops32.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.crop_to_field(g, _t(crops32), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), pos32)
    got = fields.cpu().numpy()
    e32, e64 = rel_l2(got, fields32), rel_l2(got, fields64)
    _say(f"  rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}  (f32 oracle vs f64 "
         f"{rel_l2(fields32, fields64):.3e})")
    assert e32 < TOL and e64 <= fp64_bound(fields32, fields64)


@pytest.mark.parametrize("name,born", [("small", 1), ("small", 0), ("u16", 1), ("u16", 0),
                                       ("cfg1", 1), ("cfg1", 0)])
def test_crop_to_field_herraez(gpu_ctx, name, born):
    """DEROUL 1 / VOLKOV 0: Herraez unwrapping (phaseUnwrapping_Mat) before
    the fit, against the oracle's literal restatement (the reference's sort,
    grouping and border rand() values): the GPU builds the same spanning
    tree (Boruvka on the sort positions); the unwrapped phase may differ by
    one global 2 pi k, which the fit's constant term removes. At most 20
    angles (the oracle's grouping is a Python loop)."""
    import torch
    case = make_case(name, born=born, deroul=2)
    _say(f"crop_to_field herraez case={name} born={born}")
    holos = case.acq.holos[:20]
    g = case.lib_geom()
    f32, p32, crops32 = ref.preprocess(holos, case.g, case.acq.mask, "f32", return_crops=True)
    f64, _ = ref.preprocess(holos, case.g, case.acq.mask, "f64")
    A, C = crops32.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.crop_to_field(g, _t(crops32), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), p32)
    got = fields.cpu().numpy()
    e32, e64 = rel_l2(got, f32), rel_l2(got, f64)
    _say(f"  rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}  (f32 oracle vs f64 "
         f"{rel_l2(f32, f64):.3e})")
    assert e32 < TOL and e64 <= fp64_bound(f32, f64)


@pytest.mark.parametrize("name", ["small", "u16"])
def test_arrayfire_windows(gpu_ctx, name):
    """The ArrayFire variant's windows (alpha 0.025 for the hologram, the
    (1+i) tukey2D(C, 0.025) crop window; pretraitement_hors_axe_GPU
    main.cpp:74-79): crops within 1e-5 of the FP32 oracle, speculars
    bit-exact (found on the windowed crops), fields within 1e-5."""
    import dataclasses
    import torch
    case = make_case(name)
    _say(f"arrayfire windows case={name}")
    g = case.lib_geom()
    g.alpha_holo = 0.025
    g.alpha_crop = 0.025
    go = dataclasses.replace(case.g, alpha_holo=0.025, alpha_crop=0.025)
    A, C = case.acq.holos.shape[0], g.C
    crops = _empty((A, C, C), torch.complex64)
    gpu_ctx.holo_to_crop(g, _t(case.acq.holos), crops)
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.preprocess(g, _t(case.acq.holos), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    f32, p32, c32 = ref.preprocess(case.acq.holos, go, case.acq.mask, "f32", return_crops=True)
    ec, ef = rel_l2(crops.cpu().numpy(), c32), rel_l2(fields.cpu().numpy(), f32)
    _say(f"  crops rel-L2 {ec:.3e}, fields {ef:.3e}")
    assert np.array_equal(pos.cpu().numpy(), p32)
    assert ec < TOL and ef < TOL


def test_reconstruct_plain_herraez(gpu_ctx):
    """The whole chain with Herraez unwrapping (20 angles of the 100 x 512^2
    case, Born): speculars bit-exact, index within 1e-5 of the FP32 oracle."""
    import torch
    case = make_case("cfg1", deroul=2)
    _say("reconstruct_plain herraez case=cfg1[:20]")
    holos = case.acq.holos[:20]
    g = case.lib_geom()
    f32, p32 = ref.preprocess(holos, case.g, case.acq.mask, "f32")
    idx32 = ref.reconstruct(f32, p32, case.g, "f32")[0]
    V, A = g.dim_final, holos.shape[0]
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    pos = _empty((2 * A,), torch.int32)
    gpu_ctx.reconstruct_plain(g, _t(holos), case.acq.mask, re, im, pos)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), p32)
    e32 = rel_l2(re.cpu().numpy() + 1j * im.cpu().numpy(), idx32)
    _say(f"  index rel-L2 vs f32 {e32:.3e}")
    assert e32 < TOL


def test_schedules_bit_identical(gpu_ctx):
    """The launch schedule (hologram chunk, gather window 8/16/24) changes
    no bit of the speculars or the index volume (cfg1: 100 x 512^2 ->
    256^3)."""
    import torch
    import mtd_transmission_amd as m
    case = make_case("cfg1")
    _say("schedules: hologram chunks and gather windows")
    g = case.lib_geom()
    V, A = g.dim_final, case.acq.holos.shape[0]
    holos = _t(case.acq.holos)
    outs = []
    try:
        for chunk, zw in [(0, 0), (7, 8), (0, 24), (1000, 16), (33, 24)]:
            gpu_ctx.set_schedule(chunk, zw)
            re = _empty((V, V, V), torch.float32)
            im = _empty((V, V, V), torch.float32)
            pos = _empty((2 * A,), torch.int32)
            gpu_ctx.reconstruct_plain(g, holos, case.acq.mask, re, im, pos)
            torch.cuda.synchronize()
            outs.append((re, im, pos))
        with pytest.raises(m.MTDError):
            gpu_ctx.set_schedule(0, 12)
    finally:
        gpu_ctx.set_schedule(0, 0)
    for re, im, pos in outs[1:]:
        assert torch.equal(pos, outs[0][2])
        assert torch.equal(re, outs[0][0]) and torch.equal(im, outs[0][1])


def test_specular_argmax_ties(gpu_ctx):
    """All-zero holograms among ordinary ones: every crop key of those is 0,
    the exact tie the column pass resolves with the reference's keys (first
    index, slot 0); the speculars of the whole stack equal the oracle's."""
    import torch
    case = make_case("small")
    _say("specular argmax with all-zero holograms")
    g = case.lib_geom()
    holos = case.acq.holos.copy()
    holos[1] = 0
    holos[3] = 0
    _, pos32 = ref.preprocess(holos, case.g, case.acq.mask, "f32")
    A, C = holos.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.preprocess(g, _t(holos), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    got = pos.cpu().numpy()
    assert got[1] == 0 and got[1 + A] == 0 and got[3] == 0 and got[3 + A] == 0
    assert np.array_equal(got, pos32)


@pytest.mark.parametrize("name", ["nomask", "small", "cfg1"])
def test_preprocess(gpu_ctx, name):
    import torch
    case = make_case(name)
    _say(f"preprocess case={name}")
    g = case.lib_geom()
    _, pos32, fields32 = oracle_front(case, "f32")
    _, _, fields64 = oracle_front(case, "f64")
    A, C = case.acq.holos.shape[0], g.C
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.preprocess(g, _t(case.acq.holos), case.acq.mask, pos, fields)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), pos32)
    got = fields.cpu().numpy()
    e32, e64 = rel_l2(got, fields32), rel_l2(got, fields64)
    _say(f"  fields rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert e32 < TOL and e64 <= fp64_bound(fields32, fields64)


@pytest.mark.parametrize("name", ["tiny", "small", "cfg1"])
def test_field_to_spectrum(gpu_ctx, name):
    import torch
    case = make_case(name)
    _say(f"field_to_spectrum case={name}")
    g = case.lib_geom()
    _, pos, fields32 = oracle_front(case, "f32")
    S32 = oracle_spectra(case, fields32, pos, "f32")
    S64 = oracle_spectra(case, fields32.astype(np.complex128), pos, "f64")
    A, C = fields32.shape[0], g.C
    out = _empty((A, C, C), torch.complex64)
    gpu_ctx.field_to_spectrum(g, _t(fields32), _t(pos.astype(np.int32)), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    e32, e64 = rel_l2(got, S32), rel_l2(got, S64)
    _say(f"  spectra rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert e32 < TOL and e64 <= fp64_bound(S32, S64)


@pytest.mark.parametrize("name", ["tiny", "small", "u16", "cfg1"])
def test_ewald_scatter_bit_exact(gpu_ctx, name):
    """Identical FP32 spectra in -> bit-identical counts AND values out."""
    import torch
    case = make_case(name)
    _say(f"ewald_scatter case={name}")
    g = case.lib_geom()
    _, pos, fields32 = oracle_front(case, "f32")
    S32 = oracle_spectra(case, fields32, pos, "f32")
    vol32, cnt32 = ref.ewald_scatter(S32, pos, case.g, "f32")
    V = g.dim_final
    vol = _empty((V, V, V), torch.complex64)
    cnt = _empty((V, V, V), torch.int32)
    gpu_ctx.ewald_scatter(g, _t(S32), _t(pos.astype(np.int32)), vol, cnt)
    torch.cuda.synchronize()
    gc = cnt.cpu().numpy().ravel()
    gv = vol.cpu().numpy().ravel()
    _say(f"  hits {int(cnt32.sum())} gpu {int(gc.sum())} max {int(gc.max())}")
    assert np.array_equal(gc, cnt32)
    assert np.array_equal(gv.view(np.uint32), vol32.view(np.uint32))


@pytest.mark.parametrize("name", ["tiny", "small", "cfg1"])
def test_finalize(gpu_ctx, name):
    import torch
    case = make_case(name)
    _say(f"finalize case={name}")
    g = case.lib_geom()
    _, pos, fields32 = oracle_front(case, "f32")
    S32 = oracle_spectra(case, fields32, pos, "f32")
    vol32, cnt32 = ref.ewald_scatter(S32, pos, case.g, "f32")
    idx32, otf32 = ref.finalize(vol32, cnt32, case.g, "f32")
    idx64, _ = ref.finalize(vol32.astype(np.complex128), cnt32, case.g, "f64")
    V = g.dim_final
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    otf = _empty((V, V, V), torch.uint8)
    gpu_ctx.finalize(g, _t(vol32.reshape(V, V, V)), _t(cnt32.reshape(V, V, V)), re, im, otf)
    torch.cuda.synchronize()
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    e32, e64 = rel_l2(got, idx32), rel_l2(got, idx64)
    _say(f"  index rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert np.array_equal(otf.cpu().numpy().astype(bool), otf32)
    assert e32 < TOL and e64 <= fp64_bound(idx32, idx64)


@pytest.mark.parametrize("name,c_aber", [("nomask", 1), ("small", 1), ("cfg1", 1),
                                         ("tiny", 0)])
def test_reconstruct_plain(gpu_ctx, name, c_aber):
    """Holograms to index in one call vs the oracle's two programs."""
    import torch
    case = make_case(name, c_aber=c_aber)
    _say(f"reconstruct_plain case={name}")
    g = case.lib_geom()
    fields32, pos32 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f32")
    idx32, vol32, cnt32, _ = ref.reconstruct(fields32, pos32, case.g, "f32")
    fields64, pos64 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f64")
    idx64, _, cnt64, _ = ref.reconstruct(fields64, pos64, case.g, "f64")
    assert np.array_equal(cnt32, cnt64)
    V = g.dim_final
    A = case.acq.holos.shape[0]
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    pos = _empty((2 * A,), torch.int32)
    gpu_ctx.reconstruct_plain(g, _t(case.acq.holos), case.acq.mask, re, im, pos)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), pos32)
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    e32, e64 = rel_l2(got, idx32), rel_l2(got, idx64)
    _say(f"  index rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e} (oracle f32 vs f64 {rel_l2(idx32, idx64):.3e})")
    assert e32 < TOL and e64 <= fp64_bound(idx32, idx64)


def test_accumulate_counts_match_scatter(gpu_ctx):
    """mtd_ewald_accumulate (fields in) == field_to_spectrum + ewald_scatter."""
    import torch
    case = make_case("small")
    _say("accumulate vs scatter case=small")
    g = case.lib_geom()
    _, pos, fields32 = oracle_front(case, "f32")
    A, C, V = fields32.shape[0], g.C, g.dim_final
    f = _t(fields32)
    p = _t(pos.astype(np.int32))
    S = _empty((A, C, C), torch.complex64)
    gpu_ctx.field_to_spectrum(g, f, p, S)
    v1 = _empty((V, V, V), torch.complex64)
    c1 = _empty((V, V, V), torch.int32)
    gpu_ctx.ewald_scatter(g, S, p, v1, c1)
    v2 = _empty((V, V, V), torch.complex64)
    c2 = _empty((V, V, V), torch.int32)
    gpu_ctx.ewald_accumulate(g, f, p, v2, c2)
    torch.cuda.synchronize()
    assert torch.equal(c1, c2)
    assert torch.equal(v1, v2)


def test_empty_stack_gives_zero_volume(gpu_ctx):
    """A = 0: counts and spectrum all zero (every voxel written)."""
    import torch
    case = make_case("tiny")
    _say("empty stack case=tiny")
    g = case.lib_geom()
    V = g.dim_final
    vol = torch.full((V, V, V), 7.0, dtype=torch.complex64, device="cuda")
    cnt = torch.full((V, V, V), 7, dtype=torch.int32, device="cuda")
    S = _empty((0, g.C, g.C), torch.complex64)
    gpu_ctx.ewald_scatter(g, S, _empty((0,), torch.int32), vol, cnt)
    torch.cuda.synchronize()
    assert int(cnt.abs().sum()) == 0
    assert float(vol.abs().sum()) == 0.0


def test_skipped_angle(gpu_ctx):
    """An angle whose specular lies outside the Ewald sphere contributes
    nothing (Reconstruction/src/fonctions.cpp:897)."""
    import torch
    case = make_case("tiny")
    _say("skipped angle case=tiny")
    g = case.lib_geom()
    C, V, nx = g.C, g.dim_final, g.nxmax
    rng = np.random.default_rng(3)
    S = (rng.normal(size=(2, C, C)) + 1j * rng.normal(size=(2, C, C))).astype(np.complex64)
    # corner specular: |fm0|^2 = 2*nx^2 > R^2 -> skipped
    pos = np.array([0, nx, 0, nx], dtype=np.int64)
    R = case.g.rayon
    assert 2 * nx * nx > R * R
    vol32, cnt32 = ref.ewald_scatter(S, pos, case.g, "f32")
    vol = _empty((V, V, V), torch.complex64)
    cnt = _empty((V, V, V), torch.int32)
    gpu_ctx.ewald_scatter(g, _t(S), _t(pos.astype(np.int32)), vol, cnt)
    torch.cuda.synchronize()
    P = int((ref.pupil_tables(case.g)[2]).sum())
    assert int(cnt.sum()) == P  # only the second angle
    assert np.array_equal(cnt.cpu().numpy().ravel(), cnt32)
    assert np.array_equal(vol.cpu().numpy().ravel().view(np.uint32), vol32.view(np.uint32))


def test_invalid_geometry_errors(gpu_ctx):
    import mtd_transmission_amd as m
    import torch
    case = make_case("tiny")
    g = case.lib_geom()
    g.dim_final = 16  # < 4*NXMAX: the reference would write out of bounds
    V = 16
    vol = _empty((V, V, V), torch.complex64)
    cnt = _empty((V, V, V), torch.int32)
    with pytest.raises(m.MTDError, match="DIM_FINAL"):
        gpu_ctx.ewald_scatter(g, _empty((1, g.C, g.C), torch.complex64),
                              _empty((2,), torch.int32), vol, cnt)
    g = case.lib_geom()
    g.deroul = 3  # not an mtd_deroul value
    with pytest.raises(m.MTDError, match="deroul must be"):
        gpu_ctx.holo_to_crop(g, _t(case.acq.holos), _empty((6, g.C, g.C), torch.complex64))
    g = case.lib_geom()
    g.circle_cx = 2  # crop window leaves the hologram
    with pytest.raises(m.MTDError, match="crop window"):
        gpu_ctx.holo_to_crop(g, _t(case.acq.holos), _empty((6, g.C, g.C), torch.complex64))


@pytest.mark.parametrize("pinned", [True, False])
def test_host_hologram_stack(gpu_ctx, pinned):
    """A host hologram stack (pinned or pageable) is streamed by the library,
    chunk copies overlapped with the hologram stage: crops bit-identical to
    the device stack over three chunks (150 x 1024^2 uint16, 60 per chunk),
    and the whole plain method bit-identical on the 100 x 512^2 case."""
    import torch
    import mtd_transmission_amd as m
    from cases import LAM, N0, NA
    _say(f"host hologram stack pinned={pinned}")
    n, nx, A = 1024, 128, 150
    pix = nx * LAM / (NA * n)
    g = m.geom_derive(n, nx, 512, n // 2 - round(0.3 * n), n // 2 + round(0.3 * n), NA, N0,
                      LAM, pix)
    gen = torch.Generator().manual_seed(3)
    holos_h = torch.randint(0, 60000, (A, n, n), generator=gen, dtype=torch.int32).to(torch.uint16)
    if pinned:
        holos_h = holos_h.pin_memory()
    holos_d = holos_h.cuda()
    C = g.C
    c0 = _empty((A, C, C), torch.complex64)
    c1 = _empty((A, C, C), torch.complex64)
    gpu_ctx.holo_to_crop(g, holos_d, c0)
    gpu_ctx.holo_to_crop(g, holos_h, c1)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(c0), torch.view_as_real(c1))
    case = make_case("cfg1")
    g = case.lib_geom()
    hd = _t(case.acq.holos)
    hh = torch.from_numpy(np.ascontiguousarray(case.acq.holos))
    if pinned:
        hh = hh.pin_memory()
    V = g.dim_final
    r0, i0 = _empty((V, V, V), torch.float32), _empty((V, V, V), torch.float32)
    r1, i1 = _empty((V, V, V), torch.float32), _empty((V, V, V), torch.float32)
    gpu_ctx.reconstruct_plain(g, hd, case.acq.mask, r0, i0)
    gpu_ctx.reconstruct_plain(g, hh, case.acq.mask, r1, i1)
    torch.cuda.synchronize()
    assert torch.equal(r0, r1) and torch.equal(i0, i1)


def test_host_stack_back_to_back(gpu_ctx):
    """Two host (pinned) stacks preprocessed back to back with no wait between
    the calls, each spanning two field chunks (Volkov unwrapping at crop 256:
    about 150 angles per chunk, 170 angles) and with different hologram chunk
    sizes (50, then 70): the second call's uploads must wait for the first
    call's passes that still read the staging slots. Fields and speculars are
    bit-identical to the same stacks preprocessed from device memory."""
    import torch
    import mtd_transmission_amd as m
    from cases import LAM, N0, NA
    _say("host stacks back to back, two field chunks each")
    n, nx, A = 1024, 128, 170
    pix = nx * LAM / (NA * n)
    g = m.geom_derive(n, nx, 512, n // 2 - round(0.3 * n), n // 2 + round(0.3 * n), NA, N0,
                      LAM, pix, 1.0, 1.0, 1.0, 1.0, 1, 1, m.MTD_DEROUL_VOLKOV)
    C = g.C
    gen = torch.Generator().manual_seed(11)
    hs = [torch.randint(0, 60000, (A, n, n), generator=gen, dtype=torch.int32)
          .to(torch.uint16).pin_memory() for _ in range(2)]
    ref_f, ref_p = [], []
    for h in hs:
        f = _empty((A, C, C), torch.complex64)
        p = _empty((2 * A,), torch.int32)
        gpu_ctx.preprocess(g, h.cuda(), None, p, f)
        ref_f.append(f)
        ref_p.append(p)
    torch.cuda.synchronize()
    got_f = [_empty((A, C, C), torch.complex64) for _ in range(2)]
    got_p = [_empty((2 * A,), torch.int32) for _ in range(2)]
    try:
        for k, chunk in enumerate((50, 70)):
            gpu_ctx.set_schedule(holo_chunk=chunk)
            gpu_ctx.preprocess(g, hs[k], None, got_p[k], got_f[k])
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_schedule(0, 0)
    for k in range(2):
        assert torch.equal(got_p[k], ref_p[k])
        assert torch.equal(torch.view_as_real(got_f[k]), torch.view_as_real(ref_f[k]))


def test_timing_only_one_kernel(gpu_ctx):
    """mtd_ctx_set_timing_only: the report then holds that kernel alone, with
    one entry per launch; all kernels again after the filter is cleared."""
    import torch
    case = make_case("small")
    g = case.lib_geom()
    V, A = g.dim_final, case.acq.holos.shape[0]
    holos = _t(case.acq.holos)
    re = _empty((V, V, V), torch.float32)
    im = _empty((V, V, V), torch.float32)
    try:
        gpu_ctx.timing_reset()
        gpu_ctx.set_timing(True, only="holo_rows")
        gpu_ctx.reconstruct_plain(g, holos, case.acq.mask, re, im, None)
        gpu_ctx.reconstruct_plain(g, holos, case.acq.mask, re, im, None)
        gpu_ctx.set_timing(False)
        kt = gpu_ctx.timing()
        assert set(kt) == {"holo_rows"} and kt["holo_rows"][0] >= 2
        gpu_ctx.timing_reset()
        gpu_ctx.set_timing(True)
        gpu_ctx.reconstruct_plain(g, holos, case.acq.mask, re, im, None)
        gpu_ctx.set_timing(False)
        kt = gpu_ctx.timing()
        assert {"holo_rows", "holo_cols", "ewald_gather", "vol_z"} <= set(kt)
    finally:
        gpu_ctx.set_timing(False)
        gpu_ctx.timing_reset()


@pytest.mark.parametrize("name", ["small", "cfg1"])
def test_demodulate(gpu_ctx, name):
    """mtd_demodulate (calc_Uborn2) on the library's own crops and speculars:
    within rel-L2 1e-5 of the FP32 oracle's demodulation of the oracle's
    crops, and within the FP64 bound of the FP64 oracle's."""
    import torch
    case = make_case(name)
    _say(f"demodulate case={name}")
    g = case.lib_geom()
    A, C = case.acq.holos.shape[0], g.C
    crops = _empty((A, C, C), torch.complex64)
    gpu_ctx.holo_to_crop(g, _t(case.acq.holos), crops)
    pos = _empty((2 * A,), torch.int32)
    fields = _empty((A, C, C), torch.complex64)
    gpu_ctx.preprocess(g, _t(case.acq.holos), case.acq.mask, pos, fields)
    u = _empty((A, C, C), torch.complex64)
    gpu_ctx.demodulate(g, crops, pos, u)
    torch.cuda.synchronize()
    _, p32, c32 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f32", return_crops=True)
    _, p64, c64 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f64", return_crops=True)
    assert np.array_equal(pos.cpu().numpy(), p32)
    u32 = np.stack([ref.demod(c32[a], int(p32[a]), int(p32[a + A]), "f32") for a in range(A)])
    u64 = np.stack([ref.demod(c64[a], int(p64[a]), int(p64[a + A]), "f64") for a in range(A)])
    got = u.cpu().numpy()
    e32, e64 = rel_l2(got, u32), rel_l2(got, u64)
    _say(f"  rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e}")
    assert e32 < TOL and e64 <= fp64_bound(u32, u64)
