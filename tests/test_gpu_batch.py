"""Batched form on the GPU: acquisitions of different hologram sides, crops,
volumes and angle counts in one call, eagerly and as one replayed HIP graph.

Bars: every acquisition's index volume and speculars are bit-identical to
mtd_reconstruct_plain on the same inputs (which the parity tests hold against
the oracle), after the eager first call, after each graph replay, and after
new hologram contents are written into the captured buffers; the "small"
member is also checked against the FP32 oracle (rel-L2 1e-5). The
autotuned schedules change launch shapes only, so their results are
bit-identical too."""
import numpy as np
import pytest

from cases import make_case, rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, N0, NA = 6.33e-7, 1.333, 1.2
# (hologram side, NXMAX, volume side, angles, seed): uneven angle counts,
# crops 64/128/160 with V = 2C
SHAPES = [(256, 32, 128, 21, 11), (512, 64, 256, 37, 12), (512, 80, 320, 13, 13),
          (256, 32, 128, 9, 14)]


def _say(msg):
    print(msg, flush=True)


def _acq(n, nx, V, A, seed):
    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    ph = synth.scale_phantom(synth.make_phantom(seed, 3), 0.45 * n * pix / 33.8e-6)
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph, seed=seed + 1)
    g = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    return g, holos, np.asarray(mask)


def _outputs(geoms, holos):
    import torch
    re = [torch.empty((g.dim_final,) * 3, device="cuda") for g in geoms]
    im = [torch.empty_like(r) for r in re]
    pos = [torch.empty((2 * h.shape[0],), dtype=torch.int32, device="cuda") for h in holos]
    return re, im, pos


def _plain(ctx, geoms, holos, masks):
    import torch
    re, im, pos = _outputs(geoms, holos)
    for i, g in enumerate(geoms):
        ctx.reconstruct_plain(g, holos[i], masks[i], re[i], im[i], pos[i])
    torch.cuda.synchronize()
    return [(r.clone(), j.clone(), p.clone()) for r, j, p in zip(re, im, pos)]


def _same(got, want):
    import torch
    for i, ((r, j, p), (wr, wj, wp)) in enumerate(zip(got, want)):
        assert torch.equal(p, wp), f"speculars of acquisition {i}"
        assert torch.equal(r, wr) and torch.equal(j, wj), f"volume of acquisition {i}"


def test_batch_eager_graph_and_new_contents(gpu_ctx):
    import torch
    _say("batched form: 4 acquisitions + the small oracle case")
    acqs = [_acq(*s) for s in SHAPES]
    case = make_case("small")
    geoms = [a[0] for a in acqs] + [case.lib_geom()]
    holos = [a[1] for a in acqs] + [torch.from_numpy(case.acq.holos).cuda()]
    masks = [a[2] for a in acqs] + [case.acq.mask]
    want = _plain(gpu_ctx, geoms, holos, masks)
    re, im, pos = _outputs(geoms, holos)
    ms = gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, times=True)
    torch.cuda.synchronize()
    _say(f"  eager call: per-acquisition ms {np.round(ms, 3).tolist()}")
    assert np.all(np.isfinite(ms)) and np.all(ms > 0)
    _same(list(zip(re, im, pos)), want)
    # the small member against the FP32 oracle
    f32, p32 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f32")
    idx32 = ref.reconstruct(f32, p32, case.g, "f32")[0]
    assert np.array_equal(pos[-1].cpu().numpy(), p32)
    e = rel_l2(re[-1].cpu().numpy() + 1j * im[-1].cpu().numpy(), idx32)
    _say(f"  small member index rel-L2 vs oracle f32 {e:.3e}")
    assert e < 1e-5
    # graph replays, with eager work and synchronisations between them
    for rep in range(3):
        for t in re + im:
            t.fill_(float("nan"))
        for p in pos:
            p.fill_(-1)
        torch.cuda.synchronize()
        ms = gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, times=True)
        torch.cuda.synchronize()
        _say(f"  replay {rep}: per-acquisition ms {np.round(ms, 3).tolist()}")
        _same(list(zip(re, im, pos)), want)
    # new hologram contents in the captured buffers: the replay reads them
    for h in holos:
        hv = h.view(torch.int16) if h.dtype == torch.uint16 else h
        hv.copy_(torch.roll(hv, shifts=1, dims=0))
    want2 = _plain(gpu_ctx, geoms, holos, masks)
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)
    torch.cuda.synchronize()
    _same(list(zip(re, im, pos)), want2)


def test_batch_autotune_same_results(gpu_ctx, tmp_path):
    """The autotuner's schedules (hologram chunk, gather window and batch,
    volume tile variants) leave every bit unchanged; its picks go to the
    cache file, and a new context reads them back instead of tuning."""
    import torch
    import mtd_transmission_amd as m
    _say("batched form: autotuned schedules, persisted")
    acqs = [_acq(*s) for s in SHAPES[:3]]
    geoms = [a[0] for a in acqs]
    holos = [a[1] for a in acqs]
    masks = [a[2] for a in acqs]
    want = _plain(gpu_ctx, geoms, holos, masks)
    cache = tmp_path / "tune.txt"
    ctx = m.Context(0)
    try:
        ctx.set_tune_cache(str(cache))
        re, im, pos = _outputs(geoms, holos)
        ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, autotune=True)
        torch.cuda.synchronize()
        _same(list(zip(re, im, pos)), want)
        scheds = [ctx.batch_tuned_schedule(g.n, g.C, g.dim_final) for g in geoms]
        _say(f"  tuned schedules {scheds}")
        for sc, g in zip(scheds, geoms):
            assert sc is not None and sc["from_cache"] == 0
            assert sc["holo_chunk"] > 0 and sc["gather_window"] in (8, 16, 24)
            assert sc["gather_batch"] in (8, 16) and sc["vol_rows_tile"] in (0, 1)
            assert sc["vol_cols_tile"] in (0, 1, 2)
            assert ctx.batch_tuned_chunk(g.n, g.C) == sc["holo_chunk"]
            assert ctx.batch_tuned_window(g.n, g.C) == sc["gather_window"]
        ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, autotune=True)
        torch.cuda.synchronize()
        _same(list(zip(re, im, pos)), want)
    finally:
        ctx.close()
    lines = cache.read_text().splitlines()
    _say(f"  cache file: {len(lines)} entries, first: {lines[0] if lines else None}")
    assert len(lines) == len({(g.n, g.C, g.dim_final) for g in geoms})
    assert all(ln.startswith("mtd-tune ") for ln in lines)
    ctx2 = m.Context(0)
    try:
        ctx2.set_tune_cache(str(cache))
        for sc, g in zip(scheds, geoms):
            got = ctx2.batch_tuned_schedule(g.n, g.C, g.dim_final)
            assert got is not None and got["from_cache"] == 1
            assert {k: v for k, v in got.items() if k != "from_cache"} == \
                {k: v for k, v in sc.items() if k != "from_cache"}
        re, im, pos = _outputs(geoms, holos)
        ctx2.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=False, autotune=True)
        torch.cuda.synchronize()
        _same(list(zip(re, im, pos)), want)
        assert len(cache.read_text().splitlines()) == len(lines), "cached shapes were re-tuned"
    finally:
        ctx2.close()


@pytest.mark.parametrize("shape", [(256, 32, 128, 9, 21), (512, 80, 320, 7, 22),
                                   (768, 144, 576, 5, 23), (512, 96, 384, 5, 24)])
def test_volume_tile_variants_same_bits(gpu_ctx, shape):
    """Every tile variant of the volume passes gives the default's bits."""
    import torch
    g, h, mask = _acq(*shape)
    _say(f"volume tile variants: n {g.n} C {g.C} V {g.dim_final}")
    want = _plain(gpu_ctx, [g], [h], [mask])
    try:
        for rt, ct in ((1, 0), (0, 1), (0, 2), (1, 2)):
            gpu_ctx.set_schedule_param("vol_rows_tile", rt)
            gpu_ctx.set_schedule_param("vol_cols_tile", ct)
            got = _plain(gpu_ctx, [g], [h], [mask])
            _same(got, want)
    finally:
        gpu_ctx.set_schedule_param("vol_rows_tile", 0)
        gpu_ctx.set_schedule_param("vol_cols_tile", 0)


def test_batch_graph_refuses_host_stacks(gpu_ctx):
    """A captured graph cannot hold the host-stack streaming (copy stream and
    its events): host stacks are refused in graph mode, accepted eagerly."""
    import torch
    import mtd_transmission_amd as m
    g, h, mask = _acq(*SHAPES[0])
    want = _plain(gpu_ctx, [g], [h], [mask])
    re, im, pos = _outputs([g], [h])
    with pytest.raises(m.MTDError, match="device memory"):
        gpu_ctx.reconstruct_batch([g], [h.cpu()], [mask], re, im, pos, graph=True)
    gpu_ctx.reconstruct_batch([g], [h.cpu()], [mask], re, im, pos, graph=False)
    torch.cuda.synchronize()
    _same(list(zip(re, im, pos)), want)


def test_batch_recaptures_after_workspace_moves(gpu_ctx):
    """A captured batch, then a larger plain reconstruction that makes the
    context reallocate its volume workspace: the next batch call with the
    same arguments recaptures (the old graph holds the freed addresses) and
    its results are still those of mtd_reconstruct_plain."""
    import torch
    _say("batched form: recapture after a workspace reallocation")
    acqs = [_acq(*s) for s in (SHAPES[0], SHAPES[3])]
    geoms = [a[0] for a in acqs]
    holos = [a[1] for a in acqs]
    masks = [a[2] for a in acqs]
    re, im, pos = _outputs(geoms, holos)
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)  # replay
    torch.cuda.synchronize()
    # a 512^3 volume: larger than any workspace the batch captured
    gb, hb, mb = _acq(512, 64, 512, 5, 21)
    rb, ib, pb = _outputs([gb], [hb])
    gpu_ctx.reconstruct_plain(gb, hb, mb, rb[0], ib[0], pb[0])
    torch.cuda.synchronize()
    want = _plain(gpu_ctx, geoms, holos, masks)
    for t in re + im:
        t.fill_(float("nan"))
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)  # replay
    torch.cuda.synchronize()
    _same(list(zip(re, im, pos)), want)
