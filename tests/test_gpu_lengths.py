"""Every transform length the library builds that the other parity tests do
not reach, through the plain method against the oracle: crops 64, 160, 224,
288, 352, 384 (radix-2/3/5/7/11 register FFTs and their two-step splits),
volumes 2C = 128, 320, 448, 576, 704, 768, and 256^2-1024^2 holograms.
Per case (3 angles, seeded ellipsoid phantom, Fermat spiral, Poisson noise):
speculars bit-exact; fields within rel-L2 1e-5 of the FP32 oracle and within
the FP64 bound of the FP64 oracle; the index volume within 1e-5 of the FP32
oracle (dense up to 320^3, else at 400 sampled voxels of the oracle's exact
sum) and within the FP64 bound of the FP64 oracle.

The batched form at its real shapes: the 32 acquisitions of the batched
workload (holograms 512^2-1024^2, crops 128-384, V = 2C) with 5 angles each,
one captured graph, replayed; each member against the oracle."""
import numpy as np
import pytest

import lengths as L
from cases import rel_l2
from oracle import ref

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


# (hologram side N, NXMAX): crop C = 2 NXMAX, volume V = 2C; N >= 2.5 C keeps
# the sideband window inside the hologram
CASES = [(256, 32), (512, 80), (768, 112), (768, 144), (1024, 176), (1024, 192)]


def _check_index(re, im, fields32, fields64, pos, go, seed, dense_max=320, samples=400):
    V = go.dim_final
    got_full = re.cpu().numpy().astype(np.float64) + 1j * im.cpu().numpy()
    if V <= dense_max:
        o32 = ref.reconstruct(fields32, pos, go, "f32")[0]
        o64 = ref.reconstruct(fields64, pos, go, "f64")[0]
        e32, e64 = rel_l2(got_full, o32), rel_l2(got_full, o64)
        bound = L.fp64_bound(o32, o64)
    else:
        idx = np.random.default_rng(seed).integers(0, V ** 3, size=samples)
        got = got_full.ravel()[idx]
        o32 = L.index_at(fields32, pos, go, "f32", idx)
        o64 = L.index_at(fields64, pos, go, "f64", idx)
        e32, e64 = rel_l2(got, o32), rel_l2(got, o64)
        bound = L.fp64_bound(o32, o64)
    return e32, e64, bound


# sides without a compiled register FFT (the runtime-length FFTs), as
# (N, NXMAX, V, unwrapping): direct mixed-radix (600 = 8*3*25, 182 = 2*7*13,
# 364 = 4*7*13, 104 = 8*13, 208 = 16*13) and Bluestein (crop 122 = 2*61 and
# volume 254 = 2*127: chirp-z over 256 and 512 points; hologram 526 = 2*263:
# over 2048), with and without Volkov unwrapping (its 2C-point transforms);
# crop 250 (the shipped Tomo_config's NXMAX 125) and volume 416 (the shipped
# simulator configuration's DIM_FINAL) are compiled lengths
GEN_CASES = [(512, 61, 254, 0), (512, 61, 254, 1), (600, 91, 364, 0), (526, 52, 208, 1),
             (1024, 125, 512, 1), (1024, 61, 416, 0)]


@pytest.mark.parametrize("n,nx,V,deroul", GEN_CASES)
def test_plain_runtime_lengths(gpu_ctx, n, nx, V, deroul):
    _plain_case(gpu_ctx, n, nx, V, 3, deroul)


@pytest.mark.parametrize("n,nx", CASES)
def test_plain_at_length(gpu_ctx, n, nx):
    _plain_case(gpu_ctx, n, nx, 4 * nx, 3, 0)


def _plain_case(gpu_ctx, n, nx, V, A, deroul):
    import torch
    C = 2 * nx
    _say(f"plain method N={n} crop={C} V={V} A={A} deroul={deroul}")
    holos, mask, gl, go = L.acquisition(n, nx, V, A, seed=100 + nx, deroul=deroul)
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    fields = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.preprocess(gl, holos, mask, pos, fields)
    pos2 = torch.empty_like(pos)
    gpu_ctx.reconstruct_plain(gl, holos, mask, re, im, pos2)
    torch.cuda.synchronize()
    h = holos.cpu().numpy()
    f32, p32 = ref.preprocess(h, go, mask, "f32")
    f64, p64 = ref.preprocess(h, go, mask, "f64")
    assert np.array_equal(pos.cpu().numpy(), p32) and np.array_equal(pos2.cpu().numpy(), p32)
    got_f = fields.cpu().numpy()
    ef32, ef64, fb = rel_l2(got_f, f32), rel_l2(got_f, f64), L.fp64_bound(f32, f64)
    _say(f"  fields rel-L2 vs f32 {ef32:.3e}  vs f64 {ef64:.3e} (bound {fb:.2e})")
    e32, e64, ib = _check_index(re, im, f32, f64, p32, go, seed=nx)
    _say(f"  index rel-L2 vs f32 {e32:.3e}  vs f64 {e64:.3e} (bound {ib:.2e})")
    assert ef32 < TOL and ef64 <= fb
    assert e32 < TOL and e64 <= ib


def test_batched_real_shapes(gpu_ctx):
    import torch
    import bench
    shapes = bench.batch_shapes()
    A = 5
    geoms, gos, holos, masks = [], [], [], []
    for i, s in enumerate(shapes):
        h, mask, gl, go = L.acquisition(s["n"], s["nxmax"], s["v"], A, seed=3000 + i)
        geoms.append(gl)
        gos.append(go)
        holos.append(h)
        masks.append(mask)
    _say(f"batched form: {len(shapes)} acquisitions x {A} angles, crops "
         f"{sorted({g.C for g in geoms})}")
    re = [torch.empty((g.dim_final,) * 3, device="cuda") for g in geoms]
    im = [torch.empty_like(r) for r in re]
    pos = [torch.empty((2 * A,), dtype=torch.int32, device="cuda") for _ in geoms]
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)
    for r in re:
        r.fill_(float("nan"))
    gpu_ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True)  # replay
    torch.cuda.synchronize()
    worst = 0.0
    for i, go in enumerate(gos):
        hh = holos[i].cpu().numpy()
        f32, p32 = ref.preprocess(hh, go, masks[i], "f32")
        assert np.array_equal(pos[i].cpu().numpy(), p32), f"acquisition {i}: speculars"
        V = go.dim_final
        idx = np.random.default_rng(i).integers(0, V ** 3, size=200)
        got = (re[i].view(-1)[torch.from_numpy(idx).cuda()].double().cpu().numpy()
               + 1j * im[i].view(-1)[torch.from_numpy(idx).cuda()].double().cpu().numpy())
        e = rel_l2(got, L.index_at(f32, p32, go, "f32", idx))
        worst = max(worst, e)
        assert e < TOL, f"acquisition {i} ({go.n}, C {go.C}, V {V}): index rel-L2 {e:.3e}"
    _say(f"  worst index rel-L2 vs f32 oracle {worst:.3e}")
